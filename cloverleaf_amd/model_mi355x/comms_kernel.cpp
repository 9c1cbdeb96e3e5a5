// Chunk-to-chunk halo exchange (driver/comms_kernel.h:26-41) for a mesh split
// over several ranks, one chunk each.  The flow of src/serial/comms_kernel.cpp:54-190
// on the device: the boundary tiles pack their face messages into device buffers
// (the driver's clover_pack_* -> clv_pack_message, the reference's layout), the
// messages are staged through host memory for MPI, and the received messages are
// unpacked on the device; left/right first, then bottom/top, so corner ghosts
// travel in two hops.  With one chunk every face is external and nothing moves.
#include "comms_kernel.h"

#include <memory>
#include <vector>

#include "mi355x_model.h"

void clover_allocate_buffers(global_variables &, parallel_ &) {}

namespace {

// device messages (send, receive) and host staging for one face pair, sized for
// every field at depth 2
struct FaceBuffers {
  std::unique_ptr<clover::Buffer1D<double>> snd, rcv;
  std::vector<double> host_snd, host_rcv;
  void ensure(clover::context &ctx, size_t n) {
    if (snd && snd->size >= n) return;
    snd = std::make_unique<clover::Buffer1D<double>>(ctx, n);
    rcv = std::make_unique<clover::Buffer1D<double>>(ctx, n);
    host_snd.assign(n, 0.0);
    host_rcv.assign(n, 0.0);
  }
};
FaceBuffers g_face[4];

// the tiles whose `face` lies on the chunk's face (the driver's external_tile_mask
// keeps bottom and top in one slot, driver/definitions.h:58)
bool on_chunk_face(const global_variables &globals, int tile, int face) {
  const tile_info &t = globals.chunk.tiles[tile].info;
  const chunk_type &c = globals.chunk;
  switch (face) {
    case chunk_left: return t.t_left == c.left;
    case chunk_right: return t.t_right == c.right;
    case chunk_bottom: return t.t_bottom == c.bottom;
    default: return t.t_top == c.top;
  }
}

void sync_tiles(global_variables &globals) {
  for (int tile = 0; tile < globals.config.tiles_per_chunk; ++tile)
    mi355x::check(clv_chunk_synchronize(mi355x::chunk(globals, tile)), "clover_exchange");
}

// pack every boundary tile of `face`, stage to the host and post the send/receive
void post(global_variables &globals, int face, const int fields[NUM_FIELDS], int depth, int offsets[NUM_FIELDS], int total, int tag_send,
          int tag_recv, MPI_Request &rs, MPI_Request &rr) {
  FaceBuffers &b = g_face[face];
  for (int tile = 0; tile < globals.config.tiles_per_chunk; ++tile) {
    if (!on_chunk_face(globals, tile, face)) continue;
    if (face == chunk_left) clover_pack_left(globals, *b.snd, tile, fields, depth, offsets);
    if (face == chunk_right) clover_pack_right(globals, *b.snd, tile, fields, depth, offsets);
    if (face == chunk_bottom) clover_pack_bottom(globals, *b.snd, tile, fields, depth, offsets);
    if (face == chunk_top) clover_pack_top(globals, *b.snd, tile, fields, depth, offsets);
  }
  sync_tiles(globals);
  mi355x::check(clv_copy_to_host(b.host_snd.data(), b.snd->data, (size_t)total * sizeof(double)), "clover_exchange staging");
  if (face == chunk_left) clover_send_recv_message_left(globals, b.host_snd.data(), b.host_rcv.data(), total, tag_send, tag_recv, rs, rr);
  if (face == chunk_right) clover_send_recv_message_right(globals, b.host_snd.data(), b.host_rcv.data(), total, tag_send, tag_recv, rs, rr);
  if (face == chunk_bottom)
    clover_send_recv_message_bottom(globals, b.host_snd.data(), b.host_rcv.data(), total, tag_send, tag_recv, rs, rr);
  if (face == chunk_top) clover_send_recv_message_top(globals, b.host_snd.data(), b.host_rcv.data(), total, tag_send, tag_recv, rs, rr);
}

// copy the received message to the device and unpack it into the boundary tiles
void receive(global_variables &globals, int face, const int fields[NUM_FIELDS], int depth, int offsets[NUM_FIELDS], int total) {
  FaceBuffers &b = g_face[face];
  mi355x::check(clv_copy_to_device(b.rcv->data, b.host_rcv.data(), (size_t)total * sizeof(double)), "clover_exchange staging");
  for (int tile = 0; tile < globals.config.tiles_per_chunk; ++tile) {
    if (!on_chunk_face(globals, tile, face)) continue;
    if (face == chunk_left) clover_unpack_left(globals, *b.rcv, fields, tile, depth, offsets);
    if (face == chunk_right) clover_unpack_right(globals, *b.rcv, fields, tile, depth, offsets);
    if (face == chunk_bottom) clover_unpack_bottom(globals, *b.rcv, fields, tile, depth, offsets);
    if (face == chunk_top) clover_unpack_top(globals, *b.rcv, fields, tile, depth, offsets);
  }
}

}  // namespace

void clover_exchange(global_variables &globals, const int fields[NUM_FIELDS], int depth) {
  const auto &nb = globals.chunk.chunk_neighbours;
  if (nb[chunk_left] == external_face && nb[chunk_right] == external_face && nb[chunk_bottom] == external_face &&
      nb[chunk_top] == external_face)
    return;
  // per-field message offsets, src/serial/comms_kernel.cpp:64-73
  int lr_offset[NUM_FIELDS], bt_offset[NUM_FIELDS];
  int end_lr = 0, end_bt = 0;
  for (int f = 0; f < NUM_FIELDS; ++f) {
    if (fields[f] != 1) continue;
    lr_offset[f] = end_lr;
    bt_offset[f] = end_bt;
    end_lr += depth * (globals.chunk.y_max + 5);
    end_bt += depth * (globals.chunk.x_max + 5);
  }
  for (int face = 0; face < 4; ++face)
    g_face[face].ensure(globals.context, (size_t)NUM_FIELDS * 2 * ((face <= chunk_right ? globals.chunk.y_max : globals.chunk.x_max) + 5));
  MPI_Request req[4] = {};
  int count = 0;
  if (nb[chunk_left] != external_face) {
    post(globals, chunk_left, fields, depth, lr_offset, end_lr, 1, 2, req[count], req[count + 1]);
    count += 2;
  }
  if (nb[chunk_right] != external_face) {
    post(globals, chunk_right, fields, depth, lr_offset, end_lr, 2, 1, req[count], req[count + 1]);
    count += 2;
  }
  MPI_Waitall(count, req, MPI_STATUS_IGNORE);
  if (nb[chunk_left] != external_face) receive(globals, chunk_left, fields, depth, lr_offset, end_lr);
  if (nb[chunk_right] != external_face) receive(globals, chunk_right, fields, depth, lr_offset, end_lr);
  count = 0;
  for (MPI_Request &r : req) r = {};
  if (nb[chunk_bottom] != external_face) {
    post(globals, chunk_bottom, fields, depth, bt_offset, end_bt, 3, 4, req[count], req[count + 1]);
    count += 2;
  }
  if (nb[chunk_top] != external_face) {
    post(globals, chunk_top, fields, depth, bt_offset, end_bt, 4, 3, req[count], req[count + 1]);
    count += 2;
  }
  MPI_Waitall(count, req, MPI_STATUS_IGNORE);
  if (nb[chunk_top] != external_face) receive(globals, chunk_top, fields, depth, bt_offset, end_bt);
  if (nb[chunk_bottom] != external_face) receive(globals, chunk_bottom, fields, depth, bt_offset, end_bt);
}

// MPI on the host-staged messages; the neighbour's rank is its chunk number - 1
// (src/serial/comms_kernel.cpp:192-245)
#define CLV_SEND_RECV(name, face)                                                                                                   \
  void name(global_variables &globals, double *snd, double *rcv, int total_size, int tag_send, int tag_recv, MPI_Request &req_send, \
            MPI_Request &req_recv) {                                                                                                \
    const int task = globals.chunk.chunk_neighbours[face] - 1;                                                                      \
    MPI_Isend(snd, total_size, MPI_DOUBLE, task, tag_send, MPI_COMM_WORLD, &req_send);                                              \
    MPI_Irecv(rcv, total_size, MPI_DOUBLE, task, tag_recv, MPI_COMM_WORLD, &req_recv);                                              \
  }
CLV_SEND_RECV(clover_send_recv_message_left, chunk_left)
CLV_SEND_RECV(clover_send_recv_message_right, chunk_right)
CLV_SEND_RECV(clover_send_recv_message_top, chunk_top)
CLV_SEND_RECV(clover_send_recv_message_bottom, chunk_bottom)
