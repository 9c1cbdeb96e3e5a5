// Chunk-to-chunk exchange entry points (driver/comms_kernel.h:26-41).  With one
// chunk every neighbour is an external face and clover_exchange has nothing to
// send (src/serial/comms_kernel.cpp:85-190 skips every face); a decomposed run
// stops with a message.
#include "comms_kernel.h"

#include "mi355x_model.h"

void clover_allocate_buffers(global_variables &, parallel_ &) {}

void clover_exchange(global_variables &globals, const int fields[NUM_FIELDS], int depth) {
  (void)fields;
  (void)depth;
  for (int face = 0; face < 4; ++face)
    if (globals.chunk.chunk_neighbours[face] != external_face) mi355x::unsupported("clover_exchange");
}

void clover_send_recv_message_left(global_variables &, clover::StagingBuffer1D<double>, clover::StagingBuffer1D<double>, int, int, int,
                                   MPI_Request &, MPI_Request &) {
  mi355x::unsupported("clover_send_recv_message_left");
}
void clover_send_recv_message_right(global_variables &, clover::StagingBuffer1D<double>, clover::StagingBuffer1D<double>, int, int, int,
                                    MPI_Request &, MPI_Request &) {
  mi355x::unsupported("clover_send_recv_message_right");
}
void clover_send_recv_message_top(global_variables &, clover::StagingBuffer1D<double>, clover::StagingBuffer1D<double>, int, int, int,
                                  MPI_Request &, MPI_Request &) {
  mi355x::unsupported("clover_send_recv_message_top");
}
void clover_send_recv_message_bottom(global_variables &, clover::StagingBuffer1D<double>, clover::StagingBuffer1D<double>, int, int, int,
                                     MPI_Request &, MPI_Request &) {
  mi355x::unsupported("clover_send_recv_message_bottom");
}
