// Exchange-message packing (driver/pack_kernel.h:23-47), reached from the driver's
// clover_pack_* / clover_unpack_* (driver/comms.cpp:282-930) when the mesh is
// split over several chunks.  One launch per field and face on the device
// (clv_pack_message / clv_unpack_message): the message stays in the device
// buffer the driver passes, in the reference's layout; the field's tile is found
// from the buffer.
#include "pack_kernel.h"

#include "mi355x_model.h"

namespace {
void pack(global_variables &globals, clover::Buffer2D<double> &field, clover::Buffer1D<double> &buf, int face, int depth, int field_type,
          int offset, bool unpack, const char *what) {
  clv_chunk *c = mi355x::chunk_of(globals, field.data);
  const clv_status st = unpack ? clv_unpack_message(c, field.data, field_type, face, depth, buf.data, buf.size, offset)
                               : clv_pack_message(c, field.data, field_type, face, depth, buf.data, buf.size, offset);
  mi355x::check(st, what);
}
}  // namespace

#define CLV_PACK(name, face, unpack)                                                                                                   \
  void name(global_variables &globals, int, int, int, int, clover::Buffer2D<double> &field, clover::Buffer1D<double> &buf, int, int, \
            int, int, int depth, int field_type, int buffer_offset) {                                                                \
    pack(globals, field, buf, face, depth, field_type, buffer_offset, unpack, #name);                                              \
  }
CLV_PACK(clover_pack_message_left, CLV_FACE_LEFT, false)
CLV_PACK(clover_unpack_message_left, CLV_FACE_LEFT, true)
CLV_PACK(clover_pack_message_right, CLV_FACE_RIGHT, false)
CLV_PACK(clover_unpack_message_right, CLV_FACE_RIGHT, true)
CLV_PACK(clover_pack_message_top, CLV_FACE_TOP, false)
CLV_PACK(clover_unpack_message_top, CLV_FACE_TOP, true)
CLV_PACK(clover_pack_message_bottom, CLV_FACE_BOTTOM, false)
CLV_PACK(clover_unpack_message_bottom, CLV_FACE_BOTTOM, true)
