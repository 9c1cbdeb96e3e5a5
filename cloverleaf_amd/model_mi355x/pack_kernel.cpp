// Halo message packing (driver/pack_kernel.h:25-48): only reached with more than
// one chunk, which this model does not implement.
#include "pack_kernel.h"

#include "mi355x_model.h"

#define CLV_UNSUPPORTED_PACK(name)                                                                                                   \
  void name(global_variables &, int, int, int, int, clover::Buffer2D<double> &, clover::Buffer1D<double> &, int, int, int, int, int, int, \
            int) {                                                                                                                   \
    mi355x::unsupported(#name);                                                                                                      \
  }
CLV_UNSUPPORTED_PACK(clover_pack_message_left)
CLV_UNSUPPORTED_PACK(clover_unpack_message_left)
CLV_UNSUPPORTED_PACK(clover_pack_message_right)
CLV_UNSUPPORTED_PACK(clover_unpack_message_right)
CLV_UNSUPPORTED_PACK(clover_pack_message_top)
CLV_UNSUPPORTED_PACK(clover_unpack_message_top)
CLV_UNSUPPORTED_PACK(clover_pack_message_bottom)
CLV_UNSUPPORTED_PACK(clover_unpack_message_bottom)
