// The MI355X model's side of the reference's model-plugin type contract
// (driver/definitions.h:22 includes this file; the driver instantiates these
// types in driver/definitions.cpp:4-41 and reads them in driver/shared.cpp:65-93
// and driver/visit.cpp:118-187).  Buffers are device allocations made through the
// clover_mi355x C ABI; no HIP type appears here, so the driver compiles with any
// host C++17 compiler.
//
// Handles are shallow: the driver moves and copies tiles (driver/start.cpp:65-66)
// and the kernels only need the device address, so a Buffer is a (size, pointer)
// pair without a destructor; finalise() releases every field's storage.
//
// 2D layout: x-contiguous with one row pitch shared by every 2D buffer of a tile
// and at least sizeY + 1 rows, which is what clv_chunk_wrap needs (each array
// covers (y_cells + 5) rows of pitch >= x_cells + 5).  The pitch is fixed by the
// first 2D buffer a context allocates (density0, a cell array, xrange = nx + 4,
// driver/definitions.cpp:6) as the next multiple of 16 doubles above it, so the
// one-larger node and face arrays fit the same pitch.  Element (0, 0) sits 14 elements
// past the allocation's 256-byte boundary, so element (2, j), the first interior cell
// of every row, starts a 128-byte line: a wave's 64 lanes then store whole lines (a line
// written partly by two waves costs HBM time).  With CLV_ENTRY_FUSED the model exchanges
// buffers between handles (advection), so every 2D buffer has the same offset and
// release() frees data - kLead2D.
#pragma once

#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "clover_mi355x.h"
#include "shared.h"

// runs the model's recorded timestep chain, if any (initialise.cpp), so that a host
// copy sees the arrays the reference's sequence leaves
void mi355x_flush_deferred();

namespace clover {

struct context {
  int device = 0;
  size_t pitch = 0;  // doubles per row of every 2D buffer; set by the first allocation
};

struct chunk_context {};

inline void check_clv(clv_status st, const char *what) {
  if (st != CLV_OK) {
    std::fprintf(stderr, "mi355x model: %s failed (%d): %s\n", what, st, clv_last_error());
    std::abort();
  }
}

template <typename T> struct Buffer1D {
  size_t size;
  T *data;
  Buffer1D(context &, size_t size) : size(size), data(nullptr) {
    void *p = nullptr;
    check_clv(clv_device_alloc(size * sizeof(T), &p), "Buffer1D allocation");
    data = static_cast<T *>(p);
  }
  void release() {
    check_clv(clv_device_free(data), "Buffer1D release");
    data = nullptr;
  }
  T *actual() { return data; }
  template <size_t D> [[nodiscard]] size_t extent() const {
    static_assert(D < 1);
    return size;
  }
  std::vector<T> mirrored() const {
    mi355x_flush_deferred();
    std::vector<T> buffer(size);
    check_clv(clv_copy_to_host(buffer.data(), data, size * sizeof(T)), "Buffer1D mirrored");
    return buffer;
  }
};

template <typename T> struct Buffer2D {
  static constexpr size_t kLead2D = 128 / sizeof(T) - 2;  // element (2, j) on a 128-byte line
  size_t sizeX, sizeY, pitch;
  T *data;
  Buffer2D(context &ctx, size_t sizeX, size_t sizeY) : sizeX(sizeX), sizeY(sizeY), pitch(0), data(nullptr) {
    if (ctx.pitch == 0 || ctx.pitch < sizeX) ctx.pitch = (sizeX + 1 + 15) / 16 * 16;
    pitch = ctx.pitch;
    void *p = nullptr;
    check_clv(clv_device_alloc(((sizeY + 1) * pitch + kLead2D) * sizeof(T), &p), "Buffer2D allocation");
    data = static_cast<T *>(p) + kLead2D;
  }
  void release() {
    check_clv(clv_device_free(data - kLead2D), "Buffer2D release");
    data = nullptr;
  }
  T *actual() { return data; }
  template <size_t D> [[nodiscard]] size_t extent() const {
    static_assert(D < 2);
    return D == 0 ? sizeX : sizeY;
  }
  // host copy in the order BufferMirror2D indexes it (driver/shared.h:30-37:
  // element (i, j) at [j + i * sizeY])
  std::vector<T> mirrored() const {
    mi355x_flush_deferred();
    std::vector<T> rows((sizeY + 1) * pitch);
    check_clv(clv_copy_to_host(rows.data(), data, rows.size() * sizeof(T)), "Buffer2D mirrored");
    std::vector<T> out(sizeX * sizeY);
    for (size_t i = 0; i < sizeX; ++i)
      for (size_t j = 0; j < sizeY; ++j) out[j + i * sizeY] = rows[j * pitch + i];
    return out;
  }
  clover::BufferMirror2D<T> mirrored2() { return {mirrored(), extent<0>(), extent<1>()}; }
};

template <typename T> using StagingBuffer1D = T *;

}  // namespace clover

using clover::Range1d;
using clover::Range2d;
