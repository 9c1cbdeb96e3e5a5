// Cell volume and face areas of a chunk for the per-kernel entry points and the
// reductions.  While d.geom_mesh (a chunk in CLV_GEOMETRY_MESH mode whose arrays still
// hold what initialise_chunk wrote, src/serial/initialise_chunk.cpp:78-84: volume =
// dx * dy, xarea(i, j) = celldy(j), yarea(i, j) = celldx(i) over i < nx + 4,
// j < ny + 4, zero past that) they are computed from the 1D spacings -- volume as
// celldx(i) * celldy(j), the same doubles the arrays hold -- instead of read;
// otherwise (the default) they are read.  The flag is uniform over a launch.
#pragma once

#include "clv_internal.h"

namespace clv {

__device__ __forceinline__ bool geom_in(const Dev &d, int i, int j) { return i < d.nx + 4 && j < d.ny + 4; }
__device__ __forceinline__ double geom_volume(const Dev &d, int i, int j) {
  if (d.geom_mesh) return geom_in(d, i, j) ? d.m[CLV_CELLDX][i] * d.m[CLV_CELLDY][j] : 0.0;
  return d.f[CLV_VOLUME][(int64_t)j * d.pitch + i];
}
__device__ __forceinline__ double geom_xarea(const Dev &d, int i, int j) {
  if (d.geom_mesh) return geom_in(d, i, j) ? d.m[CLV_CELLDY][j] : 0.0;
  return d.f[CLV_XAREA][(int64_t)j * d.pitch + i];
}
__device__ __forceinline__ double geom_yarea(const Dev &d, int i, int j) {
  if (d.geom_mesh) return geom_in(d, i, j) ? d.m[CLV_CELLDX][i] : 0.0;
  return d.f[CLV_YAREA][(int64_t)j * d.pitch + i];
}

}  // namespace clv
