// Host-side launchers of the gfx950 kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "rs_device.hpp"

namespace rs {

// One digit-group pass over `nf` fields (blockIdx.y = field).  `row` selects
// the contiguous lowest group; otherwise a strided column group of stride g.S.
template <typename S>
hipError_t launch_fiber_pass(int lgL, bool row, const FieldArgs *args, int nf, const Geom &g,
                             hipStream_t st);
template <typename S> hipError_t launch_xop(const FieldArgs *args, int nf, int64_t N, hipStream_t st);
template <typename S> hipError_t launch_kop(const FieldArgs *args, int nf, const Geom &g, hipStream_t st);
template <typename S>
hipError_t launch_detect(const DetArgs *args, int nf, const int64_t *bounds, int64_t npx, hipStream_t st);
hipError_t launch_fill_f64(double *p, int64_t n, double v, hipStream_t st);
template <typename S> hipError_t launch_copy(void *dst, const void *src, int64_t n, hipStream_t st);

// supported fiber lengths (log2) of the row and column passes
constexpr int kMinRowLg = 6;
constexpr int kMaxRowLgF32 = 13;
constexpr int kMaxRowLgF64 = 12;
constexpr int kMinColLg = 4;
constexpr int kMaxColLg = 9;

}  // namespace rs
