// gfx950 kernels of the split-step path other than the fiber passes
// (fiber.hpp, instantiated by fiber_inst.hip):
//
//  mask_fill_kernel (fiber.hpp) -- grating decisions of a top column pass,
//                  cached as bits for the passes of a run that share them.
//  xop_kernel   -- x-space multiply / generation as its own pass (plain form).
//  kop_kernel   -- spectrum multiply as its own pass (plain form).
//  detect_kernel-- per-pixel FP64 sums of |u|^2 over host-computed sample
//                  ranges (big-wave histogram binning), weighted accumulate.
//
// Reference counterparts: fast-wave/fwcuda/kernels.cu (propagate_convolve_step
// :40-65, apply_grating_factors :67-89, scale :160-172, square_and_downsample
// :174-208).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <mutex>

#include "device_ops.hpp"
#include "engine.hpp"
#include "fiber.hpp"
#include "rs_detector.hpp"
#include "rs_kernels.hpp"

namespace rs {

// the fiber-pass variants are compiled in their own translation units
extern template hipError_t launch_fiber_xv<float, XV_NONE>(int, bool, const FieldArgs *, int, const Geom &,
                                                           hipStream_t, hipEvent_t, hipEvent_t);
extern template hipError_t launch_fiber_xv<float, XV_MASK>(int, bool, const FieldArgs *, int, const Geom &,
                                                           hipStream_t, hipEvent_t, hipEvent_t);
extern template hipError_t launch_fiber_xv<float, XV_GEN>(int, bool, const FieldArgs *, int, const Geom &,
                                                          hipStream_t, hipEvent_t, hipEvent_t);
extern template hipError_t launch_fiber_xv<double, XV_NONE>(int, bool, const FieldArgs *, int, const Geom &,
                                                            hipStream_t, hipEvent_t, hipEvent_t);
extern template hipError_t launch_fiber_xv<double, XV_MASK>(int, bool, const FieldArgs *, int, const Geom &,
                                                            hipStream_t, hipEvent_t, hipEvent_t);
extern template hipError_t launch_fiber_xv<double, XV_GEN>(int, bool, const FieldArgs *, int, const Geom &,
                                                           hipStream_t, hipEvent_t, hipEvent_t);

// ---------------------------------------------------------------------------
// plain-form elementwise passes
// ---------------------------------------------------------------------------

template <typename S>
__global__ void __launch_bounds__(256) xop_kernel(const FieldArgs *__restrict__ args, int64_t N) {
  using T2 = typename C2T<S>::t;
  const FieldArgs &A = args[blockIdx.y];
  T2 *data = (T2 *)A.data;
  const bool gen = A.x.kind == XOP_POINT || A.x.kind == XOP_ONES || A.x.from_ones;
  for (int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x; n < N; n += (int64_t)gridDim.x * 256) {
    T2 v;
    if (gen) { v.x = 1; v.y = 0; } else v = data[n];
    data[n] = apply_xop(A.x, n, N, v);
  }
}

template <typename S>
__global__ void __launch_bounds__(256) kop_kernel(const FieldArgs *__restrict__ args, Geom g) {
  using T2 = typename C2T<S>::t;
  const FieldArgs &A = args[blockIdx.y];
  T2 *data = (T2 *)A.data;
  const int64_t mask = (int64_t(1) << g.lgR) - 1;
  const int64_t kstep = g.N >> g.lgR;
  for (int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x; p < g.N; p += (int64_t)gridDim.x * 256) {
    const int64_t k = row_kbase(g, p >> g.lgR) + kstep * (p & mask);
    data[p] = apply_kop(A.k, k, g.N, data[p]);
  }
}

// ---------------------------------------------------------------------------
// detector
// ---------------------------------------------------------------------------

// Sum of a 256-thread workgroup's values in a fixed order: each wavefront
// reduces its 64 lanes with cross-lane xor exchanges (ds_bpermute, no LDS
// storage, no barrier), then thread 0 adds the four wavefront sums, so the
// order never depends on scheduling.  Valid in thread 0; red4 holds four
// doubles and may be reused after the next barrier.
__device__ __forceinline__ double block_sum256(double s, double *red4) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if ((threadIdx.x & 63) == 0) red4[threadIdx.x >> 6] = s;
  __syncthreads();
  return (red4[0] + red4[1]) + (red4[2] + red4[3]);
}

// One workgroup per pixel; the members of the launch are visited in order
// inside the workgroup, so members that share an accumulator row add into it
// without races and in a fixed order (results do not depend on scheduling).
template <typename S>
__global__ void __launch_bounds__(256) detect_kernel(const DetArgs *__restrict__ args, int nf,
                                                     const int64_t *__restrict__ bounds) {
  using T2 = typename C2T<S>::t;
  __shared__ double red[4];
  const int k = blockIdx.x;
  const int64_t lo = bounds[k], hi = bounds[k + 1];
  for (int f = 0; f < nf; ++f) {
    const DetArgs &A = args[f];
    const T2 *u = (const T2 *)A.data;
    double s = 0.0;
    for (int64_t j = lo + threadIdx.x; j < hi; j += 256) {
      const T2 z = u[j];
      s += (double)z.x * (double)z.x + (double)z.y * (double)z.y;
    }
    const double tot = block_sum256(s, red);
    if (threadIdx.x == 0) {
      const double prev = A.accumulate ? A.acc[k] : 0.0;
      A.acc[k] = prev + A.weight * tot;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// allocation-free detector read (rs_square_and_downsample_noalloc)
// ---------------------------------------------------------------------------

// Device memory of the allocation-free detector read: the FP64 pixel sums
// (kNoAllocMaxPixels doubles), then the pixel bounds (kNoAllocMaxPixels + 1
// int64).  One copy per GPU, kept until the process ends; host code reaches
// it only through the address noalloc_scratch() looked up.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};
}  // namespace

void *noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) throw Error(1, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    int cur = -1;
    hip_check(hipGetDevice(&cur), "hipGetDevice");
    if (cur != device) throw Error(1, "noalloc_scratch: device is not current");
    void *p = nullptr;
    hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    g_scratch_ptr[device] = p;
  }
  return g_scratch_ptr[device];
}

// bounds[k] for k <= npx (one thread each, binary search over the samples)
__global__ void __launch_bounds__(256) det_bounds_kernel(DetGeom g, int64_t npx, int64_t *__restrict__ bounds) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k <= npx) bounds[k] = det_bound(g, k);
}

// acc[k] = sum of |u_j|^2 over bounds[k] <= j < bounds[k + 1] in FP64, one
// workgroup per pixel: the same sums, in the same order, as detect_kernel
// for one field written into zeroed accumulators with weight 1
template <typename S>
__global__ void __launch_bounds__(256) detect_one_kernel(const void *__restrict__ data, double *__restrict__ acc,
                                                         const int64_t *__restrict__ bounds) {
  using T2 = typename C2T<S>::t;
  __shared__ double red[4];
  const int k = blockIdx.x;
  const int64_t lo = bounds[k], hi = bounds[k + 1];
  const T2 *u = (const T2 *)data;
  double s = 0.0;
  for (int64_t j = lo + threadIdx.x; j < hi; j += 256) {
    const T2 z = u[j];
    s += (double)z.x * (double)z.x + (double)z.y * (double)z.y;
  }
  const double tot = block_sum256(s, red);
  if (threadIdx.x == 0) acc[k] = 0.0 + 1.0 * tot;
}

template <typename S>
hipError_t launch_detect_noalloc(const void *data, int64_t N, double dx, int64_t npx, double px, void *scr,
                                 hipStream_t st) {
  if (npx <= 0) return hipSuccess;
  if (npx > kNoAllocMaxPixels || npx > N) return hipErrorInvalidValue;
  double *acc = (double *)scr;
  int64_t *bounds = (int64_t *)((char *)scr + kNoAllocMaxPixels * sizeof(double));
  const DetGeom g = det_geom(N, dx, npx, px);
  hipLaunchKernelGGL(det_bounds_kernel, dim3((unsigned)((npx + 1 + 255) / 256)), dim3(256), 0, st, g, npx, bounds);
  hipLaunchKernelGGL(detect_one_kernel<S>, dim3((unsigned)npx), dim3(256), 0, st, data, acc,
                     (const int64_t *)bounds);
  return hipGetLastError();
}

__global__ void fill_f64_kernel(double *p, int64_t n, double v) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) p[i] = v;
}

template <typename S>
__global__ void copy_kernel(typename C2T<S>::t *__restrict__ dst, const typename C2T<S>::t *__restrict__ src,
                            int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

namespace {
inline unsigned grid_stride_blocks(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return (unsigned)b;
}
}  // namespace

template <typename S>
hipError_t launch_fiber_pass(int lgL, bool row, int xv, const FieldArgs *args, int nf, const Geom &g,
                             hipStream_t st, hipEvent_t t0, hipEvent_t t1) {
  switch (xv) {
    case XV_NONE: return launch_fiber_xv<S, XV_NONE>(lgL, row, args, nf, g, st, t0, t1);
    case XV_MASK: return launch_fiber_xv<S, XV_MASK>(lgL, row, args, nf, g, st, t0, t1);
    case XV_GEN: return launch_fiber_xv<S, XV_GEN>(lgL, row, args, nf, g, st, t0, t1);
    default: return hipErrorInvalidValue;
  }
}

template <typename S>
hipError_t launch_mask_fill(int lgL, const XOp &x, const Geom &g, void *bits, hipStream_t st) {
  return launch_mask_fill_t<S>(lgL, x, g, bits, st);
}

template <typename S>
hipError_t launch_xop(const FieldArgs *args, int nf, int64_t N, hipStream_t st) {
  hipLaunchKernelGGL(xop_kernel<S>, dim3(grid_stride_blocks(N), nf), dim3(256), 0, st, args, N);
  return hipGetLastError();
}

template <typename S>
hipError_t launch_kop(const FieldArgs *args, int nf, const Geom &g, hipStream_t st) {
  hipLaunchKernelGGL(kop_kernel<S>, dim3(grid_stride_blocks(g.N), nf), dim3(256), 0, st, args, g);
  return hipGetLastError();
}

template <typename S>
hipError_t launch_detect(const DetArgs *args, int nf, const int64_t *bounds, int64_t npx, hipStream_t st) {
  if (npx <= 0) return hipSuccess;
  hipLaunchKernelGGL(detect_kernel<S>, dim3((unsigned)npx), dim3(256), 0, st, args, nf, bounds);
  return hipGetLastError();
}

hipError_t launch_fill_f64(double *p, int64_t n, double v, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_f64_kernel, dim3(grid_stride_blocks(n)), dim3(256), 0, st, p, n, v);
  return hipGetLastError();
}

template <typename S>
hipError_t launch_copy(void *dst, const void *src, int64_t n, hipStream_t st) {
  using T2 = typename C2T<S>::t;
  hipLaunchKernelGGL(copy_kernel<S>, dim3(grid_stride_blocks(n)), dim3(256), 0, st, (T2 *)dst,
                     (const T2 *)src, n);
  return hipGetLastError();
}

template hipError_t launch_fiber_pass<float>(int, bool, int, const FieldArgs *, int, const Geom &, hipStream_t,
                                             hipEvent_t, hipEvent_t);
template hipError_t launch_fiber_pass<double>(int, bool, int, const FieldArgs *, int, const Geom &, hipStream_t,
                                              hipEvent_t, hipEvent_t);
template hipError_t launch_mask_fill<float>(int, const XOp &, const Geom &, void *, hipStream_t);
template hipError_t launch_mask_fill<double>(int, const XOp &, const Geom &, void *, hipStream_t);
template hipError_t launch_xop<float>(const FieldArgs *, int, int64_t, hipStream_t);
template hipError_t launch_xop<double>(const FieldArgs *, int, int64_t, hipStream_t);
template hipError_t launch_kop<float>(const FieldArgs *, int, const Geom &, hipStream_t);
template hipError_t launch_kop<double>(const FieldArgs *, int, const Geom &, hipStream_t);
template hipError_t launch_detect<float>(const DetArgs *, int, const int64_t *, int64_t, hipStream_t);
template hipError_t launch_detect<double>(const DetArgs *, int, const int64_t *, int64_t, hipStream_t);
template hipError_t launch_copy<float>(void *, const void *, int64_t, hipStream_t);
template hipError_t launch_detect_noalloc<float>(const void *, int64_t, double, int64_t, double, void *,
                                                 hipStream_t);
template hipError_t launch_detect_noalloc<double>(const void *, int64_t, double, int64_t, double, void *,
                                                  hipStream_t);
template hipError_t launch_copy<double>(void *, const void *, int64_t, hipStream_t);

}  // namespace rs
