// Host-side launchers of the gfx950 kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "rs_device.hpp"

namespace rs {

// One digit-group pass over `nf` fields (blockIdx.y = field).  `row` selects
// the contiguous lowest group; otherwise a strided column group of stride g.S.
// xv (XopVariant) must cover the x-space ops of every field of the launch
// (xop_variant below).  Non-null t0 / t1 receive the kernel's own start and
// end times (recorded by the dispatch itself, no extra stream commands).
template <typename S>
hipError_t launch_fiber_pass(int lgL, bool row, int xv, const FieldArgs *args, int nf, const Geom &g,
                             hipStream_t st, hipEvent_t t0 = nullptr, hipEvent_t t1 = nullptr);
// the same for one compile-time variant (fiber_inst.hip)
template <typename S, int XV>
hipError_t launch_fiber_xv(int lgL, bool row, const FieldArgs *args, int nf, const Geom &g, hipStream_t st,
                           hipEvent_t t0, hipEvent_t t1);

// Smallest variant that runs the x-space ops of host records a[0..nf): none,
// grating mask / scalar (single chunk), or any kind.
inline int xop_variant(const FieldArgs *a, int nf, bool top) {
  if (!top) return XV_NONE;
  int v = XV_NONE;
  for (int i = 0; i < nf; ++i) {
    const XOp &x = a[i].x;
    if (x.kind == XOP_NONE) continue;
    const bool simple = x.kind == XOP_SCALAR || (x.kind == XOP_MASK && x.single_chunk);
    v = simple ? (v == XV_GEN ? XV_GEN : XV_MASK) : XV_GEN;
  }
  return v;
}
template <typename S> hipError_t launch_xop(const FieldArgs *args, int nf, int64_t N, hipStream_t st);
// grating decisions of a top column pass of fiber length 2^lgL for the mask x
// (single chunk, finalized), as bits in the pass's sample order: N/8 bytes
template <typename S>
hipError_t launch_mask_fill(int lgL, const XOp &x, const Geom &g, void *bits, hipStream_t st);
template <typename S> hipError_t launch_kop(const FieldArgs *args, int nf, const Geom &g, hipStream_t st);
template <typename S>
hipError_t launch_detect(const DetArgs *args, int nf, const int64_t *bounds, int64_t npx, hipStream_t st);
hipError_t launch_fill_f64(double *p, int64_t n, double v, hipStream_t st);
template <typename S> hipError_t launch_copy(void *dst, const void *src, int64_t n, hipStream_t st);

// Allocation-free detector read: pixel bounds and FP64 sums of |u|^2 of one
// natural-order field into the per-GPU scratch memory `scr` (the address
// noalloc_scratch returned): sums at scr[0, npx) as doubles, bounds after
// kNoAllocMaxPixels doubles.  npx <= kNoAllocMaxPixels and npx <= N.
constexpr int64_t kNoAllocMaxPixels = 65536;
constexpr size_t kNoAllocScratchBytes = (size_t)(2 * kNoAllocMaxPixels + 1) * 8;
constexpr int kMaxDevices = 64;
// device address of the scratch memory on `device`, which must be current;
// looked up once per device (hipGetSymbolAddress)
void *noalloc_scratch(int device);
template <typename S>
hipError_t launch_detect_noalloc(const void *data, int64_t N, double dx, int64_t npx, double px, void *scr,
                                 hipStream_t st);

// supported fiber lengths (log2) of the row and column passes
constexpr int kMinRowLg = 6;
constexpr int kMaxRowLgF32 = 13;
constexpr int kMaxRowLgF64 = 12;
constexpr int kMinColLg = 4;
constexpr int kMaxColLg = 9;        // default plans: columns of up to 512 with full 128 B segments
constexpr int kMaxWideColLgF32 = 13;  // wide column passes (C * L = 16384 complex64, 128 KB LDS)
constexpr int kMaxWideColLgF64 = 12;  // (C * L = 8192 complex128)

// Stockham stage sequence of a fiber of length L: the remainder radix
// R0 (L = R0 * 16^k) at sub-length NS = 1, then radix 16 at NS = R0, 16 R0, ...
constexpr int stage_r0(int L) {
  int l = L;
  while (l > 16) l /= 16;
  return l;
}
// Radix-16 stages at sub-length NS <= kSmallNS read their twiddles
// W_{16 NS}^{k r} from a table of NS rows of kTwRow entries ([k][r], one pad
// entry per row); the tables of all such stages follow the L entries W_L^j of
// the fiber's twiddle array, in stage order.
constexpr int kSmallNS = 32;
constexpr int kTwRow = 17;
constexpr int small_tab_offset(int L, int NS) {
  int off = 0;
  for (int ns = stage_r0(L); ns < NS && ns < L; ns *= 16)
    if (ns > 1 && ns <= kSmallNS) off += ns * kTwRow;
  return off;
}
constexpr int small_tab_entries(int L) { return small_tab_offset(L, L); }
constexpr int fiber_tab_entries(int L) { return L + small_tab_entries(L); }

// fibers per workgroup of a column pass of length L (tile 0, the default):
// row segments of 256 B (512 B for complex64 fibers up to 64) and at least
// one wavefront up to L = 256; 128 B at L = 512, where 256 B would leave one
// workgroup per CU; beyond that C * L fixed at 128 KB of LDS.  Tile 1 is the
// alternative the autotuner may pick per (N, precision, batch size): twice
// as many fibers at L = 128 and 512 (and complex128 at 64), half as many
// elsewhere (64 KB wide tiles, two workgroups per CU).
constexpr int col_fibers(int L, bool f64, int tile = 0) {
  int c = 0;
  if (L >= 1024) c = (f64 ? 8192 : 16384) / L;
  else if (L == 512) c = f64 ? 8 : 16;
  else if (L <= 16) c = 64;
  else if (f64) c = L <= 32 ? 32 : 16;
  else c = L <= 64 ? 64 : 32;
  if (tile == 0) return c;
  const bool wider = L == 128 || L == 512 || (f64 && L == 64);
  return wider ? 2 * c : (c > 1 ? c / 2 : c);
}
constexpr int kColTiles = 2;

}  // namespace rs
