// Fiber pass: one digit-group pass of the four-step FFT over C fibers of
// length L per workgroup tile, staged through LDS with radix-16 Stockham
// stages, with the elementwise work of the split step fused in (x-space
// multiply in the top column pass, spectrum multiply between the row DFT and
// row iDFT).
//
// Compile-time variants: S (complex64 / complex128 fields), L and C, ROW
// (contiguous lowest group or strided column group) and XV (x-space op: none,
// grating mask / scalar with host-folded factors, or any kind).
// The instantiations are spread over several translation units
// (fiber_inst.hip built once per precision and variant).
#pragma once
#include "device_ops.hpp"
#include "rs_kernels.hpp"

namespace rs {

template <typename S, int L, int C, bool ROW, int XV>
__global__ void __launch_bounds__(C *(L / 16))
    fiber_pass(const FieldArgs *__restrict__ args, Geom g) {
  using T2 = typename C2T<S>::t;
  constexpr int TPF = L / 16;
  // fiber slab stride: odd, so that lanes working on neighbouring fibers at
  // the same offset (column passes) hit different LDS banks
  constexpr int FIB = (L + L / 16) | 1;
  __shared__ T2 lds[C * FIB];
  // complex64 row passes: the spectrum phase table, gathered at random
  // indices, is served from LDS
  constexpr bool KTAB = ROW && sizeof(S) == 4;
  __shared__ float2 ktab[KTAB ? kCisTab : 1];

  const FieldArgs &A = args[blockIdx.y];
  const T2 *__restrict__ twf = (const T2 *)g.tw_fib;
  const int t = threadIdx.x;
  // twiddle tables of the small Stockham stages, staged in LDS (first read
  // after the first stage's barrier)
  constexpr int NSTAB = small_tab_entries(L);
  __shared__ T2 stab[NSTAB > 0 ? NSTAB : 1];
  if constexpr (NSTAB > 0)
    for (int i = t; i < NSTAB; i += C * TPF) stab[i] = twf[L + i];
  if constexpr (KTAB) {
    if (A.k.on) {
      const float2 *gt = (const float2 *)g.cis_tab;
      for (int i = t; i < kCisTab; i += C * TPF) ktab[i] = gt[i];
      __syncthreads();
    }
  }
  const int c = ROW ? t / TPF : t % C;
  const int q = ROW ? t % TPF : t / C;
  T2 *fib = lds + c * FIB;

  // Sample index of v[m] in the workgroup's tile: nb + lane + m * mst, with
  // nb and mst uniform over the workgroup and lane * sizeof(T2) < 2^31 for
  // N <= 2^30.
  const int64_t ntiles = g.N / ((int64_t)L * C);
  const int64_t mst = ROW ? (int64_t)TPF : (int64_t)TPF * g.S;
  const uint32_t lane = ROW ? (uint32_t)(c * L + q) : (uint32_t)(c + (int64_t)q * g.S);
  const uint32_t lb = lane * (uint32_t)sizeof(T2);
  const int64_t cpb = ROW ? 1 : g.S / C;
  // first sample (nb) and first column of a tile
  auto tile_base = [&](int64_t tile, int64_t &nb, int64_t &c0) {
    if (ROW) {
      nb = tile * C * L;
      c0 = 0;
      return;
    }
    // XCD-aware order: workgroups are dealt round-robin to the 8 XCDs; when a
    // row segment of C columns is shorter than a 128 B line, the `segs`
    // tiles that share one line are given ids on the same XCD so that its L2
    // merges their partial-line reads and writes
    int64_t bid = tile;
    if (g.xcd_segs > 1) {
      const int64_t G = 8 * (int64_t)g.xcd_segs;
      if (bid < (ntiles / G) * G) {
        const int64_t r = bid % G;
        bid = (bid - r) + (r % 8) * g.xcd_segs + r / 8;
      }
    }
    const int64_t hi = bid / cpb;
    c0 = (bid % cpb) * C;
    nb = hi * (int64_t)L * g.S + c0;
  };

  const gptr<nvec2<S>> data = global_ptr<S>(A.data);
  const bool xop = XV != XV_NONE && g.top && A.x.kind != XOP_NONE;
  bool gen = false;
  if constexpr (XV == XV_MASK) gen = xop && A.x.from_ones;
  if constexpr (XV == XV_GEN) gen = xop && (A.x.kind == XOP_POINT || A.x.kind == XOP_ONES || A.x.from_ones);

  const int64_t tile = blockIdx.x;
  int64_t nb, c0;
  tile_base(tile, nb, c0);
  T2 v[16];
  if (!gen) {
    const gptr<nvec2<S>> p0 = data + nb;
#pragma unroll
    for (int m = 0; m < 16; ++m) v[m] = gload<S>(p0 + m * mst, lb);
  } else {
#pragma unroll
    for (int m = 0; m < 16; ++m) { v[m].x = 1; v[m].y = 0; }
  }
  {
    const int64_t col = c0 + c;
    const int64_t rr = tile * C + c;

    // inter-group twiddles: top passes keep them live from the inverse to
    // the forward DFT; middle passes (one direction) apply them on the fly
    constexpr bool KEEP_TW = !ROW && XV != XV_NONE;
    T2 tw[16];
    bool have_tw = false;
    if (!ROW && A.do_inv) {
      if constexpr (KEEP_TW) {
        inter_twiddles<T2, L>(g, q, col, tw);
        have_tw = true;
#pragma unroll
        for (int m = 0; m < 16; ++m) v[m] = cmulc(v[m], tw[m]);
      } else {
        apply_inter_twiddles<true, T2, L>(g, q, col, v);
      }
      stockham<T2, L, 1, true>(v, fib, q, twf, stab);
    }

    if constexpr (XV == XV_MASK) {
      if (xop) {
        const XOp &x = A.x;
        T2 fa, fb;
        fa.x = (S)x.fa[0]; fa.y = (S)x.fa[1];
        fb.x = (S)x.fb[0]; fb.y = (S)x.fb[1];
        if (x.kind == XOP_MASK) {
          const double omg = 1.0 - x.m_guard;
          const int32_t n0 = (int32_t)(nb + lane);
          const int32_t ms = (int32_t)mst;
#pragma unroll
          for (int m = 0; m < 16; ++m) {
            const T2 f = mask_in(x, omg, n0 + m * ms) ? fa : fb;
            v[m] = gen ? f : cmul(v[m], f);
          }
        } else {
#pragma unroll
          for (int m = 0; m < 16; ++m) v[m] = gen ? fa : cmul(v[m], fa);
        }
      }
    } else if constexpr (XV == XV_GEN) {
      if (xop) {
#pragma unroll
        for (int m = 0; m < 16; ++m) v[m] = apply_xop(A.x, nb + lane + m * mst, g.N, v[m]);
      }
    }

    if (A.do_fwd) {
      stockham<T2, L, 1, false>(v, fib, q, twf, stab);
      if (!ROW) {
        if (have_tw) {
#pragma unroll
          for (int m = 0; m < 16; ++m) v[m] = cmul(v[m], tw[m]);
        } else {
          apply_inter_twiddles<false, T2, L>(g, q, col, v);
        }
      }
    }

    if (ROW && A.k.on) {
      const int64_t kb = row_kbase(g, rr);
      const int64_t kstep = g.N >> g.lgR;
      if constexpr (sizeof(S) == 4) {
        const KopF32 kp = kop_f32_setup(A.k, ktab, 63 - __clzll((unsigned long long)g.N));
        const int k0 = (int)(kb + kstep * q);
        const int kst = (int)(kstep * TPF);
        if (kp.rlen != 0) {
#pragma unroll
          for (int m = 0; m < 16; ++m) v[m] = kop_f32<true>(kp, k0 + m * kst, v[m]);
        } else {
#pragma unroll
          for (int m = 0; m < 16; ++m) v[m] = kop_f32<false>(kp, k0 + m * kst, v[m]);
        }
      } else {
#pragma unroll
        for (int m = 0; m < 16; ++m) {
          const int64_t k = kb + kstep * (q + m * TPF);
          v[m] = apply_kop(A.k, k, g.N, v[m]);
        }
      }
    }

    if (ROW && A.do_inv) stockham<T2, L, 1, true>(v, fib, q, twf, stab);

    {
      const gptr<nvec2<S>> p0 = data + nb;
#pragma unroll
      for (int m = 0; m < 16; ++m) gstore<S>(p0 + m * mst, lb, v[m]);
    }

  }
}

// columns per workgroup for a column pass of fiber length L: 16 columns of
// complex64 (128 B per row segment) / 8 of complex128, more for short fibers
template <typename S, int L, int XV>
static hipError_t launch_col(const FieldArgs *args, int nf, const Geom &g, hipStream_t st) {
  constexpr int C = col_fibers(L, sizeof(S) == 8);
  if (g.S % C != 0) return hipErrorInvalidValue;
  const int64_t nblk = g.N / ((int64_t)L * C);
  dim3 grid((unsigned)nblk, (unsigned)nf);
  hipLaunchKernelGGL((fiber_pass<S, L, C, false, XV>), grid, dim3(C * (L / 16)), 0, st, args, g);
  return hipGetLastError();
}

// rows per workgroup: full waves and 256 threads (complex64) / 128 threads
// (complex128) for rows up to 4096 / 2048 samples, one row beyond; fields
// with fewer rows than that (single-group fields) use one row per workgroup
template <typename S, int L, int XV>
static hipError_t launch_row(const FieldArgs *args, int nf, const Geom &g, hipStream_t st) {
  constexpr int SPAN = sizeof(S) == 8 ? 2048 : 4096;
  constexpr int C = L >= SPAN ? 1 : SPAN / L;
  const int64_t rows = g.N / L;
  if (rows % C == 0) {
    dim3 grid((unsigned)(rows / C), (unsigned)nf);
    hipLaunchKernelGGL((fiber_pass<S, L, C, true, XV>), grid, dim3(C * (L / 16)), 0, st, args, g);
  } else {
    dim3 grid((unsigned)rows, (unsigned)nf);
    hipLaunchKernelGGL((fiber_pass<S, L, 1, true, XV>), grid, dim3(L / 16), 0, st, args, g);
  }
  return hipGetLastError();
}

template <typename S, int XV>
hipError_t launch_fiber_xv(int lgL, bool row, const FieldArgs *args, int nf, const Geom &g, hipStream_t st) {
  if (row) {
    switch (lgL) {
      case 6: return launch_row<S, 64, XV>(args, nf, g, st);
      case 7: return launch_row<S, 128, XV>(args, nf, g, st);
      case 8: return launch_row<S, 256, XV>(args, nf, g, st);
      case 9: return launch_row<S, 512, XV>(args, nf, g, st);
      case 10: return launch_row<S, 1024, XV>(args, nf, g, st);
      case 11: return launch_row<S, 2048, XV>(args, nf, g, st);
      case 12: return launch_row<S, 4096, XV>(args, nf, g, st);
      case 13: if constexpr (sizeof(S) == 4) return launch_row<S, 8192, XV>(args, nf, g, st);
               else return hipErrorInvalidValue;
      default: return hipErrorInvalidValue;
    }
  }
  switch (lgL) {
    case 4: return launch_col<S, 16, XV>(args, nf, g, st);
    case 5: return launch_col<S, 32, XV>(args, nf, g, st);
    case 6: return launch_col<S, 64, XV>(args, nf, g, st);
    case 7: return launch_col<S, 128, XV>(args, nf, g, st);
    case 8: return launch_col<S, 256, XV>(args, nf, g, st);
    case 9: return launch_col<S, 512, XV>(args, nf, g, st);
    case 10: return launch_col<S, 1024, XV>(args, nf, g, st);
    case 11: return launch_col<S, 2048, XV>(args, nf, g, st);
    case 12: return launch_col<S, 4096, XV>(args, nf, g, st);
    case 13: if constexpr (sizeof(S) == 4) return launch_col<S, 8192, XV>(args, nf, g, st);
             else return hipErrorInvalidValue;
    default: return hipErrorInvalidValue;
  }
}

}  // namespace rs
