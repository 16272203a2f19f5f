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
#include <hip/hip_ext.h>

#include "device_ops.hpp"
#include "rs_kernels.hpp"

namespace rs {

// Element map of a fiber-pass tile: thread t of tile `tile` holds samples
// nb + lane + m * mst (m < 16) of fiber c (column c0 + c for column passes),
// fiber positions q + m * L/16.  nb and mst are uniform over the workgroup,
// lane * sizeof(complex) < 2^31 for N <= 2^30.
// lane_at(j): the lane offset of fiber position j of this thread's fiber.
template <int L, int C, bool ROW> struct TileMap {
  int c, q;
  int64_t nb, c0, mst;
  uint32_t lane;
  int64_t cst = 0;  // column stride (column passes)
  __device__ __forceinline__ uint32_t lane_at(int j) const {
    return ROW ? (uint32_t)(c * L + j) : (uint32_t)(c + (int64_t)j * cst);
  }
  __device__ __forceinline__ TileMap(const Geom &g, int64_t tile, int t) {
    constexpr int TPF = L / 16;
    c = ROW ? t / TPF : t % C;
    q = ROW ? t % TPF : t / C;
    if (ROW) {
      nb = tile * C * L;
      c0 = 0;
      mst = TPF;
      lane = (uint32_t)(c * L + q);
      return;
    }
    cst = g.S;
    // XCD-aware order: workgroups are dealt round-robin to the 8 XCDs; when a
    // row segment of C columns is shorter than a 128 B line, the `segs`
    // tiles that share one line are given ids on the same XCD so that its L2
    // merges their partial-line reads and writes
    // (N, S, C, L and the segment count are powers of two: shifts and masks
    // of workgroup-uniform values, no 64-bit division in the prologue)
    constexpr int lgC = __builtin_ctz(C), lgLC = __builtin_ctz(L * C);
    const uint64_t ntiles = (uint64_t)g.N >> lgLC;
    uint64_t bid = (uint64_t)tile;
    if (g.xcd_segs > 1) {
      const int lgseg = __builtin_ctz((uint32_t)g.xcd_segs), lgG = 3 + lgseg;
      if (bid < ((ntiles >> lgG) << lgG)) {
        const uint64_t r = bid & ((uint64_t(1) << lgG) - 1);
        bid = (bid - r) + ((r & 7) << lgseg) + (r >> 3);
      }
    }
    const int lgcpb = __builtin_ctzll((uint64_t)g.S) - lgC;  // column tiles per row of the top group
    const int64_t hi = (int64_t)(bid >> lgcpb);
    c0 = (int64_t)((bid & ((uint64_t(1) << lgcpb) - 1)) << lgC);
    nb = hi * (int64_t)L * g.S + c0;
    mst = (int64_t)TPF * g.S;
    lane = (uint32_t)(c + (int64_t)q * g.S);
  }
};

// Stride of the fiber slabs in LDS (complex entries): at least L + L/16 (one
// pad entry per 16).  LDS serves a 64-bit access 16 lanes per cycle (32
// banks of 4 bytes; 8 lanes for complex128): G = 16 (complex64) or 8
// (complex128) lanes share a cycle.  Lanes of a column tile take consecutive
// columns, so a group holds min(C, G) fibers at G / C consecutive positions
// each: with C >= G an odd stride spreads the fibers over the banks; with
// fewer fibers the stride is G / C mod G, which puts fiber c at c G / C and
// its G / C positions after it, all G distinct (MI355X, 2^24 wide top pass
// of 2048 x 8: 378 bank-conflict cycles per wave with the odd stride, as a
// model of these groups predicts).  Rows and single fibers keep the odd
// stride.
template <typename S> constexpr int fiber_stride(int L, int C, bool row) {
  const int base = L + L / 16;
  constexpr int G = sizeof(S) == 4 ? 16 : 8;
  if (row || C == 1 || C >= G) return base | 1;
  int f = base;
  while (f % G != G / C) ++f;
  return f;
}

// NT: non-temporal field loads (top passes over fields that stream from HBM)
// SD: a row pass that runs one direction only (the plain form's FFT passes)
template <typename S, int L, int C, bool ROW, int XV, bool NT = false, bool SD = false>
__global__ void __launch_bounds__(C *(L / 16))
    fiber_pass(const FieldArgs *__restrict__ args, Geom g) {
  using T2 = typename C2T<S>::t;
  constexpr int TPF = L / 16;
  constexpr int FIB = fiber_stride<S>(L, C, ROW);
  __shared__ T2 lds[C * FIB];
  // complex64 row passes: the spectrum phase table, gathered at random
  // indices, is served from LDS
  constexpr bool KTAB = ROW && sizeof(S) == 4;
  __shared__ float2 ktab[KTAB ? kCisTab : 1];

  // column passes whose fibers span several threads share each column's
  // twiddle powers through LDS (build_col_powers) when the extra 16 C values
  // leave as many workgroups per CU as without them
  constexpr int kLdsPerCu = 160 * 1024;
  constexpr int kBase = (C * FIB + small_tab_entries(L)) * (int)sizeof(T2) + (KTAB ? kCisTab * 8 : 0);
  constexpr int kWith = kBase + 16 * C * (int)sizeof(T2);
  constexpr bool CPOW = !ROW && TPF >= 4 && kLdsPerCu / kWith == kLdsPerCu / kBase;
  __shared__ T2 cpow[CPOW ? 16 * C : 1];

  // every thread of a fiber in one wavefront: rows of up to 1024 (a row's
  // TPF threads are consecutive), column tiles of one wavefront
  // (column passes keep their fibers across wavefronts: lanes of a wave take
  // consecutive columns, for whole row segments per load; wave-local columns,
  // 64 / TPF per wave, made the 2^26 top pass slower, 216 -> 347 us at 64)
  constexpr bool WAVE = ROW ? TPF <= 64 : C * TPF <= 64;
  // fibers over several wavefronts and at least 256 long take the two-level
  // DFT (fiber_dft_2l): one workgroup barrier per direction, the rest of the
  // exchanges wave-synchronous.  Its spectral side (the frequency-ordered
  // samples: the outputs of the forward DFT, the inputs of the inverse) puts
  // fiber position j = fiber_spectral(q) + m L/16 in v[m]; the natural side
  // keeps j = q + m L/16.  The memory layout is the same either way.
  // Row passes of one direction keep the Stockham form: their spectral side
  // is in memory, where the two-level order would spread a wavefront's
  // accesses 16 samples apart (MI355X, complex128 4096-long rows at 2^24:
  // 82 against 97 us per pass).
  constexpr bool TL = !WAVE && TPF >= 16 && !SD;
  // each sub-fiber's L/256 threads (times C columns) lie in one wavefront
  static_assert(!TL || (ROW ? TPF / 16 : (TPF / 16) * C) <= 64, "two-level groups span wavefronts");
  // field accesses from scalar per-m bases (sbase): rows, passes with an
  // x-space operation, and column tiles of whole 128 B lines.  The wide
  // tiles of partial lines without one stay on vector addresses, where the
  // scalar bases measured slower (MI355X, 2^24: complex128 4096 x 2 column
  // pass 152.7 -> 173.3 us, complex64 4096 x 4 inverse 109.4 -> 117.2 us;
  // with the grating mask, 4096 x 4 106.3 -> 103.3 us, 2^26 top pass 256 x 32
  // 213.1 -> 203.5 us).
  constexpr bool SB = ROW || XV != XV_NONE || C * (int)sizeof(T2) >= 128;

  const FieldArgs &A = args[blockIdx.y];
  const T2 *__restrict__ twf = (const T2 *)g.tw_fib;
  const int t = threadIdx.x;
  constexpr int NSTAB = small_tab_entries(L);
  __shared__ T2 stab[NSTAB > 0 ? NSTAB : 1];
  const int64_t tile = g.tile0 + blockIdx.x;
  const TileMap<L, C, ROW> tm(g, tile, t);
  const int c = tm.c, q = tm.q;
  const int64_t mst = tm.mst;
  const uint32_t lane = tm.lane;
  const uint32_t lb = lane * (uint32_t)sizeof(T2);
  // position base of the spectral side (twiddles, spectrum bins) and its
  // lane offset in bytes
  const int qs = TL ? fiber_spectral<L>(q) : q;
  const uint32_t lbs = TL ? tm.lane_at(qs) * (uint32_t)sizeof(T2) : lb;
  // loads read the spectral side when the pass starts with an inverse
  // (column passes) or has no forward (row passes); stores likewise
  const bool spec_in = ROW ? !A.do_fwd : A.do_inv;
  const bool spec_out = ROW ? !A.do_inv : A.do_fwd;
  const uint32_t lb_in = spec_in ? lbs : lb, lb_out = spec_out ? lbs : lb;
  T2 *fib = lds + c * FIB;

  const gptr<nvec2<S>> data = global_ptr<S>(A.data);
  const bool xop = XV != XV_NONE && g.top && A.x.kind != XOP_NONE;
  bool gen = false;
  if constexpr (XV == XV_MASK) gen = xop && A.x.from_ones;
  if constexpr (XV == XV_GEN) gen = xop && (A.x.kind == XOP_POINT || A.x.kind == XOP_ONES || A.x.from_ones);

  const int64_t nb = tm.nb, c0 = tm.c0;
  // LDS tables: the twiddles of the small Stockham stages (first read after
  // the first exchange of a Stockham fiber over several wavefronts, so a
  // workgroup barrier of their own is needed only where the exchanges are
  // wave-synchronous or the fiber is two-level, whose cross twiddles and
  // sub-fiber stages can read them before its one barrier), the spectrum
  // phase table and the column powers.
  // complex64 passes issue the field loads first and then every table read
  // before the first LDS write; the table reads wait for the field loads
  // with their own, so such a pass pays one memory latency before its first
  // DFT instead of one per table (MI355X, 2^26: row 4096 104 -> 100 us,
  // top 64 228 -> 216 us per launch).  Column passes of 128 keep the tables
  // first (field loads early measured 217 -> 244 us there), and so do the
  // complex128 passes, which have no registers to spare for table entries
  // next to 64 in flight (their staging arrays went to scratch memory).
  constexpr bool EARLY = sizeof(S) == 4 && (ROW || L != 128);
  const bool kt = KTAB && A.k.on;
  if constexpr (!EARLY) {
    if constexpr (NSTAB > 0)
      for (int i = t; i < NSTAB; i += C * TPF) stab[i] = twf[L + i];
    if (kt) {
      const float2 *gt = (const float2 *)g.cis_tab;
      for (int i = t; i < kCisTab; i += C * TPF) ktab[i] = gt[i];
    }
    if constexpr (CPOW) build_col_powers<false, T2, L, C>(g, c0, cpow, t);
    if (CPOW || kt || ((WAVE || TL) && NSTAB > 0)) __syncthreads();
  }
  // cached grating decisions of this thread, loaded with the field (their
  // first use comes after the inverse DFT's barriers)
  uint32_t mword = 0;
  if constexpr (XV == XV_MASK && !ROW)
    if (xop && A.x.kind == XOP_MASK && A.x.mbits) mword = ((gptr<const uint16_t>)A.x.mbits)[tile * (C * TPF) + t];
  T2 v[16];
  if (!gen) {
    const gptr<nvec2<S>> p0 = data + nb;
    // top passes over fields that stream from HBM read them with the
    // non-temporal hint: each sample is read once and overwritten in place
    // (MI355X, 2^26 complex64 top pass: 218 -> 212 us per launch; on fields
    // that the last-level cache holds the hint costs, 181 -> 231 us).  A
    // compile-time choice: a run-time branch between two loads of the same
    // address lets the compiler merge them and drop the hint.
    if constexpr (NT) {
#pragma unroll
      for (int m = 0; m < 16; ++m) v[m] = gload_nt<S>(field_base<SB, S>(p0 + m * mst), lb_in);
    } else {
#pragma unroll
      for (int m = 0; m < 16; ++m) v[m] = gload<S>(field_base<SB, S>(p0 + m * mst), lb_in);
    }
  } else {
    // a generated plane wave: ones.  The empty volatile asm keeps these
    // moves inside the branch (hoisted, every thread of a reading pass ran
    // 32 of them before its loads overwrote them).
    S one = 1, zero = 0;
    asm volatile("" : "+v"(one), "+v"(zero));
#pragma unroll
    for (int m = 0; m < 16; ++m) { v[m].x = one; v[m].y = zero; }
  }
  T2 wq;
  if constexpr (CPOW) wq = col_base_twiddle<T2, L>(g, qs, c0 + c);
  if constexpr (EARLY) {
    LdsStage<T2, NSTAB, C * TPF> st;
    st.load(twf + L, t);
    LdsStage<float2, KTAB ? kCisTab : 0, C * TPF> sk;
    if (kt) sk.load((const float2 *)g.cis_tab, t);
    if constexpr (CPOW) build_col_powers<true, T2, L, C>(g, c0, cpow, t);
    st.store(stab, t);
    if (kt) sk.store(ktab, t);
    if (CPOW || kt || ((WAVE || TL) && NSTAB > 0)) __syncthreads();
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
        if constexpr (CPOW) col_twiddles<T2, C>(wq, c, cpow, tw);
        else inter_twiddles<T2, L>(g, qs, col, tw);
        have_tw = true;
#pragma unroll
        for (int m = 0; m < 16; ++m) v[m] = cmulc(v[m], tw[m]);
      } else {
        if constexpr (CPOW) apply_col_twiddles<true, T2, C>(wq, c, cpow, v);
        else apply_inter_twiddles<true, T2, L>(g, qs, col, v);
      }
      if constexpr (TL) fiber_dft_2l<T2, L, true>(v, fib, q, twf, stab);
      else stockham<T2, L, 1, true, WAVE>(v, fib, q, twf, stab);
    }

    if constexpr (XV == XV_MASK) {
      if (xop) {
        const XOp &x = A.x;
        T2 fa, fb;
        fa.x = (S)x.fa[0]; fa.y = (S)x.fa[1];
        fb.x = (S)x.fb[0]; fb.y = (S)x.fb[1];
        // a generated plane wave holds ones here, so the products are the
        // factors themselves
        if (x.kind == XOP_MASK && !ROW && x.mbits) {
          // decisions cached for this tile shape (mask_fill_kernel)
#pragma unroll
          for (int m = 0; m < 16; ++m) v[m] = cmul(v[m], ((mword >> m) & 1u) ? fa : fb);
        } else if (x.kind == XOP_MASK) {
          bool in[16];
          mask_tile(x, (int32_t)(nb + lane), (int32_t)mst, in);
#pragma unroll
          for (int m = 0; m < 16; ++m) v[m] = cmul(v[m], in[m] ? fa : fb);
        } else {
#pragma unroll
          for (int m = 0; m < 16; ++m) v[m] = cmul(v[m], fa);
        }
      }
    } else if constexpr (XV == XV_GEN) {
      if (xop) {
#pragma unroll
        for (int m = 0; m < 16; ++m) v[m] = apply_xop(A.x, nb + lane + m * mst, g.N, v[m]);
      }
    }

    if (A.do_fwd) {
      if constexpr (TL) fiber_dft_2l<T2, L, false>(v, fib, q, twf, stab);
      else stockham<T2, L, 1, false, WAVE>(v, fib, q, twf, stab);
      if (!ROW) {
        if (have_tw) {
#pragma unroll
          for (int m = 0; m < 16; ++m) v[m] = cmul(v[m], tw[m]);
        } else if constexpr (CPOW) {
          apply_col_twiddles<false, T2, C>(wq, c, cpow, v);
        } else {
          apply_inter_twiddles<false, T2, L>(g, qs, col, v);
        }
      }
    }

    if (ROW && A.k.on) {
      const int64_t kb = row_kbase(g, rr);
      const int64_t kstep = g.N >> g.lgR;
      if constexpr (sizeof(S) == 4) {
        const KopF32 kp = kop_f32_setup(A.k, ktab);
        const int k0 = (int)(kb + kstep * qs);  // < N/16
        const int kst = (int)(kstep * TPF);    // = N/16
        const KopQuad kq = kop_f32_quad(kp, k0, g.N);
        if (kp.rlen != 0) {
#pragma unroll
          for (int m = 0; m < 16; ++m) v[m] = kop_f32<true>(kp, kop_f32_u(kq, m), k0 + m * kst, v[m]);
        } else {
#pragma unroll
          for (int m = 0; m < 16; ++m) v[m] = kop_f32<false>(kp, kop_f32_u(kq, m), k0 + m * kst, v[m]);
        }
      } else {
#pragma unroll
        for (int m = 0; m < 16; ++m) {
          const int64_t k = kb + kstep * (qs + m * TPF);
          v[m] = apply_kop(A.k, k, g.N, v[m]);
        }
      }
    }

    if (ROW && A.do_inv) {
      if constexpr (TL) fiber_dft_2l<T2, L, true>(v, fib, q, twf, stab);
      else stockham<T2, L, 1, true, WAVE>(v, fib, q, twf, stab);
    }

    {
      const gptr<nvec2<S>> p0 = data + nb;
#pragma unroll
      for (int m = 0; m < 16; ++m) gstore<S>(field_base<SB, S>(p0 + m * mst), lb_out, v[m]);
    }

  }
}

// Grating decisions of a top column pass of fiber length L, in the order the
// pass visits its samples (TileMap): bit m of word [tile * threads + t].
template <int L, int C>
__global__ void __launch_bounds__(C *(L / 16)) mask_fill_kernel(XOp x, Geom g, uint16_t *__restrict__ bits) {
  const TileMap<L, C, false> tm(g, blockIdx.x, threadIdx.x);
  bool in[16];
  mask_tile(x, (int32_t)(tm.nb + tm.lane), (int32_t)tm.mst, in);
  uint32_t w = 0;
#pragma unroll
  for (int m = 0; m < 16; ++m) w |= (in[m] ? 1u : 0u) << m;
  bits[(int64_t)blockIdx.x * (C * (L / 16)) + threadIdx.x] = (uint16_t)w;
}

template <typename S, int L, int TILE>
static hipError_t launch_mask_fill_c(const XOp &x, const Geom &g, void *bits, hipStream_t st) {
  constexpr int C = col_fibers(L, sizeof(S) == 8, TILE);
  if (g.S % C != 0) return hipErrorInvalidValue;
  const int64_t nblk = g.N / ((int64_t)L * C);
  hipLaunchKernelGGL((mask_fill_kernel<L, C>), dim3((unsigned)nblk), dim3(C * (L / 16)), 0, st, x, g,
                     (uint16_t *)bits);
  return hipGetLastError();
}

template <typename S, int L>
static hipError_t launch_mask_fill_l(const XOp &x, const Geom &g, void *bits, hipStream_t st) {
  return g.tile ? launch_mask_fill_c<S, L, 1>(x, g, bits, st) : launch_mask_fill_c<S, L, 0>(x, g, bits, st);
}

template <typename S>
hipError_t launch_mask_fill_t(int lgL, const XOp &x, const Geom &g, void *bits, hipStream_t st) {
  switch (lgL) {
    case 4: return launch_mask_fill_l<S, 16>(x, g, bits, st);
    case 5: return launch_mask_fill_l<S, 32>(x, g, bits, st);
    case 6: return launch_mask_fill_l<S, 64>(x, g, bits, st);
    case 7: return launch_mask_fill_l<S, 128>(x, g, bits, st);
    case 8: return launch_mask_fill_l<S, 256>(x, g, bits, st);
    case 9: return launch_mask_fill_l<S, 512>(x, g, bits, st);
    case 10: return launch_mask_fill_l<S, 1024>(x, g, bits, st);
    case 11: return launch_mask_fill_l<S, 2048>(x, g, bits, st);
    case 12: return launch_mask_fill_l<S, 4096>(x, g, bits, st);
    case 13: return launch_mask_fill_l<S, 8192>(x, g, bits, st);
    default: return hipErrorInvalidValue;
  }
}

// tiles of `per` samples each covering the launch's span: first tile id into
// gs.tile0, count into nblk; false when the span does not split into whole
// tiles (or into whole XCD groups of tiles, when the pass regroups them)
static inline bool span_tiles(const Geom &g, int64_t per, Geom &gs, int64_t &nblk) {
  gs = g;
  const int64_t off = g.span_len ? g.span_off : 0;
  const int64_t len = g.span_len ? g.span_len : g.N;
  if (off < 0 || len <= 0 || off + len > g.N || off % per || len % per) return false;
  gs.tile0 = off / per;
  nblk = len / per;
  if (g.span_len && g.xcd_segs > 1) {
    const int64_t G = 8 * (int64_t)g.xcd_segs;
    if (gs.tile0 % G || nblk % G) return false;
  }
  return true;
}

// columns per workgroup for a column pass of fiber length L: col_fibers of
// the plan's tile variant (g.tile)
template <typename S, int L, int XV, int TILE>
static hipError_t launch_col_c(const FieldArgs *args, int nf, const Geom &g, hipStream_t st, hipEvent_t t0,
                               hipEvent_t t1) {
  constexpr int C = col_fibers(L, sizeof(S) == 8, TILE);
  if (g.S % C != 0) return hipErrorInvalidValue;
  Geom gs;
  int64_t nblk;
  if (!span_tiles(g, (int64_t)L * C, gs, nblk)) return hipErrorInvalidValue;
  dim3 grid((unsigned)nblk, (unsigned)nf);
  // non-temporal variant: top passes (an x-space operation inside) of up to
  // 512, the lengths a plan's top group takes on fields beyond the cache
  if constexpr (XV != XV_NONE && L <= 512) {
    if (g.nt) {
      hipExtLaunchKernelGGL((fiber_pass<S, L, C, false, XV, true>), grid, dim3(C * (L / 16)), 0, st, t0, t1, 0,
                            args, gs);
      return hipGetLastError();
    }
  }
  hipExtLaunchKernelGGL((fiber_pass<S, L, C, false, XV>), grid, dim3(C * (L / 16)), 0, st, t0, t1, 0, args, gs);
  return hipGetLastError();
}

template <typename S, int L, int XV>
static hipError_t launch_col(const FieldArgs *args, int nf, const Geom &g, hipStream_t st, hipEvent_t t0,
                             hipEvent_t t1) {
  return g.tile ? launch_col_c<S, L, XV, 1>(args, nf, g, st, t0, t1) : launch_col_c<S, L, XV, 0>(args, nf, g, st, t0, t1);
}

// rows per workgroup: full waves and 256 threads (complex64) / 128 threads
// (complex128) for rows up to 4096 / 2048 samples, one row beyond; fields
// with fewer rows than that (single-group fields) use one row per workgroup
template <typename S, int L, int XV, bool SD>
static hipError_t launch_row_sd(const FieldArgs *args, int nf, const Geom &g, hipStream_t st, hipEvent_t t0,
                                hipEvent_t t1) {
  constexpr int SPAN = sizeof(S) == 8 ? 2048 : 4096;
  constexpr int C = L >= SPAN ? 1 : SPAN / L;
  const int64_t rows = g.N / L;
  Geom gs;
  int64_t nblk;
  if (rows % C == 0 && span_tiles(g, (int64_t)L * C, gs, nblk)) {
    dim3 grid((unsigned)nblk, (unsigned)nf);
    hipExtLaunchKernelGGL((fiber_pass<S, L, C, true, XV, false, SD>), grid, dim3(C * (L / 16)), 0, st, t0, t1, 0,
                          args, gs);
  } else {
    if (!span_tiles(g, (int64_t)L, gs, nblk)) return hipErrorInvalidValue;
    dim3 grid((unsigned)nblk, (unsigned)nf);
    hipExtLaunchKernelGGL((fiber_pass<S, L, 1, true, XV, false, SD>), grid, dim3(L / 16), 0, st, t0, t1, 0, args,
                          gs);
  }
  return hipGetLastError();
}

// rows of one direction (Geom::one_dir) that the two-level form would take
// run the Stockham variant
template <typename S, int L, int XV>
static hipError_t launch_row(const FieldArgs *args, int nf, const Geom &g, hipStream_t st, hipEvent_t t0,
                             hipEvent_t t1) {
  if constexpr (L >= 2048 && XV == XV_NONE)
    if (g.one_dir) return launch_row_sd<S, L, XV, true>(args, nf, g, st, t0, t1);
  return launch_row_sd<S, L, XV, false>(args, nf, g, st, t0, t1);
}

template <typename S, int XV>
hipError_t launch_fiber_xv(int lgL, bool row, const FieldArgs *args, int nf, const Geom &g, hipStream_t st,
                           hipEvent_t t0, hipEvent_t t1) {
  if (row) {
    switch (lgL) {
      case 6: return launch_row<S, 64, XV>(args, nf, g, st, t0, t1);
      case 7: return launch_row<S, 128, XV>(args, nf, g, st, t0, t1);
      case 8: return launch_row<S, 256, XV>(args, nf, g, st, t0, t1);
      case 9: return launch_row<S, 512, XV>(args, nf, g, st, t0, t1);
      case 10: return launch_row<S, 1024, XV>(args, nf, g, st, t0, t1);
      case 11: return launch_row<S, 2048, XV>(args, nf, g, st, t0, t1);
      case 12: return launch_row<S, 4096, XV>(args, nf, g, st, t0, t1);
      case 13: if constexpr (sizeof(S) == 4) return launch_row<S, 8192, XV>(args, nf, g, st, t0, t1);
               else return hipErrorInvalidValue;
      default: return hipErrorInvalidValue;
    }
  }
  switch (lgL) {
    case 4: return launch_col<S, 16, XV>(args, nf, g, st, t0, t1);
    case 5: return launch_col<S, 32, XV>(args, nf, g, st, t0, t1);
    case 6: return launch_col<S, 64, XV>(args, nf, g, st, t0, t1);
    case 7: return launch_col<S, 128, XV>(args, nf, g, st, t0, t1);
    case 8: return launch_col<S, 256, XV>(args, nf, g, st, t0, t1);
    case 9: return launch_col<S, 512, XV>(args, nf, g, st, t0, t1);
    case 10: return launch_col<S, 1024, XV>(args, nf, g, st, t0, t1);
    case 11: return launch_col<S, 2048, XV>(args, nf, g, st, t0, t1);
    case 12: return launch_col<S, 4096, XV>(args, nf, g, st, t0, t1);
    case 13: if constexpr (sizeof(S) == 4) return launch_col<S, 8192, XV>(args, nf, g, st, t0, t1);
             else return hipErrorInvalidValue;
    default: return hipErrorInvalidValue;
  }
}

}  // namespace rs
