// Device building blocks of the gfx950 kernels: complex arithmetic, in-register
// DFTs, radix-16 Stockham stages through LDS, twiddle generation, and the
// elementwise operations of the split step (x-space grating / source, spectrum
// propagator).
//
// Reference counterparts: fast-wave/fwcuda/kernels.cu (propagate_analytically
// :6-28, propagate_convolve_step :40-65, apply_grating_factors :67-89) and the
// cuFFT calls of fast-wave/src/fft.cpp:28-44; numerics follow big-wave
// (propagation.py, optical_element.py).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "rs_device.hpp"
#include "rs_kernels.hpp"

namespace rs {

template <typename S> struct C2T;
template <> struct C2T<float> { using t = float2; };
template <> struct C2T<double> { using t = double2; };

template <typename T2> __device__ __forceinline__ T2 cmul(T2 a, T2 b) {
  T2 r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}
template <typename T2> __device__ __forceinline__ T2 cmulc(T2 a, T2 b) {  // a * conj(b)
  T2 r;
  r.x = a.x * b.x + a.y * b.y;
  r.y = a.y * b.x - a.x * b.y;
  return r;
}
template <typename T2> __device__ __forceinline__ T2 cadd(T2 a, T2 b) {
  T2 r; r.x = a.x + b.x; r.y = a.y + b.y; return r;
}
template <typename T2> __device__ __forceinline__ T2 csub(T2 a, T2 b) {
  T2 r; r.x = a.x - b.x; r.y = a.y - b.y; return r;
}
// multiply by -i (forward) or +i (inverse)
template <bool INV, typename T2> __device__ __forceinline__ T2 mul_mi(T2 a) {
  T2 r;
  if (INV) { r.x = -a.y; r.y = a.x; } else { r.x = a.y; r.y = -a.x; }
  return r;
}

// Field memory as global-address-space pointers (global_load / global_store
// instead of flat), addressed as a workgroup-uniform pointer plus a 32-bit
// per-lane byte offset.
template <typename T> using gptr = __attribute__((address_space(1))) T *;
template <typename S> using nvec2 = S __attribute__((ext_vector_type(2)));
template <typename S> __device__ __forceinline__ gptr<nvec2<S>> global_ptr(void *p) {
  return (gptr<nvec2<S>>)p;
}
// A workgroup-uniform base held in scalar registers as an opaque value, so
// that an access stays (scalar base + 32-bit lane offset) instead of being
// reassociated into a per-lane 64-bit address (a 64-bit multiply-add and an
// add per access on the vector unit).
template <typename S> __device__ __forceinline__ gptr<nvec2<S>> sbase(gptr<nvec2<S>> p) {
  uint64_t a = (uint64_t)p;
  asm("" : "+s"(a));
  return (gptr<nvec2<S>>)a;
}
template <bool SB, typename S> __device__ __forceinline__ gptr<nvec2<S>> field_base(gptr<nvec2<S>> p) {
  if constexpr (SB) return sbase<S>(p);
  else return p;
}
template <typename S>
__device__ __forceinline__ typename C2T<S>::t gload(gptr<nvec2<S>> p, uint32_t off_bytes) {
  const nvec2<S> r = *(gptr<nvec2<S>>)((gptr<char>)p + off_bytes);
  typename C2T<S>::t o;
  o.x = r.x;
  o.y = r.y;
  return o;
}
template <typename S>
__device__ __forceinline__ typename C2T<S>::t gload_nt(gptr<nvec2<S>> p, uint32_t off_bytes) {
  const nvec2<S> r = __builtin_nontemporal_load((gptr<nvec2<S>>)((gptr<char>)p + off_bytes));
  typename C2T<S>::t o;
  o.x = r.x;
  o.y = r.y;
  return o;
}
template <typename S>
__device__ __forceinline__ void gstore(gptr<nvec2<S>> p, uint32_t off_bytes, typename C2T<S>::t v) {
  nvec2<S> r;
  r.x = v.x;
  r.y = v.y;
  *(gptr<nvec2<S>>)((gptr<char>)p + off_bytes) = r;
}

// ---------------------------------------------------------------------------
// in-register DFTs, natural order in and out, W = exp(-2 pi i / R) (forward)
// ---------------------------------------------------------------------------

template <bool INV, typename T2> __device__ __forceinline__ void dft2(T2 &a, T2 &b) {
  T2 t = a; a = cadd(t, b); b = csub(t, b);
}

template <bool INV, typename T2> __device__ __forceinline__ void dft4(T2 &x0, T2 &x1, T2 &x2, T2 &x3) {
  T2 t0 = cadd(x0, x2), t1 = csub(x0, x2);
  T2 t2 = cadd(x1, x3), t3 = mul_mi<INV>(csub(x1, x3));
  x0 = cadd(t0, t2); x2 = csub(t0, t2);
  x1 = cadd(t1, t3); x3 = csub(t1, t3);
}

#define RS_SQRT1_2 0.70710678118654752440
#define RS_COS_PI8 0.92387953251128675613
#define RS_SIN_PI8 0.38268343236508977173

// rotate by W8^j for j=1,2,3 (forward sign; conjugate for inverse)
template <bool INV, typename T2> __device__ __forceinline__ T2 w8(T2 a, int j) {
  using S = decltype(a.x);
  const S h = (S)RS_SQRT1_2;
  T2 r;
  if (j == 1) {  // (1 - i)/sqrt2
    if (!INV) { r.x = (a.x + a.y) * h; r.y = (a.y - a.x) * h; }
    else { r.x = (a.x - a.y) * h; r.y = (a.y + a.x) * h; }
  } else if (j == 2) {
    r = mul_mi<INV>(a);
  } else {  // W8^3 = (-1 - i)/sqrt2
    if (!INV) { r.x = (a.y - a.x) * h; r.y = -(a.x + a.y) * h; }
    else { r.x = -(a.x + a.y) * h; r.y = (a.x - a.y) * h; }
  }
  return r;
}

template <bool INV, typename T2> __device__ __forceinline__ void dft8(T2 *x) {
  T2 a0 = x[0], a1 = x[2], a2 = x[4], a3 = x[6];
  T2 b0 = x[1], b1 = x[3], b2 = x[5], b3 = x[7];
  dft4<INV>(a0, a1, a2, a3);
  dft4<INV>(b0, b1, b2, b3);
  b1 = w8<INV>(b1, 1); b2 = w8<INV>(b2, 2); b3 = w8<INV>(b3, 3);
  x[0] = cadd(a0, b0); x[4] = csub(a0, b0);
  x[1] = cadd(a1, b1); x[5] = csub(a1, b1);
  x[2] = cadd(a2, b2); x[6] = csub(a2, b2);
  x[3] = cadd(a3, b3); x[7] = csub(a3, b3);
}

// W16^j for j in 0..9 as (cos, sin) of -2 pi j/16
template <bool INV, typename T2> __device__ __forceinline__ T2 w16(T2 a, int j) {
  using S = decltype(a.x);
  const S c1 = (S)RS_COS_PI8, s1 = (S)RS_SIN_PI8, h = (S)RS_SQRT1_2;
  S c, s;  // forward: multiply by c - i s
  switch (j) {
    case 0: return a;
    case 1: c = c1; s = s1; break;
    case 2: c = h; s = h; break;
    case 3: c = s1; s = c1; break;
    case 4: return mul_mi<INV>(a);
    case 6: c = -h; s = h; break;
    case 9: c = -c1; s = -s1; break;
    default: c = 1; s = 0; break;
  }
  if (INV) s = -s;
  T2 r;
  r.x = a.x * c + a.y * s;
  r.y = a.y * c - a.x * s;
  return r;
}

template <bool INV, typename T2> __device__ __forceinline__ void dft16(T2 *x) {
  // X[q + 4s] = sum_p W4^{ps} W16^{pq} sum_m x[p + 4m] W4^{mq}
  T2 g[4][4];
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    g[p][0] = x[p]; g[p][1] = x[p + 4]; g[p][2] = x[p + 8]; g[p][3] = x[p + 12];
    dft4<INV>(g[p][0], g[p][1], g[p][2], g[p][3]);
  }
#pragma unroll
  for (int p = 1; p < 4; ++p)
#pragma unroll
    for (int q = 1; q < 4; ++q) g[p][q] = w16<INV>(g[p][q], p * q);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    T2 h0 = g[0][q], h1 = g[1][q], h2 = g[2][q], h3 = g[3][q];
    dft4<INV>(h0, h1, h2, h3);
    x[q] = h0; x[q + 4] = h1; x[q + 8] = h2; x[q + 12] = h3;
  }
}

// fused multiply-add in the field precision
__device__ __forceinline__ float fmaS(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmaS(double a, double b, double c) { return __builtin_fma(a, b, c); }

// a + x * w (x * conj(w) for the inverse) as four fused multiply-adds
template <bool INV, typename T2> __device__ __forceinline__ T2 add_tw(T2 a, T2 x, T2 w) {
  T2 r;
  if (!INV) {
    r.x = fmaS(x.x, w.x, fmaS(-x.y, w.y, a.x));
    r.y = fmaS(x.x, w.y, fmaS(x.y, w.x, a.y));
  } else {
    r.x = fmaS(x.x, w.x, fmaS(x.y, w.y, a.x));
    r.y = fmaS(x.y, w.x, fmaS(-x.x, w.y, a.y));
  }
  return r;
}
// 2 a - t (t = a + y: a - y from the sum, one rounding)
template <typename T2> __device__ __forceinline__ T2 twice_minus(T2 a, T2 t) {
  using S = decltype(a.x);
  T2 r;
  r.x = fmaS((S)2, a.x, -t.x);
  r.y = fmaS((S)2, a.y, -t.y);
  return r;
}

// dft16 of x[r] * w[r] (w[0] = 1): the stage twiddles of x[p + 8] and
// x[p + 12] are fused into the first radix-4 layer, a + c w = four fused
// multiply-adds and a - c w = 2 a - (a + c w) two more, against a complex
// product and two sums (16 VALU instructions fewer per stage)
template <bool INV, typename T2> __device__ __forceinline__ void dft16_tw(T2 *x, const T2 *w) {
  T2 g[4][4];
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    const T2 a = p ? (INV ? cmulc(x[p], w[p]) : cmul(x[p], w[p])) : x[p];
    const T2 b = INV ? cmulc(x[p + 4], w[p + 4]) : cmul(x[p + 4], w[p + 4]);
    const T2 t0 = add_tw<INV>(a, x[p + 8], w[p + 8]);
    const T2 t1 = twice_minus(a, t0);
    const T2 t2 = add_tw<INV>(b, x[p + 12], w[p + 12]);
    const T2 t3 = mul_mi<INV>(twice_minus(b, t2));
    g[p][0] = cadd(t0, t2); g[p][2] = csub(t0, t2);
    g[p][1] = cadd(t1, t3); g[p][3] = csub(t1, t3);
  }
#pragma unroll
  for (int p = 1; p < 4; ++p)
#pragma unroll
    for (int q = 1; q < 4; ++q) g[p][q] = w16<INV>(g[p][q], p * q);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    T2 h0 = g[0][q], h1 = g[1][q], h2 = g[2][q], h3 = g[3][q];
    dft4<INV>(h0, h1, h2, h3);
    x[q] = h0; x[q + 4] = h1; x[q + 8] = h2; x[q + 12] = h3;
  }
}

template <int R, bool INV, typename T2> __device__ __forceinline__ void dftR(T2 *x) {
  if constexpr (R == 2) dft2<INV>(x[0], x[1]);
  else if constexpr (R == 4) dft4<INV>(x[0], x[1], x[2], x[3]);
  else if constexpr (R == 8) dft8<INV>(x);
  else dft16<INV>(x);
}

// ---------------------------------------------------------------------------
// Stockham stages over one fiber of length L held by L/16 threads.
// Thread q holds positions q + m*(L/16), m < 16, in v[m] before the first and
// after the last stage; intermediate stages go through the fiber's LDS slab.
// ---------------------------------------------------------------------------

// LDS position of fiber element i: one pad word per 16 (bank spread)
__device__ __forceinline__ int lds_pad(int i) { return i + (i >> 4); }
// lds_pad(i + d) - lds_pad(i) for the strides used below: d = r * D with D a
// power of two and i < D or i a multiple of 16 * ceil(D / 16) plus less than
// D, so the pad of the sum splits and the offset is a compile-time constant
// (ds_read / ds_write immediate offsets)
template <int D> constexpr int lds_step(int r) { return r * D + ((r * D) >> 4); }

// w[r] = W^(r e) for r < R from the table W[j] = exp(-2 pi i j / len),
// index mask len-1: reads e, 2e, 4e, 8e and forms the other powers as one
// product of two read values (error <= ~2 ulp, no long recurrences).
template <typename T2, int R>
__device__ __forceinline__ void twiddle_powers(const T2 *__restrict__ tab, uint32_t e, uint32_t mask, T2 *w) {
  w[0].x = 1;
  w[0].y = 0;
  if constexpr (R >= 2) w[1] = tab[e & mask];
  if constexpr (R >= 4) {
    w[2] = tab[(2 * e) & mask];
    w[3] = cmul(w[2], w[1]);
  }
  if constexpr (R >= 8) {
    w[4] = tab[(4 * e) & mask];
    w[5] = cmul(w[4], w[1]);
    w[6] = cmul(w[4], w[2]);
    w[7] = cmul(w[4], w[3]);
  }
  if constexpr (R >= 16) {
    w[8] = tab[(8 * e) & mask];
#pragma unroll
    for (int r = 9; r < 16; ++r) w[r] = cmul(w[8], w[r - 8]);
  }
}

// split-table twiddle W_M^e = hi[e >> sh] * lo[e & (2^sh - 1)], M = 2^lgM
template <typename T2>
__device__ __forceinline__ T2 split_twiddle(const T2 *__restrict__ hi, const T2 *__restrict__ lo, int sh,
                                            uint32_t e) {
  return cmul(hi[e >> sh], lo[e & ((1u << sh) - 1)]);
}

// Exchange ordering between the Stockham stages of a fiber.  WAVE: every
// thread of the fiber sits in one wavefront (rows of up to 1024, column tiles
// of one wavefront), whose LDS operations execute in program order, so the
// exchange needs no workgroup barrier, only a fence that keeps the compiler
// from moving LDS accesses across it.
template <bool WAVE> __device__ __forceinline__ void fiber_sync() {
  if constexpr (WAVE) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// TS: the large stages read W_L^e as entry e * TS of a table of L * TS
// entries (a sub-fiber of a longer fiber shares that fiber's table)
template <typename T2, int L, int R, int NS, bool FIRST, bool LAST, bool INV, bool WAVE, int TS = 1>
__device__ __forceinline__ void stockham_stage(T2 (&v)[16], T2 *fib, int q,
                                               const T2 *__restrict__ twf, const T2 *stab) {
  constexpr int TPF = L / 16;
  constexpr int NB = 16 / R;
  constexpr int LR = L / R;
  T2 x[NB][R];
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const int b = q + i * TPF;
    if constexpr (FIRST) {
#pragma unroll
      for (int r = 0; r < R; ++r) x[i][r] = v[i + r * NB];
      dftR<R, INV>(x[i]);
    } else {
      // R = 16 here (the remainder radix only starts a transform): b < L/R
      static_assert(R == 16, "later stages are radix 16");
      const T2 *src = fib + lds_pad(b);
#pragma unroll
      for (int r = 0; r < R; ++r) x[i][r] = src[lds_step<LR>(r)];
      // W_{NS R}^{r k} for r < R: small stages read all of them from their
      // [k][r] table in LDS; larger ones read k, 2k, 4k, 8k from W_L^j and
      // form the other powers as products of at most two of them (reading
      // all fifteen from a [k][r] table in global memory instead was slower:
      // 116 against 102 us per 2048-long row pass at 2^26, the loads' latency
      // exposed)
      T2 w[R];
      if constexpr (NS <= kSmallNS) {
        const T2 *row = stab + small_tab_offset(L, NS) + (b % NS) * kTwRow;
#pragma unroll
        for (int r = 1; r < R; ++r) w[r] = row[r];
      } else {
        const uint32_t e = (uint32_t)(b % NS) * (uint32_t)(L / (NS * R) * TS);
        twiddle_powers<T2, R>(twf, e, (uint32_t)(L * TS) - 1, w);
      }
      w[0].x = 1;
      w[0].y = 0;
      dft16_tw<INV>(x[i], w);
    }
  }
  if constexpr (LAST) {
#pragma unroll
    for (int i = 0; i < NB; ++i)
#pragma unroll
      for (int r = 0; r < R; ++r) v[i + r * NB] = x[i][r];
  } else {
    fiber_sync<WAVE>();
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      const int b = q + i * TPF;
      const int k = b % NS;
      const int o = (b / NS) * NS * R + k;
      T2 *dst = fib + lds_pad(o);
#pragma unroll
      for (int r = 0; r < R; ++r) dst[lds_step<NS>(r)] = x[i][r];
    }
    fiber_sync<WAVE>();
  }
}

// radix of the stage that starts at sub-length NS: the remainder radix
// (L = R0 * 16^k) goes first, where it needs no twiddles
template <int L, int NS> constexpr int stage_radix() {
  if (NS != 1) return 16;
  int r = 1, l = L;
  while (l > 16) l /= 16;
  r = l;  // L = r * 16^k with r in {2, 4, 8, 16}
  return r;
}

template <typename T2, int L, int NS, bool INV, bool WAVE = false, int TS = 1>
__device__ __forceinline__ void stockham(T2 (&v)[16], T2 *fib, int q, const T2 *__restrict__ twf,
                                         const T2 *stab) {
  constexpr int R = stage_radix<L, NS>();
  constexpr bool FIRST = NS == 1;
  constexpr bool LAST = NS * R == L;
  stockham_stage<T2, L, R, NS, FIRST, LAST, INV, WAVE, TS>(v, fib, q, twf, stab);
  if constexpr (!LAST) stockham<T2, L, NS * R, INV, WAVE, TS>(v, fib, q, twf, stab);
}

// Two-level DFT of a fiber of length L = 16 T (T = L / 16 threads, T >= 16)
// whose threads span several wavefronts.  With n = n2 + T n1 and
// k = k1 + 16 k2: a 16-point DFT over each thread's registers (n1), the
// twiddles W_L^{n2 k1}, one exchange across the fiber (the only workgroup
// barrier of a direction), then for each k1 a T-point Stockham DFT over n2,
// held by the G = T / 16 threads q = k1 G .. k1 G + G - 1, which lie in one
// wavefront (wave-synchronous exchanges in sub-slab k1 of the fiber's slab).
// Forward: thread q holds x[q + T m] (natural side) before and
// X[j + T m] after, with j = fiber_spectral(q) = q / G + 16 (q % G) (spectral
// side); the inverse maps the spectral side back to the natural side.
// Each thread's cross-fiber reads and writes touch the same 16 LDS words,
// so an exchange needs no barrier before its writes.
// The twiddles of the cross step are W_{16 T}^{q k1}: the [k][r] table of the
// stage at sub-length T when T <= kSmallNS (LDS), else four reads of W_L.
template <int L> __device__ __forceinline__ int fiber_spectral(int q) {
  constexpr int G = L / 256;
  return q / G + 16 * (q % G);
}

template <typename T2, int L>
__device__ __forceinline__ void cross_twiddles(int q, const T2 *__restrict__ twf, const T2 *stab, T2 (&w)[16]) {
  constexpr int T = L / 16;
  if constexpr (T <= kSmallNS) {
    const T2 *row = stab + small_tab_offset(L, T) + q * kTwRow;
#pragma unroll
    for (int r = 1; r < 16; ++r) w[r] = row[r];
  } else {
    twiddle_powers<T2, 16>(twf, (uint32_t)q, (uint32_t)L - 1, w);
  }
  w[0].x = 1;
  w[0].y = 0;
}

template <typename T2, int L, bool INV>
__device__ __forceinline__ void fiber_dft_2l(T2 (&v)[16], T2 *fib, int q, const T2 *__restrict__ twf,
                                             const T2 *stab) {
  constexpr int T = L / 16, G = T / 16, SUB = T + T / 16;
  static_assert(G >= 1, "two-level fibers hold at least 256 samples");
  const int g = q / G, l = q % G;
  T2 *sub = fib + g * SUB;
  if constexpr (!INV) {
    dft16<false>(v);
    T2 w[16];
    cross_twiddles<T2, L>(q, twf, stab, w);
#pragma unroll
    for (int k = 1; k < 16; ++k) v[k] = cmul(v[k], w[k]);
    // the words written here are the ones this thread read last (inverse
    // half of a column pass) or untouched (first use)
    fiber_sync<true>();
#pragma unroll
    for (int k = 0; k < 16; ++k) fib[k * SUB + lds_pad(q)] = v[k];
    __syncthreads();
#pragma unroll
    for (int m = 0; m < 16; ++m) v[m] = sub[lds_pad(l + G * m)];
    stockham<T2, T, 1, false, true, 16>(v, sub, l, twf, stab);
  } else {
    stockham<T2, T, 1, true, true, 16>(v, sub, l, twf, stab);
    // sub-slab g is this wavefront's own until the barrier
    fiber_sync<true>();
#pragma unroll
    for (int m = 0; m < 16; ++m) sub[lds_pad(l + G * m)] = v[m];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = fib[k * SUB + lds_pad(q)];
    T2 w[16];
    cross_twiddles<T2, L>(q, twf, stab, w);
    dft16_tw<true>(v, w);
  }
}

// ---------------------------------------------------------------------------
// x-space operations
// ---------------------------------------------------------------------------

// Grating coordinate of sample n as big-wave forms it: np.linspace per chunk
// (big-wave/optical_element.py:352-355), without contraction.
__device__ __forceinline__ double grating_x(const XOp &op, int64_t n, int64_t N) {
#pragma clang fp contract(off)
  if (op.single_chunk) return (double)n * op.xstep + op.x0;
  const int64_t c = n / op.chunk;
  const int64_t k = n - c * op.chunk;
  int64_t len = N - c * op.chunk;
  if (len > op.chunk) len = op.chunk;
  const int64_t start = c * op.chunk - N / 2;
  const double x_start = (double)start * op.dx + op.xpos;
  const double x_stop = x_start + (double)len * op.dx;
  const double step = (x_stop - x_start) / (double)len;
  return (double)k * step + x_start;
}

// np.mod(x / pitch + off, 1) as big-wave rounds it.  x * (1/pitch) differs
// from the correctly rounded quotient by a few ulp of max(|t|, 1); outside a
// guard band around the mask edges (ph = 0, ph = dc and ph = 1) that cannot
// change the decision ph < dc, inside it the exact quotient is used.
__device__ __forceinline__ double grating_phase(double x, double pitch, double inv_pitch, double off,
                                                double dc) {
#pragma clang fp contract(off)
  double t = x * inv_pitch + off;
  double ph = t - floor(t);
  const double guard = (fabs(t) + 1.0) * 1e-13;
  if (ph < guard || ph > 1.0 - guard || fabs(ph - dc) < guard) {
    t = x / pitch + off;
    ph = t - floor(t);
  }
  return ph;
}

// Any x-space op on natural index n.  The grating coordinate and mask follow
// big-wave bit for bit: x = linspace(x0, x0 + len*dx, len, endpoint=False)
// per chunk, mask = mod(x/pitch + dc/2, 1) < dc, all in FP64 without
// contraction (big-wave/optical_element.py:341-357).
template <typename T2>
__device__ __forceinline__ T2 apply_xop(const XOp &op, int64_t n, int64_t N, T2 v) {
#pragma clang fp contract(off)
  using S = decltype(v.x);
  double re = 0.0, im = 0.0;
  bool mul = true;
  switch (op.kind) {
    case XOP_MASK: {
      const double x = grating_x(op, n, N);
      if (grating_phase(x, op.pitch, op.inv_pitch, op.half_dc, op.dc) < op.dc) {
        re = op.fa[0]; im = op.fa[1];
      } else {
        re = op.fb[0]; im = op.fb[1];
      }
      break;
    }
    case XOP_DBA: {
      // (mod(x/pitch0 + dc0/2, 1) < dc0) * (mod(x/pitch1, 1) < dc1)
      // (big-wave/optical_element.py:359-375)
      const double x = grating_x(op, n, N);
      if (grating_phase(x, op.pitch, op.inv_pitch, op.half_dc, op.dc) < op.dc &&
          grating_phase(x, op.pitch1, op.inv_pitch1, 0.0, op.dc1) < op.dc1) {
        re = op.fa[0]; im = op.fa[1];
      } else {
        re = op.fb[0]; im = op.fb[1];
      }
      break;
    }
    case XOP_SCALAR:
      re = op.fa[0]; im = op.fa[1];
      break;
    case XOP_POINT: {
      const double xd = (double)n - (double)N / 2.0;
      const double x = xd * op.dx - op.xs;
      const double r = sqrt(x * x + op.z * op.z);
      const double a = r * op.kphase;
      double s, c;
      sincos(a, &s, &c);
      const double sr = sqrt(r);
      re = c / sr; im = s / sr;
      mul = false;
      break;
    }
    case XOP_ONES:
      re = 1.0; im = 0.0;
      mul = false;
      break;
    case XOP_SAMPLE: {
      // delta + i beta by np.interp(x, rx, db, left=0, right=0) over the
      // complex values, then exp(2 pi i dz / wl * db) (big-wave/optical_element.py
      // :449-470); intervals between equal pixels and exact knots use the
      // host-computed pixel factor
      const double x = ((double)n - (double)N / 2.0) * op.dx;
      const double *rx = op.s_rx;
      const int64_t len = op.s_len;
      re = 1.0;
      im = 0.0;
      if (x < rx[0] || x > rx[len - 1]) break;  // db = 0
      if (x == rx[len - 1]) {
        re = op.s_f[2 * (len - 1)];
        im = op.s_f[2 * (len - 1) + 1];
        break;
      }
      int64_t j = (int64_t)floor((x - rx[0]) * op.s_inv_px);
      if (j < 0) j = 0;
      if (j > len - 2) j = len - 2;
      while (j > 0 && rx[j] > x) --j;
      while (j < len - 2 && rx[j + 1] <= x) ++j;
      const double *db = op.s_db;
      if (rx[j] == x || (db[2 * j] == db[2 * j + 2] && db[2 * j + 1] == db[2 * j + 3])) {
        re = op.s_f[2 * j];
        im = op.s_f[2 * j + 1];
        break;
      }
      // numpy's complex interp: slope = difference * (1 / dx), no fused ops
      const double inv = 1.0 / (rx[j + 1] - rx[j]);
      const double dr = ((db[2 * j + 2] - db[2 * j]) * inv) * (x - rx[j]) + db[2 * j];
      const double di = ((db[2 * j + 3] - db[2 * j + 1]) * inv) * (x - rx[j]) + db[2 * j + 1];
      const double er = -(op.s_a * di) + 0.0 * dr;
      const double ei = op.s_a * dr + 0.0 * di;
      const double e = exp(er);
      double sn, cs;
      sincos(ei, &sn, &cs);
      re = e * cs;
      im = e * sn;
      break;
    }
    default:
      return v;
  }
  // fold the post factor in double, then cast (factors are cast to the field
  // precision as big-wave's per-chunk factor array is, optical_element.py:151-169)
  const double pr = re * op.post[0] - im * op.post[1];
  const double pi = re * op.post[1] + im * op.post[0];
  T2 f;
  f.x = (S)pr;
  f.y = (S)pi;
  return mul ? cmul(v, f) : f;
}

// Grating decisions of one thread's 16 samples n0 + m*ms (single chunk,
// factors folded on the host); in[m] selects fa.  t_m = fma(m, ms*m_a, t_0)
// with t_0 = fma(n0, m_a, m_b) is within m_guard of big-wave's x/pitch + dc/2
// (two roundings of at most an ulp of the span each, far inside the guard);
// a fraction within m_guard of an edge (0, dc or 1) is decided on the exact
// FP64 path, sample by sample (ties recur every few thousand samples when
// pitch/dx is rational).
__device__ __forceinline__ void mask_tile(const XOp &op, int32_t n0, int32_t ms, bool (&in)[16]) {
#pragma clang fp contract(off)
  const double t0 = __builtin_fma((double)n0, op.m_a, op.m_b);
  const double d = (double)ms * op.m_a;
  const double omg = 1.0 - op.m_guard;
#pragma unroll
  for (int m = 0; m < 16; ++m) {
    const double t = __builtin_fma((double)m, d, t0);
    double ph = t - floor(t);
    if (ph < op.m_guard || ph > omg || fabs(ph - op.dc) < op.m_guard) {
      const double x = (double)(n0 + m * ms) * op.xstep + op.x0;
      const double te = x / op.pitch + op.half_dc;
      ph = te - floor(te);
    }
    in[m] = ph < op.dc;
  }
}

// ---------------------------------------------------------------------------
// spectrum operations
// ---------------------------------------------------------------------------

// (cos, sin)(2 pi t) for t in [-1/2, 1/2] given in FP64 (complex128 fields).
__device__ __forceinline__ void cis_rev(double t, double *c_out, double *s_out) {
  sincospi(2.0 * t, s_out, c_out);
}

// spectrum multiply at FFT bin k: phase in revolutions computed in FP64 and
// reduced to [-1/2, 1/2] before the sine and cosine (big-wave/propagation.py:124-137).
// Complex128 fields and the plain-form spectrum kernel of both precisions.
template <typename T2>
__device__ __forceinline__ T2 apply_kop(const KOp &op, int64_t k, int64_t N, T2 v) {
  using S = decltype(v.x);
  if (k >= op.rs && k < op.re) {
    T2 z; z.x = 0; z.y = 0;
    return z;
  }
  const int64_t ks = (k < N / 2) ? k : k - N;
  const double kd = (double)ks;
  double t = op.alpha_rev * (kd * kd) + op.beta_rev;
  t -= rint(t);
  double c, s;
  cis_rev(t, &c, &s);
  T2 h;
  h.x = (S)(c * op.scale[0] - s * op.scale[1]);
  h.y = (S)(c * op.scale[1] + s * op.scale[0]);
  return cmul(v, h);
}

// Complex64 spectrum factor exp(2 pi i t) * sc of the row pass.  u = 256 t
// is formed in FP64 (k'^2 exact up to 2^53, one rounding in the fma), split
// as u = j + f with j integer and |f| <= 1/2; exp(2 pi i j / 256) comes from
// a 256-entry table held in LDS and exp(2 pi i f / 256) from its Taylor
// terms to third (sine) and second (cosine) order, truncation < 2.4e-12 and
// 9.5e-10 for |2 pi f / 256| <= 1.23e-2, both below complex64 rounding.
// Table entries and products are complex64.

struct KopF32 {
  const float2 *tab;     // kCisTab entries exp(2 pi i j / kCisTab) (LDS)
  double au, bu;         // alpha, beta in 1/kCisTab revolutions
  float s1, s3, c2, sc;  // sc*w, -sc*w^3/6, -sc*w^2/2, sc (w = phi per unit of f; KOp::t_*)
  int rs, rlen;          // cutoff band [rs, rs + rlen)
};

__device__ __forceinline__ KopF32 kop_f32_setup(const KOp &op, const float2 *tab) {
  KopF32 p;
  p.tab = tab;
  p.au = op.alpha_rev * (double)kCisTab;
  p.bu = op.beta_rev * (double)kCisTab;
  p.sc = op.t_sc;
  p.s1 = op.t_s1;
  p.s3 = op.t_s3;
  p.c2 = op.t_c2;
  p.rs = (int)op.rs;
  p.rlen = (int)(op.re - op.rs);
  return p;
}

// Row-pass spectrum factors of one thread: its 16 bins are k = k0 + m N/16
// with 0 <= k0 < N/16, so the signed bins are k0 + m' N/16 (m' = m for m < 8,
// m - 16 above) and u = U0 + U1 m' + U2 m'^2 with U0 = 256 (alpha k0^2 + beta),
// U1 = 512 alpha k0 N/16, U2 = 256 alpha (N/16)^2, evaluated by Horner's rule
// in FP64 (a few ulp of the largest term, far below complex64 rounding).
struct KopQuad {
  double u0, u1, u2;
};
__device__ __forceinline__ KopQuad kop_f32_quad(const KopF32 &p, int k0, int64_t N) {
  const double kd = (double)k0;
  const double D = (double)(N >> 4);
  KopQuad q;
  q.u0 = __builtin_fma(p.au, kd * kd, p.bu);
  q.u1 = ((2.0 * p.au) * kd) * D;
  q.u2 = (p.au * D) * D;
  return q;
}
__device__ __forceinline__ double kop_f32_u(const KopQuad &q, int m) {
  const double mm = (double)(m < 8 ? m : m - 16);
  return __builtin_fma(mm, __builtin_fma(mm, q.u2, q.u1), q.u0);
}

template <bool CUT> __device__ __forceinline__ float2 kop_f32(const KopF32 &p, double u, int k, float2 v) {
  const double j = rint(u);
  const float f = (float)(u - j);
  // j mod kCisTab from the low mantissa bits of j + 1.5 * 2^52 (|j| < 2^51)
  const int idx = __double2loint(j + 6755399441055744.0) & (kCisTab - 1);
  const float2 T = p.tab[idx];
  // |w f| <= pi / 256: the cosine's next term, (w f)^4 / 24 < 1e-9, is below
  // complex64 rounding; the sine's cubic term (< 3.1e-7) is not
  const float f2 = f * f;
  const float cs = fmaf(f2, p.c2, p.sc);
  const float sn = f * fmaf(f2, p.s3, p.s1);
  float2 h;
  h.x = T.x * cs - T.y * sn;
  h.y = T.x * sn + T.y * cs;
  if (CUT && (uint32_t)(k - p.rs) < (uint32_t)p.rlen) { h.x = 0.f; h.y = 0.f; }
  return cmul(v, h);
}

// frequency bin of row-pass position j in row rr (digit-reversed groups)
__device__ __forceinline__ int64_t row_kbase(const Geom &g, int64_t rr) {
  if (g.ngroups == 3) {
    const int64_t ka = rr >> g.lgB;
    const int64_t kb = rr & ((int64_t(1) << g.lgB) - 1);
    return ka + (kb << g.lgA);
  }
  return rr;
}

// Inter-group twiddles of one thread's 16 elements: tw[m] = W_M^{(q + m*TPF) col},
// M = L * S, as W_M^{q col} * (W_M^{TPF col})^m with the powers built from
// five split-table reads.  Every exponent is taken mod M (<= 2^30, a divisor
// of 2^32), so 32-bit products that wrap give the same index.
template <typename T2, int L>
__device__ __forceinline__ void inter_twiddles(const Geom &g, int q, int64_t col, T2 (&tw)[16]) {
  const T2 *hi = (const T2 *)g.tw_hi;
  const T2 *lo = (const T2 *)g.tw_lo;
  const uint32_t mask = (uint32_t)((int64_t)L * g.S - 1);
  constexpr uint32_t TPF = L / 16;
  const uint32_t cl = (uint32_t)col;
  const uint32_t eb = (TPF * cl) & mask;
  const T2 wq = split_twiddle(hi, lo, g.tw_sh, ((uint32_t)q * cl) & mask);
  T2 b1 = split_twiddle(hi, lo, g.tw_sh, eb);
  T2 b2 = split_twiddle(hi, lo, g.tw_sh, (2 * eb) & mask);
  T2 b4 = split_twiddle(hi, lo, g.tw_sh, (4 * eb) & mask);
  T2 b8 = split_twiddle(hi, lo, g.tw_sh, (8 * eb) & mask);
  T2 p[8];
  p[0].x = 1; p[0].y = 0;
  p[1] = b1; p[2] = b2; p[3] = cmul(b2, b1);
  p[4] = b4; p[5] = cmul(b4, b1); p[6] = cmul(b4, b2); p[7] = cmul(b4, p[3]);
  const T2 q8 = cmul(wq, b8);
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    tw[m] = cmul(wq, p[m]);
    tw[m + 8] = cmul(q8, p[m]);
  }
}

// v[m] *= tw[m] (conj(tw[m]) when CONJ) with the twiddles of inter_twiddles
// formed on the fly: only the eight powers p[m] stay live (registers)
template <bool CONJ, typename T2, int L>
__device__ __forceinline__ void apply_inter_twiddles(const Geom &g, int q, int64_t col, T2 (&v)[16]) {
  const T2 *hi = (const T2 *)g.tw_hi;
  const T2 *lo = (const T2 *)g.tw_lo;
  const uint32_t mask = (uint32_t)((int64_t)L * g.S - 1);
  constexpr uint32_t TPF = L / 16;
  const uint32_t cl = (uint32_t)col;
  const uint32_t eb = (TPF * cl) & mask;
  const T2 wq = split_twiddle(hi, lo, g.tw_sh, ((uint32_t)q * cl) & mask);
  T2 b1 = split_twiddle(hi, lo, g.tw_sh, eb);
  T2 b2 = split_twiddle(hi, lo, g.tw_sh, (2 * eb) & mask);
  T2 b4 = split_twiddle(hi, lo, g.tw_sh, (4 * eb) & mask);
  T2 b8 = split_twiddle(hi, lo, g.tw_sh, (8 * eb) & mask);
  T2 p[8];
  p[0].x = 1; p[0].y = 0;
  p[1] = b1; p[2] = b2; p[3] = cmul(b2, b1);
  p[4] = b4; p[5] = cmul(b4, b1); p[6] = cmul(b4, b2); p[7] = cmul(b4, p[3]);
  const T2 q8 = cmul(wq, b8);
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const T2 t0 = cmul(wq, p[m]);
    const T2 t1 = cmul(q8, p[m]);
    v[m] = CONJ ? cmulc(v[m], t0) : cmul(v[m], t0);
    v[m + 8] = CONJ ? cmulc(v[m + 8], t1) : cmul(v[m + 8], t1);
  }
}

// A table of N entries copied from global memory into LDS by the NT threads
// of a workgroup in two halves, load() into registers and store() from them,
// so that the reads of several tables are in flight together
template <typename T, int N, int NT>
struct LdsStage {
  static constexpr int K = N > 0 ? (N + NT - 1) / NT : 0;
  T r[K > 0 ? K : 1];
  __device__ __forceinline__ void load(const T *__restrict__ src, int t) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int i = t + k * NT;
      if (N % NT == 0 || i < N) r[k] = src[i];
    }
  }
  __device__ __forceinline__ void store(T *__restrict__ dst, int t) const {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int i = t + k * NT;
      if (N % NT == 0 || i < N) dst[i] = r[k];
    }
  }
};

// Column powers of a column-pass tile, shared through LDS: the TPF threads of
// column c need W_M^{m TPF col} (col = c0 + c, m < 16) alike, so the
// workgroup forms each of the 16 C values once, P[m * C + c], one split-table
// read each.  A thread's twiddles are then tw[m] = W_M^{q col} * P[m][c]:
// one split-table read and 15 products (three roundings from the exact
// value, against four or five for the power recurrence above).
// BATCH: every table read of the thread first, then the products; else one
// entry at a time.
template <bool BATCH, typename T2, int L, int C>
__device__ __forceinline__ void build_col_powers(const Geom &g, int64_t c0, T2 *__restrict__ P, int t) {
  const T2 *hi = (const T2 *)g.tw_hi;
  const T2 *lo = (const T2 *)g.tw_lo;
  const uint32_t mask = (uint32_t)((int64_t)L * g.S - 1);
  constexpr int TPF = L / 16;
  constexpr int NT = C * TPF;
  constexpr int K = (16 * C + NT - 1) / NT;
  const uint32_t lm = (1u << g.tw_sh) - 1;
  if constexpr (!BATCH) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int i = k * NT + t;
      if (16 * C % NT == 0 || i < 16 * C) {
        const uint32_t m = (uint32_t)(i / C), cl = (uint32_t)c0 + (uint32_t)(i % C);
        P[i] = split_twiddle(hi, lo, g.tw_sh, (m * (uint32_t)TPF * cl) & mask);
      }
    }
    return;
  }
  T2 h[K], l[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int i = k * NT + t;
    if (16 * C % NT == 0 || i < 16 * C) {
      const uint32_t m = (uint32_t)(i / C), cl = (uint32_t)c0 + (uint32_t)(i % C);
      const uint32_t e = (m * (uint32_t)TPF * cl) & mask;
      h[k] = hi[e >> g.tw_sh];
      l[k] = lo[e & lm];
    }
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int i = k * NT + t;
    if (16 * C % NT == 0 || i < 16 * C) P[i] = cmul(h[k], l[k]);
  }
}

// W_M^{q col}: a thread's own factor, read once before the powers are formed
// so that its latency overlaps theirs
template <typename T2, int L>
__device__ __forceinline__ T2 col_base_twiddle(const Geom &g, int q, int64_t col) {
  const T2 *hi = (const T2 *)g.tw_hi;
  const T2 *lo = (const T2 *)g.tw_lo;
  const uint32_t mask = (uint32_t)((int64_t)L * g.S - 1);
  return split_twiddle(hi, lo, g.tw_sh, ((uint32_t)q * (uint32_t)col) & mask);
}

template <typename T2, int C>
__device__ __forceinline__ void col_twiddles(T2 wq, int c, const T2 *__restrict__ P, T2 (&tw)[16]) {
  tw[0] = wq;
#pragma unroll
  for (int m = 1; m < 16; ++m) tw[m] = cmul(wq, P[m * C + c]);
}

template <bool CONJ, typename T2, int C>
__device__ __forceinline__ void apply_col_twiddles(T2 wq, int c, const T2 *__restrict__ P, T2 (&v)[16]) {
  v[0] = CONJ ? cmulc(v[0], wq) : cmul(v[0], wq);
#pragma unroll
  for (int m = 1; m < 16; ++m) {
    const T2 t = cmul(wq, P[m * C + c]);
    v[m] = CONJ ? cmulc(v[m], t) : cmul(v[m], t);
  }
}

}  // namespace rs
