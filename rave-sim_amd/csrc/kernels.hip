// gfx950 kernels of the split-step path.
//
//  fiber_pass   -- one digit-group pass of the four-step FFT over C fibers of
//                  length L per workgroup, staged through LDS with radix-16
//                  Stockham stages, with the elementwise work of the split
//                  step fused in (x-space multiply in the top column pass,
//                  spectrum multiply between the row DFT and row iDFT).
//  xop_kernel   -- x-space multiply / generation as its own pass (plain form).
//  kop_kernel   -- spectrum multiply as its own pass (plain form).
//  detect_kernel-- per-pixel FP64 sums of |u|^2 over host-computed sample
//                  ranges (big-wave histogram binning), weighted accumulate.
//
// Reference counterparts: fast-wave/fwcuda/kernels.cu (propagate_analytically
// :6-28, propagate_convolve_step :40-65, apply_grating_factors :67-89,
// scale :160-172, square_and_downsample :174-208) and cuFFT calls in
// fast-wave/src/fft.cpp:28-44.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

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
template <typename T2> __device__ __forceinline__ T2 mk(double re, double im) {
  T2 r; r.x = re; r.y = im; return r;
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

__device__ __forceinline__ int lds_pad(int i) { return i + (i >> 4); }

// w[r] = W^(r e) for r < R from the table W[j] = exp(-2 pi i j / len),
// index mask len-1: reads e, 2e, 4e, 8e and forms the other powers as one
// product of two read values (error <= ~2 ulp, no long recurrences).
template <typename T2, int R>
__device__ __forceinline__ void twiddle_powers(const T2 *__restrict__ tab, int64_t e, int64_t mask, T2 *w) {
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
                                            int64_t e) {
  return cmul(hi[e >> sh], lo[e & ((int64_t(1) << sh) - 1)]);
}

template <typename T2, int L, int R, int NS, bool FIRST, bool LAST, bool INV>
__device__ __forceinline__ void stockham_stage(T2 (&v)[16], T2 *fib, int q,
                                               const T2 *__restrict__ twf) {
  constexpr int TPF = L / 16;
  constexpr int NB = 16 / R;
  T2 x[NB][R];
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const int b = q + i * TPF;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if constexpr (FIRST) x[i][r] = v[i + r * NB];
      else x[i][r] = fib[lds_pad(b + r * (L / R))];
    }
    if constexpr (!FIRST) {
      // W_{NS R}^{r k} for r < R from a few table reads (k, 2k, 4k, 8k);
      // the remaining powers are products of at most two of them
      const int e = (b % NS) * (L / (NS * R));
      T2 w[R];
      twiddle_powers<T2, R>(twf, e, L - 1, w);
#pragma unroll
      for (int r = 1; r < R; ++r) x[i][r] = INV ? cmulc(x[i][r], w[r]) : cmul(x[i][r], w[r]);
    }
    dftR<R, INV>(x[i]);
  }
  if constexpr (LAST) {
#pragma unroll
    for (int i = 0; i < NB; ++i)
#pragma unroll
      for (int r = 0; r < R; ++r) v[i + r * NB] = x[i][r];
  } else {
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      const int b = q + i * TPF;
      const int k = b % NS;
      const int o = (b / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) fib[lds_pad(o + r * NS)] = x[i][r];
    }
    __syncthreads();
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

template <typename T2, int L, int NS, bool INV>
__device__ __forceinline__ void stockham(T2 (&v)[16], T2 *fib, int q, const T2 *__restrict__ twf) {
  constexpr int R = stage_radix<L, NS>();
  constexpr bool FIRST = NS == 1;
  constexpr bool LAST = NS * R == L;
  stockham_stage<T2, L, R, NS, FIRST, LAST, INV>(v, fib, q, twf);
  if constexpr (!LAST) stockham<T2, L, NS * R, INV>(v, fib, q, twf);
}

// ---------------------------------------------------------------------------
// elementwise operations
// ---------------------------------------------------------------------------

// x-space op on natural index n.  The grating coordinate and mask follow
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
      double x;
      if (op.single_chunk) {
        x = (double)n * op.xstep + op.x0;
      } else {
        const int64_t c = n / op.chunk;
        const int64_t k = n - c * op.chunk;
        int64_t len = N - c * op.chunk;
        if (len > op.chunk) len = op.chunk;
        const int64_t start = c * op.chunk - N / 2;
        const double x_start = (double)start * op.dx + op.xpos;
        const double x_stop = x_start + (double)len * op.dx;
        const double step = (x_stop - x_start) / (double)len;
        x = (double)k * step + x_start;
      }
      // t = x/pitch + dc/2 as big-wave rounds it.  x * (1/pitch) differs from
      // the correctly rounded quotient by a few ulp of t; outside a guard band
      // around the mask edges (ph = 0 and ph = dc) that cannot change the
      // decision, inside it the exact quotient is used.
      double t = x * op.inv_pitch + op.half_dc;
      double ph = t - floor(t);
      const double guard = fabs(t) * 1e-13 + 1e-300;
      if (ph < guard || ph > 1.0 - guard || fabs(ph - op.dc) < guard) {
        t = x / op.pitch + op.half_dc;
        ph = t - floor(t);
      }
      if (ph < op.dc) { re = op.fa[0]; im = op.fa[1]; }
      else { re = op.fb[0]; im = op.fb[1]; }
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

// (cos, sin)(2 pi t) for t in [-1/2, 1/2] given in FP64.  The quadrant is
// split off in FP64 (exact: t4 = 4t, q = rint(t4), f = t4 - q in [-1/2, 1/2]),
// then x = f pi/2 in [-pi/4, pi/4] goes through the Taylor series of sin
// (to x^9, truncation < 2e-9) and cos (to x^10, < 2e-10) in the field
// precision; the quadrant rotates (cos, sin) by q * 90 degrees.
template <typename S> __device__ __forceinline__ void cis_rev(double t, S *c_out, S *s_out);

template <> __device__ __forceinline__ void cis_rev<float>(double t, float *c_out, float *s_out) {
  const double t4 = 4.0 * t;
  const double q = rint(t4);
  const float x = (float)((t4 - q) * 1.5707963267948966);
  const float x2 = x * x;
  float sp = -1.0f / 362880.0f;
  sp = fmaf(sp, x2, 1.0f / 5040.0f);
  sp = fmaf(sp, x2, -1.0f / 120.0f);
  sp = fmaf(sp, x2, 1.0f / 6.0f);
  // sin x = x - x^3 (1/6 - x^2/120 + ...)
  const float sn = fmaf(-x * x2, sp, x);
  float cp = -1.0f / 3628800.0f;
  cp = fmaf(cp, x2, 1.0f / 40320.0f);
  cp = fmaf(cp, x2, -1.0f / 720.0f);
  cp = fmaf(cp, x2, 1.0f / 24.0f);
  cp = fmaf(cp, x2, -0.5f);
  const float cs = fmaf(cp, x2, 1.0f);
  const int qi = (int)q & 3;
  // rotate by q quarter turns: (c, s) -> (-s, c) per turn
  const float c1 = (qi & 1) ? -sn : cs;
  const float s1 = (qi & 1) ? cs : sn;
  *c_out = (qi & 2) ? -c1 : c1;
  *s_out = (qi & 2) ? -s1 : s1;
}

template <> __device__ __forceinline__ void cis_rev<double>(double t, double *c_out, double *s_out) {
  sincospi(2.0 * t, s_out, c_out);
}

// spectrum multiply at FFT bin k: phase in revolutions computed in FP64 and
// reduced to [-1/2, 1/2] before the cast (big-wave/propagation.py:124-137).
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
  S s, c;
  cis_rev<S>(t, &c, &s);
  T2 h;
  h.x = c * (S)op.scale[0] - s * (S)op.scale[1];
  h.y = c * (S)op.scale[1] + s * (S)op.scale[0];
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
// M = L * S, as W_M^{q col} * (W_M^{TPF col})^m with the powers built like
// twiddle_powers (five split-table reads instead of sixteen).
template <typename T2, int L>
__device__ __forceinline__ void inter_twiddles(const Geom &g, int q, int64_t col, T2 (&tw)[16]) {
  const T2 *hi = (const T2 *)g.tw_hi;
  const T2 *lo = (const T2 *)g.tw_lo;
  const int64_t mask = (int64_t)L * g.S - 1;
  constexpr int TPF = L / 16;
  const int64_t eb = ((int64_t)TPF * col) & mask;
  const T2 wq = split_twiddle(hi, lo, g.tw_sh, ((int64_t)q * col) & mask);
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

// ---------------------------------------------------------------------------
// fiber pass
// ---------------------------------------------------------------------------

template <typename S, int L, int C, bool ROW>
__global__ void __launch_bounds__(C *(L / 16))
    fiber_pass(const FieldArgs *__restrict__ args, Geom g) {
  using T2 = typename C2T<S>::t;
  constexpr int TPF = L / 16;
  // fiber slab stride: odd, so that lanes working on neighbouring fibers at
  // the same offset (column passes) hit different LDS banks
  constexpr int FIB = (L + L / 16) | 1;
  __shared__ T2 lds[C * FIB];

  const FieldArgs &A = args[blockIdx.y];
  T2 *__restrict__ data = (T2 *)A.data;
  const T2 *__restrict__ twf = (const T2 *)g.tw_fib;
  const int t = threadIdx.x;
  const int c = ROW ? t / TPF : t % C;
  const int q = ROW ? t % TPF : t / C;
  T2 *fib = lds + c * FIB;

  int64_t base, js, col = 0, rr = 0;
  if (ROW) {
    rr = (int64_t)blockIdx.x * C + c;
    base = rr * L;
    js = 1;
  } else {
    // XCD-aware order: workgroups are dealt round-robin to the 8 XCDs; when a
    // row segment of C columns is shorter than a 128 B line, the `segs`
    // workgroups that share one line are given ids on the same XCD so that
    // its L2 merges their partial-line reads and writes
    int64_t bid = blockIdx.x;
    if (g.xcd_segs > 1) {
      const int64_t G = 8 * (int64_t)g.xcd_segs;
      if (bid < ((int64_t)gridDim.x / G) * G) {
        const int64_t r = bid % G;
        bid = (bid - r) + (r % 8) * g.xcd_segs + r / 8;
      }
    }
    const int64_t cpb = g.S / C;
    const int64_t hi = bid / cpb;
    col = (bid % cpb) * C + c;
    base = hi * (int64_t)L * g.S + col;
    js = g.S;
  }

  const bool xop = g.top && A.x.kind != XOP_NONE;
  const bool gen = xop && (A.x.kind == XOP_POINT || A.x.kind == XOP_ONES || A.x.from_ones);
  T2 v[16];
  if (!gen) {
#pragma unroll
    for (int m = 0; m < 16; ++m) v[m] = data[base + (int64_t)(q + m * TPF) * js];
  } else {
#pragma unroll
    for (int m = 0; m < 16; ++m) { v[m].x = 1; v[m].y = 0; }
  }

  if (!ROW && A.do_inv) {
    T2 tw[16];
    inter_twiddles<T2, L>(g, q, col, tw);
#pragma unroll
    for (int m = 0; m < 16; ++m) v[m] = cmulc(v[m], tw[m]);
    stockham<T2, L, 1, true>(v, fib, q, twf);
  }

  if (xop) {
#pragma unroll
    for (int m = 0; m < 16; ++m) {
      const int64_t n = base + (int64_t)(q + m * TPF) * js;
      v[m] = apply_xop(A.x, n, g.N, v[m]);
    }
  }

  if (A.do_fwd) {
    stockham<T2, L, 1, false>(v, fib, q, twf);
    if (!ROW) {
      T2 tw[16];  // recomputed rather than kept live across the FFTs (registers)
      inter_twiddles<T2, L>(g, q, col, tw);
#pragma unroll
      for (int m = 0; m < 16; ++m) v[m] = cmul(v[m], tw[m]);
    }
  }

  if (ROW && A.k.on) {
    const int64_t kb = row_kbase(g, rr);
    const int64_t kstep = g.N >> g.lgR;
#pragma unroll
    for (int m = 0; m < 16; ++m) {
      const int64_t k = kb + kstep * (q + m * TPF);
      v[m] = apply_kop(A.k, k, g.N, v[m]);
    }
  }

  if (ROW && A.do_inv) stockham<T2, L, 1, true>(v, fib, q, twf);

#pragma unroll
  for (int m = 0; m < 16; ++m) data[base + (int64_t)(q + m * TPF) * js] = v[m];
}

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

// One workgroup per pixel; the members of the launch are visited in order
// inside the workgroup, so members that share an accumulator row add into it
// without races and in a fixed order (results do not depend on scheduling).
template <typename S>
__global__ void __launch_bounds__(256) detect_kernel(const DetArgs *__restrict__ args, int nf,
                                                     const int64_t *__restrict__ bounds) {
  using T2 = typename C2T<S>::t;
  __shared__ double red[256];
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
    red[threadIdx.x] = s;
    __syncthreads();
#pragma unroll
    for (int w = 128; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      const double prev = A.accumulate ? A.acc[k] : 0.0;
      A.acc[k] = prev + A.weight * red[0];
    }
    __syncthreads();
  }
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

// columns per workgroup for a column pass of fiber length L: 16 columns of
// complex64 (128 B per row segment) / 8 of complex128, more for short fibers
template <typename S, int L>
static hipError_t launch_col(const FieldArgs *args, int nf, const Geom &g, hipStream_t st) {
  constexpr int C = col_fibers(L, sizeof(S) == 8);
  if (g.S % C != 0) return hipErrorInvalidValue;
  const int64_t nblk = g.N / ((int64_t)L * C);
  dim3 grid((unsigned)nblk, (unsigned)nf);
  hipLaunchKernelGGL((fiber_pass<S, L, C, false>), grid, dim3(C * (L / 16)), 0, st, args, g);
  return hipGetLastError();
}

template <typename S, int L>
static hipError_t launch_row(const FieldArgs *args, int nf, const Geom &g, hipStream_t st) {
  constexpr int C = L >= 256 ? 1 : 256 / L;
  const int64_t rows = g.N / L;
  if (rows % C != 0) return hipErrorInvalidValue;
  dim3 grid((unsigned)(rows / C), (unsigned)nf);
  hipLaunchKernelGGL((fiber_pass<S, L, C, true>), grid, dim3(C * (L / 16)), 0, st, args, g);
  return hipGetLastError();
}

template <typename S>
hipError_t launch_fiber_pass(int lgL, bool row, const FieldArgs *args, int nf, const Geom &g,
                             hipStream_t st) {
  if (row) {
    switch (lgL) {
      case 6: return launch_row<S, 64>(args, nf, g, st);
      case 7: return launch_row<S, 128>(args, nf, g, st);
      case 8: return launch_row<S, 256>(args, nf, g, st);
      case 9: return launch_row<S, 512>(args, nf, g, st);
      case 10: return launch_row<S, 1024>(args, nf, g, st);
      case 11: return launch_row<S, 2048>(args, nf, g, st);
      case 12: return launch_row<S, 4096>(args, nf, g, st);
      case 13: if constexpr (sizeof(S) == 4) return launch_row<S, 8192>(args, nf, g, st);
               else return hipErrorInvalidValue;
      default: return hipErrorInvalidValue;
    }
  }
  switch (lgL) {
    case 4: return launch_col<S, 16>(args, nf, g, st);
    case 5: return launch_col<S, 32>(args, nf, g, st);
    case 6: return launch_col<S, 64>(args, nf, g, st);
    case 7: return launch_col<S, 128>(args, nf, g, st);
    case 8: return launch_col<S, 256>(args, nf, g, st);
    case 9: return launch_col<S, 512>(args, nf, g, st);
    case 10: return launch_col<S, 1024>(args, nf, g, st);
    case 11: return launch_col<S, 2048>(args, nf, g, st);
    case 12: return launch_col<S, 4096>(args, nf, g, st);
    case 13: if constexpr (sizeof(S) == 4) return launch_col<S, 8192>(args, nf, g, st);
             else return hipErrorInvalidValue;
    default: return hipErrorInvalidValue;
  }
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

template hipError_t launch_fiber_pass<float>(int, bool, const FieldArgs *, int, const Geom &, hipStream_t);
template hipError_t launch_fiber_pass<double>(int, bool, const FieldArgs *, int, const Geom &, hipStream_t);
template hipError_t launch_xop<float>(const FieldArgs *, int, int64_t, hipStream_t);
template hipError_t launch_xop<double>(const FieldArgs *, int, int64_t, hipStream_t);
template hipError_t launch_kop<float>(const FieldArgs *, int, const Geom &, hipStream_t);
template hipError_t launch_kop<double>(const FieldArgs *, int, const Geom &, hipStream_t);
template hipError_t launch_detect<float>(const DetArgs *, int, const int64_t *, int64_t, hipStream_t);
template hipError_t launch_detect<double>(const DetArgs *, int, const int64_t *, int64_t, hipStream_t);
template hipError_t launch_copy<float>(void *, const void *, int64_t, hipStream_t);
template hipError_t launch_copy<double>(void *, const void *, int64_t, hipStream_t);

}  // namespace rs
