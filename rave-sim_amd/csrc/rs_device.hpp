// Structures shared between the host scheduler and the gfx950 kernels.
//
// A field of N = 2^n complex samples is transformed with a digit-group
// (four-step) FFT.  The index bits are split into 1..3 groups: the lowest
// group is the contiguous "row" of length Lr, the higher groups are strided
// "columns".  One split-step propagation is
//   [top column pass: inverse col DFT of the previous slice, x-space multiply
//    (grating / substrate / source generation), forward col DFT + twiddle]
//   [middle column pass, forward]            (3 groups only)
//   [row pass: row DFT, spectrum multiply exp(i phi(k)) * cutoff * 1/N, row iDFT]
//   [middle column pass, inverse]            (3 groups only)
// so that the field never returns to natural order between slices.
#pragma once
#include <stdint.h>

namespace rs {

enum XopKind : int32_t {
  XOP_NONE = 0,
  XOP_MASK = 1,    // grating: factor = frac(x/pitch + dc/2) < dc ? fa : fb
  XOP_SCALAR = 2,  // u *= fa
  XOP_POINT = 3,   // u = exp(i kphase r) / sqrt(r)   (generated, input not read)
  XOP_ONES = 4,    // u = 1 (plane wave, generated)
  XOP_SAMPLE = 5,  // sample row: u *= exp(i a interp(x; rx, db)) (big-wave Sample.apply)
  XOP_DBA = 6,     // DBA grating: frac(x/pitch + dc/2) < dc and frac(x/pitch1) < dc1 ? fa : fb
};

// x-space (natural order) elementwise operation.  Every kind is followed by a
// multiplication with `post` (substrate factors and other scalars fold here).
struct XOp {
  int32_t kind;
  int32_t single_chunk;  // 1 when N <= chunk (grating x needs no chunk index)
  int32_t from_ones;     // MASK / SCALAR applied to a plane wave of ones (input not read)
  int32_t pad_;
  double fa[2], fb[2];
  double post[2];
  // grating (big-wave/optical_element.py:341-357)
  double pitch, dc, half_dc, xpos, dx;
  double inv_pitch;       // 1/pitch for the fast mask test (exact division near edges)
  double x0, xstep;       // single chunk: linspace start and step, computed on the host
  int64_t chunk;
  // second pitch and duty cycle of a DBA grating
  // (big-wave/optical_element.py:359-375)
  double pitch1, inv_pitch1, dc1;
  // point source (big-wave/propagation.py:141-151)
  double xs, z, kphase;
  // fast mask test (single chunk), set by finalize_xop on the host:
  // t ~ fma(n, m_a, m_b) is within m_guard of big-wave's x/pitch + dc/2, and
  // samples whose fraction lies within m_guard of an edge use the exact path
  double m_a, m_b, m_guard;
  // optional cached decisions of a top column pass (host-filled per run):
  // bit m of word [tile * threads + thread] is the decision for v[m]
  const void *mbits;
  // sample row (big-wave/optical_element.py:432-476): pixel positions rx[k],
  // delta + i beta per pixel (interleaved), the factor exp(i a db) per pixel
  // (interleaved), a = 2 pi pixel_size_z / wl; x = (n - N/2) dx
  const double *s_rx;
  const double *s_db;
  const double *s_f;
  int64_t s_len;
  double s_inv_px, s_a;
};

// compile-time x-space variants of a fiber pass: none (middle / row passes),
// grating mask or scalar with the factors folded on the host, or any kind
enum XopVariant : int32_t { XV_NONE = 0, XV_MASK = 1, XV_GEN = 2 };

// entries of the complex64 row pass's phase table exp(2 pi i j / kCisTab)
constexpr int kCisTab = 256;

// spectrum multiply in the row pass: exp(2 pi i (alpha k'^2 + beta)) * scale,
// zero for k in [rs, re) (big-wave/propagation.py:74-138).
struct KOp {
  int32_t on;
  int32_t pad_;
  double alpha_rev, beta_rev;
  int64_t rs, re;
  double scale[2];
  // complex64 row passes: Taylor coefficients of exp(i w f) * scale[0] with
  // w = 2 pi / 256 (phase per unit of the table residual f), formed on the
  // host: sc * w, -sc * w^3 / 6, -sc * w^2 / 2, sc
  float t_s1, t_s3, t_c2, t_sc;
};

struct FieldArgs {
  void *data;      // field base pointer (float2* or double2*)
  int32_t do_inv;  // column: conj twiddle + inverse DFT first; row: inverse DFT last
  int32_t do_fwd;  // forward DFT (+ twiddle for columns)
  XOp x;
  KOp k;
};

// detector launch record (one per field)
struct DetArgs {
  const void *data;
  double *acc;      // n_px accumulators of this member's output row
  double weight;
  int32_t accumulate;
  int32_t pad_;
};

// geometry of one pass over a plan
struct Geom {
  int64_t N;
  int64_t S;       // element stride of this pass's fibers (1 for the row pass)
  int32_t top;     // this pass holds the highest digit group (x-space ops allowed)
  int32_t ngroups;
  int32_t lgA, lgB, lgR;  // log2 of top / middle / row group sizes (lgB = 0 if 2 groups)
  // inter-group twiddle W_M^e, e = k*lo, M = L*S, split as hi[e >> sh] * lo[e & mask]
  int32_t tw_sh;
  int32_t xcd_segs;  // column passes: workgroups sharing one 128 B line (1 = none)
  int32_t tile;      // column passes: tile width variant (col_fibers), 0 = default
  int32_t nt;        // top column passes of up to 512: non-temporal field loads (fields streaming from HBM)
  int32_t one_dir;   // row passes: every field runs one direction only (plain-form FFT passes)
  const void *tw_hi;
  const void *tw_lo;
  const void *tw_fib;  // W_L^j, j < L, then the small-stage tables (rs_kernels.hpp)
  const void *cis_tab;  // complex64 plans: exp(2 pi i j / 256), j < 256 (row pass phase)
  // samples [span_off, span_off + span_len) of every field (span_len 0: the
  // whole field); a span is a whole number of top-digit blocks, so that the
  // pass's tiles inside it are the contiguous ids tile0 .. tile0 + grid - 1
  // (tile0 is set by the launcher)
  int64_t span_off, span_len, tile0;
};

}  // namespace rs
