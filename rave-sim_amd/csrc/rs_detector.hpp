// Detector pixel boundaries: which samples np.histogram puts into each
// pixel, as big-wave bins |u|^2 (big-wave/propagation.py:293-308).
//
// Samples sit at x_j = (j - N/2) dx.  The n_px + 1 uniform bins span
// [-(n_px/2) px, (n_px/2 + 1) px) with edges from np.linspace; a sample's bin
// follows numpy's rule (numpy/lib/_histograms_impl.py): an index from the
// scaled position, then moved down or up by one where the computed edges
// disagree with it; the extra last bin is dropped.  x_j increases with j, so
// pixel k holds the samples [bound(k), bound(k + 1)) and each bound is a
// binary search over j.
//
// One implementation for the host (engine.cpp) and the device (the detector
// kernels of the allocation-free entry point), evaluated without
// contraction so both give the same bounds bit for bit.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace rs {

struct DetGeom {
  int64_t N;
  double dx;
  int64_t nb;  // bins: n_px + 1
  double first, last, step;
};

__host__ __device__ inline DetGeom det_geom(int64_t N, double dx, int64_t npx, double px) {
#pragma clang fp contract(off)
  DetGeom g;
  g.N = N;
  g.dx = dx;
  g.nb = npx + 1;
  g.first = (-((double)npx / 2.0)) * px;
  g.last = ((double)npx / 2.0 + 1.0) * px;
  // np.linspace(first, last, nb + 1): step = delta / div, edge i = i * step + first
  g.step = (g.last - g.first) / (double)g.nb;
  return g;
}

__host__ __device__ inline double det_edge(const DetGeom &g, int64_t i) {
#pragma clang fp contract(off)
  return i == g.nb ? g.last : (double)i * g.step + g.first;
}

// bin of sample j: -1 below the range, nb + 1 above it
__host__ __device__ inline int64_t det_bin(const DetGeom &g, int64_t j) {
#pragma clang fp contract(off)
  const double x = ((double)j - (double)g.N / 2.0) * g.dx;
  if (x < g.first) return -1;
  if (x > g.last) return g.nb + 1;
  const double f = ((x - g.first) / (g.last - g.first)) * (double)g.nb;
  int64_t idx = (int64_t)f;
  if (idx == g.nb) idx -= 1;
  if (x < det_edge(g, idx)) idx -= 1;
  if (x >= det_edge(g, idx + 1) && idx != g.nb - 1) idx += 1;
  return idx;
}

// first sample j in [0, N] with det_bin(j) >= k
__host__ __device__ inline int64_t det_bound(const DetGeom &g, int64_t k) {
  int64_t lo = 0, hi = g.N;
  while (lo < hi) {
    const int64_t mid = lo + (hi - lo) / 2;
    if (det_bin(g, mid) >= k) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

}  // namespace rs
