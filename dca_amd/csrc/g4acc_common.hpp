// Shared device types and helpers of the g4acc HIP library (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace g4acc {

// Complex numbers are HIP's double2 / float2 (x = re, y = im); Cplx<Real> picks one.
template <typename Real>
struct CplxOf;
template <>
struct CplxOf<double> {
  using type = double2;
};
template <>
struct CplxOf<float> {
  using type = float2;
};
template <typename Real>
using Cplx = typename CplxOf<Real>::type;

template <typename C>
__device__ __forceinline__ C cmul(C a, C b) {
  return C{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// a * conj(b)
template <typename C>
__device__ __forceinline__ C cmulc(C a, C b) {
  return C{a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
template <typename C>
__device__ __forceinline__ C cadd(C a, C b) {
  return C{a.x + b.x, a.y + b.y};
}
template <typename C>
__device__ __forceinline__ C csub(C a, C b) {
  return C{a.x - b.x, a.y - b.y};
}
// acc += a * b
template <typename C>
__device__ __forceinline__ void cfma(C& acc, C a, C b) {
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(-a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(a.y, b.x, acc.y);
}
// acc -= a * b
template <typename C>
__device__ __forceinline__ void cfms(C& acc, C a, C b) {
  acc.x = fma(-a.x, b.x, acc.x);
  acc.x = fma(a.y, b.y, acc.x);
  acc.y = fma(-a.x, b.y, acc.y);
  acc.y = fma(-a.y, b.x, acc.y);
}
// acc += a * conj(b)
template <typename C>
__device__ __forceinline__ void cfmac(C& acc, C a, C b) {
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(a.y, b.y, acc.x);
  acc.y = fma(a.y, b.x, acc.y);
  acc.y = fma(-a.x, b.y, acc.y);
}
__device__ __forceinline__ double2 to_d(double2 a) { return a; }
__device__ __forceinline__ double2 to_d(float2 a) { return double2{a.x, a.y}; }
template <typename C>
__device__ __forceinline__ C from_d(double2 a);
template <>
__device__ __forceinline__ double2 from_d<double2>(double2 a) { return a; }
template <>
__device__ __forceinline__ float2 from_d<float2>(double2 a) {
  return float2{static_cast<float>(a.x), static_cast<float>(a.y)};
}

// One spin sector of one walker ("slot" = 2 * walker + spin), as resolved to device pointers.
struct SlotDesc {
  const double* m;     // column-major N x N, leading dimension ld
  const double* tau;   // N
  const int* site;     // N
  const int* band;     // N
  int n;
  int ld;
  double sign;         // the walker's sign (same for both spins)
  // Non-null: m holds the walker's N matrix and M = diag(expv - 1) N is formed where M is read
  // (CtauxWalker::computeM, ctaux_walker.hpp:1889-1916); expv[i] = exp(V) of vertex i.
  const double* expv;
  // General model: the vertices' left site and band (nullptr: the same as site / band, i.e.
  // density-density vertices); site / band are their right labels, which order the rows of M, and
  // the left ones order its columns (cached_ndft_base.hpp:97-103).
  const int* site_l;
  const int* band_l;
  double sign_im;      // imaginary part of a complex sign (0 for real models)
  int m_cplx;          // m holds interleaved complex entries (ld counts complex entries)
};

// Static geometry passed by value to kernels.
struct Geo {
  int nb, nc, nw, nwE, ext, n_orb, D;
  int nq, nv;
  double beta;
};

}  // namespace g4acc
