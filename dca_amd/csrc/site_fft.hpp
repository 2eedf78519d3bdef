// Site transforms of an L x L square cluster held in registers: sites labelled x + L y, momenta
// kx + L ky, exp(+-i k.r) = exp(+-2 pi i (kx x + ky y) / L).
//
// Every function here is __host__ __device__ and fully unrolled: all array indices are
// compile-time constants once the callers' loops unroll, so the arrays stay in registers. The
// same code runs on the CPU in tests/cpp/site_fft_test.cpp against a direct double sum.
#ifndef G4ACC_SITE_FFT_HPP
#define G4ACC_SITE_FFT_HPP

#include "g4acc_common.hpp"

namespace g4acc {

// cos(2 pi m / L) and sin(2 pi m / L) for L = 2, 4, 6, 8 (0 <= m < L).
__host__ __device__ constexpr double unit_cos(int L, int m) {
  return L == 6 ? (m == 0 ? 1.0 : m == 3 ? -1.0 : (m == 1 || m == 5) ? 0.5 : -0.5)
                : (m * (8 / L) == 0 ? 1.0 : m * (8 / L) == 4 ? -1.0 : (m * (8 / L) == 2 || m * (8 / L) == 6) ? 0.0
                   : (m * (8 / L) == 1 || m * (8 / L) == 7) ? 0.70710678118654752440
                                                            : -0.70710678118654752440);
}
__host__ __device__ constexpr double unit_sin(int L, int m) {
  return L == 6 ? ((m == 0 || m == 3) ? 0.0 : (m == 1 || m == 2) ? 0.86602540378443864676
                                                                 : -0.86602540378443864676)
                : ((m * (8 / L) == 0 || m * (8 / L) == 4) ? 0.0 : m * (8 / L) == 2 ? 1.0 : m * (8 / L) == 6 ? -1.0
                   : (m * (8 / L) == 1 || m * (8 / L) == 3) ? 0.70710678118654752440
                                                            : -0.70710678118654752440);
}

template <typename C>
__host__ __device__ __forceinline__ C sf_add(C a, C b) {
  return C{a.x + b.x, a.y + b.y};
}
template <typename C>
__host__ __device__ __forceinline__ C sf_sub(C a, C b) {
  return C{a.x - b.x, a.y - b.y};
}

// v * exp(-2 pi i m / L). Quarter turns are exact swaps and negations.
template <int L, typename Real>
__host__ __device__ __forceinline__ Cplx<Real> rot(Cplx<Real> v, int m) {
  m = ((m % L) + L) % L;
  if ((4 * m) % L == 0) {
    const int q = (4 * m) / L;
    return q == 0 ? v : q == 1 ? Cplx<Real>{v.y, -v.x} : q == 2 ? Cplx<Real>{-v.x, -v.y} : Cplx<Real>{-v.y, v.x};
  }
  const Real c = static_cast<Real>(unit_cos(L, m)), sn = static_cast<Real>(unit_sin(L, m));
  return Cplx<Real>{v.x * c + v.y * sn, v.y * c - v.x * sn};
}

// Bit reversal of i in log2(N) bits.
template <int N>
__host__ __device__ constexpr int bit_reverse(int i) {
  int r = 0;
  for (int n = N; n > 1; n >>= 1) {
    r = (r << 1) | (i & 1);
    i >>= 1;
  }
  return r;
}

// In place: x[o + s k] <- sum_n x[o + s n] exp(SIGN 2 pi i k n / N), N a power of two, radix-2
// decimation in time (N/2 log2 N butterflies; the twiddles of N <= 8 are +-1, +-i and
// (+-1 +-i)/sqrt(2)).
template <int N, int SIGN, typename Real, int SZ>
__host__ __device__ __forceinline__ void fft_pow2(Cplx<Real> (&x)[SZ], int o, int s) {
  using C = Cplx<Real>;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const int j = bit_reverse<N>(i);
    if (j > i) {
      const C t = x[o + s * i];
      x[o + s * i] = x[o + s * j];
      x[o + s * j] = t;
    }
  }
#pragma unroll
  for (int len = 2; len <= N; len <<= 1) {
#pragma unroll
    for (int i = 0; i < N; i += len) {
#pragma unroll
      for (int j = 0; j < len / 2; ++j) {
        // twiddle exp(SIGN 2 pi i j / len) = exp(-2 pi i m / N), m = -SIGN j N / len (mod N)
        const int m = -SIGN * j * (N / len);
        const C u = x[o + s * (i + j)];
        const C t = rot<N, Real>(x[o + s * (i + j + len / 2)], m);
        x[o + s * (i + j)] = sf_add(u, t);
        x[o + s * (i + j + len / 2)] = sf_sub(u, t);
      }
    }
  }
}

// In place: v[kx + L ky] <- sum_{x,y} v[x + L y] exp(SIGN 2 pi i (kx x + ky y) / L), L = 2, 4, 8:
// L transforms over y, then L over x (2 L (L/2) log2 L butterflies: 64 for 4 x 4, 384 for 8 x 8,
// against L^4 complex products for the dense site sum).
template <int L, int SIGN, typename Real>
__host__ __device__ __forceinline__ void site_fft_2d(Cplx<Real> (&v)[L * L]) {
  static_assert(L == 2 || L == 4 || L == 8, "radix-2 site transform");
#pragma unroll
  for (int x = 0; x < L; ++x)
    fft_pow2<L, SIGN, Real>(v, x, L);
#pragma unroll
  for (int ky = 0; ky < L; ++ky)
    fft_pow2<L, SIGN, Real>(v, L * ky, 1);
}

// The same transform as two passes of L-point sums (2 L^3 rotations), for L = 6 (no radix-2).
template <int L, int SIGN, typename Real>
__host__ __device__ __forceinline__ void site_dft_2d(Cplx<Real> (&v)[L * L]) {
  using C = Cplx<Real>;
#pragma unroll
  for (int x = 0; x < L; ++x) {
    C col[L];
#pragma unroll
    for (int y = 0; y < L; ++y)
      col[y] = v[x + L * y];
#pragma unroll
    for (int ky = 0; ky < L; ++ky) {
      C acc = col[0];
#pragma unroll
      for (int y = 1; y < L; ++y)
        acc = sf_add(acc, rot<L, Real>(col[y], -SIGN * ky * y));
      v[x + L * ky] = acc;
    }
  }
#pragma unroll
  for (int ky = 0; ky < L; ++ky) {
    C row[L];
#pragma unroll
    for (int x = 0; x < L; ++x)
      row[x] = v[x + L * ky];
#pragma unroll
    for (int kx = 0; kx < L; ++kx) {
      C acc = row[0];
#pragma unroll
      for (int x = 1; x < L; ++x)
        acc = sf_add(acc, rot<L, Real>(row[x], -SIGN * kx * x));
      v[kx + L * ky] = acc;
    }
  }
}

// One L-point transform in place: x[k] <- sum_n x[n] exp(SIGN 2 pi i k n / L), radix 2 for
// L = 2, 4, 8 and the direct sum of rotations for L = 6.
template <int L, int SIGN, typename Real>
__host__ __device__ __forceinline__ void site_transform_1d(Cplx<Real> (&x)[L]) {
  if constexpr (L == 2 || L == 4 || L == 8) {
    fft_pow2<L, SIGN, Real>(x, 0, 1);
  }
  else {
    using C = Cplx<Real>;
    C y[L];
#pragma unroll
    for (int k = 0; k < L; ++k) {
      C acc = x[0];
#pragma unroll
      for (int n = 1; n < L; ++n)
        acc = sf_add(acc, rot<L, Real>(x[n], -SIGN * k * n));
      y[k] = acc;
    }
#pragma unroll
    for (int k = 0; k < L; ++k)
      x[k] = y[k];
  }
}

// Radix-2 where it applies, the two-pass sum otherwise.
template <int L, int SIGN, typename Real>
__host__ __device__ __forceinline__ void site_transform_2d(Cplx<Real> (&v)[L * L]) {
  if constexpr (L == 2 || L == 4 || L == 8)
    site_fft_2d<L, SIGN, Real>(v);
  else
    site_dft_2d<L, SIGN, Real>(v);
}

}  // namespace g4acc

#endif
