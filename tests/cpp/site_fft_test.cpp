// CPU check of the register site transforms of the M -> G kernels (dca_amd/csrc/site_fft.hpp):
// the same __host__ __device__ code, run on the host against the direct double sum
//   v[kx + L ky] = sum_{x,y} v0[x + L y] exp(SIGN 2 pi i (kx x + ky y) / L)
// for L = 2, 4, 6, 8, both signs, in double and float, and the single L-point transform of the
// 6x6 / 8x8 site-transform kernel. Prints "OK" and exits 0 when all agree (double: 1e-13, float:
// 2e-6 of the largest output).
#include <cmath>
#include <complex>
#include <cstdio>
#include <random>

#include "../../dca_amd/csrc/site_fft.hpp"

namespace {

int failures = 0;

template <int L, int SIGN, typename Real>
void check(std::mt19937_64& rng) {
  using C = g4acc::Cplx<Real>;
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  C v[L * L];
  std::complex<double> v0[L * L];
  for (int i = 0; i < L * L; ++i) {
    v0[i] = {u(rng), u(rng)};
    v[i] = C{static_cast<Real>(v0[i].real()), static_cast<Real>(v0[i].imag())};
    v0[i] = {static_cast<double>(v[i].x), static_cast<double>(v[i].y)};
  }
  g4acc::site_transform_2d<L, SIGN, Real>(v);
  const double two_pi = 6.283185307179586476925;
  double err = 0, mx = 0;
  for (int ky = 0; ky < L; ++ky)
    for (int kx = 0; kx < L; ++kx) {
      std::complex<double> acc = 0;
      for (int y = 0; y < L; ++y)
        for (int x = 0; x < L; ++x)
          acc += v0[x + L * y] * std::polar(1.0, SIGN * two_pi * (kx * x + ky * y) / L);
      const std::complex<double> got(v[kx + L * ky].x, v[kx + L * ky].y);
      err = std::max(err, std::abs(got - acc));
      mx = std::max(mx, std::abs(acc));
    }
  const double tol = sizeof(Real) == 8 ? 1e-13 : 2e-6;
  if (!(err <= tol * mx)) {
    std::printf("FAIL L=%d sign=%d %s: max error %.3e of %.3e\n", L, SIGN,
                sizeof(Real) == 8 ? "double" : "float", err, mx);
    ++failures;
  }
}

// One L-point transform (site_transform_1d, the passes of the 6x6 / 8x8 site-transform kernel).
template <int L, int SIGN, typename Real>
void check_1d(std::mt19937_64& rng) {
  using C = g4acc::Cplx<Real>;
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  C v[L];
  std::complex<double> v0[L];
  for (int i = 0; i < L; ++i) {
    v[i] = C{static_cast<Real>(u(rng)), static_cast<Real>(u(rng))};
    v0[i] = {static_cast<double>(v[i].x), static_cast<double>(v[i].y)};
  }
  g4acc::site_transform_1d<L, SIGN, Real>(v);
  const double two_pi = 6.283185307179586476925;
  double err = 0, mx = 0;
  for (int k = 0; k < L; ++k) {
    std::complex<double> acc = 0;
    for (int n = 0; n < L; ++n)
      acc += v0[n] * std::polar(1.0, SIGN * two_pi * k * n / L);
    err = std::max(err, std::abs(std::complex<double>(v[k].x, v[k].y) - acc));
    mx = std::max(mx, std::abs(acc));
  }
  const double tol = sizeof(Real) == 8 ? 1e-14 : 1e-6;
  if (!(err <= tol * mx)) {
    std::printf("FAIL 1d L=%d sign=%d %s: max error %.3e of %.3e\n", L, SIGN,
                sizeof(Real) == 8 ? "double" : "float", err, mx);
    ++failures;
  }
}

template <int L>
void check_all(std::mt19937_64& rng) {
  for (int rep = 0; rep < 4; ++rep) {
    check<L, 1, double>(rng);
    check<L, -1, double>(rng);
    check<L, 1, float>(rng);
    check<L, -1, float>(rng);
    check_1d<L, 1, double>(rng);
    check_1d<L, -1, double>(rng);
    check_1d<L, 1, float>(rng);
    check_1d<L, -1, float>(rng);
  }
}

}  // namespace

int main() {
  std::mt19937_64 rng(0x5173);
  check_all<2>(rng);
  check_all<4>(rng);
  check_all<6>(rng);
  check_all<8>(rng);
  if (failures)
    return 1;
  std::printf("OK\n");
  return 0;
}
