// Host-only checks of the C++ TpAccumulator mirror (include/g4acc.hpp) that need no GPU: the
// reference's std::logic_error cases raised before any device call.
//   * a vertex with different left and right orbitals (non-density-density,
//     tp/ndft/cached_ndft_base.hpp:72-78) is refused;
//   * fewer single-particle than extended vertex frequencies is refused
//     (tp_accumulator_base.hpp:197-198), and the slice takes the centred frequencies
//     (tp_accumulator_base.hpp:210-221).
// Prints "OK" and exits 0 when every check passes.
#include <complex>
#include <cstdio>
#include <stdexcept>
#include <vector>

#include "g4acc.hpp"

namespace {

struct Vertex {
  double tau;
  int lb, ls, rb, rs;
  double get_tau() const { return tau; }
  int get_left_band() const { return lb; }
  int get_left_site() const { return ls; }
  int get_right_band() const { return rb; }
  int get_right_site() const { return rs; }
};

int failures = 0;

void expect(bool ok, const char* what) {
  if (!ok) {
    std::printf("FAIL %s\n", what);
    ++failures;
  }
}

}  // namespace

int main() {
  std::vector<double> tau;
  std::vector<int> site, band;

  std::vector<Vertex> dd{{0.1, 1, 2, 1, 2}, {0.7, 0, 3, 0, 3}};
  g4acc::extract_configuration(dd, tau, site, band);
  expect(tau.size() == 2 && tau[1] == 0.7 && site[0] == 2 && band[0] == 1 && site[1] == 3,
         "density-density vertices are read from the right band and site");

  for (const Vertex bad : {Vertex{0.2, 0, 1, 1, 1}, Vertex{0.2, 1, 0, 1, 1}}) {
    std::vector<Vertex> conf{dd[0], bad};
    bool thrown = false;
    try {
      g4acc::extract_configuration(conf, tau, site, band);
    }
    catch (const std::logic_error&) {
      thrown = true;
    }
    expect(thrown, "a vertex with different left and right orbitals raises std::logic_error");
  }

  {
    // Models with non-density-density vertices read both label sets.
    std::vector<int> site_l, band_l;
    std::vector<Vertex> nd{{0.2, 0, 1, 2, 3}, {0.4, 1, 0, 1, 2}};
    g4acc::extract_configuration_general(nd, tau, site, band, site_l, band_l);
    expect(site[0] == 3 && band[0] == 2 && site_l[0] == 1 && band_l[0] == 0 && site[1] == 2 &&
               band[1] == 1 && site_l[1] == 0 && band_l[1] == 1 && tau[1] == 0.4,
           "left and right labels of non-density-density vertices");
    expect(g4acc::is_complex<std::complex<double>>::value && !g4acc::is_complex<double>::value,
           "complex scalars select the general entry");
  }

  // G0(b1, s1, b2, s2, k, w) = w + 100 k + 1000 b1 + 2000 b2 + 10000 s1 (+ i s2)
  auto g0 = [](int b1, int s1, int b2, int s2, int k, int w) {
    return std::complex<double>(w + 100. * k + 1000. * b1 + 2000. * b2 + 10000. * s1, s2);
  };
  const int nb = 2, nc = 3;
  bool thrown = false;
  try {
    (void)g4acc::slice_g0(g0, nb, nc, 6, 8);
  }
  catch (const std::logic_error&) {
    thrown = true;
  }
  expect(thrown, "too few single-particle frequencies raises std::logic_error");
  const auto sl = g4acc::slice_g0(g0, nb, nc, 10, 6);
  expect(sl.size() == static_cast<std::size_t>(nb * nb * 2 * nc * 6), "slice size");
  bool all = true;
  for (int w = 0; w < 6; ++w)
    for (int k = 0; k < nc; ++k)
      for (int s = 0; s < 2; ++s)
        for (int b2 = 0; b2 < nb; ++b2)
          for (int b1 = 0; b1 < nb; ++b1)
            all &= sl[b1 + nb * (b2 + nb * (s + 2 * (k + nc * w)))] == g0(b1, s, b2, s, k, w + 2);
  expect(all, "slice is (b1, b2, s, k, wE) with the centred frequencies and equal spins");

  if (failures)
    return 1;
  std::printf("OK\n");
  return 0;
}
