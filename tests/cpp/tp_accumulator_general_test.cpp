// Drives the C++ TpAccumulator mirror (include/g4acc.hpp) on the general model: complex M and
// complex signs (SignType<complex>, type_help.hpp:299-312), vertices whose left (site, band)
// differ from their right ones (cached_ndft_base.hpp:97-103), and models that are not spin
// symmetric (tp_accumulator_cpu.hpp:668-757). The case file is written by
// tests/test_cpp_wrapper.py: geometry, model flags, measurements and the FP64 oracle's G4.
//
// usage: tp_accumulator_general_test <case file> <precision 0|1> <l_inf tolerance>
#include <array>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <vector>

#include "g4acc.hpp"

namespace {

struct Vertex {
  double tau;
  int band, site, band_l, site_l;
  double get_tau() const { return tau; }
  int get_right_band() const { return band; }
  int get_right_site() const { return site; }
  int get_left_band() const { return band_l; }
  int get_left_site() const { return site_l; }
};

template <class Scalar>
struct Matrix {
  int n = 0;
  std::vector<Scalar> v;
  const Scalar* ptr() const { return v.data(); }
  int leadingDimension() const { return n > 0 ? n : 1; }
  int nrRows() const { return n; }
};

template <class Scalar>
struct Measurement {
  std::array<Matrix<Scalar>, 2> M;
  std::array<std::vector<Vertex>, 2> config;
  std::complex<double> sign;
};

template <class T>
void read(std::ifstream& in, T* p, std::size_t n) {
  in.read(reinterpret_cast<char*>(p), static_cast<std::streamsize>(sizeof(T) * n));
  if (!in)
    throw std::runtime_error("case file truncated");
}

int fail(const std::string& what) {
  std::cout << "FAIL " << what << std::endl;
  return 1;
}

template <class Scalar>
int run(std::ifstream& in, const g4acc::Geometry& geo, const g4acc_model& model, int n_meas, int n_g4,
        int precision, double tol) {
  std::vector<Measurement<Scalar>> ms(n_meas);
  int n_max = 1;
  for (auto& m : ms) {
    int n[2];
    double sign[2];
    read(in, n, 2);
    read(in, sign, 2);
    m.sign = {sign[0], sign[1]};
    for (int s = 0; s < 2; ++s) {
      n_max = std::max(n_max, n[s]);
      m.M[s].n = n[s];
      m.M[s].v.resize(static_cast<std::size_t>(n[s]) * n[s]);
      read(in, m.M[s].v.data(), m.M[s].v.size());
      std::vector<double> tau(n[s]);
      std::vector<int> site(n[s]), band(n[s]), site_l(n[s]), band_l(n[s]);
      read(in, tau.data(), n[s]);
      read(in, site.data(), n[s]);
      read(in, band.data(), n[s]);
      if (model.non_density_density) {
        read(in, site_l.data(), n[s]);
        read(in, band_l.data(), n[s]);
      }
      else {
        site_l = site;
        band_l = band;
      }
      for (int i = 0; i < n[s]; ++i)
        m.config[s].push_back(Vertex{tau[i], band[i], site[i], band_l[i], site_l[i]});
    }
  }
  const int n_ch = static_cast<int>(geo.channels.size());
  std::vector<std::vector<std::complex<double>>> expected(n_ch, std::vector<std::complex<double>>(n_g4));
  for (auto& e : expected)
    read(in, e.data(), e.size());

  g4acc_options opt{};
  opt.precision = precision;
  opt.sum_mode = G4ACC_SUM_FP64;
  opt.max_walkers = 1;
  opt.max_vertices = n_max;
  opt.device = 0;
  // Half of the measurements on each accumulator, merged with sumTo as the threaded solver does
  // (stdthread_qmci_cluster_solver.hpp:571-572).
  g4acc::TpAccumulator acc(geo, opt, false, model), other(geo, opt, false, model);
  acc.resetAccumulation(0);
  other.resetAccumulation(0);
  for (int i = 0; i < n_meas; ++i) {
    auto& a = (i % 2) ? other : acc;
    a.accumulate(ms[i].M, ms[i].config, ms[i].sign);
    a.synchronizeCopy();
  }
  other.sumTo(acc);
  acc.finalize();
  const auto& G4 = acc.get_G4();
  for (int c = 0; c < n_ch; ++c) {
    double num = 0, den = 0;
    for (int i = 0; i < n_g4; ++i) {
      num = std::max(num, std::abs(G4[c][i] - expected[c][i]));
      den = std::max(den, std::abs(expected[c][i]));
    }
    std::printf("channel %d l_inf %.3e\n", geo.channels[c], num / den);
    if (!(num / den < tol))
      return fail("G4 mismatch");
  }
  if constexpr (std::is_same<Scalar, std::complex<double>>::value) {
    // Complex M on an accumulator built without complex_m: std::logic_error.
    g4acc_model real_model = model;
    real_model.complex_m = 0;
    g4acc::TpAccumulator r(geo, opt, false, real_model);
    r.resetAccumulation(0);
    for (const auto& m : ms) {
      if (m.M[0].n == 0 && m.M[1].n == 0)
        continue;
      try {
        r.accumulate(m.M, m.config, m.sign);
        return fail("complex M on a real accumulator did not throw");
      }
      catch (const std::logic_error&) {
      }
      break;
    }
  }
  std::cout << "OK" << std::endl;
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 4) {
    std::cerr << "usage: " << argv[0] << " <case> <precision> <tol>\n";
    return 2;
  }
  std::ifstream in(argv[1], std::ios::binary);
  int hdr[11];
  read(in, hdr, 11);
  const int nb = hdr[0], nc = hdr[1], n_freq = hdr[2], n_wex = hdr[3], n_kex = hdr[4],
            n_ch = hdr[5], n_meas = hdr[6], n_g4 = hdr[7];
  const g4acc_model model{hdr[8], hdr[9], hdr[10]};
  g4acc::Geometry geo;
  read(in, &geo.beta, 1);
  geo.n_bands = nb;
  geo.n_sites = nc;
  geo.n_freq = n_freq;
  geo.w_ex.resize(n_wex);
  geo.k_ex.resize(n_kex);
  geo.channels.resize(n_ch);
  geo.k_add.resize(nc * nc);
  geo.k_sub.resize(nc * nc);
  read(in, geo.w_ex.data(), n_wex);
  read(in, geo.k_ex.data(), n_kex);
  read(in, geo.channels.data(), n_ch);
  read(in, geo.k_add.data(), nc * nc);
  read(in, geo.k_sub.data(), nc * nc);
  int ext = 0;
  for (int w : geo.w_ex)
    ext = std::max(ext, std::abs(w));
  geo.dft.resize(nc * nc);
  geo.g0.resize(static_cast<std::size_t>(nb) * nb * 2 * nc * (n_freq + 2 * ext));
  read(in, geo.dft.data(), geo.dft.size());
  read(in, geo.g0.data(), geo.g0.size());
  const int precision = std::atoi(argv[2]);
  const double tol = std::atof(argv[3]);
  try {
    return model.complex_m ? run<std::complex<double>>(in, geo, model, n_meas, n_g4, precision, tol)
                           : run<double>(in, geo, model, n_meas, n_g4, precision, tol);
  }
  catch (const std::exception& e) {
    return fail(e.what());
  }
}
