// Drives the C++ TpAccumulator mirror (include/g4acc.hpp) the way the CT-AUX accumulator drives
// the reference class (ctaux_accumulator.hpp:278,303,394,522,544), on a case file written by
// tests/test_cpp_wrapper.py: geometry, measurements and the FP64 oracle's G4. Mock Matrix and
// Vertex types expose the reference's accessor names.
//
// usage: tp_accumulator_wrapper_test <case file> <precision 0|1> <l_inf tolerance>
#include <cmath>
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
  int band, site;
  double get_tau() const { return tau; }
  int get_right_band() const { return band; }
  int get_right_site() const { return site; }
  int get_left_band() const { return band; }
  int get_left_site() const { return site; }
};

struct Matrix {
  int n = 0;
  std::vector<double> v;
  const double* ptr() const { return v.data(); }
  int leadingDimension() const { return n > 0 ? n : 1; }
  int nrRows() const { return n; }
};

struct Measurement {
  std::array<Matrix, 2> M;
  std::array<std::vector<Vertex>, 2> config;
  signed char sign;
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

}  // namespace

int main(int argc, char** argv) {
  if (argc != 4) {
    std::cerr << "usage: " << argv[0] << " <case> <precision> <tol>\n";
    return 2;
  }
  std::ifstream in(argv[1], std::ios::binary);
  int hdr[8];
  read(in, hdr, 8);
  const int nb = hdr[0], nc = hdr[1], n_freq = hdr[2], n_wex = hdr[3], n_kex = hdr[4],
            n_ch = hdr[5], n_meas = hdr[6], n_g4 = hdr[7];
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
  std::vector<Measurement> ms(n_meas);
  int n_max = 1;
  for (auto& m : ms) {
    int n[2];
    double sign;
    read(in, n, 2);
    read(in, &sign, 1);
    m.sign = static_cast<signed char>(sign);
    for (int s = 0; s < 2; ++s) {
      n_max = std::max(n_max, n[s]);
      m.M[s].n = n[s];
      m.M[s].v.resize(static_cast<std::size_t>(n[s]) * n[s]);
      read(in, m.M[s].v.data(), m.M[s].v.size());
      std::vector<double> tau(n[s]);
      std::vector<int> site(n[s]), band(n[s]);
      read(in, tau.data(), n[s]);
      read(in, site.data(), n[s]);
      read(in, band.data(), n[s]);
      for (int i = 0; i < n[s]; ++i)
        m.config[s].push_back(Vertex{tau[i], band[i], site[i]});
    }
  }
  std::vector<std::vector<std::complex<double>>> expected(n_ch, std::vector<std::complex<double>>(n_g4));
  for (auto& e : expected)
    read(in, e.data(), e.size());

  g4acc_options opt{};
  opt.precision = std::atoi(argv[2]);
  opt.sum_mode = G4ACC_SUM_FP64;
  opt.max_walkers = 1;
  opt.max_vertices = n_max;
  opt.device = 0;
  const double tol = std::atof(argv[3]);

  g4acc::TpAccumulator acc(geo, opt), other(geo, opt);
  // accumulate before resetAccumulation: std::logic_error (tp_accumulator_gpu.hpp:329-330).
  try {
    acc.accumulate(ms[0].M, ms[0].config, ms[0].sign);
    return fail("accumulate before resetAccumulation did not throw");
  }
  catch (const std::logic_error&) {
  }
  acc.resetAccumulation(0);
  other.resetAccumulation(0);
  // Half of the measurements on each accumulator, merged with sumTo as the threaded solver does
  // (stdthread_qmci_cluster_solver.hpp:571-572).
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
    const double linf = num / den;
    std::printf("channel %d l_inf %.3e\n", geo.channels[c], linf);
    if (!(linf < tol))
      return fail("G4 mismatch");
  }
  // The same measurements through the queued drop-in: one accumulate() per measurement, run in
  // batches of three inside the handle.
  {
    g4acc_options qopt = opt;
    qopt.max_walkers = 3;
    g4acc::TpAccumulator queued(geo, qopt, true);
    queued.resetAccumulation(0);
    for (int i = 0; i < n_meas; ++i) {
      queued.accumulate(ms[i].M, ms[i].config, ms[i].sign);
      queued.synchronizeCopy();
    }
    queued.finalize();
    const auto& Q = queued.get_G4();
    for (int c = 0; c < n_ch; ++c) {
      double num = 0, den = 0;
      for (int i = 0; i < n_g4; ++i) {
        num = std::max(num, std::abs(Q[c][i] - expected[c][i]));
        den = std::max(den, std::abs(expected[c][i]));
      }
      std::printf("queued channel %d l_inf %.3e\n", geo.channels[c], num / den);
      if (!(num / den < tol))
        return fail("queued G4 mismatch");
    }
  }
  // get_G4 on the emptied source: std::logic_error (tp_accumulator_cpu.hpp:874-875).
  try {
    other.get_G4();
    return fail("get_G4 after sumTo did not throw");
  }
  catch (const std::logic_error&) {
  }
  // An unimplemented channel: std::logic_error (tp_accumulator_cpu.hpp:663-664).
  try {
    g4acc::Geometry bad = geo;
    bad.channels = {5};
    g4acc::TpAccumulator x(bad, opt);
    return fail("unsupported channel did not throw");
  }
  catch (const std::logic_error&) {
  }
  std::cout << "OK" << std::endl;
  return 0;
}
