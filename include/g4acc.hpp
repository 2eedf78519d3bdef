// C++ host mirror of DCA++'s TpAccumulator over the g4acc C ABI (g4acc.h).
//
// Drop-in for accumulator::TpAccumulator<Parameters, DistType::NONE, DeviceType> as the CT-AUX
// and CT-INT accumulators use it (ctaux_accumulator.hpp:69,278,303,339,394,522,544;
// ctint_accumulator.hpp:155,175,193,245,258,270,289): same method names, the same argument
// meaning, and the reference's exception types (std::logic_error, std::bad_alloc,
// std::runtime_error) rethrown from the ABI's status codes.
//
// Template arguments are duck-typed on the reference's types:
//   Matrix         nrRows(), leadingDimension(), ptr()           (dca::linalg::Matrix<Real, ...>)
//   Configuration  size(), operator[](i).get_tau(),
//                  operator[](i).get_right_band(), .get_right_site(),
//                  operator[](i).get_left_band(), .get_left_site()    (CT-AUX vertex_singleton,
//                  ctaux/structs/vertex_singleton.hpp:84-95; CT-INT sector entries)
// The orbital of a vertex is taken from its right band and site, as CachedNdftBase does for
// density-density interactions (tp/ndft/cached_ndft_base.hpp:97-103). A vertex whose left and
// right orbitals differ (a non-density-density interaction, which the reference handles with a
// second sorted side, cached_ndft_base.hpp:72-78) raises std::logic_error instead of giving a
// wrong G4, unless the accumulator was built for such a model (g4acc_model::non_density_density),
// which sorts the columns of M by the left labels. Complex M (Matrix<std::complex<double>>) and a
// complex sign need g4acc_model::complex_m; a model that is not spin symmetric sets
// g4acc_model::spin_symmetric = 0 (tp_accumulator_cpu.hpp:668-757).
#ifndef G4ACC_HPP
#define G4ACC_HPP

#include <algorithm>
#include <array>
#include <complex>
#include <cstddef>
#include <new>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "g4acc.h"

namespace g4acc {

inline void throw_on_error(int status, const char* message) {
  switch (status) {
    case G4ACC_OK:
      return;
    case G4ACC_OUT_OF_MEMORY:
      throw std::bad_alloc();
    case G4ACC_DEVICE_ERROR:
      throw std::runtime_error(message ? message : "g4acc device error");
    default:
      throw std::logic_error(message ? message : "g4acc error");
  }
}

// Vertex times and orbitals of one spin sector's configuration, as the accumulator reads them
// (CachedNdftBase::sortConfiguration, tp/ndft/cached_ndft_base.hpp:89-129, density-density:
// both sides use the right band and site). Throws std::logic_error for a vertex whose left and
// right band or site differ.
template <class Configuration>
void extract_configuration(const Configuration& conf, std::vector<double>& tau, std::vector<int>& site,
                           std::vector<int>& band) {
  const int n = static_cast<int>(conf.size());
  tau.resize(n);
  site.resize(n);
  band.resize(n);
  for (int i = 0; i < n; ++i) {
    const auto& v = conf[i];
    if (v.get_left_band() != v.get_right_band() || v.get_left_site() != v.get_right_site())
      throw std::logic_error(
          "g4acc: vertex " + std::to_string(i) +
          " has different left and right orbitals; only density-density vertices are supported");
    tau[i] = v.get_tau();
    band[i] = v.get_right_band();
    site[i] = v.get_right_site();
  }
}

// Both label sets of a configuration (non-density-density vertices): right labels order the rows
// of M, left labels its columns (cached_ndft_base.hpp:97-103).
template <class Configuration>
void extract_configuration_general(const Configuration& conf, std::vector<double>& tau,
                                   std::vector<int>& site, std::vector<int>& band,
                                   std::vector<int>& site_l, std::vector<int>& band_l) {
  const int n = static_cast<int>(conf.size());
  tau.resize(n);
  site.resize(n);
  band.resize(n);
  site_l.resize(n);
  band_l.resize(n);
  for (int i = 0; i < n; ++i) {
    const auto& v = conf[i];
    tau[i] = v.get_tau();
    band[i] = v.get_right_band();
    site[i] = v.get_right_site();
    band_l[i] = v.get_left_band();
    site_l[i] = v.get_left_site();
  }
}

template <class T>
struct is_complex : std::false_type {};
template <class T>
struct is_complex<std::complex<T>> : std::true_type {};

// G0 over the EXTENDED vertex frequencies, as TpAccumulatorBase::initializeG0 slices the
// cluster-excluded single-particle G0 (tp_accumulator_base.hpp:210-221), in the (b1, b2, s, k, wE)
// layout of g4acc_geometry::g0. g0_sp(b1, s1, b2, s2, k, w) reads the single-particle function
// over its n_sp_freq frequencies. Throws std::logic_error when there are fewer single-particle
// than extended vertex frequencies, as the reference constructor does (:197-198).
template <class G0Function>
std::vector<std::complex<double>> slice_g0(const G0Function& g0_sp, int n_bands, int n_sites,
                                           int n_sp_freq, int n_ext_freq) {
  if (n_sp_freq < n_ext_freq)
    throw std::logic_error("The number of single particle frequencies is too small.");
  const int off = (n_sp_freq - n_ext_freq) / 2;
  std::vector<std::complex<double>> g0;
  g0.reserve(static_cast<std::size_t>(n_bands) * n_bands * 2 * n_sites * n_ext_freq);
  for (int w = 0; w < n_ext_freq; ++w)
    for (int k = 0; k < n_sites; ++k)
      for (int s = 0; s < 2; ++s)
        for (int b2 = 0; b2 < n_bands; ++b2)
          for (int b1 = 0; b1 < n_bands; ++b1)
            g0.push_back(g0_sp(b1, s, b2, s, k, w + off));
  return g0;
}

// Owning version of g4acc_geometry (everything the reference constructor reads from
// Parameters, the static domains and G0).
struct Geometry {
  int n_bands = 1;
  int n_sites = 1;
  double beta = 1.;
  int n_freq = 2;                         // 2 * four-point-fermionic-frequencies
  std::vector<int> w_ex{0};               // FrequencyExchangeDomain::get_elements()
  std::vector<int> k_ex{0};               // MomentumExchangeDomain::get_elements()
  std::vector<int> channels;              // sorted, unique FourPointType values
  std::vector<int> k_add, k_sub;          // [i + Nc j]
  std::vector<std::complex<double>> dft;  // [k + Nc r] = exp(i k.r)
  std::vector<std::complex<double>> band_phase;  // [b + nb k], empty: no band displacement
  std::vector<std::complex<double>> g0;   // (b1, b2, s, k, wE)

  g4acc_geometry view() const {
    g4acc_geometry g{};
    g.n_bands = n_bands;
    g.n_sites = n_sites;
    g.beta = beta;
    g.n_freq = n_freq;
    g.n_w_ex = static_cast<int>(w_ex.size());
    g.w_ex = w_ex.data();
    g.n_k_ex = static_cast<int>(k_ex.size());
    g.k_ex = k_ex.data();
    g.n_channels = static_cast<int>(channels.size());
    g.channels = channels.data();
    g.k_add = k_add.data();
    g.k_sub = k_sub.data();
    g.dft = reinterpret_cast<const double*>(dft.data());
    g.band_phase = band_phase.empty() ? nullptr : reinterpret_cast<const double*>(band_phase.data());
    g.g0 = reinterpret_cast<const double*>(g0.data());
    return g;
  }
};

class TpAccumulator {
public:
  using TpComplex = std::complex<double>;
  using TpGreensFunction = std::vector<TpComplex>;  // reference order, first index fastest

  // queued = true: accumulate() hands each measurement to the handle's queue, which runs as one
  // batch of options.max_walkers measurements (g4acc_accumulate_queued).
  // model: the reference's compile-time model properties (spin symmetry, complex M, non-density-
  // density vertices); the default is a real, spin-symmetric, density-density model. The queue
  // serves the default model only.
  TpAccumulator(const Geometry& geometry, const g4acc_options& options, bool queued = false,
                g4acc_model model = g4acc_model{1, 0, 0})
      : geometry_(geometry), queued_(queued), model_(model) {
    const g4acc_geometry g = geometry_.view();
    throw_on_error(g4acc_create_model(&g, &options, &model_, &handle_), g4acc_last_error(nullptr));
  }
  ~TpAccumulator() {
    if (handle_)
      g4acc_destroy(handle_);
  }
  TpAccumulator(const TpAccumulator&) = delete;
  TpAccumulator& operator=(const TpAccumulator&) = delete;
  TpAccumulator(TpAccumulator&& other) noexcept
      : geometry_(std::move(other.geometry_)),
        queued_(other.queued_),
        model_(other.model_),
        handle_(std::exchange(other.handle_, nullptr)) {}

  // Zeroes G4 (tp_accumulator_cpu.hpp:224-229). The loop index is accepted for signature parity.
  void resetAccumulation(unsigned int /*dca_loop*/ = 0) { check(g4acc_reset(handle_)); }

  // accumulate(M_pair, configs, sign) (tp_accumulator_cpu.hpp:107-109). `memory` says where the
  // M matrices live; the configuration is always read from the host objects. Returns the
  // algorithmic flop count of the measurement.
  template <class MatrixPair, class Configuration, class Sign>
  double accumulate(const MatrixPair& M_pair, const std::array<Configuration, 2>& configs,
                    const Sign sign, int memory = G4ACC_HOST) {
    using Scalar = std::remove_cv_t<std::remove_pointer_t<decltype(M_pair[0].ptr())>>;
    if constexpr (is_complex<Scalar>::value || is_complex<Sign>::value) {
      return accumulate_general(M_pair, configs, sign, memory);
    }
    else {
      if (model_.non_density_density || model_.complex_m || !model_.spin_symmetric)
        return accumulate_general(M_pair, configs, sign, memory);
      return accumulate_real(M_pair, configs, sign, memory);
    }
  }

 private:
  // The density-density, spin-symmetric, real model (g4acc_accumulate / g4acc_accumulate_queued).
  template <class MatrixPair, class Configuration, class Sign>
  double accumulate_real(const MatrixPair& M_pair, const std::array<Configuration, 2>& configs,
                         const Sign sign, int memory) {
    g4acc_walker w{};
    for (int s = 0; s < 2; ++s) {
      const auto& conf = configs[s];
      const int n = static_cast<int>(conf.size());
      extract_configuration(conf, tau_[s], site_[s], band_[s]);
      w.n[s] = n;
      w.m[s] = n ? M_pair[s].ptr() : nullptr;
      w.ld[s] = n ? M_pair[s].leadingDimension() : 1;
      w.tau[s] = tau_[s].data();
      w.site[s] = site_[s].data();
      w.band[s] = band_[s].data();
    }
    w.sign = static_cast<double>(sign);
    const int kind = memory == G4ACC_HOST ? G4ACC_HOST : G4ACC_DEVICE_M;
    check(queued_ ? g4acc_accumulate_queued(handle_, &w, kind) : g4acc_accumulate(handle_, &w, kind));
    return flops(w.n[0], w.n[1]);
  }

 public:
  // accumulate for the general model (g4acc_accumulate_general): complex M and sign, and the
  // vertices' left labels when the model is non-density-density.
  template <class MatrixPair, class Configuration, class Sign>
  double accumulate_general(const MatrixPair& M_pair, const std::array<Configuration, 2>& configs,
                            const Sign sign, int memory = G4ACC_HOST) {
    using Scalar = std::remove_cv_t<std::remove_pointer_t<decltype(M_pair[0].ptr())>>;
    if (is_complex<Scalar>::value && !model_.complex_m)
      throw std::logic_error("g4acc: complex M needs an accumulator built with complex_m");
    g4acc_walker_general w{};
    for (int s = 0; s < 2; ++s) {
      const auto& conf = configs[s];
      const int n = static_cast<int>(conf.size());
      if (model_.non_density_density)
        extract_configuration_general(conf, tau_[s], site_[s], band_[s], site_l_[s], band_l_[s]);
      else
        extract_configuration(conf, tau_[s], site_[s], band_[s]);
      w.n[s] = n;
      w.ld[s] = n ? M_pair[s].leadingDimension() : 1;
      bool widened = false;
      if constexpr (!is_complex<Scalar>::value) {
        if (model_.complex_m && n) {
          // Real M on a complex handle: widened into a complex copy (column-major, leading
          // dimension n), on the host only.
          if (memory != G4ACC_HOST)
            throw std::logic_error("g4acc: a complex handle takes real device M only as complex entries");
          mc_[s].assign(2 * static_cast<std::size_t>(n) * n, 0.);
          const Scalar* m = M_pair[s].ptr();
          for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
              mc_[s][2 * (i + static_cast<std::size_t>(n) * j)] =
                  static_cast<double>(m[i + static_cast<std::size_t>(w.ld[s]) * j]);
          w.m[s] = mc_[s].data();
          w.ld[s] = n;
          widened = true;
        }
      }
      if (!widened)
        w.m[s] = n ? reinterpret_cast<const double*>(M_pair[s].ptr()) : nullptr;
      w.tau[s] = tau_[s].data();
      w.site[s] = site_[s].data();
      w.band[s] = band_[s].data();
      w.site_l[s] = model_.non_density_density ? site_l_[s].data() : nullptr;
      w.band_l[s] = model_.non_density_density ? band_l_[s].data() : nullptr;
    }
    const std::complex<double> z(sign);
    w.sign[0] = z.real();
    w.sign[1] = z.imag();
    const int kind = memory == G4ACC_HOST ? G4ACC_HOST : G4ACC_DEVICE_M;
    check(g4acc_accumulate_general(handle_, &w, 1, kind));
    return flops(w.n[0], w.n[1]);
  }

  // Runs the queue and waits for all work (tp_accumulator_gpu.hpp:471-484).
  void finalize() { check(g4acc_finalize(handle_)); }
  // After this returns the M matrices passed to accumulate may be overwritten
  // (ctaux_accumulator.hpp:394). Queued measurements keep waiting for their batch.
  void synchronizeCopy() { check(g4acc_sync_copy(handle_)); }
  template <class T>
  void syncStreams(const T&) {
    check(g4acc_sync(handle_));
  }

  // get_G4() (tp_accumulator_cpu.hpp:871-877): one vector per channel, in the reference order.
  const std::vector<TpGreensFunction>& get_G4() {
    const std::size_t n = g4acc_g4_size(handle_);
    G4_.assign(geometry_.channels.size(), TpGreensFunction(n));
    for (std::size_t c = 0; c < G4_.size(); ++c)
      check(g4acc_get_g4(handle_, static_cast<int>(c), reinterpret_cast<double*>(G4_[c].data()), n));
    return G4_;
  }

  // The solvers' normalisation of G4 (ctaux_cluster_solver.hpp:523-527): get_G4() divided by
  // beta^2 * accumulated_sign, in FP64.
  std::vector<TpGreensFunction> get_G4_normalized(double accumulated_sign) {
    const std::size_t n = g4acc_g4_size(handle_);
    std::vector<TpGreensFunction> out(geometry_.channels.size(), TpGreensFunction(n));
    for (std::size_t c = 0; c < out.size(); ++c)
      check(g4acc_get_g4_normalized(handle_, static_cast<int>(c), accumulated_sign,
                                    reinterpret_cast<double*>(out[c].data()), n));
    return out;
  }

  // sumTo(other) (tp_accumulator_cpu.hpp:895-904): adds into other and empties this G4.
  void sumTo(TpAccumulator& other) { check(g4acc_sum_to(handle_, other.handle_)); }

  void* get_stream() const { return g4acc_stream(handle_); }
  std::size_t deviceFingerprint() const { return g4acc_bytes(handle_); }
  g4acc_handle* handle() const { return handle_; }

private:
  void check(int status) const { throw_on_error(status, g4acc_last_error(handle_)); }

  // M -> G: 8 nwE N^2 + 8 Nc D N + 8 D N nwE + 8 Nc D^2 per spin; G4: 8 flops per product term.
  double flops(int n0, int n1) const {
    int ext = 0;
    for (int w : geometry_.w_ex)
      ext = std::max(ext, w < 0 ? -w : w);
    const double nwE = geometry_.n_freq + 2 * ext, nc = geometry_.n_sites, nb = geometry_.n_bands;
    const double D = nb * nc * nwE;
    double f = 0;
    for (int n : {n0, n1})
      f += 8 * nwE * n * n + 8 * nc * D * n + 8 * D * n * nwE + 8 * nc * D * D;
    const double g4 = nb * nb * nb * nb * nc * nc * geometry_.n_freq * geometry_.n_freq *
                      geometry_.k_ex.size() * geometry_.w_ex.size();
    return f + 8. * 6. * g4 * geometry_.channels.size();
  }

  Geometry geometry_;
  bool queued_ = false;
  g4acc_model model_{1, 0, 0};
  g4acc_handle* handle_ = nullptr;
  std::array<std::vector<double>, 2> tau_, mc_;
  std::array<std::vector<int>, 2> site_, band_, site_l_, band_l_;
  std::vector<TpGreensFunction> G4_;
};

}  // namespace g4acc

#endif  // G4ACC_HPP
