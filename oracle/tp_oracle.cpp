// FP64 CPU oracle of the DCA++ two-particle accumulator.
//
// TEST INFRASTRUCTURE ONLY: the checker for the HIP product path and the CPU baseline of the
// benchmark. Nothing in the product loads it.
//
// Each function restates the reference file:line given in its comment, in the same loop order,
// in double precision (the reference's default TpAccumulationPrecision). Its complex GEMMs use
// its own loop by default (the checker needs no external library); tpo_use_blas() switches them
// to BLAS zgemm_ from a shared library loaded at run time (the MKL runtime of the image), which
// is what the reference calls (tp/ndft/cached_ndft_cpu.hpp:214,217; space_transform_2D.hpp:80-81),
// for the CPU baseline. The reference's "cpu w_ex" stdout print inside the w_ex loop
// (tp_accumulator_cpu.hpp:460,500) is left out; everything else, including the per-call
// allocation of M_r_r_w_w (tp_accumulator_cpu.hpp:261-262), is kept.
#include "tp_oracle.h"

#include <dlfcn.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

using cd = std::complex<double>;

// BLAS zgemm_ (Fortran ABI, include/dca/linalg/blas/blas3.hpp:29), when loaded.
using zgemm_fn = void (*)(const char*, const char*, const int*, const int*, const int*, const cd*,
                          const cd*, const int*, const cd*, const int*, const cd*, cd*, const int*);
zgemm_fn g_blas_zgemm = nullptr;
std::string g_blas_name = "own loop zgemm";

// Column-major complex GEMM, C = alpha * A * op(B) + beta * C, op(B) = B ('N') or B^H ('C').
// Stands in for BLAS zgemm_ as called by linalg::matrixop::gemm (include/dca/linalg/matrixop.hpp).
void zgemm(char tb, int m, int n, int k, cd alpha, const cd* A, int lda, const cd* B, int ldb,
           cd beta, cd* C, int ldc) {
  if (g_blas_zgemm) {
    const char ta = 'N';
    g_blas_zgemm(&ta, &tb, &m, &n, &k, &alpha, A, &lda, B, &ldb, &beta, C, &ldc);
    return;
  }
  for (int j = 0; j < n; ++j) {
    cd* c = C + static_cast<size_t>(ldc) * j;
    if (beta == cd(0))
      for (int i = 0; i < m; ++i)
        c[i] = 0;
    else if (beta != cd(1))
      for (int i = 0; i < m; ++i)
        c[i] *= beta;
    for (int l = 0; l < k; ++l) {
      const cd b = tb == 'N' ? B[l + static_cast<size_t>(ldb) * j]
                             : std::conj(B[j + static_cast<size_t>(ldb) * l]);
      const cd ab = alpha * b;
      const cd* a = A + static_cast<size_t>(lda) * l;
      for (int i = 0; i < m; ++i)
        c[i] += a[i] * ab;
    }
  }
}

struct CMat {
  int rows = 0, cols = 0;
  std::vector<cd> v;
  void resize(int r, int c) {
    rows = r;
    cols = c;
    v.assign(static_cast<size_t>(r) * c, cd(0));
  }
  cd& operator()(int i, int j) { return v[i + static_cast<size_t>(rows) * j]; }
  const cd& operator()(int i, int j) const { return v[i + static_cast<size_t>(rows) * j]; }
  cd* data() { return v.data(); }
};

// details::Triple (tp/ndft/triple.hpp:23-30): ordered by (orbital, idx).
struct Triple {
  int orbital;
  int idx;
  double tau;
  bool operator<(const Triple& rhs) const {
    return orbital == rhs.orbital ? idx < rhs.idx : orbital < rhs.orbital;
  }
};

enum Channel { PH_TRANSVERSE = 1, PH_MAGNETIC = 2, PH_CHARGE = 3, PP_UP_DOWN = 4 };

}  // namespace

struct tpo_acc {
  int nb, nc, nw, nwE, ext, nq, nv;
  double beta;
  std::vector<int> wex, kex, channels, kadd, ksub;
  std::vector<cd> dft, band_phase;
  std::vector<double> w_ext;  // EXTENDED vertex frequency values
  std::vector<cd> G0;         // (b1, b2, s, k, wE)
  std::vector<cd> G;          // (b1, b2, s, k1, k2, w1E, w2E)
  std::vector<cd> Mrrww;      // last M_r_r_w_w, (r1, r2, b1, b2, s, w1E, w2E)
  std::vector<std::vector<cd>> G4;
  CMat Ga, Gb, G0M;           // band-block work spaces (tp_accumulator_cpu.hpp:208)
  bool keep_mrrww = true;     // copy M_r_r_w_w for tpo_get_Mrrww (a test hook; off for timing)
  // Parameters::model_type::lattice_type::spin_symmetric (tp_accumulator_base.hpp:167).
  bool spin_symmetric = true;

  size_t g0i(int b1, int b2, int s, int k, int w) const {
    return b1 + nb * (b2 + nb * (s + 2 * (k + static_cast<size_t>(nc) * w)));
  }
  size_t gi(int b1, int b2, int s, int k1, int k2, int w1, int w2) const {
    return b1 + nb * (b2 + nb * (s + 2 * (k1 + nc * (k2 + static_cast<size_t>(nc) *
                                                          (w1 + static_cast<size_t>(nwE) * w2)))));
  }
  size_t mi(int r1, int r2, int b1, int b2, int s, int w1, int w2) const {
    return r1 + nc * (r2 + nc * (b1 + nb * (b2 + nb * (s + 2 * (w1 + static_cast<size_t>(nwE) * w2)))));
  }
  size_t g4i(int b1, int b2, int b3, int b4, int k1, int w1, int k2, int w2, int iq, int iv) const {
    size_t idx = iv;
    idx = idx * nq + iq;
    idx = idx * nw + w2;
    idx = idx * nc + k2;
    idx = idx * nw + w1;
    idx = idx * nc + k1;
    idx = idx * nb + b4;
    idx = idx * nb + b3;
    idx = idx * nb + b2;
    idx = idx * nb + b1;
    return idx;
  }
  size_t g4_size() const {
    return static_cast<size_t>(nb) * nb * nb * nb * nc * nw * nc * nw * nq * nv;
  }

  // CachedNdft<CPU>::execute (tp/ndft/cached_ndft_cpu.hpp:106-140) with sortConfiguration
  // (tp/ndft/cached_ndft_base.hpp:89-129), computeT (:142-158), computeMMatrix (:160-174),
  // computeTSubmatrices (:176-197), executeTrimmedFT (:200-222), copyPartialResult (:239-252).
  // The rows of M are sorted by the vertices' right orbitals and the columns by their left ones
  // ("M is an inverse matrix", cached_ndft_base.hpp:97-103); site_l / band_l == nullptr is the
  // density-density case, where both sides are the same sort (:72-78). MS: double or complex.
  template <typename MS>
  double ndft(int spin, int n_v, const MS* m, int ld, const double* tau, const int* site,
              const int* band, const int* site_l, const int* band_l, std::vector<cd>& out) {
    const int n_orb = nb * nc;
    const bool nd = site_l != nullptr;
    std::vector<Triple> cfg_l(n_v), cfg_r;
    for (int l = 0; l < n_v; ++l)
      cfg_l[l] = Triple{band[l] + nb * site[l], l, tau[l]};
    std::sort(cfg_l.begin(), cfg_l.end());
    if (nd) {
      cfg_r.resize(n_v);
      for (int l = 0; l < n_v; ++l)
        cfg_r[l] = Triple{band_l[l] + nb * site_l[l], l, tau[l]};
      std::sort(cfg_r.begin(), cfg_r.end());
    }
    const std::vector<Triple>& cfg_right = nd ? cfg_r : cfg_l;
    auto ranges = [&](const std::vector<Triple>& cfg, std::vector<int>& start, std::vector<int>& end) {
      start.resize(n_orb);
      end.resize(n_orb);
      for (int orb = 0; orb < n_orb; ++orb) {
        Triple t{orb, 0, 0};
        start[orb] = std::lower_bound(cfg.begin(), cfg.end(), t) - cfg.begin();
        t.idx = n_v;
        end[orb] = std::upper_bound(cfg.begin(), cfg.end(), t) - cfg.begin();
      }
    };
    std::vector<int> start_l, end_l, start_r, end_r;
    ranges(cfg_l, start_l, end_l);
    ranges(cfg_right, start_r, end_r);

    CMat T;
    T.resize(nwE, n_v);
    for (int j = 0; j < n_v; ++j)
      for (int i = 0; i < nwE; ++i) {
        const double x = tau[j] * w_ext[i];
        T(i, j) = cd(std::cos(x), std::sin(x));
      }

    double flops = 0;
    CMat Mij, Tl, Tr, TlM, TlMTr;
    for (int orb_j = 0; orb_j < n_orb; ++orb_j) {
      const int n_j = end_r[orb_j] - start_r[orb_j];
      for (int orb_i = 0; orb_i < n_orb; ++orb_i) {
        const int n_i = end_l[orb_i] - start_l[orb_i];
        const int b1 = orb_i % nb, r1 = orb_i / nb;
        const int b2 = orb_j % nb, r2 = orb_j / nb;
        if (n_i > 0 && n_j > 0) {
          Mij.resize(n_i, n_j);
          for (int lj = start_r[orb_j]; lj < end_r[orb_j]; ++lj)
            for (int li = start_l[orb_i]; li < end_l[orb_i]; ++li)
              Mij(li - start_l[orb_i], lj - start_r[orb_j]) =
                  cd(m[cfg_l[li].idx + static_cast<size_t>(ld) * cfg_right[lj].idx]);
          Tl.resize(nwE, n_i);
          for (int li = start_l[orb_i]; li < end_l[orb_i]; ++li)
            std::copy_n(&T(0, cfg_l[li].idx), nwE, &Tl(0, li - start_l[orb_i]));
          Tr.resize(nwE, n_j);
          for (int lj = start_r[orb_j]; lj < end_r[orb_j]; ++lj)
            std::copy_n(&T(0, cfg_right[lj].idx), nwE, &Tr(0, lj - start_r[orb_j]));

          TlM.resize(nwE, n_j);
          zgemm('N', nwE, n_j, n_i, 1., Tl.data(), nwE, Mij.data(), n_i, 0., TlM.data(), nwE);
          flops += 4. * nwE * n_i * n_j;
          TlMTr.resize(nwE, nwE);
          zgemm('C', nwE, nwE, n_j, 1., TlM.data(), nwE, Tr.data(), nwE, 0., TlMTr.data(), nwE);
          flops += 8. * nwE * n_j * nwE;

          for (int w2 = 0; w2 < nwE; ++w2)
            for (int w1 = 0; w1 < nwE; ++w1)
              out[mi(r1, r2, b1, b2, spin, w1, w2)] = TlMTr(w1, w2);
        }
        else {
          for (int w2 = 0; w2 < nwE; ++w2)
            for (int w1 = 0; w1 < nwE; ++w1)
              out[mi(r1, r2, b1, b2, spin, w1, w2)] = 0;
        }
      }
    }
    return flops;
  }

  // SpaceTransform2D::execute (math/function_transform/special_transforms/space_transform_2D.hpp:60-91).
  void space_transform(std::vector<cd>& f_in) {
    const cd norm(1. / nc);
    std::vector<cd> tmp(static_cast<size_t>(nc) * nc);
    for (int w2 = 0; w2 < nwE; ++w2)
      for (int w1 = 0; w1 < nwE; ++w1)
        for (int s = 0; s < 2; ++s)
          for (int b2 = 0; b2 < nb; ++b2)
            for (int b1 = 0; b1 < nb; ++b1) {
              cd* f_r_r = &f_in[mi(0, 0, b1, b2, s, w1, w2)];
              zgemm('N', nc, nc, nc, 1., dft.data(), nc, f_r_r, nc, 0., tmp.data(), nc);
              zgemm('C', nc, nc, nc, norm, tmp.data(), nc, dft.data(), nc, 0., f_r_r, nc);
              for (int k2 = 0; k2 < nc; ++k2)
                for (int k1 = 0; k1 < nc; ++k1)
                  G[gi(b1, b2, s, k1, k2, w1, w2)] = f_r_r[k1 + nc * k2] * band_phase[b1 + nb * k1] *
                                                     std::conj(band_phase[b2 + nb * k2]);
            }
  }

  // computeGActl / computeGSingleband / computeGMultiband / matrixOperationsGMultiband
  // (tp_accumulator_cpu.hpp:288-375).
  void compute_G() {
    for (int w2 = 0; w2 < nwE; ++w2)
      for (int w1 = 0; w1 < nwE; ++w1)
        for (int k2 = 0; k2 < nc; ++k2)
          for (int k1 = 0; k1 < nc; ++k1)
            for (int s = 0; s < 2; ++s) {
              if (nb == 1) {
                const cd G0_w1 = G0[g0i(0, 0, s, k1, w1)];
                const cd G0_w2 = G0[g0i(0, 0, s, k2, w2)];
                cd& g = G[gi(0, 0, s, k1, k2, w1, w2)];
                const cd M_val = g;
                if (k2 == k1 && w2 == w1)
                  g = -G0_w1 * M_val * G0_w2 + G0_w1 * beta;
                else
                  g = -G0_w1 * M_val * G0_w2;
              }
              else {
                const cd* G0_w1 = &G0[g0i(0, 0, s, k1, w1)];
                const cd* G0_w2 = &G0[g0i(0, 0, s, k2, w2)];
                cd* Mm = &G[gi(0, 0, s, k1, k2, w1, w2)];
                G0M.resize(nb, nb);
                zgemm('N', nb, nb, nb, 1., G0_w1, nb, Mm, nb, 0., G0M.data(), nb);
                zgemm('N', nb, nb, nb, -1., G0M.data(), nb, G0_w2, nb, 0., Mm, nb);
                if (G0_w1 == G0_w2)
                  for (int b2 = 0; b2 < nb; ++b2)
                    for (int b1 = 0; b1 < nb; ++b1)
                      Mm[b1 + nb * b2] += G0_w1[b1 + nb * b2] * beta;
              }
            }
  }

  // getGSingleband (tp_accumulator_cpu.hpp:377-386).
  cd getG1(int s, int k1, int k2, int w1, int w2) const {
    return G[gi(0, 0, s, k1, k2, w1 + ext, w2 + ext)];
  }

  // getGMultiband (tp_accumulator_cpu.hpp:388-405): a band-transposed read of G_.
  void getGM(int s, int k1, int k2, int w1, int w2, CMat& out, cd sign = 0) const {
    const cd* p = &G[gi(0, 0, s, k1, k2, w1 + ext, w2 + ext)];
    for (int b2 = 0; b2 < nb; ++b2)
      for (int b1 = 0; b1 < nb; ++b1)
        out(b1, b2) = sign * out(b1, b2) + p[b2 + b1 * nb];
  }

  // updateG4Atomic (tp_accumulator_cpu.hpp:759-810).
  void atomic_update(cd* g4, int s_a, int k1_a, int k2_a, int w1_a, int w2_a, int s_b, int k1_b,
                     int k2_b, int w1_b, int w2_b, cd alpha, bool cross_legs) {
    if (nb == 1) {
      const cd a = getG1(s_a, k1_a, k2_a, w1_a, w2_a);
      const cd b = getG1(s_b, k1_b, k2_b, w1_b, w2_b);
      *g4 += alpha * a * b;
      return;
    }
    getGM(s_a, k1_a, k2_a, w1_a, w2_a, Ga);
    getGM(s_b, k1_b, k2_b, w1_b, w2_b, Gb);
    for (int b4 = 0; b4 < nb; ++b4)
      for (int b3 = 0; b3 < nb; ++b3)
        for (int b2 = 0; b2 < nb; ++b2)
          for (int b1 = 0; b1 < nb; ++b1) {
            const cd prod = cross_legs ? Ga(b1, b4) * Gb(b2, b3) : Ga(b1, b3) * Gb(b2, b4);
            *g4 += alpha * prod;
            ++g4;
          }
  }

  // updateG4SpinDifference (tp_accumulator_cpu.hpp:812-869).
  void spin_difference_update(cd* g4, double sign, int k1_a, int k2_a, int w1_a, int w2_a,
                              int k1_b, int k2_b, int w1_b, int w2_b, cd alpha,
                              bool cross_legs) {
    if (nb == 1) {
      *g4 += alpha * (getG1(0, k1_a, k2_a, w1_a, w2_a) + cd(sign) * getG1(1, k1_a, k2_a, w1_a, w2_a)) *
             (getG1(0, k1_b, k2_b, w1_b, w2_b) + cd(sign) * getG1(1, k1_b, k2_b, w1_b, w2_b));
      return;
    }
    getGM(0, k1_a, k2_a, w1_a, w2_a, Ga);
    getGM(1, k1_a, k2_a, w1_a, w2_a, Ga, sign);
    getGM(0, k1_b, k2_b, w1_b, w2_b, Gb);
    getGM(1, k1_b, k2_b, w1_b, w2_b, Gb, sign);
    for (int b4 = 0; b4 < nb; ++b4)
      for (int b3 = 0; b3 < nb; ++b3)
        for (int b2 = 0; b2 < nb; ++b2)
          for (int b1 = 0; b1 < nb; ++b1) {
            *g4 += alpha * (cross_legs ? Ga(b1, b4) * Gb(b2, b3) : Ga(b2, b1) * Gb(b3, b4));
            ++g4;
          }
  }

  // updateG4, spin-symmetric branch (tp_accumulator_cpu.hpp:407-666); the factor is the sign
  // (SignType: +-1, or a complex phase for complex models, :433-438). Returns false for a channel
  // the reference does not implement (:663-664, :752-753).
  bool update_G4(int channel_id, cd factor, double& flops) {
    if (!spin_symmetric)
      return update_G4_no_spin(channel_id, factor, flops);
    const cd sign_over_2 = 0.5 * factor;
    auto momentum_sum = [&](int k, int q) { return kadd[k + nc * q]; };
    auto q_minus_k = [&](int k, int q) { return ksub[k + nc * q]; };
    auto w_plus_w_ex = [](int w, int w_ex) { return w + w_ex; };
    const int nw_c = nw;
    auto w_ex_minus_w = [nw_c](int w, int w_ex) { return w_ex + nw_c - 1 - w; };
    const double f_atomic = 3 * std::pow(nb, 4);
    const double f_spin_diff = f_atomic + 2 * std::pow(nb, 2);
    const double n_loops = double(nw) * nw * nc * nc;
    std::vector<cd>& g4 = G4[channel_id];
    Ga.resize(nb, nb);
    Gb.resize(nb, nb);

    switch (channels[channel_id]) {
      case PH_TRANSVERSE:
        for (int iv = 0; iv < nv; ++iv) {
          const int w_ex = wex[iv];
          for (int iq = 0; iq < nq; ++iq) {
            const int k_ex = kex[iq];
            for (int w2 = 0; w2 < nw; ++w2)
              for (int k2 = 0; k2 < nc; ++k2)
                for (int w1 = 0; w1 < nw; ++w1)
                  for (int k1 = 0; k1 < nc; ++k1)
                    for (int s = 0; s < 2; ++s) {
                      getGM(s, k1, k2, w1, w2, Ga);
                      getGM(!s, momentum_sum(k2, k_ex), momentum_sum(k1, k_ex),
                            w_plus_w_ex(w2, w_ex), w_plus_w_ex(w1, w_ex), Gb);
                      for (int b4 = 0; b4 < nb; ++b4)
                        for (int b3 = 0; b3 < nb; ++b3)
                          for (int b2 = 0; b2 < nb; ++b2)
                            for (int b1 = 0; b1 < nb; ++b1)
                              g4[g4i(b1, b2, b3, b4, k1, w1, k2, w2, iq, iv)] +=
                                  -sign_over_2 * Ga(b2, b4) * Gb(b3, b1);
                    }
          }
        }
        flops += n_loops * 2 * f_atomic;
        return true;

      case PH_MAGNETIC:
        for (int iv = 0; iv < nv; ++iv) {
          const int w_ex = wex[iv];
          for (int iq = 0; iq < nq; ++iq) {
            const int k_ex = kex[iq];
            for (int w2 = 0; w2 < nw; ++w2)
              for (int k2 = 0; k2 < nc; ++k2)
                for (int w1 = 0; w1 < nw; ++w1)
                  for (int k1 = 0; k1 < nc; ++k1) {
                    for (int s1 = 0; s1 < 2; ++s1)
                      for (int s2 = 0; s2 < 2; ++s2) {
                        const cd spin_factor = (s1 == s2) ? 1 : -1;
                        getGM(s1, k1, momentum_sum(k1, k_ex), w1, w_plus_w_ex(w1, w_ex), Ga);
                        getGM(s2, momentum_sum(k2, k_ex), k2, w_plus_w_ex(w2, w_ex), w2, Gb);
                        for (int b4 = 0; b4 < nb; ++b4)
                          for (int b3 = 0; b3 < nb; ++b3)
                            for (int b2 = 0; b2 < nb; ++b2)
                              for (int b1 = 0; b1 < nb; ++b1)
                                g4[g4i(b1, b2, b3, b4, k1, w1, k2, w2, iq, iv)] +=
                                    sign_over_2 * Ga(b2, b1) * Gb(b3, b4) * spin_factor;
                      }
                    for (int s = 0; s < 2; ++s) {
                      getGM(s, k1, k2, w1, w2, Ga);
                      getGM(s, momentum_sum(k2, k_ex), momentum_sum(k1, k_ex),
                            w_plus_w_ex(w2, w_ex), w_plus_w_ex(w1, w_ex), Gb);
                      for (int b4 = 0; b4 < nb; ++b4)
                        for (int b3 = 0; b3 < nb; ++b3)
                          for (int b2 = 0; b2 < nb; ++b2)
                            for (int b1 = 0; b1 < nb; ++b1)
                              g4[g4i(b1, b2, b3, b4, k1, w1, k2, w2, iq, iv)] +=
                                  -sign_over_2 * Ga(b3, b1) * Gb(b2, b4);
                    }
                  }
          }
        }
        flops += n_loops * (f_spin_diff + 2 * f_atomic);
        return true;

      case PH_CHARGE:
        for (int iv = 0; iv < nv; ++iv) {
          const int w_ex = wex[iv];
          for (int iq = 0; iq < nq; ++iq) {
            const int k_ex = kex[iq];
            for (int w2 = 0; w2 < nw; ++w2)
              for (int k2 = 0; k2 < nc; ++k2)
                for (int w1 = 0; w1 < nw; ++w1)
                  for (int k1 = 0; k1 < nc; ++k1) {
                    cd* p = &g4[g4i(0, 0, 0, 0, k1, w1, k2, w2, iq, iv)];
                    spin_difference_update(p, 1, k1, momentum_sum(k1, k_ex), w1,
                                           w_plus_w_ex(w1, w_ex), momentum_sum(k2, k_ex), k2,
                                           w_plus_w_ex(w2, w_ex), w2, sign_over_2, false);
                    for (int s = 0; s < 2; ++s)
                      atomic_update(p, s, k1, k2, w1, w2, s, momentum_sum(k2, k_ex),
                                    momentum_sum(k1, k_ex), w_plus_w_ex(w2, w_ex),
                                    w_plus_w_ex(w1, w_ex), -sign_over_2, true);
                  }
          }
        }
        flops += n_loops * (f_spin_diff + 2 * f_atomic);
        return true;

      case PP_UP_DOWN:
        for (int iv = 0; iv < nv; ++iv) {
          const int w_ex = wex[iv];
          for (int iq = 0; iq < nq; ++iq) {
            const int k_ex = kex[iq];
            for (int w2 = 0; w2 < nw; ++w2)
              for (int k2 = 0; k2 < nc; ++k2)
                for (int w1 = 0; w1 < nw; ++w1)
                  for (int k1 = 0; k1 < nc; ++k1)
                    for (int s = 0; s < 2; ++s) {
                      getGM(s, k1, k2, w1, w2, Ga);
                      getGM(!s, q_minus_k(k1, k_ex), q_minus_k(k2, k_ex), w_ex_minus_w(w1, w_ex),
                            w_ex_minus_w(w2, w_ex), Gb);
                      for (int b4 = 0; b4 < nb; ++b4)
                        for (int b3 = 0; b3 < nb; ++b3)
                          for (int b2 = 0; b2 < nb; ++b2)
                            for (int b1 = 0; b1 < nb; ++b1)
                              g4[g4i(b1, b2, b3, b4, k1, w1, k2, w2, iq, iv)] +=
                                  sign_over_2 * Ga(b1, b3) * Gb(b2, b4);
                    }
          }
        }
        flops += n_loops * 2 * f_atomic;
        return true;

      default:
        return false;
    }
  }

  // updateG4, branch of models that are not spin symmetric (tp_accumulator_cpu.hpp:668-757): the
  // spin is part of the band index, only spin sector 0 of G is read, and there is no 1/2.
  bool update_G4_no_spin(int channel_id, cd factor, double& flops) {
    auto q_minus_k = [&](int k, int q) { return ksub[k + nc * q]; };
    auto q_plus_k = [&](int k, int q) { return kadd[k + nc * q]; };
    auto w_plus_w_ex = [](int w, int w_ex) { return w + w_ex; };
    const int nw_c = nw;
    auto w_ex_minus_w = [nw_c](int w, int w_ex) { return w_ex + nw_c - 1 - w; };
    const double f_atomic = 3 * std::pow(nb, 4);
    const double n_loops = double(nw) * nw * nc * nc;
    std::vector<cd>& g4 = G4[channel_id];
    Ga.resize(nb, nb);
    Gb.resize(nb, nb);
    switch (channels[channel_id]) {
      case PP_UP_DOWN:
        for (int iv = 0; iv < nv; ++iv) {
          const int w_ex = wex[iv];
          for (int iq = 0; iq < nq; ++iq) {
            const int k_ex = kex[iq];
            for (int w2 = 0; w2 < nw; ++w2)
              for (int k2 = 0; k2 < nc; ++k2)
                for (int w1 = 0; w1 < nw; ++w1)
                  for (int k1 = 0; k1 < nc; ++k1) {
                    // G(k2, k1, b1, b3) * G(k_ex - k2, k_ex - k1, b2, b4)
                    getGM(0, k1, k2, w1, w2, Ga);
                    getGM(0, q_minus_k(k1, k_ex), q_minus_k(k2, k_ex), w_ex_minus_w(w1, w_ex),
                          w_ex_minus_w(w2, w_ex), Gb);
                    for (int b4 = 0; b4 < nb; ++b4)
                      for (int b3 = 0; b3 < nb; ++b3)
                        for (int b2 = 0; b2 < nb; ++b2)
                          for (int b1 = 0; b1 < nb; ++b1)
                            g4[g4i(b1, b2, b3, b4, k1, w1, k2, w2, iq, iv)] += factor * Ga(b1, b3) * Gb(b2, b4);
                    // -G(k1, k_ex - k2, b1, b4) * G(k_ex - k1, k2, b2, b3)
                    getGM(0, k1, q_minus_k(k2, k_ex), w1, w_ex_minus_w(w2, w_ex), Ga);
                    getGM(0, q_minus_k(k1, k_ex), k2, w_ex_minus_w(w1, w_ex), w2, Gb);
                    for (int b4 = 0; b4 < nb; ++b4)
                      for (int b3 = 0; b3 < nb; ++b3)
                        for (int b2 = 0; b2 < nb; ++b2)
                          for (int b1 = 0; b1 < nb; ++b1)
                            g4[g4i(b1, b2, b3, b4, k1, w1, k2, w2, iq, iv)] -= factor * Ga(b1, b4) * Gb(b2, b3);
                  }
          }
        }
        flops += n_loops * 4 * f_atomic;
        return true;

      case PH_MAGNETIC:
        for (int iv = 0; iv < nv; ++iv) {
          const int w_ex = wex[iv];
          for (int iq = 0; iq < nq; ++iq) {
            const int k_ex = kex[iq];
            for (int w2 = 0; w2 < nw; ++w2)
              for (int k2 = 0; k2 < nc; ++k2)
                for (int w1 = 0; w1 < nw; ++w1)
                  for (int k1 = 0; k1 < nc; ++k1) {
                    // G(k1, k1+k_ex, s2, s1) G(k2+k_ex, k2, s3, s4)
                    getGM(0, k1, q_plus_k(k1, k_ex), w1, w_plus_w_ex(w1, w_ex), Ga);
                    getGM(0, q_plus_k(k2, k_ex), k2, w_plus_w_ex(w2, w_ex), w2, Gb);
                    for (int b4 = 0; b4 < nb; ++b4)
                      for (int b3 = 0; b3 < nb; ++b3)
                        for (int b2 = 0; b2 < nb; ++b2)
                          for (int b1 = 0; b1 < nb; ++b1)
                            g4[g4i(b1, b2, b3, b4, k1, w1, k2, w2, iq, iv)] += factor * Ga(b2, b1) * Gb(b3, b4);
                    // -G(k2+k_ex, k1+k_ex, s3, s1) G(k1, k2, s2, s4)
                    getGM(0, q_plus_k(k2, k_ex), q_plus_k(k1, k_ex), w_plus_w_ex(w2, w_ex),
                          w_plus_w_ex(w1, w_ex), Ga);
                    getGM(0, k1, k2, w1, w2, Gb);
                    for (int b4 = 0; b4 < nb; ++b4)
                      for (int b3 = 0; b3 < nb; ++b3)
                        for (int b2 = 0; b2 < nb; ++b2)
                          for (int b1 = 0; b1 < nb; ++b1)
                            g4[g4i(b1, b2, b3, b4, k1, w1, k2, w2, iq, iv)] -= factor * Ga(b3, b1) * Gb(b2, b4);
                  }
          }
        }
        flops += n_loops * std::pow(nb, 4) * 2 * 3;
        return true;

      default:
        return false;
    }
  }

  // computeM (tp_accumulator_cpu.hpp:254-276) followed by computeG (:278-286).
  // m[s]: real M, or interleaved complex M when m_complex; site_l / band_l: the vertices' left
  // labels (nullptr: density-density vertices).
  double compute_M_and_G(const int n[2], const double* const m[2], bool m_complex, const int ld[2],
                         const double* const tau[2], const int* const site[2],
                         const int* const band[2], const int* const* site_l = nullptr,
                         const int* const* band_l = nullptr) {
    double flops = 0;
    std::vector<cd> M_r_r_w_w(static_cast<size_t>(nc) * nc * nb * nb * 2 * nwE * nwE, cd(0));
    for (int s = 0; s < 2; ++s) {
      if (!n[s])
        continue;
      const int* sl = site_l ? site_l[s] : nullptr;
      const int* bl = band_l ? band_l[s] : nullptr;
      if (m_complex)
        flops += ndft(s, n[s], reinterpret_cast<const cd*>(m[s]), ld[s], tau[s], site[s], band[s], sl, bl,
                      M_r_r_w_w);
      else
        flops += ndft(s, n[s], m[s], ld[s], tau[s], site[s], band[s], sl, bl, M_r_r_w_w);
    }
    if (keep_mrrww)
      Mrrww = M_r_r_w_w;
    space_transform(M_r_r_w_w);
    compute_G();
    flops += 1e-9 * 8 * std::pow(nb, 3) * nwE * nwE * nc * nc * 2;
    return flops;
  }
};

extern "C" {

tpo_acc* tpo_create(const tpo_geometry* g, char* err, int errlen) {
  auto fail = [&](const char* msg) -> tpo_acc* {
    if (err && errlen > 0)
      std::snprintf(err, errlen, "%s", msg);
    return nullptr;
  };
  if (!g || g->nb < 1 || g->nc < 1 || g->nw < 2 || g->nw % 2 || g->n_wex < 1 || g->n_kex < 1 ||
      g->n_channels < 1 || !g->wex || !g->kex || !g->channels || !g->k_add || !g->k_sub ||
      !g->dft || !g->g0 || !(g->beta > 0))
    return fail("invalid geometry");
  auto* a = new tpo_acc();
  a->nb = g->nb;
  a->nc = g->nc;
  a->nw = g->nw;
  a->beta = g->beta;
  a->wex.assign(g->wex, g->wex + g->n_wex);
  a->kex.assign(g->kex, g->kex + g->n_kex);
  a->channels.assign(g->channels, g->channels + g->n_channels);
  a->nq = g->n_kex;
  a->nv = g->n_wex;
  // FrequencyExchangeDomain::initialize (src/phys/domains/time_and_frequency/frequency_exchange_domain.cpp:26-48).
  a->ext = 0;
  for (int w : a->wex)
    a->ext = std::max(a->ext, std::abs(w));
  for (int q : a->kex)
    if (q < 0 || q >= a->nc) {
      delete a;
      return fail("k_ex index out of range");
    }
  // vertex_frequency_domain<EXTENDED>::initialize (vertex_frequency_domain.hpp:157-185).
  a->nwE = a->nw + 2 * a->ext;
  a->w_ext.resize(a->nwE);
  for (int l = 0; l < a->nwE / 2; ++l) {
    a->w_ext[a->nwE / 2 + l] = M_PI / a->beta * (1 + 2 * l);
    a->w_ext[a->nwE / 2 - 1 - l] = -M_PI / a->beta * (1 + 2 * l);
  }
  const int nc2 = a->nc * a->nc;
  a->kadd.assign(g->k_add, g->k_add + nc2);
  a->ksub.assign(g->k_sub, g->k_sub + nc2);
  a->dft.resize(nc2);
  for (int i = 0; i < nc2; ++i)
    a->dft[i] = cd(g->dft[2 * i], g->dft[2 * i + 1]);
  a->band_phase.assign(static_cast<size_t>(a->nb) * a->nc, cd(1));
  if (g->band_phase)
    for (int i = 0; i < a->nb * a->nc; ++i)
      a->band_phase[i] = cd(g->band_phase[2 * i], g->band_phase[2 * i + 1]);
  const size_t n_g0 = static_cast<size_t>(a->nb) * a->nb * 2 * a->nc * a->nwE;
  a->G0.resize(n_g0);
  for (size_t i = 0; i < n_g0; ++i)
    a->G0[i] = cd(g->g0[2 * i], g->g0[2 * i + 1]);
  a->G.assign(static_cast<size_t>(a->nb) * a->nb * 2 * nc2 * a->nwE * a->nwE, cd(0));
  a->G4.assign(a->channels.size(), std::vector<cd>(a->g4_size(), cd(0)));
  return a;
}

void tpo_destroy(tpo_acc* a) { delete a; }

void tpo_keep_intermediate(tpo_acc* a, int keep) {
  a->keep_mrrww = keep != 0;
  if (!keep)
    std::vector<cd>().swap(a->Mrrww);
}

int tpo_use_blas(const char* path) {
  if (!path) {
    g_blas_zgemm = nullptr;
    g_blas_name = "own loop zgemm";
    return 0;
  }
  void* lib = dlopen(path, RTLD_NOW | RTLD_LOCAL);
  if (!lib)
    return 0;
  auto fn = reinterpret_cast<zgemm_fn>(dlsym(lib, "zgemm_"));
  if (!fn)
    return 0;
  // One accumulator per thread (the reference's threading model): BLAS itself single-threaded
  // (MKL's sequential threading layer, MKL_THREADING_SEQUENTIAL = 1; OpenBLAS with one thread).
  if (auto layer = reinterpret_cast<int (*)(int)>(dlsym(lib, "MKL_Set_Threading_Layer")))
    layer(1);
  else if (auto oset = reinterpret_cast<void (*)(int)>(dlsym(lib, "openblas_set_num_threads")))
    oset(1);
  g_blas_zgemm = fn;
  g_blas_name = std::string("zgemm_ from ") + path;
  return 1;
}

const char* tpo_blas_name() { return g_blas_name.c_str(); }

// resetAccumulation (tp_accumulator_cpu.hpp:224-229).
void tpo_reset(tpo_acc* a) {
  for (auto& g4 : a->G4)
    std::fill(g4.begin(), g4.end(), cd(0));
}

// accumulate (tp_accumulator_cpu.hpp:231-252).
double tpo_accumulate(tpo_acc* a, const int n[2], const double* const m[2], const int ld[2],
                      const double* const tau[2], const int* const site[2],
                      const int* const band[2], double sign) {
  double flops = 0;
  if (!(n[0] + n[1]))
    return flops;
  flops += a->compute_M_and_G(n, m, false, ld, tau, site, band);
  for (int c = 0; c < static_cast<int>(a->G4.size()); ++c) {
    double f = 0;
    if (!a->update_G4(c, cd(sign), f))
      return -1;
    flops += 1e-9 * f;
  }
  return flops;
}

void tpo_compute_G(tpo_acc* a, const int n[2], const double* const m[2], const int ld[2],
                   const double* const tau[2], const int* const site[2], const int* const band[2]) {
  a->compute_M_and_G(n, m, false, ld, tau, site, band);
}

void tpo_set_spin_symmetric(tpo_acc* a, int spin_symmetric) { a->spin_symmetric = spin_symmetric != 0; }

// accumulate (tp_accumulator_cpu.hpp:231-252) for the general model: complex M, the vertices'
// right (site, band) and left (site_l, band_l) labels, a complex sign.
double tpo_accumulate_general(tpo_acc* a, const int n[2], const double* const m[2], int m_complex,
                              const int ld[2], const double* const tau[2], const int* const site[2],
                              const int* const band[2], const int* const site_l[2],
                              const int* const band_l[2], double sign_re, double sign_im) {
  double flops = 0;
  if (!(n[0] + n[1]))
    return flops;
  flops += a->compute_M_and_G(n, m, m_complex != 0, ld, tau, site, band, site_l, band_l);
  for (int c = 0; c < static_cast<int>(a->G4.size()); ++c) {
    double f = 0;
    if (!a->update_G4(c, cd(sign_re, sign_im), f))
      return -1;
    flops += 1e-9 * f;
  }
  return flops;
}

void tpo_compute_G_general(tpo_acc* a, const int n[2], const double* const m[2], int m_complex,
                           const int ld[2], const double* const tau[2], const int* const site[2],
                           const int* const band[2], const int* const site_l[2],
                           const int* const band_l[2]) {
  a->compute_M_and_G(n, m, m_complex != 0, ld, tau, site, band, site_l, band_l);
}

static long long copy_out(const std::vector<cd>& v, double* out, long long n_complex) {
  const long long n = static_cast<long long>(v.size());
  if (!out)
    return n;
  if (n_complex < n)
    return -1;
  std::memcpy(out, v.data(), v.size() * sizeof(cd));
  return n;
}

long long tpo_get_Mrrww(const tpo_acc* a, double* out, long long n) { return copy_out(a->Mrrww, out, n); }
long long tpo_get_G(const tpo_acc* a, double* out, long long n) { return copy_out(a->G, out, n); }
long long tpo_get_G4(const tpo_acc* a, int c, double* out, long long n) {
  if (c < 0 || c >= static_cast<int>(a->G4.size()) || a->G4[c].empty())
    return -1;
  return copy_out(a->G4[c], out, n);
}
long long tpo_g4_size(const tpo_acc* a) { return static_cast<long long>(a->g4_size()); }
long long tpo_g_size(const tpo_acc* a) { return static_cast<long long>(a->G.size()); }

// sumTo (tp_accumulator_cpu.hpp:895-904).
int tpo_sum_to(tpo_acc* src, tpo_acc* dst) {
  if (src->G4.size() != dst->G4.size())
    return -1;
  for (size_t c = 0; c < src->G4.size(); ++c) {
    if (src->G4[c].size() != dst->G4[c].size())
      return -1;
    for (size_t i = 0; i < src->G4[c].size(); ++i)
      dst->G4[c][i] += src->G4[c][i];
  }
  src->G4.clear();
  return 0;
}

}  // extern "C"
