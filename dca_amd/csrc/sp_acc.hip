// Single-particle accumulator: M_r_w of every measurement, summed over measurements in FP64.
//
// The reference measures (SpAccumulator, sp_accumulator.hpp:160-221; GPU sp_accumulator_gpu.hpp)
//     M_r_w(b1, s, b2, s, r, w) += sign / Nc sum_{i, j} M_s(i, j) exp(i w (tau_i - tau_j))
//                                   [b_i = b1, b_j = b2, r_i - r_j = r]
// over the single-particle Matsubara frequencies, through a delayed NFFT (a Kaiser-Bessel window
// convolved onto an oversampled time grid, one FFT at finalize; dnfft_1d.hpp). Here the same sum
// is evaluated exactly in FP64, as the orbital-resolved NDFT at equal frequencies:
//     C(o1, o2; w) = sum_{i in o1} T(w, i) sum_{j in o2} M_ij conj(T(w, j)),  T(w, i) = e^{i w tau_i}
//     M_r_w(b1, b2, r, w) = 1/Nc sum over orbital pairs o1 = (b1, r1), o2 = (b2, r2), r1 - r2 = r
// with the vertices of a slot sorted by orbital (sort_kernel). One thread owns a frequency w and a
// group of OG row orbitals o1; it walks the sorted columns j, so C(o1, o2) for the orbital o2 of the
// current run of columns accumulates in registers and is stored when the run ends. The M entries
// of a column are the same for all lanes (one frequency per lane), so every load is a broadcast.
#include "g4acc_common.hpp"
#include "g4acc_kernels.hpp"

namespace g4acc {

// T_sp[slot][w + nW * ii] = exp(i w_w tau_{perm ii}), w_w = idx_w pi / beta with idx_w = +-(2l + 1)
// (frequency_domain.cpp:29-49), reduced exactly: idx_w tau / beta taken modulo 2, sincospi.
__global__ void __launch_bounds__(256) sp_phase_kernel(const SlotDesc* __restrict__ desc, Geo g, int n_w,
                                                       int n_max, const int* __restrict__ perm,
                                                       double2* __restrict__ T) {
  const int slot = blockIdx.y;
  const SlotDesc d = desc[slot];
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n_w * d.n)
    return;
  const int w = idx % n_w, ii = idx / n_w;
  const int v = perm[static_cast<size_t>(slot) * n_max + ii];
  const int half = n_w / 2;
  const int fidx = w >= half ? 2 * (w - half) + 1 : -(2 * (half - 1 - w) + 1);
  double x = static_cast<double>(fidx) * (d.tau[v] / g.beta);
  x -= 2.0 * rint(0.5 * x);
  double s, c;
  sincospi(x, &s, &c);
  T[static_cast<size_t>(slot) * n_w * n_max + idx] = double2{c, s};
}

// C[slot][(o1 * n_orb + o2) * n_w + w] for the thread's frequency w and its OG orbitals o1.
template <int OG>
__global__ void __launch_bounds__(64) sp_ndft_kernel(const SlotDesc* __restrict__ desc, Geo g, int n_w,
                                                     int n_max, bool squared,
                                                     const int* __restrict__ perm,
                                                     const int* __restrict__ off,
                                                     const double2* __restrict__ T,
                                                     double2* __restrict__ Cout) {
  const int slot = blockIdx.z;
  const SlotDesc d = desc[slot];
  const int w = blockIdx.x * 64 + threadIdx.x;
  const int o_first = blockIdx.y * OG;
  const int n_orb = g.n_orb;
  const bool w_ok = w < n_w;
  const int wc = w_ok ? w : n_w - 1;
  const int* P = perm + static_cast<size_t>(slot) * n_max;
  const int* O = off + static_cast<size_t>(slot) * (n_orb + 1);
  const double2* Ts = T + static_cast<size_t>(slot) * n_w * n_max;
  double2* C = Cout + static_cast<size_t>(slot) * n_orb * n_orb * n_w;
  const double* xv = d.expv;
  double2 acc[OG];
  for (int o2 = 0; o2 < n_orb; ++o2) {
#pragma unroll
    for (int k = 0; k < OG; ++k)
      acc[k] = double2{0, 0};
    for (int jj = O[o2]; jj < O[o2 + 1]; ++jj) {
      const int j = P[jj];
      const double* mcol = d.m + static_cast<size_t>(d.ld) * j;
      const double2 tj = Ts[wc + static_cast<size_t>(n_w) * jj];
#pragma unroll
      for (int k = 0; k < OG; ++k) {
        const int o1 = o_first + k;
        if (o1 >= n_orb)
          break;
        double2 b{0, 0};  // sum_{i in o1} T(w, i) M_ij
        for (int ii = O[o1]; ii < O[o1 + 1]; ++ii) {
          const int i = P[ii];
          double m = xv ? (xv[i] - 1.0) * mcol[i] : mcol[i];
          if (squared)
            m *= m;
          const double2 ti = Ts[wc + static_cast<size_t>(n_w) * ii];
          b.x = fma(ti.x, m, b.x);
          b.y = fma(ti.y, m, b.y);
        }
        cfmac(acc[k], b, tj);  // += b conj(T(w, j))
      }
    }
    if (w_ok) {
#pragma unroll
      for (int k = 0; k < OG; ++k) {
        const int o1 = o_first + k;
        if (o1 < n_orb)
          C[(static_cast<size_t>(o1) * n_orb + o2) * n_w + w] = acc[k];
      }
    }
  }
}

// acc[s][(b1 + nb (b2 + nb r)) * n_w + w] += sign / Nc sum over the slots of spin s and the
// orbital pairs (b1, r1), (b2, r2) with index(r1 - r2) = r, in slot order (deterministic).
__global__ void __launch_bounds__(256) sp_reduce_kernel(const SlotDesc* __restrict__ desc, Geo g, int n_w,
                                                        int n_slots, const int* __restrict__ rsub,
                                                        const double2* __restrict__ Cin,
                                                        double2* __restrict__ acc) {
  const int nb = g.nb, nc = g.nc, n_orb = g.n_orb;
  const size_t n_lab = static_cast<size_t>(nb) * nb * nc;
  const size_t idx = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int s = blockIdx.y;
  if (idx >= n_lab * n_w)
    return;
  const int w = static_cast<int>(idx % n_w);
  const size_t lab = idx / n_w;
  const int b1 = lab % nb, b2 = (lab / nb) % nb, r = static_cast<int>(lab / (nb * nb));
  const double inv_nc = 1.0 / nc;
  double2 sum{0, 0};
  for (int slot = s; slot < n_slots; slot += 2) {
    const double f = desc[slot].sign * inv_nc;
    if (desc[slot].n == 0)
      continue;
    const double2* C = Cin + static_cast<size_t>(slot) * n_orb * n_orb * n_w;
    for (int r1 = 0; r1 < nc; ++r1) {
      // the site r2 with r1 - r2 = r: r2 = index(r1 - r) (rsub[a + nc b] = index(r_b - r_a))
      const int r2 = rsub[r + nc * r1];
      const double2 c = C[(static_cast<size_t>(b1 + nb * r1) * n_orb + (b2 + nb * r2)) * n_w + w];
      sum.x = fma(f, c.x, sum.x);
      sum.y = fma(f, c.y, sum.y);
    }
  }
  double2* a = acc + static_cast<size_t>(s) * n_lab * n_w + idx;
  a->x += sum.x;
  a->y += sum.y;
}

void launch_sp_accumulate(const SpArgs& a, hipStream_t st) {
  const Geo& g = a.g;
  hipLaunchKernelGGL(sp_phase_kernel, dim3((a.n_w * a.n_max + 255) / 256, a.n_slots), dim3(256), 0, st, a.desc,
                     g, a.n_w, a.n_max, a.perm, a.T);
  constexpr int OG = 8;
  const dim3 grid((a.n_w + 63) / 64, (g.n_orb + OG - 1) / OG, a.n_slots);
  const size_t n_lab = static_cast<size_t>(g.nb) * g.nb * g.nc;
  const dim3 rgrid(static_cast<unsigned>((n_lab * a.n_w + 255) / 256), 2);
  for (int q = 0; q < (a.squared_too ? 2 : 1); ++q) {
    hipLaunchKernelGGL((sp_ndft_kernel<OG>), grid, dim3(64), 0, st, a.desc, g, a.n_w, a.n_max, q == 1, a.perm,
                       a.off, a.T, a.C);
    hipLaunchKernelGGL(sp_reduce_kernel, rgrid, dim3(256), 0, st, a.desc, g, a.n_w, a.n_slots, a.rsub, a.C,
                       q == 0 ? a.acc : a.acc_sqr);
  }
}

}  // namespace g4acc
