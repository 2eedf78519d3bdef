// M -> G kernels: one walker's M matrices (per spin sector) to the two-point function G.
//
// The reference computes, per spin s (tp_accumulator_cpu.hpp:254-306):
//   M~(r1,b1,w1; r2,b2,w2) = sum_{i in (b1,r1), j in (b2,r2)} e^{+i w1 t_i} M_ij e^{-i w2 t_j}
//                                           (CachedNdft, tp/ndft/cached_ndft_cpu.hpp:106-222)
//   M~(k1,k2) = (1/Nc) sum_{r1,r2} e^{i k1.r1} M~(r1,r2) e^{-i k2.r2} x band phases
//                                           (SpaceTransform2D, space_transform_2D.hpp:60-91)
//   G = -G0(k1,w1) M~ G0(k2,w2) + delta_{k1k2} delta_{w1w2} beta G0(k1,w1)   (:308-375)
//
// Here the same product is factored as four contractions of small K, each a kernel over a
// slot (= walker x spin) batch, with vertices grouped by orbital o = b + nb*r (a counting sort,
// the device equivalent of CachedNdftBase::sortConfiguration, cached_ndft_base.hpp:89-129):
//   S1  B[(o,w), j]     = sum_{i in o} T[w,i] M[i,j]                 K = |o|   (left NDFT)
//   S2  A[(b,k,w), j]   = sum_r e^{i k.r} B[(b+nb r, w), j]            K = Nc    (left site DFT)
//   S3  E[p1, (o2,w2)]  = sum_{j in o2} A[p1, j] conj(T[w2, j])        K = |o2|  (right NDFT)
//   S4  G[p1, p2]       = -G0 (1/Nc sum_r2 E[p1,(b2+nb r2, w2)] e^{-i k2.r2}) G0 + beta G0 delta
// with T[w,i] = e^{i w_w t_i} over the EXTENDED frequencies and p = b + nb*(k + Nc*w).
// The flop count per spin is 8 nwE N^2 + 8 Nc D N + 8 D N nwE + 8 Nc D^2 (vs. the reference's
// 8 nwE N^2 + 8 nwE^2 n_orb N + 16 Nc D^2).
//
// G is stored per slot as a D x D column-major matrix holding the value the reference's
// getGMultiband returns (tp_accumulator_cpu.hpp:388-405, a band-transposed read of G_):
//   Gs[(b1,k1,w1) + D (b2,k2,w2)] = G_(b2, b1, s, k1, k2, w1, w2)
// which is also the reference GPU layout (TT/tp/tp_accumulator_multiband_gpu_test.cpp:173).
#include "g4acc_common.hpp"
#include "g4acc_kernels.hpp"

namespace g4acc {

// ---------------------------------------------------------------------------------------------
// Counting sort of the vertices of each slot by orbital, stable in the vertex index: the order of
// details::Triple (tp/ndft/triple.hpp:23-30). perm[slot][rank] = vertex index, off[slot][o] =
// first rank of orbital o (n_orb + 1 entries). Invalid site/band indices raise *err_flag.
__global__ void __launch_bounds__(256) sort_kernel(const SlotDesc* __restrict__ desc, Geo g,
                                                   int n_max, const int* __restrict__ site_map,
                                                   int* __restrict__ perm,
                                                   int* __restrict__ off, int* __restrict__ err_flag,
                                                   bool left) {
  extern __shared__ int lds[];
  int* orb = lds;               // n_max
  int* cnt = lds + n_max;       // n_orb + 1
  const int slot = blockIdx.x;
  const SlotDesc d = desc[slot];
  const int n = d.n;
  int* P = perm + static_cast<size_t>(slot) * n_max;
  int* O = off + static_cast<size_t>(slot) * (g.n_orb + 1);
  for (int o = threadIdx.x; o <= g.n_orb; o += blockDim.x)
    cnt[o] = 0;
  __syncthreads();
  const int* bands = left && d.band_l ? d.band_l : d.band;
  const int* sites = left && d.site_l ? d.site_l : d.site;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const int b = bands[i], r = sites[i];
    int o = 0;
    if (b < 0 || b >= g.nb || r < 0 || r >= g.nc)
      atomicOr(err_flag, 1);
    else
      o = b + g.nb * (site_map ? site_map[r] : r);
    orb[i] = o;
    atomicAdd(&cnt[o + 1], 1);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int o = 1; o <= g.n_orb; ++o)
      cnt[o] += cnt[o - 1];
  }
  __syncthreads();
  for (int o = threadIdx.x; o <= g.n_orb; o += blockDim.x)
    O[o] = cnt[o];
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const int o = orb[i];
    int rank = cnt[o];
    for (int j = 0; j < i; ++j)
      rank += (orb[j] == o);
    P[rank] = i;
  }
}

// T[slot][w + nwE * ii] = exp(i w_w tau_{perm[ii]}), w_w = pi (2w - nwE + 1) / beta.
// The phase is reduced exactly: w_w tau = pi * (2w - nwE + 1) * (tau / beta), taken modulo 2 in
// FP64 and evaluated with sincospi, so its error does not grow with |w tau|.
template <typename Real>
__global__ void __launch_bounds__(256) phase_kernel(const SlotDesc* __restrict__ desc, Geo g,
                                                    int n_max, const int* __restrict__ perm,
                                                    Cplx<Real>* __restrict__ T) {
  const int slot = blockIdx.y;
  const SlotDesc d = desc[slot];
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= g.nwE * d.n)
    return;
  const int w = idx % g.nwE, ii = idx / g.nwE;
  const int v = perm[static_cast<size_t>(slot) * n_max + ii];
  const double t = d.tau[v] / g.beta;
  double x = static_cast<double>(2 * w - g.nwE + 1) * t;
  x -= 2.0 * rint(0.5 * x);
  double s, c;
  sincospi(x, &s, &c);
  T[static_cast<size_t>(slot) * g.nwE * n_max + idx] = Cplx<Real>{static_cast<Real>(c), static_cast<Real>(s)};
}

// S1: B[slot][q + D jj] = sum_{ii in orbital o} T[w, ii] M[perm ii, perm2 jj], q = o + n_orb w.
// Rows of M in the order of perm (right labels), columns in the order of perm2 (left labels; the
// same sort for density-density vertices). M real, or complex (d.m_cplx).
template <typename Real>
__global__ void __launch_bounds__(256) ndft_left_kernel(const SlotDesc* __restrict__ desc, Geo g,
                                                        int n_max, int slot0,
                                                        const int* __restrict__ perm,
                                                        const int* __restrict__ off,
                                                        const int* __restrict__ perm2,
                                                        const Cplx<Real>* __restrict__ T,
                                                        Cplx<Real>* __restrict__ B) {
  using C = Cplx<Real>;
  const int slot = slot0 + blockIdx.z;
  const SlotDesc d = desc[slot];
  const int jj = blockIdx.y;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (jj >= d.n || q >= g.D)
    return;
  const int o = q % g.n_orb, w = q / g.n_orb;
  const int* P = perm + static_cast<size_t>(slot) * n_max;
  const int* O = off + static_cast<size_t>(slot) * (g.n_orb + 1);
  const C* Ts = T + static_cast<size_t>(slot) * g.nwE * n_max;
  const int j = perm2[static_cast<size_t>(slot) * n_max + jj];
  C acc{0, 0};
  if (d.m_cplx) {
    const double2* mcol = reinterpret_cast<const double2*>(d.m) + static_cast<size_t>(d.ld) * j;
#pragma unroll 4
    for (int ii = O[o]; ii < O[o + 1]; ++ii) {
      const double2 m = mcol[P[ii]];
      cfma(acc, Ts[w + static_cast<size_t>(g.nwE) * ii], C{static_cast<Real>(m.x), static_cast<Real>(m.y)});
    }
  }
  else {
    const double* mcol = d.m + static_cast<size_t>(d.ld) * j;
    const double* xv = d.expv;  // N and exp(V): M_ij = (exp(V_i) - 1) N_ij, formed in FP64
#pragma unroll 4
    for (int ii = O[o]; ii < O[o + 1]; ++ii) {
      const int i = P[ii];
      const Real m = static_cast<Real>(xv ? (xv[i] - 1.0) * mcol[i] : mcol[i]);
      const C t = Ts[w + static_cast<size_t>(g.nwE) * ii];
      acc.x = fma(t.x, m, acc.x);
      acc.y = fma(t.y, m, acc.y);
    }
  }
  B[static_cast<size_t>(blockIdx.z) * g.D * n_max + q + static_cast<size_t>(g.D) * jj] = acc;
}

// S2: A[slot][p + D jj] = sum_r dft[k + Nc r] B[(b + nb r) + n_orb w + D jj], p = b + nb (k + Nc w).
template <typename Real>
__global__ void __launch_bounds__(256) dft_left_kernel(const SlotDesc* __restrict__ desc, Geo g,
                                                       int n_max, int slot0,
                                                       const Cplx<Real>* __restrict__ dft,
                                                       const Cplx<Real>* __restrict__ B,
                                                       Cplx<Real>* __restrict__ A) {
  using C = Cplx<Real>;
  const int slot = slot0 + blockIdx.z;
  const int jj = blockIdx.y;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (jj >= desc[slot].n || p >= g.D)
    return;
  const int b = p % g.nb, k = (p / g.nb) % g.nc, w = p / g.n_orb;
  const size_t base = static_cast<size_t>(blockIdx.z) * g.D * n_max + static_cast<size_t>(g.D) * jj;
  const C* Bc = B + base + b + static_cast<size_t>(g.n_orb) * w;
  C acc{0, 0};
#pragma unroll 8
  for (int r = 0; r < g.nc; ++r)
    cfma(acc, dft[k + g.nc * r], Bc[g.nb * r]);
  A[base + p] = acc;
}

// S3: E[slot][p1 + D q2] = sum_{jj in orbital o2} A[p1 + D jj] conj(T[w2, jj]), q2 = o2 + n_orb w2,
// over the column sort (off, T are the column order's arrays).
template <typename Real>
__global__ void __launch_bounds__(256) ndft_right_kernel(const SlotDesc* __restrict__ desc, Geo g,
                                                         int n_max, int slot0,
                                                         const int* __restrict__ off,
                                                         const Cplx<Real>* __restrict__ T,
                                                         const Cplx<Real>* __restrict__ A,
                                                         Cplx<Real>* __restrict__ E) {
  using C = Cplx<Real>;
  const int slot = slot0 + blockIdx.z;
  const int q2 = blockIdx.y;
  const int p1 = blockIdx.x * blockDim.x + threadIdx.x;
  if (p1 >= g.D)
    return;
  const int o2 = q2 % g.n_orb, w2 = q2 / g.n_orb;
  const int* O = off + static_cast<size_t>(slot) * (g.n_orb + 1);
  const C* Ts = T + static_cast<size_t>(slot) * g.nwE * n_max;
  const C* As = A + static_cast<size_t>(blockIdx.z) * g.D * n_max;
  C acc{0, 0};
#pragma unroll 4
  for (int jj = O[o2]; jj < O[o2 + 1]; ++jj)
    cfmac(acc, As[p1 + static_cast<size_t>(g.D) * jj], Ts[w2 + static_cast<size_t>(g.nwE) * jj]);
  E[static_cast<size_t>(blockIdx.z) * g.D * g.D + p1 + static_cast<size_t>(g.D) * q2] = acc;
}

// S4: right site DFT, band phases and G = -G0 M~ G0 + beta G0 delta for one (k1,w1,k2,w2) band
// block per thread; stores the band-transposed block (see the file comment).
// NB bands as a compile-time constant: the nb x nb blocks M~(e, f) and the G0 factors stay in
// registers (with a run-time band count they went to scratch memory).
template <typename Real, int NB>
__global__ void __launch_bounds__(256) dft_right_G_kernel(Geo g, int slot0,
                                                          const Cplx<Real>* __restrict__ dft,
                                                          const double2* __restrict__ bphase,
                                                          const double2* __restrict__ G0,
                                                          const Cplx<Real>* __restrict__ E,
                                                          Cplx<Real>* __restrict__ G) {
  using C = Cplx<Real>;
  const int slot = slot0 + blockIdx.z;
  const int s = slot & 1;
  const int KW = g.nc * g.nwE;
  // Workgroups are dealt to the 8 XCDs round-robin by id; a slot's grid is remapped so that each
  // XCD takes a contiguous range of (kw1 block, kw2), and the Nc workgroups of one frequency w2,
  // which read the same columns of E, share its L2.
  unsigned bx = blockIdx.x, by = blockIdx.y;
  const unsigned per_slot = gridDim.x * gridDim.y;
  if (per_slot % 8 == 0) {
    const unsigned lin = blockIdx.x + gridDim.x * blockIdx.y;
    const unsigned lid = (lin % 8) * (per_slot / 8) + lin / 8;
    bx = lid % gridDim.x;
    by = lid / gridDim.x;
  }
  const int kw1 = static_cast<int>(bx * blockDim.x + threadIdx.x);
  const int kw2 = static_cast<int>(by);
  if (kw1 >= KW)
    return;
  const int k1 = kw1 % g.nc, w1 = kw1 / g.nc;
  const int k2 = kw2 % g.nc, w2 = kw2 / g.nc;
  const C* Es = E + static_cast<size_t>(blockIdx.z) * g.D * g.D;
  const double inv_nc = 1.0 / g.nc;
  double2 mt[NB][NB];
#pragma unroll
  for (int e = 0; e < NB; ++e) {
    const size_t row = e + static_cast<size_t>(NB) * kw1;
#pragma unroll
    for (int f = 0; f < NB; ++f) {
      C acc{0, 0};
#pragma unroll 8
      for (int r2 = 0; r2 < g.nc; ++r2)
        cfmac(acc, Es[row + static_cast<size_t>(g.D) * (f + NB * r2 + static_cast<size_t>(g.n_orb) * w2)],
              dft[k2 + g.nc * r2]);
      double2 v = to_d(acc);
      v.x *= inv_nc;
      v.y *= inv_nc;
      if (bphase)
        v = cmulc(cmul(v, bphase[e + NB * k1]), bphase[f + NB * k2]);
      mt[e][f] = v;
    }
  }
  // G0 index (b1, b2, s, k, w): the left factors G0(c, e; k1, w1) per thread, the right ones
  // G0(f, dd; k2, w2) shared by the workgroup.
  auto g0 = [&](int b1, int b2, int k, int w) {
    return G0[b1 + NB * (b2 + NB * (s + 2 * (k + static_cast<size_t>(g.nc) * w)))];
  };
  double2 gl[NB][NB], tmp[NB][NB];
#pragma unroll
  for (int c = 0; c < NB; ++c)
#pragma unroll
    for (int e = 0; e < NB; ++e)
      gl[c][e] = g0(c, e, k1, w1);
#pragma unroll
  for (int e = 0; e < NB; ++e)
#pragma unroll
    for (int dd = 0; dd < NB; ++dd) {
      double2 t{0, 0};
#pragma unroll
      for (int f = 0; f < NB; ++f)
        cfma(t, mt[e][f], g0(f, dd, k2, w2));
      tmp[e][dd] = t;
    }
  C* Gs = G + static_cast<size_t>(slot) * g.D * g.D;
  const bool diag = (k1 == k2) && (w1 == w2);
#pragma unroll
  for (int c = 0; c < NB; ++c)
#pragma unroll
    for (int dd = 0; dd < NB; ++dd) {
      double2 acc{0, 0};
#pragma unroll
      for (int e = 0; e < NB; ++e)
        cfma(acc, gl[c][e], tmp[e][dd]);
      acc.x = -acc.x;
      acc.y = -acc.y;
      if (diag) {
        acc.x = fma(gl[c][dd].x, g.beta, acc.x);
        acc.y = fma(gl[c][dd].y, g.beta, acc.y);
      }
      // value G_(c, dd, s, k1, k2, w1, w2) is stored at row (dd, k1, w1), column (c, k2, w2)
      Gs[dd + static_cast<size_t>(NB) * kw1 + static_cast<size_t>(g.D) * (c + static_cast<size_t>(NB) * kw2)] =
          from_d<C>(acc);
    }
}

// M = diag(exp(V) - 1) N: the M matrix a CT-AUX walker hands to the accumulator
// (CtauxWalker::computeM, ctaux_walker.hpp:1889-1916, which forms exp(V_i) - 1 per vertex and
// scales the rows of N, src/linalg/lapack/multiply_diagonal_gpu.cu). One thread per entry, rows
// fastest (coalesced columns).
__global__ void __launch_bounds__(256) compute_m_kernel(int n, const double* __restrict__ exp_v,
                                                        const double* __restrict__ N, int ldn,
                                                        double* __restrict__ M, int ldm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= n)
    return;
  M[i + static_cast<size_t>(ldm) * j] = (exp_v[i] - 1.0) * N[i + static_cast<size_t>(ldn) * j];
}

// ---------------------------------------------------------------------------------------------
// Host launchers.

void launch_compute_m(int n, const double* exp_v, const double* N, int ldn, double* M, int ldm,
                      hipStream_t st) {
  hipLaunchKernelGGL(compute_m_kernel, dim3((n + 255) / 256, n), dim3(256), 0, st, n, exp_v, N, ldn, M, ldm);
}

void launch_sort(const SlotDesc* desc, const Geo& g, int n_slots, int n_max, const int* site_map,
                 int* perm, int* off, int* err_flag, hipStream_t st, bool left) {
  const size_t lds = sizeof(int) * (n_max + g.n_orb + 1);
  hipLaunchKernelGGL(sort_kernel, dim3(n_slots), dim3(256), lds, st, desc, g, n_max, site_map, perm, off,
                     err_flag, left);
}

template <typename Real>
void launch_phases(const SlotDesc* desc, const Geo& g, int n_slots, int n_max, const int* perm,
                   Cplx<Real>* T, hipStream_t st) {
  hipLaunchKernelGGL(phase_kernel<Real>, dim3((g.nwE * n_max + 255) / 256, n_slots), dim3(256), 0, st, desc, g,
                     n_max, perm, T);
}

template <typename Real>
void launch_mtog(const MtoGArgs<Real>& a, hipStream_t st) {
  const Geo& g = a.g;
  launch_phases<Real>(a.desc, g, a.n_slots, a.n_max, a.perm, a.T, st);
  const bool two_sorts = a.perm2 != a.perm;
  if (two_sorts)
    launch_phases<Real>(a.desc, g, a.n_slots, a.n_max, a.perm2, a.T2, st);
  // The register-fused kernels read one sort and real M; the host never selects them otherwise.
  if (a.fused && !two_sorts && launch_mtog_fused<Real>(a, st))
    return;
  for (int s0 = 0; s0 < a.n_slots; s0 += a.chunk) {
    const int ns = (a.n_slots - s0) < a.chunk ? (a.n_slots - s0) : a.chunk;
    hipLaunchKernelGGL(ndft_left_kernel<Real>, dim3((g.D + 255) / 256, a.n_max, ns), dim3(256), 0, st,
                       a.desc, g, a.n_max, s0, a.perm, a.off, a.perm2, a.T, a.B);
    hipLaunchKernelGGL(dft_left_kernel<Real>, dim3((g.D + 255) / 256, a.n_max, ns), dim3(256), 0, st,
                       a.desc, g, a.n_max, s0, a.dft, a.B, a.A);
    hipLaunchKernelGGL(ndft_right_kernel<Real>, dim3((g.D + 255) / 256, g.D, ns), dim3(256), 0, st,
                       a.desc, g, a.n_max, s0, a.off2, a.T2, a.A, a.E);
    const int KW = g.nc * g.nwE;
    const dim3 grid((KW + 255) / 256, KW, ns);
    switch (g.nb) {
      case 1: hipLaunchKernelGGL((dft_right_G_kernel<Real, 1>), grid, dim3(256), 0, st, g, s0, a.dft, a.bphase, a.G0, a.E, a.G); break;
      case 2: hipLaunchKernelGGL((dft_right_G_kernel<Real, 2>), grid, dim3(256), 0, st, g, s0, a.dft, a.bphase, a.G0, a.E, a.G); break;
      case 3: hipLaunchKernelGGL((dft_right_G_kernel<Real, 3>), grid, dim3(256), 0, st, g, s0, a.dft, a.bphase, a.G0, a.E, a.G); break;
      default: hipLaunchKernelGGL((dft_right_G_kernel<Real, 4>), grid, dim3(256), 0, st, g, s0, a.dft, a.bphase, a.G0, a.E, a.G); break;
    }
  }
}

template void launch_mtog<double>(const MtoGArgs<double>&, hipStream_t);
template void launch_mtog<float>(const MtoGArgs<float>&, hipStream_t);
template void launch_phases<double>(const SlotDesc*, const Geo&, int, int, const int*, Cplx<double>*, hipStream_t);
template void launch_phases<float>(const SlotDesc*, const Geo&, int, int, const int*, Cplx<float>*, hipStream_t);

}  // namespace g4acc
