// Single-band M -> G with register-resident intermediates (n_bands == 1, Nc a template constant).
//
// Same contractions as mtog_kernels.hip (S1..S4), fused so that the orbital-resolved
// intermediates never reach memory:
//   left  (per slot, thread = (w, j)):
//     B[r]          = sum_{i in site r} T[w, i] M[perm i, perm j]          r = 0..Nc-1 in registers
//     A[(k, w), j]  = sum_r dft[k, r] B[r]                                 written once
//   right (per slot, thread = (p1, a group of TW frequencies w2)):
//     E[t][r]       = sum_{j in site r} A[p1, j] conj(T[w2_t, j])          in registers
//     M~(p1,(k2,w2))= (1/Nc) sum_r E[t][r] conj(dft[k2, r]) x band phases
//     G(p1,p2)      = -G0(p1) M~ G0(p2) + delta_{p1 p2} beta G0(p1)        FP64 epilogue, stored once
// T[w, j] = exp(i w_w tau_{perm j}) and the site runs of the sorted vertices come from
// sort_kernel / phase_kernel. Per slot this is 8 nwE N^2 + 8 Nc D N + 8 D N nwE + 8 Nc D^2 flops
// and one write of G: the E matrix (D x D) of the unfused path is never stored.
#include <type_traits>

#include "g4acc_common.hpp"
#include "g4acc_kernels.hpp"
#include "site_fft.hpp"

namespace g4acc {

// Side of the square cluster whose site DFT factorises into two L-point passes (0: none).
template <int NC>
constexpr int grid_side() {
  return NC == 4 ? 2 : NC == 16 ? 4 : NC == 36 ? 6 : NC == 64 ? 8 : 0;
}

// SEP: the sites are labelled x + L y (square L x L cluster) and the left site transform runs
// as an in-register FFT (site_fft.hpp); output kx + L ky goes to the caller's momentum kmap[.].
template <typename Real, int NC, bool SEP>
__global__ void __launch_bounds__(256) mtog_left_fused(const SlotDesc* __restrict__ desc, Geo g,
                                                       int n_max, int slot0,
                                                       const int* __restrict__ perm,
                                                       const int* __restrict__ off,
                                                       const Cplx<Real>* __restrict__ T,
                                                       const Cplx<Real>* __restrict__ dft,
                                                       const int* __restrict__ kmap,
                                                       Cplx<Real>* __restrict__ A) {
  using C = Cplx<Real>;
  const int slot = slot0 + blockIdx.z;
  const SlotDesc d = desc[slot];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int w = t % g.nwE, jj = t / g.nwE;
  if (jj >= d.n)
    return;
  const int* P = perm + static_cast<size_t>(slot) * n_max;
  const int* O = off + static_cast<size_t>(slot) * (NC + 1);
  const C* Ts = T + static_cast<size_t>(slot) * g.nwE * n_max;
  const double* mcol = d.m + static_cast<size_t>(d.ld) * P[jj];
  const double* xv = d.expv;  // N and exp(V): M_ij = (exp(V_i) - 1) N_ij, formed in FP64
  C b[NC];
#pragma unroll
  for (int r = 0; r < NC; ++r) {
    C acc{0, 0};
    for (int ii = O[r]; ii < O[r + 1]; ++ii) {
      const int i = P[ii];
      const Real m = static_cast<Real>(xv ? (xv[i] - 1.0) * mcol[i] : mcol[i]);
      const C tw = Ts[w + static_cast<size_t>(g.nwE) * ii];
      acc.x = fma(tw.x, m, acc.x);
      acc.y = fma(tw.y, m, acc.y);
    }
    b[r] = acc;
  }
  C* Ac = A + static_cast<size_t>(blockIdx.z) * g.D * n_max + static_cast<size_t>(g.D) * jj + NC * w;
  if constexpr (SEP) {
    constexpr int L = grid_side<NC>();
    site_transform_2d<L, 1, Real>(b);
#pragma unroll
    for (int kk = 0; kk < NC; ++kk)
      Ac[kmap[kk]] = b[kk];
  }
  else {
#pragma unroll 4
    for (int k = 0; k < NC; ++k) {
      C acc{0, 0};
#pragma unroll
      for (int r = 0; r < NC; ++r)
        cfma(acc, dft[k + NC * r], b[r]);
      Ac[k] = acc;
    }
  }
}


// Per chunk of JV sorted vertices the workgroup stages in LDS the A rows of its 64 p1 (all four
// waves read the same rows) and the phases of its 4 TW frequencies; both are loaded
// cooperatively, eight A elements per thread in flight, and the next chunk is loaded into
// registers while the current one is consumed. The cluster DFT matrix and the G0 slice of the
// workgroup's frequencies are staged once.
template <typename Real, int NC, int TW, bool SEP>
__global__ void __launch_bounds__(256) mtog_right_fused(const SlotDesc* __restrict__ desc, Geo g,
                                                        int n_max, int slot0,
                                                        const int* __restrict__ off,
                                                        const Cplx<Real>* __restrict__ T,
                                                        const Cplx<Real>* __restrict__ A,
                                                        const Cplx<Real>* __restrict__ dft,
                                                        const double2* __restrict__ bphase,
                                                        const double2* __restrict__ G0,
                                                        const int* __restrict__ kmap,
                                                        Cplx<Real>* __restrict__ G) {
  using C = Cplx<Real>;
  constexpr int L = SEP ? grid_side<NC>() : 0;
  constexpr int WT = 4 * TW;                  // frequencies per workgroup
  constexpr int JV = 16;                      // vertices per staged chunk
  constexpr int NA = JV * 64 / 256;           // A elements loaded per thread and chunk
  constexpr int NT = (JV * WT + 255) / 256;   // phases loaded per thread and chunk
  __shared__ C Al[JV * 64];
  __shared__ C Tl[JV * WT];
  __shared__ C Dl[NC * NC];
  __shared__ double2 G0l[NC * WT];
  __shared__ int Kl[NC];                      // caller momentum of each (kx, ky) (separable DFT)
  const int slot = slot0 + blockIdx.z;
  const int s = slot & 1;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int p1 = blockIdx.x * 64 + lane;
  const int w2wg = blockIdx.y * WT;           // first frequency of the workgroup
  const int w2b = w2wg + wave * TW;           // first frequency of this wave
  const int n = desc[slot].n;
  const bool active = p1 < g.D && w2b < g.nwE;
  const int* O = off + static_cast<size_t>(slot) * (NC + 1);
  const C* Ts = T + static_cast<size_t>(slot) * g.nwE * n_max;
  const C* Ab = A + static_cast<size_t>(blockIdx.z) * g.D * n_max + static_cast<size_t>(blockIdx.x) * 64;
  const int rows = g.D - blockIdx.x * 64;     // valid p1 rows of the workgroup

  for (int i = threadIdx.x; i < NC * NC; i += 256)
    Dl[i] = dft[i];
  if (threadIdx.x < NC)
    Kl[threadIdx.x] = kmap ? kmap[threadIdx.x] : threadIdx.x;
  for (int i = threadIdx.x; i < NC * WT; i += 256) {
    const int k = i % NC, wl = i / NC, w = w2wg + wl;
    G0l[i] = w < g.nwE ? G0[s + 2 * (k + static_cast<size_t>(NC) * w)] : double2{0, 0};
  }

  C pa[NA], pt[NT];
  auto prefetch = [&](int c0) {
#pragma unroll
    for (int u = 0; u < NA; ++u) {
      const int i = threadIdx.x + 256 * u, jl = i / 64, pl = i % 64;
      pa[u] = (c0 + jl < n && pl < rows) ? Ab[pl + static_cast<size_t>(g.D) * (c0 + jl)] : C{0, 0};
    }
#pragma unroll
    for (int u = 0; u < NT; ++u) {
      const int i = threadIdx.x + 256 * u, wl = i % WT, jl = i / WT, w = w2wg + wl;
      pt[u] = (i < JV * WT && c0 + jl < n && w < g.nwE) ? Ts[w + static_cast<size_t>(g.nwE) * (c0 + jl)]
                                                       : C{0, 0};
    }
  };

  C e[TW][NC];
#pragma unroll
  for (int tt = 0; tt < TW; ++tt)
#pragma unroll
    for (int r = 0; r < NC; ++r)
      e[tt][r] = C{0, 0};
  const int tw_ok = g.nwE - w2b;  // frequencies of this wave inside the domain
  if (n > 0)
    prefetch(0);
  for (int c0 = 0; c0 < n; c0 += JV) {
    const int c1 = c0 + JV < n ? c0 + JV : n;
    __syncthreads();
#pragma unroll
    for (int u = 0; u < NA; ++u)
      Al[threadIdx.x + 256 * u] = pa[u];
#pragma unroll
    for (int u = 0; u < NT; ++u)
      if (threadIdx.x + 256 * u < JV * WT)
        Tl[threadIdx.x + 256 * u] = pt[u];
    __syncthreads();
    if (c1 < n)
      prefetch(c1);
    if (!active)
      continue;
#pragma unroll
    for (int r = 0; r < NC; ++r) {
      const int j0 = O[r] > c0 ? O[r] : c0;
      const int j1 = O[r + 1] < c1 ? O[r + 1] : c1;
#pragma unroll 4
      for (int jj = j0; jj < j1; ++jj) {
        const C av = Al[(jj - c0) * 64 + lane];
        const C* tl = Tl + (jj - c0) * WT + wave * TW;
#pragma unroll
        for (int tt = 0; tt < TW; ++tt)
          if (tt < tw_ok)
            cfmac(e[tt][r], av, tl[tt]);
      }
    }
  }
  if (n == 0)
    __syncthreads();  // Dl / G0l visible (no chunk loop ran)
  if (!active)
    return;
  // Epilogue: right site DFT, band phases, G = -G0 M~ G0 + beta G0 delta (FP64).
  const int k1 = p1 % NC, w1 = p1 / NC;
  const double2 ga = G0[s + 2 * (k1 + static_cast<size_t>(NC) * w1)];
  const double inv_nc = 1.0 / NC;
  const double2 gneg{-ga.x * inv_nc, -ga.y * inv_nc};  // -G0(p1) / Nc, folded once per row
  C* Gs = G + static_cast<size_t>(slot) * g.D * g.D + p1;
  if constexpr (L > 0) {
    {
      // Square cluster: sites are labelled x + L y, so the DFT over them is two L-point passes;
      // output kx + L ky is the caller's momentum Kl[kx + L ky].
#pragma unroll
      for (int tt = 0; tt < TW; ++tt) {
        const int w2 = w2b + tt;
        if (w2 >= g.nwE)
          break;
        site_transform_2d<L, -1, Real>(e[tt]);
#pragma unroll
        for (int kk = 0; kk < NC; ++kk) {
          const int k2 = Kl[kk];
          double2 m = to_d(e[tt][kk]);
          if (bphase)
            m = cmulc(cmul(m, bphase[k1]), bphase[k2]);
          double2 v = cmul(cmul(gneg, m), G0l[k2 + NC * (w2 - w2wg)]);
          const int p2 = k2 + NC * w2;
          if (p1 == p2) {
            v.x = fma(ga.x, g.beta, v.x);
            v.y = fma(ga.y, g.beta, v.y);
          }
          Gs[static_cast<size_t>(g.D) * p2] = from_d<C>(v);
        }
      }
      return;
    }
  }
#pragma unroll
  for (int tt = 0; tt < TW; ++tt) {
    const int w2 = w2b + tt;
    if (w2 >= g.nwE)
      break;
#pragma unroll 2
    for (int k2 = 0; k2 < NC; ++k2) {
      C acc{0, 0};
#pragma unroll
      for (int r = 0; r < NC; ++r)
        cfmac(acc, e[tt][r], Dl[k2 + NC * r]);
      double2 m = to_d(acc);
      if (bphase)
        m = cmulc(cmul(m, bphase[k1]), bphase[k2]);
      double2 v = cmul(cmul(gneg, m), G0l[k2 + NC * (w2 - w2wg)]);
      const int p2 = k2 + NC * w2;
      if (p1 == p2) {
        v.x = fma(ga.x, g.beta, v.x);
        v.y = fma(ga.y, g.beta, v.y);
      }
      Gs[static_cast<size_t>(g.D) * p2] = from_d<C>(v);
    }
  }
}

typedef float mtog_f32x4 __attribute__((ext_vector_type(4)));

// Steps of four sorted vertices a slot may need (sum over sites of ceil(n_r / 4)) that the
// matrix-core right kernel tracks; vertex capacities beyond mtog_mfma_max_vertices(nc) use the
// vector kernel.
constexpr int mtog_mfma_max_steps() { return 128; }
constexpr int mtog_mfma_max_vertices(int nc) { return 4 * (mtog_mfma_max_steps() - nc); }

// Right half on the matrix cores (complex64, Nc <= 16). Two waves share a tile of 16
// frequencies w2 by 16 rows p1; with SPLIT = 2 each takes half of the sites. For its sites r a
// wave accumulates
//     E_r[w2][p1] = sum_{j in site r} conj(T[w2, j]) A[p1, j]
// with v_mfma_f32_16x16x4f32: one step takes four sorted vertices of one site (the last step of
// a site is padded with zeros) and issues four real products. Lane l supplies the operands
// conj-T[w0 + l%16][j0 + l/16] and A[p0 + l%16][j0 + l/16], and ends up holding
// E_r[w0 + 4(l/16) + i][p0 + l%16] (i = 0..3). The two waves of a pair then swap halves through
// LDS so that each holds all sites for two of the four i: the site DFT and the FP64 epilogue are
// lane-local, and each store of one (i, k2) writes 16 consecutive rows (128 bytes) in each of four
// columns of G. Splitting the sites halves the registers that hold them (four waves per SIMD for
// Nc = 16). Operands are loaded PF steps ahead of their use; the workgroup walks all frequency
// tiles of its rows, so the step table is built once per slot.
template <int NC, bool SEP>
__global__ void __launch_bounds__(256) mtog_right_mfma(const SlotDesc* __restrict__ desc, Geo g,
                                                       int n_max, int slot0,
                                                       const int* __restrict__ off,
                                                       const float2* __restrict__ T,
                                                       const float2* __restrict__ A,
                                                       const float2* __restrict__ dft,
                                                       const double2* __restrict__ bphase,
                                                       const double2* __restrict__ G0,
                                                       const int* __restrict__ kmap,
                                                       float2* __restrict__ G) {
  using C = float2;
  constexpr int L = SEP ? grid_side<NC>() : 0;
  constexpr int PF = 4;
  constexpr int kSteps = mtog_mfma_max_steps();
  constexpr int SPLIT = NC == 16 ? 2 : 1;  // waves per tile
  constexpr int NH = NC / SPLIT;           // sites per wave
  constexpr int NI = 4 / SPLIT;            // frequencies per lane in the epilogue
  constexpr int TPW = 4 / SPLIT;           // row tiles per workgroup
  __shared__ C Dl[SEP ? 1 : NC * NC];
  __shared__ double2 G0l[NC * 16];
  __shared__ int Kl[NC];
  __shared__ int So[NC + 1];  // vertex range of each site
  __shared__ int Sp[NC + 1];  // first step of each site (steps of four vertices, per site)
  // Site halves in transit: [wave][NH sites][NI i][re, im][lane].
  __shared__ float Xl[SPLIT > 1 ? 4 * 64 * NH * NI * 2 : 1];
  const int slot = slot0 + blockIdx.z;
  const int s = slot & 1;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int half = SPLIT > 1 ? (wave & 1) : 0;
  const int p0 = blockIdx.x * (16 * TPW) + 16 * (SPLIT > 1 ? wave >> 1 : wave);
  const int* O = off + static_cast<size_t>(slot) * (NC + 1);
  const C* Ts = T + static_cast<size_t>(slot) * g.nwE * n_max;
  const C* Ab = A + static_cast<size_t>(blockIdx.z) * g.D * n_max;

  if constexpr (!SEP)
    for (int i = threadIdx.x; i < NC * NC; i += 256)
      Dl[i] = dft[i];
  if (threadIdx.x < NC)
    Kl[threadIdx.x] = kmap ? kmap[threadIdx.x] : threadIdx.x;
  if (threadIdx.x <= NC)
    So[threadIdx.x] = O[threadIdx.x];
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int r = 0; r < NC; ++r) {
      Sp[r] = acc;
      acc += (So[r + 1] - So[r] + 3) >> 2;
    }
    Sp[NC] = acc;
  }
  __syncthreads();
  // This wave's sites [r_lo, r_lo + NH) and their steps [q_lo, q_lo + n_steps).
  const int r_lo = half * NH;
  const int q_lo = Sp[r_lo], n_steps = Sp[r_lo + NH] - q_lo;
  // Step table across the lanes: step q_lo + q (q = lane + 64 h) is entry h of lane q % 64,
  // packed as first vertex (bits 0-19), vertices in the step (20-22) and, for the last step of a
  // site, 1 + (site - r_lo) (24-28): the accumulated tile is moved to that site after the step.
  int tab[kSteps / 64];
#pragma unroll
  for (int h = 0; h < kSteps / 64; ++h) {
    const int q = lane + 64 * h;
    int e = 0;
    if (q < n_steps) {
      const int qg = q_lo + q;
      int r = r_lo;
      while (Sp[r + 1] <= qg)
        ++r;
      const int base = So[r] + 4 * (qg - Sp[r]);
      const int cnt = min(4, So[r + 1] - base);
      e = base | (cnt << 20) | (qg == Sp[r + 1] - 1 ? (r - r_lo + 1) << 24 : 0);
    }
    tab[h] = e;
  }
  auto entry = [&](int q) {  // q wave-uniform
    int e = 0;
#pragma unroll
    for (int h = 0; h < kSteps / 64; ++h) {
      const int v = __builtin_amdgcn_readlane(tab[h], q & 63);
      e = (q >> 6) == h ? v : e;
    }
    return q < n_steps ? e : 0;
  };
  const int jl = lane >> 4, pl = p0 + (lane & 15);
  const bool p_ok = pl < g.D;
  const int pc = p_ok ? pl : g.D - 1;

  const int n_wt = (g.nwE + 15) / 16;
  for (int wt = blockIdx.y; wt < n_wt; wt += gridDim.y) {
    const int w0 = wt * 16;
    __syncthreads();  // the previous tile's epilogue has read G0l and Xl
    for (int i = threadIdx.x; i < NC * 16; i += 256) {
      const int k = i % NC, w = w0 + i / NC;
      G0l[i] = w < g.nwE ? G0[s + 2 * (k + static_cast<size_t>(NC) * w)] : double2{0, 0};
    }
    const int wl = w0 + (lane & 15);
    const bool w_ok = wl < g.nwE;
    // Branch-free loads at clamped (in-bounds) addresses; lanes past the step's vertices are
    // zeroed when the operands are consumed.
    const int wc = w_ok ? wl : g.nwE - 1;
    auto load = [&](int e, C& ta, C& aa) {
      const int v = (e & 0xfffff) + jl;
      const int vc = jl < ((e >> 20) & 7) ? v : 0;
      ta = Ts[wc + static_cast<size_t>(g.nwE) * vc];
      aa = Ab[pc + static_cast<size_t>(g.D) * vc];
    };
    // The products of the current site accumulate in one register tile (cr, ci); after the last
    // step of a site the tile is moved to that site's slot (er, ei). The ring of loaded operands
    // is indexed statically (the loop is unrolled by PF), so a load's destination is never moved
    // while it is in flight.
    mtog_f32x4 er[NH], ei[NH];
#pragma unroll
    for (int r = 0; r < NH; ++r)
      er[r] = ei[r] = mtog_f32x4{0, 0, 0, 0};
    mtog_f32x4 cr{0, 0, 0, 0}, ci{0, 0, 0, 0};
    C qt[PF], qa[PF];
    int qe[PF];
#pragma unroll
    for (int k = 0; k < PF; ++k) {
      qe[k] = entry(k);
      load(qe[k], qt[k], qa[k]);
    }
    for (int g0 = 0; g0 < n_steps; g0 += PF) {
#pragma unroll
      for (int k = 0; k < PF; ++k) {
        if (g0 + k >= n_steps)
          break;
        const int ce = qe[k];
        const bool lok = jl < ((ce >> 20) & 7);
        const bool tok = lok && w_ok, aok = lok && p_ok;
        const C ct{tok ? qt[k].x : 0.f, tok ? qt[k].y : 0.f};
        const C ca{aok ? qa[k].x : 0.f, aok ? qa[k].y : 0.f};
        qe[k] = entry(g0 + k + PF);
        load(qe[k], qt[k], qa[k]);
        // E += conj(T) A = (Tx Ax + Ty Ay) + i (Tx Ay - Ty Ax)
        cr = __builtin_amdgcn_mfma_f32_16x16x4f32(ct.x, ca.x, cr, 0, 0, 0);
        ci = __builtin_amdgcn_mfma_f32_16x16x4f32(ct.x, ca.y, ci, 0, 0, 0);
        cr = __builtin_amdgcn_mfma_f32_16x16x4f32(ct.y, ca.y, cr, 0, 0, 0);
        ci = __builtin_amdgcn_mfma_f32_16x16x4f32(-ct.y, ca.x, ci, 0, 0, 0);
        const int fs = ce >> 24;
        if (fs != 0) {
          switch (fs - 1) {  // a branch, not per-site selects on every step
#define MTOG_SITE_STORE(R)  \
  case R:                   \
    if constexpr (R < NH) { \
      er[R] = cr;           \
      ei[R] = ci;           \
    }                       \
    break;
            MTOG_SITE_STORE(0) MTOG_SITE_STORE(1) MTOG_SITE_STORE(2) MTOG_SITE_STORE(3)
            MTOG_SITE_STORE(4) MTOG_SITE_STORE(5) MTOG_SITE_STORE(6) MTOG_SITE_STORE(7)
            MTOG_SITE_STORE(8) MTOG_SITE_STORE(9) MTOG_SITE_STORE(10) MTOG_SITE_STORE(11)
            MTOG_SITE_STORE(12) MTOG_SITE_STORE(13) MTOG_SITE_STORE(14) MTOG_SITE_STORE(15)
#undef MTOG_SITE_STORE
            default: break;
          }
          cr = ci = mtog_f32x4{0, 0, 0, 0};
        }
      }
    }
    // Gather all sites for this lane's NI frequencies: i = NI * half + ii.
    float xr[NC][NI], xi[NC][NI];
    if constexpr (SPLIT > 1) {
      // Send this wave's sites at the partner's frequencies (i = NI (1 - half) + ii); receive the
      // partner's sites at this wave's frequencies (i = NI half + ii). Indices stay static: the
      // half only selects values.
      float* mine = Xl + wave * (NH * NI * 2 * 64) + lane;  // lane fastest: no bank conflicts
      const float* theirs = Xl + (wave ^ 1) * (NH * NI * 2 * 64) + lane;
#pragma unroll
      for (int r = 0; r < NH; ++r)
#pragma unroll
        for (int ii = 0; ii < NI; ++ii) {
          mine[64 * ((r * NI + ii) * 2)] = half ? er[r][ii] : er[r][NI + ii];
          mine[64 * ((r * NI + ii) * 2 + 1)] = half ? ei[r][ii] : ei[r][NI + ii];
        }
      __syncthreads();
#pragma unroll
      for (int r = 0; r < NH; ++r)
#pragma unroll
        for (int ii = 0; ii < NI; ++ii) {
          const float own_r = half ? er[r][NI + ii] : er[r][ii];
          const float own_i = half ? ei[r][NI + ii] : ei[r][ii];
          const float oth_r = theirs[64 * ((r * NI + ii) * 2)], oth_i = theirs[64 * ((r * NI + ii) * 2 + 1)];
          xr[r][ii] = half ? oth_r : own_r;  // sites 0..NH-1 belong to half 0
          xi[r][ii] = half ? oth_i : own_i;
          xr[NH + r][ii] = half ? own_r : oth_r;
          xi[NH + r][ii] = half ? own_i : oth_i;
        }
    }
    else {
#pragma unroll
      for (int r = 0; r < NC; ++r)
#pragma unroll
        for (int ii = 0; ii < NI; ++ii) {
          xr[r][ii] = er[r][ii];
          xi[r][ii] = ei[r][ii];
        }
    }
    if (!p_ok)
      continue;  // every wave walks the same tiles (barriers above)
    // Epilogue per held frequency: site DFT, band phases, G = -G0 M~ G0 + beta G0 delta (FP64).
    const int p1 = pl;
    const int k1 = p1 % NC, w1 = p1 / NC;
    const double2 ga = G0[s + 2 * (k1 + static_cast<size_t>(NC) * w1)];
    const double inv_nc = 1.0 / NC;
    const double2 gneg{-ga.x * inv_nc, -ga.y * inv_nc};  // -G0(p1) / Nc, folded once per row
    C* Gs = G + static_cast<size_t>(slot) * g.D * g.D + p1;
#pragma unroll
    for (int ii = 0; ii < NI; ++ii) {
      const int w2 = w0 + 4 * jl + NI * half + ii;
      if (w2 >= g.nwE)
        break;
      C e[NC];
#pragma unroll
      for (int r = 0; r < NC; ++r)
        e[r] = C{xr[r][ii], xi[r][ii]};
      if constexpr (L > 0)
        site_transform_2d<L, -1, float>(e);
      else {
        C f[NC];
#pragma unroll
        for (int k2 = 0; k2 < NC; ++k2) {
          C acc{0, 0};
#pragma unroll
          for (int r = 0; r < NC; ++r)
            cfmac(acc, e[r], Dl[k2 + NC * r]);
          f[k2] = acc;
        }
#pragma unroll
        for (int k2 = 0; k2 < NC; ++k2)
          e[k2] = f[k2];
      }
#pragma unroll
      for (int kk = 0; kk < NC; ++kk) {
        const int k2 = L > 0 ? Kl[kk] : kk;
        double2 m = to_d(e[kk]);
        if (bphase)
          m = cmulc(cmul(m, bphase[k1]), bphase[k2]);
        double2 v = cmul(cmul(gneg, m), G0l[k2 + NC * (w2 - w0)]);
        const int p2 = k2 + NC * w2;
        if (p1 == p2) {
          v.x = fma(ga.x, g.beta, v.x);
          v.y = fma(ga.y, g.beta, v.y);
        }
        Gs[static_cast<size_t>(g.D) * p2] = from_d<C>(v);
      }
    }
  }
}

namespace {

template <typename Real, int NC, int TW>
void launch_fused_nc(const MtoGArgs<Real>& a, int s0, int ns, hipStream_t st) {
  // The in-register FFT site transforms are instantiated for the 2x2, 4x4 and 8x8 tori (radix 2);
  // 6x6 keeps the dense site sums (its two-pass L = 6 form needs more registers than it saves).
  constexpr bool kSep = NC == 4 || NC == 16 || NC == 64;
  const Geo& g = a.g;
  const bool sep = kSep && a.grid_L == grid_side<NC>() && a.kmap;
  if constexpr (NC == 36 || NC == 64) {
    if (a.large && a.grid_L == grid_side<NC>() && a.kmap) {
      launch_mtog_large<Real, NC>(a, s0, ns, st);
      return;
    }
  }
  const dim3 lgrid((g.nwE * a.n_max + 255) / 256, 1, ns);
  if (sep)
    hipLaunchKernelGGL((mtog_left_fused<Real, NC, kSep>), lgrid, dim3(256), 0, st, a.desc, g, a.n_max, s0,
                       a.perm, a.off, a.T, a.dft, a.kmap, a.A);
  else
    hipLaunchKernelGGL((mtog_left_fused<Real, NC, false>), lgrid, dim3(256), 0, st, a.desc, g, a.n_max, s0,
                       a.perm, a.off, a.T, a.dft, nullptr, a.A);
  // Matrix-core right kernel (complex64): Nc <= 8, or 16 sites of a 4x4 torus (the dense
  // 16-point DFT epilogue would not fit in registers), with its step table in range.
  if constexpr (std::is_same<Real, float>::value && NC <= 16) {
    if (a.mfma && (NC <= 8 || sep) && a.n_max <= mtog_mfma_max_vertices(NC)) {
      constexpr bool kSepM = NC == 4 || NC == 16;
      const int rows = NC == 16 ? 32 : 64;  // p1 rows per workgroup (two waves per tile for Nc = 16)
      const dim3 grid((g.D + rows - 1) / rows, 1, ns);  // each workgroup walks all frequency tiles
      if (sep)
        hipLaunchKernelGGL((mtog_right_mfma<NC, kSepM>), grid, dim3(256), 0, st, a.desc, g, a.n_max, s0, a.off,
                           a.T, a.A, a.dft, a.bphase, a.G0, a.kmap, a.G);
      else if constexpr (NC <= 8)
        hipLaunchKernelGGL((mtog_right_mfma<NC, false>), grid, dim3(256), 0, st, a.desc, g, a.n_max, s0, a.off,
                           a.T, a.A, a.dft, a.bphase, a.G0, nullptr, a.G);
      return;
    }
  }
  const dim3 grid((g.D + 63) / 64, (g.nwE + 4 * TW - 1) / (4 * TW), ns);
  // The complex64 8x8 FFT epilogue needs 256 VGPRs (one wave per SIMD) against 162 for the dense
  // 64-point DFT (three waves), and ran 2.3x slower; FP64 keeps the FFT (both forms at one wave).
  constexpr bool kSepRight = kSep && (NC != 64 || std::is_same<Real, double>::value);
  if (sep && kSepRight)
    hipLaunchKernelGGL((mtog_right_fused<Real, NC, TW, kSepRight>), grid, dim3(256), 0, st, a.desc, g, a.n_max,
                       s0, a.off, a.T, a.A, a.dft, a.bphase, a.G0, a.kmap, a.G);
  else
    hipLaunchKernelGGL((mtog_right_fused<Real, NC, TW, false>), grid, dim3(256), 0, st, a.desc, g, a.n_max, s0,
                       a.off, a.T, a.A, a.dft, a.bphase, a.G0, nullptr, a.G);
}

}  // namespace

template <typename Real>
bool launch_mtog_fused(const MtoGArgs<Real>& a, hipStream_t st) {
  const Geo& g = a.g;
  if (g.nb != 1)
    return false;
  constexpr bool dp = sizeof(Real) == 8;
  auto run = [&](auto launcher) {
    for (int s0 = 0; s0 < a.n_slots; s0 += a.chunk) {
      const int ns = (a.n_slots - s0) < a.chunk ? (a.n_slots - s0) : a.chunk;
      launcher(s0, ns);
    }
    return true;
  };
  const int tw = a.tw;
#define G4_RUN(NC_, TW_) run([&](int s0, int ns) { launch_fused_nc<Real, NC_, TW_>(a, s0, ns, st); })
  switch (g.nc) {
    case 4:
      return tw == 4 ? G4_RUN(4, 4) : G4_RUN(4, 8);
    case 8:
      if (tw == 2)
        return G4_RUN(8, 2);
      return (dp || tw == 4) ? G4_RUN(8, 4) : G4_RUN(8, 8);
    case 16:
      if (tw == 1)
        return G4_RUN(16, 1);
      return (dp || tw == 2) ? G4_RUN(16, 2) : G4_RUN(16, 4);
    case 36:
      return (dp || tw != 2) ? G4_RUN(36, 1) : G4_RUN(36, 2);
    case 64:
      return G4_RUN(64, 1);
    default:
      return false;
  }
#undef G4_RUN
}

bool mtog_fused_covers(const Geo& g) {
  return g.nb == 1 && (g.nc == 4 || g.nc == 8 || g.nc == 16 || g.nc == 36 || g.nc == 64);
}

template bool launch_mtog_fused<double>(const MtoGArgs<double>&, hipStream_t);
template bool launch_mtog_fused<float>(const MtoGArgs<float>&, hipStream_t);

}  // namespace g4acc
