// M -> G for the large single-band square clusters (6x6 and 8x8) on the matrix cores. Left half:
//
//   mtog_permute_m  M in sorted order (Ms), in the working precision
//   mtog_s1_mfma    S1: B[jj][r][w] = sum_{ii in site r} Ms[ii][jj] T[w, ii]
//   mtog_left_site  S2: A[(k, w)][jj] = sum_r exp(+i k.r) B[jj][r][w]
//
// Right half, in two kernels (FP64, or FP32 with E as a workspace):
//
//   mtog_s3_mfma  S3 on the matrix cores: for every site r of the column sort
//                     E[w2][r][p1] = sum_{jj in site r} conj(T[w2, jj]) A[p1, jj]
//                 written to a D x D workspace (per slot of the chunk);
//   mtog_site_G   the site transform over r and the G epilogue (FP64)
//                     G[p1, (k2, w2)] = -G0(p1) / Nc sum_r exp(-i k2.r) E[w2][r][p1] G0(k2, w2)
//                                       + beta G0(p1) delta_{p1 p2}
//                 as two L-point passes through LDS (sites labelled x + L y).
//
// Keeping E in registers, as the fused right kernels do, needs Nc complex sums per (p1, w2) pair
// and thread; at Nc = 64 that leaves one or three waves per SIMD and no room for matrix-core tiles.
// Written out, E costs one D x D write and read per slot, and both kernels run at full occupancy:
// S3 as 16 x 16 x 4 matrix-core steps of four vertices, the site transform with eight (or six)
// values per thread. In FP32 the default is one kernel, mtog_right_block, that keeps E of a
// 16 x 16 (p1, w2) block in registers and LDS instead.
#include <type_traits>

#include "g4acc_common.hpp"
#include "g4acc_kernels.hpp"
#include "site_fft.hpp"

namespace g4acc {

namespace {

template <typename Real>
struct Mfma16;
// Output row of register i of a lane in lane group jl = lane / 16 (the column is lane % 16). The
// two instructions differ (measured on gfx950): FP32 holds four consecutive rows per lane, FP64
// rows strided by four.
template <>
struct Mfma16<float> {
  typedef float acc_t __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int jl, int i) { return 4 * jl + i; }
};
template <>
struct Mfma16<double> {
  typedef double acc_t __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int jl, int i) { return jl + 4 * i; }
};

}  // namespace

// One workgroup owns 64 rows p1 (16 per wave) and 64 frequencies w2 (four 16 x 16 tiles per wave)
// and walks the sites. A step takes four vertices of one site: lane l supplies conj-T[w][jj0 + l/16]
// for w = w0 + 16 t + l%16 and A[p0 + l%16][jj0 + l/16], and after the site holds
// E[w0 + 16 t + row(l/16, i)][p0 + l%16], i = 0..3, which it stores as 16-entry runs of p1.
// Loads past the site's vertices or the tile's edges read clamped, in-bounds addresses and are
// zeroed before use.
template <typename Real, int NC>
__global__ void __launch_bounds__(256) mtog_s3_mfma(Geo g, int n_max, int slot0,
                                                    const int* __restrict__ off,
                                                    const Cplx<Real>* __restrict__ T,
                                                    const Cplx<Real>* __restrict__ A,
                                                    Cplx<Real>* __restrict__ E) {
  using C = Cplx<Real>;
  using V = typename Mfma16<Real>::acc_t;
  const int slot = slot0 + blockIdx.z;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int D = g.D, nwE = g.nwE;
  const int* O = off + static_cast<size_t>(slot) * (NC + 1);
  const C* Ts = T + static_cast<size_t>(slot) * nwE * n_max;
  const C* Ab = A + static_cast<size_t>(blockIdx.z) * D * n_max;
  C* Eb = E + static_cast<size_t>(blockIdx.z) * D * D;
  const int jl = lane >> 4;
  const int pl = blockIdx.x * 64 + 16 * wave + (lane & 15);
  const bool p_ok = pl < D;
  const int pc = p_ok ? pl : D - 1;
  const int w0 = blockIdx.y * 64;
  int wc[4];
  bool w_ok[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int w = w0 + 16 * t + (lane & 15);
    w_ok[t] = w < nwE;
    wc[t] = w_ok[t] ? w : nwE - 1;
  }
  for (int r = 0; r < NC; ++r) {
    const int b = O[r], e = O[r + 1];
    V cr[4], ci[4];
#pragma unroll
    for (int t = 0; t < 4; ++t)
      cr[t] = ci[t] = V{0, 0, 0, 0};
    for (int j0 = b; j0 < e; j0 += 4) {
      const int jj = j0 + jl;
      const bool j_ok = jj < e;
      const int jc = j_ok ? jj : b;
      C a = Ab[pc + static_cast<size_t>(D) * jc];
      if (!(j_ok && p_ok))
        a = C{0, 0};
      C tt[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        tt[t] = Ts[wc[t] + static_cast<size_t>(nwE) * jc];
        if (!(j_ok && w_ok[t]))
          tt[t] = C{0, 0};
      }
      // E += conj(T) A = (Tx Ax + Ty Ay) + i (Tx Ay - Ty Ax)
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        cr[t] = Mfma16<Real>::mma(tt[t].x, a.x, cr[t]);
        ci[t] = Mfma16<Real>::mma(tt[t].x, a.y, ci[t]);
        cr[t] = Mfma16<Real>::mma(tt[t].y, a.y, cr[t]);
        ci[t] = Mfma16<Real>::mma(-tt[t].y, a.x, ci[t]);
      }
    }
    if (p_ok) {
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int w = w0 + 16 * t + Mfma16<Real>::row(jl, i);
          if (w < nwE)
            Eb[pl + static_cast<size_t>(D) * (r + NC * w)] = C{cr[t][i], ci[t][i]};
        }
    }
  }
}

// One workgroup owns 64 rows p1 (one per lane) and one frequency w2. Pass 1: wave q transforms
// the rows y = q, q + 4, ... over x into LDS; pass 2: it transforms the columns kx = q, q + 4, ...
// over y and writes G for the L momenta (kx, ky) of each.
template <typename Real, int L>
__global__ void __launch_bounds__(256) mtog_site_G(Geo g, int slot0, const Cplx<Real>* __restrict__ E,
                                                   const double2* __restrict__ bphase,
                                                   const double2* __restrict__ G0,
                                                   const int* __restrict__ kmap,
                                                   Cplx<Real>* __restrict__ G) {
  using C = Cplx<Real>;
  constexpr int NC = L * L;
  __shared__ C F[NC * 64];  // [kx + L y][lane]
  __shared__ int Kl[NC];
  const int slot = slot0 + blockIdx.z;
  const int s = slot & 1;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int D = g.D;
  const int p1 = blockIdx.x * 64 + lane;
  const bool p_ok = p1 < D;
  const int pc = p_ok ? p1 : D - 1;
  const int w2 = blockIdx.y;
  const C* Eb = E + static_cast<size_t>(blockIdx.z) * D * D + static_cast<size_t>(D) * NC * w2 + pc;
  if (threadIdx.x < NC)
    Kl[threadIdx.x] = kmap[threadIdx.x];
  for (int y = wave; y < L; y += 4) {
    C v[L];
#pragma unroll
    for (int x = 0; x < L; ++x)
      v[x] = Eb[static_cast<size_t>(D) * (x + L * y)];
    site_transform_1d<L, -1, Real>(v);
#pragma unroll
    for (int kx = 0; kx < L; ++kx)
      F[(kx + L * y) * 64 + lane] = v[kx];
  }
  __syncthreads();
  if (!p_ok)
    return;
  const int k1 = p1 % NC, w1 = p1 / NC;
  const double2 ga = G0[s + 2 * (k1 + static_cast<size_t>(NC) * w1)];
  const double inv_nc = 1.0 / NC;
  const double2 gneg{-ga.x * inv_nc, -ga.y * inv_nc};  // -G0(p1) / Nc
  C* Gs = G + static_cast<size_t>(slot) * D * D + p1;
  for (int kx = wave; kx < L; kx += 4) {
    C v[L];
#pragma unroll
    for (int y = 0; y < L; ++y)
      v[y] = F[(kx + L * y) * 64 + lane];
    site_transform_1d<L, -1, Real>(v);
#pragma unroll
    for (int ky = 0; ky < L; ++ky) {
      const int k2 = Kl[kx + L * ky];
      double2 m = to_d(v[ky]);
      if (bphase)
        m = cmulc(cmul(m, bphase[k1]), bphase[k2]);
      double2 val = cmul(cmul(gneg, m), G0[s + 2 * (k2 + static_cast<size_t>(NC) * w2)]);
      const int p2 = k2 + NC * w2;
      if (p1 == p2) {
        val.x = fma(ga.x, g.beta, val.x);
        val.y = fma(ga.y, g.beta, val.y);
      }
      Gs[static_cast<size_t>(D) * p2] = from_d<C>(val);
    }
  }
}

// The right half in one kernel (FP32): S3, the site transform and the G epilogue of a 16 x 16 block
// of (p1, w2) pairs. Wave y owns the row of sites x + L y, x = 0..L-1, whose vertices are one
// contiguous range of the column sort. It walks that range as one stream of steps of four
// vertices (a step never spans two sites): lane l supplies conj-T[w0 + l%16][jj0 + l/16] and
// A[p0 + l%16][jj0 + l/16] to a 16 x 16 x 4 matrix-core product, with the loads PF steps ahead of
// the products and consecutive steps alternating between two accumulator sets. At the end of
// site x its tile goes to the registers of x (lane l holds the pairs (w2 local, p1 local) =
// (row(l/16, i), l%16), i = 0..3) through a branch on x, which is uniform, and after the row the
// transform over x runs in registers. The transform over y needs every row: two momenta kx at a
// time go through LDS (2 x L x 256 values, 32 KiB at L = 8; with the block's G0 values 48 KiB, so
// two workgroups share a CU and one's epilogue overlaps the other's products), and each thread
// transforms (pair, kx) tasks over y and writes G in runs of 16 consecutive p1 (128 bytes) after
// the FP64 epilogue -G0(p1) M G0(p2) / Nc + beta G0(p1) delta.
// Workgroups that share a p1 block (and so read the same 16 rows of A) are given consecutive
// logical ids on one XCD when the block count allows it, so A is read from HBM about once.
// Offsets within a slot's A, T and G fit in 32 bits (the handle selects this kernel only then),
// which keeps the per-step address arithmetic to one multiply-add per load.
template <int L>
__global__ void __launch_bounds__(64 * L, 4) mtog_right_block(Geo g, int n_max, int slot0, int n_wb, int xcd_map,
                                                            const int* __restrict__ off,
                                                            const Cplx<float>* __restrict__ T,
                                                            const Cplx<float>* __restrict__ A,
                                                            const double2* __restrict__ bphase,
                                                            const double2* __restrict__ G0,
                                                            const int* __restrict__ kmap,
                                                            Cplx<float>* __restrict__ G) {
  using C = Cplx<float>;
  using V = Mfma16<float>::acc_t;
  constexpr int NC = L * L, NT = 64 * L, KR = 2;
  constexpr int PF = 4;  // steps in flight
  __shared__ C X[KR * L * 256];  // [kx - kx0][y][w2 local * 16 + p1 local]
  __shared__ double2 Gb[16 * NC];  // G0(k2, w2) of the block's 16 frequencies w2, k2 fastest
  __shared__ double2 Ga[16];       // G0(p1) of the block's 16 rows p1
  __shared__ int Kl[NC];
  __shared__ int Ol[NC + 1];  // the slot's site offsets
  const int slot = slot0 + blockIdx.z;
  const int s = slot & 1;
  const int lane = threadIdx.x & 63;
  const int y = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int D = g.D, nwE = g.nwE;
  unsigned lid = blockIdx.x;
  if (xcd_map)
    lid = (blockIdx.x % 8) * (gridDim.x / 8) + blockIdx.x / 8;
  const int p0 = static_cast<int>(lid / n_wb) * 16, w0 = static_cast<int>(lid % n_wb) * 16;
  if (threadIdx.x < NC)
    Kl[threadIdx.x] = kmap[threadIdx.x];
  // Site offsets through LDS: reading them in the step loop then waits on LDS only, not on the
  // vector loads in flight.
  if (threadIdx.x <= NC)
    Ol[threadIdx.x] = off[static_cast<size_t>(slot) * (NC + 1) + threadIdx.x];
  for (int t = threadIdx.x; t < 16 * NC; t += NT) {
    const int w2 = w0 + t / NC;
    Gb[t] = w2 < nwE ? G0[s + 2 * (static_cast<size_t>(t % NC) + static_cast<size_t>(NC) * w2)] : double2{0, 0};
  }
  if (threadIdx.x < 16)
    Ga[threadIdx.x] = p0 + threadIdx.x < D ? G0[s + 2 * static_cast<size_t>(p0 + threadIdx.x)] : double2{0, 0};
  __syncthreads();

  const int jl = lane >> 4;
  const int pl = p0 + (lane & 15), wl = w0 + (lane & 15);
  // Per-slot bases (uniform) and this lane's element offsets within them.
  const C* Ab = A + static_cast<size_t>(blockIdx.z) * D * n_max;
  const C* Tb = T + static_cast<size_t>(slot) * nwE * n_max;
  const unsigned a_lane = static_cast<unsigned>(pl < D ? pl : D - 1);
  const unsigned t_lane = static_cast<unsigned>(wl < nwE ? wl : nwE - 1);
  const int* Orow = Ol + L * y;
  // Sites of the row with vertices, as a bit mask (uniform).
  unsigned full = 0;
#pragma unroll
  for (int x = 0; x < L; ++x)
    if (Orow[x] < Orow[x + 1])
      full |= 1u << x;
  full = __builtin_amdgcn_readfirstlane(full);

  // Step cursor (uniform across the wave): site cx (L: done), its last vertex cl, next step's
  // first vertex cj. Advancing has no loop.
  int cx = L, cj = 0, cl = 0;
  auto open_site = [&](int from) {
    const unsigned rest = from < L ? full >> from : 0u;
    cx = rest ? from + __builtin_ctz(rest) : L;
    if (cx < L) {
      cj = __builtin_amdgcn_readfirstlane(Orow[cx]);
      cl = __builtin_amdgcn_readfirstlane(Orow[cx + 1]) - 1;
    }
  };
  open_site(0);
  C xs[PF], ts[PF];
  int rs[PF];  // site of the step in the ring position (L: none)
  float ms[PF];  // 1 for vertices of the site, 0 past its end
  auto fetch = [&](int k) {
    rs[k] = cx;
    const int jj = cj + jl;
    // Past the site's last vertex the lane reads that vertex again (finite data) and is masked.
    ms[k] = cx < L && jj <= cl ? 1.0f : 0.0f;
    const unsigned jc = static_cast<unsigned>(jj <= cl ? jj : cl);
    xs[k] = Ab[a_lane + jc * static_cast<unsigned>(D)];
    ts[k] = Tb[t_lane + jc * static_cast<unsigned>(nwE)];
    if (cx < L) {
      cj += 4;
      if (cj > cl)
        open_site(cx + 1);
    }
  };
#pragma unroll
  for (int k = 0; k < PF; ++k)
    fetch(k);

  C e[4][L];  // [pair i][site x], then [pair i][kx]
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int x = 0; x < L; ++x)
      e[i][x] = C{0, 0};
  V cr0{0, 0, 0, 0}, ci0{0, 0, 0, 0}, cr1{0, 0, 0, 0}, ci1{0, 0, 0, 0};
  int acc_x = rs[0];
  auto flush = [&]() {
    C t[4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
      t[i] = C{cr0[i] + cr1[i], ci0[i] + ci1[i]};
    switch (acc_x) {
#define G4ACC_SITE_CASE(X)          \
  case X:                           \
    if constexpr (X < L)            \
      for (int i = 0; i < 4; ++i)   \
        e[i][X < L ? X : 0] = t[i]; \
    break;
      G4ACC_SITE_CASE(0)
      G4ACC_SITE_CASE(1)
      G4ACC_SITE_CASE(2)
      G4ACC_SITE_CASE(3)
      G4ACC_SITE_CASE(4)
      G4ACC_SITE_CASE(5)
      G4ACC_SITE_CASE(6)
      G4ACC_SITE_CASE(7)
#undef G4ACC_SITE_CASE
      default:
        break;
    }
    cr0 = ci0 = cr1 = ci1 = V{0, 0, 0, 0};
  };
  static_assert(L <= 8, "site cases above");
  // Every ring position issues its two loads on every pass, also past the stream's end (dummy
  // steps on a valid address), so the loads in flight are the same on every path through the
  // loop and the wait before a step covers only that step's loads. Steps past the end skip the
  // products.
  while (rs[0] < L) {
#pragma unroll
    for (int k = 0; k < PF; ++k) {
      if (rs[k] < L) {
        if (rs[k] != acc_x) {
          flush();
          acc_x = rs[k];
        }
        // Vertices past the site's end: A scaled by 0, not selected in place, which would make
        // the compiler wait for every load in flight.
        const C xa{xs[k].x * ms[k], xs[k].y * ms[k]}, ta = ts[k];
        // E += conj(T) A = (Tx Ax + Ty Ay) + i (Tx Ay - Ty Ax)
        if (k % 2 == 0) {
          cr0 = Mfma16<float>::mma(ta.x, xa.x, cr0);
          ci0 = Mfma16<float>::mma(ta.x, xa.y, ci0);
          cr0 = Mfma16<float>::mma(ta.y, xa.y, cr0);
          ci0 = Mfma16<float>::mma(-ta.y, xa.x, ci0);
        }
        else {
          cr1 = Mfma16<float>::mma(ta.x, xa.x, cr1);
          ci1 = Mfma16<float>::mma(ta.x, xa.y, ci1);
          cr1 = Mfma16<float>::mma(ta.y, xa.y, cr1);
          ci1 = Mfma16<float>::mma(-ta.y, xa.x, ci1);
        }
      }
      fetch(k);
    }
  }
  if (acc_x < L)
    flush();
#pragma unroll
  for (int i = 0; i < 4; ++i)
    site_transform_1d<L, -1, float>(e[i]);

  C* Gs = G + static_cast<size_t>(slot) * D * D;
  const double inv_nc = 1.0 / NC;
#pragma unroll
  for (int kx0 = 0; kx0 < L; kx0 += KR) {
    if (kx0 > 0)
      __syncthreads();  // the previous momenta have been read
#pragma unroll
    for (int kk = 0; kk < KR; ++kk)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        X[(kk * L + y) * 256 + Mfma16<float>::row(jl, i) * 16 + (lane & 15)] = e[i][kx0 + kk];
    __syncthreads();
    for (int t = threadIdx.x; t < KR * 256; t += NT) {
      const int pr = t & 255, kk = t >> 8;
      const int p1 = p0 + (pr & 15), w2 = w0 + (pr >> 4);
      if (p1 >= D || w2 >= nwE)
        continue;
      const int kx = kx0 + kk;
      const double2 ga = Ga[pr & 15];                       // G0(k1, w1), p1 = k1 + Nc w1
      const double2 gneg{-ga.x * inv_nc, -ga.y * inv_nc};  // -G0(p1) / Nc
      C v[L];
#pragma unroll
      for (int yy = 0; yy < L; ++yy)
        v[yy] = X[(kk * L + yy) * 256 + pr];
      site_transform_1d<L, -1, float>(v);
      const int k1 = p1 % NC;
#pragma unroll
      for (int ky = 0; ky < L; ++ky) {
        const int k2 = Kl[kx + L * ky];
        const unsigned p2 = static_cast<unsigned>(k2 + NC * w2);
        double2 m = to_d(v[ky]);
        if (bphase)
          m = cmulc(cmul(m, bphase[k1]), bphase[k2]);
        double2 val = cmul(cmul(gneg, m), Gb[k2 + NC * (pr >> 4)]);
        if (static_cast<unsigned>(p1) == p2) {
          val.x = fma(ga.x, g.beta, val.x);
          val.y = fma(ga.y, g.beta, val.y);
        }
        Gs[static_cast<unsigned>(p1) + p2 * static_cast<unsigned>(D)] = from_d<C>(val);
      }
    }
  }
}

// Left half, step 0: M of the slot in sorted order, Ms[ii][jj] = M[perm ii, perm jj] (row-major,
// stride n_max, working precision), with M = diag(exp(V) - 1) N formed here for the CT-AUX
// hand-off. 64 x 64 tiles through LDS: reads run down M's columns, writes along Ms's rows.
template <typename Real>
__global__ void __launch_bounds__(256) mtog_permute_m(const SlotDesc* __restrict__ desc, int n_max, int slot0,
                                                      const int* __restrict__ perm, Real* __restrict__ Ms) {
  __shared__ Real tile[64][65];
  const int slot = slot0 + blockIdx.z;
  const SlotDesc d = desc[slot];
  // Workgroups are dealt to the 8 XCDs round-robin by id; a slot's tiles are remapped so that each
  // XCD takes a contiguous range (column blocks slowest), and the tiles that gather from the same 64
  // columns of M, whose cache lines they share, run on one XCD.
  unsigned bx = blockIdx.x, by = blockIdx.y;
  const unsigned per_slot = gridDim.x * gridDim.y;
  if (per_slot % 8 == 0) {
    const unsigned lin = blockIdx.x + gridDim.x * blockIdx.y;
    const unsigned lid = (lin % 8) * (per_slot / 8) + lin / 8;
    bx = lid % gridDim.x;
    by = lid / gridDim.x;
  }
  const int ii0 = static_cast<int>(bx) * 64, jj0 = static_cast<int>(by) * 64;
  if (ii0 >= d.n || jj0 >= d.n)
    return;
  const int* P = perm + static_cast<size_t>(slot) * n_max;
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int a = e & 63, b = e >> 6;  // a: row within the tile (fastest in M's column)
    const int ii = ii0 + a, jj = jj0 + b;
    Real v = 0;
    if (ii < d.n && jj < d.n) {
      const int i = P[ii];
      const double m = d.m[i + static_cast<size_t>(d.ld) * P[jj]];
      v = static_cast<Real>(d.expv ? (d.expv[i] - 1.0) * m : m);
    }
    tile[b][a] = v;
  }
  __syncthreads();
  Real* out = Ms + static_cast<size_t>(blockIdx.z) * n_max * n_max;
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int b = e & 63, a = e >> 6;  // b: column within the tile (fastest in Ms's row)
    const int ii = ii0 + a, jj = jj0 + b;
    if (ii < d.n && jj < d.n)
      out[jj + static_cast<size_t>(n_max) * ii] = tile[b][a];
  }
}

// Left half, step 1 (S1 on the matrix cores): for every site r of the sort,
//     B[jj][r][w] = sum_{ii in site r} Ms[ii][jj] T[w, ii].
// A workgroup owns 64 columns jj (16 per wave) and 64 frequencies w (four 16 x 16 tiles per wave)
// and walks the sites; a step takes four vertices of one site. Lane l supplies Ms[ii0 + l/16][jj]
// for jj = j0 + l%16 and T[w0 + 16 t + l%16][ii0 + l/16], and holds
// B[j0 + row(l/16, i)][r][w0 + 16 t + l%16], stored as 16-entry runs of w.
template <typename Real, int NC>
__global__ void __launch_bounds__(256) mtog_s1_mfma(const SlotDesc* __restrict__ desc, Geo g, int n_max,
                                                    int slot0, const int* __restrict__ off,
                                                    const Cplx<Real>* __restrict__ T,
                                                    const Real* __restrict__ Ms, Cplx<Real>* __restrict__ B) {
  using C = Cplx<Real>;
  using V = typename Mfma16<Real>::acc_t;
  const int slot = slot0 + blockIdx.z;
  const int n = desc[slot].n;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int j0 = blockIdx.x * 64 + 16 * wave;
  if (blockIdx.x * 64 >= n)
    return;
  const int nwE = g.nwE;
  const int* O = off + static_cast<size_t>(slot) * (NC + 1);
  const C* Ts = T + static_cast<size_t>(slot) * nwE * n_max;
  const Real* Mz = Ms + static_cast<size_t>(blockIdx.z) * n_max * n_max;
  C* Bz = B + static_cast<size_t>(blockIdx.z) * g.D * n_max;
  const int kl = lane >> 4;
  const int jq = j0 + (lane & 15);
  const bool j_ok = jq < n;
  const int jc = j_ok ? jq : n - 1;
  const int w0 = blockIdx.y * 64;
  int wc[4];
  bool w_ok[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int w = w0 + 16 * t + (lane & 15);
    w_ok[t] = w < nwE;
    wc[t] = w_ok[t] ? w : nwE - 1;
  }
  // The sites' steps form one stream (four vertices of one site per step), walked with two steps
  // in flight: the loads of step i + 2 are issued while step i is multiplied. Every load is issued
  // (clamped indices, finite data; past the stream's end a dummy step) and no load sits under a
  // branch, so the wait before a step covers only that step's loads. The lane's M value (row j,
  // step vertex i) is scaled by 0 past the site's vertices or the column count, which zeroes those
  // products; columns past nwE (clamped T reads) are computed but not stored. A site's B rows are
  // stored when the stream moves past it (zeros for sites without vertices).
  __shared__ int Os[NC + 1];
  if (threadIdx.x <= NC)
    Os[threadIdx.x] = O[threadIdx.x];
  __syncthreads();
  int cs = NC, cj = 0, ce = 0;  // cursor (uniform): site of the next step, its first vertex, site end
  auto open_site = [&](int from) {
    cs = NC;
    for (int r = from; r < NC; ++r)
      if (Os[r] < Os[r + 1]) {
        cs = r;
        cj = Os[r];
        ce = Os[r + 1];
        break;
      }
  };
  open_site(0);
  Real mv[2];
  C tv[2][4];
  int sv[2];  // site of the step in the ring position (NC: none)
  auto fetch = [&](int k) {
    sv[k] = cs;
    const int ii = cj + kl;
    const bool i_ok = cs < NC && ii < ce;
    const int ic = i_ok ? ii : 0;
    mv[k] = Mz[jc + static_cast<size_t>(n_max) * ic] * static_cast<Real>(i_ok && j_ok ? 1 : 0);
#pragma unroll
    for (int t = 0; t < 4; ++t)
      tv[k][t] = Ts[wc[t] + static_cast<size_t>(nwE) * ic];
    if (cs < NC) {
      cj += 4;
      if (cj >= ce)
        open_site(cs + 1);
    }
  };
  V cr[4], ci[4];
#pragma unroll
  for (int t = 0; t < 4; ++t)
    cr[t] = ci[t] = V{0, 0, 0, 0};
  int cur = 0;  // site being accumulated
  // Stores the accumulated site cur and zeros for the sites without vertices up to (excluding) s.
  auto store_to = [&](int s) {
    for (; cur < s; ++cur) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int jj = j0 + Mfma16<Real>::row(kl, i);
        if (jj < n)
#pragma unroll
          for (int t = 0; t < 4; ++t)
            if (w_ok[t])
              Bz[w0 + 16 * t + (lane & 15) + static_cast<size_t>(nwE) * (cur + NC * static_cast<size_t>(jj))] =
                  C{cr[t][i], ci[t][i]};
      }
#pragma unroll
      for (int t = 0; t < 4; ++t)
        cr[t] = ci[t] = V{0, 0, 0, 0};
    }
  };
  fetch(0);
  fetch(1);
  while (sv[0] < NC) {
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      if (sv[k] < NC) {
        if (sv[k] != cur)
          store_to(sv[k]);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          cr[t] = Mfma16<Real>::mma(mv[k], tv[k][t].x, cr[t]);
          ci[t] = Mfma16<Real>::mma(mv[k], tv[k][t].y, ci[t]);
        }
      }
      fetch(k);
    }
  }
  store_to(NC);
}

// Left half, step 2: the site transform A[(k, w)][jj] = sum_r exp(+i k.r) B[jj][r][w] for one
// column jj and 64 frequencies w (one per lane), sites labelled x + L y: pass 1 over x (wave q takes
// the rows y = q, q + 4, ...), pass 2 over y, then the block of A (NC consecutive momenta per
// frequency, caller momentum order) is copied out through LDS as one contiguous run.
template <typename Real, int L>
__global__ void __launch_bounds__(256) mtog_left_site(const SlotDesc* __restrict__ desc, Geo g, int n_max,
                                                      int slot0, const Cplx<Real>* __restrict__ B,
                                                      const int* __restrict__ kmap, Cplx<Real>* __restrict__ A) {
  using C = Cplx<Real>;
  constexpr int NC = L * L;
  __shared__ C F[NC * 64];  // [kx + L y][lane], then [k][lane]
  __shared__ int Kl[NC];
  const int slot = slot0 + blockIdx.z;
  const int jj = blockIdx.y;
  if (jj >= desc[slot].n)
    return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nwE = g.nwE;
  const int w0 = blockIdx.x * 64;
  const int w = w0 + lane;
  const int wc = w < nwE ? w : nwE - 1;
  const C* Bj = B + static_cast<size_t>(blockIdx.z) * g.D * n_max + static_cast<size_t>(nwE) * NC * jj + wc;
  if (threadIdx.x < NC)
    Kl[threadIdx.x] = kmap[threadIdx.x];
  for (int y = wave; y < L; y += 4) {
    C v[L];
#pragma unroll
    for (int x = 0; x < L; ++x)
      v[x] = Bj[static_cast<size_t>(nwE) * (x + L * y)];
    site_transform_1d<L, 1, Real>(v);
#pragma unroll
    for (int kx = 0; kx < L; ++kx)
      F[(kx + L * y) * 64 + lane] = v[kx];
  }
  __syncthreads();
  constexpr int KX = (L + 3) / 4;  // columns kx per wave
  C v[KX][L];
#pragma unroll
  for (int c = 0; c < KX; ++c) {
    const int kx = wave + 4 * c;
    if (kx < L)
#pragma unroll
      for (int y = 0; y < L; ++y)
        v[c][y] = F[(kx + L * y) * 64 + lane];
  }
  __syncthreads();  // F is reused for the output block
#pragma unroll
  for (int c = 0; c < KX; ++c) {
    const int kx = wave + 4 * c;
    if (kx < L) {
      site_transform_1d<L, 1, Real>(v[c]);
#pragma unroll
      for (int ky = 0; ky < L; ++ky)
        F[Kl[kx + L * ky] * 64 + lane] = v[c][ky];
    }
  }
  __syncthreads();
  const int nw = min(64, nwE - w0);
  C* Aj = A + static_cast<size_t>(blockIdx.z) * g.D * n_max + static_cast<size_t>(g.D) * jj +
          static_cast<size_t>(NC) * w0;
  for (int e = threadIdx.x; e < NC * nw; e += 256)
    Aj[e] = F[(e % NC) * 64 + e / NC];
}

template <typename Real, int NC>
void launch_mtog_large(const MtoGArgs<Real>& a, int s0, int ns, hipStream_t st) {
  constexpr int L = NC == 64 ? 8 : 6;
  static_assert(L * L == NC, "6x6 and 8x8 clusters");
  const Geo& g = a.g;
  Real* Ms = reinterpret_cast<Real*>(a.Ms);
  const int nt = (a.n_max + 63) / 64;
  hipLaunchKernelGGL((mtog_permute_m<Real>), dim3(nt, nt, ns), dim3(256), 0, st, a.desc, a.n_max, s0, a.perm, Ms);
  hipLaunchKernelGGL((mtog_s1_mfma<Real, NC>), dim3(nt, (g.nwE + 63) / 64, ns), dim3(256), 0, st, a.desc, g,
                     a.n_max, s0, a.off, a.T, Ms, a.B);
  hipLaunchKernelGGL((mtog_left_site<Real, L>), dim3((g.nwE + 63) / 64, a.n_max, ns), dim3(256), 0, st, a.desc,
                     g, a.n_max, s0, a.B, a.kmap, a.A);
  if constexpr (std::is_same<Real, float>::value) {
    // Set by the handle only where a slot's A, T and G fit 32-bit offsets.
    if (a.right_fused) {
      const int n_wb = (g.nwE + 15) / 16;
      const unsigned nblk = static_cast<unsigned>((g.D + 15) / 16) * n_wb;
      hipLaunchKernelGGL((mtog_right_block<L>), dim3(nblk, 1, ns), dim3(64 * L), 0, st, g, a.n_max, s0,
                         n_wb, nblk % 8 == 0 ? 1 : 0, a.off, a.T, a.A, a.bphase, a.G0, a.kmap, a.G);
      return;
    }
  }
  const dim3 g3((g.D + 63) / 64, (g.nwE + 63) / 64, ns);
  hipLaunchKernelGGL((mtog_s3_mfma<Real, NC>), g3, dim3(256), 0, st, g, a.n_max, s0, a.off, a.T, a.A, a.E);
  const dim3 gs((g.D + 63) / 64, g.nwE, ns);
  hipLaunchKernelGGL((mtog_site_G<Real, L>), gs, dim3(256), 0, st, g, s0, a.E, a.bphase, a.G0, a.kmap, a.G);
}

template void launch_mtog_large<float, 36>(const MtoGArgs<float>&, int, int, hipStream_t);
template void launch_mtog_large<float, 64>(const MtoGArgs<float>&, int, int, hipStream_t);
template void launch_mtog_large<double, 36>(const MtoGArgs<double>&, int, int, hipStream_t);
template void launch_mtog_large<double, 64>(const MtoGArgs<double>&, int, int, hipStream_t);

}  // namespace g4acc
