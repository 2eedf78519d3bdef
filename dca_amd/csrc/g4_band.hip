// Multiband G4 update, LDS-tiled: all channels of a batch of walkers in one launch (n_bands > 1).
//
// Same sums as g4_kernels.hip (tp_accumulator_cpu.hpp:407-666, band-transposed reads of
// getGMultiband, :377-405). A workgroup owns the tile
//     { (b1, b2, b3, b4, k1, w1, k2, w2, q, v) : b1 fixed; all b2, b3, b4, k1, k2;  w1, w2, q, v fixed }
// and per walker stages in LDS the three G blocks (both spins) that every entry of the tile reads,
// each (nb Nc) x (nb Nc) with rows b + nb k:
//   X = G[(., w1), (., w2)],  Y = G[(., w2+v), (., w1+v)],  Z = G[(., v-w1), (., v-w2)]
// plus the rank-one factors Ga_s(b2,b1;k1) = G_s[(b2,k1,w1),(b1,k1+q,w1+v)] and
// Gb_s(b3,b4;k2) = G_s[(b3,k2+q,w2+v),(b4,k2,w2)]. A thread owns (k1, k2, b2, b3) and loops b4:
//   MAGNETIC    += a (Ga_0 - Ga_1)(b2,b1) (Gb_0 - Gb_1)(b3,b4) - a sum_s X_s(b3,b1) Y_s(b2,b4)
//   CHARGE      += a (Ga_0 + Ga_1)(b2,b1) (Gb_0 + Gb_1)(b3,b4) - a sum_s X_s(b1,b4) Y_s(b2,b3)
//   PP_UP_DOWN  += a sum_s X_s(b1,b3) Z_{1-s}(b2,b4)
//   TRANSVERSE  -= a sum_s X_s(b2,b4) Y_{1-s}(b3,b1)
// with X(b,b') = X(b,b'; k1,k2), Y(b,b') = Y(b,b'; k2+q, k1+q), Z(b,b') = Z(b,b'; q-k1, q-k2).
// The factors that do not depend on b4 are read once per walker. Each walker's terms are formed in
// the G precision and added to FP64 totals; G4 is read and written once per launch.
#include "g4_tile_common.hpp"

namespace g4acc {

namespace {

constexpr int kBandPre = 7;  // staged G elements per thread and walker (upper bound)

constexpr int kBandThreads = 768;  // largest workgroup (keeps 168 VGPRs per thread)

template <typename Real, int NB, int MASK>
__global__ void __launch_bounds__(kBandThreads) g4_band_kernel(TiledArgs<Real> a) {
  using C = Cplx<Real>;
  constexpr bool kMag = (MASK & (1 << PH_MAGNETIC)) != 0;
  constexpr bool kChg = (MASK & (1 << PH_CHARGE)) != 0;
  constexpr bool kPP = (MASK & (1 << PP_UP_DOWN)) != 0;
  constexpr bool kTr = (MASK & (1 << PH_TRANSVERSE)) != 0;
  extern __shared__ unsigned char smem_raw[];
  C* lds = reinterpret_cast<C*>(smem_raw);
  const Geo& g = a.g;
  const int nc = g.nc, nw = g.nw, e = g.ext;
  const int R = NB * nc, ldb = R + 1;  // block rows b + NB k, padded column stride
  const int bsz = R * ldb, nbk = NB * NB * nc;
  const size_t D = g.D, DD = D * D;
  // Tile: b1 fastest (the workgroups of one tile stage the same blocks), then q, v, w1, w2. The
  // workgroups are dealt to the 8 XCDs round-robin by id; each XCD gets a contiguous range of
  // tiles, so the workgroups that stage the same blocks share its L2.
  int t = blockIdx.x;
  if (gridDim.x % 8 == 0)
    t = static_cast<int>((blockIdx.x % 8) * (gridDim.x / 8) + blockIdx.x / 8);
  const int b1 = t % NB;
  t /= NB;
  const int iq = t % g.nq;
  t /= g.nq;
  const int iv = t % g.nv;
  t /= g.nv;
  const int w1 = t % nw, w2 = t / nw;
  const int q = a.k_ex[iq], nu = a.w_ex[iv];
  const int w1e = w1 + e, w2e = w2 + e, w1n = w1 + nu + e, w2n = w2 + nu + e;
  const int w1r = nu + nw - 1 - w1 + e, w2r = nu + nw - 1 - w2 + e;
  // LDS: X0 X1 Y0 Y1 Z0 Z1 blocks ([col][row], row fastest), then Ga0 Ga1 Gb0 Gb1.
  C* Xl = lds;
  C* Yl = Xl + 2 * bsz;
  C* Zl = Yl + 2 * bsz;
  C* GAl = Zl + 2 * bsz;
  C* GBl = GAl + 2 * nbk;

  // Staging map: element i = tid + TT j of the 6 R^2 + 4 NB^2 Nc staged values of a walker.
  const int TT = blockDim.x;
  const int n_stage = 6 * R * R + 4 * nbk;
  int src[kBandPre], dst[kBandPre];
#pragma unroll
  for (int j = 0; j < kBandPre; ++j) {
    const int i = threadIdx.x + TT * j;
    src[j] = -1;
    dst[j] = 0;
    if (i < 6 * R * R) {
      const int blk = i / (R * R), idx = i % (R * R), r = idx % R, c = idx / R;
      const int sp = blk & 1, kind = blk >> 1;  // 0: X, 1: Y, 2: Z
      const int we_r = kind == 0 ? w1e : kind == 1 ? w2n : w1r;
      const int we_c = kind == 0 ? w2e : kind == 1 ? w1n : w2r;
      src[j] = static_cast<int>(sp * DD + (static_cast<size_t>(R) * we_r + r) + D * (static_cast<size_t>(R) * we_c + c));
      dst[j] = blk * bsz + c * ldb + r;
    }
    else if (i < n_stage) {
      const int f = i - 6 * R * R, part = f / nbk, idx = f % nbk;
      const int sp = part & 1;
      const int bA = idx % NB, bB = (idx / NB) % NB, k = idx / (NB * NB);
      size_t row, col;
      if (part < 2) {  // Ga_s(bA = b2, bB = b1; k1 = k): G_s[(b2, k1, w1), (b1, k1+q, w1+v)]
        row = bA + NB * (k + static_cast<size_t>(nc) * w1e);
        col = bB + NB * (a.k_add[k + nc * q] + static_cast<size_t>(nc) * w1n);
      }
      else {           // Gb_s(bA = b3, bB = b4; k2 = k): G_s[(b3, k2+q, w2+v), (b4, k2, w2)]
        row = bA + NB * (a.k_add[k + nc * q] + static_cast<size_t>(nc) * w2n);
        col = bB + NB * (k + static_cast<size_t>(nc) * w2e);
      }
      src[j] = static_cast<int>(sp * DD + row + D * col);
      dst[j] = 6 * bsz + part * nbk + idx;
    }
  }

  // The thread's entries: (k1, k2, b2, b3) and all b4.
  const int n_own = nc * nc * NB * NB;
  const bool ok = static_cast<int>(threadIdx.x) < n_own;
  const int tp = ok ? threadIdx.x : 0;
  const int k1 = tp % nc, k2 = (tp / nc) % nc, b2 = (tp / (nc * nc)) % NB, b3 = tp / (nc * nc * NB);
  const int kq1 = a.k_add[k1 + nc * q], kq2 = a.k_add[k2 + nc * q];
  const int qk1 = a.k_sub[k1 + nc * q], qk2 = a.k_sub[k2 + nc * q];
  auto xi = [&](int bA, int bB) { return (bB + NB * k2) * ldb + bA + NB * k1; };    // X(bA,bB;k1,k2)
  auto yi = [&](int bA, int bB) { return (bB + NB * kq1) * ldb + bA + NB * kq2; };  // Y(bA,bB;k2+q,k1+q)
  auto zi = [&](int bA, int bB) { return (bB + NB * qk2) * ldb + bA + NB * qk1; };  // Z(bA,bB;q-k1,q-k2)

  // FP64 totals per entry b4 and term. Each walker's terms are formed in the G precision (both
  // spins summed) and added here, so no partial sums are kept across walkers.
  double2 tot[NB][4];
#pragma unroll
  for (int b4 = 0; b4 < NB; ++b4)
#pragma unroll
    for (int c = 0; c < 4; ++c)
      tot[b4][c] = double2{0, 0};
  auto add = [](double2& t, Real x, Real y) {
    t.x += static_cast<double>(x);
    t.y += static_cast<double>(y);
  };

  // Staged values of walkers m + 1 .. m + DEPTH are loaded while walker m is consumed (DEPTH
  // register sets: two for complex64 G, one for complex128, whose second set would not fit in
  // registers). Every load is issued unconditionally (idle staging slots re-read element 0, walkers
  // past the batch re-read the last one) and the loop is unrolled by DEPTH, so the loads in flight
  // are counted exactly and staging waits only for the walker it writes.
  constexpr int DEPTH = sizeof(Real) == 4 ? 2 : 1;
  const int n_walk = a.n_walkers;
  bool use[kBandPre];
#pragma unroll
  for (int j = 0; j < kBandPre; ++j) {
    use[j] = src[j] >= 0;
    if (!use[j])
      src[j] = 0;
  }
  C pre[DEPTH][kBandPre];
  auto prefetch = [&](int m, C (&pr)[kBandPre]) {
    const C* H = a.G + (2 * static_cast<size_t>(m < n_walk ? m : n_walk - 1)) * DD;
#pragma unroll
    for (int j = 0; j < kBandPre; ++j)
      pr[j] = H[src[j]];
  };
  auto stage = [&](const C (&cur)[kBandPre]) {
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kBandPre; ++j)
      if (use[j])
        lds[dst[j]] = cur[j];
    __syncthreads();
  };
  auto consume = [&](int m) {
    const Real alpha = static_cast<Real>(0.5 * a.desc[2 * m].sign);
    // The LDS values are read in batches (the b4-independent factors, then per b4 its eight
    // values) ahead of the products that use them: interleaving reads and products made each
    // read wait for the one before.
    C x31[2], y23[2], x13[2], y31[2];
    C ga0{0, 0}, ga1{0, 0};
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const C* X = Xl + s * bsz;
      const C* Y = Yl + s * bsz;
      if (kMag)
        x31[s] = X[xi(b3, b1)];
      if (kChg)
        y23[s] = Y[yi(b2, b3)];
      if (kPP)
        x13[s] = X[xi(b1, b3)];
      if (kTr)
        y31[s] = Yl[(1 - s) * bsz + yi(b3, b1)];
    }
    if (kMag || kChg) {
      ga0 = GAl[b2 + NB * (b1 + NB * k1)];
      ga1 = GAl[nbk + b2 + NB * (b1 + NB * k1)];
    }
    const C du{alpha * (ga0.x - ga1.x), alpha * (ga0.y - ga1.y)};
    const C su{alpha * (ga0.x + ga1.x), alpha * (ga0.y + ga1.y)};
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      if (kMag)
        x31[s] = C{alpha * x31[s].x, alpha * x31[s].y};
      if (kChg)
        y23[s] = C{alpha * y23[s].x, alpha * y23[s].y};
      if (kPP)
        x13[s] = C{alpha * x13[s].x, alpha * x13[s].y};
      if (kTr)
        y31[s] = C{alpha * y31[s].x, alpha * y31[s].y};
    }
#pragma unroll
    for (int b4 = 0; b4 < NB; ++b4) {
      C yb[2], xb[2], zb[2], xt[2], gb0{0, 0}, gb1{0, 0};
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        if (kMag)
          yb[s] = Yl[s * bsz + yi(b2, b4)];
        if (kChg)
          xb[s] = Xl[s * bsz + xi(b1, b4)];
        if (kPP)
          zb[s] = Zl[(1 - s) * bsz + zi(b2, b4)];
        if (kTr)
          xt[s] = Xl[s * bsz + xi(b2, b4)];
      }
      if (kMag || kChg) {
        gb0 = GBl[b3 + NB * (b4 + NB * k2)];
        gb1 = GBl[nbk + b3 + NB * (b4 + NB * k2)];
      }
      if (kMag) {
        C t = cmul(du, C{gb0.x - gb1.x, gb0.y - gb1.y});
        cfms(t, x31[0], yb[0]);
        cfms(t, x31[1], yb[1]);
        add(tot[b4][0], t.x, t.y);
      }
      if (kChg) {
        C t = cmul(su, C{gb0.x + gb1.x, gb0.y + gb1.y});
        cfms(t, xb[0], y23[0]);
        cfms(t, xb[1], y23[1]);
        add(tot[b4][1], t.x, t.y);
      }
      if (kPP) {
        C t = cmul(x13[0], zb[0]);
        cfma(t, x13[1], zb[1]);
        add(tot[b4][2], t.x, t.y);
      }
      if (kTr) {
        C t = cmul(xt[0], y31[0]);
        cfma(t, xt[1], y31[1]);
        add(tot[b4][3], -t.x, -t.y);
      }
    }
  };
  if (n_walk > 0) {
#pragma unroll
    for (int d = 0; d < DEPTH; ++d)
      prefetch(d, pre[d]);
    for (int m = 0; m < n_walk; m += DEPTH) {
#pragma unroll
      for (int d = 0; d < DEPTH; ++d) {
        stage(pre[d]);
        prefetch(m + d + DEPTH, pre[d]);
        if (ok && m + d < n_walk)
          consume(m + d);
      }
    }
  }
  if (!ok)
    return;
  const size_t S = static_cast<size_t>(NB) * NB * NB * NB * nc * nw * nc * nw;
  const size_t qv = static_cast<size_t>(iq) + static_cast<size_t>(g.nq) * iv;
#pragma unroll
  for (int b4 = 0; b4 < NB; ++b4) {
    const size_t idx =
        b1 + NB * (b2 + NB * (b3 + NB * (b4 + static_cast<size_t>(NB) *
                                              (k1 + static_cast<size_t>(nc) *
                                                        (w1 + static_cast<size_t>(nw) *
                                                                  (k2 + static_cast<size_t>(nc) * w2))))));
    // Totals are indexed by term (static), the outputs by the channel's position in G4 order.
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int type = k == 0 ? PH_MAGNETIC : k == 1 ? PH_CHARGE : k == 2 ? PP_UP_DOWN : PH_TRANSVERSE;
      if (!(MASK & (1 << type)))
        continue;
      int c = 0;
      while (a.ch[c] != type)
        ++c;
      g4_store_t(a.g4[c], qv * S + idx, tot[b4][k], a.sum_mode);
    }
  }
}

template <typename Real, int NB>
bool launch_band_nb(TiledArgs<Real> a, hipStream_t st) {
  const Geo& g = a.g;
  int mask = 0;
  for (int c = 0; c < a.n_ch; ++c)
    mask |= 1 << a.ch[c];
  const int R = NB * g.nc;
  const int threads = ((g.nc * g.nc * NB * NB + 63) / 64) * 64;
  const int n_stage = 6 * R * R + 4 * NB * NB * g.nc;
  if (threads > kBandThreads || n_stage > kBandPre * threads ||
      2 * static_cast<size_t>(g.D) * g.D >= (size_t(1) << 31))
    return false;
  const size_t lds = sizeof(Cplx<Real>) * (6 * static_cast<size_t>(R) * (R + 1) + 4 * NB * NB * g.nc);
  const unsigned tiles = static_cast<unsigned>(g.nw) * g.nw * g.nv * g.nq * NB;
  constexpr int M = 1 << PH_MAGNETIC, Q = 1 << PH_CHARGE, P = 1 << PP_UP_DOWN, T = 1 << PH_TRANSVERSE;
  switch (mask) {
    case M | Q | P:
      hipLaunchKernelGGL((g4_band_kernel<Real, NB, M | Q | P>), dim3(tiles), dim3(threads), lds, st, a);
      return true;
    case T | M | Q | P:
      hipLaunchKernelGGL((g4_band_kernel<Real, NB, T | M | Q | P>), dim3(tiles), dim3(threads), lds, st, a);
      return true;
    default:
      return false;
  }
}

}  // namespace

template <typename Real>
bool launch_g4_band(TiledArgs<Real> a, hipStream_t st) {
  switch (a.g.nb) {
    case 2: return launch_band_nb<Real, 2>(a, st);
    case 3: return launch_band_nb<Real, 3>(a, st);
    case 4: return launch_band_nb<Real, 4>(a, st);
    default: return false;
  }
}

template bool launch_g4_band<double>(TiledArgs<double>, hipStream_t);
template bool launch_g4_band<float>(TiledArgs<float>, hipStream_t);

}  // namespace g4acc
