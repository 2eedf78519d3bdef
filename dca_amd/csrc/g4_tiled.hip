// Single-band G4 update, LDS-tiled: all channels of a batch of walkers in one launch.
//
// Same sums as g4_kernels.hip (tp_accumulator_cpu.hpp:407-666) for n_bands == 1, organised around
// what the entries of one tile share. A workgroup owns the tile
//     { (k1, w1, k2, w2, q, v) : all k1, k2;  w1, w2 fixed;  v fixed;  q in a group of QG }
// and, per walker and spin, reads
//   X(k1,k2)    = G[(k1, w1), (k2, w2)]                        directly (column segments)
//   Yb(a,b)     = G[(a, w2+v), (b, w1+v)]                      into LDS: Y^q(k1,k2) = Yb(k2+q, k1+q)
//   Zb(a,b)     = G[(a, v-w1), (b, v-w2)]                      into LDS: Z^q(k1,k2) = Zb(q-k1, q-k2)
//   du,su(k1)   = a (u_0 -+ u_1),  u_s = G_s[(k1,w1), (k1+q, w1+v)]   precomputed (uv_kernel)
//   dv,sv(k2)   =    v_0 -+ v_1,   v_s = G_s[(k2+q,w2+v), (k2,w2)]
// so every G element a tile needs is fetched once per walker and reused by the QG momenta.
//   MAGNETIC    += du dv - a (X_0 Y_0 + X_1 Y_1)
//   CHARGE      += su sv - a (X_0 Y_0 + X_1 Y_1)
//   PP_UP_DOWN  += a (X_0 Z_1 + X_1 Z_0)
//   TRANSVERSE  -= a (X_0 Y_1 + X_1 Y_0)
// Products are formed in the G precision and summed per term type over 8 walkers before they are
// added to FP64 accumulators; G4 is read and written once per launch (FP64, or FP32 value + Kahan
// compensation).
#include <algorithm>
#include <type_traits>
#include <utility>

#include "g4_tile_common.hpp"

namespace g4acc {

// Two values read together from LDS (one 16-byte read for complex64).
template <typename C>
struct alignas(2 * sizeof(C)) CPair {
  C a, b;
};

// A 16-byte pair is stored whole (one ds_write_b128); a 32-byte one member by member, since a
// 32-byte aligned temporary goes through scratch memory.
template <typename C>
__device__ __forceinline__ void store_pair(CPair<C>& dst, C a, C b) {
  if constexpr (sizeof(C) == 8) {
    dst = CPair<C>{a, b};
  }
  else {
    dst.a = a;
    dst.b = b;
  }
}

// UV[((m * 4 + plane) * nqv + qv) * P + p], planes: a(u0-u1), a(u0+u1), v0-v1, v0+v1.
template <typename Real>
__global__ void __launch_bounds__(256) uv_kernel(TiledArgs<Real> a) {
  using C = Cplx<Real>;
  const Geo& g = a.g;
  const int P = g.nc * g.nw;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P)
    return;
  const int qv = blockIdx.y, m = blockIdx.z;
  const int nqv = g.nq * g.nv;
  const int q = a.k_ex[qv % g.nq], nu = a.w_ex[qv / g.nq];
  const int k = p % g.nc, w = p / g.nc;
  const int kq = a.k_add[k + g.nc * q];
  const size_t D = g.D;
  const C* G0 = a.G + (2 * static_cast<size_t>(m)) * D * D;
  const C* G1 = G0 + D * D;
  const size_t ru = k + static_cast<size_t>(g.nc) * (w + g.ext);
  const size_t cu = kq + static_cast<size_t>(g.nc) * (w + nu + g.ext);
  const C u0 = G0[ru + D * cu], u1 = G1[ru + D * cu];
  const C v0 = G0[cu + D * ru], v1 = G1[cu + D * ru];
  const Real alpha = static_cast<Real>(0.5 * a.desc[2 * m].sign);
  C* out = a.UV + (static_cast<size_t>(m) * 4 * nqv + qv) * P + p;
  const size_t plane = static_cast<size_t>(nqv) * P;
  out[0] = C{alpha * (u0.x - u1.x), alpha * (u0.y - u1.y)};
  out[plane] = C{alpha * (u0.x + u1.x), alpha * (u0.y + u1.y)};
  out[2 * plane] = C{v0.x - v1.x, v0.y - v1.y};
  out[3 * plane] = C{v0.x + v1.x, v0.y + v1.y};
}

// MASK: compile-time channel set (bit 1 << type), or 0 to read the channel flags at run time.
// RX: the rank-one terms of MAGNETIC and CHARGE come precomputed for the batch (a.R, g4_rank1.hip)
// and are added at the G4 store; the kernel then stages no factors and forms no rank-one products.
template <typename Real, int TT, int PPT, int QG, int NCH, int MASK, bool RX>
__global__ void __launch_bounds__(TT) g4_tiled_kernel(TiledArgs<Real> a) {
  using C = Cplx<Real>;
  extern __shared__ unsigned char smem_raw[];
  C* lds = reinterpret_cast<C*>(smem_raw);
  const Geo& g = a.g;
  const int nc = g.nc, nw = g.nw, e = g.ext, P = nc * nw;
  const size_t D = g.D, DD = D * D;
  int w1, w2, z;
  if (!tile_coords(a, w1, w2, z))
    return;
  const int iv = z / a.nqg, q0 = (z % a.nqg) * QG;
  const int nu = a.w_ex[iv];
  const int ld = nc + 1, bs = nc * ld;
  constexpr bool kStatic = MASK != 0;
  constexpr int kY = (1 << PH_MAGNETIC) | (1 << PH_CHARGE) | (1 << 1);
  const bool needY = kStatic ? (MASK & kY) != 0 : a.need_y;
  const bool needZ = kStatic ? (MASK & (1 << PP_UP_DOWN)) != 0 : a.need_z;
  // Both spins of a staged element side by side, and the two factors a thread reads together.
  using P2 = CPair<C>;
  P2* Yl = reinterpret_cast<P2*>(lds);  // [b][a] (a fastest): {spin 0, spin 1}
  P2* Zl = Yl + (needY ? bs : 0);
  P2* Ul = Zl + (needZ ? bs : 0);       // [side][qi][k]: side 0 {a(u0-u1), a(u0+u1)}, 1 {v0-v1, v0+v1}
  const size_t xoff = nc * static_cast<size_t>(w1 + e) + D * (nc * static_cast<size_t>(w2 + e));
  const size_t yoff = nc * static_cast<size_t>(w2 + nu + e) + D * (nc * static_cast<size_t>(w1 + nu + e));
  const size_t zoff = nc * static_cast<size_t>(nu + nw - 1 - w1 + e) +
                      D * (nc * static_cast<size_t>(nu + nw - 1 - w2 + e));
  const int nqv = g.nq * g.nv;

  int k1[PPT], k2[PPT];
  int yi[PPT][QG], zi[PPT][QG];
  bool ok[PPT];
  int qv[QG];
  bool qok[QG];
#pragma unroll
  for (int qi = 0; qi < QG; ++qi) {
    qok[qi] = q0 + qi < g.nq;
    qv[qi] = (qok[qi] ? q0 + qi : 0) + g.nq * iv;
  }
#pragma unroll
  for (int j = 0; j < PPT; ++j) {
    const int pi = threadIdx.x + TT * j;
    ok[j] = pi < nc * nc;
    k1[j] = ok[j] ? pi % nc : 0;
    k2[j] = ok[j] ? pi / nc : 0;
#pragma unroll
    for (int qi = 0; qi < QG; ++qi) {
      const int q = a.k_ex[qok[qi] ? q0 + qi : 0];
      yi[j][qi] = a.k_add[k1[j] + nc * q] * ld + a.k_add[k2[j] + nc * q];
      zi[j][qi] = a.k_sub[k2[j] + nc * q] * ld + a.k_sub[k1[j] + nc * q];
    }
  }
  // The rank-one factors: one staged element per thread (4 QG nc <= 256).
  const int n_uv = 4 * QG * nc;
  const int u_plane = threadIdx.x / (QG * nc), u_qi = (threadIdx.x / nc) % QG, u_k = threadIdx.x % nc;
  const bool u_ok = threadIdx.x < n_uv;
  const int u_dst = 2 * (((u_plane >> 1) * QG + u_qi) * nc + u_k) + (u_plane & 1);
  const size_t plane = static_cast<size_t>(nqv) * P;
  const size_t u_src = u_plane * plane + static_cast<size_t>(qv[0] + u_qi) * P + u_k +
                       static_cast<size_t>(nc) * (u_plane < 2 ? w1 : w2);
  const bool u_valid = !RX && u_ok && (q0 + u_qi < g.nq);

  // Channel outputs are FP64 totals. With complex64 G the walker terms are first summed in FP32
  // per term type for FLUSH walkers (relative rounding ~2^-24 per step), so the per-walker work
  // is fused multiply-adds only:
  //   xs = a(X0 Y0 + X1 Y1), pz = a(X0 Z1 + X1 Z0), tr = a(X0 Y1 + X1 Y0), dm = du dv, dc = su sv
  //   MAGNETIC = dm - xs, CHARGE = dc - xs, PP = pz, TRANSVERSE = -tr,
  // and the FP64 additions happen once per FLUSH walkers. With complex128 G every walker's
  // channel term goes straight into the totals.
  // 1024-thread tiles of two pairs per thread with precomputed rank-one terms (6x6 batches) run a
  // two-walker pipeline (below): each walker's terms go straight into the FP64 totals, whose
  // registers would otherwise hold FP32 partials.
  constexpr bool kPipe = RX && TT == 1024 && PPT == 2;
  // 1024-thread tiles of four pairs per thread (8x8) with complex64 G: every prefetch load is
  // issued (clamped indices and walker) and the thread's X entries of the current walker are
  // loaded before the next walker's prefetch, so the products wait for X only. Each walker's
  // terms go straight into the FP64 totals, whose registers would otherwise hold FP32 partials.
  constexpr bool kXFirst = !RX && TT == 1024 && PPT == 4 && sizeof(Real) == 4;
  constexpr bool kChunk = sizeof(Real) == 4 && !kPipe && !kXFirst;
  constexpr int FLUSH = 8;
  constexpr int PC = kChunk ? PPT : 1, QC = kChunk ? QG : 1;
  // Registers: with many pairs per thread (or 1024-thread workgroups, 128 VGPRs) X is read in place.
  constexpr bool kPrefetchX = PPT * (TT / 256) <= 4 && !(TT > 256 && PPT > 1);
  bool has_trans = a.has_trans, has_mag = false, has_chg = false;
#pragma unroll
  for (int c = 0; c < NCH; ++c) {
    has_mag |= a.ch[c] == PH_MAGNETIC;
    has_chg |= a.ch[c] == PH_CHARGE;
  }
  if constexpr (kStatic) {
    has_trans = (MASK & (1 << 1)) != 0;
    has_mag = (MASK & (1 << PH_MAGNETIC)) != 0;
    has_chg = (MASK & (1 << PH_CHARGE)) != 0;
  }
  double2 tot[PPT][QG][NCH];
  C xs[PC][QC], pz[PC][QC], tr[PC][QC], dm[PC][QC], dc[PC][QC];
#pragma unroll
  for (int j = 0; j < PPT; ++j)
#pragma unroll
    for (int qi = 0; qi < QG; ++qi)
#pragma unroll
      for (int c = 0; c < NCH; ++c)
        tot[j][qi][c] = double2{0, 0};
#pragma unroll
  for (int j = 0; j < PC; ++j)
#pragma unroll
    for (int qi = 0; qi < QC; ++qi)
      xs[j][qi] = pz[j][qi] = tr[j][qi] = dm[j][qi] = dc[j][qi] = C{0, 0};
  if constexpr (kPipe) {
    // X, Y and Z (both spins) of walkers m + 1 and m + 2 are in flight while walker m is consumed.
    // The elements a walker stages form one flat list, PS per thread, each stored with both spins
    // in one 16-byte write; their source offsets (within the walker's G, below 2^31) and LDS
    // places are fixed for the tile. Every load is issued (idle slots re-read element 0, walkers
    // past the batch re-read the last one) and the loop is unrolled by two, so the loads in
    // flight are counted exactly and staging waits only for the walker it writes.
    constexpr int PS = 4;  // 3 nc^2 <= PS TT (the launcher takes this path for Nc <= 36 only)
    constexpr int NBK = 1 + ((MASK & kY) != 0) + ((MASK & (1 << PP_UP_DOWN)) != 0);
    P2* Xl = Yl + (NBK - 1) * bs;  // after Y and Z
    const int ncc = nc * nc, n_items = NBK * ncc, n_walk = a.n_walkers;
    int psrc[PS], pdst[PS];
    bool puse[PS];
#pragma unroll
    for (int j = 0; j < PS; ++j) {
      const int f = threadIdx.x + TT * j;
      puse[j] = f < n_items;
      const int blk = puse[j] ? f / ncc : 0, idx = f % ncc;
      // Blocks in LDS order: Y (if staged), Z (if staged), X.
      const bool isx = blk == NBK - 1, isy = !isx && needY && blk == 0;
      const size_t boff = isx ? xoff : (isy ? yoff : zoff);
      psrc[j] = puse[j] ? static_cast<int>(boff + (idx % nc) + D * (idx / nc)) : 0;
      pdst[j] = blk * bs + (idx / nc) * ld + idx % nc;
    }
    C pre[2][PS][2];
    auto fetch = [&](int m, C (&pr)[PS][2]) {
      const C* H = a.G + (2 * static_cast<size_t>(m < n_walk ? m : n_walk - 1)) * DD;
#pragma unroll
      for (int j = 0; j < PS; ++j) {
        pr[j][0] = H[psrc[j]];
        pr[j][1] = H[DD + psrc[j]];
      }
    };
    auto stage = [&](const C (&cur)[PS][2]) {
      __syncthreads();
#pragma unroll
      for (int j = 0; j < PS; ++j)
        if (puse[j])
          store_pair(Yl[pdst[j]], cur[j][0], cur[j][1]);
      __syncthreads();
    };
    auto consume = [&](int m) {
      const Real alpha = static_cast<Real>(0.5 * a.desc[2 * m].sign);
#pragma unroll
      for (int j = 0; j < PPT; ++j) {
        if (!ok[j])
          continue;
        const P2 xx = Xl[k2[j] * ld + k1[j]];
        const C x0{alpha * xx.a.x, alpha * xx.a.y}, x1{alpha * xx.b.x, alpha * xx.b.y};
#pragma unroll
        for (int qi = 0; qi < QG; ++qi) {
          if (!qok[qi])
            continue;
          C cross{0, 0}, ppz{0, 0};
          if (needY) {
            const P2 yy = Yl[yi[j][qi]];
            cross = cmul(x0, yy.a);
            cfma(cross, x1, yy.b);
          }
          if (needZ) {
            const P2 zz = Zl[zi[j][qi]];
            ppz = cmul(x0, zz.b);
            cfma(ppz, x1, zz.a);
          }
#pragma unroll
          for (int c = 0; c < NCH; ++c) {
            const int ch = a.ch[c];
            const C t = ch == PP_UP_DOWN ? ppz : C{-cross.x, -cross.y};
            tot[j][qi][c].x += static_cast<double>(t.x);
            tot[j][qi][c].y += static_cast<double>(t.y);
          }
        }
      }
    };
    if (n_walk > 0) {
      fetch(0, pre[0]);
      fetch(1, pre[1]);
      for (int m = 0; m < n_walk; m += 2) {
#pragma unroll
        for (int d = 0; d < 2; ++d) {
          stage(pre[d]);
          fetch(m + d + 2, pre[d]);
          if (m + d < n_walk)
            consume(m + d);
        }
      }
    }
  }
  else {
  // Everything walker m needs is loaded into registers while walker m - 1 is consumed: the
  // staged blocks (each thread stages at most PPT elements per block, nc^2 <= 256 PPT), the
  // rank-one factors and (for few pairs per thread) the thread's own X entries.
  C pre[4][PPT], pre_uv, pre_x[2][kPrefetchX ? PPT : 1];
  auto prefetch = [&](int m) {
    if constexpr (kXFirst) {
      const int mc = m < a.n_walkers ? m : a.n_walkers - 1;
      const C* H0 = a.G + (2 * static_cast<size_t>(mc)) * DD;
      const C* H1 = H0 + DD;
#pragma unroll
      for (int j = 0; j < PPT; ++j) {
        const int idx = threadIdx.x + TT * j;
        const int idc = idx < nc * nc ? idx : 0;
        const size_t go = (idc % nc) + D * (idc / nc);
        if (needY) {
          pre[0][j] = H0[yoff + go];
          pre[1][j] = H1[yoff + go];
        }
        if (needZ) {
          pre[2][j] = H0[zoff + go];
          pre[3][j] = H1[zoff + go];
        }
      }
      pre_uv = a.UV[static_cast<size_t>(mc) * 4 * plane + (u_valid ? u_src : 0)];
      return;
    }
    const C* H0 = a.G + (2 * static_cast<size_t>(m)) * DD;
    const C* H1 = H0 + DD;
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
      const int idx = threadIdx.x + TT * j;
      if (idx < nc * nc) {
        const size_t go = (idx % nc) + D * (idx / nc);
        if (needY) {
          pre[0][j] = H0[yoff + go];
          pre[1][j] = H1[yoff + go];
        }
        if (needZ) {
          pre[2][j] = H0[zoff + go];
          pre[3][j] = H1[zoff + go];
        }
      }
      if constexpr (kPrefetchX) {
        if (ok[j]) {
          const size_t xo = xoff + k1[j] + D * k2[j];
          pre_x[0][j] = H0[xo];
          pre_x[1][j] = H1[xo];
        }
      }
    }
    if (u_valid)
      pre_uv = a.UV[static_cast<size_t>(m) * 4 * plane + u_src];
  };
  if (a.n_walkers > 0)
    prefetch(0);
  for (int m = 0; m < a.n_walkers; ++m) {
    __syncthreads();
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
      const int idx = threadIdx.x + TT * j;
      if (idx < nc * nc) {
        const int li = (idx / nc) * ld + (idx % nc);
        if (needY)
          store_pair(Yl[li], pre[0][j], pre[1][j]);
        if (needZ)
          store_pair(Zl[li], pre[2][j], pre[3][j]);
      }
    }
    if (u_valid)
      reinterpret_cast<C*>(Ul)[u_dst] = pre_uv;
    C xw[2][kPrefetchX ? PPT : 1];
    if constexpr (kPrefetchX) {
#pragma unroll
      for (int j = 0; j < PPT; ++j) {
        xw[0][j] = pre_x[0][j];
        xw[1][j] = pre_x[1][j];
      }
    }
    __syncthreads();
    const Real alpha = static_cast<Real>(0.5 * a.desc[2 * m].sign);
    const C* G0m = a.G + (2 * static_cast<size_t>(m)) * DD;
    C xr[kXFirst ? PPT : 1][2];
    if constexpr (kXFirst) {
#pragma unroll
      for (int j = 0; j < PPT; ++j) {
        const size_t xo = xoff + k1[j] + D * k2[j];  // pair 0 for idle pairs
        xr[j][0] = G0m[xo];
        xr[j][1] = G0m[DD + xo];
      }
      prefetch(m + 1);
    }
    else if (m + 1 < a.n_walkers) {
      prefetch(m + 1);
    }
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
      if (!ok[j])
        continue;
      C x0, x1;
      if constexpr (kXFirst) {
        x0 = xr[j][0];
        x1 = xr[j][1];
      }
      else if constexpr (kPrefetchX) {
        x0 = xw[0][j];
        x1 = xw[1][j];
      }
      else {
        const size_t xo = xoff + k1[j] + D * k2[j];
        x0 = G0m[xo];
        x1 = G0m[DD + xo];
      }
      x0.x *= alpha;
      x0.y *= alpha;
      x1.x *= alpha;
      x1.y *= alpha;
#pragma unroll
      for (int qi = 0; qi < QG; ++qi) {
        if (!qok[qi])
          continue;
        const P2* uq = Ul + qi * nc;
        if constexpr (kChunk) {
          const int jj = kChunk ? j : 0, qq = kChunk ? qi : 0;
          if (needY) {
            const P2 yy = Yl[yi[j][qi]];
            const C y0 = yy.a, y1 = yy.b;
            cfma(xs[jj][qq], x0, y0);
            cfma(xs[jj][qq], x1, y1);
            if (has_trans) {
              cfma(tr[jj][qq], x0, y1);
              cfma(tr[jj][qq], x1, y0);
            }
          }
          if (needZ) {
            const P2 zz = Zl[zi[j][qi]];
            cfma(pz[jj][qq], x0, zz.b);
            cfma(pz[jj][qq], x1, zz.a);
          }
          if (!RX && (has_mag || has_chg)) {
            const P2 uu = uq[k1[j]], vv = uq[QG * nc + k2[j]];
            if (has_mag)
              cfma(dm[jj][qq], uu.a, vv.a);
            if (has_chg)
              cfma(dc[jj][qq], uu.b, vv.b);
          }
        }
        else {
          C cross{0, 0}, ppz{0, 0}, trans{0, 0};
          if (needY) {
            const C y0 = Yl[yi[j][qi]].a, y1 = Yl[yi[j][qi]].b;  // members: no 32-byte temporary
            cross = cmul(x0, y0);
            cfma(cross, x1, y1);
            if (has_trans) {
              trans = cmul(x0, y1);
              cfma(trans, x1, y0);
            }
          }
          if (needZ) {
            ppz = cmul(x0, Zl[zi[j][qi]].b);
            cfma(ppz, x1, Zl[zi[j][qi]].a);
          }
#pragma unroll
          for (int c = 0; c < NCH; ++c) {
            const int ch = a.ch[c];
            C t;
            if (ch == PH_MAGNETIC) {
              t = csub(C{0, 0}, cross);
              if (!RX)
                cfma(t, uq[k1[j]].a, uq[QG * nc + k2[j]].a);
            }
            else if (ch == PH_CHARGE) {
              t = csub(C{0, 0}, cross);
              if (!RX)
                cfma(t, uq[k1[j]].b, uq[QG * nc + k2[j]].b);
            }
            else if (ch == PP_UP_DOWN)
              t = ppz;
            else
              t = C{-trans.x, -trans.y};
            tot[j][qi][c].x += static_cast<double>(t.x);
            tot[j][qi][c].y += static_cast<double>(t.y);
          }
        }
      }
    }
    if constexpr (kChunk) {
      if ((m + 1) % FLUSH == 0 || m + 1 == a.n_walkers) {
#pragma unroll
        for (int j = 0; j < PC; ++j)
#pragma unroll
          for (int qi = 0; qi < QC; ++qi) {
#pragma unroll
            for (int c = 0; c < NCH; ++c) {
              const int ch = a.ch[c];
              C t;
              if (ch == PH_MAGNETIC)
                t = csub(dm[j][qi], xs[j][qi]);
              else if (ch == PH_CHARGE)
                t = csub(dc[j][qi], xs[j][qi]);
              else if (ch == PP_UP_DOWN)
                t = pz[j][qi];
              else
                t = C{-tr[j][qi].x, -tr[j][qi].y};
              tot[j][qi][c].x += static_cast<double>(t.x);
              tot[j][qi][c].y += static_cast<double>(t.y);
            }
            xs[j][qi] = pz[j][qi] = tr[j][qi] = dm[j][qi] = dc[j][qi] = C{0, 0};
          }
      }
    }
  }
  }
  auto store = [&](int c, size_t idx, double2 v) {
    if constexpr (RX) {
      if (a.r1_slot[c] >= 0) {
        const size_t plane = static_cast<size_t>(nqv) * P * P;
        const float2 r = static_cast<const float2*>(a.R)[a.r1_slot[c] * plane + idx];
        v.x += static_cast<double>(r.x);
        v.y += static_cast<double>(r.y);
      }
    }
    g4_store_t(a.g4[c], idx, v, a.sum_mode);
  };
  auto g4_index = [&](int ka, int wa, int kb, int wb, int qvi) {
    return ka + static_cast<size_t>(nc) *
                    (wa + static_cast<size_t>(nw) * (kb + static_cast<size_t>(nc) * (wb + static_cast<size_t>(nw) * qvi)));
  };
#pragma unroll
  for (int j = 0; j < PPT; ++j) {
    if (!ok[j])
      continue;
#pragma unroll
    for (int qi = 0; qi < QG; ++qi) {
      if (!qok[qi])
        continue;
      const size_t idx = g4_index(k1[j], w1, k2[j], w2, qv[qi]);
#pragma unroll
      for (int c = 0; c < NCH; ++c)
        store(c, idx, tot[j][qi][c]);
    }
  }
  if (a.sym && w1 != w2) {
    // Mirror tile (w2, w1): G4(k2, w2, k1, w1) = G4(k1, w1, k2, w2), transposed through LDS (the
    // staging area is free now) so that the stores run along k1 of the mirror tile.
    double2* tl = reinterpret_cast<double2*>(lds);
#pragma unroll
    for (int qi = 0; qi < QG; ++qi) {
      if (!qok[qi])
        continue;
#pragma unroll
      for (int c = 0; c < NCH; ++c) {
        __syncthreads();
#pragma unroll
        for (int j = 0; j < PPT; ++j)
          if (ok[j])
            tl[k2[j] + nc * k1[j]] = tot[j][qi][c];
        __syncthreads();
#pragma unroll
        for (int j = 0; j < PPT; ++j)
          if (ok[j])
            store(c, g4_index(k1[j], w2, k2[j], w1, qv[qi]), tl[k1[j] + nc * k2[j]]);
      }
    }
  }
}

namespace {

template <typename Real, int TT, int PPT, int QG, int NCH, int MASK>
void launch_tiled_fixed(TiledArgs<Real> a, size_t lds, hipStream_t st) {
  const unsigned tiles = tile_schedule(a);
  constexpr bool kRX = std::is_same<Real, float>::value && MASK != 0 &&
                       (MASK & ((1 << PH_MAGNETIC) | (1 << PH_CHARGE))) != 0;
  if constexpr (kRX) {
    if (a.R) {
      launch_g4_rank1(a, st);
      if constexpr (TT == 1024 && PPT == 2) {
        // The pipelined tile stages X beside Y and Z (both spins each).
        constexpr int kY = (1 << PH_MAGNETIC) | (1 << PH_CHARGE) | (1 << PH_TRANSVERSE);
        constexpr int NBK = 1 + ((MASK & kY) != 0) + ((MASK & (1 << PP_UP_DOWN)) != 0);
        lds = sizeof(Cplx<Real>) * 2 * NBK * static_cast<size_t>(a.g.nc) * (a.g.nc + 1);
      }
      hipLaunchKernelGGL((g4_tiled_kernel<Real, TT, PPT, QG, NCH, MASK, true>), dim3(tiles), dim3(TT), lds, st, a);
      return;
    }
  }
  hipLaunchKernelGGL((g4_tiled_kernel<Real, TT, PPT, QG, NCH, MASK, false>), dim3(tiles), dim3(TT), lds, st, a);
}

// Channel sets of the measured workloads get their own instantiations (registers sized for the
// terms they need); any other set reads the channel flags at run time.
template <typename Real, int TT, int PPT, int QG>
bool launch_tiled_nch(const TiledArgs<Real>& a, size_t lds, hipStream_t st) {
  int mask = 0;
  for (int c = 0; c < a.n_ch; ++c)
    mask |= 1 << a.ch[c];
  constexpr int M = 1 << PH_MAGNETIC, Q = 1 << PH_CHARGE, P = 1 << PP_UP_DOWN;
  switch (mask) {
    case M | Q | P: launch_tiled_fixed<Real, TT, PPT, QG, 3, M | Q | P>(a, lds, st); return true;
    case M | P: launch_tiled_fixed<Real, TT, PPT, QG, 2, M | P>(a, lds, st); return true;
    case Q: launch_tiled_fixed<Real, TT, PPT, QG, 1, Q>(a, lds, st); return true;
    case M: launch_tiled_fixed<Real, TT, PPT, QG, 1, M>(a, lds, st); return true;
    default: break;
  }
  switch (a.n_ch) {
    case 1: launch_tiled_fixed<Real, TT, PPT, QG, 1, 0>(a, lds, st); return true;
    case 2: launch_tiled_fixed<Real, TT, PPT, QG, 2, 0>(a, lds, st); return true;
    case 3: launch_tiled_fixed<Real, TT, PPT, QG, 3, 0>(a, lds, st); return true;
    case 4: launch_tiled_fixed<Real, TT, PPT, QG, 4, 0>(a, lds, st); return true;
    default: return false;
  }
}

// Workgroup shape: 256 threads while a thread owns at most 2 of the nc^2 (k1, k2) pairs,
// 1024 threads (at most 4 pairs each) for larger clusters.
struct TileShape {
  int threads, ppt;
};
TileShape tile_shape(int nc) {
  const int pairs = nc * nc;
  if (pairs <= 256)
    return {256, 1};
  if (pairs <= 512)
    return {256, 2};
  if (pairs <= 1024)
    return {1024, 1};
  if (pairs <= 2048)
    return {1024, 2};
  if (pairs <= 4096)
    return {1024, 4};
  return {0, 0};
}

}  // namespace

int tiled_pairs_per_thread(int nc) { return tile_shape(nc).ppt; }

// Tile (w1, w2) of G4 reads the G blocks X = (w1, w2), Y = (w2 + v, w1 + v) and
// Z = (v + nw - 1 - w1, v + nw - 1 - w2) (extended frequencies). The transpose of a tile holds its
// Y block as X, its mirror (w -> nw - 1 - w) its Z block, and the neighbour along the diagonal
// the next v. So a group is `strip` consecutive positions on the diagonals w1 - w2 = +d and -d
// together with their mirror images; the workgroups resident on one XCD at a time cover a few
// groups and find most of their blocks in its L2.
std::vector<int> diagonal_tile_order(int nw, int strip) {
  std::vector<int> out;
  out.reserve(static_cast<size_t>(nw) * nw);
  std::vector<char> seen(static_cast<size_t>(nw) * nw, 0);
  for (int d = 0; d < nw; ++d)
    for (int p0 = 0; p0 < nw - d; p0 += strip)
      for (int sgn = 0; sgn < 2; ++sgn) {
        if (d == 0 && sgn == 1)
          continue;
        for (int p = p0; p < std::min(p0 + strip, nw - d); ++p)
          for (int mir = 0; mir < 2; ++mir) {
            int w2 = p, w1 = p + d;
            if (sgn)
              std::swap(w1, w2);
            if (mir) {
              w1 = nw - 1 - w1;
              w2 = nw - 1 - w2;
            }
            char& s = seen[w1 + static_cast<size_t>(nw) * w2];
            if (s)
              continue;
            s = 1;
            out.push_back(w1 | (w2 << 16));
          }
      }
  return out;
}

size_t tiled_lds_bytes(const Geo& g, bool need_y, bool need_z, size_t cplx_bytes, int qg) {
  const size_t bs = static_cast<size_t>(g.nc) * (g.nc + 1);
  return cplx_bytes * (bs * 2 * ((need_y ? 1 : 0) + (need_z ? 1 : 0)) + 4 * static_cast<size_t>(qg) * g.nc);
}

template <typename Real>
bool launch_g4_tiled(TiledArgs<Real> a, hipStream_t st) {
  const Geo& g = a.g;
  const TileShape ts = tile_shape(g.nc);
  const int ppt = ts.ppt;
  if (g.nb != 1 || ppt == 0 || a.n_ch < 1 || a.n_ch > 4)
    return false;
  // Momenta per workgroup: as many as fit 16 FP64 accumulators per thread (and as the
  // instantiated shapes provide: 4 for one pair per thread, 2 for two, else 1).
  int qg = g.nq >= 4 ? 4 : (g.nq >= 2 ? 2 : 1);
  if (ppt == 2 && qg > 2)
    qg = 2;
  if (ppt > 2)
    qg = 1;
  // FP64 accumulators per thread: 16 for 256-thread workgroups, 4 for 1024-thread ones (8 when
  // the batch's rank-one terms come precomputed, which frees the factor staging and partials).
  int mask = 0;
  for (int c = 0; c < a.n_ch; ++c)
    mask |= 1 << a.ch[c];
  constexpr int kM = 1 << PH_MAGNETIC, kQ = 1 << PH_CHARGE, kP = 1 << PP_UP_DOWN;
  // The pipelined 1024 x 2 tile (the rank-one path of that shape) addresses a walker's G with
  // 32-bit offsets and stages at most 4 x 1024 elements per walker (3 nc^2: Nc <= 36).
  if (ts.threads == 1024 && ppt == 2 &&
      (2 * static_cast<size_t>(g.D) * g.D >= (size_t(1) << 31) || 3 * g.nc * g.nc > 4 * 1024))
    a.R = nullptr;
  const bool rx = std::is_same<Real, float>::value && a.R != nullptr &&
                  (mask == (kM | kQ | kP) || mask == (kM | kP) || mask == kQ || mask == kM);
  const int acc_budget = ts.threads == 256 ? 16 : (rx ? 8 : 4);
  while (qg > 1 && ppt * qg * a.n_ch > acc_budget)
    qg /= 2;
  while (a.qg_cap > 0 && qg > a.qg_cap)
    qg /= 2;
  if (ppt * qg * a.n_ch > 16 || 4 * qg * g.nc > ts.threads)
    return false;  // accumulators would not fit in registers / one staged factor per thread
  a.nqg = (g.nq + qg - 1) / qg;
  const size_t lds = tiled_lds_bytes(g, a.need_y, a.need_z, sizeof(Cplx<Real>), qg);
  const dim3 ug((g.nc * g.nw + 255) / 256, g.nq * g.nv, a.n_walkers);
  hipLaunchKernelGGL(uv_kernel<Real>, ug, dim3(256), 0, st, a);
  if constexpr (std::is_same<Real, float>::value) {
    if (a.mfma && ts.threads == 256 && ppt == 1) {
      TiledArgs<float> b = a;
      b.sym = false;  // the matrix-core tile writes no mirror tiles
      if (launch_g4_mfma(b, qg, st))
        return true;
    }
  }
  auto by_qg = [&](auto tt, auto pp) {
    constexpr int TT = decltype(tt)::value, PP = decltype(pp)::value;
    if (qg == 4)
      return launch_tiled_nch<Real, TT, PP, (PP == 1 ? 4 : 1)>(a, lds, st);
    if (qg == 2)
      return launch_tiled_nch<Real, TT, PP, (PP <= 2 ? 2 : 1)>(a, lds, st);
    return launch_tiled_nch<Real, TT, PP, 1>(a, lds, st);
  };
  using I256 = std::integral_constant<int, 256>;
  using I1024 = std::integral_constant<int, 1024>;
  if (ts.threads == 256)
    return ppt == 1 ? by_qg(I256{}, std::integral_constant<int, 1>{})
                    : by_qg(I256{}, std::integral_constant<int, 2>{});
  switch (ppt) {
    case 1: return by_qg(I1024{}, std::integral_constant<int, 1>{});
    case 2: return by_qg(I1024{}, std::integral_constant<int, 2>{});
    case 4: return by_qg(I1024{}, std::integral_constant<int, 4>{});
    default: return false;
  }
}

template bool launch_g4_tiled<double>(TiledArgs<double>, hipStream_t);
template bool launch_g4_tiled<float>(TiledArgs<float>, hipStream_t);

}  // namespace g4acc
