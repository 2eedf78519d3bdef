// Single-band G4 update with the rank-one terms on the matrix cores (v_mfma_f32_16x16x4f32).
//
// Same tile, staging and crossed terms as g4_tiled.hip (tp_accumulator_cpu.hpp:407-666 for
// n_bands == 1), for clusters of up to 16 sites: one (k1, k2) pair per thread of a 256-thread
// workgroup that owns the tile (w1, w2, v, momentum group). The rank-one terms of the magnetic
// and charge channels,
//   R^q(k1, k2) = sum_m du_m(k1) dv_m(k2)      (magnetic; su, sv for charge)
//   du = a (u_0 - u_1), su = a (u_0 + u_1),  u_s = G_s[(k1,w1), (k1+q, w1+v)]
//   dv =    v_0 - v_1,  sv =    v_0 + v_1,   v_s = G_s[(k2+q,w2+v), (k2,w2)]
// are, for each q of the tile, a 16 x 16 complex product over the walkers. Every 4 walkers each
// wave issues four MFMAs per tile it owns (re += ar br - ai bi, im += ar bi + ai br), reading
// the factors (uv_kernel) straight from memory. The crossed terms stay on the vector ALUs:
//   xs = a (X_0 Y_0 + X_1 Y_1),  pz = a (X_0 Z_1 + X_1 Z_0),  tr = a (X_0 Y_1 + X_1 Y_0)
//   MAGNETIC += R_mag - xs,  CHARGE += R_chg - xs,  PP_UP_DOWN += pz,  TRANSVERSE -= tr.
// Both kinds of FP32 sums are added to FP64 totals every 8 walkers (the rank-one totals live in
// LDS, [tile][element][lane] so that a wave's flush touches consecutive 16-byte slots); G4 is read
// and written once per launch.
#include "g4_tile_common.hpp"

namespace g4acc {

namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kFlush = 8;  // walkers per FP32 partial sum (a multiple of the MFMA depth, 4)

// v_mfma_f32_16x16x4f32 lane layout: lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16];
// element r of lane l's result is D[4 (l / 16) + r][l % 16].
__device__ __forceinline__ int mfma_lane(int row, int col) { return col + 16 * (row >> 2); }

template <int QG, int MASK>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
g4_mfma_kernel(TiledArgs<float> a) {
  using C = Cplx<float>;
  constexpr bool kMag = (MASK & (1 << PH_MAGNETIC)) != 0;
  constexpr bool kChg = (MASK & (1 << PH_CHARGE)) != 0;
  constexpr bool kPP = (MASK & (1 << PP_UP_DOWN)) != 0;
  constexpr bool kTr = (MASK & (1 << PH_TRANSVERSE)) != 0;
  constexpr bool kY = kMag || kChg || kTr;
  constexpr int ND = (kMag ? 1 : 0) + (kChg ? 1 : 0);  // rank-one tiles per momentum
  constexpr int NT = QG * ND;                          // rank-one tiles per workgroup
  constexpr int NS = (NT + 3) / 4;                     // rank-one tiles per wave
  static_assert(ND > 0, "the matrix-core form needs a magnetic or charge channel");

  extern __shared__ unsigned char smem_raw[];
  const Geo& g = a.g;
  const int nc = g.nc, nw = g.nw, e = g.ext, P = nc * nw;
  const size_t D = g.D, DD = D * D;
  int w1, w2, z;
  if (!tile_coords(a, w1, w2, z))
    return;
  const int iv = z / a.nqg, q0 = (z % a.nqg) * QG;
  const int nu = a.w_ex[iv];
  const int ld = nc + 1, bs = nc * ld;
  // Staging buffer (Y [spin][b][a] with a fastest, then Z likewise), then the rank-one totals
  // [tile][4][lane].
  C* Yl = reinterpret_cast<C*>(smem_raw);
  C* Zl = Yl + (kY ? 2 * bs : 0);
  double2* Rl = reinterpret_cast<double2*>(Zl + (kPP ? 2 * bs : 0));
  const size_t xoff = nc * static_cast<size_t>(w1 + e) + D * (nc * static_cast<size_t>(w2 + e));
  const size_t yoff = nc * static_cast<size_t>(w2 + nu + e) + D * (nc * static_cast<size_t>(w1 + nu + e));
  const size_t zoff = nc * static_cast<size_t>(nu + nw - 1 - w1 + e) +
                      D * (nc * static_cast<size_t>(nu + nw - 1 - w2 + e));
  const int nqv = g.nq * g.nv;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

  const bool ok = tid < nc * nc;
  const int k1 = ok ? tid % nc : 0, k2 = ok ? tid / nc : 0;
  int yi[QG], zi[QG], qv[QG];
  bool qok[QG];
#pragma unroll
  for (int qi = 0; qi < QG; ++qi) {
    qok[qi] = q0 + qi < g.nq;
    qv[qi] = (qok[qi] ? q0 + qi : 0) + g.nq * iv;
    const int q = a.k_ex[qok[qi] ? q0 + qi : 0];
    yi[qi] = a.k_add[k1 + nc * q] * ld + a.k_add[k2 + nc * q];
    zi[qi] = a.k_sub[k2 + nc * q] * ld + a.k_sub[k1 + nc * q];
  }

  // Rank-one tiles of this wave: tile t = wave + 4 s is momentum t / ND, channel t % ND
  // (magnetic first). For the walker group starting at m0, lane l supplies walker m0 + l / 16
  // and index l % 16 of both factors.
  const int li = lane & 15, lk = lane >> 4;
  const size_t plane = static_cast<size_t>(nqv) * P;
  unsigned ua[NS], vb[NS];  // the factor buffer holds fewer than 2^32 entries (host-checked)
  bool tv[NS], fv[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const int t = wave + 4 * s;
    tv[s] = t < NT;
    const int qi = tv[s] ? t / ND : 0;
    const int pa = (kMag && t % ND == 0) ? 0 : 1;
    ua[s] = static_cast<unsigned>(pa * plane + static_cast<size_t>(qv[qi]) * P + li + nc * w1);
    vb[s] = static_cast<unsigned>((pa + 2) * plane + static_cast<size_t>(qv[qi]) * P + li + nc * w2);
    fv[s] = tv[s] && qok[qi] && li < nc;
    if (tv[s])
#pragma unroll
      for (int r = 0; r < 4; ++r)
        Rl[(t * 4 + r) * 64 + lane] = double2{0, 0};
  }
  f32x4 accr[NS], acci[NS];
  C opa[NS], opb[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    accr[s] = f32x4{0, 0, 0, 0};
    acci[s] = f32x4{0, 0, 0, 0};
  }
  auto load_factors = [&](int m0) {
    const int mw = m0 + lk;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      opa[s] = opb[s] = C{0, 0};
      if (fv[s] && mw < a.n_walkers) {
        const unsigned wo = static_cast<unsigned>(mw * 4 * plane);
        opa[s] = a.UV[wo + ua[s]];
        opb[s] = a.UV[wo + vb[s]];
      }
    }
  };
  auto flush_rank_one = [&]() {
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      if (!tv[s])
        continue;
      double2* dst = Rl + (wave + 4 * s) * 4 * 64 + lane;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        double2 v = dst[64 * r];
        v.x += static_cast<double>(accr[s][r]);
        v.y += static_cast<double>(acci[s][r]);
        dst[64 * r] = v;
      }
      accr[s] = f32x4{0, 0, 0, 0};
      acci[s] = f32x4{0, 0, 0, 0};
    }
  };

  C xs[QG], pz[QG], tr[QG];
  double2 txs[QG], tpz[QG], ttr[QG];
#pragma unroll
  for (int qi = 0; qi < QG; ++qi) {
    xs[qi] = pz[qi] = tr[qi] = C{0, 0};
    txs[qi] = tpz[qi] = ttr[qi] = double2{0, 0};
  }
  // The staged blocks (one element per thread and block) and the thread's X entries of walker
  // m + 1 are loaded while walker m is consumed.
  C pre[4], pre_x[2];
  auto prefetch = [&](int m) {
    const C* H0 = a.G + (2 * static_cast<size_t>(m)) * DD;
    const C* H1 = H0 + DD;
    if (ok) {
      const size_t go = k1 + D * k2;
      if (kY) {
        pre[0] = H0[yoff + go];
        pre[1] = H1[yoff + go];
      }
      if (kPP) {
        pre[2] = H0[zoff + go];
        pre[3] = H1[zoff + go];
      }
      const size_t xo = xoff + go;
      pre_x[0] = H0[xo];
      pre_x[1] = H1[xo];
    }
  };
  if (a.n_walkers > 0) {
    prefetch(0);
    load_factors(0);
  }
  for (int m = 0; m < a.n_walkers; ++m) {
    __syncthreads();
    if (ok) {
      const int li2 = k2 * ld + k1;
      if (kY) {
        Yl[li2] = pre[0];
        Yl[bs + li2] = pre[1];
      }
      if (kPP) {
        Zl[li2] = pre[2];
        Zl[bs + li2] = pre[3];
      }
    }
    C x0 = pre_x[0], x1 = pre_x[1];
    __syncthreads();
    if (m + 1 < a.n_walkers)
      prefetch(m + 1);
    if (m % 4 == 0) {
      // Rank-one products of walkers m .. m + 3 (missing walkers supply zero factors).
      // Independent accumulators back to back, so no MFMA waits on its predecessor's result.
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        if (!tv[s])
          continue;
        accr[s] = __builtin_amdgcn_mfma_f32_16x16x4f32(opa[s].x, opb[s].x, accr[s], 0, 0, 0);
        acci[s] = __builtin_amdgcn_mfma_f32_16x16x4f32(opa[s].x, opb[s].y, acci[s], 0, 0, 0);
      }
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        if (!tv[s])
          continue;
        accr[s] = __builtin_amdgcn_mfma_f32_16x16x4f32(-opa[s].y, opb[s].y, accr[s], 0, 0, 0);
        acci[s] = __builtin_amdgcn_mfma_f32_16x16x4f32(opa[s].y, opb[s].x, acci[s], 0, 0, 0);
      }
      if (m + 4 < a.n_walkers)
        load_factors(m + 4);
    }
    // The groups issued at walkers m - 6 and m - 2 are complete by now: fold them into FP64.
    if (m % kFlush == kFlush - 2)
      flush_rank_one();
    if (ok) {
      const float alpha = static_cast<float>(0.5 * a.desc[2 * m].sign);
      x0.x *= alpha;
      x0.y *= alpha;
      x1.x *= alpha;
      x1.y *= alpha;
#pragma unroll
      for (int qi = 0; qi < QG; ++qi) {
        if (!qok[qi])
          continue;
        if (kY) {
          const C y0 = Yl[yi[qi]], y1 = Yl[bs + yi[qi]];
          if (kMag || kChg) {
            cfma(xs[qi], x0, y0);
            cfma(xs[qi], x1, y1);
          }
          if (kTr) {
            cfma(tr[qi], x0, y1);
            cfma(tr[qi], x1, y0);
          }
        }
        if (kPP) {
          const C z0 = Zl[zi[qi]], z1 = Zl[bs + zi[qi]];
          cfma(pz[qi], x0, z1);
          cfma(pz[qi], x1, z0);
        }
      }
      if ((m + 1) % kFlush == 0 || m + 1 == a.n_walkers) {
#pragma unroll
        for (int qi = 0; qi < QG; ++qi) {
          if (kMag || kChg) {
            txs[qi].x += static_cast<double>(xs[qi].x);
            txs[qi].y += static_cast<double>(xs[qi].y);
          }
          if (kPP) {
            tpz[qi].x += static_cast<double>(pz[qi].x);
            tpz[qi].y += static_cast<double>(pz[qi].y);
          }
          if (kTr) {
            ttr[qi].x += static_cast<double>(tr[qi].x);
            ttr[qi].y += static_cast<double>(tr[qi].y);
          }
          xs[qi] = pz[qi] = tr[qi] = C{0, 0};
        }
      }
    }
  }
  flush_rank_one();  // groups issued after the last flush
  __syncthreads();   // every wave's rank-one totals are in LDS
  if (!ok)
    return;
  const int rl = mfma_lane(k1, k2), rr = k1 & 3;
#pragma unroll
  for (int qi = 0; qi < QG; ++qi) {
    if (!qok[qi])
      continue;
    const size_t idx = k1 + static_cast<size_t>(nc) *
                                (w1 + static_cast<size_t>(nw) *
                                          (k2 + static_cast<size_t>(nc) *
                                                    (w2 + static_cast<size_t>(nw) * qv[qi])));
    for (int c = 0; c < a.n_ch; ++c) {
      const int ch = a.ch[c];
      double2 v;
      if (ch == PH_MAGNETIC || ch == PH_CHARGE) {
        const int t = qi * ND + ((ch == PH_CHARGE && kMag) ? 1 : 0);
        const double2 r = Rl[(t * 4 + rr) * 64 + rl];
        v = double2{r.x - txs[qi].x, r.y - txs[qi].y};
      }
      else if (ch == PP_UP_DOWN)
        v = tpz[qi];
      else
        v = double2{-ttr[qi].x, -ttr[qi].y};
      g4_store_t(a.g4[c], idx, v, a.sum_mode);
    }
  }
}

template <int QG, int MASK>
void launch_mfma_fixed(TiledArgs<float> a, hipStream_t st) {
  constexpr int ND = ((MASK >> PH_MAGNETIC) & 1) + ((MASK >> PH_CHARGE) & 1);
  const Geo& g = a.g;
  const size_t bs = static_cast<size_t>(g.nc) * (g.nc + 1);
  const size_t lds = sizeof(Cplx<float>) * bs * 2 *
                         (((MASK & ((1 << PH_MAGNETIC) | (1 << PH_CHARGE) | (1 << PH_TRANSVERSE))) ? 1 : 0) +
                          ((MASK & (1 << PP_UP_DOWN)) ? 1 : 0)) +
                     sizeof(double2) * 4 * 64 * QG * ND;
  const unsigned tiles = tile_schedule(a);
  hipLaunchKernelGGL((g4_mfma_kernel<QG, MASK>), dim3(tiles), dim3(256), lds, st, a);
}

template <int QG>
bool launch_mfma_q(const TiledArgs<float>& a, hipStream_t st) {
  int mask = 0;
  for (int c = 0; c < a.n_ch; ++c)
    mask |= 1 << a.ch[c];
  constexpr int M = 1 << PH_MAGNETIC, Q = 1 << PH_CHARGE, P = 1 << PP_UP_DOWN;
  switch (mask) {
    case M | Q | P: launch_mfma_fixed<QG, M | Q | P>(a, st); return true;
    case M | P: launch_mfma_fixed<QG, M | P>(a, st); return true;
    case M | Q: launch_mfma_fixed<QG, M | Q>(a, st); return true;
    case Q | P: launch_mfma_fixed<QG, Q | P>(a, st); return true;
    case M: launch_mfma_fixed<QG, M>(a, st); return true;
    case Q: launch_mfma_fixed<QG, Q>(a, st); return true;
    default: return false;
  }
}

}  // namespace

bool launch_g4_mfma(const TiledArgs<float>& a, int qg, hipStream_t st) {
  const Geo& g = a.g;
  const size_t uv_entries = static_cast<size_t>(a.n_walkers) * 4 * g.nq * g.nv * g.nc * g.nw;
  if (g.nb != 1 || g.nc > 16 || uv_entries >= (static_cast<size_t>(1) << 32))
    return false;
  switch (qg) {
    case 1: return launch_mfma_q<1>(a, st);
    case 2: return launch_mfma_q<2>(a, st);
    case 4: return launch_mfma_q<4>(a, st);
    default: return false;
  }
}

}  // namespace g4acc
