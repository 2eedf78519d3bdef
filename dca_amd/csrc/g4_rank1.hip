// Rank-one G4 terms of a batch on the matrix cores (single band, complex64 G).
//
// For MAGNETIC and CHARGE the G4 update of one walker holds a term that factorises over the two
// momenta (tp_accumulator_cpu.hpp:491-575, the s1/s2 sums of G(k1,k1+q)G(k2+q,k2)):
//     R_c(p1, p2; q, v) += U_c[m](p1; q, v) V_c[m](p2; q, v),   p = k + Nc w,
// with U/V the factors uv_kernel writes: MAGNETIC a(u0 - u1), v0 - v1; CHARGE a(u0 + u1), v0 + v1.
// Summed over the walkers of a batch this is one complex GEMM with K = walkers per (channel, q, v):
// R_c = V_c^T U_c. This kernel computes it with v_mfma_f32_16x16x4f32 (four walkers per step,
// four real products per complex step), adds every eight walkers' FP32 products to FP64 sums, and
// stores R once as complex64 in G4 order (p1 fastest, then p2, then (q, v)), so the tile kernel
// adds it to its FP64 totals at its single G4 store instead of forming the rank-one products per
// walker and entry on the vector ALUs.
//
// A wave owns 16 rows p1 by 64 columns p2 (four 16 x 16 tiles sharing the U operand). Lane l
// supplies V[m0 + l/16][p2 + l%16] (A operand) and U[m0 + l/16][p1 + l%16] (B operand) and holds
// R(p2 = p2_t + 4(l/16) + i, p1 = p1_0 + l%16), i = 0..3, so a store writes 16 consecutive p1.
#include "g4acc_common.hpp"
#include "g4acc_kernels.hpp"

namespace g4acc {

typedef float r1_f32x4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256) g4_rank1_kernel(const float2* __restrict__ UV, int n_walkers,
                                                      int P, int nqv, const int* __restrict__ planes,
                                                      float2* __restrict__ R) {
  constexpr int NT = 4;  // 16 x 16 tiles per wave along p2
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int c = blockIdx.z / nqv, qv = blockIdx.z % nqv;
  const int pu = planes[c];  // U plane; V is plane pu + 2
  const int p1 = blockIdx.x * 64 + wave * 16 + (lane & 15);
  const int p2_0 = blockIdx.y * 64;
  const int kl = lane >> 4;
  const bool p1_ok = p1 < P;
  const size_t row = static_cast<size_t>(nqv) * P;  // between planes of one walker
  const float2* Ub = UV + static_cast<size_t>(pu) * row + static_cast<size_t>(qv) * P;
  const float2* Vb = Ub + 2 * row;
  const size_t wstride = 4 * row;  // between walkers

  double dre[NT][4], dim[NT][4];
  r1_f32x4 fre[NT], fim[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    fre[t] = fim[t] = r1_f32x4{0, 0, 0, 0};
#pragma unroll
    for (int i = 0; i < 4; ++i)
      dre[t][i] = dim[t][i] = 0.0;
  }
  auto load = [&](int m0, float2& u, float2 (&v)[NT]) {
    const int m = m0 + kl;
    const bool ok = m < n_walkers;
    const size_t mo = static_cast<size_t>(ok ? m : 0) * wstride;
    u = ok && p1_ok ? Ub[mo + p1] : float2{0, 0};
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int p2 = p2_0 + 16 * t + (lane & 15);
      v[t] = ok && p2 < P ? Vb[mo + p2] : float2{0, 0};
    }
  };
  float2 u, v[NT];
  load(0, u, v);
  for (int m0 = 0; m0 < n_walkers; m0 += 4) {
    const float2 cu = u;
    float2 cv[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t)
      cv[t] = v[t];
    if (m0 + 4 < n_walkers)
      load(m0 + 4, u, v);
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      // R += V U = (Vx Ux - Vy Uy) + i (Vy Ux + Vx Uy)
      fre[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(cv[t].x, cu.x, fre[t], 0, 0, 0);
      fim[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(cv[t].y, cu.x, fim[t], 0, 0, 0);
      fre[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(-cv[t].y, cu.y, fre[t], 0, 0, 0);
      fim[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(cv[t].x, cu.y, fim[t], 0, 0, 0);
    }
    if ((m0 + 4) % 8 == 0 || m0 + 4 >= n_walkers) {  // eight walkers per FP32 partial sum
#pragma unroll
      for (int t = 0; t < NT; ++t) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          dre[t][i] += static_cast<double>(fre[t][i]);
          dim[t][i] += static_cast<double>(fim[t][i]);
        }
        fre[t] = fim[t] = r1_f32x4{0, 0, 0, 0};
      }
    }
  }
  if (!p1_ok)
    return;
  float2* Rc = R + static_cast<size_t>(blockIdx.z) * P * P;  // (c, qv) plane
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int p2 = p2_0 + 16 * t + 4 * kl + i;
      if (p2 < P)
        Rc[static_cast<size_t>(p2) * P + p1] = float2{static_cast<float>(dre[t][i]), static_cast<float>(dim[t][i])};
    }
}

void launch_g4_rank1(const TiledArgs<float>& a, hipStream_t st) {
  const int P = a.g.nc * a.g.nw, nqv = a.g.nq * a.g.nv;
  const dim3 grid((P + 63) / 64, (P + 63) / 64, a.n_r1 * nqv);
  hipLaunchKernelGGL(g4_rank1_kernel, grid, dim3(256), 0, st, a.UV, a.n_walkers, P, nqv, a.r1_planes,
                     static_cast<float2*>(a.R));
}

}  // namespace g4acc
