// G4 update kernels: G4(b1,b2,b3,b4,k1,w1,k2,w2,k_ex,w_ex) += sum over the walkers of a batch of
// sign/2 * (products of two G values), per channel, as in the reference's spin-symmetric
// updateG4 (tp_accumulator_cpu.hpp:407-666):
//
//   MAGNETIC  (:491-546)  a [(Ga_0 - Ga_1)(b2,b1) (Gb_0 - Gb_1)(b3,b4)
//                            - sum_s G_s(b3,b1; k1,k2; w1,w2) G_s(b2,b4; k2+q,k1+q; w2+v,w1+v)]
//   CHARGE    (:548-575)  a [(Ga_0 + Ga_1)(b2,b1) (Gb_0 + Gb_1)(b3,b4)
//                            - sum_s G_s(b1,b4; k1,k2; w1,w2) G_s(b2,b3; k2+q,k1+q; w2+v,w1+v)]
//   PP_UP_DOWN(:628-661)  a sum_s G_s(b1,b3; k1,k2; w1,w2) G_{1-s}(b2,b4; q-k1,q-k2; v-w1,v-w2)
//   TRANSVERSE(:454-489) -a sum_s G_s(b2,b4; k1,k2; w1,w2) G_{1-s}(b3,b1; k2+q,k1+q; w2+v,w1+v)
// with a = sign/2, Ga_s = G_s(k1, k1+q; w1, w1+v), Gb_s = G_s(k2+q, k2; w2+v, w2), k+q = add(k,q),
// q-k = subtract(k,q), w+v -> w + v and v-w -> v + nw - 1 - w (compact indices, shifted by the
// extension into the EXTENDED domain of G).
//
// Models that are not spin symmetric (SS = false; the spin is part of the band index) read spin
// sector 0 only, with the whole sign (tp_accumulator_cpu.hpp:668-757):
//   PP_UP_DOWN(:670-705)  s [G_0(b1,b3; k1,k2; w1,w2) G_0(b2,b4; q-k1,q-k2; v-w1,v-w2)
//                            - G_0(b1,b4; k1,q-k2; w1,v-w2) G_0(b2,b3; q-k1,k2; v-w1,w2)]
//   MAGNETIC  (:707-750)  s [G_0(b2,b1; k1,k1+q; w1,w1+v) G_0(b3,b4; k2+q,k2; w2+v,w2)
//                            - G_0(b3,b1; k2+q,k1+q; w2+v,w1+v) G_0(b2,b4; k1,k2; w1,w2)]
// The sign s is a complex phase for complex models (SignType, type_help.hpp:299-312).
//
// Each launch sums all walkers of the batch for an entry in FP64 registers and touches the entry
// of G4 once (FP64 read-modify-write, or FP32 value + FP32 Kahan compensation).
#include "g4acc_common.hpp"
#include "g4acc_kernels.hpp"

namespace g4acc {

constexpr int PH_TRANSVERSE = 1, PH_MAGNETIC = 2, PH_CHARGE = 3, PP_UP_DOWN = 4;

template <typename Real>
struct GView {
  const Cplx<Real>* G;
  int nb, nc, D;
  size_t slot_stride;
  __device__ __forceinline__ double2 operator()(int slot, int b1, int b2, int k1, int k2, int w1e,
                                                int w2e) const {
    const size_t row = b1 + static_cast<size_t>(nb) * (k1 + static_cast<size_t>(nc) * w1e);
    const size_t col = b2 + static_cast<size_t>(nb) * (k2 + static_cast<size_t>(nc) * w2e);
    return to_d(G[slot * slot_stride + row + D * col]);
  }
};

__device__ __forceinline__ void g4_store(void* g4, size_t i, double2 acc, int sum_mode) {
  if (sum_mode == SUM_FP64) {
    double2* p = static_cast<double2*>(g4) + i;
    double2 v = *p;
    v.x += acc.x;
    v.y += acc.y;
    *p = v;
  }
  else {
    // (value, compensation) pairs; the represented sum is value - compensation.
    float4* p = static_cast<float4*>(g4) + i;
    float4 v = *p;
    const float hx = static_cast<float>(acc.x), hy = static_cast<float>(acc.y);
    const float lx = static_cast<float>(acc.x - hx), ly = static_cast<float>(acc.y - hy);
    float yx = hx - v.z, yy = hy - v.w;
    float tx = v.x + yx, ty = v.y + yy;
    v.z = ((tx - v.x) - yx) - lx;
    v.w = ((ty - v.y) - yy) - ly;
    v.x = tx;
    v.y = ty;
    *p = v;
  }
}

template <typename Real, int CH, bool SS>
__global__ void __launch_bounds__(256) g4_generic_kernel(G4Args<Real> a) {
  const Geo& g = a.g;
  const int nb = g.nb, nc = g.nc, nw = g.nw;
  const size_t S = static_cast<size_t>(nb) * nb * nb * nb * nc * nw * nc * nw;
  const size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= S)
    return;
  const int iq = blockIdx.y % g.nq, iv = blockIdx.y / g.nq;
  size_t r = e;
  const int b1 = r % nb; r /= nb;
  const int b2 = r % nb; r /= nb;
  const int b3 = r % nb; r /= nb;
  const int b4 = r % nb; r /= nb;
  const int k1 = r % nc; r /= nc;
  const int w1 = r % nw; r /= nw;
  const int k2 = r % nc; r /= nc;
  const int w2 = static_cast<int>(r);
  const int q = a.k_ex[iq], nu = a.w_ex[iv];
  const int ext = g.ext;
  const int w1e = w1 + ext, w2e = w2 + ext;
  GView<Real> L{a.G, nb, nc, g.D, static_cast<size_t>(g.D) * g.D};

  double2 acc{0, 0};
  for (int m = 0; m < a.n_walkers; ++m) {
    const double f = SS ? 0.5 : 1.0;
    const double2 alpha{f * a.desc[2 * m].sign, f * a.desc[2 * m].sign_im};
    const int s0 = 2 * m, s1 = 2 * m + 1;
    double2 c{0, 0};
    if constexpr (!SS) {
      if constexpr (CH == PP_UP_DOWN) {
        const int qk1 = a.k_sub[k1 + nc * q], qk2 = a.k_sub[k2 + nc * q];
        const int w1r = nu + nw - 1 - w1 + ext, w2r = nu + nw - 1 - w2 + ext;
        c = csub(cmul(L(s0, b1, b3, k1, k2, w1e, w2e), L(s0, b2, b4, qk1, qk2, w1r, w2r)),
                 cmul(L(s0, b1, b4, k1, qk2, w1e, w2r), L(s0, b2, b3, qk1, k2, w1r, w2e)));
      }
      else {  // PH_MAGNETIC
        const int kq1 = a.k_add[k1 + nc * q], kq2 = a.k_add[k2 + nc * q];
        const int w1n = w1 + nu + ext, w2n = w2 + nu + ext;
        c = csub(cmul(L(s0, b2, b1, k1, kq1, w1e, w1n), L(s0, b3, b4, kq2, k2, w2n, w2e)),
                 cmul(L(s0, b3, b1, kq2, kq1, w2n, w1n), L(s0, b2, b4, k1, k2, w1e, w2e)));
      }
    }
    else if constexpr (CH == PH_MAGNETIC || CH == PH_CHARGE) {
      const int kq1 = a.k_add[k1 + nc * q], kq2 = a.k_add[k2 + nc * q];
      const int w1n = w1 + nu + ext, w2n = w2 + nu + ext;
      double2 ga0 = L(s0, b2, b1, k1, kq1, w1e, w1n), ga1 = L(s1, b2, b1, k1, kq1, w1e, w1n);
      double2 gb0 = L(s0, b3, b4, kq2, k2, w2n, w2e), gb1 = L(s1, b3, b4, kq2, k2, w2n, w2e);
      double2 ga, gb;
      if constexpr (CH == PH_MAGNETIC) {
        ga = csub(ga0, ga1);
        gb = csub(gb0, gb1);
      }
      else {
        ga = cadd(ga0, ga1);
        gb = cadd(gb0, gb1);
      }
      c = cmul(ga, gb);
      for (int sl = s0; sl <= s1; ++sl) {
        double2 x, y;
        if constexpr (CH == PH_MAGNETIC) {
          x = L(sl, b3, b1, k1, k2, w1e, w2e);
          y = L(sl, b2, b4, kq2, kq1, w2n, w1n);
        }
        else {
          x = L(sl, b1, b4, k1, k2, w1e, w2e);
          y = L(sl, b2, b3, kq2, kq1, w2n, w1n);
        }
        c = csub(c, cmul(x, y));
      }
    }
    else if constexpr (CH == PP_UP_DOWN) {
      const int qk1 = a.k_sub[k1 + nc * q], qk2 = a.k_sub[k2 + nc * q];
      const int w1r = nu + nw - 1 - w1 + ext, w2r = nu + nw - 1 - w2 + ext;
      c = cadd(cmul(L(s0, b1, b3, k1, k2, w1e, w2e), L(s1, b2, b4, qk1, qk2, w1r, w2r)),
               cmul(L(s1, b1, b3, k1, k2, w1e, w2e), L(s0, b2, b4, qk1, qk2, w1r, w2r)));
    }
    else {  // PH_TRANSVERSE
      const int kq1 = a.k_add[k1 + nc * q], kq2 = a.k_add[k2 + nc * q];
      const int w1n = w1 + nu + ext, w2n = w2 + nu + ext;
      c = cadd(cmul(L(s0, b2, b4, k1, k2, w1e, w2e), L(s1, b3, b1, kq2, kq1, w2n, w1n)),
               cmul(L(s1, b2, b4, k1, k2, w1e, w2e), L(s0, b3, b1, kq2, kq1, w2n, w1n)));
      c.x = -c.x;
      c.y = -c.y;
    }
    acc.x = fma(alpha.x, c.x, fma(-alpha.y, c.y, acc.x));
    acc.y = fma(alpha.x, c.y, fma(alpha.y, c.x, acc.y));
  }
  g4_store(a.g4, static_cast<size_t>(blockIdx.y) * S + e, acc, a.sum_mode);
}

template <typename Real>
void launch_g4(const G4Args<Real>& a, hipStream_t st) {
  const Geo& g = a.g;
  const size_t S = static_cast<size_t>(g.nb) * g.nb * g.nb * g.nb * g.nc * g.nw * g.nc * g.nw;
  dim3 grid(static_cast<unsigned>((S + 255) / 256), g.nq * g.nv);
  if (!a.spin_symmetric) {  // the host admits only the pp and magnetic channels here
    if (a.channel == PP_UP_DOWN)
      hipLaunchKernelGGL((g4_generic_kernel<Real, PP_UP_DOWN, false>), grid, dim3(256), 0, st, a);
    else if (a.channel == PH_MAGNETIC)
      hipLaunchKernelGGL((g4_generic_kernel<Real, PH_MAGNETIC, false>), grid, dim3(256), 0, st, a);
    return;
  }
  switch (a.channel) {
    case PH_MAGNETIC:
      hipLaunchKernelGGL((g4_generic_kernel<Real, PH_MAGNETIC, true>), grid, dim3(256), 0, st, a);
      break;
    case PH_CHARGE:
      hipLaunchKernelGGL((g4_generic_kernel<Real, PH_CHARGE, true>), grid, dim3(256), 0, st, a);
      break;
    case PP_UP_DOWN:
      hipLaunchKernelGGL((g4_generic_kernel<Real, PP_UP_DOWN, true>), grid, dim3(256), 0, st, a);
      break;
    case PH_TRANSVERSE:
      hipLaunchKernelGGL((g4_generic_kernel<Real, PH_TRANSVERSE, true>), grid, dim3(256), 0, st, a);
      break;
    default:
      break;
  }
}
template void launch_g4<double>(const G4Args<double>&, hipStream_t);
template void launch_g4<float>(const G4Args<float>&, hipStream_t);

// ---------------------------------------------------------------------------------------------
__global__ void g4_add_kernel(void* dst, const void* src, size_t n, int sum_mode) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  if (sum_mode == SUM_FP64) {
    const double2 v = static_cast<const double2*>(src)[i];
    g4_store(dst, i, v, SUM_FP64);
  }
  else {
    const float4 v = static_cast<const float4*>(src)[i];
    g4_store(dst, i, double2{static_cast<double>(v.x) - v.z, static_cast<double>(v.y) - v.w},
             SUM_FP32_KAHAN);
  }
}

__global__ void g4_export_kernel(const void* src, double2* dst, size_t n, int sum_mode) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  if (sum_mode == SUM_FP64) {
    dst[i] = static_cast<const double2*>(src)[i];
  }
  else {
    const float4 v = static_cast<const float4*>(src)[i];
    dst[i] = double2{static_cast<double>(v.x) - v.z, static_cast<double>(v.y) - v.w};
  }
}

__global__ void zero16_kernel(double2* dst, size_t n) {
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x)
    dst[i] = double2{0, 0};
}

void launch_g4_add(void* dst, const void* src, size_t n, int sum_mode, hipStream_t st) {
  hipLaunchKernelGGL(g4_add_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, st,
                     dst, src, n, sum_mode);
}
void launch_g4_export(const void* src, double2* dst, size_t n, int sum_mode, hipStream_t st) {
  hipLaunchKernelGGL(g4_export_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, st,
                     src, dst, n, sum_mode);
}
void launch_zero16(void* dst, size_t n, hipStream_t st) {
  const size_t blocks = (n + 255) / 256;
  hipLaunchKernelGGL(zero16_kernel, dim3(static_cast<unsigned>(blocks < 16384 ? blocks : 16384)),
                     dim3(256), 0, st, static_cast<double2*>(dst), n);
}

// ---------------------------------------------------------------------------------------------
// Reference layout G_(b1,b2,s,k1,k2,w1,w2) <-> device slots 0 (s=0) and 1 (s=1) of one walker,
// Gs[(b2,k1,w1) + D (b1,k2,w2)] = G_(b1,b2,s,k1,k2,w1,w2).
__device__ __forceinline__ void ref_to_dev(size_t i, const Geo& g, int& slot, size_t& off) {
  size_t r = i;
  const int b1 = r % g.nb; r /= g.nb;
  const int b2 = r % g.nb; r /= g.nb;
  const int s = r % 2; r /= 2;
  const int k1 = r % g.nc; r /= g.nc;
  const int k2 = r % g.nc; r /= g.nc;
  const int w1 = r % g.nwE; r /= g.nwE;
  const int w2 = static_cast<int>(r);
  const size_t row = b2 + static_cast<size_t>(g.nb) * (k1 + static_cast<size_t>(g.nc) * w1);
  const size_t col = b1 + static_cast<size_t>(g.nb) * (k2 + static_cast<size_t>(g.nc) * w2);
  slot = s;
  off = row + static_cast<size_t>(g.D) * col;
}

template <typename Real>
__global__ void import_G_kernel(const double2* ref, Cplx<Real>* G, Geo g) {
  const size_t n = 2 * static_cast<size_t>(g.D) * g.D;
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  int slot;
  size_t off;
  ref_to_dev(i, g, slot, off);
  G[slot * static_cast<size_t>(g.D) * g.D + off] = from_d<Cplx<Real>>(ref[i]);
}

template <typename Real>
__global__ void export_G_kernel(const Cplx<Real>* G, double2* ref, Geo g) {
  const size_t n = 2 * static_cast<size_t>(g.D) * g.D;
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  int slot;
  size_t off;
  ref_to_dev(i, g, slot, off);
  ref[i] = to_d(G[slot * static_cast<size_t>(g.D) * g.D + off]);
}

template <typename Real>
void launch_import_G(const double2* ref, Cplx<Real>* G, const Geo& g, hipStream_t st) {
  const size_t n = 2 * static_cast<size_t>(g.D) * g.D;
  hipLaunchKernelGGL(import_G_kernel<Real>, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0,
                     st, ref, G, g);
}
template <typename Real>
void launch_export_G(const Cplx<Real>* G, double2* ref, const Geo& g, hipStream_t st) {
  const size_t n = 2 * static_cast<size_t>(g.D) * g.D;
  hipLaunchKernelGGL(export_G_kernel<Real>, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0,
                     st, G, ref, g);
}
template void launch_import_G<double>(const double2*, double2*, const Geo&, hipStream_t);
template void launch_import_G<float>(const double2*, float2*, const Geo&, hipStream_t);
template void launch_export_G<double>(const double2*, double2*, const Geo&, hipStream_t);
template void launch_export_G<float>(const float2*, double2*, const Geo&, hipStream_t);

}  // namespace g4acc
