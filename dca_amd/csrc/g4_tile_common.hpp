// Device helpers shared by the single-band G4 tile kernels (g4_tiled.hip, g4_mfma.hip).
#ifndef G4ACC_G4_TILE_COMMON_HPP
#define G4ACC_G4_TILE_COMMON_HPP

#include "g4acc_common.hpp"
#include "g4acc_kernels.hpp"

namespace g4acc {

// FourPointType values (four_point_type.hpp); anything else is PARTICLE_HOLE_TRANSVERSE (1).
constexpr int PH_TRANSVERSE = 1, PH_MAGNETIC = 2, PH_CHARGE = 3, PP_UP_DOWN = 4;

// G4[i] += acc, in FP64 or as an FP32 value plus an FP32 Kahan compensation.
__device__ __forceinline__ void g4_store_t(void* g4, size_t i, double2 acc, int sum_mode) {
  if (sum_mode == SUM_FP64) {
    double2* p = static_cast<double2*>(g4) + i;
    double2 v = *p;
    v.x += acc.x;
    v.y += acc.y;
    *p = v;
  }
  else {
    float4* p = static_cast<float4*>(g4) + i;
    float4 v = *p;
    const float hx = static_cast<float>(acc.x), hy = static_cast<float>(acc.y);
    const float lx = static_cast<float>(acc.x - hx), ly = static_cast<float>(acc.y - hy);
    const float yx = hx - v.z, yy = hy - v.w;
    const float tx = v.x + yx, ty = v.y + yy;
    v.z = ((tx - v.x) - yx) - lx;
    v.w = ((ty - v.y) - yy) - ly;
    v.x = tx;
    v.y = ty;
    *p = v;
  }
}

// The (w1, w2, z) tile of this workgroup, z = (exchange frequency, momentum group) with the
// momentum group fastest. Returns false for the padding tiles of the super-tile schedule.
template <typename Real>
__device__ __forceinline__ bool tile_coords(const TiledArgs<Real>& a, int& w1, int& w2, int& z) {
  const int nw = a.g.nw;
  const int nz = a.g.nv * a.nqg;
  if (a.sym) {
    // Transpose symmetry: the nw (nw + 1) / 2 tiles w1 <= w2 in triangular order, w2 slowest.
    const unsigned L = blockIdx.x / nz;
    z = blockIdx.x % nz;
    int r = static_cast<int>((sqrt(8.0 * L + 1.0) - 1.0) * 0.5);
    while (static_cast<unsigned>((r + 1) * (r + 2) / 2) <= L)
      ++r;
    while (static_cast<unsigned>(r * (r + 1) / 2) > L)
      --r;
    w2 = r;
    w1 = static_cast<int>(L) - r * (r + 1) / 2;
    return w2 < nw;
  }
  if (a.tile_perm != nullptr) {
    // Schedule 2: logical tiles in the order of the pair table (exchange frequency and momentum
    // group innermost), each XCD taking a contiguous range.
    unsigned L = blockIdx.x;
    if (a.n_xcd > 1)
      L = (blockIdx.x % a.n_xcd) * (gridDim.x / a.n_xcd) + blockIdx.x / a.n_xcd;
    const int packed = a.tile_perm[L / nz];
    z = L % nz;
    w1 = packed & 0xffff;
    w2 = packed >> 16;
  }
  else if (a.n_xcd > 0) {
    // Workgroups are dealt to the XCDs round-robin by id; give each XCD a contiguous range of
    // logical tiles, ordered by super-tiles of ST x ST (w1, w2) blocks with the exchange
    // frequency and momentum group innermost.
    const int ST = a.super_tile, nsw = a.n_super;
    const unsigned n_tiles = gridDim.x;
    const unsigned xcd = blockIdx.x % a.n_xcd, local = blockIdx.x / a.n_xcd;
    const unsigned L = xcd * (n_tiles / a.n_xcd) + local;
    const unsigned per_super = static_cast<unsigned>(ST) * ST * nz;
    const unsigned sup = L / per_super, rem = L % per_super;
    z = rem % nz;
    w1 = (sup % nsw) * ST + (rem / nz) % ST;
    w2 = (sup / nsw) * ST + (rem / nz) / ST;
  }
  else {
    // Natural order: w1 fastest, then w2, then (v, momentum group).
    w1 = blockIdx.x % nw;
    w2 = (blockIdx.x / nw) % nw;
    z = blockIdx.x / (nw * nw);
  }
  return w1 < nw && w2 < nw;
}

// Grid size of a tile launch; fills in the schedule fields of a.
template <typename Real>
inline unsigned tile_schedule(TiledArgs<Real>& a) {
  const Geo& g = a.g;
  if (a.sym) {
    a.tile_perm = nullptr;
    a.super_tile = a.n_super = 1;
    a.n_xcd = 0;
    return static_cast<unsigned>(g.nw) * (g.nw + 1) / 2 * g.nv * a.nqg;
  }
  if (a.tile_order == 2 && a.tile_perm != nullptr) {
    const unsigned tiles = static_cast<unsigned>(g.nw) * g.nw * g.nv * a.nqg;
    a.super_tile = a.n_super = 1;
    a.n_xcd = tiles % 8 == 0 ? 8 : 1;
    return tiles;
  }
  a.tile_perm = nullptr;
  if (a.tile_order == 1) {
    a.super_tile = g.nw < 8 ? g.nw : 8;
    a.n_super = (g.nw + a.super_tile - 1) / a.super_tile;
    const unsigned tiles =
        static_cast<unsigned>(a.n_super * a.super_tile) * (a.n_super * a.super_tile) * g.nv * a.nqg;
    a.n_xcd = tiles % 8 == 0 ? 8 : 1;
    return tiles;
  }
  a.super_tile = a.n_super = 1;
  a.n_xcd = 0;
  return static_cast<unsigned>(g.nw) * g.nw * g.nv * a.nqg;
}

}  // namespace g4acc

#endif
