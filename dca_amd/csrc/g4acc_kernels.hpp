// Host-side declarations of the g4acc kernel launchers.
#pragma once

#include "g4acc_common.hpp"

namespace g4acc {

template <typename Real>
struct MtoGArgs {
  Geo g;
  const SlotDesc* desc;   // n_slots device descriptors
  int n_slots;            // 2 * walkers
  int n_max;              // vertex capacity per slot
  int chunk;              // slots whose S1..S4 workspaces exist at once
  bool fused;             // use the register-fused single-band kernels
  int tw;                 // frequencies per thread in the fused right kernel (0: default)
  int mfma;               // fused right kernel on the matrix cores (complex64, Nc <= 16)
  // 6x6 and 8x8 clusters: the whole M -> G on the matrix cores with the B, E and Ms workspaces
  // (mtog_large.hip), instead of the register-fused kernels.
  bool large = false;
  // FP32 only: S3, the right site transform and the epilogue in one kernel, E kept in registers
  // and LDS (no E workspace).
  bool right_fused = false;
  void* Ms = nullptr;     // [chunk][n_max][n_max] M in sorted order, working precision
  const int* perm;        // [n_slots][n_max]  rows of M, sorted by the vertices' right labels
  const int* off;         // [n_slots][n_orb + 1]
  Cplx<Real>* T;          // [n_slots][n_max][nwE]
  // Columns of M, sorted by the vertices' left labels, and their phases: the same arrays as
  // perm / off / T unless the vertices are non-density-density (four-stage kernels only).
  const int* perm2;
  const int* off2;
  Cplx<Real>* T2;
  Cplx<Real>* B;          // [chunk][n_max][D]
  Cplx<Real>* A;          // [chunk][n_max][D]
  Cplx<Real>* E;          // [chunk][D][D]
  Cplx<Real>* G;          // [n_slots][D][D]
  const Cplx<Real>* dft;  // [Nc][Nc] (k fastest)
  const double2* bphase;  // [Nc][nb] or nullptr
  const double2* G0;      // (b1, b2, s, k, wE)
  // Square L x L clusters (L = 2, 4, 6, 8): sites are relabelled x + L y and the right site DFT
  // runs as two L-point passes; kmap[kx + L ky] is the caller's momentum index. L = 0: dense DFT.
  int grid_L;
  const int* kmap;
};

// M = diag(exp_v - 1) N for one n x n matrix (column-major, leading dimensions ldn, ldm).
void launch_compute_m(int n, const double* exp_v, const double* N, int ldn, double* M, int ldm,
                      hipStream_t st);

// site_map (or nullptr): the internal label of each caller site (sites sorted by it).
// left = true sorts by the vertices' left labels (where a descriptor has them).
void launch_sort(const SlotDesc* desc, const Geo& g, int n_slots, int n_max, const int* site_map,
                 int* perm, int* off, int* err_flag, hipStream_t st, bool left = false);
// Phases T[slot][w + nwE ii] = exp(i w_w tau_{perm ii}) of the vertices in the order of perm.
template <typename Real>
void launch_phases(const SlotDesc* desc, const Geo& g, int n_slots, int n_max, const int* perm,
                   Cplx<Real>* T, hipStream_t st);

template <typename Real>
void launch_mtog(const MtoGArgs<Real>& a, hipStream_t st);
// Register-fused single-band M -> G (mtog_fused.hip); false if the geometry is not covered.
template <typename Real>
bool launch_mtog_fused(const MtoGArgs<Real>& a, hipStream_t st);
bool mtog_fused_covers(const Geo& g);
// M -> G for single-band 6x6 (NC = 36) and 8x8 (NC = 64) clusters with site labels x + L y, on
// slots [s0, s0 + ns) of a chunk (needs the Ms, B and E workspaces).
template <typename Real, int NC>
void launch_mtog_large(const MtoGArgs<Real>& a, int s0, int ns, hipStream_t st);

enum SumMode { SUM_FP64 = 0, SUM_FP32_KAHAN = 1 };

template <typename Real>
struct G4Args {
  Geo g;
  int channel;                 // FourPointType value
  const Cplx<Real>* G;         // [n_slots][D][D]
  const SlotDesc* desc;        // signs
  int n_walkers;
  const int* k_add;            // device [Nc][Nc]
  const int* k_sub;
  const int* k_ex;             // device [nq]
  const int* w_ex;             // device [nv]
  int sum_mode;
  void* g4;                    // double2[size] (FP64) or float4[size] (value, compensation)
  // false: the updateG4 branch of models that are not spin symmetric (pp and magnetic, spin
  // sector 0 only, tp_accumulator_cpu.hpp:668-757)
  bool spin_symmetric = true;
};

template <typename Real>
void launch_g4(const G4Args<Real>& a, hipStream_t st);

// Single-band, all channels in one launch (g4_tiled.hip). Returns false when the geometry is
// outside what the tiled kernel handles (the caller then uses launch_g4 per channel).
template <typename Real>
struct TiledArgs {
  Geo g;
  const Cplx<Real>* G;         // [n_slots][D][D]
  Cplx<Real>* UV;              // [walkers][4][nq*nv][Nc*nw]
  const SlotDesc* desc;
  int n_walkers;
  int n_ch;
  int ch[4];                   // channel types in G4 order
  void* g4[4];
  const int* k_add;
  const int* k_sub;
  const int* k_ex;
  const int* w_ex;
  int nqg;                     // momentum groups per exchange frequency (set by the launcher)
  int tile_order;              // 0: natural, 1: XCD-aware super-tiles
  int qg_cap;                  // largest momentum group (0: as many as registers allow)
  int super_tile, n_super, n_xcd;  // tile schedule (set by the launcher)
  const int* tile_perm;        // schedule 2: (w1 | w2 << 16) in dispatch order, nw * nw entries
  int sum_mode;
  int mfma;                    // 1: rank-one terms on the matrix cores where a kernel provides it
  // Rank-one terms precomputed for the batch (complex64 G, 256-thread tiles): R[c][q,v][p2][p1] for
  // the n_r1 channels with r1_planes[c] their U plane, and r1_slot[ch] the R plane of G4 channel
  // slot ch (-1: none). R == nullptr: the tile kernel forms them itself.
  void* R;
  int n_r1;
  const int* r1_planes;  // device, n_r1 entries
  int r1_slot[4];
  bool need_y, need_z, has_trans;
  // Transpose symmetry (one exchange momentum q = 0, one exchange frequency v = 0, channels among
  // magnetic, charge and transverse): G4 over ((k1, w1), (k2, w2)) is then symmetric, so only the
  // tiles w1 <= w2 are computed and each also writes its mirror tile (w2, w1).
  bool sym = false;
};

template <typename Real>
bool launch_g4_tiled(TiledArgs<Real> a, hipStream_t st);
// The matrix-core form of the single-band tile kernel (FP32 G, Nc <= 16, one pair per thread);
// false if it does not cover the geometry or channel set. The caller has run uv_kernel and set
// a.nqg for momentum groups of qg.
bool launch_g4_mfma(const TiledArgs<float>& a, int qg, hipStream_t st);
// R of the batch from the rank-one factors (g4_rank1.hip); the caller has run uv_kernel.
void launch_g4_rank1(const TiledArgs<float>& a, hipStream_t st);
// Multiband LDS-tiled G4 update (g4_band.hip); false if the band count, cluster or channel set
// is not covered (the generic kernel then runs).
template <typename Real>
bool launch_g4_band(TiledArgs<Real> a, hipStream_t st);
int tiled_pairs_per_thread(int nc);
// Tile schedule 2: (w1 | w2 << 16) for all nw^2 frequency pairs, in dispatch order.
std::vector<int> diagonal_tile_order(int nw, int strip);
// LDS of one tile workgroup with momentum groups of qg (staged blocks and rank-one factors).
size_t tiled_lds_bytes(const Geo& g, bool need_y, bool need_z, size_t cplx_bytes, int qg);

// Single-particle accumulator (sp_acc.hip): exact FP64 M_r_w of the batch's slots added to acc
// (and M_ij^2's to acc_sqr when squared_too), acc laid out [s][(b1 + nb (b2 + nb r)) n_w + w].
struct SpArgs {
  Geo g;
  const SlotDesc* desc;
  int n_slots, n_max, n_w;     // n_w = 2 * single-particle frequencies
  const int* perm;             // sort of the slots' vertices by caller orbital b + nb r
  const int* off;
  double2* T;                  // [n_slots][n_max][n_w] phases
  double2* C;                  // [n_slots][n_orb][n_orb][n_w] orbital-pair sums
  const int* rsub;             // [Nc][Nc], rsub[a + Nc b] = index(r_b - r_a)
  double2* acc;
  double2* acc_sqr;
  bool squared_too;
};
void launch_sp_accumulate(const SpArgs& a, hipStream_t st);

// dst += src over n complex entries of a G4 buffer in the given sum mode.
void launch_g4_add(void* dst, const void* src, size_t n, int sum_mode, hipStream_t st);
// G4 buffer (any sum mode) -> complex double.
void launch_g4_export(const void* src, double2* dst, size_t n, int sum_mode, hipStream_t st);
// Fill n 16-byte entries with zeros (a kernel, not hipMemsetAsync, so it is safe to capture).
void launch_zero16(void* dst, size_t n, hipStream_t st);
// Reference-layout G (b1,b2,s,k1,k2,w1E,w2E) complex double -> per-slot device layout of one walker.
template <typename Real>
void launch_import_G(const double2* ref, Cplx<Real>* G, const Geo& g, hipStream_t st);
template <typename Real>
void launch_export_G(const Cplx<Real>* G, double2* ref, const Geo& g, hipStream_t st);

}  // namespace g4acc
