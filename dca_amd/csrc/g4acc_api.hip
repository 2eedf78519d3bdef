// C ABI of the g4acc library: handle life cycle, argument checking, host/device staging of the
// walkers' inputs and the launch sequence of one batch. Nothing here throws across the ABI.
#include "../../include/g4acc.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "g4acc_common.hpp"
#include "g4acc_kernels.hpp"

using namespace g4acc;

namespace {

thread_local std::string g_create_error;

size_t workspace_budget_bytes() { return size_t(16) << 30; }

}  // namespace

// Workspace of g4acc_compute_G_noalloc. Each GPU holds its own copy at its own address until the
// process ends. In host code this name is only a stand-in for the symbol: kernels and copies get
// the per-device address that hipGetSymbolAddress returns (scratch_address below).
static __device__ __align__(256) unsigned char scratch[G4ACC_NOALLOC_SCRATCH_BYTES];

namespace {

std::mutex g_scratch_mu;
std::vector<void*> g_scratch_ptr;  // by device ordinal; null until resolved on that device

// The device address of `scratch` on `device`, which must be the current device. Resolved once per
// device (the first g4acc_create there) and kept for the life of the process.
int scratch_address(int device, void** out, std::string& err) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (device >= static_cast<int>(g_scratch_ptr.size()))
    g_scratch_ptr.resize(static_cast<size_t>(device) + 1, nullptr);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return e == hipSuccess ? G4ACC_DEVICE_ERROR : (e == hipErrorOutOfMemory ? G4ACC_OUT_OF_MEMORY
                                                                                  : G4ACC_DEVICE_ERROR);
    }
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return G4ACC_OK;
}

}  // namespace

struct g4acc_handle {
  int device = 0;
  hipStream_t stream = nullptr;
  Geo g{};
  int precision = G4ACC_FP64;
  int sum_mode = G4ACC_SUM_FP64;
  int max_walkers = 1, max_vertices = 1, chunk = 1;
  int mtog_chunk = 0;  // slots per M -> G pass when below chunk (tuning key; 0: chunk)
  std::vector<int> channels, w_ex, k_ex;
  size_t g4_entries = 0;   // per channel
  bool initialized = false;
  bool g4_present = true;
  std::string err;

  // device memory
  std::vector<void*> allocations;
  size_t bytes = 0;
  int *d_kadd = nullptr, *d_ksub = nullptr, *d_kex = nullptr, *d_wex = nullptr;
  int* d_tperm = nullptr;          // G4 tile schedule 2 (diagonal strips)
  // Square-cluster site labels (x + L y) and the caller momentum of each (kx, ky); L = 0: none.
  int grid_L = 0;
  int *d_site_map = nullptr, *d_kmap = nullptr;
  void* d_dft = nullptr;           // Cplx<Real>[Nc*Nc]
  double2* d_bphase = nullptr;     // [nb*Nc] or null
  double2* d_G0 = nullptr;
  std::vector<void*> d_g4;         // per channel, 16 B per entry
  int *d_perm = nullptr, *d_off = nullptr, *d_err = nullptr;
  void *d_T = nullptr, *d_B = nullptr, *d_A = nullptr, *d_E = nullptr, *d_G = nullptr;
  void* d_Ms = nullptr;            // M in sorted order (6x6 / 8x8 matrix-core M -> G)
  void* d_UV = nullptr;            // rank-one factors of the tiled single-band G4 kernel
  // Model (g4acc_create_model). The general model (any field off its default) runs the
  // four-stage M -> G and the per-entry G4 kernels.
  int spin_symmetric = 1, complex_m = 0, non_dd = 0;
  bool general() const { return !spin_symmetric || complex_m || non_dd; }
  int *d_perm2 = nullptr, *d_off2 = nullptr;  // column sort by left labels (non-density-density)
  void* d_T2 = nullptr;                       // and its phases
  bool fused_mtog = false;         // register-fused single-band M -> G (no E workspace)
  bool use_tiled = false;
  bool use_band = false;           // LDS-tiled multiband G4 kernel (G4ACC_GENERIC_G4=1 disables)
  int tile_order = -1;             // G4 tile schedule (G4ACC_TILE_ORDER); -1 = by batch size
  SlotDesc* d_desc = nullptr;
  SlotDesc* h_desc = nullptr;      // pinned
  double* d_stage_f = nullptr;     // device staging of host inputs (M, tau)
  int* d_stage_i = nullptr;        // (site, band)
  double* h_stage_f = nullptr;     // pinned
  int* h_stage_i = nullptr;
  size_t stage_f_cap = 0, stage_i_cap = 0;
  // Queue of g4acc_accumulate_queued: owned copies of host inputs; device inputs are copied
  // into d_qm (M), d_qtau, d_qsb (site, band) slot by slot.
  struct Queued {
    int n[2] = {0, 0};
    double sign = 1;
    std::vector<double> m[2], tau[2];
    std::vector<int> site[2], band[2];
  };
  std::vector<Queued> queue;
  int queue_kind = G4ACC_HOST;
  double* d_qm = nullptr;
  double* d_qtau = nullptr;
  int* d_qsb = nullptr;
  hipEvent_t staged = nullptr;     // staging buffers free once this event completes
  bool staged_pending = false;
  // Captured launch sequences, one per batch size (G4ACC_GRAPH=0 disables them).
  bool use_graph = true;
  int mtog_tw = 0;                 // tuning: frequencies per thread of the fused M -> G kernel
  int mtog_mfma = 0;               // tuning: fused right M -> G kernel on the matrix cores
  // 6x6 / 8x8 square clusters, the matrix-core M -> G of mtog_large.hip (large_ok): 1 = S3 into
  // the E workspace plus the site transform kernel (large_e: E was allocated), 2 = FP32 only, S3,
  // site transform and epilogue in one kernel with E in LDS (the FP32 default).
  int mtog_large = 0;
  bool large_ok = false, large_e = false, right_fused_ok = false;
  int g4_qg = 0;                   // tuning: momenta per G4 workgroup (0: register limit)
  int g4_mfma = 0;                 // tuning: rank-one G4 terms on the matrix cores (FP32 G, Nc <= 16)
  // tuning: transpose-symmetric G4 tiles (one q = 0, one v = 0, channels among magnetic, charge,
  // transverse); sym_ok when the geometry allows it.
  int g4_sym = 0;
  bool sym_ok = false;
  // tuning: rank-one MAGNETIC / CHARGE terms of a batch as one matrix-core GEMM (g4_rank1.hip),
  // stored in d_R (complex64, one G4-sized plane per channel) and added by the tile kernel.
  int g4_rank1 = 0;
  void* d_R = nullptr;
  // This device's address of the static workspace of g4acc_compute_G_noalloc.
  unsigned char* d_scratch = nullptr;
  std::string scratch_err;
  int* d_r1_planes = nullptr;
  int n_r1 = 0, r1_slot[4] = {-1, -1, -1, -1};
  // Single-particle accumulator (g4acc_sp_enable): sorted by caller orbital, exact M_r_w sums.
  bool sp = false;
  bool sp_sqr = false;
  int sp_nw = 0;
  int *d_sp_perm = nullptr, *d_sp_off = nullptr, *d_rsub = nullptr;
  double2 *d_sp_T = nullptr, *d_sp_C = nullptr, *d_sp_acc = nullptr, *d_sp_acc_sqr = nullptr;
  size_t sp_entries = 0;  // per spin: nb^2 Nc n_w
  std::vector<std::pair<int, hipGraphExec_t>> graphs;  // key: (walkers << 2) | stage
  // stage timing (profiling hook)
  bool timing = false;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  double t_mtog_ms = 0, t_g4_ms = 0;
  long long timed_batches = 0;

  int fail(int code, const std::string& m) {
    err = m;
    return code;
  }
  int check(hipError_t e, const char* what) {
    if (e == hipSuccess)
      return G4ACC_OK;
    err = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? G4ACC_OUT_OF_MEMORY : G4ACC_DEVICE_ERROR;
  }
  template <typename T>
  int alloc(T** p, size_t count, const char* what) {
    void* q = nullptr;
    const size_t b = std::max<size_t>(count * sizeof(T), 16);
    hipError_t e = hipMalloc(&q, b);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      err = std::string("hipMalloc(") + what + ", " + std::to_string(b) + " B) failed: " +
            hipGetErrorString(e);
      return e == hipErrorOutOfMemory ? G4ACC_OUT_OF_MEMORY : G4ACC_DEVICE_ERROR;
    }
    allocations.push_back(q);
    bytes += b;
    *p = static_cast<T*>(q);
    return G4ACC_OK;
  }
  size_t cplx_bytes() const { return precision == G4ACC_FP64 ? 16 : 8; }
};

#define G4_TRY(expr)         \
  do {                       \
    const int _rc = (expr);  \
    if (_rc != G4ACC_OK)     \
      return _rc;            \
  } while (0)

namespace {

int validate_geometry(const g4acc_geometry* gm, std::string& why) {
  if (!gm) {
    why = "geometry is NULL";
    return G4ACC_INVALID_ARG;
  }
  if (gm->n_bands < 1 || gm->n_bands > 4) {
    why = "n_bands must be in [1, 4]";
    return G4ACC_INVALID_ARG;
  }
  if (gm->n_sites < 1 || gm->n_freq < 2 || gm->n_freq % 2 || !(gm->beta > 0)) {
    why = "n_sites >= 1, even n_freq >= 2 and beta > 0 are required";
    return G4ACC_INVALID_ARG;
  }
  if (gm->n_w_ex < 1 || !gm->w_ex || gm->n_k_ex < 1 || !gm->k_ex || gm->n_channels < 1 ||
      !gm->channels || !gm->k_add || !gm->k_sub || !gm->dft || !gm->g0) {
    why = "missing exchange lists, channels, k tables, dft or g0";
    return G4ACC_INVALID_ARG;
  }
  for (int i = 0; i < gm->n_k_ex; ++i)
    if (gm->k_ex[i] < 0 || gm->k_ex[i] >= gm->n_sites) {
      why = "k_ex index out of range";
      return G4ACC_INVALID_ARG;
    }
  const int nc2 = gm->n_sites * gm->n_sites;
  for (int i = 0; i < nc2; ++i)
    if (gm->k_add[i] < 0 || gm->k_add[i] >= gm->n_sites || gm->k_sub[i] < 0 ||
        gm->k_sub[i] >= gm->n_sites) {
      why = "k_add / k_sub entry out of range";
      return G4ACC_INVALID_ARG;
    }
  for (int i = 0; i < gm->n_channels; ++i) {
    const int c = gm->channels[i];
    if (c < G4ACC_PARTICLE_HOLE_TRANSVERSE || c > G4ACC_PARTICLE_PARTICLE_UP_DOWN) {
      why = "Specified four point type not implemented.";
      return G4ACC_UNSUPPORTED_CHANNEL;
    }
    if (i > 0 && gm->channels[i - 1] >= c) {
      why = "channels must be sorted and unique";
      return G4ACC_INVALID_ARG;
    }
  }
  return G4ACC_OK;
}

template <typename Real>
void upload_dft(std::vector<unsigned char>& buf, const double* dft, int nc) {
  using C = Cplx<Real>;
  buf.resize(sizeof(C) * nc * nc);
  C* p = reinterpret_cast<C*>(buf.data());
  for (int i = 0; i < nc * nc; ++i)
    p[i] = C{static_cast<Real>(dft[2 * i]), static_cast<Real>(dft[2 * i + 1])};
}

// Finds whether the cluster DFT dft[k + nc r] = exp(i k.r) is that of an L x L torus: site and
// momentum coordinates (x, y), (kx, ky) in Z_L with dft = exp(2 pi i (kx x + ky y) / L). Then
// xy[r] = x + L y and kxy_to_k[kx + L ky] = k. Only L = 2, 4, 6, 8 are used (the separable right
// site DFT is instantiated for them). Returns L, or 0.
int detect_grid(int nc, const double* dft, std::vector<int>& xy, std::vector<int>& kxy_to_k) {
  const int L = static_cast<int>(std::lround(std::sqrt(static_cast<double>(nc))));
  if (L * L != nc || (L != 2 && L != 4 && L != 6 && L != 8))
    return 0;
  const double two_pi = 6.283185307179586476925;
  std::vector<int> P(static_cast<size_t>(nc) * nc);  // P[k + nc r]: exponent in Z_L
  for (int i = 0; i < nc * nc; ++i) {
    const double re = dft[2 * i], im = dft[2 * i + 1];
    int m = static_cast<int>(std::lround(std::atan2(im, re) * L / two_pi));
    m = ((m % L) + L) % L;
    if (std::abs(re - std::cos(two_pi * m / L)) > 1e-9 || std::abs(im - std::sin(two_pi * m / L)) > 1e-9)
      return 0;
    P[i] = m;
  }
  std::vector<int> kx(nc), ky(nc), seen(nc);
  for (int r1 = 0; r1 < nc; ++r1)
    for (int r2 = 0; r2 < nc; ++r2) {
      // Candidate unit sites (1,0) and (0,1): k -> (P[k][r1], P[k][r2]) must be a bijection.
      std::fill(seen.begin(), seen.end(), 0);
      bool bij = true;
      for (int k = 0; k < nc && bij; ++k) {
        kx[k] = P[k + nc * r1];
        ky[k] = P[k + nc * r2];
        bij = !seen[kx[k] + L * ky[k]]++;
      }
      if (!bij)
        continue;
      // Every site then has coordinates (x, y) with P[k][r] = kx x + ky y for all k.
      std::vector<int> xy_try(nc, -1), used(nc, 0);
      bool all = true;
      for (int r = 0; r < nc && all; ++r) {
        for (int c = 0; c < nc && xy_try[r] < 0; ++c) {
          const int x = c % L, y = c / L;
          bool fits = true;
          for (int k = 0; k < nc && fits; ++k)
            fits = P[k + nc * r] == (kx[k] * x + ky[k] * y) % L;
          if (fits)
            xy_try[r] = c;
        }
        all = xy_try[r] >= 0 && !used[xy_try[r]]++;
      }
      if (!all)
        continue;
      xy = xy_try;
      kxy_to_k.assign(nc, 0);
      for (int k = 0; k < nc; ++k)
        kxy_to_k[kx[k] + L * ky[k]] = k;
      return L;
    }
  return 0;
}

// One walker's inputs as the staging reads them. The real walker (g4acc_walker), the CT-AUX
// hand-off (N with exp(V), g4acc_accumulate_batch_n) and the general model's walker all map onto it.
struct WalkerIn {
  int n[2];
  const double* m[2];
  int ld[2];
  const double* tau[2];
  const int* site[2];
  const int* band[2];
  const int* site_l[2];    // nullptr: the vertices are density-density
  const int* band_l[2];
  bool n_form;             // m holds N and M = diag(exp(V) - 1) N with expv[s] = exp(V)
  const double* expv[2];
  double sign, sign_im;
  bool m_cplx;
};

WalkerIn walker_in(const g4acc_walker& w, const double* const* exp_v = nullptr) {
  WalkerIn x{};
  for (int s = 0; s < 2; ++s) {
    x.n[s] = w.n[s];
    x.m[s] = w.m[s];
    x.ld[s] = w.ld[s];
    x.tau[s] = w.tau[s];
    x.site[s] = w.site[s];
    x.band[s] = w.band[s];
    x.expv[s] = exp_v ? exp_v[s] : nullptr;
  }
  x.n_form = exp_v != nullptr;
  x.sign = w.sign;
  return x;
}

WalkerIn walker_in(const g4acc_handle* h, const g4acc_walker_general& w) {
  WalkerIn x{};
  for (int s = 0; s < 2; ++s) {
    x.n[s] = w.n[s];
    x.m[s] = w.m[s];
    x.ld[s] = w.ld[s];
    x.tau[s] = w.tau[s];
    x.site[s] = w.site[s];
    x.band[s] = w.band[s];
    x.site_l[s] = h->non_dd ? w.site_l[s] : nullptr;
    x.band_l[s] = h->non_dd ? w.band_l[s] : nullptr;
  }
  x.sign = w.sign[0];
  x.sign_im = w.sign[1];
  x.m_cplx = h->complex_m != 0;
  return x;
}

// Resolves the walkers of one chunk into slot descriptors; host inputs are packed into the
// pinned staging buffer and copied once. Returns the number of slots. exp(V) values live where M
// does; left labels where the configuration does.
int stage_walkers(g4acc_handle* h, const WalkerIn* ws, int nw_, int memory, hipStream_t st) {
  const int nb = h->g.nb, nc = h->g.nc;
  // The pinned descriptor and staging buffers are reused: wait until the previous copies are done.
  if (h->staged_pending) {
    G4_TRY(h->check(hipEventSynchronize(h->staged), "hipEventSynchronize"));
    h->staged_pending = false;
  }
  size_t of = 0, oi = 0;
  for (int w = 0; w < nw_; ++w) {
    const WalkerIn& wk = ws[w];
    for (int s = 0; s < 2; ++s) {
      SlotDesc& d = h->h_desc[2 * w + s];
      const int n = wk.n[s];
      d.n = n;
      d.sign = wk.sign;
      d.sign_im = wk.sign_im;
      d.m_cplx = wk.m_cplx ? 1 : 0;
      d.expv = nullptr;
      d.site_l = d.band_l = nullptr;
      if (n == 0) {
        d.m = nullptr;
        d.tau = nullptr;
        d.site = nullptr;
        d.band = nullptr;
        d.ld = 1;
        continue;
      }
      const size_t per = wk.m_cplx ? 2 : 1;  // doubles per element of M
      if (!(memory & 1)) {  // M in host memory: packed column by column
        of += of & (per - 1);  // complex entries start on a 16-byte boundary
        double* fm = h->h_stage_f + of;
        for (int j = 0; j < n; ++j)
          std::memcpy(fm + per * n * j, wk.m[s] + per * wk.ld[s] * j, sizeof(double) * per * n);
        d.m = h->d_stage_f + of;
        d.ld = n;
        of += per * n * n;
      }
      else {
        d.m = wk.m[s];
        d.ld = wk.ld[s];
      }
      if (wk.n_form) {
        const double* xv = wk.expv[s];
        if (!xv)
          return h->fail(G4ACC_INVALID_ARG, "exp_v missing for a non-empty spin sector");
        if (!(memory & 1)) {
          std::memcpy(h->h_stage_f + of, xv, sizeof(double) * n);
          d.expv = h->d_stage_f + of;
          of += n;
        }
        else
          d.expv = xv;
      }
      const bool left = wk.site_l[s] != nullptr;
      if (!(memory & 2)) {  // configuration in host memory: validated and packed
        std::memcpy(h->h_stage_f + of, wk.tau[s], sizeof(double) * n);
        int* im = h->h_stage_i + oi;
        for (int i = 0; i < n; ++i) {
          const int r = wk.site[s][i], b = wk.band[s][i];
          if (r < 0 || r >= nc || b < 0 || b >= nb)
            return h->fail(G4ACC_INVALID_ARG, "vertex site or band index out of range");
          im[i] = r;
          im[n + i] = b;
          if (left) {
            const int rl = wk.site_l[s][i], bl = wk.band_l[s][i];
            if (rl < 0 || rl >= nc || bl < 0 || bl >= nb)
              return h->fail(G4ACC_INVALID_ARG, "vertex left site or band index out of range");
            im[2 * n + i] = rl;
            im[3 * n + i] = bl;
          }
        }
        d.tau = h->d_stage_f + of;
        d.site = h->d_stage_i + oi;
        d.band = h->d_stage_i + oi + n;
        if (left) {
          d.site_l = h->d_stage_i + oi + 2 * n;
          d.band_l = h->d_stage_i + oi + 3 * n;
        }
        of += n;
        oi += (left ? 4 : 2) * static_cast<size_t>(n);
      }
      else {
        d.tau = wk.tau[s];
        d.site = wk.site[s];
        d.band = wk.band[s];
        d.site_l = wk.site_l[s];
        d.band_l = wk.band_l[s];
      }
    }
  }
  const int n_slots = 2 * nw_;
  if (of)
    G4_TRY(h->check(hipMemcpyAsync(h->d_stage_f, h->h_stage_f, of * sizeof(double),
                                   hipMemcpyHostToDevice, st), "hipMemcpyAsync(M, tau)"));
  if (oi)
    G4_TRY(h->check(hipMemcpyAsync(h->d_stage_i, h->h_stage_i, oi * sizeof(int),
                                   hipMemcpyHostToDevice, st), "hipMemcpyAsync(config)"));
  G4_TRY(h->check(hipMemcpyAsync(h->d_desc, h->h_desc, sizeof(SlotDesc) * n_slots,
                                 hipMemcpyHostToDevice, st), "hipMemcpyAsync(desc)"));
  G4_TRY(h->check(hipEventRecord(h->staged, st), "hipEventRecord"));
  h->staged_pending = true;
  return G4ACC_OK;
}

template <typename Real>
MtoGArgs<Real> mtog_args(g4acc_handle* h, int n_slots) {
  MtoGArgs<Real> a;
  a.g = h->g;
  a.desc = h->d_desc;
  a.n_slots = n_slots;
  a.n_max = h->max_vertices;
  a.chunk = h->mtog_chunk > 0 && h->mtog_chunk < h->chunk ? h->mtog_chunk : h->chunk;
  a.fused = h->fused_mtog;
  a.tw = h->mtog_tw;
  a.mfma = h->mtog_mfma;
  a.large = h->mtog_large != 0;
  a.right_fused = h->mtog_large == 2;
  a.Ms = h->d_Ms;
  a.perm = h->d_perm;
  a.off = h->d_off;
  a.T = static_cast<Cplx<Real>*>(h->d_T);
  a.perm2 = h->non_dd ? h->d_perm2 : h->d_perm;
  a.off2 = h->non_dd ? h->d_off2 : h->d_off;
  a.T2 = static_cast<Cplx<Real>*>(h->non_dd ? h->d_T2 : h->d_T);
  a.B = static_cast<Cplx<Real>*>(h->d_B);
  a.A = static_cast<Cplx<Real>*>(h->d_A);
  a.E = static_cast<Cplx<Real>*>(h->d_E);
  a.G = static_cast<Cplx<Real>*>(h->d_G);
  a.dft = static_cast<const Cplx<Real>*>(h->d_dft);
  a.bphase = h->d_bphase;
  a.G0 = h->d_G0;
  a.grid_L = h->grid_L;
  a.kmap = h->d_kmap;
  return a;
}

template <typename Real>
void run_mtog(g4acc_handle* h, int n_slots, hipStream_t st) {
  launch_sort(h->d_desc, h->g, n_slots, h->max_vertices, h->d_site_map, h->d_perm, h->d_off, h->d_err, st);
  if (h->non_dd)
    launch_sort(h->d_desc, h->g, n_slots, h->max_vertices, h->d_site_map, h->d_perm2, h->d_off2, h->d_err, st,
                true);
  launch_mtog<Real>(mtog_args<Real>(h, n_slots), st);
}

template <typename Real>
void run_g4(g4acc_handle* h, int n_walkers) {
  if (h->use_tiled) {
    TiledArgs<Real> a;
    a.g = h->g;
    a.G = static_cast<const Cplx<Real>*>(h->d_G);
    a.UV = static_cast<Cplx<Real>*>(h->d_UV);
    a.desc = h->d_desc;
    a.n_walkers = n_walkers;
    a.n_ch = static_cast<int>(h->channels.size());
    a.need_y = a.need_z = a.has_trans = false;
    for (int c = 0; c < a.n_ch; ++c) {
      a.ch[c] = h->channels[c];
      a.g4[c] = h->d_g4[c];
      a.need_y |= a.ch[c] != G4ACC_PARTICLE_PARTICLE_UP_DOWN;
      a.need_z |= a.ch[c] == G4ACC_PARTICLE_PARTICLE_UP_DOWN;
      a.has_trans |= a.ch[c] == G4ACC_PARTICLE_HOLE_TRANSVERSE;
    }
    a.k_add = h->d_kadd;
    a.k_sub = h->d_ksub;
    a.k_ex = h->d_kex;
    a.w_ex = h->d_wex;
    a.nqg = 1;
    a.qg_cap = h->g4_qg;
    // XCD-aware super-tiles pay off when many walkers re-read G; a single walker's launch is
    // dominated by the G4 read-modify-write, which the natural order streams best.
    a.tile_order = h->tile_order >= 0 ? h->tile_order : (n_walkers > 1 ? 2 : 0);
    a.tile_perm = h->d_tperm;
    a.sum_mode = h->sum_mode;
    a.mfma = h->g4_mfma;
    a.sym = h->g4_sym != 0;
    const bool r1 = std::is_same<Real, float>::value && h->g4_rank1 && h->d_R;
    a.R = r1 ? h->d_R : nullptr;
    a.n_r1 = h->n_r1;
    a.r1_planes = h->d_r1_planes;
    for (int c = 0; c < 4; ++c)
      a.r1_slot[c] = h->r1_slot[c];
    if (launch_g4_tiled<Real>(a, h->stream))
      return;
  }
  else if (h->g.nb > 1 && h->use_band && n_walkers >= 4) {
    // The staged blocks pay off over several walkers; a single walker's entries are cheaper to
    // gather directly (generic kernel).
    TiledArgs<Real> a{};
    a.g = h->g;
    a.G = static_cast<const Cplx<Real>*>(h->d_G);
    a.desc = h->d_desc;
    a.n_walkers = n_walkers;
    a.n_ch = static_cast<int>(h->channels.size());
    for (int c = 0; c < a.n_ch; ++c) {
      a.ch[c] = h->channels[c];
      a.g4[c] = h->d_g4[c];
    }
    a.k_add = h->d_kadd;
    a.k_sub = h->d_ksub;
    a.k_ex = h->d_kex;
    a.w_ex = h->d_wex;
    a.sum_mode = h->sum_mode;
    if (launch_g4_band<Real>(a, h->stream))
      return;
  }
  for (size_t c = 0; c < h->channels.size(); ++c) {
    G4Args<Real> a;
    a.g = h->g;
    a.channel = h->channels[c];
    a.G = static_cast<const Cplx<Real>*>(h->d_G);
    a.desc = h->d_desc;
    a.n_walkers = n_walkers;
    a.k_add = h->d_kadd;
    a.k_sub = h->d_ksub;
    a.k_ex = h->d_kex;
    a.w_ex = h->d_wex;
    a.sum_mode = h->sum_mode;
    a.g4 = h->d_g4[c];
    a.spin_symmetric = h->spin_symmetric != 0;
    launch_g4<Real>(a, h->stream);
  }
}

void launch_sp(g4acc_handle* h, int n_walkers, hipStream_t st) {
  const int n_slots = 2 * n_walkers;
  launch_sort(h->d_desc, h->g, n_slots, h->max_vertices, nullptr, h->d_sp_perm, h->d_sp_off, h->d_err, st);
  SpArgs a;
  a.g = h->g;
  a.desc = h->d_desc;
  a.n_slots = n_slots;
  a.n_max = h->max_vertices;
  a.n_w = h->sp_nw;
  a.perm = h->d_sp_perm;
  a.off = h->d_sp_off;
  a.T = h->d_sp_T;
  a.C = h->d_sp_C;
  a.rsub = h->d_rsub;
  a.acc = h->d_sp_acc;
  a.acc_sqr = h->d_sp_acc_sqr;
  a.squared_too = h->sp_sqr;
  launch_sp_accumulate(a, st);
}

void launch_stage(g4acc_handle* h, int n_walkers, int stage, hipStream_t st) {
  if (stage == 2) {
    launch_sp(h, n_walkers, st);
    return;
  }
  if (stage == 0) {
    if (h->precision == G4ACC_FP64)
      run_mtog<double>(h, 2 * n_walkers, st);
    else
      run_mtog<float>(h, 2 * n_walkers, st);
  }
  else if (h->precision == G4ACC_FP64)
    run_g4<double>(h, n_walkers);
  else
    run_g4<float>(h, n_walkers);
}

// The kernels of one stage read everything batch-specific (vertex counts, pointers, signs) from
// the slot descriptors, so each stage is captured once per batch size into a hipGraph and
// replayed; only the descriptor copy precedes each replay. The captured sequences hold kernel
// launches only (no memset nodes).
int run_stage_graph(g4acc_handle* h, int n_walkers, int stage, hipStream_t st) {
  const int key = (n_walkers << 2) | stage;
  hipGraphExec_t exec = nullptr;
  for (auto& e : h->graphs)
    if (e.first == key)
      exec = e.second;
  if (!exec) {
    hipGraph_t graph = nullptr;
    G4_TRY(h->check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture"));
    launch_stage(h, n_walkers, stage, st);
    const hipError_t launched = hipGetLastError();
    const hipError_t ended = hipStreamEndCapture(st, &graph);
    if (launched != hipSuccess || ended != hipSuccess || !graph) {
      if (graph)
        (void)hipGraphDestroy(graph);
      return h->check(launched != hipSuccess ? launched : ended, "graph capture");
    }
    const hipError_t inst = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    G4_TRY(h->check(inst, "hipGraphInstantiate"));
    h->graphs.emplace_back(key, exec);
  }
  return h->check(hipGraphLaunch(exec, st), "hipGraphLaunch");
}

int run_stage(g4acc_handle* h, int n_walkers, int stage, hipStream_t st) {
  if (h->use_graph && !h->timing)
    return run_stage_graph(h, n_walkers, stage, st);
  launch_stage(h, n_walkers, stage, st);
  return h->check(hipGetLastError(), "kernel launch");
}

// Runs the staged batch: M -> G, then the G4 update, on the handle's stream.
int run_batch(g4acc_handle* h, int n_walkers) {
  hipStream_t st = h->stream;
  if (h->timing)
    G4_TRY(h->check(hipEventRecord(h->ev[0], st), "hipEventRecord"));
  G4_TRY(run_stage(h, n_walkers, 0, st));
  if (h->timing)
    G4_TRY(h->check(hipEventRecord(h->ev[1], st), "hipEventRecord"));
  if (h->timing)
    G4_TRY(h->check(hipEventRecord(h->ev[2], h->stream), "hipEventRecord"));
  G4_TRY(run_stage(h, n_walkers, 1, h->stream));
  if (h->sp)
    G4_TRY(run_stage(h, n_walkers, 2, h->stream));
  if (h->timing) {
    G4_TRY(h->check(hipEventRecord(h->ev[3], h->stream), "hipEventRecord"));
    G4_TRY(h->check(hipEventSynchronize(h->ev[3]), "hipEventSynchronize"));
    float a = 0, b = 0;
    G4_TRY(h->check(hipEventElapsedTime(&a, h->ev[0], h->ev[1]), "hipEventElapsedTime"));
    G4_TRY(h->check(hipEventElapsedTime(&b, h->ev[2], h->ev[3]), "hipEventElapsedTime"));
    h->t_mtog_ms += a;
    h->t_g4_ms += b;
    ++h->timed_batches;
  }
  return G4ACC_OK;
}

// Waits for all work of the handle.
int quiesce(g4acc_handle* h) {
  G4_TRY(h->check(hipStreamSynchronize(h->stream), "hipStreamSynchronize"));
  h->staged_pending = false;
  return G4ACC_OK;
}

// Real walkers carry a sign of +-1; the general model's any finite (complex) sign.
int check_walker(g4acc_handle* h, const WalkerIn& w, bool general = false) {
  for (int s = 0; s < 2; ++s) {
    if (w.n[s] < 0 || w.n[s] > h->max_vertices)
      return h->fail(G4ACC_INVALID_ARG, "vertex count " + std::to_string(w.n[s]) +
                                            " outside [0, max_vertices=" +
                                            std::to_string(h->max_vertices) + "]");
    if (w.n[s] > 0 && (!w.m[s] || !w.tau[s] || !w.site[s] || !w.band[s] || w.ld[s] < w.n[s]))
      return h->fail(G4ACC_INVALID_ARG, "walker arrays missing or ld < n");
    if (w.n[s] > 0 && (!w.site_l[s]) != (!w.band_l[s]))
      return h->fail(G4ACC_INVALID_ARG, "left sites and left bands must be given together");
  }
  if (general) {
    if (!std::isfinite(w.sign) || !std::isfinite(w.sign_im))
      return h->fail(G4ACC_INVALID_ARG, "sign must be finite");
  }
  else if (!(w.sign == 1.0 || w.sign == -1.0))
    return h->fail(G4ACC_INVALID_ARG, "sign must be +1 or -1");
  return G4ACC_OK;
}
int check_walker(g4acc_handle* h, const g4acc_walker& w) { return check_walker(h, walker_in(w)); }

}  // namespace

extern "C" {

int g4acc_create(const g4acc_geometry* gm, const g4acc_options* opt, g4acc_handle** out) {
  const g4acc_model model{1, 0, 0};
  return g4acc_create_model(gm, opt, &model, out);
}

int g4acc_create_model(const g4acc_geometry* gm, const g4acc_options* opt, const g4acc_model* model,
                       g4acc_handle** out) {
  g_create_error.clear();
  if (!out) {
    g_create_error = "out is NULL";
    return G4ACC_INVALID_ARG;
  }
  *out = nullptr;
  std::string why;
  int rc = validate_geometry(gm, why);
  if (rc != G4ACC_OK) {
    g_create_error = why;
    return rc;
  }
  if (!opt || opt->max_walkers < 1 || opt->max_vertices < 1 ||
      (opt->precision != G4ACC_FP64 && opt->precision != G4ACC_FP32) ||
      (opt->sum_mode != G4ACC_SUM_FP64 && opt->sum_mode != G4ACC_SUM_FP32_KAHAN)) {
    g_create_error = "invalid options";
    return G4ACC_INVALID_ARG;
  }
  if (!model || (model->spin_symmetric != 0 && model->spin_symmetric != 1) ||
      (model->complex_m != 0 && model->complex_m != 1) ||
      (model->non_density_density != 0 && model->non_density_density != 1)) {
    g_create_error = "invalid model (each field 0 or 1)";
    return G4ACC_INVALID_ARG;
  }
  if (!model->spin_symmetric)
    for (int i = 0; i < gm->n_channels; ++i)
      if (gm->channels[i] != G4ACC_PARTICLE_PARTICLE_UP_DOWN && gm->channels[i] != G4ACC_PARTICLE_HOLE_MAGNETIC) {
        // tp_accumulator_cpu.hpp:752-753
        g_create_error = "Specified four point type not implemented.";
        return G4ACC_UNSUPPORTED_CHANNEL;
      }
  int n_dev = 0;
  if (hipGetDeviceCount(&n_dev) != hipSuccess || n_dev < 1) {
    (void)hipGetLastError();
    g_create_error = "no HIP device available";
    return G4ACC_DEVICE_ERROR;
  }
  if (opt->device < 0 || opt->device >= n_dev) {
    g_create_error = "device ordinal out of range";
    return G4ACC_INVALID_ARG;
  }
  auto* h = new (std::nothrow) g4acc_handle();
  if (!h) {
    g_create_error = "host allocation failed";
    return G4ACC_OUT_OF_MEMORY;
  }
  auto bail = [&](int code) {
    g_create_error = h->err;
    g4acc_destroy(h);
    return code;
  };
  h->device = opt->device;
  h->spin_symmetric = model->spin_symmetric;
  h->complex_m = model->complex_m;
  h->non_dd = model->non_density_density;
  if ((rc = h->check(hipSetDevice(h->device), "hipSetDevice")) != G4ACC_OK)
    return bail(rc);
  {
    // Only g4acc_compute_G_noalloc needs the address; if it cannot be resolved, that entry point
    // reports why and everything else works as before.
    void* p = nullptr;
    if (scratch_address(h->device, &p, h->scratch_err) == G4ACC_OK)
      h->d_scratch = static_cast<unsigned char*>(p);
    (void)hipGetLastError();
  }
  if ((rc = h->check(hipStreamCreate(&h->stream), "hipStreamCreate")) != G4ACC_OK)
    return bail(rc);
  if ((rc = h->check(hipEventCreateWithFlags(&h->staged, hipEventDisableTiming), "hipEventCreate")) != G4ACC_OK)
    return bail(rc);
  for (auto& e : h->ev)
    if ((rc = h->check(hipEventCreate(&e), "hipEventCreate")) != G4ACC_OK)
      return bail(rc);

  Geo& g = h->g;
  g.nb = gm->n_bands;
  g.nc = gm->n_sites;
  g.nw = gm->n_freq;
  g.beta = gm->beta;
  g.nq = gm->n_k_ex;
  g.nv = gm->n_w_ex;
  g.ext = 0;
  for (int i = 0; i < gm->n_w_ex; ++i)
    g.ext = std::max(g.ext, std::abs(gm->w_ex[i]));
  g.nwE = g.nw + 2 * g.ext;
  g.n_orb = g.nb * g.nc;
  g.D = g.n_orb * g.nwE;
  h->precision = opt->precision;
  h->sum_mode = opt->sum_mode;
  h->max_walkers = opt->max_walkers;
  h->max_vertices = opt->max_vertices;
  h->channels.assign(gm->channels, gm->channels + gm->n_channels);
  h->w_ex.assign(gm->w_ex, gm->w_ex + gm->n_w_ex);
  h->k_ex.assign(gm->k_ex, gm->k_ex + gm->n_k_ex);
  h->g4_entries = static_cast<size_t>(g.nb) * g.nb * g.nb * g.nb * g.nc * g.nw * g.nc * g.nw *
                  g.nq * g.nv;

  const int nc2 = g.nc * g.nc;
  const size_t D = g.D, W2 = 2 * static_cast<size_t>(h->max_walkers), NM = h->max_vertices;
  const size_t cb = h->cplx_bytes();
  // Slots whose S1..S4 workspaces (B, A: D x N; E: D x D) exist at once. The fused single-band
  // kernels need only A.
  h->fused_mtog = mtog_fused_covers(g) && !h->general() && getenv("G4ACC_UNFUSED_MTOG") == nullptr;
  // 6x6 and 8x8 tori: the right half runs as matrix-core S3 into E plus the site transform.
  std::vector<int> xy, kxy_to_k;
  if (getenv("G4ACC_DENSE_DFT") == nullptr)
    h->grid_L = detect_grid(g.nc, gm->dft, xy, kxy_to_k);
  h->large_ok = h->fused_mtog && (g.nc == 36 || g.nc == 64) && h->grid_L * h->grid_L == g.nc;
  // The FP32 right kernel with E in registers and LDS indexes a slot's A and G with 32-bit
  // offsets. E is allocated for FP64, where that kernel does not exist, for larger shapes, or on
  // request.
  h->right_fused_ok = h->large_ok && cb == 8 && D * D < (size_t(1) << 32) && D * NM < (size_t(1) << 32);
  h->large_e = h->large_ok && (!h->right_fused_ok || getenv("G4ACC_MTOG_E") != nullptr);
  h->mtog_large = !h->large_ok ? 0 : h->right_fused_ok ? 2 : 1;
  // Per slot of a chunk: A (D x N); the matrix-core 6x6 / 8x8 path adds B (D x N), the sorted M
  // (N x N real) and E (D x D) when allocated; the four-stage path B and E.
  const size_t per_slot = !h->fused_mtog ? cb * (2 * D * NM + D * D)
                          : h->large_ok  ? cb * (2 * D * NM + (h->large_e ? D * D : 0)) + cb / 2 * NM * NM
                                         : cb * D * NM;
  h->chunk = static_cast<int>(std::max<size_t>(1, std::min<size_t>(W2, workspace_budget_bytes() / per_slot)));
  // A whole batch that fits 1.5x the budget runs as one chunk: a tail chunk of a few slots leaves
  // most of the GPU idle in the left-half kernels.
  if (static_cast<size_t>(h->chunk) < W2 && W2 * per_slot <= workspace_budget_bytes() / 2 * 3)
    h->chunk = static_cast<int>(W2);
  const size_t CH = h->chunk;

  if ((rc = h->alloc(&h->d_kadd, nc2, "k_add")) || (rc = h->alloc(&h->d_ksub, nc2, "k_sub")) ||
      (rc = h->alloc(&h->d_kex, g.nq, "k_ex")) || (rc = h->alloc(&h->d_wex, g.nv, "w_ex")) ||
      (rc = h->alloc(&h->d_tperm, static_cast<size_t>(g.nw) * g.nw, "tile order")) ||
      (rc = h->alloc(reinterpret_cast<unsigned char**>(&h->d_dft), cb * nc2, "dft")) ||
      (rc = h->alloc(&h->d_G0, static_cast<size_t>(g.nb) * g.nb * 2 * g.nc * g.nwE, "G0")) ||
      (rc = h->alloc(&h->d_perm, W2 * NM, "perm")) ||
      (rc = h->alloc(&h->d_off, W2 * (g.n_orb + 1), "off")) || (rc = h->alloc(&h->d_err, 1, "err")) ||
      (rc = h->alloc(reinterpret_cast<unsigned char**>(&h->d_T), cb * W2 * NM * g.nwE, "T")) ||
      (rc = h->alloc(reinterpret_cast<unsigned char**>(&h->d_B),
                     h->fused_mtog && !h->large_ok ? 1 : cb * CH * NM * D, "B")) ||
      (rc = h->alloc(reinterpret_cast<unsigned char**>(&h->d_Ms), h->large_ok ? cb / 2 * CH * NM * NM : 1,
                     "sorted M")) ||
      (rc = h->alloc(reinterpret_cast<unsigned char**>(&h->d_A), cb * CH * NM * D, "A")) ||
      (rc = h->alloc(reinterpret_cast<unsigned char**>(&h->d_E),
                     h->fused_mtog && !h->large_e ? 1 : cb * CH * D * D, "E")) ||
      (rc = h->alloc(reinterpret_cast<unsigned char**>(&h->d_G), cb * W2 * D * D, "G")) ||
      (rc = h->alloc(&h->d_desc, W2, "desc")))
    return bail(rc);
  if (gm->band_phase && (rc = h->alloc(&h->d_bphase, static_cast<size_t>(g.nb) * g.nc, "band_phase")))
    return bail(rc);
  // Non-density-density vertices: a second sort (columns of M by left labels) and its phases.
  if (h->non_dd &&
      ((rc = h->alloc(&h->d_perm2, W2 * NM, "perm2")) || (rc = h->alloc(&h->d_off2, W2 * (g.n_orb + 1), "off2")) ||
       (rc = h->alloc(reinterpret_cast<unsigned char**>(&h->d_T2), cb * W2 * NM * g.nwE, "T2"))))
    return bail(rc);
  // The LDS-tiled G4 kernel covers one band, Nc >= 8 and tiles whose staged G blocks fit the
  // workgroup's LDS.
  {
    bool need_y = false, need_z = false;
    for (int c : h->channels) {
      need_y |= c != G4ACC_PARTICLE_PARTICLE_UP_DOWN;
      need_z |= c == G4ACC_PARTICLE_PARTICLE_UP_DOWN;
    }
    int max_lds = 0;
    if (hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, h->device) != hipSuccess)
      max_lds = 65536;
    const size_t lds = tiled_lds_bytes(g, need_y, need_z, cb, 4);
    if (const char* to = getenv("G4ACC_TILE_ORDER"))
      h->tile_order = atoi(to);
    if (const char* gr = getenv("G4ACC_GRAPH"))
      h->use_graph = atoi(gr) != 0;
    // The tile kernels take a real sign and the spin-symmetric channels: the general model runs
    // the per-entry kernel.
    h->use_tiled = g.nb == 1 && g.nc >= 8 && tiled_pairs_per_thread(g.nc) > 0 && !h->general() &&
                   lds <= static_cast<size_t>(max_lds) && getenv("G4ACC_GENERIC_G4") == nullptr;
    const size_t R = static_cast<size_t>(g.nb) * g.nc;
    const size_t band_lds = cb * (6 * R * (R + 1) + 4 * static_cast<size_t>(g.nb) * g.nb * g.nc);
    h->use_band = g.nb > 1 && band_lds <= static_cast<size_t>(max_lds) && !h->general() &&
                  getenv("G4ACC_GENERIC_G4") == nullptr;
    if (h->use_tiled &&
        (rc = h->alloc(reinterpret_cast<unsigned char**>(&h->d_UV),
                       cb * W2 / 2 * 4 * g.nq * g.nv * g.nc * g.nw, "UV")))
      return bail(rc);
  }
  // Transpose symmetry of G4 over ((k1, w1), (k2, w2)): q = 0 (k + q = k for every k), v = 0 and
  // channels whose terms pair G(p1, p2) with G(p2, p1) only.
  {
    bool ok = h->use_tiled && g.nq == 1 && g.nv == 1 && gm->w_ex[0] == 0;
    for (int k = 0; ok && k < g.nc; ++k)
      ok = gm->k_add[k + g.nc * gm->k_ex[0]] == k;
    for (int c : h->channels)
      ok = ok && (c == G4ACC_PARTICLE_HOLE_MAGNETIC || c == G4ACC_PARTICLE_HOLE_CHARGE ||
                  c == G4ACC_PARTICLE_HOLE_TRANSVERSE);
    h->sym_ok = ok && getenv("G4ACC_NO_SYM") == nullptr;
    h->g4_sym = h->sym_ok ? 1 : 0;
  }
  h->d_g4.assign(h->channels.size(), nullptr);
  for (auto& p : h->d_g4)
    if ((rc = h->alloc(reinterpret_cast<double2**>(&p), h->g4_entries, "G4")))
      return bail(rc);
  // Per slot: M (or N; complex M takes two doubles per entry), tau and exp(V); right and left labels.
  h->stage_f_cap = W2 * ((h->complex_m ? 2 : 1) * NM * NM + 2 * NM);
  h->stage_i_cap = W2 * (h->non_dd ? 4 : 2) * NM;
  if ((rc = h->alloc(&h->d_stage_f, h->stage_f_cap, "stage_f")) ||
      (rc = h->alloc(&h->d_stage_i, h->stage_i_cap, "stage_i")))
    return bail(rc);
  if ((rc = h->check(hipHostMalloc(reinterpret_cast<void**>(&h->h_desc), sizeof(SlotDesc) * W2),
                     "hipHostMalloc")) ||
      (rc = h->check(hipHostMalloc(reinterpret_cast<void**>(&h->h_stage_f), sizeof(double) * h->stage_f_cap),
                     "hipHostMalloc")) ||
      (rc = h->check(hipHostMalloc(reinterpret_cast<void**>(&h->h_stage_i), sizeof(int) * h->stage_i_cap),
                     "hipHostMalloc")))
    return bail(rc);

  // Constant tables. On a square cluster the sites are relabelled x + L y (the sort orders the
  // vertices by that label and the DFT columns follow it), so the right site DFT can run as two
  // L-point passes (G4ACC_DENSE_DFT=1 keeps the caller's labels and the dense DFT).
  const std::vector<int> tperm = diagonal_tile_order(g.nw, 2);
  std::vector<double> dft_lab(gm->dft, gm->dft + 2 * static_cast<size_t>(nc2));
  if (h->grid_L) {
    for (int r = 0; r < g.nc; ++r)
      for (int k = 0; k < g.nc; ++k) {
        dft_lab[2 * (k + static_cast<size_t>(g.nc) * xy[r])] = gm->dft[2 * (k + static_cast<size_t>(g.nc) * r)];
        dft_lab[2 * (k + static_cast<size_t>(g.nc) * xy[r]) + 1] =
            gm->dft[2 * (k + static_cast<size_t>(g.nc) * r) + 1];
      }
    if ((rc = h->alloc(&h->d_site_map, g.nc, "site map")) || (rc = h->alloc(&h->d_kmap, g.nc, "k map")) ||
        (rc = h->check(hipMemcpy(h->d_site_map, xy.data(), sizeof(int) * g.nc, hipMemcpyHostToDevice), "copy")) ||
        (rc = h->check(hipMemcpy(h->d_kmap, kxy_to_k.data(), sizeof(int) * g.nc, hipMemcpyHostToDevice), "copy")))
      return bail(rc);
  }
  std::vector<unsigned char> dftbuf;
  if (h->precision == G4ACC_FP64)
    upload_dft<double>(dftbuf, dft_lab.data(), g.nc);
  else
    upload_dft<float>(dftbuf, dft_lab.data(), g.nc);
  if ((rc = h->check(hipMemcpy(h->d_kadd, gm->k_add, sizeof(int) * nc2, hipMemcpyHostToDevice), "copy")) ||
      (rc = h->check(hipMemcpy(h->d_ksub, gm->k_sub, sizeof(int) * nc2, hipMemcpyHostToDevice), "copy")) ||
      (rc = h->check(hipMemcpy(h->d_kex, gm->k_ex, sizeof(int) * g.nq, hipMemcpyHostToDevice), "copy")) ||
      (rc = h->check(hipMemcpy(h->d_wex, gm->w_ex, sizeof(int) * g.nv, hipMemcpyHostToDevice), "copy")) ||
      (rc = h->check(hipMemcpy(h->d_tperm, tperm.data(), sizeof(int) * tperm.size(), hipMemcpyHostToDevice),
                     "copy")) ||
      (rc = h->check(hipMemcpy(h->d_dft, dftbuf.data(), dftbuf.size(), hipMemcpyHostToDevice), "copy")) ||
      (rc = h->check(hipMemcpy(h->d_G0, gm->g0,
                               sizeof(double2) * g.nb * g.nb * 2 * g.nc * g.nwE, hipMemcpyHostToDevice),
                     "copy")) ||
      (rc = h->check(hipMemset(h->d_err, 0, sizeof(int)), "memset")))
    return bail(rc);
  if (gm->band_phase &&
      (rc = h->check(hipMemcpy(h->d_bphase, gm->band_phase, sizeof(double2) * g.nb * g.nc,
                               hipMemcpyHostToDevice), "copy")))
    return bail(rc);
  // Until resetAccumulation, accumulate is refused (tp_accumulator_gpu.hpp:329-330); G4 starts at 0.
  for (auto* p : h->d_g4)
    launch_zero16(p, h->g4_entries, h->stream);
  if ((rc = h->check(hipStreamSynchronize(h->stream), "hipStreamSynchronize")))
    return bail(rc);
  // Default launch shapes for callers that do not tune (the C++ drop-in): the matrix-core forms
  // where they were measured faster, complex64 G on 16-site clusters with batches of 8 or more.
  if (h->precision != G4ACC_FP64 && g.nb == 1 && g.nc == 16 && !h->general()) {
    h->mtog_mfma = 1;
    int mask = 0;
    for (int c : h->channels)
      mask |= 1 << c;
    constexpr int M = 1 << G4ACC_PARTICLE_HOLE_MAGNETIC, Q = 1 << G4ACC_PARTICLE_HOLE_CHARGE,
                  P = 1 << G4ACC_PARTICLE_PARTICLE_UP_DOWN;
    if (h->use_tiled && h->max_walkers >= 8 && (mask == (M | Q | P) || mask == (M | P) || mask == M || mask == Q) &&
        (rc = g4acc_set_tuning(h, "g4_rank1", 1)) != G4ACC_OK)
      return bail(rc);
  }
  *out = h;
  return G4ACC_OK;
}

int g4acc_destroy(g4acc_handle* h) {
  if (!h)
    return G4ACC_OK;
  if (h->stream)
    (void)hipStreamSynchronize(h->stream);
  for (auto& e : h->graphs)
    (void)hipGraphExecDestroy(e.second);
  for (void* p : h->allocations)
    (void)hipFree(p);
  if (h->h_desc)
    (void)hipHostFree(h->h_desc);
  if (h->h_stage_f)
    (void)hipHostFree(h->h_stage_f);
  if (h->h_stage_i)
    (void)hipHostFree(h->h_stage_i);
  if (h->staged)
    (void)hipEventDestroy(h->staged);
  for (auto e : h->ev)
    if (e)
      (void)hipEventDestroy(e);
  if (h->stream)
    (void)hipStreamDestroy(h->stream);
  delete h;
  return G4ACC_OK;
}

int g4acc_reset(g4acc_handle* h) {
  if (!h)
    return G4ACC_INVALID_ARG;
  G4_TRY(h->check(hipSetDevice(h->device), "hipSetDevice"));
  h->queue.clear();
  for (auto* p : h->d_g4)
    launch_zero16(p, h->g4_entries, h->stream);
  if (h->sp) {
    launch_zero16(h->d_sp_acc, 2 * h->sp_entries, h->stream);
    if (h->sp_sqr)
      launch_zero16(h->d_sp_acc_sqr, 2 * h->sp_entries, h->stream);
  }
  h->initialized = true;
  h->g4_present = true;
  return h->check(hipGetLastError(), "zero G4");
}

}  // extern "C"

namespace {

// accumulate_batch over walkers whose M is given as it is (exp_v == nullptr) or as N with the
// vertices' exp(V) values (exp_v[2 i + s]).
int check_batch(g4acc_handle* h, int n_walkers, const void* walkers, int memory) {
  if (!h)
    return G4ACC_INVALID_ARG;
  if (!h->initialized)
    return h->fail(G4ACC_NOT_INITIALIZED, "accumulate called before resetAccumulation");
  if (n_walkers < 0 || (n_walkers > 0 && !walkers) ||
      (memory != G4ACC_HOST && memory != G4ACC_DEVICE_M && memory != G4ACC_DEVICE))
    return h->fail(G4ACC_INVALID_ARG, "invalid batch arguments");
  return G4ACC_OK;
}

// Runs the non-empty walkers in chunks of max_walkers (empty configurations are no-ops,
// tp_accumulator_cpu.hpp:238-239).
int accumulate_walkers(g4acc_handle* h, const std::vector<WalkerIn>& all, int memory) {
  if (!h->g4_present)
    return G4ACC_OK;  // after sumTo the reference's G4 vector is empty: nothing is accumulated
  G4_TRY(h->check(hipSetDevice(h->device), "hipSetDevice"));
  std::vector<WalkerIn> todo;
  todo.reserve(all.size());
  for (const WalkerIn& w : all)
    if (w.n[0] + w.n[1] > 0)
      todo.push_back(w);
  for (size_t i0 = 0; i0 < todo.size(); i0 += h->max_walkers) {
    const int nw_ = static_cast<int>(std::min<size_t>(h->max_walkers, todo.size() - i0));
    G4_TRY(stage_walkers(h, todo.data() + i0, nw_, memory, h->stream));
    G4_TRY(run_batch(h, nw_));
  }
  return G4ACC_OK;
}

int accumulate_batch_impl(g4acc_handle* h, const g4acc_walker* walkers, const double* const* exp_v,
                          int n_walkers, int memory) {
  G4_TRY(check_batch(h, n_walkers, walkers, memory));
  std::vector<WalkerIn> all;
  all.reserve(n_walkers);
  for (int i = 0; i < n_walkers; ++i) {
    all.push_back(walker_in(walkers[i], exp_v ? exp_v + 2 * i : nullptr));
    G4_TRY(check_walker(h, all.back()));
  }
  return accumulate_walkers(h, all, memory);
}

}  // namespace

extern "C" {

int g4acc_accumulate_batch(g4acc_handle* h, const g4acc_walker* walkers, int n_walkers, int memory) {
  return accumulate_batch_impl(h, walkers, nullptr, n_walkers, memory);
}

int g4acc_accumulate_general(g4acc_handle* h, const g4acc_walker_general* walkers, int n_walkers,
                             int memory) {
  G4_TRY(check_batch(h, n_walkers, walkers, memory));
  std::vector<WalkerIn> all;
  all.reserve(n_walkers);
  for (int i = 0; i < n_walkers; ++i) {
    all.push_back(walker_in(h, walkers[i]));
    G4_TRY(check_walker(h, all.back(), true));
  }
  return accumulate_walkers(h, all, memory);
}

int g4acc_accumulate_batch_n(g4acc_handle* h, const g4acc_walker* walkers, const double* const* exp_v,
                             int n_walkers, int memory) {
  if (h && n_walkers > 0 && !exp_v)
    return h->fail(G4ACC_INVALID_ARG, "exp_v is NULL");
  return accumulate_batch_impl(h, walkers, exp_v, n_walkers, memory);
}

int g4acc_compute_m(g4acc_handle* h, int n, const double* exp_v, const double* N, int ldn, double* M,
                    int ldm) {
  if (!h)
    return G4ACC_INVALID_ARG;
  if (n < 0 || (n > 0 && (!exp_v || !N || !M || ldn < n || ldm < n)))
    return h->fail(G4ACC_INVALID_ARG, "compute_m: invalid arguments");
  if (n == 0)
    return G4ACC_OK;
  G4_TRY(h->check(hipSetDevice(h->device), "hipSetDevice"));
  launch_compute_m(n, exp_v, N, ldn, M, ldm, h->stream);
  return h->check(hipGetLastError(), "compute_m launch");
}

int g4acc_accumulate(g4acc_handle* h, const g4acc_walker* walker, int memory) {
  return g4acc_accumulate_batch(h, walker, 1, memory);
}

int g4acc_flush(g4acc_handle* h) {
  if (!h)
    return G4ACC_INVALID_ARG;
  if (h->queue.empty())
    return G4ACC_OK;
  const size_t NM = h->max_vertices, NM2 = NM * NM;
  const int kind = h->queue_kind;
  std::vector<g4acc_walker> ws(h->queue.size());
  for (size_t i = 0; i < ws.size(); ++i) {
    const auto& q = h->queue[i];
    g4acc_walker& w = ws[i];
    w.sign = q.sign;
    for (int s = 0; s < 2; ++s) {
      const size_t sl = 2 * i + s;
      w.n[s] = q.n[s];
      w.ld[s] = q.n[s] > 0 ? q.n[s] : 1;
      if (q.n[s] == 0) {
        w.m[s] = nullptr;
        w.tau[s] = nullptr;
        w.site[s] = w.band[s] = nullptr;
        continue;
      }
      w.m[s] = kind == G4ACC_HOST ? q.m[s].data() : h->d_qm + sl * NM2;
      if (kind == G4ACC_DEVICE) {
        w.tau[s] = h->d_qtau + sl * NM;
        w.site[s] = h->d_qsb + 2 * sl * NM;
        w.band[s] = h->d_qsb + (2 * sl + 1) * NM;
      }
      else {
        w.tau[s] = q.tau[s].data();
        w.site[s] = q.site[s].data();
        w.band[s] = q.band[s].data();
      }
    }
  }
  // The batch reads host copies before it returns and device copies in stream order, so the
  // queue slots are free for the next measurements.
  const int rc = g4acc_accumulate_batch(h, ws.data(), static_cast<int>(ws.size()), kind);
  h->queue.clear();
  return rc;
}

int g4acc_accumulate_queued(g4acc_handle* h, const g4acc_walker* walker, int memory) {
  if (!h || !walker)
    return G4ACC_INVALID_ARG;
  if (!h->initialized)
    return h->fail(G4ACC_NOT_INITIALIZED, "accumulate called before resetAccumulation");
  if (memory != G4ACC_HOST && memory != G4ACC_DEVICE_M && memory != G4ACC_DEVICE)
    return h->fail(G4ACC_INVALID_ARG, "invalid memory kind");
  if (!h->g4_present)
    return G4ACC_OK;
  G4_TRY(h->check(hipSetDevice(h->device), "hipSetDevice"));
  const g4acc_walker& w = *walker;
  G4_TRY(check_walker(h, w));
  if (w.n[0] + w.n[1] == 0)
    return G4ACC_OK;  // empty configuration: no-op (tp_accumulator_cpu.hpp:238-239)
  if (!h->queue.empty() && h->queue_kind != memory)
    G4_TRY(g4acc_flush(h));
  const size_t NM = h->max_vertices, NM2 = NM * NM, W2 = 2 * static_cast<size_t>(h->max_walkers);
  if (memory != G4ACC_HOST && !h->d_qm)
    G4_TRY(h->alloc(&h->d_qm, W2 * NM2, "queue M"));
  if (memory == G4ACC_DEVICE && !h->d_qtau) {
    G4_TRY(h->alloc(&h->d_qtau, W2 * NM, "queue tau"));
    G4_TRY(h->alloc(&h->d_qsb, 2 * W2 * NM, "queue site/band"));
  }
  const size_t slot = h->queue.size();
  g4acc_handle::Queued q;
  q.sign = w.sign;
  for (int s = 0; s < 2; ++s) {
    const int n = w.n[s];
    q.n[s] = n;
    if (n == 0)
      continue;
    const size_t sl = 2 * slot + s;
    if (memory == G4ACC_HOST) {
      q.m[s].resize(static_cast<size_t>(n) * n);
      for (int j = 0; j < n; ++j)
        std::memcpy(q.m[s].data() + static_cast<size_t>(n) * j, w.m[s] + static_cast<size_t>(w.ld[s]) * j,
                    sizeof(double) * n);
    }
    else
      G4_TRY(h->check(hipMemcpy2DAsync(h->d_qm + sl * NM2, sizeof(double) * n, w.m[s],
                                       sizeof(double) * w.ld[s], sizeof(double) * n, n,
                                       hipMemcpyDeviceToDevice, h->stream), "queue M copy"));
    if (memory == G4ACC_DEVICE) {
      G4_TRY(h->check(hipMemcpyAsync(h->d_qtau + sl * NM, w.tau[s], sizeof(double) * n,
                                     hipMemcpyDeviceToDevice, h->stream), "queue tau copy"));
      G4_TRY(h->check(hipMemcpyAsync(h->d_qsb + 2 * sl * NM, w.site[s], sizeof(int) * n,
                                     hipMemcpyDeviceToDevice, h->stream), "queue site copy"));
      G4_TRY(h->check(hipMemcpyAsync(h->d_qsb + (2 * sl + 1) * NM, w.band[s], sizeof(int) * n,
                                     hipMemcpyDeviceToDevice, h->stream), "queue band copy"));
    }
    else {
      q.tau[s].assign(w.tau[s], w.tau[s] + n);
      q.site[s].assign(w.site[s], w.site[s] + n);
      q.band[s].assign(w.band[s], w.band[s] + n);
    }
  }
  h->queue.push_back(std::move(q));
  h->queue_kind = memory;
  if (static_cast<int>(h->queue.size()) == h->max_walkers)
    return g4acc_flush(h);
  return G4ACC_OK;
}

int g4acc_accumulate_single(g4acc_handle* h, const g4acc_walker* walker, int memory) {
  return g4acc_accumulate_queued(h, walker, memory);
}

int g4acc_sync_copy(g4acc_handle* h) {
  if (!h)
    return G4ACC_INVALID_ARG;
  G4_TRY(h->check(hipSetDevice(h->device), "hipSetDevice"));
  return h->check(hipStreamSynchronize(h->stream), "hipStreamSynchronize");
}

int g4acc_sync(g4acc_handle* h) {
  if (!h)
    return G4ACC_INVALID_ARG;
  G4_TRY(h->check(hipSetDevice(h->device), "hipSetDevice"));
  G4_TRY(g4acc_flush(h));
  G4_TRY(quiesce(h));
  int flag = 0;
  G4_TRY(h->check(hipMemcpy(&flag, h->d_err, sizeof(int), hipMemcpyDeviceToHost), "read error flag"));
  if (flag) {
    G4_TRY(h->check(hipMemset(h->d_err, 0, sizeof(int)), "clear error flag"));
    return h->fail(G4ACC_INVALID_ARG, "a device vertex had a site or band index out of range");
  }
  return G4ACC_OK;
}

int g4acc_finalize(g4acc_handle* h) { return g4acc_sync(h); }

int g4acc_get_g4(g4acc_handle* h, int channel_index, double* out, size_t n_complex) {
  if (!h)
    return G4ACC_INVALID_ARG;
  if (!h->g4_present)
    return h->fail(G4ACC_EMPTY, "There is no G4 stored in this class.");
  if (channel_index < 0 || channel_index >= static_cast<int>(h->channels.size()) || !out ||
      n_complex < h->g4_entries)
    return h->fail(G4ACC_INVALID_ARG, "bad channel index or output buffer too small");
  G4_TRY(g4acc_sync(h));
  void* src = h->d_g4[channel_index];
  if (h->sum_mode == G4ACC_SUM_FP64)
    return h->check(hipMemcpy(out, src, sizeof(double2) * h->g4_entries, hipMemcpyDeviceToHost), "copy G4");
  double2* tmp = nullptr;
  G4_TRY(h->check(hipMalloc(&tmp, sizeof(double2) * h->g4_entries), "hipMalloc(export)"));
  launch_g4_export(src, tmp, h->g4_entries, h->sum_mode, h->stream);
  int rc = h->check(hipStreamSynchronize(h->stream), "export G4");
  if (rc == G4ACC_OK)
    rc = h->check(hipMemcpy(out, tmp, sizeof(double2) * h->g4_entries, hipMemcpyDeviceToHost), "copy G4");
  (void)hipFree(tmp);
  return rc;
}

int g4acc_get_g4_normalized(g4acc_handle* h, int channel_index, double accumulated_sign, double* out,
                            size_t n_complex) {
  if (!h)
    return G4ACC_INVALID_ARG;
  if (accumulated_sign == 0.0)
    return h->fail(G4ACC_INVALID_ARG, "accumulated sign is zero");
  G4_TRY(g4acc_get_g4(h, channel_index, out, n_complex));
  const double f = 1.0 / (h->g.beta * h->g.beta * accumulated_sign);
  for (size_t i = 0; i < 2 * h->g4_entries; ++i)
    out[i] *= f;
  return G4ACC_OK;
}

int g4acc_sum_to(g4acc_handle* src, g4acc_handle* dst) {
  if (!src || !dst || src == dst)
    return G4ACC_INVALID_ARG;
  if (src->channels != dst->channels || src->g4_entries != dst->g4_entries ||
      src->sum_mode != dst->sum_mode || src->device != dst->device)
    return src->fail(G4ACC_INVALID_ARG, "Objects accumulate different number of channels.");
  if (!src->g4_present || !dst->g4_present)
    return src->fail(G4ACC_EMPTY, "There is no G4 stored in this class.");
  G4_TRY(src->check(hipSetDevice(src->device), "hipSetDevice"));
  G4_TRY(g4acc_flush(src));
  G4_TRY(src->check(hipStreamSynchronize(src->stream), "hipStreamSynchronize"));
  for (size_t c = 0; c < src->channels.size(); ++c)
    launch_g4_add(dst->d_g4[c], src->d_g4[c], src->g4_entries, src->sum_mode, dst->stream);
  // The single-particle sums are added and kept in the source (SpAccumulator::sumTo,
  // sp_accumulator_gpu.hpp:231-236).
  if (src->sp && dst->sp && src->sp_entries == dst->sp_entries) {
    launch_g4_add(dst->d_sp_acc, src->d_sp_acc, 2 * src->sp_entries, G4ACC_SUM_FP64, dst->stream);
    if (src->sp_sqr && dst->sp_sqr)
      launch_g4_add(dst->d_sp_acc_sqr, src->d_sp_acc_sqr, 2 * src->sp_entries, G4ACC_SUM_FP64, dst->stream);
  }
  G4_TRY(dst->check(hipStreamSynchronize(dst->stream), "sum_to"));
  src->g4_present = false;
  return G4ACC_OK;
}

size_t g4acc_g4_size(const g4acc_handle* h) { return h ? h->g4_entries : 0; }
size_t g4acc_bytes(const g4acc_handle* h) { return h ? h->bytes : 0; }
void* g4acc_stream(g4acc_handle* h) { return h ? static_cast<void*>(h->stream) : nullptr; }

const char* g4acc_last_error(const g4acc_handle* h) {
  return h ? h->err.c_str() : g_create_error.c_str();
}

size_t g4acc_g_size(const g4acc_handle* h) {
  return h ? 2 * static_cast<size_t>(h->g.D) * h->g.D : 0;
}

}  // extern "C"

namespace {

// M -> G of one walker, exported in the reference layout to host memory.
int compute_G_impl(g4acc_handle* h, const WalkerIn& w, bool general, int memory, double* out,
                   size_t n_complex) {
  if (n_complex < g4acc_g_size(h))
    return h->fail(G4ACC_INVALID_ARG, "output buffer too small");
  if (memory != G4ACC_HOST && memory != G4ACC_DEVICE_M && memory != G4ACC_DEVICE)
    return h->fail(G4ACC_INVALID_ARG, "invalid memory kind");
  G4_TRY(h->check(hipSetDevice(h->device), "hipSetDevice"));
  G4_TRY(check_walker(h, w, general));
  G4_TRY(quiesce(h));
  G4_TRY(stage_walkers(h, &w, 1, memory, h->stream));
  if (h->precision == G4ACC_FP64)
    run_mtog<double>(h, 2, h->stream);
  else
    run_mtog<float>(h, 2, h->stream);
  double2* tmp = nullptr;
  const size_t n = g4acc_g_size(h);
  G4_TRY(h->check(hipMalloc(&tmp, sizeof(double2) * n), "hipMalloc(export)"));
  if (h->precision == G4ACC_FP64)
    launch_export_G<double>(static_cast<const double2*>(h->d_G), tmp, h->g, h->stream);
  else
    launch_export_G<float>(static_cast<const float2*>(h->d_G), tmp, h->g, h->stream);
  int rc = g4acc_sync(h);
  if (rc == G4ACC_OK)
    rc = h->check(hipMemcpy(out, tmp, sizeof(double2) * n, hipMemcpyDeviceToHost), "copy G");
  (void)hipFree(tmp);
  return rc;
}

}  // namespace

extern "C" {

int g4acc_compute_G(g4acc_handle* h, const g4acc_walker* walker, int memory, double* out,
                    size_t n_complex) {
  if (!h || !walker || !out)
    return G4ACC_INVALID_ARG;
  return compute_G_impl(h, walker_in(*walker), false, memory, out, n_complex);
}

int g4acc_compute_G_general(g4acc_handle* h, const g4acc_walker_general* walker, int memory, double* out,
                            size_t n_complex) {
  if (!h || !walker || !out)
    return G4ACC_INVALID_ARG;
  return compute_G_impl(h, walker_in(h, *walker), true, memory, out, n_complex);
}

}  // extern "C"

namespace {

// Byte offsets of the intermediates of one walker's M -> G inside the static workspace.
struct NoAllocLayout {
  size_t desc, perm, off, T, B, A, E, G, total;
};

NoAllocLayout noalloc_layout(const g4acc_handle* h, int n_max) {
  const size_t cb = h->cplx_bytes(), D = h->g.D, N = n_max;
  auto up = [](size_t b) { return (b + 255) / 256 * 256; };
  NoAllocLayout l;
  size_t o = 0;
  l.desc = o;
  o += up(2 * sizeof(SlotDesc));
  l.perm = o;
  o += up(2 * N * sizeof(int));
  l.off = o;
  o += up(2 * (static_cast<size_t>(h->g.n_orb) + 1) * sizeof(int));
  l.T = o;
  o += up(2 * N * h->g.nwE * cb);
  l.A = o;
  o += up(2 * N * D * cb);
  // The four-stage (multiband) form also keeps the time-transformed M and the left-transformed E.
  l.B = l.E = 0;
  if (!h->fused_mtog) {
    l.B = o;
    o += up(2 * N * D * cb);
    l.E = o;
    o += up(2 * D * D * cb);
  }
  l.G = o;
  o += up(2 * D * D * cb);
  l.total = o;
  return l;
}

template <typename Real>
void launch_mtog_noalloc(g4acc_handle* h, const NoAllocLayout& l, int n_max, double* G_dev) {
  unsigned char* s = h->d_scratch;
  int* perm = reinterpret_cast<int*>(s + l.perm);
  int* off = reinterpret_cast<int*>(s + l.off);
  MtoGArgs<Real> a = mtog_args<Real>(h, 2);
  a.large = false;  // the matrix-core 6x6 / 8x8 path needs the handle's B, E and Ms workspaces
  a.desc = reinterpret_cast<const SlotDesc*>(s + l.desc);
  a.n_max = n_max;
  a.chunk = 2;  // both spin sectors in one pass over the workspace
  a.perm = a.perm2 = perm;
  a.off = a.off2 = off;
  a.T = a.T2 = reinterpret_cast<Cplx<Real>*>(s + l.T);
  a.A = reinterpret_cast<Cplx<Real>*>(s + l.A);
  a.B = h->fused_mtog ? nullptr : reinterpret_cast<Cplx<Real>*>(s + l.B);
  a.E = h->fused_mtog ? nullptr : reinterpret_cast<Cplx<Real>*>(s + l.E);
  a.G = reinterpret_cast<Cplx<Real>*>(s + l.G);
  launch_sort(a.desc, h->g, 2, n_max, h->d_site_map, perm, off, h->d_err, h->stream);
  launch_mtog<Real>(a, h->stream);
  launch_export_G<Real>(a.G, reinterpret_cast<double2*>(G_dev), h->g, h->stream);
}

}  // namespace

extern "C" {

int g4acc_compute_G_noalloc(g4acc_handle* h, const g4acc_walker* walker, double* G_dev, size_t n_complex) {
  if (!h || !walker || !G_dev)
    return G4ACC_INVALID_ARG;
  if (n_complex < g4acc_g_size(h))
    return h->fail(G4ACC_INVALID_ARG, "output buffer too small");
  if (!h->d_scratch)
    return h->fail(G4ACC_DEVICE_ERROR, "the static workspace has no device address: " + h->scratch_err);
  if (h->general())
    return h->fail(G4ACC_INVALID_ARG, "compute_G_noalloc serves handles of the default model only");
  G4_TRY(h->check(hipSetDevice(h->device), "hipSetDevice"));
  G4_TRY(check_walker(h, *walker));
  const int n_max = std::max(1, std::max(walker->n[0], walker->n[1]));
  const NoAllocLayout l = noalloc_layout(h, n_max);
  if (l.total > G4ACC_NOALLOC_SCRATCH_BYTES)
    return h->fail(G4ACC_INVALID_ARG, "compute_G_noalloc: this measurement needs " + std::to_string(l.total) +
                                          " B of workspace, more than the static " +
                                          std::to_string(G4ACC_NOALLOC_SCRATCH_BYTES) + " B");
  // The pinned descriptor buffer is reused: wait until its previous copy has been read. Measurements
  // queued by g4acc_accumulate_queued stay queued (their copies are not touched here).
  if (h->staged_pending) {
    G4_TRY(h->check(hipEventSynchronize(h->staged), "hipEventSynchronize"));
    h->staged_pending = false;
  }
  for (int s = 0; s < 2; ++s) {
    SlotDesc& d = h->h_desc[s];
    d.n = walker->n[s];
    d.ld = walker->ld[s];
    d.sign = walker->sign;
    d.expv = nullptr;
    d.site_l = d.band_l = nullptr;
    d.sign_im = 0;
    d.m_cplx = 0;
    d.m = d.n ? walker->m[s] : nullptr;
    d.tau = d.n ? walker->tau[s] : nullptr;
    d.site = d.n ? walker->site[s] : nullptr;
    d.band = d.n ? walker->band[s] : nullptr;
  }
  G4_TRY(h->check(hipMemcpyAsync(h->d_scratch + l.desc, h->h_desc, 2 * sizeof(SlotDesc), hipMemcpyHostToDevice,
                                 h->stream), "hipMemcpyAsync(desc)"));
  G4_TRY(h->check(hipEventRecord(h->staged, h->stream), "hipEventRecord"));
  h->staged_pending = true;
  if (h->precision == G4ACC_FP64)
    launch_mtog_noalloc<double>(h, l, n_max, G_dev);
  else
    launch_mtog_noalloc<float>(h, l, n_max, G_dev);
  return h->check(hipGetLastError(), "compute_G_noalloc launch");
}

int g4acc_accumulate_G(g4acc_handle* h, const double* G_ref, double sign) {
  if (!h || !G_ref)
    return G4ACC_INVALID_ARG;
  if (!h->initialized)
    return h->fail(G4ACC_NOT_INITIALIZED, "accumulate called before resetAccumulation");
  if (!(sign == 1.0 || sign == -1.0))
    return h->fail(G4ACC_INVALID_ARG, "sign must be +1 or -1");
  if (!h->g4_present)
    return G4ACC_OK;
  G4_TRY(h->check(hipSetDevice(h->device), "hipSetDevice"));
  G4_TRY(quiesce(h));
  const size_t n = g4acc_g_size(h);
  double2* tmp = nullptr;
  G4_TRY(h->check(hipMalloc(&tmp, sizeof(double2) * n), "hipMalloc(import)"));
  int rc = h->check(hipMemcpyAsync(tmp, G_ref, sizeof(double2) * n, hipMemcpyHostToDevice, h->stream),
                    "copy G");
  if (rc == G4ACC_OK) {
    h->h_desc[0] = SlotDesc{nullptr, nullptr, nullptr, nullptr, 1, 1, sign, nullptr, nullptr, nullptr, 0.0, 0};
    h->h_desc[1] = h->h_desc[0];
    rc = h->check(hipMemcpyAsync(h->d_desc, h->h_desc, 2 * sizeof(SlotDesc), hipMemcpyHostToDevice,
                                 h->stream), "copy desc");
  }
  if (rc == G4ACC_OK) {
    if (h->precision == G4ACC_FP64) {
      launch_import_G<double>(tmp, static_cast<double2*>(h->d_G), h->g, h->stream);
      run_g4<double>(h, 1);
    }
    else {
      launch_import_G<float>(tmp, static_cast<float2*>(h->d_G), h->g, h->stream);
      run_g4<float>(h, 1);
    }
    rc = h->check(hipStreamSynchronize(h->stream), "accumulate_G");
  }
  (void)hipFree(tmp);
  return rc;
}

int g4acc_set_tuning(g4acc_handle* h, const char* key, int value) {
  if (!h || !key)
    return G4ACC_INVALID_ARG;
  const std::string k(key);
  if (k == "mtog_tw" && (value == 0 || value == 1 || value == 2 || value == 4 || value == 8))
    h->mtog_tw = value;
  else if (k == "mtog_mfma" && (value == 0 || value == 1))
    h->mtog_mfma = value;
  else if (k == "mtog_large" &&
           (value == 0 || (value == 1 && h->large_e) || (value == 2 && h->right_fused_ok)))
    h->mtog_large = value;
  else if (k == "mtog_chunk" && value >= 0)
    h->mtog_chunk = value;
  else if (k == "g4_sym" && (value == 0 || (value == 1 && h->sym_ok)))
    h->g4_sym = value;
  else if (k == "g4_qg" && (value == 0 || value == 1 || value == 2 || value == 4))
    h->g4_qg = value;
  else if (k == "g4_tile_order" && value >= -1 && value <= 2)
    h->tile_order = value;
  else if (k == "g4_mfma" && (value == 0 || value == 1))
    h->g4_mfma = value;
  else if (k == "g4_tile_strip" && value >= 1 && value <= 16) {
    // Diagonal positions per group of schedule 2 (table uploaded now, outside any batch).
    const std::vector<int> tperm = diagonal_tile_order(h->g.nw, value);
    int rs;
    if ((rs = h->check(hipSetDevice(h->device), "hipSetDevice")) ||
        (rs = h->check(hipMemcpy(h->d_tperm, tperm.data(), tperm.size() * sizeof(int), hipMemcpyHostToDevice),
                       "hipMemcpy")))
      return rs;
  }
  else if (k == "g4_rank1" && (value == 0 || value == 1)) {
    if (value && !h->d_R && h->use_tiled && h->precision != G4ACC_FP64) {
      int rs;
      if ((rs = h->check(hipSetDevice(h->device), "hipSetDevice")))
        return rs;
      // One complex64 plane of G4's (k1, w1, k2, w2, q, v) size per rank-one channel.
      std::vector<int> planes;
      for (size_t c = 0; c < h->channels.size(); ++c) {
        const int ch = h->channels[c];
        if (ch == G4ACC_PARTICLE_HOLE_MAGNETIC || ch == G4ACC_PARTICLE_HOLE_CHARGE) {
          h->r1_slot[c] = static_cast<int>(planes.size());
          planes.push_back(ch == G4ACC_PARTICLE_HOLE_MAGNETIC ? 0 : 1);  // U plane of uv_kernel
        }
      }
      if (!planes.empty()) {
        int rc;
        if ((rc = h->alloc(reinterpret_cast<float2**>(&h->d_R), planes.size() * h->g4_entries, "R")) ||
            (rc = h->alloc(&h->d_r1_planes, planes.size(), "r1_planes")))
          return rc;
        if ((rc = h->check(hipMemcpy(h->d_r1_planes, planes.data(), planes.size() * sizeof(int),
                                     hipMemcpyHostToDevice), "hipMemcpy")))
          return rc;
        h->n_r1 = static_cast<int>(planes.size());
      }
    }
    h->g4_rank1 = value;
  }
  else if (k == "graph" && (value == 0 || value == 1))
    h->use_graph = value != 0;
  else
    return h->fail(G4ACC_INVALID_ARG, "unknown tuning key or value: " + k);
  // Captured sequences bake the launch configuration in: drop them.
  G4_TRY(quiesce(h));
  for (auto& e : h->graphs)
    (void)hipGraphExecDestroy(e.second);
  h->graphs.clear();
  return G4ACC_OK;
}

int g4acc_sp_enable(g4acc_handle* h, int n_sp_freq, const int* r_sub, int accumulate_m_sqr) {
  if (!h)
    return G4ACC_INVALID_ARG;
  if (h->sp)
    return h->fail(G4ACC_INVALID_ARG, "the single-particle accumulator is already enabled");
  if (h->general())
    return h->fail(G4ACC_INVALID_ARG, "the single-particle accumulator serves handles of the default model only");
  if (n_sp_freq < 1 || !r_sub)
    return h->fail(G4ACC_INVALID_ARG, "n_sp_freq >= 1 and r_sub are required");
  const int nc = h->g.nc;
  for (int i = 0; i < nc * nc; ++i)
    if (r_sub[i] < 0 || r_sub[i] >= nc)
      return h->fail(G4ACC_INVALID_ARG, "r_sub entry out of range");
  G4_TRY(h->check(hipSetDevice(h->device), "hipSetDevice"));
  G4_TRY(quiesce(h));
  const size_t W2 = 2 * static_cast<size_t>(h->max_walkers), NM = h->max_vertices;
  const size_t n_w = 2 * static_cast<size_t>(n_sp_freq), n_orb = h->g.n_orb;
  const size_t entries = static_cast<size_t>(h->g.nb) * h->g.nb * nc * n_w;
  G4_TRY(h->alloc(&h->d_sp_perm, W2 * NM, "sp perm"));
  G4_TRY(h->alloc(&h->d_sp_off, W2 * (n_orb + 1), "sp off"));
  G4_TRY(h->alloc(&h->d_rsub, static_cast<size_t>(nc) * nc, "r_sub"));
  G4_TRY(h->alloc(&h->d_sp_T, W2 * NM * n_w, "sp phases"));
  G4_TRY(h->alloc(&h->d_sp_C, W2 * n_orb * n_orb * n_w, "sp orbital sums"));
  G4_TRY(h->alloc(&h->d_sp_acc, 2 * entries, "sp M_r_w"));
  if (accumulate_m_sqr)
    G4_TRY(h->alloc(&h->d_sp_acc_sqr, 2 * entries, "sp M_r_w_sqr"));
  G4_TRY(h->check(hipMemcpy(h->d_rsub, r_sub, sizeof(int) * nc * nc, hipMemcpyHostToDevice), "copy r_sub"));
  h->sp_nw = static_cast<int>(n_w);
  h->sp_entries = entries;
  h->sp_sqr = accumulate_m_sqr != 0;
  launch_zero16(h->d_sp_acc, 2 * entries, h->stream);
  if (h->sp_sqr)
    launch_zero16(h->d_sp_acc_sqr, 2 * entries, h->stream);
  G4_TRY(h->check(hipStreamSynchronize(h->stream), "sp zero"));
  h->sp = true;
  for (auto& e : h->graphs)
    (void)hipGraphExecDestroy(e.second);
  h->graphs.clear();
  return G4ACC_OK;
}

size_t g4acc_sp_size(const g4acc_handle* h) { return h && h->sp ? 2 * 2 * h->sp_entries : 0; }

int g4acc_sp_get_m_r_w(g4acc_handle* h, int squared, double* out, size_t n_complex) {
  if (!h)
    return G4ACC_INVALID_ARG;
  if (!h->sp)
    return h->fail(G4ACC_NOT_INITIALIZED, "the single-particle accumulator is not enabled");
  if (squared && !h->sp_sqr)
    return h->fail(G4ACC_INVALID_ARG, "M_r_w_sqr was not requested (accumulate_m_sqr = 0)");
  if (!out || n_complex < g4acc_sp_size(h))
    return h->fail(G4ACC_INVALID_ARG, "output buffer too small");
  G4_TRY(g4acc_sync(h));
  const size_t E = h->sp_entries;
  std::vector<double2> acc(2 * E);
  G4_TRY(h->check(hipMemcpy(acc.data(), squared ? h->d_sp_acc_sqr : h->d_sp_acc, sizeof(double2) * 2 * E,
                            hipMemcpyDeviceToHost), "copy M_r_w"));
  // [s][(b1 + nb (b2 + nb r)) n_w + w] -> (b1, s, b2, s', r, w), zero for s != s'.
  const int nb = h->g.nb, nc = h->g.nc, nw = h->sp_nw;
  std::memset(out, 0, sizeof(double2) * g4acc_sp_size(h));
  for (int s = 0; s < 2; ++s)
    for (int r = 0; r < nc; ++r)
      for (int b2 = 0; b2 < nb; ++b2)
        for (int b1 = 0; b1 < nb; ++b1)
          for (int w = 0; w < nw; ++w) {
            const double2 v = acc[s * E + (b1 + nb * (b2 + static_cast<size_t>(nb) * r)) * nw + w];
            const size_t o = b1 + nb * (s + 2 * (b2 + static_cast<size_t>(nb) * (s + 2 * (r + static_cast<size_t>(nc) * w))));
            out[2 * o] = v.x;
            out[2 * o + 1] = v.y;
          }
  return G4ACC_OK;
}

int g4acc_set_timing(g4acc_handle* h, int enable) {
  if (!h)
    return G4ACC_INVALID_ARG;
  h->timing = enable != 0;
  h->t_mtog_ms = h->t_g4_ms = 0;
  h->timed_batches = 0;
  return G4ACC_OK;
}

int g4acc_get_timing(const g4acc_handle* h, double* mtog_ms, double* g4_ms, long long* batches) {
  if (!h)
    return G4ACC_INVALID_ARG;
  if (mtog_ms)
    *mtog_ms = h->t_mtog_ms;
  if (g4_ms)
    *g4_ms = h->t_g4_ms;
  if (batches)
    *batches = h->timed_batches;
  return G4ACC_OK;
}

}  // extern "C"
