/* g4acc: MI355X (gfx950) two-particle Green's function accumulator, C ABI.
 *
 * Drop-in boundary for DCA++'s TpAccumulator<Parameters, DistType, DeviceType>
 * (include/dca/phys/dca_step/cluster_solver/shared_tools/accumulation/tp/tp_accumulator_base.hpp:50-51,
 *  CPU: .../tp/tp_accumulator_cpu.hpp:56, GPU: .../tp/tp_accumulator_gpu.hpp:59).
 * Plain C types only; no function throws; every call returns a g4acc_status.
 *
 * Layouts (first index fastest, as dca::func::function stores them):
 *   G0  (b1, b2, s, k, wE) complex double, wE over the EXTENDED vertex frequencies
 *   G4  (b1, b2, b3, b4, k1, w1, k2, w2, k_ex, w_ex) complex double per channel
 *   G   (b1, b2, s, k1, k2, w1E, w2E) complex double (TpAccumulatorBase::G_)
 *   M   column-major N x N real double per spin sector, leading dimension ld
 * Complex values are interleaved (re, im).
 */
#ifndef G4ACC_H
#define G4ACC_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum g4acc_status {
  G4ACC_OK = 0,
  G4ACC_INVALID_ARG = 1,          /* std::logic_error / std::invalid_argument in the reference */
  G4ACC_UNSUPPORTED_CHANNEL = 2,  /* tp_accumulator_cpu.hpp:663-664, tp_accumulator_gpu.hpp:456-468 */
  G4ACC_NOT_INITIALIZED = 3,      /* accumulate before resetAccumulation, tp_accumulator_gpu.hpp:329-330 */
  G4ACC_OUT_OF_MEMORY = 4,        /* std::bad_alloc, tp_accumulator_gpu.hpp:314-317 */
  G4ACC_DEVICE_ERROR = 5,         /* std::runtime_error on a failed launch, tp_accumulator_kernels.cu:695-700 */
  G4ACC_EMPTY = 6                 /* get_G4 with no G4 stored, tp_accumulator_cpu.hpp:874-875 */
} g4acc_status;

/* FourPointType values (include/dca/phys/four_point_type.hpp:27-33). */
typedef enum g4acc_channel {
  G4ACC_PARTICLE_HOLE_TRANSVERSE = 1,
  G4ACC_PARTICLE_HOLE_MAGNETIC = 2,
  G4ACC_PARTICLE_HOLE_CHARGE = 3,
  G4ACC_PARTICLE_PARTICLE_UP_DOWN = 4
} g4acc_channel;

/* Arithmetic of the M -> G step. The reference's choice is TpAccumulationPrecision
 * (include/dca/config/mc_options.hpp.in:28, float or double). */
typedef enum g4acc_precision {
  G4ACC_FP64 = 0,   /* complex128 G, FP64 products and sums */
  G4ACC_FP32 = 1    /* complex64 G, FP32 products with FP32 accumulation; phases, the G epilogue
                       and the G4 sums stay FP64 */
} g4acc_precision;

/* Running sum of G4 over measurements. */
typedef enum g4acc_sum_mode {
  G4ACC_SUM_FP64 = 0,        /* FP64 G4 */
  G4ACC_SUM_FP32_KAHAN = 1   /* FP32 G4 with a per-entry FP32 Kahan compensation term */
} g4acc_sum_mode;

/* Where a walker's arrays live. Host arrays are copied into the library's pinned staging buffer
 * before the call returns; device arrays are read asynchronously until g4acc_sync. */
typedef enum g4acc_memory {
  G4ACC_HOST = 0,      /* M and the configuration (tau, site, band) in host memory */
  G4ACC_DEVICE_M = 1,  /* M in device memory, configuration in host memory (a GPU walker's
                          computeM output with the host-side configuration, as the reference's
                          GPU accumulator takes them, tp_accumulator_gpu.hpp:321-355) */
  G4ACC_DEVICE = 3     /* M and the configuration in device memory */
} g4acc_memory;

/* Everything the reference constructor reads from Parameters, the static domains and G0
 * (TpAccumulatorBase ctor, tp_accumulator_base.hpp:184-221). */
typedef struct g4acc_geometry {
  int n_bands;             /* BANDS of the model */
  int n_sites;             /* cluster sites Nc (= momenta) */
  double beta;
  int n_freq;              /* COMPACT vertex frequencies = 2 * four-point-fermionic-frequencies */
  int n_w_ex;              /* FrequencyExchangeDomain elements */
  const int* w_ex;
  int n_k_ex;              /* MomentumExchangeDomain elements (k indices) */
  const int* k_ex;
  int n_channels;          /* channels, sorted and unique (four_point_parameters.hpp:194-196) */
  const int* channels;
  const int* k_add;        /* Nc*Nc, k_add[i + Nc*j] = index(k_i + k_j) */
  const int* k_sub;        /* Nc*Nc, k_sub[i + Nc*j] = index(k_j - k_i) */
  const double* dft;       /* Nc*Nc complex, dft[k + Nc*r] = exp(i k.r) */
  const double* band_phase;/* n_bands*Nc complex exp(i k.a_b) at [b + n_bands*k], or NULL */
  const double* g0;        /* n_bands^2 * 2 * Nc * (n_freq + 2 max|w_ex|) complex, (b1,b2,s,k,wE) */
} g4acc_geometry;

typedef struct g4acc_options {
  int precision;      /* g4acc_precision */
  int sum_mode;       /* g4acc_sum_mode */
  int max_walkers;    /* walkers per batched launch (>= 1) */
  int max_vertices;   /* largest N per spin sector */
  int device;         /* HIP device ordinal */
} g4acc_options;

/* One measurement: what a walker hands to TpAccumulator::accumulate(M_pair, configs, sign)
 * (tp_accumulator_cpu.hpp:107-109). Density-density vertices: one band and one site each
 * (Vertex::get_left_band()/get_right_band() agree). */
typedef struct g4acc_walker {
  int n[2];                 /* vertices per spin sector (configs[s].size()) */
  const double* m[2];       /* M_pair[s], column-major */
  int ld[2];                /* leading dimensions of m[s] (>= n[s]) */
  const double* tau[2];     /* vertex imaginary times in [0, beta) */
  const int* site[2];       /* vertex cluster sites (r indices) */
  const int* band[2];       /* vertex bands */
  double sign;              /* SignType: +1 / -1 */
} g4acc_walker;

/* Model properties the reference fixes at compile time from Parameters. g4acc_create is
 * g4acc_create_model with {1, 0, 0}. */
typedef struct g4acc_model {
  int spin_symmetric;       /* lattice_type::spin_symmetric (tp_accumulator_base.hpp:167). 0: the
                               updateG4 branch of models whose spin is part of the band index
                               (tp_accumulator_cpu.hpp:668-757): G4 from spin sector 0 with the
                               whole sign; only the pp and magnetic channels exist there */
  int complex_m;            /* M complex (complex models; Matrix<std::complex<double>>) */
  int non_density_density;  /* vertices whose left and right (site, band) differ
                               (HasInitializeNonDensityInteractionMethod, tp_accumulator_base.hpp:169-170):
                               rows of M follow the right labels, columns the left ones
                               (cached_ndft_base.hpp:72-78, 97-103) */
} g4acc_model;

/* One measurement of the general model (accumulate(M_pair, configs, sign) with any SpScalar,
 * tp_accumulator_cpu.hpp:231-252). */
typedef struct g4acc_walker_general {
  int n[2];                 /* vertices per spin sector */
  const double* m[2];       /* M_pair[s], column-major; interleaved (re, im) when complex_m */
  int ld[2];                /* leading dimensions in elements (complex entries when complex_m) */
  const double* tau[2];     /* vertex times in [0, beta) */
  const int* site[2];       /* Vertex::get_right_site() */
  const int* band[2];       /* Vertex::get_right_band() */
  const int* site_l[2];     /* Vertex::get_left_site() (read when non_density_density; NULL = site) */
  const int* band_l[2];     /* Vertex::get_left_band() (read when non_density_density; NULL = band) */
  double sign[2];           /* SignType (type_help.hpp:299-312): (re, im); a complex phase for
                               complex models, (+-1, 0) otherwise */
} g4acc_walker_general;

typedef struct g4acc_handle g4acc_handle;

/* Constructor (tp_accumulator_base.hpp:184-208, tp_accumulator_gpu_base.hpp:132-152). */
int g4acc_create(const g4acc_geometry* geometry, const g4acc_options* options, g4acc_handle** out);
/* Constructor for a model with the properties in *model. A handle whose model is not
 * {1, 0, 0} runs the four-stage M -> G and the per-entry G4 kernels. With spin_symmetric = 0, a
 * channel other than pp and magnetic is refused (G4ACC_UNSUPPORTED_CHANNEL, as the reference
 * throws std::logic_error, tp_accumulator_cpu.hpp:752-753). */
int g4acc_create_model(const g4acc_geometry* geometry, const g4acc_options* options,
                       const g4acc_model* model, g4acc_handle** out);
/* Destructor. */
int g4acc_destroy(g4acc_handle* h);
/* resetAccumulation (tp_accumulator_cpu.hpp:224-229, tp_accumulator_gpu.hpp:271-296). */
int g4acc_reset(g4acc_handle* h);
/* accumulate(M_pair, configs, sign) (tp_accumulator_cpu.hpp:231-252, tp_accumulator_gpu.hpp:321-355).
 * An empty configuration (n[0] + n[1] == 0) is a no-op, as on the CPU (:238-239).
 * Caller buffers are borrowed until g4acc_sync (the reference's synchronizeCopy). */
int g4acc_accumulate(g4acc_handle* h, const g4acc_walker* walker, int memory);
/* Many walkers' measurements in one batched launch sequence; equivalent to calling
 * g4acc_accumulate for each walker. n_walkers may exceed max_walkers (processed in chunks). */
int g4acc_accumulate_batch(g4acc_handle* h, const g4acc_walker* walkers, int n_walkers, int memory);
/* g4acc_accumulate_batch for measurements of the general model (complex M, a complex sign, left
 * and right vertex labels), on a handle from g4acc_create_model. Complex M needs complex_m; left
 * labels are read only with non_density_density. Any finite sign is accepted. */
int g4acc_accumulate_general(g4acc_handle* h, const g4acc_walker_general* walkers, int n_walkers,
                             int memory);
/* g4acc_accumulate_batch for walkers that hand over N and the vertices' exp(V) instead of M, as a
 * CT-AUX walker holds them (CtauxWalker::computeM, ctaux_walker.hpp:1889-1916): walkers[i].m[s]
 * points to N_s (leading dimension ld[s]) and exp_v[2 i + s] to the n[s] values exp(V_j) of its
 * vertices (CV::exp_V), in the same memory as N. M_s = diag(exp_v - 1) N_s is formed in FP64 where
 * the M -> G step reads it, so M never reaches memory; results equal g4acc_accumulate_batch on
 * that M bit for bit. */
int g4acc_accumulate_batch_n(g4acc_handle* h, const g4acc_walker* walkers, const double* const* exp_v,
                             int n_walkers, int memory);
/* CtauxWalker::computeM on the device (multiplyDiagonalLeft, src/linalg/lapack/multiply_diagonal_gpu.cu):
 * M = diag(exp_v - 1) N for one n x n spin sector; exp_v, N and M are device pointers, column-major
 * with leading dimensions ldn and ldm; runs on the handle's stream (g4acc_stream). */
int g4acc_compute_m(g4acc_handle* h, int n, const double* exp_v, const double* N, int ldn, double* M,
                    int ldm);
/* accumulate(M_pair, configs, sign) with batching inside the handle, for callers that hand over
 * one measurement at a time (the reference launches each one on its own,
 * tp_accumulator_gpu.hpp:321-355). The measurement is copied into the handle's queue: host
 * arrays before the call returns, device arrays by a device copy on the handle's stream. The
 * queue runs as one batch when it holds max_walkers measurements, and at g4acc_flush,
 * g4acc_sync, g4acc_finalize, g4acc_get_g4 and g4acc_sum_to; g4acc_reset discards it. Results
 * equal g4acc_accumulate_batch over the same measurements in the same chunks. */
int g4acc_accumulate_queued(g4acc_handle* h, const g4acc_walker* walker, int memory);
/* Same as g4acc_accumulate_queued: the one-measurement drop-in for accumulate(M, configuration,
 * sign) (tp_accumulator_gpu.hpp:321-355) that batches internally. */
int g4acc_accumulate_single(g4acc_handle* h, const g4acc_walker* walker, int memory);
/* Runs the queued measurements now. */
int g4acc_flush(g4acc_handle* h);
/* synchronizeCopy() for queued measurements: returns once the caller's buffers have been copied
 * (device copies complete); does not run the queue. */
int g4acc_sync_copy(g4acc_handle* h);
/* finalize() (tp_accumulator_gpu.hpp:471-484): runs the queue and waits for all work. */
int g4acc_finalize(g4acc_handle* h);
/* syncStreams(): runs the queue and waits for all work; caller buffers may then be reused. */
int g4acc_sync(g4acc_handle* h);
/* get_G4()[channel_index] (tp_accumulator_cpu.hpp:871-877): copies n_complex = g4acc_g4_size()
 * complex doubles in the reference order. */
int g4acc_get_g4(g4acc_handle* h, int channel_index, double* out, size_t n_complex);
/* The solvers' G4 normalisation after the accumulation (ctaux_cluster_solver.hpp:523-527,
 * ctint_cluster_solver.hpp:511-516): get_g4 divided by beta^2 * accumulated_sign, in FP64.
 * accumulated_sign is the solver's sum of the measurement signs (0: G4ACC_INVALID_ARG). */
int g4acc_get_g4_normalized(g4acc_handle* h, int channel_index, double accumulated_sign, double* out,
                            size_t n_complex);
/* sumTo(other) (tp_accumulator_cpu.hpp:895-904): dst.G4 += src.G4 on the device, then src.G4
 * is cleared (get_g4 on src returns G4ACC_EMPTY until the next g4acc_reset). */
int g4acc_sum_to(g4acc_handle* src, g4acc_handle* dst);
/* Complex entries per channel of G4. */
size_t g4acc_g4_size(const g4acc_handle* h);
/* Device bytes held by the handle. */
size_t g4acc_bytes(const g4acc_handle* h);
/* get_stream() (tp_accumulator_gpu.hpp): the hipStream_t the handle launches on. */
void* g4acc_stream(g4acc_handle* h);
/* Message of the last failed call on h (or of the last failed g4acc_create when h is NULL). */
const char* g4acc_last_error(const g4acc_handle* h);

/* Testing entry (the reference's get_G_Debug, tp_accumulator_cpu.hpp:120-123): runs only the
 * M -> G step of one walker and copies G in the reference layout (b1,b2,s,k1,k2,w1E,w2E). */
int g4acc_compute_G(g4acc_handle* h, const g4acc_walker* walker, int memory, double* out,
                    size_t n_complex);
/* g4acc_compute_G for a measurement of the general model (get_G_Debug of a complex or
 * non-density-density model). */
int g4acc_compute_G_general(g4acc_handle* h, const g4acc_walker_general* walker, int memory, double* out,
                            size_t n_complex);
/* Complex entries of G in the reference layout. */
size_t g4acc_g_size(const g4acc_handle* h);
/* No counterpart in the reference. For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): g4acc_compute_G of one
 * walker whose M and configuration are in device memory (G4ACC_DEVICE), with G written in the same
 * layout (FP64 complex, n_complex >= g4acc_g_size(h)) to the caller's device buffer G_dev. The call
 * allocates nothing: every intermediate (slot descriptors, vertex sort, phases, the left-transformed
 * M, G in the handle's precision) lives in one static device array of
 * G4ACC_NOALLOC_SCRATCH_BYTES that each GPU the library runs on holds until the process ends.
 * Measurements whose intermediates exceed it are refused with G4ACC_INVALID_ARG. The handle's own
 * workspaces and G4 are not touched. All calls on one GPU share that array, which keeps the last
 * call's contents: calls that overlap on one GPU (two handles, two streams, two host threads) race.
 * Runs on the handle's stream and returns without waiting; G_dev is complete after g4acc_sync. */
#define G4ACC_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int g4acc_compute_G_noalloc(g4acc_handle* h, const g4acc_walker* walker, double* G_dev, size_t n_complex);
/* Testing entry (the minimum slice of the path): accumulates one measurement whose G is given
 * in the reference layout (b1,b2,s,k1,k2,w1E,w2E), host memory, skipping M -> G. Same effect on
 * G4 as the reference's updateG4 loop over channels (tp_accumulator_cpu.hpp:248-249). */
int g4acc_accumulate_G(g4acc_handle* h, const double* G, double sign);

/* The single-particle accumulator beside the two-particle one: SpAccumulator<Parameters, GPU>
 * (sp_accumulator_gpu.hpp; constructor :131-141, accumulate :165-190, finalize :215-229,
 * sumTo :231-236). Once enabled, every accumulate call of the handle also adds
 *   M_r_w(b1, s, b2, s, r, w) += sign / Nc sum_{i, j} M_s(i, j) exp(i w (tau_i - tau_j))
 *                                 [b_i = b1, b_j = b2, r_i - r_j = r]
 * over the 2 n_sp_freq single-particle Matsubara frequencies w = +-(2l + 1) pi / beta
 * (frequency_domain.cpp:29-49), and its M_ij^2 counterpart when accumulate_m_sqr != 0. The
 * reference approximates this sum with a delayed NFFT (dnfft_1d.hpp); here it is evaluated
 * exactly in FP64 (the two agree to the NFFT's approximation, below 1e-9 from 64 frequencies on).
 * r_sub[a + Nc b] = index(r_b - r_a) is the cluster's site subtraction. g4acc_reset clears it,
 * g4acc_sum_to adds it like G4. */
int g4acc_sp_enable(g4acc_handle* h, int n_sp_freq, const int* r_sub, int accumulate_m_sqr);
/* get_sign_times_M_r_w() (squared = 0) or get_sign_times_M_r_w_sqr() (squared = 1) after
 * finalize: n_complex >= g4acc_sp_size(h) complex doubles in the layout (b1, s1, b2, s2, r, w),
 * first index fastest, zero for s1 != s2 (MFunction of sp_accumulator.hpp). */
int g4acc_sp_get_m_r_w(g4acc_handle* h, int squared, double* out, size_t n_complex);
/* Complex entries of M_r_w (0 when the single-particle accumulator is not enabled). */
size_t g4acc_sp_size(const g4acc_handle* h);

/* Profiling hook (the reference's Profiler scopes "accumulate"/"updateG4",
 * tp_accumulator_cpu.hpp:236,416): when enabled, every batch records HIP events around its
 * M -> G and G4 stages and waits for them; times accumulate until the next enable call. */
int g4acc_set_timing(g4acc_handle* h, int enable);

/* Launch-shape tuning (no counterpart in the reference, whose launch shapes are fixed; set by
 * the autotuner's table, dca_amd/autotune.py). Keys: "mtog_tw" (frequencies per thread of the
 * fused single-band M -> G kernel: 0 default, 1, 2, 4, 8), "mtog_mfma" (1: its right half on the
 * matrix cores where available, FP32 G, Nc <= 8 or a 4x4 torus, max_vertices <= 4 (128 - Nc);
 * 0: vector ALUs), "g4_qg" (momenta per G4 workgroup:
 * 0 = as many as registers allow, 1, 2, 4), "g4_tile_order" (-1 by batch size: 2 for batches, 0 for
 * one walker; 0 natural,
 * 1 XCD-aware super-tiles, 2 XCD-aware diagonal strips with their mirror images),
 * "g4_tile_strip" (diagonal positions per group of schedule 2: 1..16, default 2), "g4_mfma"
 * (1: rank-one G4 terms on the matrix cores where available, FP32 G and Nc <= 16; 0: vector ALUs),
 * "g4_rank1" (1: the batch's rank-one MAGNETIC / CHARGE terms as one matrix-core GEMM into a
 * complex64 buffer of G4's size per such channel, allocated when the key is set; FP32 G,
 * 256-thread tiles, measured channel sets; 0: formed per walker in the tile kernel),
 * "graph" (1: replay captured batch sequences, 0: launch eagerly), "mtog_large" (6x6 and 8x8 tori:
 * 0 register-fused M -> G kernels; 1 matrix-core S1/S3 with a D x D workspace per slot of a chunk
 * for S3's output, the FP64 default, in FP32 only when the environment variable G4ACC_MTOG_E was
 * set at creation; 2 FP32 only, S3, the right site transform and the epilogue in one kernel with
 * that output kept in LDS, the FP32 default), "mtog_chunk" (slots per M -> G pass, below the
 * number the workspace was sized for; 0: that number, a whole batch where its workspace fits
 * 24 GiB), "g4_sym" (1: transpose-symmetric G4 tiles where the geometry allows them, the default
 * there; 0: all tiles).
 * Defaults: mtog_mfma = 1 and, with max_walkers >= 8 and a measured channel set, g4_rank1 = 1 for
 * complex64 G on single-band 16-site clusters; 0 elsewhere.
 * Results do not depend on the tuning beyond FP32 summation order. */
int g4acc_set_tuning(g4acc_handle* h, const char* key, int value);
int g4acc_get_timing(const g4acc_handle* h, double* mtog_ms, double* g4_ms, long long* batches);

#ifdef __cplusplus
}
#endif

#endif /* G4ACC_H */
