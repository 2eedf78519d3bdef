/* FP64 CPU oracle of the DCA++ two-particle (tp) accumulator.
 *
 * TEST INFRASTRUCTURE ONLY. This is the checker, never the product: only
 * tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py load it.
 *
 * It restates, loop for loop, the reference CPU accumulator
 *   include/dca/phys/dca_step/cluster_solver/shared_tools/accumulation/tp/tp_accumulator_cpu.hpp
 *   .../tp/tp_accumulator_base.hpp, .../tp/ndft/cached_ndft_base.hpp, .../tp/ndft/cached_ndft_cpu.hpp
 *   include/dca/math/function_transform/special_transforms/space_transform_2D.hpp
 * in plain C++17 with its own blocked complex GEMM (no BLAS).
 *
 * Pins (tests/test_oracle.py): the NDFT against its defining double sum and a factorising rank-one
 * M, the single-vertex closed form, the direct M -> G definition, M = 0 closed forms of the
 * transverse, magnetic, charge and pp channels for one and three bands, linearity, sumTo, empty
 * input. Parity unpinned: the multiband band-slot placement of each channel's terms rests on the
 * M = 0 closed forms, which are derived from the reference's source text (the reference's G4
 * baseline files are missing blobs).
 *
 * Layouts (first index fastest, as dca::func::function stores them):
 *   G0   (b1, b2, s, k, wE)                 wE over the EXTENDED vertex frequencies
 *   G    (b1, b2, s, k1, k2, w1E, w2E)      the reference's TpAccumulatorBase::G_
 *   G4   (b1, b2, b3, b4, k1, w1, k2, w2, k_ex, w_ex) per channel, w over COMPACT frequencies
 *   M    column-major N x N real (or complex) matrix per spin (leading dimension ld)
 * Complex arrays are interleaved (re, im) doubles.
 */
#ifndef TP_ORACLE_H
#define TP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tpo_geometry {
  int nb;             /* bands */
  int nc;             /* cluster sites */
  double beta;        /* inverse temperature */
  int nw;             /* compact vertex frequencies (= 2 * four-point-fermionic-frequencies) */
  int n_wex;          /* exchange frequencies */
  const int* wex;     /* exchange frequency indices (may be negative for a single transfer) */
  int n_kex;          /* exchange momenta */
  const int* kex;     /* exchange momentum indices into the k cluster */
  int n_channels;     /* channels in G4 order (sorted, unique FourPointType values) */
  const int* channels;
  const int* k_add;   /* nc*nc, k_add[i + nc*j] = index(k_i + k_j) */
  const int* k_sub;   /* nc*nc, k_sub[i + nc*j] = index(k_j - k_i) */
  const double* dft;  /* nc*nc complex, dft[k + nc*r] = exp(i k.r) */
  const double* band_phase; /* nb*nc complex, exp(i k.a_b) at [b + nb*k]; NULL = no displacements */
  const double* g0;   /* nb*nb*2*nc*nwE complex, layout (b1, b2, s, k, wE) */
} tpo_geometry;

typedef struct tpo_acc tpo_acc;

/* Returns NULL and writes a message to err (if non-NULL, size errlen) on failure. */
tpo_acc* tpo_create(const tpo_geometry* g, char* err, int errlen);
void tpo_destroy(tpo_acc* a);
/* keep = 0: do not copy the M_r_r_w_w intermediate for tpo_get_Mrrww (timing runs). */
void tpo_keep_intermediate(tpo_acc* a, int keep);
/* Complex GEMMs through BLAS zgemm_ from the shared library at path (single-threaded), or the
 * oracle's own loop for path == NULL. Returns 1 if zgemm_ was loaded. Process-wide. */
int tpo_use_blas(const char* path);
/* Which GEMM the oracle uses. */
const char* tpo_blas_name(void);
void tpo_reset(tpo_acc* a);

/* One measurement. n[s] vertices for spin s with times tau[s], sites site[s], bands band[s];
 * m[s] column-major with leading dimension ld[s]. Returns the reference's flop counter
 * (tp_accumulator_cpu.hpp:241-249), or a negative value for an unsupported channel. */
double tpo_accumulate(tpo_acc* a, const int n[2], const double* const m[2], const int ld[2],
                      const double* const tau[2], const int* const site[2],
                      const int* const band[2], double sign);

/* Parameters::model_type::lattice_type::spin_symmetric (tp_accumulator_base.hpp:167); 0 selects
 * the non-spin-symmetric updateG4 branch (tp_accumulator_cpu.hpp:668-757: pp and magnetic only,
 * reading spin sector 0). Default 1. */
void tpo_set_spin_symmetric(tpo_acc* a, int spin_symmetric);

/* accumulate for the general model: m[s] real (m_complex = 0) or interleaved complex; each vertex
 * has right labels (site, band) and left labels (site_l, band_l; NULL arrays: the same, i.e.
 * density-density vertices, cached_ndft_base.hpp:72-78); the sign is a complex phase. */
double tpo_accumulate_general(tpo_acc* a, const int n[2], const double* const m[2], int m_complex,
                              const int ld[2], const double* const tau[2], const int* const site[2],
                              const int* const band[2], const int* const site_l[2],
                              const int* const band_l[2], double sign_re, double sign_im);
void tpo_compute_G_general(tpo_acc* a, const int n[2], const double* const m[2], int m_complex,
                           const int ld[2], const double* const tau[2], const int* const site[2],
                           const int* const band[2], const int* const site_l[2],
                           const int* const band_l[2]);

/* Only the M -> G step (computeM + computeG). The result is readable with tpo_get_G. */
void tpo_compute_G(tpo_acc* a, const int n[2], const double* const m[2], const int ld[2],
                   const double* const tau[2], const int* const site[2], const int* const band[2]);

/* The M_r_r_w_w intermediate of the last M -> G step, layout (r1, r2, b1, b2, s, w1E, w2E). */
long long tpo_get_Mrrww(const tpo_acc* a, double* out, long long n_complex);
long long tpo_get_G(const tpo_acc* a, double* out, long long n_complex);
long long tpo_get_G4(const tpo_acc* a, int channel_index, double* out, long long n_complex);
long long tpo_g4_size(const tpo_acc* a);   /* complex entries per channel */
long long tpo_g_size(const tpo_acc* a);    /* complex entries of G */

/* dst.G4 += src.G4, then src's G4 is cleared (tp_accumulator_cpu.hpp:895-904). 0 on success. */
int tpo_sum_to(tpo_acc* src, tpo_acc* dst);

#ifdef __cplusplus
}
#endif

#endif
