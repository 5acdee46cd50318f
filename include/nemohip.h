/*
 * libnemohip: MI355X (gfx950) implementation of the NEMO-3D sp3d5s*
 * eigensolver hot path behind a plain C ABI.
 *
 * Conventions (same as the reference operator boundary):
 *   - complex numbers are interleaved {double re, im} (nml_dcscalar /
 *     `complex`, NEMO_3D/nml/include, layout-compatible with nh_cplx);
 *   - vectors are in Hamiltonian row order: atom-major, 10 (spinless) or 20
 *     (spin up orbitals 0-9 then spin down 10-19) rows per atom;
 *   - the operator is y = (H - shift I) x and y is fully overwritten
 *     (Hmult_spds_complete, NEMO_3D/nemo3d/src/base/h_cvectr_mult.c:262-667).
 * Every entry point returns an nh_status instead of aborting (the
 * reference calls die()); nh_last_error() gives the message.
 */
#ifndef NEMOHIP_H
#define NEMOHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nh_cplx {
  double re, im;
} nh_cplx;

typedef enum nh_status {
  NH_OK = 0,
  NH_ERR_ARG = 1,     /* invalid argument or shape */
  NH_ERR_HIP = 2,     /* HIP runtime error */
  NH_ERR_STATE = 3,   /* call out of order (e.g. H·v before a build) */
  NH_ERR_NOMEM = 4,   /* device allocation failed */
  NH_ERR_NOCONV = 5   /* solver stopped before convergence */
} nh_status;

/* Band models (Dev.band_model; basis_init, ham3d_final.c:576-620). */
enum { NH_BANDS_10_SP3D5SS_NOSPIN = 10, NH_BANDS_20_SP3D5SS_SPIN = 20 };

/* Storage layouts of the assembled Hamiltonian. */
enum {
  NH_LAYOUT_BLOCKS = 0, /* explicit 10x10 real block per (atom, neighbour) */
  NH_LAYOUT_SK = 1,     /* per-bond direction cosines + scaled integrals */
  NH_LAYOUT_MATRIX_FREE = 2 /* bond geometry only; SK + strain every call */
};
/* Arithmetic of the stored values and of the per-block products. */
enum { NH_PREC_FP64 = 0, NH_PREC_FP32 = 1 };
/* Where vector arguments live. */
enum { NH_PTR_HOST = 0, NH_PTR_DEVICE = 1 };

/* Model: what setup_hamiltonian (ham3d_final.c:731-1230) hands to the
 * assembly.  All arrays are host arrays, copied at context creation. */
typedef struct nh_model {
  int32_t nb;            /* 10 or 20 */
  int32_t n_atoms;
  int32_t n_species;
  const double *parmat;  /* [S][S][35] indexed [cation][anion] (param_init,
                            mat3d.c:133-307) */
  const double *strnmat; /* [S][S][33] indexed [cation][anion] */
  const int32_t *species;   /* [A] */
  const int32_t *is_cation; /* [A] 1 on the cation role */
  const int32_t *nbr;       /* [A][4] neighbour atom, -1 dangling */
  const int32_t *bond_q;    /* [A][4][3] ideal this->nbr vector, units a/4 */
  const double *disp;       /* [A][3] displacement from the ideal site, nm */
  double lattice[3];        /* cubic cell lengths, nm */
  int32_t passivate;        /* 1: PassivateDanglingBond */
  double db_shift;          /* DanglingBondEnergyShift, eV */
  const double *phi;        /* [A] potential subtracted from the diagonal,
                               or NULL (Phys.PotentialFile) */
} nh_model;

typedef struct nh_context nh_context;

/* Context: copies the model to the device (setup_hamiltonian's outputs). */
nh_status nh_context_create(const nh_model *model, int device,
                            nh_context **out);
/* Releases the context, including the solvers' device workspace: buffers a
 * solve releases stay allocated in the context for the next solve (a free
 * of tens of GB can stall for a second or more). */
void nh_context_destroy(nh_context *ctx);
int64_t nh_context_dim(const nh_context *ctx); /* N = n_atoms * nb */

/* Kernel variants (chosen by measured throughput; the same products to
 * rounding: the compiler contracts the block's multiply-adds differently
 * in the two).  NH_SET_HV_STAGED: 1 stages the neighbours'
 * vector entries in LDS (stored-block layout: per workgroup; SK layout: the
 * four lanes of an atom load each neighbour's entries together), 0 reads
 * them from the caches by every lane, -1 (default) picks by layout (SK:
 * staged, stored blocks: direct). */
enum { NH_SET_HV_STAGED = 1 };
/* NH_SET_RING_STEPS: Lanczos steps between eigenvector accumulations in the
 * eigenvector pass (ring length; 8, 16, 32, 64 or 128), or 0 (default): 128,
 * halved down to 8 until the ring takes at most half the free memory.
 * The eigenvectors agree to rounding whatever the length. */
enum { NH_SET_RING_STEPS = 2 };
/* NH_SET_BLOCK_HV: products of several vectors (block H·V): -1 (default)
 * the matrix-core kernel for the row-owner block layout and 0 for the SK
 * and matrix-free layouts (the fastest measured there), 0 one column per
 * grid row, 1 all columns in one launch with the bond records read once
 * (SK and matrix-free), 2 the matrix-core kernel (FP32 SK or block
 * layouts, 8 columns). */
enum { NH_SET_BLOCK_HV = 3 };
/* NH_SET_SK_ROLES: workgroups of the SK and matrix-free products: 1 holds
 * cation-role and anion-role atoms of the same cells together (both rows of
 * a bond read its record in one workgroup), 0 one role per workgroup (the
 * two roles of the same cells back to back on one XCD), -1 the default
 * (1 with FP64 products, 0 with FP32: the faster of the two measured).
 * The products are identical either way. */
enum { NH_SET_SK_ROLES = 4 };
/* NH_SET_SK_LANES: lanes per atom of the SK and matrix-free Lanczos
 * products of 20-band models: 4 (a lane per bond, both spin blocks) or 8
 * (a lane per bond and spin block: half the registers per lane); -1 the
 * default (8 for matrix-free FP32 products, 4 otherwise: the faster of the
 * two measured). */
enum { NH_SET_SK_LANES = 5 };
/* NH_SET_LZ_VECTORS: storage of the Lanczos vectors r_k of a Lanczos run
 * with FP32 SK or matrix-free products: 0 FP64 (with an FP32 copy for the
 * products' gathers), 1 FP32: the step runs in Paige's order, u_k = H q_k
 * - beta_k q_{k-1} computed in FP64 and rounded once to FP32, r_{k+1} =
 * u_k - (alpha_k / beta_k) r_k as two FP32 fmas with the coefficient split
 * into two FP32 parts, and the product of step k
 * forms r_k from u_{k-1} and r_{k-1} as it reads them, so a step is one
 * launch (no FP64 vector in either pass; alpha, beta and the sums stay
 * FP64); -1 the default (1 where the products use the
 * LDS gather: measured faster, with the same number of steps).  Either
 * way the eigenvector pass retraces the eigenvalue pass; other layouts
 * and FP64 products always keep FP64 vectors. */
enum { NH_SET_LZ_VECTORS = 6 };
nh_status nh_context_set(nh_context *ctx, int what, int value);

/* Assemble H on the device in the given layout and precision.  Replaces
 * the first-call assembly of Hmult_spds_col (h_cvectr_mult.c:1134-1886)
 * with its storage into HamZincBlende (Hmult_spds_col_prep_store,
 * h_cvectr_mult.c:1888-2082). */
nh_status nh_ham_build(nh_context *ctx, int layout, int precision);

/* Switch the arithmetic of later products between FP64 and FP32 without
 * re-assembling (the FP64 values are always kept; the FP32 copies are made
 * on first use).  The mixed-precision solvers run their bulk iterations in
 * FP32 and their final refinement in FP64 on one context.  No reference
 * counterpart (the reference picks FP32 storage at compile time,
 * HamMemorySave). */
nh_status nh_ham_set_precision(nh_context *ctx, int precision);

/* y = (H - shift I) x.  Replaces Hmult_spds_complete
 * (h_cvectr_mult.c:262-667) / Hmult_spds_col_stored (Hmult_stored.c:73-172). */
nh_status nh_hv(nh_context *ctx, double shift, const nh_cplx *x, nh_cplx *y,
                int ptr_kind);

/* Y = (H - shift I) X for ncol column vectors stored with leading
 * dimension ld (>= N) in elements.  Replaces Hmatmult_spds_complete
 * (h_cmat_mult.c:85-...) / Hmatmult_spds_col_stored
 * (Hmult_stored.c:2359-2474). */
nh_status nh_hV(nh_context *ctx, double shift, const nh_cplx *X, nh_cplx *Y,
                int ncol, int64_t ld, int ptr_kind);

/* Copy the assembled FP64 Hamiltonian back (row-owner layout): onsite
 * [A][17] = 10 diagonal energies, 6 surface s-p couplings
 * (1,2),(1,3),(1,4),(2,3),(2,4),(3,4), spin-orbit lambda; blocks
 * [A][4][10][10] with y_i += blocks[i][n] x_nbr(i,n) per spin.  For the
 * parity tests; requires a build in NH_LAYOUT_BLOCKS / NH_PREC_FP64. */
nh_status nh_ham_export(nh_context *ctx, double *onsite, double *blocks);

/* ---- device buffers -------------------------------------------------------
 * The library owns its device memory (it uses the system ROCm runtime, so
 * it does not share allocations with other HIP runtimes in the process).
 * Solver workspace is kept in a per-context pool between solves (no
 * reference counterpart: the reference allocates host memory per solve);
 * every allocation of the library, nh_device_alloc included, first hands
 * the idle pool blocks of all contexts on the device back when the device
 * is full.  nh_context_mem_info reports the device's free bytes and the
 * idle pool bytes of all contexts on it (either pointer may be NULL). */
nh_status nh_context_mem_info(nh_context *ctx, int64_t *free_bytes,
                              int64_t *pool_idle_bytes);
nh_status nh_device_alloc(nh_context *ctx, int64_t bytes, void **dptr);
nh_status nh_device_free(nh_context *ctx, void *dptr);
nh_status nh_memcpy_h2d(nh_context *ctx, void *dst, const void *src,
                        int64_t bytes);
nh_status nh_memcpy_d2h(nh_context *ctx, void *dst, const void *src,
                        int64_t bytes);

/* ---- the reference's random start vectors (host, no device work) -------
 * NH_RANDOM_RAND: out[j] = (rand() / RAND_MAX, 0), j < n, from glibc's
 * rand() after srand(seed[0]) (lanczos_initialize Random,
 * eigsys3d_par.c:125-131; n = p N gives the block start of
 * lanczos_blk.c:2220-2223).  NH_RANDOM_ZLARNV: LAPACK zlarnv(1, seed, n,
 * out), seed[0..3] in 0..4095 with seed[3] odd (qtracemin_driver.c:
 * 105-116); the seed is not updated. */
enum { NH_RANDOM_RAND = 0, NH_RANDOM_ZLARNV = 1 };
nh_status nh_random_start(int kind, const int32_t *seed, int64_t n,
                          nh_cplx *out);

/* ---- vector kernels (device pointers, length n complex) ---------------
 * Reductions accumulate in FP64 (vect_dot_vect, h_cvectr_mult.c:138-154;
 * Lanczos inner products, lanczos.c:679-754). */
nh_status nh_dot(nh_context *ctx, int64_t n, const nh_cplx *x,
                 const nh_cplx *y, nh_cplx *out /* host: sum conj(x) y */);
/* nh_dot without any allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * partial sums live in a fixed file-scope device array (2 x 2048 + 2
 * doubles, 32 KB per device, held until the process ends) whose address
 * is resolved once per device by nh_context_create; any n is accepted (at
 * most 2048 partial sums are formed).  x, y: device pointers; out: host.
 * Every call on one device shares that array, so calls that overlap on one
 * device (other host threads, other contexts of the device) race. */
nh_status nh_dot_noalloc(nh_context *ctx, int64_t n, const nh_cplx *x,
                         const nh_cplx *y, nh_cplx *out);
nh_status nh_nrm2(nh_context *ctx, int64_t n, const nh_cplx *x,
                  double *out /* host */);
/* y = a x + b y (a, b complex) */
nh_status nh_axpby(nh_context *ctx, int64_t n, nh_cplx a, const nh_cplx *x,
                   nh_cplx b, nh_cplx *y);

/* ---- solvers ----------------------------------------------------------- */

/* Lanczos options: the Lanczos_struct fields of the reference deck
 * (XMLtoC/apps/ioNemo3D/src/Lanczos_struct.H:8-30). */
typedef struct nh_lanczos_opts {
  int32_t max_iter;             /* MaxIter */
  double conv_tolerance;        /* ConvTolerance */
  int32_t num_eig_vb;           /* NumEigReq_vb */
  int32_t num_eig_cb;           /* NumEigReq_cb */
  double erange_vb[2];          /* Erange_vb */
  double erange_cb[2];          /* Erange_cb */
  int32_t conv_check_start;     /* ConvCheckStartIter */
  int32_t conv_check_skip;      /* ConvCheckSkipRate */
  int32_t start;                /* 0 CyclicRandom, 1 Const1, 2 user vector
                                   (host), 3 user vector (device), 4 Random
                                   ((rand() / RAND_MAX, 0) per entry after
                                   srand(random_seed), eigsys3d_par.c:
                                   125-131); normalised in device memory */
  int32_t cyclic_len;           /* CyclicString length */
  const double *cyclic_string;  /* CyclicString values */
  const nh_cplx *start_vector;  /* start == 2 / 3: host / device vector of
                                   length N */
  int32_t stop_rule;            /* 0: stop once num_eig converged values lie
                                   in each window (lanczos.c:1231-1232);
                                   1: stop once the num_eig Ritz values
                                   nearest the gap are all converged */
  int32_t want_vectors;         /* compute eigenvectors (second pass) */
  uint32_t random_seed;         /* start == 4: srand() seed (glibc rand) */
  /* restart (restoreState, lanczos.c:1905-1960): with restore_len > 0 the
   * eigenvalue pass is skipped and the tridiagonal alpha[0..len),
   * beta[0..len) of an earlier run (nh_eig_result.tridiag_*, host arrays)
   * is analysed instead; the eigenvector pass then starts from the start
   * vector these options describe, which must be that run's */
  const double *restore_alpha;
  const double *restore_beta;
  int32_t restore_len;
} nh_lanczos_opts;

typedef struct nh_eig_result {
  int32_t n_found;       /* converged pairs written */
  int32_t iterations;    /* Lanczos iterations of the eigenvalue pass */
  int32_t matvecs;       /* H·v products, all passes */
  double *eigenvalues;   /* caller array, capacity `capacity` */
  double *residuals;     /* caller array: ||H psi - E psi|| / ||psi|| */
  nh_cplx *eigenvectors; /* caller array capacity*N (host or device
                            memory), or NULL */
  int32_t capacity;
  double seconds;        /* wall time of the solve */
  double hv_ms;          /* mean duration of the fused H·v kernel (HIP
                            events on the solver stream, sampled launches) */
  int32_t hv_timed;      /* number of timed H·v launches */
  /* optional: the tridiagonal of the eigenvalue pass (alpha[k], beta[k]),
     caller arrays of capacity tridiag_cap, like the reference's saved
     nd_tridiag_bin (lanczos.c:1692-1809) */
  double *tridiag_alpha;
  double *tridiag_beta;
  int32_t tridiag_cap;
  /* Lanczos: the same for the fused H·v kernel of the eigenvector pass
     (it also applies the three-term step) */
  double hv_ms_vec;
  int32_t hv_timed_vec;
} nh_eig_result;

/* Lanczos without reorthogonalisation, eigenvalues from the tridiagonal
 * in the vb/cb windows, eigenvectors by a second pass
 * (eigensolve_lanczos, NEMO_3D/eigen/src/lanczos.c:1823-2089). */
nh_status nh_eigensolve_lanczos(nh_context *ctx, const nh_lanczos_opts *opts,
                                nh_eig_result *res);

/* Block Lanczos options: the BLanczos deck block (BL_size, ...) plus the
 * Lanczos windows. */
typedef struct nh_blanczos_opts {
  int32_t block_size;           /* BL_size (1..16) */
  int32_t max_iter;             /* block steps */
  double conv_tolerance;        /* residual estimate ||B_m s_m|| <= tol*||T|| */
  int32_t num_eig_vb;
  int32_t num_eig_cb;
  double erange_vb[2];
  double erange_cb[2];
  int32_t conv_check_start;     /* block steps before the first check */
  int32_t conv_check_skip;      /* block steps between checks */
  uint64_t seed;                /* start block without start_block: p vectors
                                   (rand() / RAND_MAX, 0) in turn after
                                   srand(seed) (lanczos_blk.c:2214-2229) */
  const nh_cplx *start_block;   /* optional host block [block_size][N] */
  int32_t stop_rule;            /* as nh_lanczos_opts.stop_rule; counts
                                   distinct levels (degenerate copies once) */
  int32_t want_vectors;
  int32_t deflation_on;         /* DeflationOn: drop a block vector whose
                                   norm after orthogonalisation falls below
                                   deflation_tol (deflation_test,
                                   lanczos_blk.c:225-268); the block then
                                   shrinks */
  double deflation_tol;         /* DefTolerance (the deck default is 1e-7) */
} nh_blanczos_opts;

/* Block Lanczos without reorthogonalisation (eigensolve_blk_lanczos,
 * lanczos_blk.c:2145-...): block three-term recurrence with a QR of each
 * new block, Ritz values of the block tridiagonal in the windows, Ritz
 * vectors by a second pass, FP64 residuals.  Degenerate copies of a level
 * (Kramers pairs) count once and give one eigenvector, as in the
 * single-vector solver. */
nh_status nh_eigensolve_block_lanczos(nh_context *ctx,
                                      const nh_blanczos_opts *opts,
                                      nh_eig_result *res);

/* PQ-TRACEMIN options (PQTracemin deck block: NumEigVal, RefEigVal,
 * Tolerance, cgtol, maxi, cgmaxim). */
typedef struct nh_tracemin_opts {
  int32_t num_eig;            /* NumEigVal: eigenvalues nearest ref_eig */
  int32_t block_size;         /* trace-minimisation block (>= num_eig), 0: 2x */
  double ref_eig;             /* RefEigVal sigma */
  double tolerance;           /* ||H x - lambda x|| <= tolerance * hnorm */
  double hnorm;               /* scale of the residual test (e.g. ||H||) */
  double cgtol;               /* inner CG relative stopping bound */
  int32_t max_outer;          /* maxi */
  int32_t cg_max;             /* cgmaxim */
  uint64_t seed;              /* random block: zlarnv(1, iseed) with iseed =
                                 {103 + seed, 101 seed, 7, 57} mod 4096, the
                                 reference's seed of rank `seed`
                                 (qtracemin_driver.c:105-116) */
  const nh_cplx *start_block; /* optional host block [block_size][N] */
} nh_tracemin_opts;

/* PQ-TRACEMIN on (H - sigma)^2 (qtracemin_driver.c:9-947,
 * pqtracemin_functions.f90:182-...): eigenpairs of H nearest sigma with
 * FP64 Ritz values and residuals.  Eigenvectors are returned when
 * res->eigenvectors is set. */
nh_status nh_eigensolve_tracemin(nh_context *ctx, const nh_tracemin_opts *opts,
                                 nh_eig_result *res);

/* PC-TRACEMIN options (PCTracemin deck block: NumEigVal, WantEigRange,
 * AllEigRange, ChebyDeg, Tolerance, maxi; ctracemin_driver.c:13-40). */
typedef struct nh_pctracemin_opts {
  int32_t num_eig;            /* NumEigVal: eigenpairs wanted in want_range */
  int32_t block_size;         /* block (num_eig..16), 0: num_eig + 4 */
  double want_range[2];       /* WantEigRange [a, b] */
  double all_range[2];        /* AllEigRange: bounds of the whole spectrum */
  int32_t cheb_degree;        /* ChebyDeg k of T_k(1 - gamma (H-a)(H-b)) */
  double tolerance;           /* ||H x - lambda x|| <= tolerance * hnorm */
  double hnorm;               /* scale of the test; <= 0: max |AllEigRange| */
  int32_t max_outer;          /* maxi */
  uint64_t seed;
  const nh_cplx *start_block; /* optional host block [block_size][N] */
} nh_pctracemin_opts;

/* PC-TRACEMIN (ctracemin_driver.c:10-..., map_chebyshev_shift /
 * zgtracemin, pctracemin_functions.f90:154-332): subspace iteration on
 * A = T_k(q(H)) + 2I, q(x) = 1 - gamma (x-a)(x-b), gamma from the spectrum
 * bounds as in the reference, Rayleigh-Ritz of H on the block and FP64
 * residuals; returns the num_eig Ritz pairs in [a, b] (both members of a
 * Kramers pair count). */
nh_status nh_eigensolve_pctracemin(nh_context *ctx,
                                   const nh_pctracemin_opts *opts,
                                   nh_eig_result *res);

/* FP64 refinement of approximate eigenpairs, the residual-refinement stage
 * of the mixed-precision solver (Ritz pairs of a run with FP32 products
 * are eigenpairs of the FP32-rounded operator).  Each round builds a block
 * Krylov basis [X, H X, ..., H^depth X] (Arnoldi, classical Gram-Schmidt
 * applied twice, Cholesky QR) and keeps, for every current eigenvalue, the
 * nearest FP64 Ritz pair of the basis; rounds stop once every
 * ||H x - e x|| <= tolerance * hnorm (||x|| = 1).  X [k][N] (host or
 * device per ptr_kind, k <= 16) is overwritten with the refined vectors,
 * eigenvalues (in: estimates) with the Ritz values, residuals (out) with
 * ||H x - e x||.  The context must hold the FP64 operator.  No reference
 * counterpart: the reference reports the residuals of its FP64 Ritz pairs
 * (compute_ritz_residual_norm_deviation, nemo3d/src/base/h_cvectr_mult.c:
 * 78-...) and refines nothing.  Returns NH_ERR_NOCONV when
 * max_rounds end first (results still written). */
typedef struct nh_refine_opts {
  int32_t depth;       /* Krylov blocks per round, 1..32 */
  int32_t max_rounds;
  double tolerance;
  double hnorm;
} nh_refine_opts;
nh_status nh_refine_eigenpairs(nh_context *ctx, const nh_refine_opts *opts,
                               int32_t k, nh_cplx *X, int ptr_kind,
                               double *eigenvalues, double *residuals,
                               int32_t *rounds, int32_t *matvecs);

/* ---- valence force field ------------------------------------------------
 * Keating VFF (calc_estrain / min_estrain, strain3d.c:157-475,1361-2245):
 * alpha, beta, a0 are [S][S] tables (Keating constants and unstrained
 * lattice constant of each species pair).  Atoms with fewer than four
 * neighbours stay fixed. */
/* Energy and gradient at displacements disp[A][3] (host arrays). */
nh_status nh_vff_energy_grad(nh_context *ctx, const double *alpha,
                             const double *beta, const double *a0,
                             const double *disp, double *grad_out,
                             double *energy);
/* Relax the context's displacements in place by conjugate gradients until
 * ||grad|| <= gtol or max_iter; disp_out (host [A][3], may be NULL)
 * receives the result; info[5] = {iterations, E0, E, |g0|, |g|}.  Rebuild
 * the Hamiltonian afterwards. */
nh_status nh_vff_relax(nh_context *ctx, const double *alpha,
                       const double *beta, const double *a0, int max_iter,
                       double gtol, double *disp_out, double *info);

/* The strain models and boundaries of the reference's strain block
 * (Strain_struct: StrainModel, Keating):
 *   NH_VFF_KEATING          alpha, beta constant (VFF_keating);
 *   NH_VFF_KEATING_STRAINED alpha and beta scaled with the bond's strain,
 *     StrainParam::alpha_eff / beta_eff (StrainParam.c:60-214) with the
 *     acoustic anharmonicity constants aa, ba, ca (the material's
 *     strain_VFF[0..2], StrainParam.c:265-280); the effective constants
 *     are held fixed in the gradient, as calc_estrain does
 *     (strain3d.c:1687-1700, 1835-1860);
 *   fix_surface 1: atoms with fewer than four neighbours stay fixed
 *     (FixSurfaceAll); 0: every atom moves (AllFree).
 * aa, ba, ca may be NULL for NH_VFF_KEATING. */
enum { NH_VFF_KEATING = 0, NH_VFF_KEATING_STRAINED = 1 };
typedef struct nh_vff_params {
  const double *alpha, *beta, *a0; /* [S][S] */
  const double *aa, *ba, *ca;      /* [S][S] */
  int32_t model;
  int32_t fix_surface;
} nh_vff_params;
/* nh_vff_energy_grad / nh_vff_relax with the model and boundary given. */
nh_status nh_vff_energy_grad_model(nh_context *ctx, const nh_vff_params *p,
                                   const double *disp, double *grad_out,
                                   double *energy);
nh_status nh_vff_relax_model(nh_context *ctx, const nh_vff_params *p,
                             int max_iter, double gtol, double *disp_out,
                             double *info);

/* ---- learned warm start: 3-D CNN on the cation grid ----------------------
 * L conv layers 3x3x3 (width 32, ReLU) + 1x1x1 head; FP32 master weights,
 * BF16 weights in the products, FP32 accumulation.  Grids are
 * [gx][gy][gz][channels] (channels last).  No counterpart in the
 * reference: it supplies the start block that the reference fills with
 * random numbers (lanczos_blk.c:2214-2229, qtracemin_driver.c:112-116). */
typedef struct nh_cnn nh_cnn;
nh_status nh_cnn_create(nh_context *ctx, int c_in, int width, int layers,
                        int c_out, uint64_t seed, nh_cnn **out);
void nh_cnn_destroy(nh_cnn *cnn);
int64_t nh_cnn_num_params(const nh_cnn *cnn);
nh_status nh_cnn_get_params(nh_cnn *cnn, float *host);
nh_status nh_cnn_set_params(nh_cnn *cnn, const float *host);
/* out[gx*gy*gz][c_out] (host) */
nh_status nh_cnn_forward(nh_cnn *cnn, const float *feat, int gx, int gy,
                         int gz, float *out);
/* gradient of sum(out * dout) with respect to the parameters (host) */
nh_status nh_cnn_backward(nh_cnn *cnn, const float *feat, int gx, int gy,
                          int gz, const float *dout, float *grad);
/* One Adam step on one structure of the context.  targets [n_targets][N]
 * (host or device, n_targets <= 16): eigenvectors of the plain solver, with their
 * Kramers partners; cluster[n_targets] labels each target's degenerate
 * subspace and out_cluster[n_states] (n_states <= 8) the subspace each
 * predicted column should span.  Loss = mean over columns of
 * 1 - |projection onto the subspace|^2 / |column|^2 (phase- and
 * degeneracy-invariant).  lr = 0: no update.  grad (host [n_params], or
 * NULL) receives the parameter gradient. */
nh_status nh_cnn_train_step(nh_cnn *cnn, const float *feat, int gx, int gy,
                            int gz, const int32_t *atom_point,
                            const int32_t *atom_slot, const nh_cplx *targets,
                            int n_targets, const int32_t *cluster,
                            int n_states, const int32_t *out_cluster,
                            double lr, double *loss, float *grad);
/* Rebind the network to another context (another structure) on the same
 * device: training runs over many structures with one network. */
nh_status nh_cnn_bind(nh_cnn *cnn, nh_context *ctx);
/* Start block [n_states][N] predicted for the context's structure. */
nh_status nh_cnn_start_block(nh_cnn *cnn, const float *feat, int gx, int gy,
                             int gz, const int32_t *atom_point,
                             const int32_t *atom_slot, int n_states,
                             nh_cplx *block, int ptr_kind);

/* ---- operator callbacks with the reference signatures -------------------
 * For the reference's unchanged solver drivers.  Host vectors; the context
 * is bound with nh_shim_bind (argList[0] / `d` are not read).  Like the
 * functions they replace these abort with a message on error. */
nh_status nh_shim_bind(nh_context *ctx);
/* The wave-function table (d->WF: energies and vectors [nwf][N], host) that
 * matvect_mult's spectral projection indexes through eval_scoperti. */
nh_status nh_shim_set_wavefunctions(int nwf, const double *energy,
                                    const nh_cplx *psi);
/* matvect_mult (h_cvectr_mult.c:205-258): argList[1] = &shift,
 * argList[2] = &Nproj, argList[3] = &eval_scoperti when Nproj > 0;
 * y = (H - shift) x - sum_j E_j <psi_j|x> psi_j. */
void nh_matvect_mult(const int **argList, nh_cplx *y, const nh_cplx *x);
/* matmat_mult (h_cmat_mult.c:21-83): argList[1] = &shift,
 * argList[2] = &num_col; x, y are nml_dcmatrix row-pointer arrays (one row
 * per vector).  Nproj is not applied, as in the reference. */
void nh_matmat_mult(const int **argList, nh_cplx **y, nh_cplx *const *x);
/* Hmult_spds_complete (h_cvectr_mult.c:262) and Hmatmult_spds_complete
 * (h_cmat_mult.c:85); `d` is ignored. */
void nh_Hmult_spds_complete(nh_cplx *y, void *d, double shift, nh_cplx *x);
void nh_Hmatmult_spds_complete(nh_cplx **Y, void *d, double shift,
                               nh_cplx **X, int num_col);

const char *nh_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
