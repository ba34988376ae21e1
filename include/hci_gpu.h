/*
 * hci_gpu.h -- C ABI of the MI355X (gfx950) HCI variational hot path.
 *
 * The library builds the sparse variational Hamiltonian (connection search over
 * alpha/beta determinant bitstrings + Slater-Condon elements) and runs the
 * Davidson eigensolver on it.  It replaces, on the host side of the reference
 * (aaholmes/risq):
 *   - gen_sparse_ham_fast_lookup   src/var/fast_ham.rs:527-575   -> hci_build_h
 *   - SparseMatUpperTri + MatrixOperations::matrix_vector_prod
 *                                  src/var/sparse.rs:24-88,
 *                                  src/var/eigenvalues/matrix_operations.rs:14-27
 *                                                                -> hci_sigma
 *   - Davidson::new (DPR, lowest root) as driven by sparse_optimize
 *                                  src/var/eigenvalues/algorithms/davidson.rs:106-277,
 *                                  src/var/davidson.rs:43-113    -> hci_davidson
 *   - Ham::ham_diag / ham_off_diag_no_excite (element consumers: selection, PT2)
 *                                  src/ham/mod.rs:92-124,256-303 -> hci_ham_diag,
 *                                                                   hci_elements
 *
 * Conventions
 *   - Every function returns 0 (HCI_OK) on success and a negative code on
 *     failure; hci_davidson returns HCI_NOT_CONVERGED (1) when the iteration
 *     limit is hit (the reference's Err(DavidsonError)).  No exception or abort
 *     crosses the ABI.  hci_last_error gives the message of the context's last
 *     failure.
 *   - Determinants are 4 uint64 words each: [up_lo, up_hi, dn_lo, dn_hi]
 *     (the reference's Config{up: u128, dn: u128}, src/wf/det.rs:16-19), bit p
 *     = spatial orbital p.  All determinants of one call must carry the same
 *     numbers of up and down electrons, and must be distinct.
 *   - Integral tables use the reference layout (src/utils/ints.rs:38-68):
 *     one_body[combine_2(p,q)], two_body[combine_4(p,q,r,s)], 1-based.
 *   - Arrays passed in are borrowed for the duration of the call.  `mem`
 *     arguments say where they live: HCI_MEM_HOST or HCI_MEM_DEVICE (device
 *     pointers on the context's GPU).  Handles and device buffers are owned by
 *     the library.  Calls on one context are serialised by the caller; the
 *     library uses one HIP stream per context.
 */
#ifndef HCI_GPU_H
#define HCI_GPU_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HCI_OK 0
#define HCI_NOT_CONVERGED 1
#define HCI_ERR_ARG (-1)
#define HCI_ERR_HIP (-2)
#define HCI_ERR_OOM (-3)
#define HCI_ERR_STATE (-4)

#define HCI_MEM_HOST 0
#define HCI_MEM_DEVICE 1

/* value storage of the sparse H */
#define HCI_PREC_F64 64 /* int32 col + float64 value (12 B / nonzero) */
#define HCI_PREC_F32 32 /* int32 col + float32 hi + float32 lo (hi read alone in
                           mixed sigmas, hi+lo in refinement sigmas) */
#define HCI_PREC_F16 16 /* HCI_STORE_TILED only: FP32 hi + lo as for HCI_PREC_F32 plus an
                           FP16 copy of every value (2 B more per entry), read alone in
                           the first phase of mixed = 2 Davidson runs */

/* how the matrix is held on the device (both keep the reference's pair set;
 * rows/columns are renumbered internally in alpha-major string order) */
#define HCI_STORE_FULL 0  /* both triangles: sigma without atomics, bitwise
                             reproducible, 2x the value memory */
#define HCI_STORE_UPPER 1 /* upper triangle only: sigma scatters the transpose
                             half with FP64 atomics */
#define HCI_STORE_TILED 2 /* both triangles in string-group tiles: a tile of rows
                             of one alpha (or beta) string group holds its
                             entries per column group with 16-bit local row and
                             column indices; sigma stages x slices in LDS and
                             sums rows in LDS with atomics in no fixed order, so
                             products (and Davidson iteration counts) are not
                             bitwise reproducible from run to run (they agree to
                             rounding).  An incremental build rebuilds the whole
                             tiled matrix; pairs among the previous
                             determinants keep the single rule they were stored
                             under.  Without FP16 copies, a dense window (at
                             least 2048 entries) whose rows and columns are two
                             tiles' whole row ranges is stored once for both
                             triangles (its transposed sums go through
                             per-window partial sums).  No hci_csr_row. */

/* which singles are stored (doubles are always stored iff H != 0,
 * src/var/off_diag.rs:33) */
#define HCI_SINGLES_ABS_GT_1E12 0 /* src/var/fast_ham.rs:505,510 (n > 100 path) */
#define HCI_SINGLES_NONZERO 1     /* src/var/off_diag.rs:33 (n <= 100 path) */

/* sigma precision */
#define HCI_SIGMA_FULL 0 /* stored values at full precision */
#define HCI_SIGMA_F32 1  /* FP32 values (hi word), FP64 x/y and accumulation */
#define HCI_SIGMA_F16 2  /* FP16 values (HCI_PREC_F16 storage), FP64 x/y and accumulation */

/* Davidson correction */
#define HCI_PRECOND_DPR 0 /* t = r / (lambda - H_ii), davidson.rs:391-408 */
#define HCI_PRECOND_MLP 1 /* learned MLP, DPR fallback when it does not reduce |r| */

typedef struct hci_ctx hci_ctx;
typedef struct hci_csr hci_csr;

typedef struct hci_build_opts {
    int value_prec;  /* HCI_PREC_F64 | HCI_PREC_F32 | HCI_PREC_F16 (tiled) */
    int single_rule; /* HCI_SINGLES_* */
    int mem;         /* where dets/diag live */
    int storage;     /* HCI_STORE_FULL | HCI_STORE_UPPER | HCI_STORE_TILED */
} hci_build_opts;

typedef struct hci_build_stats {
    int64_t nnz_u;          /* stored upper-triangle off-diagonals */
    int64_t n_alpha;        /* unique alpha strings */
    int64_t n_beta;         /* unique beta strings */
    int64_t alpha_pairs;    /* single-connected alpha-string pairs (ordered) */
    double seconds;         /* wall time of the call (device work included) */
    int64_t mirror_pairs;   /* HCI_STORE_TILED: dense window pairs stored as one half */
    int64_t stored_entries; /* entry slots held on the device (both triangles for FULL
                               and TILED; TILED's alignment slots included, and its
                               explicit zeros: singles the storage rule rejects whose
                               element is not zero for every occupation) */
    int64_t matrix_bytes;   /* device bytes of the entry arrays: indices and values
                               (FP64, FP32 hi + lo, FP16 copies), plus FULL/UPPER's
                               row pointers */
} hci_build_stats;

#define HCI_DAV_MAX_SUB 32
#define HCI_DAV_MAX_KEEP 8

typedef struct hci_dav_opts {
    double tol;        /* residual norm (reference coeff_eps = 1e-4, src/var/mod.rs:89) */
    double etol;       /* eigenvalue change (reference energy_eps = 1e-8, src/var/mod.rs:90) */
    int max_iters;     /* 50 (davidson.rs:75) */
    int max_sub;       /* subspace limit, reference nvalues + 10 = 11 (davidson.rs:64-68);
                          at most HCI_DAV_MAX_SUB */
    int precond;       /* HCI_PRECOND_* */
    int mixed;         /* 0: full-precision sigmas; 1: FP32 sigmas, then FP64 refinement;
                          2 (HCI_PREC_F16 storage): FP16 sigmas, then FP32, then refinement */
    double refine_tol; /* mixed: residual below which refinement starts; 0 (default):
                          when the FP32 iteration has converged or stalls */
    int mem;           /* where init / coeffs live */
    int restart_keep;  /* vectors kept at a restart: 1 = the Ritz vector, whose product
                          is recomputed (reference, davidson.rs:258-269); m >= 2 = the
                          m - 1 lowest Ritz vectors plus the previous iteration's Ritz
                          vector, formed inside the subspace with no new product;
                          m < max_sub, m <= HCI_DAV_MAX_KEEP */
    double f16_tol;    /* mixed = 2: residual below which the FP16 phase gives way to FP32
                          sigmas (restart from the Ritz vector), also when the FP16
                          iteration converges or stalls; 0 (default): 3 tol */
} hci_dav_opts;

typedef struct hci_dav_stats {
    int iters;          /* outer iterations */
    int n_sigma_f32;    /* FP32-value sigma products */
    int n_sigma_full;   /* full-precision sigma products */
    int converged;
    int mlp_applied;    /* iterations whose correction came from the MLP */
    int mlp_fallbacks;  /* iterations where the MLP was rejected for DPR */
    int refine_iter;    /* iteration at which refinement started (-1: never) */
    double residual;    /* final |r| */
    double seconds;
    double sigma_ms;    /* summed device time of the sigma products (HIP events) */
    int n_sigma_f16;    /* FP16-value sigma products (mixed = 2) */
    int f16_iter;       /* iteration at which FP32 sigmas took over (-1: never) */
    int orth_second;    /* corrections that needed the second Gram-Schmidt pass */
} hci_dav_stats;

/* context ---------------------------------------------------------------- */
int hci_create(int device, hci_ctx** out);
/* frees the context's integral tables, work vectors, recycled entry arrays
 * and the build work buffers it keeps between builds (at most 1 GB) */
void hci_destroy(hci_ctx* ctx);
const char* hci_last_error(const hci_ctx* ctx);
int hci_device_sync(hci_ctx* ctx);

/* Integrals (replaces Ham.ints, src/ham/read_ints.rs:33-42).  Copied into
 * library-owned device memory; n1/n2 are the table lengths, at least the
 * reference's allocation for n_orb (read_ints.rs:78-80).  n_orb <= 128. */
int hci_set_integrals(hci_ctx* ctx, int n_orb, double e_core, const double* one_body,
                      size_t n1, const double* two_body, size_t n2);

/* Ham::ham_diag for n determinants (src/ham/mod.rs:92-124). */
int hci_ham_diag(hci_ctx* ctx, const uint64_t* dets, int64_t n, double* diag, int mem);

/* Ham::ham_off_diag_no_excite(det1[m], det2[m]) (src/ham/mod.rs:256-303);
 * degree[m] (nullable) = excitation degree, 3 meaning "more than double". */
int hci_elements(hci_ctx* ctx, const uint64_t* det1, const uint64_t* det2, int64_t m,
                 double* value, int8_t* degree, int mem);

/* Sparse H build (gen_sparse_ham_fast_lookup, src/var/fast_ham.rs:527-575).
 * Rows in the caller's determinant order, upper triangle only, each row sorted
 * by column with no duplicates (sparse.rs:44-56); diagonal taken from `diag`
 * (new_sparse_ham / expand_sparse_ham_rows, src/wf/mod.rs:1646-1683).
 * Incremental form: with prev != NULL (built over dets[0..n_prev)) and
 * n_stored == n_prev, prev's rows are kept and only pairs with at least one
 * index >= n_stored are added (fast_ham.rs:397,429,444-447,480), under
 * opts->single_rule; the previous pairs keep the rule of the build that added
 * them (at most 8 alternations of the rule along a chain of builds).  prev
 * must have the same storage kind.  With prev == NULL, n_stored must be 0. */
int hci_build_h(hci_ctx* ctx, const uint64_t* dets, const double* diag, int64_t n,
                const hci_csr* prev, int64_t n_stored, const hci_build_opts* opts,
                hci_csr** out, hci_build_stats* stats);

int hci_csr_info(const hci_csr* h, int64_t* n, int64_t* nnz_u, int* value_prec);
/* one row of the upper triangle in the caller's order: the columns j > row
 * (ascending) and values of H[row, j]; *count receives the row length, at most
 * cap entries are written.  HCI_STORE_FULL only. */
int hci_csr_row(const hci_csr* h, int64_t row, int32_t* cols, double* vals, int64_t cap,
                int64_t* count);
/* copies to host arrays: row_ptr[n+1], col[nnz_u], val[nnz_u] (stored values
 * widened to double: hi+lo for HCI_PREC_F32), diag[n]; any may be NULL */
int hci_csr_download(const hci_csr* h, int64_t* row_ptr, int32_t* col, double* val,
                     double* diag);
/* MatrixOperations::set_diagonal (src/var/eigenvalues/matrix_operations.rs:14-27):
 * replaces the stored diagonal with diag[n] (caller order).  The diagonal is
 * read back with hci_csr_download, nrows = ncols = n with hci_csr_info. */
int hci_csr_set_diagonal(hci_csr* h, const double* diag, int mem);
void hci_csr_free(hci_csr* h);
/* Frees h like hci_csr_free but keeps an HCI_STORE_TILED matrix's entry arrays
 * in its context for the next build there (device memory freed and allocated
 * again is cleared by the GPU kernel driver first: seconds at the 100-GB scale).  The next
 * TILED build takes what fits and frees the rest; any other build, or
 * hci_destroy, frees them.  No counterpart in the reference, whose sparse
 * matrix is a Rust Vec freed on drop (src/var/sparse.rs:24-36). */
int hci_csr_recycle(hci_csr* h);

/* y[:,c] = H x[:,c] for ncol column-major vectors of length n
 * (matrix_vector_prod / matrix_matrix_prod, src/var/sparse.rs:64-88). */
int hci_sigma(const hci_csr* h, int ncol, const double* x, double* y, int sigma_mode, int mem);

/* hci_sigma without any allocation: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  x and y are device
 * pointers (ncol column-major vectors of length n).  The two internal-order
 * vectors live in a fixed device array of the library (32 MiB per GPU, held
 * until the process ends), so n may be at most 2097152 (2^21); larger
 * matrices are rejected with HCI_ERR_ARG (use hci_sigma).  Every call on one
 * GPU shares that array: calls that overlap on one GPU (several contexts or
 * threads) race.  Returns after the product is complete. */
int hci_sigma_noalloc(const hci_csr* h, int ncol, const double* x, double* y, int sigma_mode);

/* Davidson, lowest root (Davidson::new via sparse_optimize).  init (nullable)
 * is a length-n start vector (the warm start of src/var/davidson.rs:71-83);
 * coeffs receives the eigenvector.  Returns HCI_OK, HCI_NOT_CONVERGED, or an
 * error.  On non-convergence e0/coeffs are left unchanged (the reference keeps
 * the previous coefficients, src/var/davidson.rs:109-111). */
void hci_dav_default_opts(hci_dav_opts* opts);
int hci_davidson(hci_ctx* ctx, const hci_csr* h, const double* init, const hci_dav_opts* opts,
                 double* e0, double* coeffs, hci_dav_stats* stats);

/* Learned preconditioner (no reference counterpart: it replaces the diagonal
 * correction of davidson.rs:391-408 when opts.precond == HCI_PRECOND_MLP).
 * Network 24 -> 128 -> 128 -> 1 (SiLU), output o gives the bounded factor
 * g = 1 + tanh(o) / 2 in (1/2, 3/2) that scales the diagonal correction:
 * t_p = g_p r_p / (lambda - H_pp).  Parameters are FP32, row-major
 * W1[24][128], b1[128], W2[128][128], b2[128], W3[128], b3 (HCI_MLP_NPARAM
 * floats); features are standardised with feat_mean / feat_std.  Products run
 * on BF16 matrix cores with FP32 accumulation.  Training rows: X[n][24] raw
 * features, m[n] the scaled diagonal correction, y[n] the scaled exact
 * correction; loss = mean((g m - y)^2). */
#define HCI_MLP_NFEAT 24
#define HCI_MLP_HIDDEN 128
#define HCI_MLP_NPARAM 19841

typedef struct hci_mlp_train_opts {
    int epochs;
    int batch;
    float lr;          /* Adam learning rate */
    uint64_t seed;     /* initialisation / shuffling */
    int standardize;   /* 1: set feat_mean/feat_std from the training rows */
    int mem;           /* where X/m/y live */
} hci_mlp_train_opts;

typedef struct hci_mlp_train_stats {
    double first_epoch_loss;
    double last_epoch_loss;
    double seconds;
} hci_mlp_train_stats;

int hci_mlp_init(hci_ctx* ctx, uint64_t seed);
int hci_mlp_set_params(hci_ctx* ctx, const float* params, const float* feat_mean, const float* feat_std);
int hci_mlp_get_params(hci_ctx* ctx, float* params, float* feat_mean, float* feat_std);
int hci_mlp_forward(hci_ctx* ctx, const float* X, int64_t n, float* out, int mem);
int hci_mlp_loss_grad(hci_ctx* ctx, const float* X, const float* m, const float* y, int64_t n, float* grad,
                      double* loss, int mem);
void hci_mlp_default_train_opts(hci_mlp_train_opts* opts);
int hci_mlp_train(hci_ctx* ctx, const float* X, const float* m, const float* y, int64_t n,
                  const hci_mlp_train_opts* opts, hci_mlp_train_stats* stats);
/* Runs Davidson (diagonal correction) on h and records training rows at every
 * iteration: features of up to per_iter determinants (an evenly strided
 * sample), m = scaled diagonal correction, y = scaled exact correction, the
 * part of x_exact (caller order, the converged eigenvector) outside the
 * current search space.  Writes at most cap rows; *count = rows written. */
int hci_mlp_collect(hci_ctx* ctx, const hci_csr* h, const double* x_exact, const hci_dav_opts* opts,
                    int64_t per_iter, float* X, float* m, float* y, int64_t cap, int64_t* count);

/* Heat-bath selection (replaces VarWf::find_new_dets / iterate_excites,
 * src/wf/mod.rs:1418-1631, with the ExciteGenerator of src/excite/init.rs:40-249
 * and init_eps of src/wf/eps.rs:44-171).
 * hci_excite_init builds the sorted excitation lists for the current integrals
 * (core = the n_core lowest orbitals by h_pp) and returns in *eps_start the
 * first epsilon of the schedule for the Hartree-Fock start (z_sym = 1).
 * hci_select writes the determinants (4 words each) reached from dets with
 * |H| >= eps / |coeff| that are not in dets, without duplicates, in the
 * reference's discovery order; singles (exact element > eps / |coeff|) only if
 * singles != 0.  *count = the number found; if it exceeds cap the call fails
 * with HCI_ERR_ARG and nothing is written (call again with cap >= *count).
 * The caller adds each one's spin-flipped partner (src/wf/mod.rs:99-116).
 * With diag (the space's diagonal) and out_diag non-null, out_diag receives
 * each new determinant's diagonal updated from its first discoverer's in
 * O(N) (Det::new_diag_opp/same/sing, src/wf/det.rs:271-420; the reference
 * stores these in Det.diag, src/wf/mod.rs:1525,1569,1618); both may be NULL.
 * hci_excite_info returns the largest magnitudes of the three lists. */
int hci_excite_init(hci_ctx* ctx, int n_core, int n_up, int n_dn, double eps_var, double* eps_start);
int hci_excite_info(const hci_ctx* ctx, double* max_opp, double* max_same, double* max_sing);
int hci_select(hci_ctx* ctx, const uint64_t* dets, const double* coeffs, const double* diag, int64_t n,
               double eps, int singles, uint64_t* out, double* out_diag, int64_t cap, int64_t* count, int mem);

/* Deterministic second-order (Epstein-Nesbet) energy of the space (dets,
 * coeffs) with variational energy e0 (dtm_pt, src/pt/mod.rs:81-122): every
 * determinant a outside the space reached from some i by an excitation with
 * |stored magnitude| * |c_i| >= eps (the lists of hci_excite_init; singles also
 * need |H_ai c_i| >= eps, src/pt/mod.rs:97-102) contributes
 * (sum_i H_ai c_i)^2 / (e0 - H_aa).  *e2 receives the sum and *n_ext (nullable)
 * the number of such determinants.  H_aa is computed from the integrals (the
 * reference updates the parent's diagonal, src/wf/det.rs:249-260; the two agree
 * to rounding).  When the candidate excitations outnumber `batch` (0: 2^28,
 * about 27 GB of scratch), the external determinants are split into
 * ceil(candidates / batch) parts by a hash of their strings (at most 2^20), and
 * each part walks every source determinant again, keeping its own.  Needs
 * hci_excite_init on the same integrals. */
int hci_pt2_dtm(hci_ctx* ctx, const uint64_t* dets, const double* coeffs, int64_t n, double e0, double eps,
                int64_t batch, double* e2, int64_t* n_ext, int mem);

/* The screened product H|psi> on the external space: replaces
 * Wf::approx_matmul_external_dtm_only (src/wf/mod.rs:186-339), the
 * deterministic part of the semistochastic PT2 (src/semistoch/mod.rs:630).
 * For each determinant i (in order) with local = eps / |c_i|, the excitation
 * lists of hci_excite_init are walked as the reference walks them (a list is
 * entered when its maximum is >= local and cut at the first entry below it;
 * singles are then kept when |H_ai| >= local) and every target a outside
 * `dets` receives H_ai c_i.  Outputs, in first-discovery order
 * (add_det_with_coeff, src/wf/mod.rs:144-165): the external determinants
 * (out, count x 4 words), their coefficients sum_i H_ai c_i summed in
 * discovery order (out_coeffs), and (out_diag, nullable; needs diag) the
 * diagonal its first discoverer gives it (Det::new_diag, src/wf/det.rs:249-420).
 * sum_remaining (nullable, n values): c_i^2 times the sum of the
 * sum_remaining_h_squared of every list entry where determinant i's walk was
 * cut (the sampler weights the reference returns).  *count receives the number
 * of external determinants; when it exceeds cap nothing is written but
 * sum_remaining and the call fails with *count set.  Every candidate
 * excitation is held on the device at once (about 52 B each plus sort
 * scratch; no splitting into parts as in hci_pt2_dtm). */
int hci_matmul_external(hci_ctx* ctx, const uint64_t* dets, const double* coeffs, const double* diag, int64_t n,
                        double eps, uint64_t* out, double* out_coeffs, double* out_diag, int64_t cap, int64_t* count,
                        double* sum_remaining, int mem);

/* The deterministic stage of the reference's default PT2,
 * new_semistoch_enpt2_dtm_diag_singles (src/semistoch/mod.rs:262-318; chosen by
 * perturbative when use_new_semistoch is set, src/pt/mod.rs:52-56, the default
 * of src/config.rs:253).  For each determinant i and each of its valence
 * electron pairs and electrons (valence_elecs_and_epairs order), the list of
 * hci_excite_init is read skipping targets already occupied or inside the
 * space (excites_from_det_and_orbs, src/excite/iterator.rs:113-130); the first
 * remaining entry with |stored magnitude| * |c_i| < eps ends it and yields a
 * sample record (DetOrbSample, src/stoch/mod.rs:104-124, the input of
 * generate_screened_sampler, :155).  Every entry before it gives
 * h = H_ai c_i: rel_probs[i] = max |h| (the off-diagonal sampler's weights),
 * and h joins h_psi unless it is a single with |h| < eps.  *e2 receives
 * pt(h_psi, e0) = sum_a h_psi_a^2 / (e0 - H_aa) with H_aa the diagonal the
 * first discoverer of a gives it (Det::new_diag, src/wf/det.rs:249-420; needs
 * the space's diagonal `diag`), *n_ext the size of h_psi.  rel_probs (n
 * values) may be NULL.  samples may be NULL; otherwise at most cap records are
 * written in the reference's push order, and a larger need fails the call
 * with *n_samples set.  Needs hci_excite_init on the same integrals. */
typedef struct hci_det_orb_sample {
    int64_t det;           /* index of the source determinant i */
    int32_t kind;          /* 0 opposite-spin pair (p up, q down), 1 same-spin pair p < q, 2 single p */
    int32_t is_alpha;      /* 1 alpha, 0 beta, -1 none (opposite-spin pair) */
    int32_t p, q;          /* initial orbitals (q = -1 for a single) */
    double sum_abs_h;      /* sum_remaining_abs_h of the entry that ended the list */
    double sum_h_squared;  /* its sum_remaining_h_squared */
    double sum_abs_hc;     /* |c_i| * sum_abs_h */
    double sum_hc_squared; /* c_i^2 * sum_h_squared */
} hci_det_orb_sample;

int hci_semistoch_dtm(hci_ctx* ctx, const uint64_t* dets, const double* coeffs, const double* diag, int64_t n,
                      double e0, double eps, double* e2, int64_t* n_ext, double* rel_probs,
                      hci_det_orb_sample* samples, int64_t cap, int64_t* n_samples, int mem);

/* The external products that also set up a screened sampler: replaces the
 * other Wf::approx_matmul_external_* variants (src/wf/mod.rs), selected by
 * `variant` (their callers: src/wf/operations.rs:85-111,
 * src/semistoch/mod.rs:69-73,555).  For each determinant i (in order) with
 * local = eps / |c_i|, the lists of hci_excite_init are walked in the order
 * of valence_elecs_and_epairs (opposite-spin pairs, same-spin pairs up then
 * down, electrons up then down) and every candidate target outside `dets`
 * receives H_ai c_i; a list that ends below local yields a DetOrbSample record
 * (hci_det_orb_sample, src/stoch/mod.rs:104-124).
 *   HCI_EXT_SEPARATE (separate_doubles_and_singles, :341-533): a kind's lists
 *     are walked only when its maximum is >= local; each is cut at its first
 *     entry below local (record), valid or not; valid entries before the cut
 *     are candidates, singles included without a test of their element.
 *   HCI_EXT_SKIP_SINGLES (skip_singles, :535-711): doubles only; occupied
 *     targets skipped first; entries >= local are candidates; the first entry
 *     below local whose target is outside the space ends the list (record);
 *     entries below local inside the space are passed over.
 *   HCI_EXT_SEMISTOCH_SINGLES (semistoch_singles, :714-1033): doubles as
 *     SKIP_SINGLES; singles lists walked when max_sing >= local: a valid
 *     external entry >= local is a candidate when |H_ai| >= local, otherwise
 *     the list's first record; below local the list ends if it has a record,
 *     else at its first valid external entry with a record whose four sums
 *     are scaled by 1e-6 (single_sample_factor).  The oversampling statistics
 *     the reference prints are not computed.
 *   HCI_EXT_NO_SINGLES (no_singles, :1035-1193): doubles as SEPARATE but
 *     without the maxima test; every valence electron gives a record from the
 *     first entry of its singles list (fails with HCI_ERR_ARG when those
 *     lists are empty, where the reference's index panics).
 * Outputs as hci_matmul_external (first-discovery order, first discoverer's
 * new_diag in out_diag); records in the reference's push order, at most scap
 * (samples may be NULL: *n_samples still receives the count).  A capacity
 * too small for either fails the call with *count / *n_samples set and
 * nothing else written. */
#define HCI_EXT_SEPARATE 1
#define HCI_EXT_SKIP_SINGLES 2
#define HCI_EXT_SEMISTOCH_SINGLES 3
#define HCI_EXT_NO_SINGLES 4
int hci_matmul_external_sampled(hci_ctx* ctx, int variant, const uint64_t* dets, const double* coeffs,
                                const double* diag, int64_t n, double eps, uint64_t* out, double* out_coeffs,
                                double* out_diag, int64_t cap, int64_t* count, hci_det_orb_sample* samples,
                                int64_t scap, int64_t* n_samples, int mem);

/* Device-memory and timing plumbing (no reference counterpart: for callers
 * that keep inputs resident in HBM, and for timing on the context's stream).
 * Events are 16 slots per context recorded on the context stream. */
int hci_dev_alloc(hci_ctx* ctx, size_t bytes, void** ptr);
int hci_dev_free(hci_ctx* ctx, void* ptr);
int hci_memcpy(hci_ctx* ctx, void* dst, const void* src, size_t bytes, int dst_mem, int src_mem);
int hci_event_record(hci_ctx* ctx, int slot);
int hci_event_elapsed_ms(hci_ctx* ctx, int slot_a, int slot_b, float* ms);

#ifdef __cplusplus
}
#endif
#endif
