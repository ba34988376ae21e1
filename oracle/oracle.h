/*
 * CPU oracle for the HCI variational hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain C++ restatement of the reference algorithm (aaholmes/risq, Rust), used
 * by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg as the
 * result to compare against.  Nothing in the product library (risq_amd/) links
 * or calls it.
 *
 * Determinant words: every determinant is four uint64 words
 *   [up_lo, up_hi, dn_lo, dn_hi]   (bit p = spatial orbital p, 0-based)
 * which is the reference's Config{up: u128, dn: u128} (src/wf/det.rs:16-19).
 *
 * Integral tables use the reference layout: one_body[combine_2(p,q)],
 * two_body[combine_4(p,q,r,s)] with 1-based indices (src/utils/ints.rs:38-68,
 * src/ham/read_ints.rs:78-80).
 */
#ifndef HCI_ORACLE_H
#define HCI_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_ham or_ham;
typedef struct or_sparse or_sparse;

/* bit / sign helpers (src/utils/bits.rs:171-180, src/utils/ints.rs:38-97) */
int or_parity(uint64_t lo, uint64_t hi);
int or_permute(const uint64_t* d1 /*2 words*/, const uint64_t* d2 /*2 words*/);
int or_permute_2(const uint64_t* d1, const uint64_t* d2, const int32_t* v /*4*/);
uint64_t or_combine_2(int32_t p, int32_t q);
uint64_t or_combine_4(int32_t p, int32_t q, int32_t r, int32_t s);

/* integrals (src/ham/read_ints.rs:33-42) */
or_ham* or_ham_create(int32_t n_orb, double e_core, const double* one_body, size_t n1,
                      const double* two_body, size_t n2);
void or_ham_free(or_ham*);

/* Slater-Condon rules (src/ham/mod.rs:92-303) */
void or_ham_diag(const or_ham*, const uint64_t* dets, int64_t n, double* out);
/* value[m] = ham_off_diag_no_excite(det1[m], det2[m]); degree[m] = excitation
 * degree (0,1,2 or 3 meaning "more than double") */
void or_off_diag(const or_ham*, const uint64_t* det1, const uint64_t* det2, int64_t m,
                 double* value, int8_t* degree);

/* Sparse H (src/var/sparse.rs, src/var/fast_ham.rs, src/wf/mod.rs:1635-1683).
 * or_sparse_create allocates n empty rows with the given diagonal.
 * or_build_fast_lookup follows gen_sparse_ham_fast_lookup over dets[0..n): rows
 * are grown to n (expand_sparse_ham_rows), pairs with at least one index >=
 * n_stored are added, then sort_remove_duplicates runs. */
or_sparse* or_sparse_create(int64_t n, const double* diag);
int or_build_fast_lookup(or_sparse*, const or_ham*, const uint64_t* dets, int64_t n,
                         const double* diag, int64_t n_stored);
/* or_build_small follows gen_sparse_ham_fast with opp_algo = same_algo = 1
 * (src/var/ham_gen.rs:60-292,374-433,621-637), the reference's build while the
 * space holds <= 100 determinants: singles are stored iff H != 0, and the space
 * must be closed under spin flip (returns -2 otherwise; the reference panics). */
int or_build_small(or_sparse*, const or_ham*, const uint64_t* dets, int64_t n, const double* diag,
                   int64_t n_stored);
int64_t or_sparse_n(const or_sparse*);
int64_t or_sparse_nnz(const or_sparse*);
/* row_ptr[n+1], col[nnz], val[nnz], diag[n] */
void or_sparse_export(const or_sparse*, int64_t* row_ptr, int32_t* col, double* val,
                      double* diag);
void or_sparse_free(or_sparse*);

/* symmetric upper-triangle SpMV in the reference's loop order (src/var/sparse.rs:64-78) */
void or_spmv(const or_sparse*, const double* x, double* y);
/* the same on plain CSR arrays (used for CPU timing of a GPU-built H) */
void or_spmv_csr(int64_t n, const int64_t* row_ptr, const int32_t* col, const double* val,
                 const double* diag, const double* x, double* y);

/* Davidson, lowest root, DPR correction, reference control flow
 * (src/var/eigenvalues/algorithms/davidson.rs:106-277).
 * init may be NULL.  Returns 0 on convergence, 1 on non-convergence. */
int or_davidson(const or_sparse*, const double* init, double tol, double etol, int max_iters,
                double* e0, double* coeffs, int* iters, int* n_sigma);
int or_davidson_csr(int64_t n, const int64_t* row_ptr, const int32_t* col, const double* val,
                    const double* diag, const double* init, double tol, double etol,
                    int max_iters, double* e0, double* coeffs, int* iters, int* n_sigma);

/* symmetric eigen-decomposition of a small dense matrix (row-major k*k); values
 * ascending, vectors as columns of vecs (row-major k*k). */
void or_sym_eigen(int k, const double* a, double* vals, double* vecs);

#ifdef __cplusplus
}
#endif
#endif
