// Symmetric sigma product y = H x from the upper-triangle CSR.
//
// Replaces SparseMatUpperTri::matrix_vector_prod (reference src/var/sparse.rs:64-78):
//   y_i = d_i x_i + sum_{j>i} H_ij x_j + sum_{j<i} H_ji x_j.
// Work split: one wavefront per chunk of kSpmvChunk stored nonzeros (balanced
// regardless of the skewed row lengths of an HCI Hamiltonian).  For every
// nonzero the transpose term H_ij x_i is added to y_j with a hardware FP64
// atomic; the row terms H_ij x_j are summed across the lanes that share a row
// (segmented shuffle reduction) and added to y_i with one atomic per row
// segment.  y starts as d o x.  HBM traffic is one pass over (col, value) of
// the upper triangle plus the vectors; reordered FP64 sums make the result
// agree with the reference loop to rounding, not bitwise.
#include "hci_internal.h"

namespace {

__global__ void k_chunk_rows(const int64_t* rp, int64_t n, int64_t nnz, int64_t n_chunks,
                             int32_t* chunk_row) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c > n_chunks) return;
    int64_t e = c * (int64_t)kSpmvChunk;
    if (e > nnz - 1) e = nnz - 1;
    // largest r in [0, n-1] with rp[r] <= e
    int64_t lo = 0, hi = n - 1;
    while (lo < hi) {
        int64_t mid = (lo + hi + 1) >> 1;
        if (rp[mid] <= e)
            lo = mid;
        else
            hi = mid - 1;
    }
    chunk_row[c] = (int32_t)lo;
}

__global__ void k_diag_mul(const double* d, const double* x, double* y, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = d[i] * x[i];
}

// MODE 0: FP64 values; 1: FP32 hi word; 2: FP32 hi + lo
template <int MODE>
__device__ __forceinline__ double load_val(const double* val, const float* hi, const float* lo,
                                           int64_t e) {
    if (MODE == 0) return val[e];
    if (MODE == 1) return (double)hi[e];
    return (double)hi[e] + (double)lo[e];
}

template <int MODE, bool ABS>
__global__ void __launch_bounds__(256)
    k_spmv_upper(int64_t nnz, int64_t n_chunks, const int64_t* __restrict__ rp,
                 const int32_t* __restrict__ chunk_row, const int32_t* __restrict__ col,
                 const double* __restrict__ val, const float* __restrict__ hi,
                 const float* __restrict__ lo, const double* __restrict__ x, double* y) {
    const int lane = threadIdx.x & 63;
    const int64_t c = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (c >= n_chunks) return;
    const int64_t e0 = c * (int64_t)kSpmvChunk;
    const int64_t e1 = min(e0 + (int64_t)kSpmvChunk, nnz);
    int32_t r_lo = chunk_row[c];
    const int32_t r_hi = chunk_row[c + 1];
    for (int64_t eb = e0; eb < e1; eb += 64) {
        const int64_t e = eb + lane;
        const bool valid = e < e1;
        int32_t r = -1;
        double s = 0.0;
        if (valid) {
            int32_t a = r_lo, b = r_hi;
            while (a < b) {
                const int32_t mid = (a + b + 1) >> 1;
                if (rp[mid] <= e)
                    a = mid;
                else
                    b = mid - 1;
            }
            r = a;
            double v = load_val<MODE>(val, hi, lo, e);
            if (ABS) v = fabs(v);
            const int32_t j = col[e];
            const double xr = ABS ? 1.0 : x[r];
            const double xj = ABS ? 1.0 : x[j];
            unsafeAtomicAdd(&y[j], v * xr);
            s = v * xj;
        }
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const double t = __shfl_down(s, o);
            const int32_t rr = __shfl_down(r, o);
            if (lane + o < 64 && rr == r) s += t;
        }
        const int32_t rprev = __shfl_up(r, 1);
        if (valid && (lane == 0 || rprev != r)) unsafeAtomicAdd(&y[r], s);
        const int32_t rlast = __shfl(r, 63);
        if (rlast >= 0) r_lo = rlast;
    }
}

}  // namespace

int spmv_prepare(hci_csr* h) {
    hci_ctx* c = h->ctx;
    hipFree(h->d_chunk_row);
    h->d_chunk_row = nullptr;
    h->n_chunks = (h->nnz + kSpmvChunk - 1) / kSpmvChunk;
    HIP_TRY(c, hipMalloc(&h->d_chunk_row, (size_t)(h->n_chunks + 1) * 4 + 16));
    if (h->nnz > 0)
        hipLaunchKernelGGL(k_chunk_rows, grid_for(h->n_chunks + 1, 256), 256, 0, c->stream, h->d_row_ptr,
                           h->n, h->nnz, h->n_chunks, h->d_chunk_row);
    HIP_TRY(c, hipGetLastError());
    return HCI_OK;
}

int spmv_device(const hci_csr* h, const double* x, double* y, int mode) {
    hci_ctx* c = h->ctx;
    hipStream_t s = c->stream;
    hipLaunchKernelGGL(k_diag_mul, grid_for(h->n, 256), 256, 0, s, h->d_diag, x, y, h->n);
    if (h->nnz > 0) {
        const unsigned grid = grid_for(h->n_chunks, 4);
        if (h->prec == HCI_PREC_F64)
            hipLaunchKernelGGL((k_spmv_upper<0, false>), grid, 256, 0, s, h->nnz, h->n_chunks, h->d_row_ptr,
                               h->d_chunk_row, h->d_col, h->d_val, nullptr, nullptr, x, y);
        else if (mode == HCI_SIGMA_F32)
            hipLaunchKernelGGL((k_spmv_upper<1, false>), grid, 256, 0, s, h->nnz, h->n_chunks, h->d_row_ptr,
                               h->d_chunk_row, h->d_col, nullptr, h->d_hi, nullptr, x, y);
        else
            hipLaunchKernelGGL((k_spmv_upper<2, false>), grid, 256, 0, s, h->nnz, h->n_chunks, h->d_row_ptr,
                               h->d_chunk_row, h->d_col, nullptr, h->d_hi, h->d_lo, x, y);
    }
    HIP_TRY(c, hipGetLastError());
    return HCI_OK;
}

int csr_rownorm(hci_csr* h) {
    if (h->d_rownorm) return HCI_OK;
    hci_ctx* c = h->ctx;
    hipStream_t s = c->stream;
    HIP_TRY(c, hipMalloc(&h->d_rownorm, (size_t)h->n * 8 + 16));
    HIP_TRY(c, hipMemsetAsync(h->d_rownorm, 0, (size_t)h->n * 8, s));
    if (h->nnz > 0) {
        const unsigned grid = grid_for(h->n_chunks, 4);
        if (h->prec == HCI_PREC_F64)
            hipLaunchKernelGGL((k_spmv_upper<0, true>), grid, 256, 0, s, h->nnz, h->n_chunks, h->d_row_ptr,
                               h->d_chunk_row, h->d_col, h->d_val, nullptr, nullptr, nullptr, h->d_rownorm);
        else
            hipLaunchKernelGGL((k_spmv_upper<2, true>), grid, 256, 0, s, h->nnz, h->n_chunks, h->d_row_ptr,
                               h->d_chunk_row, h->d_col, nullptr, h->d_hi, h->d_lo, nullptr, h->d_rownorm);
    }
    HIP_TRY(c, hipGetLastError());
    return HCI_OK;
}

extern "C" int hci_sigma(const hci_csr* h, int ncol, const double* x, double* y, int sigma_mode,
                         int mem) {
    if (!h) return HCI_ERR_ARG;
    hci_ctx* c = h->ctx;
    if (ncol < 0 || (ncol > 0 && (!x || !y))) return hci_fail(c, HCI_ERR_ARG, "hci_sigma: bad arguments");
    if (sigma_mode != HCI_SIGMA_FULL && sigma_mode != HCI_SIGMA_F32)
        return hci_fail(c, HCI_ERR_ARG, "hci_sigma: bad sigma_mode");
    HIP_TRY(c, hipSetDevice(c->device));
    const size_t bytes = (size_t)h->n * 8;
    if (mem == HCI_MEM_DEVICE) {
        for (int k = 0; k < ncol; ++k) {
            int rc = spmv_device(h, x + (size_t)k * h->n, y + (size_t)k * h->n, sigma_mode);
            if (rc) return rc;
        }
        return HCI_OK;
    }
    double *dx = nullptr, *dy = nullptr;
    HIP_TRY(c, hipMalloc(&dx, bytes + 16));
    HIP_TRY(c, hipMalloc(&dy, bytes + 16));
    int rc = HCI_OK;
    for (int k = 0; k < ncol && rc == HCI_OK; ++k) {
        if (hipMemcpyAsync(dx, x + (size_t)k * h->n, bytes, hipMemcpyHostToDevice, c->stream) != hipSuccess) {
            rc = hci_fail(c, HCI_ERR_HIP, "hci_sigma: copy in");
            break;
        }
        rc = spmv_device(h, dx, dy, sigma_mode);
        if (rc == HCI_OK &&
            hipMemcpyAsync(y + (size_t)k * h->n, dy, bytes, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
            rc = hci_fail(c, HCI_ERR_HIP, "hci_sigma: copy out");
    }
    hipStreamSynchronize(c->stream);
    hipFree(dx);
    hipFree(dy);
    return rc;
}
