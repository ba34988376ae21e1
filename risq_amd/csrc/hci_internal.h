// Internal host-side structures of the HCI library.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>
#include <vector>

#include "../../include/hci_gpu.h"
#include "hci_device.h"

struct MlpState;  // learned preconditioner (mlp.hip)

struct hci_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    // integrals
    int n_orb = 0;
    double e_core = 0.0;
    double* d_two = nullptr;
    double* d_h1 = nullptr;
    double* d_J3 = nullptr;
    double* d_D3 = nullptr;
    size_t n_two = 0;
    MlpState* mlp = nullptr;
    hipEvent_t events[16] = {};

    DevInts ints() const {
        DevInts t;
        t.n = n_orb;
        t.nuc = e_core;
        t.two = d_two;
        t.h1 = d_h1;
        t.J3 = d_J3;
        t.D3 = d_D3;
        return t;
    }
};

struct hci_csr {
    hci_ctx* ctx = nullptr;
    int64_t n = 0;
    int64_t nnz = 0;
    int prec = HCI_PREC_F64;
    int64_t* d_row_ptr = nullptr;  // n+1
    int32_t* d_col = nullptr;      // nnz
    double* d_val = nullptr;       // nnz (HCI_PREC_F64)
    float* d_hi = nullptr;         // nnz (HCI_PREC_F32)
    float* d_lo = nullptr;         // nnz (HCI_PREC_F32)
    double* d_diag = nullptr;      // n
    // nnz-balanced SpMV work split: first row of every chunk of kSpmvChunk nonzeros
    int32_t* d_chunk_row = nullptr;
    int64_t n_chunks = 0;
    // per-row absolute sums (feature for the learned preconditioner), lazily built
    double* d_rownorm = nullptr;
};

// status helpers ------------------------------------------------------------
int hci_fail(hci_ctx* ctx, int code, const std::string& msg);

#define HIP_TRY(ctx, expr)                                                              \
    do {                                                                                \
        hipError_t _e = (expr);                                                         \
        if (_e != hipSuccess)                                                           \
            return hci_fail((ctx), _e == hipErrorOutOfMemory ? HCI_ERR_OOM : HCI_ERR_HIP, \
                            std::string(#expr) + ": " + hipGetErrorString(_e));         \
    } while (0)

// SpMV (spmv.hip) ---------------------------------------------------------------
constexpr int kSpmvChunk = 1024;  // nonzeros per wave in the nnz-balanced SpMV
int spmv_prepare(hci_csr* h);      // builds d_chunk_row
// y = H x on device vectors, on the context stream; mode = HCI_SIGMA_*
int spmv_device(const hci_csr* h, const double* x, double* y, int mode);
int csr_rownorm(hci_csr* h);       // fills d_rownorm[i] = sum_j |H_ij| (both triangles)

// launch helpers
inline unsigned grid_for(int64_t n, int block) {
    int64_t g = (n + block - 1) / block;
    return (unsigned)(g < 1 ? 1 : g);
}
