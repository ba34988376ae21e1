// Davidson eigensolver (lowest root) on the GPU.
//
// Control flow follows Davidson::new with nvalues = 1 and DPR correction
// (reference src/var/eigenvalues/algorithms/davidson.rs:106-277):
//   start vectors e0, e1 (or the warm-start vector alone), W = H V, P = V^T W;
//   each iteration: lowest Ritz pair of P, residual r = W y - lambda V y,
//   converged iff |r| < tol and |lambda - lambda_prev| < etol (or n <= 2);
//   while the subspace has room: t = r / (lambda - H_ii) (davidson.rs:391-408;
//   a component whose denominator is exactly zero is set to 0, where the
//   reference's 0/0 would put a NaN into the basis -- e.g. when the Ritz
//   vector is an exact unit vector of an unconnected determinant),
//   stop if |t| < 1e-9, orthonormalise t against V (stop if the remainder
//   is < 1e-9), append t to V and H t to W, extend P; when full, restart from
//   the Ritz vector (davidson.rs:258-269).
// All vector work stays in FP64 on the device: tall-skinny products V^T w and
// V y are single HBM passes with deterministic two-stage block reductions.
// The reference's modified Gram-Schmidt (modified_gram_schmidt.rs:22-42) is
// done as classical Gram-Schmidt applied twice (two HBM passes instead of k),
// which gives the same orthogonality to FP64 rounding.  The <= 11 x 11
// projected eigenproblem is solved on the host with cyclic Jacobi.
// Thick restart (restart_keep = m >= 2): the new basis is V Q, where the
// columns of Q (in the k-dimensional coefficient space) are the m - 1 lowest
// Ritz vectors and the previous iteration's Ritz coefficients orthonormalised
// against them.  W Q and Q^T P Q follow without any new product, because the
// previous Ritz vector lies in span(V) (the basis only grew since it was
// formed).
// Mixed precision: sigmas read FP32 values (FP64 vectors and accumulation)
// until the FP32 iteration converges (or the residual drops below refine_tol,
// or stalls), then the solver restarts from the Ritz vector with
// full-precision sigmas, so the converged eigenvalue is that of the
// full-precision H.
#include <chrono>
#include <cmath>

#include "hci_internal.h"
#include "mlp.h"

namespace {

constexpr int kMaxSub = HCI_DAV_MAX_SUB;
constexpr int kMaxKeep = HCI_DAV_MAX_KEEP;
constexpr int kRedBlocks = 1024;
constexpr int kRedThreads = 256;
constexpr int kRitzSlot = 2 + 2 * kMaxSub;               // k_ritz_dots' values at KM = kMaxSub
constexpr int kRedValues = 3 * (kMaxSub + 1) + kRitzSlot;  // the device reduction slots

struct Coef {
    double c[kMaxSub];
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
    return v;
}

// Block-reduce NV values per thread into partial[blockIdx.x * NV + k].
template <int NV>
__device__ __forceinline__ void block_store(double (&acc)[NV], int nv, double* partial) {
    __shared__ double sh[kRedThreads / 64][NV];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        if (k < nv) {
            const double s = wave_sum(acc[k]);
            if (lane == 0) sh[w][k] = s;
        }
    }
    __syncthreads();
    if (threadIdx.x < nv) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < kRedThreads / 64; ++q) s += sh[q][threadIdx.x];
        partial[(size_t)blockIdx.x * nv + threadIdx.x] = s;
    }
}

// The vector kernels below are instantiated for KM = 8, 12, 16, 24 and 32
// columns at most (the solve's subspace size rounded up): their per-thread arrays take
// registers in proportion to KM, and the occupancy that the registers leave
// decides how many rows' loads are in flight.  Rows are 32-bit offsets from
// per-column bases (n < 2^31), so a column's address stays in scalar
// registers (k_ritz_dots<16>: 149 VGPRs against 222 with a 64-bit address per
// column and thread).

// partial[b*k + c] = sum over this block's rows of A[c][i] * b[i], c < k <= KM + 1
template <int KM>
__global__ void __launch_bounds__(kRedThreads)
    k_multidot(const double* __restrict__ A, int64_t ld, int k, const double* __restrict__ b, int64_t n,
               double* partial) {
    double acc[KM + 1];
#pragma unroll
    for (int c = 0; c <= KM; ++c) acc[c] = 0.0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < (uint32_t)n; i += gridDim.x * blockDim.x) {
        const double bi = b[i];
#pragma unroll
        for (int c = 0; c <= KM; ++c)
            if (c < k) acc[c] += (A + (size_t)c * ld)[i] * bi;
    }
    block_store<KM + 1>(acc, k, partial);
}

// out[c] = sum_b partial[b*k + c], fixed order (deterministic); one block per c
__global__ void __launch_bounds__(kRedThreads) k_reduce(const double* partial, int nblocks, int k, double* out) {
    __shared__ double sh[kRedThreads];
    const int c = blockIdx.x;
    double s = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x) s += partial[(size_t)b * k + c];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = sh[0];
}

// x = V y, r = W y - lam x (W = H V), t = r / (lam - d); partial sums of r.r
// and t.t.  x and r are written only where given (most steps need t alone);
// k <= KM
template <int KM>
__global__ void __launch_bounds__(kRedThreads)
    k_ritz(const double* __restrict__ V, const double* __restrict__ W, int64_t ld, int k, Coef y,
           double lam, const double* __restrict__ d, int64_t n, double* x, double* r, double* t,
           double* partial) {
    double acc[2] = {0.0, 0.0};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < (uint32_t)n; i += gridDim.x * blockDim.x) {
        double xv = 0.0, wv = 0.0;
#pragma unroll
        for (int c = 0; c < KM; ++c)
            if (c < k) {
                xv += (V + (size_t)c * ld)[i] * y.c[c];
                wv += (W + (size_t)c * ld)[i] * y.c[c];
            }
        const double rv = wv - xv * lam;
        const double den = lam - d[i];
        const double tv = den != 0.0 ? rv / den : 0.0;
        if (x) x[i] = xv;
        if (r) r[i] = rv;
        t[i] = tv;
        acc[0] += rv * rv;
        acc[1] += tv * tv;
    }
    block_store<2>(acc, 2, partial);
}

// k_ritz that also sums V^T t and W^T t (W = H V) while V and W pass through
// its registers: partial slots r.r, t.t, V[c].t (2 + c), W[c].t (2 + KM + c).
// With the diagonal correction these give the Gram-Schmidt coefficients of t
// and, through W^T t' = W^T t - (V^T W) c, the new column of the projected
// matrix, so neither needs a pass over V of its own
template <int KM>
__global__ void __launch_bounds__(kRedThreads)
    k_ritz_dots(const double* __restrict__ V, const double* __restrict__ W, int64_t ld, int k, Coef y,
                double lam, const double* __restrict__ d, int64_t n, double* x, double* r, double* t,
                double* partial) {
    double acc[2 + 2 * KM];
#pragma unroll
    for (int c = 0; c < 2 + 2 * KM; ++c) acc[c] = 0.0;
    // rows as 32-bit offsets from per-column bases (n < 2^31): the column
    // bases stay scalar instead of one 64-bit address per column per thread
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < (uint32_t)n; i += gridDim.x * blockDim.x) {
        double vc[KM], wc[KM];
        double xv = 0.0, wv = 0.0;
#pragma unroll
        for (int c = 0; c < KM; ++c)
            if (c < k) {
                vc[c] = (V + (size_t)c * ld)[i];
                wc[c] = (W + (size_t)c * ld)[i];
                xv += vc[c] * y.c[c];
                wv += wc[c] * y.c[c];
            }
        const double rv = wv - xv * lam;
        const double den = lam - d[i];
        const double tv = den != 0.0 ? rv / den : 0.0;
        if (x) x[i] = xv;
        if (r) r[i] = rv;
        t[i] = tv;
        acc[0] += rv * rv;
        acc[1] += tv * tv;
#pragma unroll
        for (int c = 0; c < KM; ++c)
            if (c < k) {
                acc[2 + c] += vc[c] * tv;
                acc[2 + KM + c] += wc[c] * tv;
            }
    }
    block_store<2 + 2 * KM>(acc, 2 + 2 * KM, partial);
}

// t -= sum_c V[c] h[c]; then partial dots of V[c].t (c < k) and t.t (slot k); k <= KM.
// The coefficients come from hd (device memory, a previous reduction) when
// it is given, else from h.
template <int KM>
__global__ void __launch_bounds__(kRedThreads)
    k_orth_sub_dot(const double* __restrict__ V, int64_t ld, int k, Coef h, const double* __restrict__ hd,
                   double* t, int64_t n, double* partial) {
    if (hd) {
#pragma unroll
        for (int c = 0; c < KM; ++c)
            if (c < k) h.c[c] = hd[c];
    }
    double acc[KM + 1];
#pragma unroll
    for (int c = 0; c <= KM; ++c) acc[c] = 0.0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < (uint32_t)n; i += gridDim.x * blockDim.x) {
        double tv = t[i];
        double vc[KM];
#pragma unroll
        for (int c = 0; c < KM; ++c)
            if (c < k) {
                vc[c] = (V + (size_t)c * ld)[i];
                tv -= vc[c] * h.c[c];
            }
        t[i] = tv;
#pragma unroll
        for (int c = 0; c < KM; ++c)
            if (c < k) acc[c] += vc[c] * tv;
        acc[KM] += tv * tv;
    }
    // move t.t into slot k for the reduction layout (static indices only)
#pragma unroll
    for (int c = 0; c < KM; ++c)
        if (c == k) acc[c] = acc[KM];
    block_store<KM + 1>(acc, k + 1, partial);
}

// the column bound KM of the vector kernels for a subspace of km columns: the
// smallest of 8, 12, 16, 24 and 32 that is >= km
inline int km_bound(int km) { return km <= 8 ? 8 : km <= 12 ? 12 : km <= 16 ? 16 : km <= 24 ? 24 : 32; }
#define HCI_LAUNCH_KM(kern, km, grid, block, stream, ...)                                   \
    do {                                                                                    \
        switch (km_bound(km)) {                                                             \
            case 8: hipLaunchKernelGGL(kern<8>, grid, block, 0, stream, __VA_ARGS__); break;   \
            case 12: hipLaunchKernelGGL(kern<12>, grid, block, 0, stream, __VA_ARGS__); break; \
            case 16: hipLaunchKernelGGL(kern<16>, grid, block, 0, stream, __VA_ARGS__); break; \
            case 24: hipLaunchKernelGGL(kern<24>, grid, block, 0, stream, __VA_ARGS__); break; \
            default: hipLaunchKernelGGL(kern<32>, grid, block, 0, stream, __VA_ARGS__); break; \
        }                                                                                   \
    } while (0)

// In place: V[:, j] = sum_c V[:, c] q[c * kMaxKeep + j] for j < m (same for W).
// Each thread reads its row of all k columns before writing m of them.
__global__ void __launch_bounds__(kRedThreads)
    k_rotate(double* V, double* W, int64_t ld, int k, int m, const double* __restrict__ q, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double vo[kMaxKeep], wo[kMaxKeep];
#pragma unroll
        for (int j = 0; j < kMaxKeep; ++j) vo[j] = wo[j] = 0.0;
        for (int c = 0; c < k; ++c) {
            const double v = V[(size_t)c * ld + i], w = W[(size_t)c * ld + i];
#pragma unroll
            for (int j = 0; j < kMaxKeep; ++j) {
                const double qc = q[c * kMaxKeep + j];
                vo[j] += v * qc;
                wo[j] += w * qc;
            }
        }
#pragma unroll
        for (int j = 0; j < kMaxKeep; ++j)
            if (j < m) {
                V[(size_t)j * ld + i] = vo[j];
                W[(size_t)j * ld + i] = wo[j];
            }
    }
}

__global__ void k_scale(const double* t, double s, double* out, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = t[i] * s;
}

__global__ void k_unit(double* v, int64_t n, int64_t pos) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = (i == pos) ? 1.0 : 0.0;
}

// Cyclic Jacobi for the small symmetric projected matrix (row-major k x k).
// Eigenvalues ascending; vecs column c is the eigenvector of vals[c].
void jacobi_eigen(int k, const std::vector<double>& a_in, std::vector<double>& vals,
                  std::vector<double>& vecs) {
    std::vector<double> a(a_in), v((size_t)k * k, 0.0);
    for (int i = 0; i < k; ++i) v[(size_t)i * k + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, diag = 0.0;
        for (int p = 0; p < k; ++p) {
            diag += a[(size_t)p * k + p] * a[(size_t)p * k + p];
            for (int q = p + 1; q < k; ++q) off += a[(size_t)p * k + q] * a[(size_t)p * k + q];
        }
        if (off <= 1e-32 * diag || off < 1e-300) break;
        for (int p = 0; p < k; ++p)
            for (int q = p + 1; q < k; ++q) {
                const double apq = a[(size_t)p * k + q];
                if (apq == 0.0) continue;
                const double theta = (a[(size_t)q * k + q] - a[(size_t)p * k + p]) / (2.0 * apq);
                const double tt = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double cs = 1.0 / std::sqrt(tt * tt + 1.0), sn = tt * cs;
                for (int r = 0; r < k; ++r) {
                    const double arp = a[(size_t)r * k + p], arq = a[(size_t)r * k + q];
                    a[(size_t)r * k + p] = cs * arp - sn * arq;
                    a[(size_t)r * k + q] = sn * arp + cs * arq;
                }
                for (int r = 0; r < k; ++r) {
                    const double apr = a[(size_t)p * k + r], aqr = a[(size_t)q * k + r];
                    a[(size_t)p * k + r] = cs * apr - sn * aqr;
                    a[(size_t)q * k + r] = sn * apr + cs * aqr;
                }
                for (int r = 0; r < k; ++r) {
                    const double vrp = v[(size_t)r * k + p], vrq = v[(size_t)r * k + q];
                    v[(size_t)r * k + p] = cs * vrp - sn * vrq;
                    v[(size_t)r * k + q] = sn * vrp + cs * vrq;
                }
            }
    }
    std::vector<int> idx(k);
    for (int i = 0; i < k; ++i) idx[i] = i;
    std::sort(idx.begin(), idx.end(), [&](int x, int y) { return a[(size_t)x * k + x] < a[(size_t)y * k + y]; });
    vals.assign(k, 0.0);
    vecs.assign((size_t)k * k, 0.0);
    for (int c = 0; c < k; ++c) {
        vals[c] = a[(size_t)idx[c] * k + idx[c]];
        for (int r = 0; r < k; ++r) vecs[(size_t)r * k + c] = v[(size_t)r * k + idx[c]];
    }
}

struct DavWork {
    hci_ctx* c = nullptr;
    int64_t n = 0;
    int max_sub = 0;
    double* V = nullptr;
    double* W = nullptr;
    double* x = nullptr;
    double* q = nullptr;    // restart coefficients, kMaxSub x kMaxKeep
    double* h_q = nullptr;  // pinned staging for q
    double* r = nullptr;
    double* t = nullptr;
    double* partial = nullptr;
    // device reductions: 3 slots of kMaxSub + 1 values, then k_ritz_dots'
    // kRitzSlot values; h_red is their pinned host copy (same layout)
    double* red = nullptr;
    double* h_red = nullptr;
    // carves every buffer out of the context's work block (kept between calls)
    int carve(int nsub) {
        const size_t vb = ((size_t)n * 8 + 255) / 256 * 256;
        const size_t sizes[] = {vb * (size_t)nsub, vb * (size_t)nsub, vb, vb, vb,
                                (size_t)kMaxSub * kMaxKeep * 8, (size_t)kRedBlocks * kRitzSlot * 8,
                                (size_t)kRedValues * 8};
        size_t total = 0;
        for (size_t b : sizes) total += (b + 255) / 256 * 256;
        const size_t host = (size_t)kMaxSub * kMaxKeep * 8 + (size_t)kRedValues * 8 + 256;
        if (int rc = work_reserve(c, total, host)) return rc;
        double** dst[] = {&V, &W, &x, &r, &t, &q, &partial, &red};
        char* p = (char*)c->work_dev;
        for (int k = 0; k < 8; ++k) {
            *dst[k] = (double*)p;
            p += (sizes[k] + 255) / 256 * 256;
        }
        h_q = (double*)c->work_host;
        h_red = h_q + (size_t)kMaxSub * kMaxKeep;
        return HCI_OK;
    }
    unsigned red_grid() const {
        int64_t g = (n + kRedThreads - 1) / kRedThreads;
        return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, kRedBlocks));
    }
    double* col(double* M, int c) const { return M + (size_t)c * n; }
    // reductions: results land in h_red[0..k)
    // reduction into device memory only (no host round trip)
    void reduce_dev(int k, double* dst) {
        hipLaunchKernelGGL(k_reduce, k, kRedThreads, 0, c->stream, partial, (int)red_grid(), k, dst);
    }
    int reduce(int k) {
        hipLaunchKernelGGL(k_reduce, k, kRedThreads, 0, c->stream, partial, (int)red_grid(), k, red);
        HIP_TRY(c, hipMemcpyAsync(h_red, red, (size_t)k * 8, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        return HCI_OK;
    }
    // k_ritz_dots' values into their own slot, copied to the host
    int reduce_ritz(int k) {
        double* dst = red + 3 * (kMaxSub + 1);
        hipLaunchKernelGGL(k_reduce, k, kRedThreads, 0, c->stream, partial, (int)red_grid(), k, dst);
        HIP_TRY(c, hipMemcpyAsync(h_red + 3 * (kMaxSub + 1), dst, (size_t)k * 8, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        return HCI_OK;
    }
    int multidot(const double* A, int k, const double* b) {
        HCI_LAUNCH_KM(k_multidot, max_sub, red_grid(), kRedThreads, c->stream, A, n, k, b, n, partial);
        return reduce(k);
    }
};

// Training-row recorder for the learned preconditioner (hci_mlp_collect).
struct Recorder {
    const double* e_int = nullptr;  // converged eigenvector, internal order
    double* e_perp = nullptr;       // its part outside the search space
    float* feat = nullptr;          // n x HCI_MLP_NFEAT
    float *X = nullptr, *m = nullptr, *y = nullptr;  // staged rows (device)
    int64_t per_iter = 0, cap = 0, count = 0;
    uint64_t iter = 0;
};

__global__ void k_collect(const float* feat, const double* t, const double* e, double ms, double ys, int64_t n,
                          int64_t per, int64_t offset, int64_t stride, int64_t base, float* X, float* m, float* y) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= per) return;
    const int64_t p = (offset + j * stride) % n;
    for (int f = 0; f < HCI_MLP_NFEAT; ++f) X[(base + j) * HCI_MLP_NFEAT + f] = feat[p * HCI_MLP_NFEAT + f];
    m[base + j] = (float)(t[p] * ms);
    y[base + j] = (float)(e[p] * ys);
}

int record_rows(hci_ctx* c, const hci_csr* h, DavWork& w, Recorder& rec, double lam, double err, double tnorm,
                int dim_sub) {
    const int64_t N = w.n;
    hipStream_t s = c->stream;
    if (rec.count >= rec.cap) return HCI_OK;
    // exact correction: x_exact aligned with the Ritz vector, minus its part in span(V)
    int rc = w.multidot(w.x, 1, rec.e_int);
    if (rc) return rc;
    const double sgn = w.h_red[0] < 0 ? -1.0 : 1.0;
    hipLaunchKernelGGL(k_scale, grid_for(N, 256), 256, 0, s, rec.e_int, sgn, rec.e_perp, N);
    if ((rc = w.multidot(w.V, dim_sub, rec.e_perp))) return rc;
    Coef hc{};
    for (int k = 0; k < dim_sub; ++k) hc.c[k] = w.h_red[k];
    HCI_LAUNCH_KM(k_orth_sub_dot, w.max_sub, w.red_grid(), kRedThreads, s, w.V, N, dim_sub, hc, nullptr, rec.e_perp,
                  N, w.partial);
    if ((rc = w.reduce(dim_sub + 1))) return rc;
    const double enorm = std::sqrt(w.h_red[dim_sub]);
    if (!(enorm > 0.0) || !(tnorm > 0.0)) return HCI_OK;
    if ((rc = w.multidot(w.t, 1, rec.e_perp))) return rc;
    const double sgn2 = w.h_red[0] < 0 ? -1.0 : 1.0;
    if ((rc = mlp_features(c, h, lam, err, tnorm, w.r, w.x, w.t, rec.feat))) return rc;
    const int64_t per = std::min<int64_t>(std::min<int64_t>(rec.per_iter, N), rec.cap - rec.count);
    const int64_t stride = std::max<int64_t>(1, N / std::max<int64_t>(per, 1));
    const int64_t offset = (int64_t)((rec.iter * 0x9E3779B97F4A7C15ull) % (uint64_t)N);
    const double sn = std::sqrt((double)N);
    hipLaunchKernelGGL(k_collect, grid_for(per, 256), 256, 0, s, rec.feat, w.t, rec.e_perp, sn / tnorm,
                       sgn2 * sn / enorm, N, per, offset, stride, rec.count, rec.X, rec.m, rec.y);
    HIP_TRY(c, hipGetLastError());
    rec.count += per;
    ++rec.iter;
    return HCI_OK;
}

}  // namespace

int work_reserve(hci_ctx* c, size_t dev_bytes, size_t host_bytes) {
    if (c->work_dev_bytes < dev_bytes) {
        if (c->work_dev) {
            HIP_TRY(c, hipStreamSynchronize(c->stream));
            hipFree(c->work_dev);
        }
        c->work_dev = nullptr;
        c->work_dev_bytes = 0;
        HIP_TRY(c, hipMalloc(&c->work_dev, dev_bytes));
        c->work_dev_bytes = dev_bytes;
    }
    if (c->work_host_bytes < host_bytes) {
        if (c->work_host) {
            HIP_TRY(c, hipStreamSynchronize(c->stream));
            hipHostFree(c->work_host);
        }
        c->work_host = nullptr;
        c->work_host_bytes = 0;
        HIP_TRY(c, hipHostMalloc(&c->work_host, host_bytes));
        c->work_host_bytes = host_bytes;
    }
    return HCI_OK;
}

void work_release(hci_ctx* c) {
    if (c->work_dev || c->work_host) hipStreamSynchronize(c->stream);
    if (c->work_dev) hipFree(c->work_dev);
    if (c->work_host) hipHostFree(c->work_host);
    c->work_dev = c->work_host = nullptr;
    c->work_dev_bytes = c->work_host_bytes = 0;
}

extern "C" void hci_dav_default_opts(hci_dav_opts* o) {
    if (!o) return;
    o->tol = 1e-4;
    o->etol = 1e-8;
    o->max_iters = 50;
    o->max_sub = 11;
    o->precond = HCI_PRECOND_DPR;
    o->mixed = 0;
    o->refine_tol = 0.0;
    o->mem = HCI_MEM_HOST;
    o->restart_keep = 1;
    o->f16_tol = 0.0;
}

static int davidson_run(hci_ctx* c, const hci_csr* h, const double* init, const hci_dav_opts* opts,
                        double* e0, double* coeffs, hci_dav_stats* stats, Recorder* rec) {
    if (!c || !h || !opts || !e0 || !coeffs) return c ? hci_fail(c, HCI_ERR_ARG, "hci_davidson: null argument") : HCI_ERR_ARG;
    if (h->ctx != c) return hci_fail(c, HCI_ERR_ARG, "hci_davidson: matrix from another context");
    const int64_t N = h->n;
    if (N < 1) return hci_fail(c, HCI_ERR_ARG, "hci_davidson: empty matrix");
    if (opts->max_sub < 2 || opts->max_sub > kMaxSub)
        return hci_fail(c, HCI_ERR_ARG, "hci_davidson: max_sub must be in [2, 32]");
    if (opts->restart_keep < 1 || opts->restart_keep > kMaxKeep ||
        (opts->restart_keep > 1 && opts->restart_keep >= opts->max_sub))
        return hci_fail(c, HCI_ERR_ARG, "hci_davidson: restart_keep must be in [1, min(8, max_sub - 1)]");
    if (opts->precond == HCI_PRECOND_MLP && !mlp_ready(c))
        return hci_fail(c, HCI_ERR_STATE, "hci_davidson: MLP preconditioner requested but no network loaded");
    if (opts->mixed == 2 && h->prec != HCI_PREC_F16)
        return hci_fail(c, HCI_ERR_ARG, "hci_davidson: mixed = 2 needs HCI_PREC_F16 storage");
    const bool use_f32 = opts->mixed && (h->prec == HCI_PREC_F32 || h->prec == HCI_PREC_F16);
    const bool use_f16 = opts->mixed == 2;
    // FP16 rounding (2^-11 of each element) leaves the residual a few times
    // tol from the full-precision one: FP32 sigmas take over at 3 tol
    const double f16_tol = opts->f16_tol > 0.0 ? opts->f16_tol : 3.0 * opts->tol;
    auto t0 = std::chrono::steady_clock::now();
    HIP_TRY(c, hipSetDevice(c->device));
    hipStream_t s = c->stream;
    // reference: max_search_space = nvalues + 10 if that is < dim, else dim
    const int max_sub = (int)std::min<int64_t>(opts->max_sub, N);
    const int init_dim = init ? 1 : 2;
    DavWork w;
    w.c = c;
    w.n = N;
    w.max_sub = max_sub;
    const size_t vb = (size_t)N * 8;
    if (int rc = w.carve(std::max(max_sub, 2))) return rc;

    hci_dav_stats st{};
    st.refine_iter = -1;
    st.f16_iter = -1;
    int sigma_mode = use_f16 ? HCI_SIGMA_F16 : use_f32 ? HCI_SIGMA_F32 : HCI_SIGMA_FULL;
    // every sigma product is bracketed by events on the context stream
    std::vector<hipEvent_t> ev;
    struct EvFree {
        std::vector<hipEvent_t>& e;
        ~EvFree() {
            for (auto x : e) hipEventDestroy(x);
        }
    } ev_free{ev};
    auto sigma = [&](const double* v, double* out) {
        hipEvent_t a, b;
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess)
            return hci_fail(c, HCI_ERR_HIP, "hci_davidson: event creation");
        ev.push_back(a);
        ev.push_back(b);
        hipEventRecord(a, s);
        int rc = spmv_device(h, v, out, sigma_mode);
        hipEventRecord(b, s);
        if (sigma_mode == HCI_SIGMA_F16)
            ++st.n_sigma_f16;
        else if (sigma_mode == HCI_SIGMA_F32)
            ++st.n_sigma_f32;
        else
            ++st.n_sigma_full;
        return rc;
    };
#define DT(expr)                   \
    do {                           \
        int _rc = (expr);          \
        if (_rc) return _rc;       \
    } while (0)

    int dim_sub = std::min<int64_t>(init_dim, N);
    if (init) {
        // caller order -> internal order
        HIP_TRY(c, hipMemcpyAsync(w.t, init, vb, opts->mem == HCI_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
        permute_to_internal(h, w.t, w.V);
    } else {
        // identity columns of the caller's first determinants (davidson.rs:325-327)
        int32_t hp[2] = {0, 0};
        HIP_TRY(c, hipMemcpy(hp, h->d_pos, (size_t)dim_sub * 4, hipMemcpyDeviceToHost));
        for (int k = 0; k < dim_sub; ++k)
            hipLaunchKernelGGL(k_unit, grid_for(N, 256), 256, 0, s, w.col(w.V, k), N, (int64_t)hp[k]);
    }
    std::vector<double> P;  // row-major dim_sub x dim_sub
    auto rebuild_projection = [&]() -> int {
        P.assign((size_t)dim_sub * dim_sub, 0.0);
        for (int k = 0; k < dim_sub; ++k) {
            DT(sigma(w.col(w.V, k), w.col(w.W, k)));
        }
        for (int k = 0; k < dim_sub; ++k) {
            DT(w.multidot(w.V, dim_sub, w.col(w.W, k)));
            for (int r = 0; r < dim_sub; ++r) P[(size_t)r * dim_sub + k] = w.h_red[r];
        }
        // symmetrise the way the reference fills it (upper block from V^T w, mirrored)
        for (int r = 0; r < dim_sub; ++r)
            for (int q = r + 1; q < dim_sub; ++q) P[(size_t)q * dim_sub + r] = P[(size_t)r * dim_sub + q];
        return HCI_OK;
    };
    DT(rebuild_projection());

    double last_eig = 0.0;
    double lam = 0.0, err = 0.0;
    bool converged = false;
    bool mlp_prev = false;
    double err_prev = 1e300, best_err = 1e300;
    int since_best = 0;
    int it_now = 0;
    std::vector<double> y_prev;  // previous Ritz coefficients in the current basis
    // the next step of the precision ladder (FP16 -> FP32 -> full-precision
    // values): the search space restarts from the current Ritz vector w.x.
    // Only full-precision sigmas can declare convergence: every exit of a
    // reduced-precision phase (its own convergence, a stall, or the 1e-9
    // floors on the correction) comes here instead
    auto next_precision = [&]() -> int {
        if (sigma_mode == HCI_SIGMA_F16) {
            sigma_mode = HCI_SIGMA_F32;
            st.f16_iter = it_now;
        } else {
            sigma_mode = HCI_SIGMA_FULL;
            st.refine_iter = it_now;
        }
        dim_sub = 1;
        HIP_TRY(c, hipMemcpyAsync(w.V, w.x, vb, hipMemcpyDeviceToDevice, s));
        DT(rebuild_projection());
        err_prev = 1e300;
        best_err = 1e300;
        since_best = 0;
        mlp_prev = false;
        y_prev.assign(1, 1.0);
        return HCI_OK;
    };
    MlpState* net = (opts->precond == HCI_PRECOND_MLP) ? c->mlp : nullptr;
    if (net) DT(csr_rownorm(const_cast<hci_csr*>(h)));
    int it;
    for (it = 0; it < opts->max_iters; ++it) {
        it_now = it;
        std::vector<double> vals, vecs;
        jacobi_eigen(dim_sub, P, vals, vecs);
        lam = vals[0];
        Coef y{};
        for (int k = 0; k < dim_sub; ++k) y.c[k] = vecs[(size_t)k * dim_sub];
        // the Ritz vector x and residual r are written when the step reads
        // them (network features, recorded rows) or leaves the search space
        // (ensure_x: convergence, the next precision, a restart to x); with
        // the diagonal correction t goes straight to the next basis column
        const bool want_xr = net || rec;
        double* t_out = (!net && dim_sub < max_sub) ? w.col(w.V, dim_sub) : w.t;
        // diagonal-correction steps (no network, no recorder) take V^T t and
        // W^T t from the Ritz pass itself: rd[2 + c] and rd[2 + km + c]
        const bool fused = !net && !rec && dim_sub < max_sub;
        const int km = km_bound(w.max_sub);
        const double* rd = w.h_red + 3 * (kMaxSub + 1);
        if (fused) {
            HCI_LAUNCH_KM(k_ritz_dots, w.max_sub, w.red_grid(), kRedThreads, s, w.V, w.W, N, dim_sub, y, lam,
                          h->d_diag, N, nullptr, nullptr, t_out, w.partial);
            DT(w.reduce_ritz(2 + 2 * km));
        } else {
            HCI_LAUNCH_KM(k_ritz, w.max_sub, w.red_grid(), kRedThreads, s, w.V, w.W, N, dim_sub, y, lam, h->d_diag,
                          N, want_xr ? w.x : nullptr, want_xr ? w.r : nullptr, t_out, w.partial);
            DT(w.reduce(2));
        }
        const double rr = fused ? rd[0] : w.h_red[0], tt = fused ? rd[1] : w.h_red[1];
        auto ensure_x = [&]() {
            if (!want_xr)
                HCI_LAUNCH_KM(k_ritz, w.max_sub, w.red_grid(), kRedThreads, s, w.V, w.W, N, dim_sub, y, lam,
                              h->d_diag, N, w.x, w.r, w.t, w.partial);
            return hipGetLastError() == hipSuccess ? HCI_OK : hci_fail(c, HCI_ERR_HIP, "k_ritz launch failed");
        };
        err = std::sqrt(rr);
        double tnorm = std::sqrt(tt);
        const bool conv = N <= 2 || (err < opts->tol && std::fabs(lam - last_eig) < opts->etol);
        if (conv && sigma_mode == HCI_SIGMA_FULL) {
            DT(ensure_x());
            converged = true;
            break;
        }
        // an MLP correction is judged by what it did to the residual: the step
        // after a network correction that did not lower the residual norm
        // takes the diagonal correction; the network is tried again after it
        const bool failed = net && mlp_prev && !(err < err_prev);
        if (failed) ++st.mlp_fallbacks;
        const bool mlp_now = net && !failed;
        last_eig = lam;
        // stall counter: iterations without a 30% residual improvement, counted
        // only near convergence (an early plateau while the ground state is
        // still being found is not a precision floor)
        if (err < best_err * 0.7) {
            best_err = err;
            since_best = 0;
        } else if (err < 100.0 * opts->tol) {
            ++since_best;
        }
        // FP32-value sigmas converge on the rounded matrix; refine once that
        // iteration converges, the residual reaches refine_tol, or it stalls
        // FP16-value sigmas first (mixed = 2): FP32 ones take over from the
        // Ritz vector once the FP16 iteration converges, reaches f16_tol, or
        // stalls at the FP16 rounding floor
        if ((sigma_mode == HCI_SIGMA_F16 && (conv || err < f16_tol || since_best >= 8)) ||
            (sigma_mode == HCI_SIGMA_F32 && (conv || err < opts->refine_tol || since_best >= 25))) {
            DT(ensure_x());
            DT(next_precision());
            continue;
        }
        err_prev = err;
        if (rec && dim_sub + 1 <= max_sub) DT(record_rows(c, h, w, *rec, lam, err, tnorm, dim_sub));
        if (dim_sub + 1 <= max_sub) {
            double* t = w.col(w.V, dim_sub);
            if (mlp_now) {
                DT(mlp_correction(c, h, lam, err, tnorm, w.r, w.x, w.t /*dpr*/, t));
                ++st.mlp_applied;
                mlp_prev = true;
            } else {
                if (t_out != t) HIP_TRY(c, hipMemcpyAsync(t, w.t, vb, hipMemcpyDeviceToDevice, s));
                mlp_prev = false;
            }
            // classical Gram-Schmidt with the coefficients kept on the device:
            // t = V[:, dim_sub], so one pass over V[:, 0..dim_sub] gives V^T t
            // and t.t (slot A); an orthogonalisation pass reads the previous
            // slot and reduces into the next (B, then C).  The second pass runs
            // only when the first left less than 1/sqrt(2) of t's norm (the
            // DGKS test: otherwise one pass is orthogonal to working precision)
            constexpr int kSlot = kMaxSub + 1;
            double* rA = w.red;
            double* rB = w.red + kSlot;
            double* rC = w.red + 2 * kSlot;
            Coef hc{};
            if (fused) {
                // V^T t from the Ritz pass: the first pass needs no dot pass
                for (int q = 0; q < dim_sub; ++q) hc.c[q] = rd[2 + q];
                HCI_LAUNCH_KM(k_orth_sub_dot, w.max_sub, w.red_grid(), kRedThreads, s, w.V, N, dim_sub, hc, nullptr,
                              t, N, w.partial);
                w.reduce_dev(dim_sub + 1, rB);
                w.h_red[dim_sub] = tt;  // slot A's t.t
                HIP_TRY(c, hipMemcpyAsync(w.h_red + kSlot, rB, kSlot * 8, hipMemcpyDeviceToHost, s));
            } else {
                HCI_LAUNCH_KM(k_multidot, w.max_sub, w.red_grid(), kRedThreads, s, w.V, N, dim_sub + 1, t, N,
                              w.partial);
                w.reduce_dev(dim_sub + 1, rA);
                HCI_LAUNCH_KM(k_orth_sub_dot, w.max_sub, w.red_grid(), kRedThreads, s, w.V, N, dim_sub, hc, rA, t,
                              N, w.partial);
                w.reduce_dev(dim_sub + 1, rB);
                HIP_TRY(c, hipMemcpyAsync(w.h_red, w.red, 2 * kSlot * 8, hipMemcpyDeviceToHost, s));
            }
            HIP_TRY(c, hipStreamSynchronize(s));
            if (mlp_now) tnorm = std::sqrt(w.h_red[dim_sub]);
            double nrm = std::sqrt(w.h_red[kSlot + dim_sub]);
            const bool second = !(nrm * nrm >= 0.5 * w.h_red[dim_sub]);
            if (second) {
                HCI_LAUNCH_KM(k_orth_sub_dot, w.max_sub, w.red_grid(), kRedThreads, s, w.V, N, dim_sub, hc, rB, t,
                              N, w.partial);
                w.reduce_dev(dim_sub + 1, rC);
                HIP_TRY(c, hipMemcpyAsync(w.h_red + 2 * kSlot, rC, kSlot * 8, hipMemcpyDeviceToHost, s));
                HIP_TRY(c, hipStreamSynchronize(s));
                nrm = std::sqrt(w.h_red[2 * kSlot + dim_sub]);
                ++st.orth_second;
            }
            // correction floors (davidson.rs:210-218, modified_gram_schmidt.rs:29)
            if (tnorm < 1e-9 || nrm < 1e-9) {
                DT(ensure_x());
                if (sigma_mode != HCI_SIGMA_FULL) {
                    DT(next_precision());
                    continue;
                }
                converged = true;
                break;
            }
            hipLaunchKernelGGL(k_scale, grid_for(N, 256), 256, 0, s, t, 1.0 / nrm, t, N);
            DT(sigma(t, w.col(w.W, dim_sub)));
            const int nd = dim_sub + 1;
            std::vector<double> pcol((size_t)nd);
            if (fused && !second) {
                // V^T w_new = W^T v_new = (W^T t - (V^T W) c) / |t'| for the old
                // basis (W = H V, H symmetric; c = V^T t, the one pass's
                // coefficients); only v_new . w_new needs a pass (two vectors)
                for (int r = 0; r < dim_sub; ++r) {
                    double a = rd[2 + km + r];
                    for (int q = 0; q < dim_sub; ++q) a -= P[(size_t)r * dim_sub + q] * hc.c[q];
                    pcol[r] = a / nrm;
                }
                DT(w.multidot(t, 1, w.col(w.W, dim_sub)));
                pcol[dim_sub] = w.h_red[0];
            } else {
                DT(w.multidot(w.V, dim_sub + 1, w.col(w.W, dim_sub)));
                for (int r = 0; r < nd; ++r) pcol[r] = w.h_red[r];
            }
            std::vector<double> np((size_t)nd * nd, 0.0);
            for (int r = 0; r < dim_sub; ++r)
                for (int q = 0; q < dim_sub; ++q) np[(size_t)r * nd + q] = P[(size_t)r * dim_sub + q];
            for (int r = 0; r < nd; ++r) {
                np[(size_t)r * nd + dim_sub] = pcol[r];
                np[(size_t)dim_sub * nd + r] = pcol[r];
            }
            P.swap(np);
            dim_sub = nd;
        } else if (opts->restart_keep >= 2) {
            // thick restart inside the subspace: Q = [y_0 .. y_{l-1}, y_prev orthonormalised]
            const int k = dim_sub;
            const int l = std::min(opts->restart_keep - 1, k - 1);
            std::vector<double> Q((size_t)k * kMaxKeep, 0.0);  // Q[c * kMaxKeep + j]
            for (int j = 0; j < l; ++j)
                for (int c = 0; c < k; ++c) Q[(size_t)c * kMaxKeep + j] = vecs[(size_t)c * k + j];
            int m = l;
            if (!y_prev.empty()) {
                std::vector<double> u(k, 0.0);
                for (size_t c = 0; c < y_prev.size() && (int)c < k; ++c) u[c] = y_prev[c];
                for (int pass = 0; pass < 2; ++pass)
                    for (int j = 0; j < l; ++j) {
                        double d = 0.0;
                        for (int c = 0; c < k; ++c) d += Q[(size_t)c * kMaxKeep + j] * u[c];
                        for (int c = 0; c < k; ++c) u[c] -= d * Q[(size_t)c * kMaxKeep + j];
                    }
                double nu = 0.0;
                for (int c = 0; c < k; ++c) nu += u[c] * u[c];
                nu = std::sqrt(nu);
                if (nu > 1e-8) {
                    for (int c = 0; c < k; ++c) Q[(size_t)c * kMaxKeep + m] = u[c] / nu;
                    ++m;
                }
            }
            // projected matrix in the new basis: Q^T P Q
            std::vector<double> PQ((size_t)k * m, 0.0), np((size_t)m * m, 0.0);
            for (int r = 0; r < k; ++r)
                for (int j = 0; j < m; ++j) {
                    double a = 0.0;
                    for (int c = 0; c < k; ++c) a += P[(size_t)r * k + c] * Q[(size_t)c * kMaxKeep + j];
                    PQ[(size_t)r * m + j] = a;
                }
            for (int i2 = 0; i2 < m; ++i2)
                for (int j = 0; j < m; ++j) {
                    double a = 0.0;
                    for (int r = 0; r < k; ++r) a += Q[(size_t)r * kMaxKeep + i2] * PQ[(size_t)r * m + j];
                    np[(size_t)i2 * m + j] = a;
                }
            for (int i2 = 0; i2 < m; ++i2)
                for (int j = i2 + 1; j < m; ++j) np[(size_t)j * m + i2] = np[(size_t)i2 * m + j];
            std::copy(Q.begin(), Q.end(), w.h_q);
            HIP_TRY(c, hipMemcpyAsync(w.q, w.h_q, Q.size() * 8, hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL(k_rotate, w.red_grid(), kRedThreads, 0, s, w.V, w.W, N, k, m, w.q, N);
            HIP_TRY(c, hipGetLastError());
            HIP_TRY(c, hipStreamSynchronize(s));  // h_q is reused by the next restart
            P.swap(np);
            dim_sub = m;
            // the current Ritz vector is the first new column
            y.c[0] = 1.0;
            for (int c2 = 1; c2 < kMaxSub; ++c2) y.c[c2] = 0.0;
            mlp_prev = false;
        } else {
            // restart from the Ritz vector (davidson.rs:258-269)
            DT(ensure_x());
            dim_sub = 1;
            HIP_TRY(c, hipMemcpyAsync(w.V, w.x, vb, hipMemcpyDeviceToDevice, s));
            DT(rebuild_projection());
            y.c[0] = 1.0;
            mlp_prev = false;
        }
        y_prev.assign(y.c, y.c + std::min(dim_sub, kMaxSub));
    }
#undef DT
    st.iters = it;
    st.converged = converged ? 1 : 0;
    st.residual = err;
    if (converged) {
        *e0 = lam;
        permute_to_caller(h, w.x, w.r);
        HIP_TRY(c, hipMemcpyAsync(coeffs, w.r, vb, opts->mem == HCI_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
    }
    HIP_TRY(c, hipStreamSynchronize(s));
    st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (size_t k = 0; k + 1 < ev.size(); k += 2) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ev[k], ev[k + 1]) == hipSuccess) st.sigma_ms += ms;
    }
    if (stats) *stats = st;
    return converged ? HCI_OK : HCI_NOT_CONVERGED;
}

extern "C" int hci_davidson(hci_ctx* c, const hci_csr* h, const double* init, const hci_dav_opts* opts,
                            double* e0, double* coeffs, hci_dav_stats* stats) {
    return davidson_run(c, h, init, opts, e0, coeffs, stats, nullptr);
}

extern "C" int hci_mlp_collect(hci_ctx* c, const hci_csr* h, const double* x_exact, const hci_dav_opts* opts,
                               int64_t per_iter, float* X, float* m, float* y, int64_t cap, int64_t* count) {
    if (!c || !h || !x_exact || !opts || !X || !m || !y || !count || cap < 0 || per_iter < 1)
        return c ? hci_fail(c, HCI_ERR_ARG, "hci_mlp_collect: bad arguments") : HCI_ERR_ARG;
    if (h->ctx != c) return hci_fail(c, HCI_ERR_ARG, "hci_mlp_collect: matrix from another context");
    HIP_TRY(c, hipSetDevice(c->device));
    const int64_t N = h->n;
    const size_t F = HCI_MLP_NFEAT;
    Recorder rec;
    rec.per_iter = per_iter;
    rec.cap = cap;
    double* e_dev = nullptr;
    double* e_int = nullptr;
    struct Free {
        std::vector<void*> p;
        ~Free() {
            for (void* q : p) hipFree(q);
        }
    } fr;
    auto alloc = [&](void** q, size_t bytes) {
        hipError_t e = hipMalloc(q, bytes + 16);
        if (e == hipSuccess) fr.p.push_back(*q);
        return e;
    };
    HIP_TRY(c, alloc((void**)&e_dev, (size_t)N * 8));
    HIP_TRY(c, alloc((void**)&e_int, (size_t)N * 8));
    HIP_TRY(c, alloc((void**)&rec.e_perp, (size_t)N * 8));
    HIP_TRY(c, alloc((void**)&rec.feat, (size_t)N * F * 4));
    HIP_TRY(c, alloc((void**)&rec.X, (size_t)cap * F * 4));
    HIP_TRY(c, alloc((void**)&rec.m, (size_t)cap * 4));
    HIP_TRY(c, alloc((void**)&rec.y, (size_t)cap * 4));
    HIP_TRY(c, hipMemcpyAsync(e_dev, x_exact, (size_t)N * 8,
                              opts->mem == HCI_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, c->stream));
    permute_to_internal(h, e_dev, e_int);
    rec.e_int = e_int;
    hci_dav_opts o = *opts;
    o.precond = HCI_PRECOND_DPR;
    o.mem = HCI_MEM_HOST;
    std::vector<double> coeffs((size_t)N);
    double e0 = 0.0;
    int rc = davidson_run(c, h, nullptr, &o, &e0, coeffs.data(), nullptr, &rec);
    if (rc < 0) return rc;
    *count = rec.count;
    if (rec.count > 0) {
        HIP_TRY(c, hipMemcpy(X, rec.X, (size_t)rec.count * F * 4, hipMemcpyDeviceToHost));
        HIP_TRY(c, hipMemcpy(m, rec.m, (size_t)rec.count * 4, hipMemcpyDeviceToHost));
        HIP_TRY(c, hipMemcpy(y, rec.y, (size_t)rec.count * 4, hipMemcpyDeviceToHost));
    }
    return HCI_OK;
}
