// Learned Davidson preconditioner: a 24 -> 128 -> 128 -> 1 SiLU network whose
// output o_p gives a bounded positive factor g_p = 1 + tanh(o_p) / 2 in
// (1/2, 3/2) that rescales the diagonal correction, t_p = g_p * r_p / (lambda - H_pp),
// for every determinant p.  Training pairs are recorded from the library's own
// Davidson runs on small systems: per-determinant features and the exact
// correction, i.e. the part of the converged eigenvector not yet in the
// search space.  The Davidson loop falls back to the diagonal correction
// whenever a network step does not reduce the residual norm, and the converged
// FP64 eigenvalue decides convergence, so the network can change the iteration
// count but not the answer.
//
// All matrix products (forward and backward) run on the matrix cores through
// v_mfma_f32_16x16x16_bf16: operands are rounded to bfloat16 on load, products
// accumulate in FP32.  Master weights, Adam moments and activations are FP32.
// Training runs layer by layer (k_gemm_bf16); inference inside Davidson runs
// all three layers in one kernel (k_mlp_infer) with the bfloat16 weights in
// LDS and each wave's activations in registers.
#include <chrono>
#include <cmath>
#include <cstring>

#include "mlp.h"

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int F = HCI_MLP_NFEAT;
constexpr int Hd = HCI_MLP_HIDDEN;
constexpr float kGAmp = 0.5f;  // g = 1 + kGAmp tanh(out)
// parameter offsets (row-major weights: W[in][out])
constexpr int oW1 = 0, ob1 = oW1 + F * Hd, oW2 = ob1 + Hd, ob2 = oW2 + Hd * Hd, oW3 = ob2 + Hd, ob3 = oW3 + Hd;
static_assert(ob3 + 1 == HCI_MLP_NPARAM, "parameter layout");

// round-to-nearest-even FP32 -> bfloat16 (finite inputs)
__device__ __forceinline__ short to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

// C(m, n) = sum_k A(m, k) B(k, n) (+ bias(n)) (+ C if accumulate), with
// A(m, k) = A[m*sam + k*sak], B(k, n) = B[k*sbk + n*sbn], C(m, n) = C[m*scm + n*scn].
// One wavefront per 16x16 tile of C, one MFMA per 16-deep slice of K.
// Operand layout of v_mfma_f32_16x16x16_bf16: lane l holds A[l%16][4(l/16)..+3]
// and B[4(l/16)..+3][l%16]; the result lane l holds C[4(l/16)+i][l%16], i<4.
__global__ void __launch_bounds__(256)
    k_gemm_bf16(int M, int N, int K, const float* __restrict__ A, int64_t sam, int64_t sak,
                const float* __restrict__ B, int64_t sbk, int64_t sbn, float* __restrict__ C, int64_t scm,
                int64_t scn, const float* __restrict__ bias, int accumulate) {
    const int lane = threadIdx.x & 63;
    const int64_t w = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int tiles_n = (N + 15) >> 4;
    const int64_t tm = w / tiles_n;
    const int tn = (int)(w % tiles_n);
    if (tm * 16 >= M) return;
    const int64_t m0 = tm * 16;
    const int n0 = tn * 16;
    const int row = lane & 15, kq = (lane >> 4) * 4;
    const int64_t am = m0 + row;
    const int bn = n0 + row;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += 16) {
        bf16x4 a, b;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int k = k0 + kq + i;
            a[i] = (am < M && k < K) ? to_bf16(A[am * sam + (int64_t)k * sak]) : (short)0;
            b[i] = (bn < N && k < K) ? to_bf16(B[(int64_t)k * sbk + (int64_t)bn * sbn]) : (short)0;
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int64_t mm = m0 + kq + i;
        const int nn = n0 + row;
        if (mm < M && nn < N) {
            float v = acc[i] + (bias ? bias[nn] : 0.f);
            float* dst = C + mm * scm + (int64_t)nn * scn;
            *dst = accumulate ? *dst + v : v;
        }
    }
}

__device__ __forceinline__ float sigmoid(float z) { return 1.f / (1.f + __expf(-z)); }

__global__ void k_silu_fwd(const float* Z, float* H, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) H[i] = Z[i] * sigmoid(Z[i]);
}

__global__ void k_silu_bwd(const float* dH, const float* Z, float* dZ, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float s = sigmoid(Z[i]);
    dZ[i] = dH[i] * s * (1.f + Z[i] * (1.f - s));
}

// The network's scale factor is bounded and positive: g = 1 + kGAmp tanh(out).
// A positive factor keeps every component of the diagonal correction's sign.
// loss = (1/n) sum_i (g_i m_i - y_i)^2 ; dout_i = 2 (g_i m_i - y_i) m_i g'_i / n
__global__ void k_loss(const float* out, const float* m, const float* y, int64_t n, float* dout, double* partial) {
    __shared__ double sh[256];
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float th = tanhf(out[i]);
        const float r = (1.f + kGAmp * th) * m[i] - y[i];
        acc += (double)r * r;
        dout[i] = 2.f * r * m[i] * kGAmp * (1.f - th * th) / (float)n;
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// out[c] = sum_r X[r*cols + c]; one block per column, fixed order
__global__ void k_colsum(const float* X, int64_t rows, int cols, float* out) {
    __shared__ float sh[256];
    const int c = blockIdx.x;
    float acc = 0.f;
    for (int64_t r = threadIdx.x; r < rows; r += blockDim.x) acc += X[r * cols + c];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = sh[0];
}

__global__ void k_adam(float* p, const float* g, float* m1, float* m2, int n, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    m1[i] = b1 * m1[i] + (1.f - b1) * gi;
    m2[i] = b2 * m2[i] + (1.f - b2) * gi * gi;
    p[i] -= lr * (m1[i] / c1) / (sqrtf(m2[i] / c2) + eps);
}

__global__ void k_normalize(const float* X, const float* mean, const float* stdv, int64_t n, float* Xn) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * F) return;
    const int f = (int)(i % F);
    Xn[i] = (X[i] - mean[f]) / stdv[f];
}

__global__ void k_gather_rows(const float* X, const float* m, const float* y, const int64_t* idx, int64_t b,
                              float* Xb, float* mb, float* yb) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= b * F) return;
    const int64_t r = i / F;
    const int f = (int)(i % F);
    const int64_t src = idx[r];
    Xb[i] = X[src * F + f];
    if (f == 0) {
        mb[r] = m[src];
        yb[r] = y[src];
    }
}

// Per-determinant features (internal order), see feature_doc below.
template <class S>
__global__ void k_features(const uint64_t* words, const double* diag, const double* rownorm, const double* r,
                           const double* x, const double* t, double lam, double rs, double ts, double xs,
                           int64_t n, int n_up, int n_dn, int f0, float* X) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const S up = load_string<S>(words + 4 * p), dn = load_string<S>(words + 4 * p + 2);
    const double d = diag[p] - lam;
    float o[F];
    o[0] = (float)d;
    o[1] = (float)log(fabs(d) + 1e-6);
    o[2] = (float)(r[p] * rs);
    o[3] = (float)(t[p] * ts);
    o[4] = (float)(x[p] * xs);
    o[5] = (float)rownorm[p];
    o[6] = (float)(rownorm[p] / (fabs(d) + 1e-3));
    for (int k = 0; k < 16; ++k) {
        const int orb = f0 + k;
        const int bu = (int)((up >> orb) & (S)1), bd = (int)((dn >> orb) & (S)1);
        o[7 + k] = 0.5f * (float)(bu + bd);
    }
    const S hfu = lowmask<S>(n_up), hfd = lowmask<S>(n_dn);
    o[23] = 0.125f * (float)(popc(up ^ hfu) + popc(dn ^ hfd));
    // a row is 96 bytes: six 16-byte stores
    float4* out = (float4*)(X + p * F);
#pragma unroll
    for (int j = 0; j < F / 4; ++j) out[j] = make_float4(o[4 * j], o[4 * j + 1], o[4 * j + 2], o[4 * j + 3]);
}

// bfloat16 weight image for k_mlp_infer, in its LDS layout: W1 and W2
// transposed (row n holds the weights into output n, so a lane's 4 K values
// are contiguous), rows padded to kW1Row / kW2Row values against LDS bank
// conflicts (W1's K padded from 24 to 32 with zeros), then W3.
constexpr int kK1 = 32;
constexpr int kW1Row = 40, kW2Row = Hd + 8;
constexpr int kImgW1 = Hd * kW1Row, kImgW2 = Hd * kW2Row, kImg = kImgW1 + kImgW2 + Hd;

__device__ __forceinline__ float from_bf16(short h) { return __uint_as_float((uint32_t)(uint16_t)h << 16); }

__global__ void k_prep_weights(const float* P, short* img) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kImgW1) {
        const int n = i / kW1Row, k = i % kW1Row;
        img[i] = k < F ? to_bf16(P[oW1 + k * Hd + n]) : (short)0;
    }
    if (i < kImgW2) {
        const int n = i / kW2Row, k = i % kW2Row;
        img[kImgW1 + i] = k < Hd ? to_bf16(P[oW2 + k * Hd + n]) : (short)0;
    }
    if (i < Hd) img[kImgW1 + kImgW2 + i] = to_bf16(P[oW3 + i]);
}


// Fused inference: each wave takes 32 rows at a time: standardised features
// (bf16, read straight into MFMA fragments) -> Z1 = X W1 + b1 (2 MFMA K-steps
// x 8 tiles per 16 rows) -> SiLU -> H1 (bf16, registers) -> Z2 = H1 W2 + b2
// (8 x 8 MFMAs per 16 rows) -> SiLU -> out = H2 . W3 + b3 (bf16 operands, FP32
// sums over the 16 lanes of a row group).  Writes out (if given) and
// t = (1 + kGAmp tanh(out)) * dpr (if given).  The same roundings as the
// layer-by-layer path: operands to bf16 (round to nearest even), FP32
// accumulation.
//
// Eight waves per block share one LDS copy of the weights (46 KB), so two
// blocks fit a CU; at most 128 VGPRs per lane give four waves per SIMD.  The
// next 32 rows' features and this task's dpr are loaded while the current
// rows are computed.
constexpr int kInfWaves = 8;
constexpr int kInfMaxBlocks = 2048;  // 524,288 rows per pass of the grid
// SiLU with the hardware reciprocal (1 ulp) for inference, whose activations
// are rounded to bfloat16 next anyway
__device__ __forceinline__ float silu_fast(float z) { return z * __builtin_amdgcn_rcpf(1.f + __expf(-z)); }
// round to nearest even with the hardware conversion (same result as to_bf16)
__device__ __forceinline__ short bf16_rne(float f) { return __builtin_bit_cast(short, (__bf16)f); }

// raw features of 32 rows: lane holds X[row 16 h + r16][16 ks + kq + j]
struct XRaw {
    f32x4 v[2][kK1 / 16];
};
__device__ __forceinline__ void load_x(XRaw& x, const float* __restrict__ X, int64_t n, int64_t g0, int r16,
                                       int kq) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int64_t row = g0 + 16 * h + r16;
#pragma unroll
        for (int ks = 0; ks < kK1; ks += 16) {
            const int k0 = ks + kq;
            x.v[h][ks / 16] = (row < n && k0 < F) ? *(const f32x4*)&X[row * F + k0] : f32x4{0.f, 0.f, 0.f, 0.f};
        }
    }
}

// One wave per 32 determinants (two 16-row halves sharing each weight
// fragment).  Layer 1 is computed transposed (W1^T X^T), which leaves every
// lane holding H[row][16 ks + 4 (lane / 16) + i] -- exactly its A fragment of
// layer 2's K-step ks -- so the hidden activations never leave registers.
__global__ void __launch_bounds__(64 * kInfWaves, 4)
    k_mlp_infer(const float* __restrict__ X, int64_t n, const float* __restrict__ mean,
                const float* __restrict__ stdv, const short* __restrict__ img, const float* __restrict__ P,
                const double* __restrict__ dpr, float* out, double* t) {
    __shared__ __align__(16) short s_w[kImgW1 + kImgW2];
    __shared__ float s_b1[Hd], s_b2[Hd], s_w3[Hd], s_mu[kK1], s_sd[kK1];
    const int tid = threadIdx.x;
    for (int i = tid; i < (kImgW1 + kImgW2) / 2; i += blockDim.x) ((int*)s_w)[i] = ((const int*)img)[i];
    for (int i = tid; i < Hd; i += blockDim.x) {
        s_b1[i] = P[ob1 + i];
        s_b2[i] = P[ob2 + i];
        s_w3[i] = from_bf16(img[kImgW1 + kImgW2 + i]);
    }
    for (int i = tid; i < kK1; i += blockDim.x) {
        s_mu[i] = i < F ? mean[i] : 0.f;
        s_sd[i] = i < F ? stdv[i] : 1.f;
    }
    __syncthreads();
    const float b3 = P[ob3];
    const short* s_w1 = s_w;
    const short* s_w2 = s_w + kImgW1;
    const int lane = tid & 63, wv = tid >> 6;
    const int r16 = lane & 15, kq = (lane >> 4) * 4;
    // the row this lane writes: lanes with r16 < 8 of group kq take row
    // 16 (r16 / 4) + kq + r16 % 4, whose sum they hold after the reduction
    const int oh = (r16 >> 2) & 1, oi = r16 & 3;
    const int orow = 16 * oh + kq + oi;
    const int64_t stride = (int64_t)gridDim.x * kInfWaves * 32;
    int64_t g0 = ((int64_t)blockIdx.x * kInfWaves + wv) * 32;
    XRaw xr;
    if (g0 < n) load_x(xr, X, n, g0, r16, kq);
    for (; g0 < n; g0 += stride) {
        // the weights and biases are read from LDS where they are used: held
        // across tasks they would take registers the occupancy needs
        __asm__ volatile("" ::: "memory");
        double dp = 0.0;
        if (t && r16 < 8 && g0 + orow < n) dp = dpr[g0 + orow];
        // X fragments (B operand of the transposed layer 1): X[row][ks + kq + j]
        bf16x4 xf[2][kK1 / 16];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int64_t row = g0 + 16 * h + r16;
#pragma unroll
            for (int ks = 0; ks < kK1; ks += 16) {
                const int k0 = ks + kq;
                bf16x4 b;
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    b[j] = k0 + j < F && row < n ? bf16_rne((xr.v[h][ks / 16][j] - s_mu[k0 + j]) / s_sd[k0 + j])
                                                 : (short)0;
                xf[h][ks / 16] = b;
            }
        }
        if (g0 + stride < n) load_x(xr, X, n, g0 + stride, r16, kq);
        // layer 1, transposed: lane holds z1[row 16 h + r16][hidden 16 nt + kq + i]
        bf16x4 a2[2][Hd / 16];
#pragma unroll
        for (int nt = 0; nt < Hd / 16; ++nt) {
            f32x4 c0 = f32x4{0.f, 0.f, 0.f, 0.f}, c1 = c0;
#pragma unroll
            for (int ks = 0; ks < kK1; ks += 16) {
                const bf16x4 w = *(const bf16x4*)&s_w1[(nt * 16 + r16) * kW1Row + ks + kq];
                c0 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w, xf[0][ks / 16], c0, 0, 0, 0);
                c1 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w, xf[1][ks / 16], c1, 0, 0, 0);
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float bias = s_b1[nt * 16 + kq + i];
                a2[0][nt][i] = bf16_rne(silu_fast(c0[i] + bias));
                a2[1][nt][i] = bf16_rne(silu_fast(c1[i] + bias));
            }
        }
        // layer 2: lane holds z2[row 16 h + kq + i][hidden 16 nt + r16]
        float o[2][4] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
#pragma unroll 2
        for (int nt = 0; nt < Hd / 16; ++nt) {
            f32x4 c0 = f32x4{0.f, 0.f, 0.f, 0.f}, c1 = c0;
#pragma unroll
            for (int ks = 0; ks < Hd / 16; ++ks) {
                const bf16x4 w = *(const bf16x4*)&s_w2[(nt * 16 + r16) * kW2Row + ks * 16 + kq];
                c0 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a2[0][ks], w, c0, 0, 0, 0);
                c1 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a2[1][ks], w, c1, 0, 0, 0);
            }
            const int col = nt * 16 + r16;
            const float bias = s_b2[col], w3 = s_w3[col];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                o[0][i] += from_bf16(bf16_rne(silu_fast(c0[i] + bias))) * w3;
                o[1][i] += from_bf16(bf16_rne(silu_fast(c1[i] + bias))) * w3;
            }
        }
        // butterfly over the 16 lanes of each group: every lane holds the sums
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int d = 1; d < 16; d <<= 1) o[h][i] += __shfl_xor(o[h][i], d);
        float mine = o[0][0];
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (oh == h && oi == i) mine = o[h][i];
        const int64_t row = g0 + orow;
        if (r16 < 8 && row < n) {
            const float v = mine + b3;
            if (out) out[row] = v;
            if (t) t[row] = (double)(1.f + kGAmp * tanhf(v)) * dp;
        }
    }
}

unsigned gemm_grid(int64_t M, int N) { return grid_for(((M + 15) / 16) * ((N + 15) / 16), 4); }

}  // namespace

struct MlpState {
    float* P = nullptr;      // parameters
    float* G = nullptr;      // gradients
    float* A1 = nullptr;     // Adam moments
    float* A2 = nullptr;
    float* mean = nullptr;   // feature standardisation
    float* stdv = nullptr;
    int64_t cap = 0;         // rows of the work buffers
    float *Xn = nullptr, *Z1 = nullptr, *H1 = nullptr, *Z2 = nullptr, *H2 = nullptr, *out = nullptr;
    float *dout = nullptr, *dH = nullptr, *dZ = nullptr, *Xb = nullptr, *mb = nullptr, *yb = nullptr;
    int64_t* idx = nullptr;
    double* partial = nullptr;
    float* feat = nullptr;   // Davidson-time features (n x F) and outputs
    float* gout = nullptr;
    short* img = nullptr;    // bfloat16 weight image for k_mlp_infer
    int64_t feat_cap = 0;
    bool ready = false;      // parameters and standardisation loaded
    int64_t adam_t = 0;

    void free_work() {
        for (float* q : {Xn, Z1, H1, Z2, H2, out, dout, dH, dZ, Xb, mb, yb}) hipFree(q);
        hipFree(idx);
        Xn = Z1 = H1 = Z2 = H2 = out = dout = dH = dZ = Xb = mb = yb = nullptr;
        idx = nullptr;
        cap = 0;
    }
    ~MlpState() {
        free_work();
        for (float* q : {P, G, A1, A2, mean, stdv, feat, gout}) hipFree(q);
        hipFree(partial);
        hipFree(img);
    }
};

namespace {

int ensure_state(hci_ctx* c) {
    if (c->mlp) return HCI_OK;
    MlpState* s = new MlpState();
    c->mlp = s;
    HIP_TRY(c, hipMalloc(&s->P, HCI_MLP_NPARAM * 4));
    HIP_TRY(c, hipMalloc(&s->G, HCI_MLP_NPARAM * 4));
    HIP_TRY(c, hipMalloc(&s->A1, HCI_MLP_NPARAM * 4));
    HIP_TRY(c, hipMalloc(&s->A2, HCI_MLP_NPARAM * 4));
    HIP_TRY(c, hipMalloc(&s->mean, F * 4));
    HIP_TRY(c, hipMalloc(&s->stdv, F * 4));
    HIP_TRY(c, hipMalloc(&s->partial, 1024 * 8));
    HIP_TRY(c, hipMalloc(&s->img, (size_t)kImg * 2 + 16));
    HIP_TRY(c, hipMemset(s->A1, 0, HCI_MLP_NPARAM * 4));
    HIP_TRY(c, hipMemset(s->A2, 0, HCI_MLP_NPARAM * 4));
    std::vector<float> one(F, 1.f), zero(F, 0.f);
    HIP_TRY(c, hipMemcpy(s->mean, zero.data(), F * 4, hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(s->stdv, one.data(), F * 4, hipMemcpyHostToDevice));
    return HCI_OK;
}

int ensure_work(hci_ctx* c, int64_t rows) {
    MlpState* s = c->mlp;
    if (rows <= s->cap) return HCI_OK;
    s->free_work();
    const size_t r = (size_t)rows;
    HIP_TRY(c, hipMalloc(&s->Xn, r * F * 4));
    HIP_TRY(c, hipMalloc(&s->Xb, r * F * 4));
    for (float** q : {&s->Z1, &s->H1, &s->Z2, &s->H2, &s->dH, &s->dZ}) HIP_TRY(c, hipMalloc(q, r * Hd * 4));
    for (float** q : {&s->out, &s->dout, &s->mb, &s->yb}) HIP_TRY(c, hipMalloc(q, r * 4));
    HIP_TRY(c, hipMalloc(&s->idx, r * 8));
    s->cap = rows;
    return HCI_OK;
}

void gemm(hipStream_t st, int64_t M, int N, int K, const float* A, int64_t sam, int64_t sak, const float* B,
          int64_t sbk, int64_t sbn, float* C, int64_t scm, int64_t scn, const float* bias, int acc) {
    hipLaunchKernelGGL(k_gemm_bf16, gemm_grid(M, N), 256, 0, st, (int)M, N, K, A, sam, sak, B, sbk, sbn, C, scm,
                       scn, bias, acc);
}

// forward on rows already standardised in s->Xn[0..n)
void forward(hci_ctx* c, int64_t n) {
    MlpState* s = c->mlp;
    hipStream_t st = c->stream;
    const float* P = s->P;
    gemm(st, n, Hd, F, s->Xn, F, 1, P + oW1, Hd, 1, s->Z1, Hd, 1, P + ob1, 0);
    hipLaunchKernelGGL(k_silu_fwd, grid_for(n * Hd, 256), 256, 0, st, s->Z1, s->H1, n * Hd);
    gemm(st, n, Hd, Hd, s->H1, Hd, 1, P + oW2, Hd, 1, s->Z2, Hd, 1, P + ob2, 0);
    hipLaunchKernelGGL(k_silu_fwd, grid_for(n * Hd, 256), 256, 0, st, s->Z2, s->H2, n * Hd);
    gemm(st, n, 1, Hd, s->H2, Hd, 1, P + oW3, 1, 1, s->out, 1, 1, P + ob3, 0);
}

// fused inference on raw features X (n rows): out and/or t = g * dpr
void infer(hci_ctx* c, const float* X, int64_t n, float* out, const double* dpr, double* t) {
    MlpState* s = c->mlp;
    hipStream_t st = c->stream;
    hipLaunchKernelGGL(k_prep_weights, grid_for(kImgW2, 256), 256, 0, st, s->P, s->img);
    const unsigned grid = (unsigned)std::min<int64_t>(grid_for((n + 31) / 32, kInfWaves), kInfMaxBlocks);
    hipLaunchKernelGGL(k_mlp_infer, grid, 64 * kInfWaves, 0, st, X, n, s->mean, s->stdv, s->img, s->P, dpr, out, t);
}

// loss and gradients for the batch in Xn / mb / yb (n rows); returns the loss
int backward(hci_ctx* c, int64_t n, double* loss) {
    MlpState* s = c->mlp;
    hipStream_t st = c->stream;
    const float* P = s->P;
    float* G = s->G;
    const unsigned lg = (unsigned)std::min<int64_t>(grid_for(n, 256), 1024);
    hipLaunchKernelGGL(k_loss, lg, 256, 0, st, s->out, s->mb, s->yb, n, s->dout, s->partial);
    // layer 3
    gemm(st, Hd, 1, (int)n, s->H2, 1, Hd, s->dout, 1, 0, G + oW3, 1, 1, nullptr, 0);
    hipLaunchKernelGGL(k_colsum, 1, 256, 0, st, s->dout, n, 1, G + ob3);
    gemm(st, n, Hd, 1, s->dout, 1, 0, P + oW3, 0, 1, s->dH, Hd, 1, nullptr, 0);
    hipLaunchKernelGGL(k_silu_bwd, grid_for(n * Hd, 256), 256, 0, st, s->dH, s->Z2, s->dZ, n * Hd);
    // layer 2
    gemm(st, Hd, Hd, (int)n, s->H1, 1, Hd, s->dZ, Hd, 1, G + oW2, Hd, 1, nullptr, 0);
    hipLaunchKernelGGL(k_colsum, Hd, 256, 0, st, s->dZ, n, Hd, G + ob2);
    gemm(st, n, Hd, Hd, s->dZ, Hd, 1, P + oW2, 1, Hd, s->dH, Hd, 1, nullptr, 0);
    hipLaunchKernelGGL(k_silu_bwd, grid_for(n * Hd, 256), 256, 0, st, s->dH, s->Z1, s->dZ, n * Hd);
    // layer 1
    gemm(st, F, Hd, (int)n, s->Xn, 1, F, s->dZ, Hd, 1, G + oW1, Hd, 1, nullptr, 0);
    hipLaunchKernelGGL(k_colsum, Hd, 256, 0, st, s->dZ, n, Hd, G + ob1);
    HIP_TRY(c, hipGetLastError());
    if (loss) {
        std::vector<double> part(lg);
        HIP_TRY(c, hipMemcpyAsync(part.data(), s->partial, lg * 8, hipMemcpyDeviceToHost, st));
        HIP_TRY(c, hipStreamSynchronize(st));
        double t = 0.0;
        for (double v : part) t += v;
        *loss = t / (double)n;
    }
    return HCI_OK;
}

// host or device input -> device pointer (copy when host)
struct DevIn {
    const float* p = nullptr;
    float* own = nullptr;
    ~DevIn() { hipFree(own); }
};

int dev_in(hci_ctx* c, const float* src, size_t count, int mem, DevIn& d) {
    if (mem == HCI_MEM_DEVICE) {
        d.p = src;
        return HCI_OK;
    }
    HIP_TRY(c, hipMalloc(&d.own, count * 4 + 16));
    HIP_TRY(c, hipMemcpyAsync(d.own, src, count * 4, hipMemcpyHostToDevice, c->stream));
    d.p = d.own;
    return HCI_OK;
}

// splitmix64 -> uniform in [0, 1)
struct Rng {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uni() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
};

}  // namespace

bool mlp_ready(const hci_ctx* c) { return c && c->mlp && c->mlp->ready; }

void mlp_free_state(hci_ctx* c) {
    if (c && c->mlp) {
        delete c->mlp;
        c->mlp = nullptr;
    }
}

int mlp_correction(hci_ctx* c, const hci_csr* h, double lam, double rnorm, double tnorm, const double* r,
                   const double* x, const double* dpr, double* t) {
    MlpState* s = c->mlp;
    const int64_t n = h->n;
    if (n > s->feat_cap) {
        hipFree(s->feat);
        hipFree(s->gout);
        s->feat = s->gout = nullptr;
        s->feat_cap = 0;
        HIP_TRY(c, hipMalloc(&s->feat, (size_t)n * F * 4));
        HIP_TRY(c, hipMalloc(&s->gout, (size_t)n * 4));
        s->feat_cap = n;
    }
    int rc = mlp_features(c, h, lam, rnorm, tnorm, r, x, dpr, s->feat);
    if (rc) return rc;
    infer(c, s->feat, n, nullptr, dpr, t);
    HIP_TRY(c, hipGetLastError());
    return HCI_OK;
}

int mlp_features(hci_ctx* c, const hci_csr* h, double lam, double rnorm, double tnorm, const double* r,
                 const double* x, const double* dpr, float* X) {
    int rc = csr_rownorm(const_cast<hci_csr*>(h));
    if (rc) return rc;
    const int64_t n = h->n;
    if (h->n_up < 0) {  // electron counts read once per matrix (a read here would stall the stream)
        uint64_t w0[4];
        HIP_TRY(c, hipMemcpy(w0, h->d_words, 32, hipMemcpyDeviceToHost));
        hci_csr* hm = const_cast<hci_csr*>(h);
        hm->n_up = __builtin_popcountll(w0[0]) + __builtin_popcountll(w0[1]);
        hm->n_dn = __builtin_popcountll(w0[2]) + __builtin_popcountll(w0[3]);
    }
    const int n_up = h->n_up, n_dn = h->n_dn;
    int f0 = (n_up + n_dn) / 2 - 8;
    f0 = std::max(0, std::min(f0, c->n_orb - 16));
    const double sn = std::sqrt((double)n);
    const double rs = rnorm > 0 ? sn / rnorm : 0.0, ts = tnorm > 0 ? sn / tnorm : 0.0;
    if (c->n_orb <= 64)
        hipLaunchKernelGGL(k_features<uint64_t>, grid_for(n, 256), 256, 0, c->stream, h->d_words, h->d_diag,
                           h->d_rownorm, r, x, dpr, lam, rs, ts, sn, n, n_up, n_dn, f0, X);
    else
        hipLaunchKernelGGL(k_features<u128>, grid_for(n, 256), 256, 0, c->stream, h->d_words, h->d_diag,
                           h->d_rownorm, r, x, dpr, lam, rs, ts, sn, n, n_up, n_dn, f0, X);
    HIP_TRY(c, hipGetLastError());
    return HCI_OK;
}

extern "C" {

int hci_mlp_init(hci_ctx* c, uint64_t seed) {
    if (!c) return HCI_ERR_ARG;
    HIP_TRY(c, hipSetDevice(c->device));
    int rc = ensure_state(c);
    if (rc) return rc;
    // Kaiming-uniform weights, zero biases, last layer small so g starts near 1
    std::vector<float> p(HCI_MLP_NPARAM, 0.f);
    Rng g{seed};
    auto fill = [&](int off, int fan_in, int count, double gain) {
        const double a = gain * std::sqrt(6.0 / fan_in);
        for (int i = 0; i < count; ++i) p[off + i] = (float)((2.0 * g.uni() - 1.0) * a);
    };
    fill(oW1, F, F * Hd, 1.0);
    fill(oW2, Hd, Hd * Hd, 1.0);
    fill(oW3, Hd, Hd, 0.05);
    HIP_TRY(c, hipMemcpy(c->mlp->P, p.data(), HCI_MLP_NPARAM * 4, hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemset(c->mlp->A1, 0, HCI_MLP_NPARAM * 4));
    HIP_TRY(c, hipMemset(c->mlp->A2, 0, HCI_MLP_NPARAM * 4));
    c->mlp->adam_t = 0;
    c->mlp->ready = true;
    return HCI_OK;
}

int hci_mlp_set_params(hci_ctx* c, const float* params, const float* feat_mean, const float* feat_std) {
    if (!c || !params || !feat_mean || !feat_std) return c ? hci_fail(c, HCI_ERR_ARG, "hci_mlp_set_params: null") : HCI_ERR_ARG;
    HIP_TRY(c, hipSetDevice(c->device));
    int rc = ensure_state(c);
    if (rc) return rc;
    for (int f = 0; f < F; ++f)
        if (!(feat_std[f] > 0.f)) return hci_fail(c, HCI_ERR_ARG, "hci_mlp_set_params: feature std must be > 0");
    HIP_TRY(c, hipMemcpy(c->mlp->P, params, HCI_MLP_NPARAM * 4, hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->mlp->mean, feat_mean, F * 4, hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->mlp->stdv, feat_std, F * 4, hipMemcpyHostToDevice));
    c->mlp->ready = true;
    return HCI_OK;
}

int hci_mlp_get_params(hci_ctx* c, float* params, float* feat_mean, float* feat_std) {
    if (!c) return HCI_ERR_ARG;
    if (!c->mlp) return hci_fail(c, HCI_ERR_STATE, "hci_mlp_get_params: no network");
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (params) HIP_TRY(c, hipMemcpy(params, c->mlp->P, HCI_MLP_NPARAM * 4, hipMemcpyDeviceToHost));
    if (feat_mean) HIP_TRY(c, hipMemcpy(feat_mean, c->mlp->mean, F * 4, hipMemcpyDeviceToHost));
    if (feat_std) HIP_TRY(c, hipMemcpy(feat_std, c->mlp->stdv, F * 4, hipMemcpyDeviceToHost));
    return HCI_OK;
}

int hci_mlp_forward(hci_ctx* c, const float* X, int64_t n, float* out, int mem) {
    if (!c || n < 0 || (n > 0 && (!X || !out))) return c ? hci_fail(c, HCI_ERR_ARG, "hci_mlp_forward: bad arguments") : HCI_ERR_ARG;
    if (!mlp_ready(c)) return hci_fail(c, HCI_ERR_STATE, "hci_mlp_forward: no network");
    HIP_TRY(c, hipSetDevice(c->device));
    DevIn dx;
    int rc = dev_in(c, X, (size_t)n * F, mem, dx);
    if (rc) return rc;
    if (n == 0) return HCI_OK;
    float* dout = out;
    float* own = nullptr;
    if (mem != HCI_MEM_DEVICE) {
        HIP_TRY(c, hipMalloc(&own, (size_t)n * 4 + 16));
        dout = own;
    }
    infer(c, dx.p, n, dout, nullptr, nullptr);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && own) e = hipMemcpyAsync(out, own, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    hipFree(own);
    HIP_TRY(c, e);
    return HCI_OK;
}

int hci_mlp_loss_grad(hci_ctx* c, const float* X, const float* m, const float* y, int64_t n, float* grad,
                      double* loss, int mem) {
    if (!c || n < 1 || !X || !m || !y) return c ? hci_fail(c, HCI_ERR_ARG, "hci_mlp_loss_grad: bad arguments") : HCI_ERR_ARG;
    if (!mlp_ready(c)) return hci_fail(c, HCI_ERR_STATE, "hci_mlp_loss_grad: no network");
    if (n > (1 << 20)) return hci_fail(c, HCI_ERR_ARG, "hci_mlp_loss_grad: at most 2^20 rows");
    HIP_TRY(c, hipSetDevice(c->device));
    DevIn dx, dm, dy;
    int rc = dev_in(c, X, (size_t)n * F, mem, dx);
    if (!rc) rc = dev_in(c, m, (size_t)n, mem, dm);
    if (!rc) rc = dev_in(c, y, (size_t)n, mem, dy);
    if (!rc) rc = ensure_work(c, n);
    if (rc) return rc;
    MlpState* s = c->mlp;
    hipLaunchKernelGGL(k_normalize, grid_for(n * F, 256), 256, 0, c->stream, dx.p, s->mean, s->stdv, n, s->Xn);
    HIP_TRY(c, hipMemcpyAsync(s->mb, dm.p, (size_t)n * 4, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(s->yb, dy.p, (size_t)n * 4, hipMemcpyDeviceToDevice, c->stream));
    forward(c, n);
    rc = backward(c, n, loss);
    if (rc) return rc;
    if (grad) HIP_TRY(c, hipMemcpy(grad, s->G, HCI_MLP_NPARAM * 4, hipMemcpyDeviceToHost));
    return HCI_OK;
}

void hci_mlp_default_train_opts(hci_mlp_train_opts* o) {
    if (!o) return;
    o->epochs = 20;
    o->batch = 4096;
    o->lr = 1e-3f;
    o->seed = 7;
    o->standardize = 1;
    o->mem = HCI_MEM_HOST;
}

int hci_mlp_train(hci_ctx* c, const float* X, const float* m, const float* y, int64_t n,
                  const hci_mlp_train_opts* o, hci_mlp_train_stats* stats) {
    if (!c || !o || n < 1 || !X || !m || !y) return c ? hci_fail(c, HCI_ERR_ARG, "hci_mlp_train: bad arguments") : HCI_ERR_ARG;
    if (o->batch < 16 || o->batch > (1 << 16) || o->epochs < 0)
        return hci_fail(c, HCI_ERR_ARG, "hci_mlp_train: batch in [16, 65536], epochs >= 0");
    auto t0 = std::chrono::steady_clock::now();
    HIP_TRY(c, hipSetDevice(c->device));
    if (!mlp_ready(c)) {
        int rc = hci_mlp_init(c, o->seed);
        if (rc) return rc;
    }
    DevIn dx, dm, dy;
    int rc = dev_in(c, X, (size_t)n * F, o->mem, dx);
    if (!rc) rc = dev_in(c, m, (size_t)n, o->mem, dm);
    if (!rc) rc = dev_in(c, y, (size_t)n, o->mem, dy);
    if (!rc) rc = ensure_work(c, o->batch);
    if (rc) return rc;
    MlpState* s = c->mlp;
    hipStream_t st = c->stream;
    if (o->standardize) {
        // feature means / standard deviations from (a sample of) the training set
        const int64_t ns = std::min<int64_t>(n, 1 << 20);
        std::vector<float> hx((size_t)ns * F);
        HIP_TRY(c, hipMemcpyAsync(hx.data(), dx.p, hx.size() * 4, hipMemcpyDeviceToHost, st));
        HIP_TRY(c, hipStreamSynchronize(st));
        std::vector<double> mu(F, 0.0), var(F, 0.0);
        for (int64_t i = 0; i < ns; ++i)
            for (int f = 0; f < F; ++f) mu[f] += hx[(size_t)i * F + f];
        for (int f = 0; f < F; ++f) mu[f] /= (double)ns;
        for (int64_t i = 0; i < ns; ++i)
            for (int f = 0; f < F; ++f) {
                const double d = hx[(size_t)i * F + f] - mu[f];
                var[f] += d * d;
            }
        std::vector<float> fm(F), fs(F);
        for (int f = 0; f < F; ++f) {
            fm[f] = (float)mu[f];
            fs[f] = (float)std::max(std::sqrt(var[f] / (double)ns), 1e-3);
        }
        HIP_TRY(c, hipMemcpy(s->mean, fm.data(), F * 4, hipMemcpyHostToDevice));
        HIP_TRY(c, hipMemcpy(s->stdv, fs.data(), F * 4, hipMemcpyHostToDevice));
    }
    Rng rng{o->seed * 0x2545F4914F6CDD1Dull + 1};
    std::vector<int64_t> order((size_t)n);
    for (int64_t i = 0; i < n; ++i) order[i] = i;
    double first = -1.0, last = 0.0;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int ep = 0; ep < o->epochs; ++ep) {
        for (int64_t i = n - 1; i > 0; --i) std::swap(order[i], order[(int64_t)(rng.uni() * (i + 1))]);
        double ep_loss = 0.0;
        int64_t nb = 0;
        for (int64_t a = 0; a + o->batch <= n || (a == 0 && n < o->batch); a += o->batch) {
            const int64_t b = std::min<int64_t>(o->batch, n - a);
            HIP_TRY(c, hipMemcpyAsync(s->idx, order.data() + a, (size_t)b * 8, hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL(k_gather_rows, grid_for(b * F, 256), 256, 0, st, dx.p, dm.p, dy.p, s->idx, b, s->Xb,
                               s->mb, s->yb);
            hipLaunchKernelGGL(k_normalize, grid_for(b * F, 256), 256, 0, st, s->Xb, s->mean, s->stdv, b, s->Xn);
            forward(c, b);
            double l = 0.0;
            rc = backward(c, b, &l);
            if (rc) return rc;
            ++s->adam_t;
            const float c1 = 1.f - std::pow(b1, (float)s->adam_t), c2 = 1.f - std::pow(b2, (float)s->adam_t);
            hipLaunchKernelGGL(k_adam, grid_for(HCI_MLP_NPARAM, 256), 256, 0, st, s->P, s->G, s->A1, s->A2,
                               HCI_MLP_NPARAM, o->lr, b1, b2, eps, c1, c2);
            ep_loss += l;
            ++nb;
        }
        last = ep_loss / std::max<int64_t>(nb, 1);
        if (first < 0) first = last;
    }
    HIP_TRY(c, hipStreamSynchronize(st));
    s->ready = true;
    if (stats) {
        stats->first_epoch_loss = first;
        stats->last_epoch_loss = last;
        stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    return HCI_OK;
}

}  // extern "C"
