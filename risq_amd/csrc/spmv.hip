// Symmetric sigma product y = H x.
//
// Replaces SparseMatUpperTri::matrix_vector_prod (reference src/var/sparse.rs:64-78):
//   y_i = d_i x_i + sum_{j>i} H_ij x_j + sum_{j<i} H_ji x_j.
// Vectors are in the matrix's internal (alpha-major) order.  Work split: one
// wavefront per chunk of kSpmvChunk stored entries, balanced regardless of the
// skewed row lengths of an HCI Hamiltonian; 64 consecutive entries per step,
// each lane finds its row by binary search inside the chunk's row range, and
// lanes of one row are summed by a segmented shuffle reduction.
//   FULL storage: rows are complete, so no atomics: row-aligned work items
//     (k_spmv_items) give every row one owner.  Results are bitwise
//     reproducible.
//   UPPER storage: the transpose term H_pq x_p goes to y_q with a hardware FP64
//     atomic (order-dependent rounding).
//   TILED storage: k_spmv_win, x slices staged in LDS, rows summed in LDS.
// y starts as d o x in both cases.
#include <type_traits>

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
// The matrix is streamed once per product: non-temporal loads keep it from
// evicting the x vector out of L2.
template <int MODE>
__device__ __forceinline__ double load_val(const double* val, const float* hi, const float* lo,
                                           int64_t e) {
    if (MODE == 0) return __builtin_nontemporal_load(val + e);
    if (MODE == 1) return (double)__builtin_nontemporal_load(hi + e);
    return (double)__builtin_nontemporal_load(hi + e) + (double)__builtin_nontemporal_load(lo + e);
}

__device__ __forceinline__ int32_t load_col(const int32_t* col, int64_t e) {
    return __builtin_nontemporal_load(col + e);
}

// FULL storage: one wavefront per row-aligned work item.  A long row streams
// through the wave with per-lane partial sums (four entries in flight per
// lane) and one reduction at the end; a pack of short rows is processed 64
// entries per step with a segmented reduction, a row that continues into the
// next step being carried in registers.  Every row has exactly one owner:
// y_p = d_p x_p + sum_q H_pq x_q is written once, with no atomics.
template <int MODE, bool ABS>
__global__ void __launch_bounds__(256)
    k_spmv_items(int64_t n_items, const int32_t* __restrict__ item_row, const int64_t* __restrict__ rp,
                 const int32_t* __restrict__ col, const double* __restrict__ val,
                 const float* __restrict__ hi, const float* __restrict__ lo,
                 const double* __restrict__ diag, const double* __restrict__ x, double* __restrict__ y) {
    const int lane = threadIdx.x & 63;
    const int64_t k = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (k >= n_items) return;
    const int32_t ra = item_row[k], rb = item_row[k + 1];
    const int64_t e0 = rp[ra], e1 = rp[rb];
    if (rb - ra == 1 && e1 - e0 > 64) {
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
        int64_t e = e0 + lane;
        for (; e + 192 < e1; e += 256) {
            const double v0 = load_val<MODE>(val, hi, lo, e);
            const double v1 = load_val<MODE>(val, hi, lo, e + 64);
            const double v2 = load_val<MODE>(val, hi, lo, e + 128);
            const double v3 = load_val<MODE>(val, hi, lo, e + 192);
            if (ABS) {
                s0 += fabs(v0);
                s1 += fabs(v1);
                s2 += fabs(v2);
                s3 += fabs(v3);
            } else {
                const int32_t c0 = load_col(col, e), c1 = load_col(col, e + 64), c2 = load_col(col, e + 128),
                              c3 = load_col(col, e + 192);
                s0 += v0 * x[c0];
                s1 += v1 * x[c1];
                s2 += v2 * x[c2];
                s3 += v3 * x[c3];
            }
        }
        for (; e < e1; e += 64) {
            const double v0 = load_val<MODE>(val, hi, lo, e);
            s0 += ABS ? fabs(v0) : v0 * x[load_col(col, e)];
        }
        double s = (s0 + s1) + (s2 + s3);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o);
        if (lane == 0) y[ra] = ABS ? s : diag[ra] * x[ra] + s;
        return;
    }
    // rows of the pack with no entries
    for (int32_t r = ra + lane; r < rb; r += 64)
        if (rp[r] == rp[r + 1]) y[r] = ABS ? 0.0 : diag[r] * x[r];
    int32_t r_lo = ra;
    int32_t open_row = -1;  // row continuing from the previous step (wave-uniform)
    double open_sum = 0.0;
    for (int64_t eb = e0; eb < e1; eb += 64) {
        const int64_t e = eb + lane;
        const bool valid = e < e1;
        int32_t r = -1;
        double s = 0.0;
        if (valid) {
            int32_t a = r_lo, b = rb - 1;
            while (a < b) {
                const int32_t mid = (a + b + 1) >> 1;
                if (rp[mid] <= e)
                    a = mid;
                else
                    b = mid - 1;
            }
            r = a;
            const double v = load_val<MODE>(val, hi, lo, e);
            s = ABS ? fabs(v) : v * x[load_col(col, e)];
        }
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const double t = __shfl_down(s, o);
            const int32_t rr = __shfl_down(r, o);
            if (lane + o < 64 && rr == r) s += t;
        }
        const int32_t rprev = __shfl_up(r, 1);
        const bool head = valid && (lane == 0 || rprev != r);
        if (lane == 0 && r == open_row) s += open_sum;
        const int last_valid = (int)min((int64_t)63, e1 - 1 - eb);
        const int32_t r_last = __shfl(r, last_valid);
        const bool last_open = rp[r_last + 1] > min(eb + 64, e1);
        const uint64_t heads = __ballot(head);
        const uint64_t upto = last_valid == 63 ? ~0ull : ((2ull << last_valid) - 1);
        const int head_last = 63 - __clzll(heads & upto);
        if (head && !(r == r_last && last_open)) y[r] = ABS ? s : diag[r] * x[r] + s;
        const double s_last = __shfl(s, head_last);
        open_row = last_open ? r_last : -1;
        open_sum = last_open ? s_last : 0.0;
        r_lo = r_last;
    }
}

template <int MODE>
__global__ void __launch_bounds__(256)
    k_spmv_upper(int64_t nnz, int64_t n_chunks, const int64_t* __restrict__ rp,
                 const int32_t* __restrict__ chunk_row, const int32_t* __restrict__ col,
                 const double* __restrict__ val, const float* __restrict__ hi,
                 const float* __restrict__ lo, const double* __restrict__ x, double* y, bool absval) {
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
            const int32_t j = col[e];
            if (absval) {
                v = fabs(v);
                unsafeAtomicAdd(&y[j], v);
                s = v;
            } else {
                unsafeAtomicAdd(&y[j], v * x[r]);
                s = v * x[j];
            }
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

// TILED storage: one workgroup per 256-row tile, one wavefront per window at
// a time (windows handed out through an LDS counter).  The wave stages its
// window's x slice (<= 256 values) in its own LDS region, then streams the
// window's entries 8 per lane (16-byte index and value vectors, two vectors
// in flight per lane) and adds H_pq x_q to the tile's LDS row accumulator with
// hardware FP64 LDS atomics.  In each of a lane's 8 steps the wave handles 64
// consecutive entries of the window (column-major for kinds 0-2), so the x
// reads mostly broadcast and the row atomics mostly hit distinct rows.  The tile then writes its rows once:
// y_p = d_p x_p + acc_p (alpha-major part) or acc_p (beta-major part).
typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

// 8 consecutive entries of one lane, values kept in their stored width until use
template <int MODE>
struct Oct {
    u16x8 u;
    f64x2 d[MODE == 0 ? 4 : 1];
    f32x4 h[MODE == 0 || MODE == 3 ? 1 : 2];
    f32x4 l[MODE == 2 ? 2 : 1];
    u16x8 q;  // MODE 3: FP16 values
    // entries base + off .. base + off + 7: base is wave-uniform (scalar
    // address arithmetic), off a 32-bit per-lane offset
    __device__ __forceinline__ void load(const uint16_t* idx, const double* val, const float* hi, const float* lo,
                                         int64_t base, uint32_t off) {
        // MODE 3 passes the FP16 copies through val: their base in 2-byte units
        const double* v = MODE == 3 ? (const double*)((const uint16_t*)val + base) : val + base;
        load(idx + base, v, hi + base, lo + base, off);
    }
    __device__ __forceinline__ void load(const uint16_t* idx, const double* val, const float* hi, const float* lo,
                                         uint32_t e) {
        u = __builtin_nontemporal_load((const u16x8*)(idx + e));
        if (MODE == 3) {
            // MODE 3 passes the FP16 copies in place of the FP64 values
            q = __builtin_nontemporal_load((const u16x8*)((const uint16_t*)val + e));
        } else if (MODE == 0) {
#pragma unroll
            for (int k = 0; k < 4; ++k) d[k] = __builtin_nontemporal_load((const f64x2*)(val + e + 2 * k));
        } else {
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                h[k] = __builtin_nontemporal_load((const f32x4*)(hi + e + 4 * k));
                if (MODE == 2) l[k] = __builtin_nontemporal_load((const f32x4*)(lo + e + 4 * k));
            }
        }
    }
    // MODE 3 in FP32 (its accumulator type), the others in FP64
    __device__ __forceinline__ typename std::conditional<MODE == 3, float, double>::type value(int k) const {
        if (MODE == 3) return (float)__builtin_bit_cast(_Float16, (uint16_t)q[k]);
        if (MODE == 0) return d[k >> 1][k & 1];
        if (MODE == 1) return (double)h[k >> 2][k & 3];
        return (double)h[k >> 2][k & 3] + (double)l[k >> 2][k & 3];
    }
};

// x as the product reads it: FP64, or FP32 with FP16 values (MODE 3), whose
// own rounding is 2^-11
template <int MODE>
using XT = typename std::conditional<MODE == 3, float, double>::type;
// the tile's row accumulators: FP64 for every value precision (ds_add_f64
// runs the FP16-value product at 3.3 ms on the 2e6-determinant workload,
// ds_add_f32 at 21 ms)
template <int MODE>
using AT = double;

// H_pq x_q for value k of a group: FP16 values times FP32 x in FP32 (exact
// to 2^-24, far below the values' own 2^-11), the others in FP64
template <int MODE, class X>
__device__ __forceinline__ double term(const Oct<MODE>& o, int k, X xv) {
    if constexpr (MODE == 3)
        return (double)(o.value(k) * xv);
    else
        return o.value(k) * (double)xv;
}

// One lane's group of 8 entries (those with bit q of `valid`): H_pq x_q into
// row p's accumulator and, for the stored half of a mirror pair (PR), H_pq x_p
// into column q's transposed sum.  The x reads of a batch of entries are
// issued before its first product, so a wave keeps several LDS reads in
// flight instead of waiting on each one.
template <bool ABS, bool PR, int MODE, class X>
__device__ __forceinline__ void add_group(double* acc, const X* xw, const X* xr, double* tw, const Oct<MODE>& o,
                                          uint32_t valid) {
    if (ABS) {
#pragma unroll
        for (int q = 0; q < 8; ++q)
            if (valid >> q & 1u) {
                const double a = fabs((double)o.value(q));
                atomicAdd(&acc[o.u[q] >> 8], a);
                if (PR) atomicAdd(&tw[o.u[q] & 0xFFu], a);
            }
        return;
    }
    // batches of B reads: 8 with FP32 x and no transposed sums, 4 otherwise
    // (the registers they hold keep the occupancy up)
    constexpr int B = (MODE == 3 && !PR) ? 8 : 4;
#pragma unroll
    for (int h = 0; h < 8; h += B) {
        X xc[B], xp[PR ? B : 1];
#pragma unroll
        for (int q = 0; q < B; ++q) {
            xc[q] = xw[o.u[h + q] & 0xFFu];
            if (PR) xp[q] = xr[o.u[h + q] >> 8];
        }
#pragma unroll
        for (int q = 0; q < B; ++q)
            if (valid >> (h + q) & 1u) {
                atomicAdd(&acc[o.u[h + q] >> 8], term(o, h + q, xc[q]));
                if (PR) atomicAdd(&tw[o.u[h + q] & 0xFFu], term(o, h + q, xp[PR ? q : 0]));
            }
    }
}

// PAIRS: the matrix has mirror pairs (their stored halves' code and LDS are
// left out of the other instantiations, which keep their occupancy)
template <int MODE, bool ABS, bool PAIRS>
__global__ void __launch_bounds__(kTileThreads)
    k_spmv_win(const TileTask* __restrict__ tasks, const Window* __restrict__ wins,
               const uint16_t* __restrict__ idx, const double* __restrict__ val, const float* __restrict__ hi,
               const float* __restrict__ lo, const double* __restrict__ x, const double* __restrict__ diag,
               double* __restrict__ y, const double* __restrict__ xdiag, double* __restrict__ yb_out,
               const double* __restrict__ yb_in, const int32_t* __restrict__ binv, double* __restrict__ part) {
    __shared__ AT<MODE> acc[kTileRows];
    __shared__ XT<MODE> xs[kTileWaves][kWinCols];
    __shared__ XT<MODE> xr[PAIRS ? kTileRows : 1];           // x of the tile's own rows (mirror pairs)
    __shared__ AT<MODE> ts[kTileWaves][PAIRS ? kWinCols : 1];  // a wave's transposed sums (mirror pairs)
    __shared__ int next, next_small;
    const TileTask t = tasks[blockIdx.x];
    const int tid = threadIdx.x, lane = tid & 63;
    XT<MODE>* xw = xs[tid >> 6];
    const XT<MODE>* xx = (const XT<MODE>*)x;
    for (int r = tid; r < kTileRows; r += kTileThreads) acc[r] = 0;
    if (PAIRS && !ABS)
        for (int r = tid; r < t.nr; r += kTileThreads) xr[r] = xx[t.r0 + r];
    if (tid == 0) {
        next = 0;
        next_small = 0;
    }
    __syncthreads();
    // the tile's epilogue terms (diagonal term and the beta-major part's sum of
    // the same row) are requested now and wait in registers, so their loads,
    // one of them a dependent gather, overlap the windows: one row per thread
    static_assert(kTileRows == kTileThreads, "one epilogue row per thread");
    double pre_d = 0.0, pre_b = 0.0;
    if (!yb_out && tid < t.nr) {
        const int64_t p = (int64_t)t.r0 + tid;
        if (diag && !ABS) pre_d = diag[p] * xdiag[p];  // xdiag: x in FP64
        if (yb_in) pre_b = yb_in[binv[p]];
    }
    const int n_w = t.w_split - t.w_begin;
    auto grab = [&]() {
        int k = 0;
        if (lane == 0) k = atomicAdd(&next, 1);
        return __builtin_amdgcn_readfirstlane(k);
    };
    int k = grab();
    Window W{};
    if (k < n_w) W = wins[t.w_begin + k];
    while (k < n_w) {
        // the next window's descriptor is requested before this window's data
        const int kn = grab();
        Window Wn{};
        if (kn < n_w) Wn = wins[t.w_begin + kn];
        // full chunks: lane l's group of chunk c holds entries 512 c + 64 q + l;
        // the tail of m entries: group g holds entries g, g + R, ..., g + 7R
        const int cnt = W.cnt(), nc = W.nc();
        const int n_full = cnt / kWinChunk;
        const int m = cnt - n_full * kWinChunk;
        const int R = (m + kWinAlign - 1) / kWinAlign;
        const int64_t tail = W.o() + (int64_t)n_full * kWinChunk;
        const uint32_t mine = kWinAlign * lane;
        const int32_t sp = PAIRS ? W.sp() : -1;
        // the first two groups' entries and the x slice are requested together
        Oct<MODE> o0{}, o1{};
        if (n_full > 0)
            o0.load(idx, val, hi, lo, W.o(), mine);
        else if (lane < R)
            o0.load(idx, val, hi, lo, tail, mine);
        if (n_full > 1) o1.load(idx, val, hi, lo, W.o() + kWinChunk, mine);
        if (!ABS) {
            XT<MODE> xv[kWinCols / 64];
#pragma unroll
            for (int j = 0; j < kWinCols / 64; ++j) {
                const int c = lane + 64 * j;
                xv[j] = c < nc ? (xx + W.c0)[(uint32_t)c] : (XT<MODE>)0;
            }
#pragma unroll
            for (int j = 0; j < kWinCols / 64; ++j) xw[lane + 64 * j] = xv[j];
        }
        // the stored half of a mirror pair also sums its transposed products
        AT<MODE>* tw = ts[PAIRS ? tid >> 6 : 0];
        if (PAIRS && sp >= 0) {
#pragma unroll
            for (int j = 0; j < kWinCols / 64; ++j) tw[lane + 64 * j] = 0;
        }
        // the slice is written and read by this wave only: order its LDS
        // writes before the reads below
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // the stream of one window; PR: a stored mirror half (its entries are
        // sorted by diagonal, so the 64 entries of a step fall on distinct
        // rows and distinct columns, mostly, for both accumulators)
        auto stream = [&](auto pr_tag) {
            constexpr bool PR = decltype(pr_tag)::value;
            int c = 0;
            for (; c + 1 < n_full; c += 2) {
                add_group<ABS, PR>(acc, xw, xr, tw, o0, 0xFFu);
                add_group<ABS, PR>(acc, xw, xr, tw, o1, 0xFFu);
                if (c + 2 < n_full) o0.load(idx, val, hi, lo, W.o() + (int64_t)(c + 2) * kWinChunk, mine);
                if (c + 3 < n_full) o1.load(idx, val, hi, lo, W.o() + (int64_t)(c + 3) * kWinChunk, mine);
            }
            if (c < n_full) {  // one full chunk left, in o0 (odd count) ...
                add_group<ABS, PR>(acc, xw, xr, tw, o0, 0xFFu);
                if (lane < R) o0.load(idx, val, hi, lo, tail, mine);
            } else if (n_full > 0 && lane < R) {  // ... or none: the tail is still to load
                o0.load(idx, val, hi, lo, tail, mine);
            }
            if (lane < R) {
                uint32_t valid = 0;
#pragma unroll
                for (int q = 0; q < 8; ++q) valid |= (lane + q * R < m ? 1u : 0u) << q;
                add_group<ABS, PR>(acc, xw, xr, tw, o0, valid);
            }
        };
        if (PAIRS && sp >= 0) {
            stream(std::true_type{});
            // the transposed sums to the pair's slot, for the receiving tile
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
            for (int j = 0; j < kWinCols / 64; ++j) {
                const int cc = lane + 64 * j;
                if (cc < nc) part[(size_t)sp * kWinCols + cc] = (double)tw[cc];
            }
        } else {
            stream(std::false_type{});
        }
        // reads of this slice precede the next window's staging
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        W = Wn;
        k = kn;
    }
    // the tile's small windows (at most kSmallWin or kSmallWinDense entries,
    // chosen at build time), one per lane: x is read where an entry needs it
    // instead of staging the window's slice
    const int n_s = t.w_end - t.w_split;
    for (;;) {
        int kb = 0;
        if (lane == 0) kb = atomicAdd(&next_small, 64);
        kb = __shfl(kb, 0);
        if (kb >= n_s) break;
        if (kb + lane < n_s) {
            const Window Ws = wins[t.w_split + kb + lane];
            const int m = Ws.cnt();
            const int R = (m + kWinAlign - 1) / kWinAlign;
            for (int g = 0; g < R; ++g) {
                Oct<MODE> o;
                o.load(idx, val, hi, lo, Ws.o(), (uint32_t)(kWinAlign * g));
#pragma unroll
                for (int q = 0; q < 8; ++q)
                    if (g + q * R < m) {
                        const uint32_t u = o.u[q];
                        const AT<MODE> v = o.value(q);
                        if (ABS)
                            atomicAdd(&acc[u >> 8], fabs(v));
                        else
                            atomicAdd(&acc[u >> 8], v * (AT<MODE>)xx[Ws.c0 + (int)(u & 0xFFu)]);
                    }
            }
        }
    }
    __syncthreads();
    if (tid < t.nr) {
        const int64_t p = (int64_t)t.r0 + tid;
        if (yb_out)  // beta-major rows, in their own order
            yb_out[p] = (double)acc[tid];
        else  // alpha-major rows: d_p x_p + acc_p, then the beta-major sum
            y[p] = ((diag && !ABS) ? pre_d + (double)acc[tid] : (double)acc[tid]) + (yb_in ? pre_b : 0.0);
    }
}

// The receiving side of the mirror pairs: one workgroup per receiving tile
// adds, row by row, the transposed sums of the slots that target it (in slot
// order) into its part's rows: y for alpha-major tiles, the beta-major row
// sums (before the alpha-major launch reads them) for beta-major tiles.  The
// tiles of one launch have disjoint rows.
__global__ void __launch_bounds__(256)
    k_pair_reduce(const PairSeg* __restrict__ segs, const int32_t* __restrict__ slots,
                  const double* __restrict__ part, double* __restrict__ y) {
    const PairSeg g = segs[blockIdx.x];
    for (int r = threadIdx.x; r < g.nr; r += blockDim.x) {
        double sum = 0.0;
        for (int32_t k = g.s_begin; k < g.s_end; ++k) sum += part[(size_t)slots[k] * kWinCols + r];
        y[(int64_t)g.r0 + r] += sum;
    }
}

template <class T>
__global__ void k_gather_x(const int32_t* perm, const T* src, int64_t n, T* dst) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) dst[p] = src[perm[p]];
}

__global__ void k_to_f32(const double* x, int64_t n, float* out) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) out[p] = (float)x[p];
}

__global__ void k_gather(const int32_t* perm, const double* src, int64_t n, double* dst) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) dst[p] = src[perm[p]];
}

__global__ void k_scatter(const int32_t* perm, const double* src, int64_t n, double* dst) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) dst[perm[p]] = src[p];
}

}  // namespace

void permute_to_internal(const hci_csr* h, const double* x_caller, double* x_int) {
    hipLaunchKernelGGL(k_gather, grid_for(h->n, 256), 256, 0, h->ctx->stream, h->d_perm, x_caller, h->n, x_int);
}

void permute_to_caller(const hci_csr* h, const double* x_int, double* x_caller) {
    hipLaunchKernelGGL(k_scatter, grid_for(h->n, 256), 256, 0, h->ctx->stream, h->d_perm, x_int, h->n, x_caller);
}

int spmv_prepare(hci_csr* h) {
    hci_ctx* c = h->ctx;
    hipFree(h->d_chunk_row);
    hipFree(h->d_item_row);
    h->d_chunk_row = nullptr;
    h->d_item_row = nullptr;
    if (h->storage == HCI_STORE_TILED) return HCI_OK;
    if (h->storage == HCI_STORE_UPPER) {
        h->n_chunks = (h->nnz_s + kSpmvChunk - 1) / kSpmvChunk;
        HIP_TRY(c, hipMalloc(&h->d_chunk_row, (size_t)(h->n_chunks + 1) * 4 + 16));
        if (h->nnz_s > 0)
            hipLaunchKernelGGL(k_chunk_rows, grid_for(h->n_chunks + 1, 256), 256, 0, c->stream, h->d_row_ptr,
                               h->n, h->nnz_s, h->n_chunks, h->d_chunk_row);
        HIP_TRY(c, hipGetLastError());
        return HCI_OK;
    }
    // FULL: pack consecutive rows into items of <= kShortRows entries; a row
    // longer than that is an item of its own
    std::vector<int64_t> rp((size_t)h->n + 1);
    HIP_TRY(c, hipMemcpyAsync(rp.data(), h->d_row_ptr, rp.size() * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    std::vector<int32_t> items;
    items.reserve((size_t)(h->nnz_s / kShortRows + h->n / 64 + 2));
    int64_t r = 0;
    while (r < h->n) {
        items.push_back((int32_t)r);
        const int64_t start = rp[r];
        if (rp[r + 1] - start > kShortRows) {
            ++r;
            continue;
        }
        int64_t q = r + 1;
        while (q < h->n && rp[q + 1] - start <= kShortRows && q - r < 4096) ++q;
        r = q;
    }
    items.push_back((int32_t)h->n);
    h->n_items = (int64_t)items.size() - 1;
    HIP_TRY(c, hipMalloc(&h->d_item_row, items.size() * 4 + 16));
    HIP_TRY(c, hipMemcpy(h->d_item_row, items.data(), items.size() * 4, hipMemcpyHostToDevice));
    return HCI_OK;
}

namespace {
// the two part launches of a TILED product, with the mirror pairs' sums:
// beta-major tiles write their row sums in beta-major order, the beta-major
// receivers add theirs, then alpha-major tiles write y (each row with its
// beta-major sum, read through binv) and the alpha-major receivers add theirs
template <int MODE, bool PAIRS>
void launch_win_parts(const hci_csr* h, hipStream_t s, const double* vals, const double* xa, const double* xb,
                      const double* x, double* y, bool absval) {
    double* part = PAIRS ? h->d_pair_part : nullptr;
    const bool beta = h->n_task_b > 0;
    if (beta) {
        if (absval)
            hipLaunchKernelGGL((k_spmv_win<MODE, true, PAIRS>), (unsigned)h->n_task_b, kTileThreads, 0, s,
                               h->d_task_b, h->d_win, h->d_idx16, vals, h->d_hi, h->d_lo, nullptr, nullptr, y, nullptr,
                               h->d_yb, nullptr, nullptr, part);
        else
            hipLaunchKernelGGL((k_spmv_win<MODE, false, PAIRS>), (unsigned)h->n_task_b, kTileThreads, 0, s,
                               h->d_task_b, h->d_win, h->d_idx16, vals, h->d_hi, h->d_lo, xb, nullptr, y, nullptr,
                               h->d_yb, nullptr, nullptr, part);
        if (PAIRS && h->n_pair_seg > h->n_pair_seg0)
            hipLaunchKernelGGL(k_pair_reduce, (unsigned)(h->n_pair_seg - h->n_pair_seg0), 256, 0, s,
                               h->d_pair_seg + h->n_pair_seg0, h->d_pair_slot, h->d_pair_part, h->d_yb);
    }
    const double* yb = beta ? h->d_yb : nullptr;
    if (absval)
        hipLaunchKernelGGL((k_spmv_win<MODE, true, PAIRS>), (unsigned)h->n_task_a, kTileThreads, 0, s, h->d_task_a,
                           h->d_win, h->d_idx16, vals, h->d_hi, h->d_lo, nullptr, nullptr, y, nullptr, nullptr, yb,
                           h->d_binv, part);
    else
        hipLaunchKernelGGL((k_spmv_win<MODE, false, PAIRS>), (unsigned)h->n_task_a, kTileThreads, 0, s, h->d_task_a,
                           h->d_win, h->d_idx16, vals, h->d_hi, h->d_lo, xa, h->d_diag, y, x, nullptr, yb, h->d_binv,
                           part);
    if (PAIRS && h->n_pair_seg0 > 0)
        hipLaunchKernelGGL(k_pair_reduce, (unsigned)h->n_pair_seg0, 256, 0, s, h->d_pair_seg, h->d_pair_slot,
                           h->d_pair_part, y);
}

template <int MODE>
void launch_spmv(const hci_csr* h, const double* x, double* y, bool absval) {
    hipStream_t s = h->ctx->stream;
    if (h->storage == HCI_STORE_TILED) {
        // MODE 3 reads the FP16 copies through the value pointer, and FP32
        // copies of x (in both orders) through the x pointers
        const double* vals = MODE == 3 ? (const double*)h->d_h16 : h->d_val;
        const double* xa = x;
        double* xb = h->d_xb;
        if (MODE == 3 && !absval) {
            hipLaunchKernelGGL(k_to_f32, grid_for(h->n, 256), 256, 0, s, x, h->n, h->d_x32);
            xa = (const double*)h->d_x32;
            xb = (double*)h->d_xb32;
        }
        // x gathered to beta-major order for the beta-major part
        if (h->n_task_b > 0 && !absval)
            hipLaunchKernelGGL(k_gather_x<XT<MODE>>, grid_for(h->n, 256), 256, 0, s, h->d_bpos,
                               (const XT<MODE>*)xa, h->n, (XT<MODE>*)xb);
        // a matrix with stored mirror halves (never with FP16 copies) takes
        // the instantiation with their code
        if (MODE != 3 && h->n_pair > 0)
            launch_win_parts<MODE, MODE != 3>(h, s, vals, xa, xb, x, y, absval);
        else
            launch_win_parts<MODE, false>(h, s, vals, xa, xb, x, y, absval);
        return;
    }
    if constexpr (MODE == 3) return;  // FP16 copies exist in tiled storage only
    if (h->storage == HCI_STORE_FULL) {
        const unsigned grid = grid_for(h->n_items, 4);
        if (absval)
            hipLaunchKernelGGL((k_spmv_items<MODE, true>), grid, 256, 0, s, h->n_items, h->d_item_row, h->d_row_ptr,
                               h->d_col, h->d_val, h->d_hi, h->d_lo, h->d_diag, x, y);
        else
            hipLaunchKernelGGL((k_spmv_items<MODE, false>), grid, 256, 0, s, h->n_items, h->d_item_row,
                               h->d_row_ptr, h->d_col, h->d_val, h->d_hi, h->d_lo, h->d_diag, x, y);
    } else {
        if (!absval)
            hipLaunchKernelGGL(k_diag_mul, grid_for(h->n, 256), 256, 0, s, h->d_diag, x, y, h->n);
        if (h->nnz_s > 0)
            hipLaunchKernelGGL((k_spmv_upper<MODE>), grid_for(h->n_chunks, 4), 256, 0, s, h->nnz_s, h->n_chunks,
                               h->d_row_ptr, h->d_chunk_row, h->d_col, h->d_val, h->d_hi, h->d_lo, x, y, absval);
    }
}
}  // namespace

int spmv_device(const hci_csr* h, const double* x, double* y, int mode) {
    hci_ctx* c = h->ctx;
    if (mode == HCI_SIGMA_F16 && !h->d_h16) return hci_fail(c, HCI_ERR_ARG, "sigma: FP16 mode needs HCI_PREC_F16 storage");
    if (h->prec == HCI_PREC_F64)
        launch_spmv<0>(h, x, y, false);
    else if (mode == HCI_SIGMA_F16)
        launch_spmv<3>(h, x, y, false);
    else if (mode == HCI_SIGMA_F32)
        launch_spmv<1>(h, x, y, false);
    else
        launch_spmv<2>(h, x, y, false);
    HIP_TRY(c, hipGetLastError());
    return HCI_OK;
}

int csr_rownorm(hci_csr* h) {
    if (h->d_rownorm) return HCI_OK;
    hci_ctx* c = h->ctx;
    HIP_TRY(c, hipMalloc(&h->d_rownorm, (size_t)h->n * 8 + 16));
    HIP_TRY(c, hipMemsetAsync(h->d_rownorm, 0, (size_t)h->n * 8, c->stream));
    // a feature of the learned correction (read as FP32): the FP16 copies,
    // where they exist, are accurate enough and half the bytes
    if (h->d_h16)
        launch_spmv<3>(h, nullptr, h->d_rownorm, true);
    else if (h->prec == HCI_PREC_F64)
        launch_spmv<0>(h, nullptr, h->d_rownorm, true);
    else
        launch_spmv<2>(h, nullptr, h->d_rownorm, true);
    HIP_TRY(c, hipGetLastError());
    return HCI_OK;
}

extern "C" int hci_sigma(const hci_csr* h, int ncol, const double* x, double* y, int sigma_mode,
                         int mem) {
    if (!h) return HCI_ERR_ARG;
    hci_ctx* c = h->ctx;
    if (ncol < 0 || (ncol > 0 && (!x || !y))) return hci_fail(c, HCI_ERR_ARG, "hci_sigma: bad arguments");
    if (sigma_mode != HCI_SIGMA_FULL && sigma_mode != HCI_SIGMA_F32 && sigma_mode != HCI_SIGMA_F16)
        return hci_fail(c, HCI_ERR_ARG, "hci_sigma: bad sigma_mode");
    if (sigma_mode == HCI_SIGMA_F16 && !h->d_h16)
        return hci_fail(c, HCI_ERR_ARG, "hci_sigma: HCI_SIGMA_F16 needs HCI_PREC_F16 storage");
    HIP_TRY(c, hipSetDevice(c->device));
    const size_t bytes = (size_t)h->n * 8;
    // four vectors of the context's work block (kept between calls)
    const size_t vb = (bytes + 255) / 256 * 256;
    if (int rc0 = work_reserve(c, 4 * vb, 256)) return rc0;
    double* xi = (double*)c->work_dev;
    double* yi = (double*)((char*)c->work_dev + vb);
    double* dx = (double*)((char*)c->work_dev + 2 * vb);
    double* dy = (double*)((char*)c->work_dev + 3 * vb);
    int rc = HCI_OK;
    auto cleanup = [&]() { hipStreamSynchronize(c->stream); };
    for (int k = 0; k < ncol && rc == HCI_OK; ++k) {
        const double* xk = x + (size_t)k * h->n;
        double* yk = y + (size_t)k * h->n;
        if (mem == HCI_MEM_HOST) {
            if (hipMemcpyAsync(dx, xk, bytes, hipMemcpyHostToDevice, c->stream) != hipSuccess) {
                rc = hci_fail(c, HCI_ERR_HIP, "hci_sigma: copy in");
                break;
            }
            xk = dx;
        }
        permute_to_internal(h, xk, xi);
        rc = spmv_device(h, xi, yi, sigma_mode);
        if (rc) break;
        permute_to_caller(h, yi, mem == HCI_MEM_HOST ? dy : yk);
        if (mem == HCI_MEM_HOST &&
            hipMemcpyAsync(yk, dy, bytes, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
            rc = hci_fail(c, HCI_ERR_HIP, "hci_sigma: copy out");
    }
    cleanup();
    return rc;
}

// ---- allocation-free sigma -------------------------------------------------
// hci_sigma_noalloc keeps its two internal-order vectors in this file-scope
// array instead of the context's work block: callers that may not allocate
// (graph capture, serving loops) get a product that calls no allocator.  In
// host code `scratch` names a host-side stand-in: its device address comes
// from hipGetSymbolAddress, once per context (each GPU holds its own copy).
namespace {
constexpr int64_t kNoAllocMaxDets = (int64_t)1 << 21;  // two FP64 vectors: 32 MiB
__device__ __align__(256) double scratch[2 * kNoAllocMaxDets];
}  // namespace

int sigma_scratch_init(hci_ctx* c) {
    void* p = nullptr;
    HIP_TRY(c, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    c->scratch = (double*)p;
    return HCI_OK;
}

extern "C" int hci_sigma_noalloc(const hci_csr* h, int ncol, const double* x, double* y, int sigma_mode) {
    if (!h) return HCI_ERR_ARG;
    hci_ctx* c = h->ctx;
    if (ncol < 0 || (ncol > 0 && (!x || !y))) return hci_fail(c, HCI_ERR_ARG, "hci_sigma_noalloc: bad arguments");
    if (sigma_mode != HCI_SIGMA_FULL && sigma_mode != HCI_SIGMA_F32 && sigma_mode != HCI_SIGMA_F16)
        return hci_fail(c, HCI_ERR_ARG, "hci_sigma_noalloc: bad sigma_mode");
    if (sigma_mode == HCI_SIGMA_F16 && !h->d_h16)
        return hci_fail(c, HCI_ERR_ARG, "hci_sigma_noalloc: HCI_SIGMA_F16 needs HCI_PREC_F16 storage");
    if (h->n > kNoAllocMaxDets)
        return hci_fail(c, HCI_ERR_ARG, "hci_sigma_noalloc: the matrix has more than 2097152 determinants "
                                        "(the size of the library's fixed scratch); use hci_sigma");
    if (!c->scratch) return hci_fail(c, HCI_ERR_STATE, "hci_sigma_noalloc: scratch address not initialised");
    HIP_TRY(c, hipSetDevice(c->device));
    double* xi = c->scratch;
    double* yi = c->scratch + kNoAllocMaxDets;
    for (int k = 0; k < ncol; ++k) {
        permute_to_internal(h, x + (size_t)k * h->n, xi);
        if (int rc = spmv_device(h, xi, yi, sigma_mode)) return rc;
        permute_to_caller(h, yi, y + (size_t)k * h->n);
    }
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return HCI_OK;
}

// MatrixOperations::set_diagonal (src/var/eigenvalues/matrix_operations.rs:14-27):
// the caller-order diagonal replaces the matrix's
extern "C" int hci_csr_set_diagonal(hci_csr* h, const double* diag, int mem) {
    if (!h) return HCI_ERR_ARG;
    hci_ctx* c = h->ctx;
    if (!diag) return hci_fail(c, HCI_ERR_ARG, "hci_csr_set_diagonal: null diagonal");
    HIP_TRY(c, hipSetDevice(c->device));
    const double* src = diag;
    if (mem == HCI_MEM_HOST) {
        if (int rc = work_reserve(c, (size_t)h->n * 8 + 256, 256)) return rc;
        HIP_TRY(c, hipMemcpyAsync(c->work_dev, diag, (size_t)h->n * 8, hipMemcpyHostToDevice, c->stream));
        src = (const double*)c->work_dev;
    }
    permute_to_internal(h, src, h->d_diag);
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return HCI_OK;
}
