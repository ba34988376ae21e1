// Sparse variational Hamiltonian build on the GPU.
//
// Replaces gen_sparse_ham_fast_lookup (reference src/var/fast_ham.rs:527-575).
// The reference groups determinants by alpha string in hash maps and then runs
// three searches: an O(N_new * N) singles scan (:473-517), opposite-spin
// doubles over alpha-single-connected string pairs (:369-407) and same-spin
// doubles (:410-470).  Here the same pair set is found with sorted buckets:
//
//   1. rank alpha and beta strings (radix sort + head flags + scan), then sort
//      determinants by (alpha rank, beta rank) and by (beta rank, alpha rank):
//      every alpha group and every beta group is a contiguous range;
//   2. alpha strings one electron apart share exactly one "remove-one" key; a
//      sort of (key, group) lists every single-connected alpha-group pair;
//   3. one wavefront per determinant i scans three candidate ranges, 64
//      candidates per step, and keeps the pairs with j > i:
//        own alpha group : beta popcount(XOR) 2 -> beta single, 4 -> beta double
//        own beta group  : alpha popcount(XOR) 2 -> alpha single, 4 -> alpha double
//        single-connected alpha groups : beta popcount(XOR) 2 -> opposite-spin double
//      The three ranges are disjoint, so no pair is seen twice by one row.
//      Kept pairs are compacted with a wave ballot + popcount prefix.
//   4. a count pass sizes every row, a scan gives row_ptr, and a fill pass
//      writes the same sequence straight into the final arrays.
// Rows and columns are stored in the internal alpha-major order (the order of
// step 1), so a row's columns come in a few contiguous runs (its own alpha
// group, its neighbour groups) and sigma products gather x from few cache
// lines.  FULL storage keeps both triangles (every pair in both rows), UPPER
// keeps the pairs with p < q.  Elements are always evaluated with the
// lower caller index first, as the reference does.
//
// Storage rules are the reference's: singles iff |H| > 1e-12 (fast_ham.rs:505,510)
// or H != 0 for the n <= 100 path, doubles iff H != 0 (off_diag.rs:33).
#include <cstring>

#include <rocprim/rocprim.hpp>

#include <algorithm>
#include <chrono>

#include "hci_internal.h"

namespace {

constexpr int kWaves = 4;
constexpr int kBlock = 64 * kWaves;

template <class S>
struct BuildView {
    int32_t N;
    const int32_t* ga_of;   // alpha group of each determinant (caller order)
    const int32_t* gb_of;   // beta group of each determinant
    const S* ua;            // unique alpha strings, ascending
    const int32_t* ga_off;  // alpha group g = alpha-sorted positions [ga_off[g], ga_off[g+1])
    const S* sa_dn;         // beta string at each alpha-sorted (= internal) position
    const int32_t* sa_idx;  // caller index at each internal position
    const int32_t* gb_off;  // beta group ranges in beta-sorted order
    const S* sb_up;         // alpha string at each beta-sorted position
    const int32_t* sb_idx;  // caller index at each beta-sorted position
    const int32_t* sb_pos;  // internal position at each beta-sorted position
    const int32_t* nbr_off;  // single-connected alpha groups of group g: nbr[nbr_off[g]..nbr_off[g+1])
    const int32_t* nbr;      // ascending within each group
    // tiled builds: for every internal position q, its beta string with one
    // electron removed (n_beta keys), sorted by key within each alpha group;
    // group g owns [ga_off[g]*n_beta, ga_off[g+1]*n_beta)
    const S* bkey;
    const int32_t* bkq;
    int n_beta;
    // the same for alpha strings at beta-sorted positions, sorted within each
    // beta group (built only when every double vanishes)
    const S* akey;
    const int32_t* akq;
    int n_alpha;
};

__device__ __forceinline__ bool keep_single(double v, int rule) {
    return rule == HCI_SINGLES_NONZERO ? (v != 0.0) : (fabs(v) > 1e-12);
}

// TILED builds place pairs by pattern; a single the storage rule rejects is
// written as an explicit zero, which the product adds harmlessly and the
// download skips (no stored element is ever exactly zero).  Singles whose
// element vanishes for every occupation (DevInts::sok) are not placed at all
__device__ __forceinline__ double single_or_zero(double v, int rule) { return keep_single(v, rule) ? v : 0.0; }

// the rule of a pair whose larger caller index is m
__device__ __forceinline__ int rule_of(const SingleRules& R, int32_t m) {
    for (int k = 0; k + 1 < R.n; ++k)
        if (m < R.end[k]) return R.rule[k];
    return R.rule[R.n - 1];
}

template <class S>
__device__ __forceinline__ void swap_if(bool c, S& a, S& b) {
    const S t = a;
    a = c ? b : a;
    b = c ? t : b;
}

// Output of the fill pass: FP64 values, or FP32 hi + lo words.
struct FillOut {
    int32_t* col;
    double* val;
    float* hi;
    float* lo;
};

__device__ __forceinline__ void put(const FillOut& o, int64_t e, int32_t q, double v) {
    o.col[e] = q;
    if (o.val) {
        o.val[e] = v;
    } else {
        const float h = (float)v;
        o.hi[e] = h;
        o.lo[e] = (float)(v - (double)h);
    }
}

// TILED fill target: the entries of window w go to woff[w] + position, in
// the order of the enumeration; k_interleave then moves them to their slots
struct TileOut {
    uint16_t* idx;
    double* val;
    float* hi;
    float* lo;
    const int64_t* woff;
};

// streaming stores: the entries are not read again by this kernel, and
// keeping them out of the cache leaves it to the integrals and key tables
__device__ __forceinline__ void put16(const TileOut& o, int64_t e, uint32_t u, double v) {
    __builtin_nontemporal_store((uint16_t)u, o.idx + e);
    if (o.val) {
        __builtin_nontemporal_store(v, o.val + e);
    } else {
        const float h = (float)v;
        __builtin_nontemporal_store(h, o.hi + e);
        __builtin_nontemporal_store((float)(v - (double)h), o.lo + e);
    }
}

// One wavefront per internal row p (caller determinant i = sa_idx[p]).
// A candidate q (caller j) is kept iff its pair is connected, its element passes
// the storage rule, max(i, j) >= n_stored, and (UPPER) p < q.
// FILL == false: row_cnt[p] = kept + prev_cnt[p];  FILL == true: writes them at
// row_ptr[p] + prev_cnt[p] onwards.
template <class S, bool FILL, bool FULL>
__global__ void __launch_bounds__(kBlock)
    k_connect(BuildView<S> B, DevInts T, int32_t n_stored, int single_rule, int64_t* row_cnt,
              const int64_t* row_ptr, const int64_t* prev_cnt, FillOut out) {
    const int lane = threadIdx.x & 63;
    const int64_t p64 = (int64_t)blockIdx.x * kWaves + (threadIdx.x >> 6);
    if (p64 >= B.N) return;  // wave-uniform
    const int32_t p = (int32_t)p64;
    const uint64_t lt_mask = lane == 0 ? 0ull : (~0ull >> (64 - lane));

    const int32_t i = B.sa_idx[p];
    const int32_t g = B.ga_of[i];
    const int32_t h = B.gb_of[i];
    const S up_i = B.ua[g];
    const S dn_i = B.sa_dn[p];
    const int64_t pc = prev_cnt ? prev_cnt[p] : 0;
    const int64_t base = FILL ? row_ptr[p] + pc : 0;
    int64_t cnt = 0;

#define HCI_EMIT(keep, q, v)                                         \
    do {                                                             \
        const uint64_t _m = __ballot(keep);                          \
        if (FILL && (keep)) put(out, base + cnt + __popcll(_m & lt_mask), (q), (v)); \
        cnt += __popcll(_m);                                         \
    } while (0)

    // (1) same alpha string: beta singles and beta doubles
    {
        const int32_t q0 = B.ga_off[g], q1 = B.ga_off[g + 1];
        for (int32_t qb = q0; qb < q1; qb += 64) {
            const int32_t q = qb + lane;
            bool keep = false;
            double v = 0.0;
            if (q < q1 && (FULL || q > p)) {
                const int32_t j = B.sa_idx[q];
                if (max(i, j) >= n_stored) {
                    S a = dn_i, b = B.sa_dn[q];
                    const int c = popc(a ^ b);
                    swap_if(j < i, a, b);
                    if (c == 2) {
                        v = elem_single<S>(T, a, b, up_i);
                        keep = keep_single(v, single_rule);
                    } else if (c == 4) {
                        v = elem_same_double<S>(T, a, b);
                        keep = (v != 0.0);
                    }
                }
            }
            HCI_EMIT(keep, q, v);
        }
    }
    // (2) same beta string: alpha singles and alpha doubles
    {
        const int32_t q0 = B.gb_off[h], q1 = B.gb_off[h + 1];
        for (int32_t qb = q0; qb < q1; qb += 64) {
            const int32_t qq = qb + lane;
            bool keep = false;
            double v = 0.0;
            int32_t q = 0;
            if (qq < q1) {
                q = B.sb_pos[qq];
                const int32_t j = B.sb_idx[qq];
                if ((FULL || q > p) && max(i, j) >= n_stored) {
                    S a = up_i, b = B.sb_up[qq];
                    const int c = popc(a ^ b);
                    swap_if(j < i, a, b);
                    if (c == 2) {
                        v = elem_single<S>(T, a, b, dn_i);
                        keep = keep_single(v, single_rule);
                    } else if (c == 4) {
                        v = elem_same_double<S>(T, a, b);
                        keep = (v != 0.0);
                    }
                }
            }
            HCI_EMIT(keep, q, v);
        }
    }
    // (3) alpha strings one electron away: opposite-spin doubles.  The ranges of
    // up to 64 neighbour groups are flattened; each lane finds its group by a
    // binary search over the lanes' exclusive prefix sums.
    {
        const int32_t n0 = B.nbr_off[g], n1 = B.nbr_off[g + 1];
        for (int32_t nb = n0; nb < n1; nb += 64) {
            const int m = min(64, n1 - nb);
            int32_t gid = 0, st = 0, sz = 0;
            if (lane < m) {
                gid = B.nbr[nb + lane];
                st = B.ga_off[gid];
                sz = B.ga_off[gid + 1] - st;
                if (!FULL && st + sz <= p + 1) sz = 0;  // UPPER: whole group below the row
            }
            int32_t inc = sz;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int32_t t = __shfl_up(inc, o);
                if (lane >= o) inc += t;
            }
            const int32_t total = __shfl(inc, 63);
            const int32_t exc = inc - sz;
            for (int32_t fb = 0; fb < total; fb += 64) {
                const int32_t f = fb + lane;
                int lo = 0, hi = m - 1;
#pragma unroll
                for (int step = 0; step < 6; ++step) {
                    const int mid = (lo + hi + 1) >> 1;
                    const int32_t e = __shfl(exc, mid);
                    if (f >= e)
                        lo = mid;
                    else
                        hi = mid - 1;
                }
                // the largest lane whose prefix is <= f owns f (empty lanes
                // share the prefix of the next lane and are never the largest)
                const int32_t gL = __shfl(gid, lo);
                const int32_t sL = __shfl(st, lo);
                const int32_t eL = __shfl(exc, lo);
                bool keep = false;
                double v = 0.0;
                int32_t q = 0;
                if (f < total) {
                    q = sL + (f - eL);
                    const int32_t j = B.sa_idx[q];
                    if ((FULL || q > p) && max(i, j) >= n_stored) {
                        const S dj = B.sa_dn[q];
                        if (popc(dn_i ^ dj) == 2) {
                            S u1 = up_i, u2 = B.ua[gL], d1 = dn_i, d2 = dj;
                            swap_if(j < i, u1, u2);
                            swap_if(j < i, d1, d2);
                            v = elem_opp_double<S>(T, u1, u2, d1, d2);
                            keep = (v != 0.0);
                        }
                    }
                }
                HCI_EMIT(keep, q, v);
            }
        }
    }
#undef HCI_EMIT
    if (!FILL && lane == 0) row_cnt[p] = cnt + pc;
}

// ---- TILED storage --------------------------------------------------------------

// One wavefront per sub-block.  Every kept pair goes to the window of its
// column slice (local column >> 8); the wave keeps one running count per
// window in LDS.  A step whose kept pairs share one window (kinds 0, 1 and 2
// enumerate pairs column-major, f = col * nr + row, so nearly all do) places
// them in lane order; a step that spreads over several windows (kind 3 emits
// in key order) takes positions with LDS atomics.  The count pass leaves the
// per-window totals; the fill pass stores each pair at its window's offset +
// position.  Positions within a window are a permutation of its slots; the
// product does not depend on their order.
// SOK: some singles vanish for every occupation and are looked up in
// DevInts::sok (instantiated apart: the lookup costs registers that the
// builds of integrals where every single can be nonzero do not pay)
template <class S, bool FILL, bool SOK>
__global__ void __launch_bounds__(kBlock)
    k_tile_connect(const SubBlock* __restrict__ sbs, int64_t n_sb, BuildView<S> B, const S* ub, DevInts T,
                   SingleRules rules, int dbl, int64_t* win_cnt, TileOut out, unsigned long long* nnz_out,
                   const int8_t* __restrict__ pairf) {
    __shared__ int32_t s_cur[kWaves][kWinPerSb];
    __shared__ int64_t s_off[kWaves][kWinPerSb];
    const int lane = threadIdx.x & 63;
    const int64_t w = (int64_t)blockIdx.x * kWaves + (threadIdx.x >> 6);
    if (w >= n_sb) return;
    const uint64_t lt_mask = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    auto single_ok = [&](S from, S to) { return !SOK || single_possible(T, from, to); };
    const SubBlock sb = sbs[w];
    const int n_slices = (sb.nc + kWinCols - 1) / kWinCols;
    if (FILL && pairf) {  // every window held by its mirror: nothing to search
        bool all = true;
        for (int k = 0; k < n_slices; ++k) all = all && pairf[sb.win_begin + k] == 2;
        if (all) return;
    }
    int32_t* wcur = s_cur[threadIdx.x >> 6];
    int64_t* woff = s_off[threadIdx.x >> 6];
    if (lane < kWinPerSb) wcur[lane] = 0;
    if (FILL && lane < n_slices) woff[lane] = out.woff[sb.win_begin + lane];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    unsigned long long nz = 0;  // FILL: nonzero entries written (pairs of a stored mirror half count twice)
    // FILL: a window whose mirror holds its entries is not written
    auto dropped = [&](int32_t cl) { return FILL && pairf && pairf[sb.win_begin + cl / kWinCols] == 2; };
    auto emit = [&](bool keep, int32_t rl, int32_t cl, double v) {
        int8_t f = 0;
        if (FILL && pairf && keep) f = pairf[sb.win_begin + cl / kWinCols];
        keep = keep && f != 2;
        const uint64_t mk = __ballot(keep);
        if (!mk) return;
        if (FILL) nz += __popcll(__ballot(keep && v != 0.0)) + __popcll(__ballot(keep && v != 0.0 && f == 1));
        const int slice = cl / kWinCols;
        const int s0 = __shfl(slice, (int)__builtin_ctzll(mk));
        const uint64_t ms = __ballot(keep && slice == s0);
        int32_t pos = 0;
        if (ms == mk) {
            pos = wcur[s0] + __popcll(ms & lt_mask);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            if (lane == (int)__builtin_ctzll(mk)) wcur[s0] += __popcll(ms);
        } else if (keep) {
            pos = atomicAdd(&wcur[slice], 1);
        }
        if (FILL && keep)
            put16(out, woff[slice] + pos, ((uint32_t)(rl + sb.rb) << 8) | (uint32_t)(cl % kWinCols), v);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
    if (sb.kind >= 3) {
        // kind 3: alpha-major rows against a single-connected alpha group;
        // kind 4: alpha-major rows against their own alpha group (beta
        // singles); kind 5: beta-major rows against their own beta group
        // (alpha singles).  Work item g = key * nr + row: the row's string
        // (beta for kinds 3 and 4, alpha for 5) minus its key-th electron;
        // every determinant of the column group whose string contains that
        // key has an equal or single-connected string
        const bool bm = sb.kind == 5;
        const int nk = bm ? B.n_alpha : B.n_beta;
        const S* keys = bm ? B.akey : B.bkey;
        const int32_t* kq = bm ? B.akq : B.bkq;
        const int32_t* goff = bm ? B.gb_off : B.ga_off;
        const S* str = bm ? B.sb_up : B.sa_dn;
        const int32_t* cidx = bm ? B.sb_idx : B.sa_idx;
        const int32_t G = sb.nr * nk;
        const int64_t seg_lo = (int64_t)goff[sb.g_col] * nk, seg_hi = (int64_t)goff[sb.g_col + 1] * nk;
        const S other_row = bm ? ub[sb.g_row] : B.ua[sb.g_row];
        const S ua_col = bm ? other_row : B.ua[sb.g_col];
        // kind 3: the alpha single between the two groups (same beta string)
        const bool asingle = sb.kind == 3 && single_ok(other_row, ua_col);
        for (int32_t gb = 0; gb < G; gb += 64) {
            const int32_t g = gb + lane;
            int32_t rl = 0, kk = 0, i = 0;
            S b = 0, key = 0;
            int64_t cur = seg_hi;
            if (g < G) {
                kk = g / sb.nr;
                rl = g - kk * sb.nr;
                const int32_t p = sb.r0 + rl;
                i = cidx[p];
                b = str[p];
                S m = b;
                for (int t = 0; t < kk; ++t) m &= m - (S)1;
                key = b & ~(((S)1) << ctz(m));
                int64_t lo = seg_lo, hi = seg_hi;
                while (lo < hi) {
                    const int64_t mid = (lo + hi) >> 1;
                    if (keys[mid] < key)
                        lo = mid + 1;
                    else
                        hi = mid;
                }
                cur = lo;
            }
            // the walk is software-pipelined: the next key and position are
            // loaded while the current match is placed
            S kn = 0;
            int32_t qn = 0;
            if (cur < seg_hi) {
                kn = keys[cur];
                qn = kq[cur];
            }
            while (true) {
                const bool pending = cur < seg_hi && kn == key;
                if (!__ballot(pending)) break;
                bool keep = false;
                double v = 0.0;
                int32_t cl = 0;
                if (pending) {
                    const int32_t q = qn;
                    ++cur;
                    if (cur < seg_hi) {
                        kn = keys[cur];
                        qn = kq[cur];
                    }
                    if (q >= sb.c0 && q < sb.c0 + sb.nc) {
                        cl = q - sb.c0;
                        const int32_t j = cidx[q];
                        S da = b, db = str[q];
                        const bool same = (da == db);
                        if (sb.kind == 3) {
                            // an equal beta string matches every key: count it once
                            keep = (same ? kk == 0 && asingle : dbl != 0) && !dropped(cl);
                            if (FILL && keep) {
                                S u1 = other_row, u2 = ua_col;
                                swap_if(j < i, u1, u2);
                                swap_if(j < i, da, db);
                                v = same ? single_or_zero(elem_single<S>(T, u1, u2, da), rule_of(rules, max(i, j)))
                                         : elem_opp_double<S>(T, u1, u2, da, db);
                            }
                        } else {
                            // strings are unique within a group: equal only on the diagonal
                            keep = !same && single_ok(da, db) && !dropped(cl);
                            if (FILL && keep) {
                                swap_if(j < i, da, db);
                                v = single_or_zero(elem_single<S>(T, da, db, other_row), rule_of(rules, max(i, j)));
                            }
                        }
                    }
                }
                emit(keep, rl, cl, v);
            }
        }
        if (!FILL && lane < n_slices) win_cnt[sb.win_begin + lane] = wcur[lane];
        if (FILL && lane == 0) atomicAdd(nnz_out, nz);
        return;
    }
    // kinds 0-2: every (row, column) pair, column by column.  Kind 2 holds
    // beta-major positions and compares alpha strings; kinds 0 and 1 compare
    // beta strings.  A lane keeps its rows' strings in registers: with 64 or
    // more rows it owns rows lane + 64 k and the wave walks the columns one
    // at a time (each column string is one uniform load); with fewer, lane
    // takes row lane % nr of column lane / nr, 64 / nr columns per step.
    const bool bpart = sb.kind == 2;
    const S* rstr = bpart ? B.sb_up : B.sa_dn;
    const int32_t* ridx = bpart ? B.sb_idx : B.sa_idx;
    // kind 1: the alpha single between the two groups (same beta string)
    const bool asingle = sb.kind == 1 && single_ok(B.ua[sb.g_row], B.ua[sb.g_col]);
    auto pair_val = [&](bool ok, S a, int32_t i, S b, int32_t j, bool& keep, double& v) {
        keep = false;
        v = 0.0;
        if (ok) {
            const int c = popc(a ^ b);
            if (sb.kind == 2) {
                // same beta string: alpha doubles, or alpha singles when every
                // double vanishes (kinds 1 and 3 are then not built)
                keep = dbl ? c == 4 : (c == 2 && single_ok(a, b));
                if (FILL && keep) {
                    swap_if(j < i, a, b);
                    v = dbl ? elem_same_double<S>(T, a, b)
                            : single_or_zero(elem_single<S>(T, a, b, ub[sb.g_row]), rule_of(rules, max(i, j)));
                }
            } else if (sb.kind == 0) {
                keep = (c == 2 && single_ok(a, b)) || (c == 4 && dbl);
                if (FILL && keep) {
                    const S up = B.ua[sb.g_row];
                    swap_if(j < i, a, b);
                    v = c == 2 ? single_or_zero(elem_single<S>(T, a, b, up), rule_of(rules, max(i, j)))
                               : elem_same_double<S>(T, a, b);
                }
            } else {
                keep = (c == 0 && asingle) || (c == 2 && dbl);
                if (FILL && keep) {
                    S ua = B.ua[sb.g_row], uc = B.ua[sb.g_col];
                    swap_if(j < i, ua, uc);
                    swap_if(j < i, a, b);
                    v = c == 0 ? single_or_zero(elem_single<S>(T, ua, uc, a), rule_of(rules, max(i, j)))
                               : elem_opp_double<S>(T, ua, uc, a, b);
                }
            }
        }
    };
    if (sb.nr >= 64) {
        // every step stays in the window of its column, and the windows are
        // visited in order: the running count of the current window is a
        // wave-uniform register (no LDS cursor), stored when its last column
        // is done
        auto emit_col = [&](bool keep, int32_t rl, int32_t cl, double v, int32_t& run, int64_t base, int wgt) {
            const uint64_t mk = __ballot(keep);
            if (FILL) {
                nz += (unsigned long long)wgt * __popcll(__ballot(keep && v != 0.0));
                if (keep)
                    put16(out, base + run + __popcll(mk & lt_mask), ((uint32_t)(rl + sb.rb) << 8) | (uint32_t)(cl % kWinCols),
                          v);
            }
            run += __popcll(mk);
        };
        constexpr int KR = kTileRows / 64;
        S rs[KR];
        int32_t ri[KR];
#pragma unroll
        for (int k = 0; k < KR; ++k) {
            const int32_t rl = lane + 64 * k;
            rs[k] = 0;
            ri[k] = 0;
            if (rl < sb.nr) {
                rs[k] = rstr[sb.r0 + rl];
                ri[k] = ridx[sb.r0 + rl];
            }
        }
        for (int32_t w0 = 0; w0 < sb.nc; w0 += kWinCols) {
            const int slice = w0 / kWinCols;
            const int8_t f = (FILL && pairf) ? pairf[sb.win_begin + slice] : 0;
            if (f == 2) continue;  // the mirror window holds these entries
            const int wgt = f == 1 ? 2 : 1;
            const int64_t base = FILL ? out.woff[sb.win_begin + slice] : 0;
            int32_t run = 0;
            const int32_t w1 = min(sb.nc, w0 + kWinCols);
            for (int32_t cl = w0; cl < w1; ++cl) {
                const int32_t q = sb.c0 + cl;
                const S b = rstr[q];
                const int32_t j = ridx[q];
#pragma unroll
                for (int k = 0; k < KR; ++k) {
                    if (64 * k >= sb.nr) continue;
                    bool keep;
                    double v;
                    pair_val(lane + 64 * k < sb.nr, rs[k], ri[k], b, j, keep, v);
                    emit_col(keep, lane + 64 * k, cl, v, run, base, wgt);
                }
            }
            if (!FILL && lane == 0) win_cnt[sb.win_begin + slice] = run;
        }
        if (FILL && lane == 0) atomicAdd(nnz_out, nz);
        return;
    } else {
        const int32_t per = 64 / sb.nr;
        const int32_t rl = lane % sb.nr, co = lane / sb.nr;
        const bool lane_ok = co < per;
        S a = 0;
        int32_t i = 0;
        if (lane_ok) {
            a = rstr[sb.r0 + rl];
            i = ridx[sb.r0 + rl];
        }
        for (int32_t cb = 0; cb < sb.nc; cb += per) {
            const int32_t cl = cb + co;
            const bool ok = lane_ok && cl < sb.nc && !dropped(cl);
            S b = 0;
            int32_t j = 0;
            if (ok) {
                b = rstr[sb.c0 + cl];
                j = ridx[sb.c0 + cl];
            }
            bool keep;
            double v;
            pair_val(ok, a, i, b, j, keep, v);
            emit(keep, rl, ok ? cl : 0, v);
        }
    }
    if (!FILL && lane < n_slices) win_cnt[sb.win_begin + lane] = wcur[lane];
    if (FILL && lane == 0) atomicAdd(nnz_out, nz);
}

// window k of sub-block b covers columns [c0 + 256 s, ...) of b
__global__ void k_win_init(const SubBlock* sbs, int64_t n_sb, int32_t* wc0, int32_t* wnc) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n_sb) return;
    const SubBlock sb = sbs[b];
    for (int32_t s = 0, k = sb.win_begin; s < sb.nc; s += kWinCols, ++k) {
        wc0[k] = sb.c0 + s;
        wnc[k] = min(kWinCols, sb.nc - s);
    }
}

// slots per window (multiple of kWinAlign) and non-empty flags; entry n_win is 0
__global__ void k_win_sizes(const int64_t* wc, int64_t n_win, int64_t* sz, int32_t* flag) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k > n_win) return;
    const int64_t c = k < n_win ? wc[k] : 0;
    sz[k] = (c + kWinAlign - 1) / kWinAlign * kWinAlign;
    flag[k] = c > 0 ? 1 : 0;
}

__global__ void k_win_compact(const int64_t* wc, const int64_t* woff, const int32_t* nidx, const int32_t* wc0,
                              const int32_t* wnc, int64_t n_win, const int8_t* pairf, const int32_t* pslot,
                              Window* out) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_win || wc[k] == 0) return;
    Window w;
    w.off_sp = woff[k];
    if (pairf && pairf[k] == 1) w.off_sp |= (int64_t)(pslot[k] + 1) << 40;
    w.c0 = wc0[k];
    w.cnt_nc = (uint32_t)wc[k] | ((uint32_t)wnc[k] << 20);
    out[nidx[k]] = w;
}

// ---- mirror pairs ---------------------------------------------------------------

// row range and part of every window (the tile that holds it); one wave per tile
constexpr int kPairWaves = 4;
__global__ void __launch_bounds__(64 * kPairWaves)
    k_win_rows(const TileTask* __restrict__ tasks, int64_t n_task, int8_t part, int32_t* wr0, int32_t* wnr,
               int8_t* wpart) {
    const int lane = threadIdx.x & 63;
    const int64_t t = (int64_t)blockIdx.x * kPairWaves + (threadIdx.x >> 6);
    if (t >= n_task) return;
    const TileTask T = tasks[t];
    for (int32_t k = T.w_begin + lane; k < T.w_end; k += 64) {
        wr0[k] = T.r0;
        wnr[k] = T.nr;
        wpart[k] = part;
    }
}

// key (part, first row, first column) of every window dense enough to pair
__global__ void k_mirror_keys(const int64_t* wc, const int32_t* wc0, const int32_t* wr0, const int8_t* wpart,
                              int64_t n, uint64_t* key, int32_t* idx) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    key[k] = wc[k] >= kPairMinEntries
                 ? ((uint64_t)wpart[k] << 62) | ((uint64_t)(uint32_t)wr0[k] << 31) | (uint64_t)(uint32_t)wc0[k]
                 : ~0ull;
    idx[k] = (int32_t)k;
}

// A dense window (rows: one tile's range R, columns C) is paired with the
// window whose rows are C and whose columns are R, when both are the only
// windows with their keys and their sizes and counts agree; the half above
// the diagonal (R before C) is kept (flag 1), the other dropped (flag 2).
__global__ void k_mirror_match(const uint64_t* __restrict__ skey, const int32_t* __restrict__ sidx, int64_t n,
                               const int64_t* wc, const int32_t* wc0, const int32_t* wnc, const int32_t* wr0,
                               const int32_t* wnr, int8_t* pairf) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || skey[i] == ~0ull) return;
    const uint64_t key = skey[i];
    if ((i > 0 && skey[i - 1] == key) || (i + 1 < n && skey[i + 1] == key)) return;
    const int32_t k = sidx[i];
    const int32_t R0 = wr0[k], C0 = wc0[k];
    if (R0 == C0) return;  // the diagonal block pairs with itself
    const uint64_t target = (key >> 62 << 62) | ((uint64_t)(uint32_t)C0 << 31) | (uint64_t)(uint32_t)R0;
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (skey[mid] < target)
            lo = mid + 1;
        else
            hi = mid;
    }
    if (lo >= n || skey[lo] != target) return;
    if (lo + 1 < n && skey[lo + 1] == target) return;
    const int32_t k2 = sidx[lo];
    if (wnc[k2] != wnr[k] || wnc[k] != wnr[k2] || wc[k2] != wc[k]) return;
    pairf[k] = R0 < C0 ? 1 : 2;
}

// dropped windows hold no entries; kept ones are numbered (flag0[k] = 1)
__global__ void k_mirror_drop(const int8_t* pairf, int64_t n, int64_t* wc, int32_t* kept) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k > n) return;
    const int8_t f = k < n ? pairf[k] : 0;
    if (f == 2) wc[k] = 0;
    kept[k] = f == 1 ? 1 : 0;
}

// receiving key of each kept window's slot: the tile whose rows are its
// columns (part, first row, rows)
__global__ void k_mirror_recv(const int8_t* pairf, const int32_t* pslot, const int32_t* wc0, const int32_t* wnc,
                              const int8_t* wpart, int64_t n, uint64_t* rkey, int32_t* rslot) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n || pairf[k] != 1) return;
    const int32_t sl = pslot[k];
    rkey[sl] = ((uint64_t)wpart[k] << 62) | ((uint64_t)(uint32_t)wc0[k] << 31) | (uint64_t)(uint32_t)wnc[k];
    rslot[sl] = sl;
}

__global__ void k_mirror_heads(const uint64_t* rkey, int64_t n, int32_t* head) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    head[i] = i < n && (i == 0 || rkey[i] != rkey[i - 1]) ? 1 : 0;
}

__global__ void k_mirror_segs(const uint64_t* rkey, const int32_t* head, const int32_t* hpos, int64_t n,
                              PairSeg* seg) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !head[i]) return;
    const int32_t g = hpos[i];
    PairSeg p;
    p.part = (int32_t)(rkey[i] >> 62);
    p.r0 = (int32_t)((rkey[i] >> 31) & 0x7FFFFFFFull);
    p.nr = (int32_t)(rkey[i] & 0x7FFFFFFFull);
    p.s_begin = (int32_t)i;
    p.s_end = (int32_t)n;
    seg[g] = p;
}

// each segment ends where the next begins (the last at n, set above)
__global__ void k_mirror_seg_end(PairSeg* seg, int32_t n_seg) {
    const int32_t g = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (g + 1 < n_seg) seg[g].s_end = seg[g + 1].s_begin;
}

// the stored halves by slot: their (compacted) window indices, and the
// largest entry count among them (for the sort's scratch)
__global__ void k_pair_list(const int8_t* pairf, const int32_t* pslot, const int32_t* nidx, const int64_t* wc,
                            int64_t n, int32_t* list, int* max_cnt) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n || pairf[k] != 1) return;
    list[pslot[k]] = nidx[k];
    atomicMax(max_cnt, (int)wc[k]);
}

// Each stored half's entries sorted by diagonal ((column - row) mod 256),
// stably, one wavefront per window: a histogram of the diagonals, its scan,
// then the entries in fill order, 64 at a time, each lane ranked among the
// lanes of its diagonal (eight ballots), scattered into the wave's scratch
// and copied back in place.  Values travel as 8-byte words (an FP64 value,
// or FP32 hi | lo << 32).
constexpr int kDiagWaves = 4;
template <bool F64>
__global__ void __launch_bounds__(64 * kDiagWaves)
    k_pair_diag(const Window* __restrict__ wins, const int32_t* __restrict__ list, int64_t n_pair, uint16_t* idx,
                double* val, float* hi, float* lo, uint16_t* sidx, uint64_t* sval, int64_t per_wave) {
    __shared__ int32_t s_cur[kDiagWaves][256];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int32_t* cur = s_cur[wv];
    const int64_t gw = (int64_t)blockIdx.x * kDiagWaves + wv, nw = (int64_t)gridDim.x * kDiagWaves;
    uint16_t* wi = sidx + gw * per_wave;
    uint64_t* wv8 = sval + gw * per_wave;
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    auto diag_of = [](uint32_t u) { return ((u & 0xFFu) - (u >> 8)) & 0xFFu; };
    for (int64_t p = gw; p < n_pair; p += nw) {
        const Window W = wins[list[p]];
        const int64_t o = W.o();
        const int cnt = W.cnt();
#pragma unroll
        for (int j = 0; j < 4; ++j) cur[lane + 64 * j] = 0;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (int j = lane; j < cnt; j += 64) atomicAdd(&cur[diag_of(idx[o + j])], 1);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // exclusive scan: lane l holds diagonals 4 l .. 4 l + 3
        int c4[4], sum = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            c4[j] = cur[4 * lane + j];
            sum += c4[j];
        }
        int inc = sum;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int t = __shfl_up(inc, off);
            if (lane >= off) inc += t;
        }
        int run = inc - sum;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            cur[4 * lane + j] = run;
            run += c4[j];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (int base = 0; base < cnt; base += 64) {
            const int j = base + lane;
            const bool ok = j < cnt;
            uint32_t u = 0, d = 0;
            uint64_t v = 0;
            if (ok) {
                u = idx[o + j];
                d = diag_of(u);
                v = F64 ? __builtin_bit_cast(uint64_t, val[o + j])
                        : (uint64_t)__builtin_bit_cast(uint32_t, hi[o + j]) |
                              ((uint64_t)__builtin_bit_cast(uint32_t, lo[o + j]) << 32);
            }
            uint64_t same = __ballot(ok);
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                const uint64_t m = __ballot(ok && ((d >> b) & 1u));
                same &= ((d >> b) & 1u) ? m : ~m;
            }
            const int leader = ok ? (int)__builtin_ctzll(same) : lane;
            int first = 0;
            if (ok && leader == lane) first = atomicAdd(&cur[d], (int)__popcll(same));
            first = __shfl(first, leader);
            if (ok) {
                const int q = first + (int)__popcll(same & below);
                wi[q] = (uint16_t)u;
                wv8[q] = v;
            }
        }
        // the wave's scratch writes complete before its lanes read them back
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        for (int j = lane; j < cnt; j += 64) {
            idx[o + j] = wi[j];
            const uint64_t v = wv8[j];
            if (F64) {
                val[o + j] = __builtin_bit_cast(double, v);
            } else {
                hi[o + j] = __builtin_bit_cast(float, (uint32_t)v);
                lo[o + j] = __builtin_bit_cast(float, (uint32_t)(v >> 32));
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
}

// In place, one wavefront per window: every 512-entry piece (the last one
// shorter) is read in enumeration order into the wave's LDS and written back
// slot by slot (the inverse of win_slot, whose slots stay inside the piece's
// own range), so both passes over memory are contiguous.  A window
// of at most 8 entries is one group whose slots are its positions.  With
// FP16 storage each value's FP16 copy is written at its slot on the way.
constexpr int kIlWaves = 4;
constexpr int kIlStage = kWinChunk + kWinChunk / 8;  // staged entries with the bank padding
// FP16 copy of a value (HCI_PREC_F16), rounded to nearest from FP32 hi + lo
__device__ __forceinline__ uint16_t fp16_of(float hi, float lo) {
    return __builtin_bit_cast(uint16_t, (_Float16)((double)hi + (double)lo));
}
typedef unsigned short il_u16x8 __attribute__((ext_vector_type(8)));
typedef float il_f32x4 __attribute__((ext_vector_type(4)));
typedef double il_f64x2 __attribute__((ext_vector_type(2)));
template <bool F64>
__global__ void __launch_bounds__(64 * kIlWaves)
    k_interleave(const Window* __restrict__ wins, int64_t n_win, uint16_t* idx, double* val, float* hi,
                 float* lo, uint16_t* h16) {
    // one staging area per wave: FP64 values, or FP32 hi and lo words, in
    // the same bytes (a build writes one or the other).  Entry j sits at
    // j + 8 (j / 64): the slot-order reads (entries 64 apart) then hit 64
    // different banks, and each lane's 8 consecutive entries stay contiguous.
    // Each lane moves 8 entries with 16-byte vectors both ways.
    __shared__ __align__(16) uint16_t s_idx[kIlWaves][kIlStage];
    __shared__ __align__(16) double s_val[kIlWaves][kIlStage];
    uint16_t* si = s_idx[threadIdx.x >> 6];
    double* sv = s_val[threadIdx.x >> 6];
    float* sh = (float*)sv;
    float* sl = sh + kIlStage;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int lp = 8 * lane + 8 * (lane >> 3);  // staged position of this lane's first entry
    for (int64_t k = (int64_t)blockIdx.x * kIlWaves + wv; k < n_win; k += (int64_t)gridDim.x * kIlWaves) {
    const Window W = wins[k];
    const int cnt = W.cnt();
    if (cnt <= kWinAlign) {  // one group: its slots are its positions
        if (!F64 && h16 && lane < cnt) h16[W.o() + lane] = fp16_of(hi[W.o() + lane], lo[W.o() + lane]);
        continue;
    }
    for (int base = 0; base < cnt; base += kWinChunk) {
        const int n = min(kWinChunk, cnt - base);
        // fill order in: lane l takes entries 8 l .. 8 l + 7 (slots up to a
        // multiple of 8 past the last entry belong to the window)
        const int64_t e0 = W.o() + base + 8 * lane;
        if (8 * lane < n) {
            *(il_u16x8*)(si + lp) = *(const il_u16x8*)(idx + e0);
            if constexpr (F64) {
#pragma unroll
                for (int v = 0; v < 4; ++v) *(il_f64x2*)(sv + lp + 2 * v) = *(const il_f64x2*)(val + e0 + 2 * v);
            } else {
#pragma unroll
                for (int v = 0; v < 2; ++v) {
                    *(il_f32x4*)(sh + lp + 4 * v) = *(const il_f32x4*)(hi + e0 + 4 * v);
                    *(il_f32x4*)(sl + lp + 4 * v) = *(const il_f32x4*)(lo + e0 + 4 * v);
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // slot order out: lane l writes slots 8 l .. 8 l + 7, which hold
        // entries 64 q + l of a full piece, or l + q R of the last one (slots
        // past its entries are written as zeros)
        const bool full = n == kWinChunk;
        const int R = (n + kWinAlign - 1) / kWinAlign;
        if (full || lane < R) {
            il_u16x8 ui, uh;
            il_f32x4 h4[2], l4[2];
            il_f64x2 d2[4];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int j = full ? q * 64 + lane : lane + q * R;
                const bool ok = j < n;
                const int pj = j + ((j >> 6) << 3);
                ui[q] = ok ? si[pj] : (unsigned short)0;
                if constexpr (F64) {
                    d2[q >> 1][q & 1] = ok ? sv[pj] : 0.0;
                } else {
                    const float hq = ok ? sh[pj] : 0.f, lq = ok ? sl[pj] : 0.f;
                    h4[q >> 2][q & 3] = hq;
                    l4[q >> 2][q & 3] = lq;
                    uh[q] = fp16_of(hq, lq);
                }
            }
            const int64_t e = W.o() + base + 8 * lane;
            *(il_u16x8*)(idx + e) = ui;
            if constexpr (F64) {
#pragma unroll
                for (int v = 0; v < 4; ++v) *(il_f64x2*)(val + e + 2 * v) = d2[v];
            } else {
#pragma unroll
                for (int v = 0; v < 2; ++v) {
                    *(il_f32x4*)(hi + e + 4 * v) = h4[v];
                    *(il_f32x4*)(lo + e + 4 * v) = l4[v];
                }
                if (h16) *(il_u16x8*)(h16 + e) = uh;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    }
}

__global__ void k_task_remap(TileTask* t, int64_t n, const int32_t* nidx) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    t[k].w_begin = nidx[t[k].w_begin];
    t[k].w_end = nidx[t[k].w_end];
}

// entries of each tile (for the order in which tiles are launched)
// inv[perm[q]] = q
__global__ void k_invert_perm(const int32_t* __restrict__ perm, int64_t n, int32_t* inv) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) inv[perm[q]] = (int32_t)q;
}

// sort key of each tile: heavier tiles first (an ascending sort of
// INT64_MAX - work), ties in position order (the sort is stable)
__global__ void k_task_work(const TileTask* t, int64_t n, const Window* __restrict__ wins, uint64_t* key,
                            int32_t* ord) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int64_t w = 0;
    for (int32_t i = t[k].w_begin; i < t[k].w_end; ++i) w += wins[i].cnt();
    w += t[k].nr;  // rows cost their write even without entries
    key[k] = (uint64_t)(INT64_MAX - w);
    ord[k] = (int32_t)k;
}

__global__ void k_gather_tasks(const TileTask* t, const int32_t* ord, int64_t n, TileTask* out) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) out[k] = t[ord[k]];
}

// Stable partition of each tile's windows into those of more than kSmallWin
// entries, then the rest; sets w_split.  One wavefront per tile.
constexpr int kPartWaves = 4;
__global__ void __launch_bounds__(64 * kPartWaves)
    k_win_partition(TileTask* tasks, int64_t n_task, const Window* __restrict__ in, Window* __restrict__ out,
                    int small_win) {
    const int lane = threadIdx.x & 63;
    const int64_t k = (int64_t)blockIdx.x * kPartWaves + (threadIdx.x >> 6);
    if (k >= n_task) return;
    const int32_t b = tasks[k].w_begin, e = tasks[k].w_end;
    int32_t n_big = 0;
    for (int32_t w = b; w < e; w += 64)
        n_big += __popcll(__ballot(w + lane < e && in[w + lane].cnt() > small_win));
    const uint64_t below = (1ull << lane) - 1;
    int32_t pb = b, ps = b + n_big;
    for (int32_t w = b; w < e; w += 64) {
        const bool valid = w + lane < e;
        Window W{};
        if (valid) W = in[w + lane];
        const bool big = valid && W.cnt() > small_win;
        const uint64_t mb = __ballot(big), ms = __ballot(valid && !big);
        if (big) out[pb + __popcll(mb & below)] = W;
        if (valid && !big) out[ps + __popcll(ms & below)] = W;
        pb += __popcll(mb);
        ps += __popcll(ms);
    }
    if (lane == 0) tasks[k].w_split = b + n_big;
}

// ---- sorting / bucketing helpers ------------------------------------------------

template <class S>
__global__ void k_load(const uint64_t* w, int64_t n, S* up, S* dn, int32_t* iota) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    up[i] = load_string<S>(w + 4 * i);
    dn[i] = load_string<S>(w + 4 * i + 2);
    iota[i] = (int32_t)i;
}

// err[0] |= 1: electron count differs from det 0; |= 2: bits beyond n_orb
template <class S>
__global__ void k_validate(const S* up, const S* dn, int64_t n, int n_orb, int* err) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int e = 0;
    if (popc(up[i]) != popc(up[0]) || popc(dn[i]) != popc(dn[0])) e |= 1;
    if (n_orb < (int)(8 * sizeof(S))) {
        const S lim = ((S)1) << n_orb;
        if (up[i] >= lim || dn[i] >= lim) e |= 2;
    }
    if (e) atomicOr(err, e);
}

template <class S>
__global__ void k_heads(const S* sorted, int64_t n, int32_t* flag) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    flag[p] = (p == 0 || sorted[p] != sorted[p - 1]) ? 1 : 0;
}

template <class S>
__global__ void k_rank_scatter(const S* sorted, const int32_t* perm, const int32_t* flag,
                               const int32_t* scan, int64_t n, int32_t* rank_of, S* uniq) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int32_t r = scan[p] - 1;
    rank_of[perm[p]] = r;
    if (flag[p]) uniq[r] = sorted[p];
}

__global__ void k_pair_keys(const int32_t* a, const int32_t* b, int64_t n, uint64_t* key,
                            int32_t* iota) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    key[i] = ((uint64_t)(uint32_t)a[i] << 32) | (uint32_t)b[i];
    iota[i] = (int32_t)i;
}

// Gather the other-spin string in sorted order and mark group starts.
template <class S>
__global__ void k_sorted_arrays(const uint64_t* skey, const int32_t* perm, int64_t n,
                                const S* other, S* s_other, int32_t* g_off, int* dup) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    s_other[p] = other[perm[p]];
    const uint64_t k = skey[p];
    const int32_t gid = (int32_t)(k >> 32);
    if (p == 0 || (int32_t)(skey[p - 1] >> 32) != gid) g_off[gid] = (int32_t)p;
    if (p > 0 && skey[p - 1] == k) atomicOr(dup, 1);
}

template <class S>
__global__ void k_r1_keys(const S* ua, int32_t U, int n_alpha, S* key, int32_t* val) {
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= U) return;
    S s = ua[g];
    S m = s;
    int t = 0;
    while (m) {
        const int o = ctz(m);
        m &= m - (S)1;
        key[g * n_alpha + t] = s & ~(((S)1) << o);
        val[g * n_alpha + t] = (int32_t)g;
        ++t;
    }
}

__global__ void k_run_starts(const int32_t* flag, const int32_t* scan, int64_t n, int32_t* run_start) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    if (flag[e]) run_start[scan[e] - 1] = (int32_t)e;
}

__global__ void k_nbr_count(const int32_t* val, const int32_t* scan, const int32_t* run_start,
                            int64_t n, int32_t* cnt) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int32_t r = scan[e] - 1;
    const int32_t len = run_start[r + 1] - run_start[r];
    if (len > 1) atomicAdd(&cnt[val[e]], len - 1);
}

__global__ void k_nbr_fill(const int32_t* val, const int32_t* scan, const int32_t* run_start,
                           int64_t n, const int32_t* off, int32_t* cursor, int32_t* nbr) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int32_t r = scan[e] - 1;
    const int32_t a = run_start[r], b = run_start[r + 1];
    const int32_t g = val[e];
    for (int32_t x = a; x < b; ++x) {
        if (x == (int32_t)e) continue;
        nbr[off[g] + atomicAdd(&cursor[g], 1)] = val[x];
    }
}

// Incremental form: entries of the previous build are renumbered from its
// internal order to the new one through the caller indices.
// FULL: previous row r becomes new row pos[pperm[r]] as a whole, in order.
__global__ void k_prev_count_full(const int64_t* prp, const int32_t* pperm, int32_t n_prev,
                                  const int32_t* pos, int64_t* prev_cnt) {
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_prev) return;
    prev_cnt[pos[pperm[r]]] = prp[r + 1] - prp[r];
}

// UPPER: an entry lands in row min(p, q) of the new order.
__global__ void k_prev_count_upper(const int64_t* prp, const int32_t* pcol, const int32_t* pperm,
                                   int32_t n_prev, const int32_t* pos, int64_t* prev_cnt) {
    const int lane = threadIdx.x & 63;
    const int64_t r = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (r >= n_prev) return;
    const int32_t p = pos[pperm[r]];
    for (int64_t e = prp[r] + lane; e < prp[r + 1]; e += 64) {
        const int32_t q = pos[pperm[pcol[e]]];
        atomicAdd((unsigned long long*)&prev_cnt[min(p, q)], 1ull);
    }
}

__global__ void k_prev_fill(const int64_t* prp, const int32_t* pcol, const double* pval,
                            const float* phi, const float* plo, const int32_t* pperm, int32_t n_prev,
                            const int32_t* pos, const int64_t* row_ptr, int64_t* cursor, bool full,
                            FillOut out) {
    const int lane = threadIdx.x & 63;
    const int64_t r = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (r >= n_prev) return;
    const int32_t p = pos[pperm[r]];
    const int64_t a = prp[r];
    for (int64_t e = a + lane; e < prp[r + 1]; e += 64) {
        const int32_t q = pos[pperm[pcol[e]]];
        const double v = pval ? pval[e] : ((double)phi[e] + (double)plo[e]);
        if (full) {
            put(out, row_ptr[p] + (e - a), q, v);
        } else {
            const int32_t row = min(p, q);
            const int64_t k = (int64_t)atomicAdd((unsigned long long*)&cursor[row], 1ull);
            put(out, row_ptr[row] + k, max(p, q), v);
        }
    }
}

template <class S>
__global__ void k_beta_r1(const S* sa_dn, int64_t n, int nbk, S* key, int32_t* kq, const int32_t* gof,
                          const int32_t* idx, int gshift) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const S b = sa_dn[q];
    // with gof: the position's group index above bit gshift, so that one
    // stable sort of the whole array orders it by group, then by key
    const S g = gof ? ((S)(uint32_t)gof[idx[q]]) << gshift : (S)0;
    S m = b;
    for (int t = 0; t < nbk; ++t) {
        const int o = ctz(m);
        m &= m - (S)1;
        key[q * nbk + t] = g | (b & ~(((S)1) << o));
        kq[q * nbk + t] = (int32_t)q;
    }
}

// clears the group bits k_beta_r1 set above bit gshift
template <class S>
__global__ void k_low_bits(S* key, int64_t n, int gshift) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) key[i] &= (((S)1) << gshift) - (S)1;
}

__global__ void k_scale_i32(const int32_t* a, int64_t n, int f, int32_t* out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = a[i] * f;
}

__global__ void k_split(const double* v, int64_t n, float* hi, float* lo) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const double x = v[e];
    const float h = (float)x;
    hi[e] = h;
    lo[e] = (float)(x - (double)h);
}

__global__ void k_gather_words(const int32_t* perm, const uint64_t* w, int64_t n, uint64_t* out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 4 * n) return;
    out[i] = w[4 * (int64_t)perm[i >> 2] + (i & 3)];
}

__global__ void k_inverse_perm(const int32_t* perm, int64_t n, int32_t* pos) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) pos[perm[p]] = (int32_t)p;
}

__global__ void k_gather_i32(const int32_t* idx, const int32_t* src, int64_t n, int32_t* dst) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) dst[p] = src[idx[p]];
}

__global__ void k_gather_f64(const int32_t* idx, const double* src, int64_t n, double* dst) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) dst[p] = src[idx[p]];
}

// ---- host helpers -----------------------------------------------------------------

struct Scratch {
    void* p = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t need) {
        if (need <= bytes) return hipSuccess;
        hipFree(p);
        p = nullptr;
        bytes = 0;
        hipError_t e = hipMalloc(&p, need);
        if (e == hipSuccess) bytes = need;
        return e;
    }
    ~Scratch() { hipFree(p); }
};

// a build's work buffers, from the context's cache of them (scratch_take /
// scratch_give): released when the object goes out of scope or by release()
struct DevBufs {
    hci_ctx* c;
    std::vector<std::pair<void*, size_t>> ptrs;
    explicit DevBufs(hci_ctx* ctx) : c(ctx) {}
    DevBufs(const DevBufs&) = delete;
    DevBufs& operator=(const DevBufs&) = delete;
    template <class T>
    hipError_t alloc(T** p, size_t count) {
        *p = nullptr;
        const size_t bytes = count * sizeof(T) + 16;
        hipError_t e = scratch_take(c, (void**)p, bytes);
        if (e == hipSuccess) ptrs.push_back({(void*)*p, bytes});
        return e;
    }
    void release(void* p) {
        for (auto& q : ptrs)
            if (q.first == p) {
                scratch_give(c, q.first, q.second);
                q.first = nullptr;
            }
    }
    ~DevBufs() {
        for (auto& q : ptrs) scratch_give(c, q.first, q.second);
    }
};

template <class K, class V>
hipError_t sort_pairs(Scratch& tmp, const K* kin, K* kout, const V* vin, V* vout, size_t n,
                      int end_bit, hipStream_t s) {
    size_t need = 0;
    hipError_t e = rocprim::radix_sort_pairs(nullptr, need, kin, kout, vin, vout, n, 0, end_bit, s);
    if (e != hipSuccess) return e;
    if ((e = tmp.ensure(need)) != hipSuccess) return e;
    return rocprim::radix_sort_pairs(tmp.p, need, kin, kout, vin, vout, n, 0, end_bit, s);
}

template <class T>
hipError_t inclusive_scan(Scratch& tmp, const T* in, T* out, size_t n, hipStream_t s) {
    size_t need = 0;
    hipError_t e = rocprim::inclusive_scan(nullptr, need, in, out, n, rocprim::plus<T>(), s);
    if (e != hipSuccess) return e;
    if ((e = tmp.ensure(need)) != hipSuccess) return e;
    return rocprim::inclusive_scan(tmp.p, need, in, out, n, rocprim::plus<T>(), s);
}

template <class T>
hipError_t exclusive_scan(Scratch& tmp, const T* in, T* out, size_t n, hipStream_t s) {
    size_t need = 0;
    hipError_t e = rocprim::exclusive_scan(nullptr, need, in, out, (T)0, n, rocprim::plus<T>(), s);
    if (e != hipSuccess) return e;
    if ((e = tmp.ensure(need)) != hipSuccess) return e;
    return rocprim::exclusive_scan(tmp.p, need, in, out, (T)0, n, rocprim::plus<T>(), s);
}

int bits_for(int64_t x) {
    int b = 1;
    while (b < 63 && ((int64_t)1 << b) <= x) ++b;
    return b;
}

#define BT(expr) HIP_TRY(c, expr)

// Rank the strings of one spin: rank_of[i] and the ascending unique list.
template <class S>
int rank_strings(hci_ctx* c, Scratch& tmp, DevBufs& db, const S* str, int64_t N, int n_orb,
                 const int32_t* iota, int32_t* rank_of, S** uniq, int32_t* n_uniq) {
    hipStream_t s = c->stream;
    S *sk = nullptr;
    int32_t *perm = nullptr, *flag = nullptr, *scan = nullptr;
    BT(db.alloc(&sk, N));
    BT(db.alloc(&perm, N));
    BT(db.alloc(&flag, N));
    BT(db.alloc(&scan, N));
    BT(sort_pairs(tmp, str, sk, iota, perm, (size_t)N, n_orb, s));
    hipLaunchKernelGGL(k_heads<S>, grid_for(N, 256), 256, 0, s, sk, N, flag);
    BT(inclusive_scan(tmp, flag, scan, (size_t)N, s));
    int32_t U = 0;
    BT(hipMemcpyAsync(&U, scan + N - 1, 4, hipMemcpyDeviceToHost, s));
    BT(hipStreamSynchronize(s));
    BT(db.alloc(uniq, (size_t)U));
    hipLaunchKernelGGL(k_rank_scatter<S>, grid_for(N, 256), 256, 0, s, sk, perm, flag, scan, N,
                       rank_of, *uniq);
    BT(hipGetLastError());
    db.release(sk);
    db.release(perm);
    db.release(flag);
    db.release(scan);
    *n_uniq = U;
    return HCI_OK;
}

// Sub-block bookkeeping on the host from the group tables, then count and
// fill passes on the device.  Parts: alpha-major tiles of every alpha group
// (columns: the group itself and its single-connected groups) and beta-major
// tiles of every beta group (columns: the group itself).
constexpr int32_t kKeyedCols = 48;  // column groups wider than this use key lookups

// Column groups of a tile of group g, ascending: in the alpha-major part (0)
// the group itself and its single-connected groups, or (every double
// vanishing) the group itself alone; in the beta-major part (1) the group
// itself.  A group of one determinant has no pairs of its own.
struct ColWalk {
    int part, dbl;
    int32_t g;
    const int32_t* off;
    const int32_t* nbr;
    int32_t k, k_end;  // neighbour cursor
    bool self_done;
    __device__ ColWalk(int part_, int dbl_, int32_t g_, const int32_t* off_, const int32_t* nbr_off,
                       const int32_t* nbr_)
        : part(part_), dbl(dbl_), g(g_), off(off_), nbr(nbr_), k(0), k_end(0), self_done(false) {
        if (part == 0 && dbl) {
            k = nbr_off[g];
            k_end = nbr_off[g + 1];
        }
        if (off[g + 1] - off[g] < 2 && !(part == 0 && dbl)) self_done = true;  // nothing inside a 1-det group
    }
    // next column group, or -1; *kind receives its sub-block kind
    __device__ int32_t next(int* kind) {
        const int32_t nb = k < k_end ? nbr[k] : INT32_MAX;
        int32_t cg;
        if (!self_done && g < nb) {
            cg = g;
            self_done = true;
        } else if (k < k_end) {
            cg = nb;
            ++k;
        } else {
            return -1;
        }
        const int32_t size = off[cg + 1] - off[cg];
        if (part == 0)
            *kind = dbl ? (cg == g ? 0 : (size > kKeyedCols ? 3 : 1)) : (size > kKeyedCols ? 4 : 0);
        else
            *kind = (!dbl && size > kKeyedCols) ? 5 : 2;
        return cg;
    }
};

// windows and sub-blocks of one tile of each group
__global__ void k_group_counts(int part, int32_t U, const int32_t* __restrict__ off, const int32_t* __restrict__ nbr_off,
                               const int32_t* __restrict__ nbr, int dbl, int32_t* wt, int32_t* sbt) {
    const int32_t g = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (g >= U) return;
    ColWalk cw(part, dbl, g, off, nbr_off, nbr);
    int32_t w = 0, b = 0;
    int kind;
    for (int32_t cg = cw.next(&kind); cg >= 0; cg = cw.next(&kind)) {
        const int32_t size = off[cg + 1] - off[cg];
        w += (size + kWinCols - 1) / kWinCols;
        b += (size + kTileCols - 1) / kTileCols;
    }
    wt[g] = w;
    sbt[g] = b;
}

// every sub-block of group g's tiles at sbbase[g], windows from wbase[g]:
// tile by tile (256 rows; a group of at most 256 rows is one tile, shared with
// its neighbours at row offset rb[g]), column group by column group, 4096
// columns at a time.  One wave per group, one lane per tile: every tile of a
// group has the same columns, so tile t starts t tiles' worth of sub-blocks
// and windows after the group's first.
constexpr int kGenWaves = 4;
__global__ void __launch_bounds__(64 * kGenWaves)
    k_gen_subblocks(int part, int32_t U, const int32_t* __restrict__ off, const int32_t* __restrict__ nbr_off,
                    const int32_t* __restrict__ nbr, int dbl, const int64_t* __restrict__ wbase,
                    const int64_t* __restrict__ sbbase, const int32_t* __restrict__ rb,
                    const int32_t* __restrict__ wt, const int32_t* __restrict__ sbt, SubBlock* out) {
    const int64_t gw = (int64_t)blockIdx.x * kGenWaves + (threadIdx.x >> 6);
    if (gw >= U) return;
    const int32_t g = (int32_t)gw;
    const int lane = threadIdx.x & 63;
    const int32_t g0 = off[g], g1 = off[g + 1];
    const int32_t n_tiles = (g1 - g0 + kTileRows - 1) / kTileRows;
    for (int32_t t = lane; t < n_tiles; t += 64) {
        const int32_t r0 = g0 + t * kTileRows;
        int64_t w = wbase[g] + (int64_t)t * wt[g], k = sbbase[g] + (int64_t)t * sbt[g];
        const int32_t nr = min(kTileRows, g1 - r0);
        const int32_t row_base = (g1 - g0 > kTileRows) ? 0 : rb[g];
        ColWalk cw(part, dbl, g, off, nbr_off, nbr);
        int kind;
        for (int32_t cg = cw.next(&kind); cg >= 0; cg = cw.next(&kind)) {
            const int32_t c1 = off[cg + 1];
            for (int32_t cs = off[cg]; cs < c1; cs += kTileCols) {
                SubBlock b;
                b.r0 = r0;
                b.nr = nr;
                b.c0 = cs;
                b.nc = min(kTileCols, c1 - cs);
                b.g_row = g;
                b.g_col = cg;
                b.kind = kind;
                b.win_begin = (int32_t)w;
                b.rb = row_base;
                out[k++] = b;
                w += (b.nc + kWinCols - 1) / kWinCols;
            }
        }
    }
}

static int group_bits(int32_t n_groups) {
    int b = 0;
    while (b < 31 && (1 << b) < n_groups) ++b;
    return b;
}

// One-word keys whose group index fits above their n_orb string bits are
// sorted within groups by one stable radix sort of (group, key) over the
// whole array: the order a per-group sort gives, at a fraction of its cost
// when there are many small groups.
template <class S>
bool grouped_sort_fits(int n_orb, int32_t n_groups) {
    return sizeof(S) == 8 && n_orb < 64 && n_orb + group_bits(n_groups) <= 64;
}

template <class S>
hipError_t sort_grouped(hci_ctx* c, Scratch& tmp, S* k_in, S* k_out, int32_t* q_in, int32_t* q_out, int64_t n,
                 int32_t n_groups) {
    if constexpr (sizeof(S) == 8) {
        const hipError_t e = sort_pairs(tmp, k_in, k_out, q_in, q_out, (size_t)n, c->n_orb + group_bits(n_groups),
                                        c->stream);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_low_bits<S>, grid_for(n, 256), 256, 0, c->stream, k_out, n, c->n_orb);
        return hipGetLastError();
    }
    return hipErrorInvalidValue;
}

template <class S>
int build_tiled(hci_ctx* c, Scratch& tmp, BuildView<S> B, const S* ub, int32_t Ua, int32_t Ub,
                int32_t n_pairs, const hci_build_opts& o, const SingleRules& rules, hci_csr* H) {
    hipStream_t s = c->stream;
    const int64_t N = B.N;
    // beta-minus-one-electron keys, sorted within each alpha group
    int nbk = 0;
    {
        uint64_t w[2];
        HIP_TRY(c, hipMemcpy(w, B.sa_dn, sizeof(S) <= 8 ? 8 : 16, hipMemcpyDeviceToHost));
        nbk = __builtin_popcountll(w[0]) + (sizeof(S) > 8 ? __builtin_popcountll(w[1]) : 0);
    }
    DevBufs kb(c);
    const int64_t NK = N * (int64_t)std::max(nbk, 1);
    S *k_in = nullptr, *k_out = nullptr;
    int32_t *q_in = nullptr, *q_out = nullptr, *seg = nullptr;
    HIP_TRY(c, kb.alloc(&k_in, (size_t)NK));
    HIP_TRY(c, kb.alloc(&k_out, (size_t)NK));
    HIP_TRY(c, kb.alloc(&q_in, (size_t)NK));
    HIP_TRY(c, kb.alloc(&q_out, (size_t)NK));
    HIP_TRY(c, kb.alloc(&seg, (size_t)Ua + 1));
    if (nbk > 0) {
        const bool grouped = grouped_sort_fits<S>(c->n_orb, Ua);
        hipLaunchKernelGGL(k_beta_r1<S>, grid_for(N, 256), 256, 0, s, B.sa_dn, N, nbk, k_in, q_in,
                           grouped ? B.ga_of : nullptr, B.sa_idx, c->n_orb);
        if (grouped) {
            HIP_TRY(c, sort_grouped(c, tmp, k_in, k_out, q_in, q_out, NK, Ua));
        } else {
            hipLaunchKernelGGL(k_scale_i32, grid_for(Ua + 1, 256), 256, 0, s, B.ga_off, (int64_t)Ua + 1, nbk, seg);
            size_t need = 0;
            HIP_TRY(c, rocprim::segmented_radix_sort_pairs(nullptr, need, k_in, k_out, q_in, q_out, (size_t)NK,
                                                           (unsigned)Ua, seg, seg + 1, 0, c->n_orb, s));
            HIP_TRY(c, tmp.ensure(need));
            HIP_TRY(c, rocprim::segmented_radix_sort_pairs(tmp.p, need, k_in, k_out, q_in, q_out, (size_t)NK,
                                                           (unsigned)Ua, seg, seg + 1, 0, c->n_orb, s));
        }
    }
    B.bkey = k_out;
    B.bkq = q_out;
    B.n_beta = nbk;
    const int dbl = c->doubles_vanish ? 0 : 1;  // no same-beta alpha doubles to place otherwise
    if (!dbl) {
        // alpha-minus-one-electron keys, sorted within each beta group: same-beta
        // alpha singles are then found inside beta groups (kind 5)
        int nak = 0;
        {
            uint64_t w[2];
            HIP_TRY(c, hipMemcpy(w, B.sb_up, sizeof(S) <= 8 ? 8 : 16, hipMemcpyDeviceToHost));
            nak = __builtin_popcountll(w[0]) + (sizeof(S) > 8 ? __builtin_popcountll(w[1]) : 0);
        }
        const int64_t NA = N * (int64_t)std::max(nak, 1);
        S *a_in = nullptr, *a_out = nullptr;
        int32_t *aq_in = nullptr, *aq_out = nullptr, *aseg = nullptr;
        HIP_TRY(c, kb.alloc(&a_in, (size_t)NA));
        HIP_TRY(c, kb.alloc(&a_out, (size_t)NA));
        HIP_TRY(c, kb.alloc(&aq_in, (size_t)NA));
        HIP_TRY(c, kb.alloc(&aq_out, (size_t)NA));
        HIP_TRY(c, kb.alloc(&aseg, (size_t)Ub + 1));
        if (nak > 0) {
            const bool grouped = grouped_sort_fits<S>(c->n_orb, Ub);
            hipLaunchKernelGGL(k_beta_r1<S>, grid_for(N, 256), 256, 0, s, B.sb_up, N, nak, a_in, aq_in,
                               grouped ? B.gb_of : nullptr, B.sb_idx, c->n_orb);
            if (grouped) {
                HIP_TRY(c, sort_grouped(c, tmp, a_in, a_out, aq_in, aq_out, NA, Ub));
            } else {
                hipLaunchKernelGGL(k_scale_i32, grid_for(Ub + 1, 256), 256, 0, s, B.gb_off, (int64_t)Ub + 1, nak,
                                   aseg);
                size_t need = 0;
                HIP_TRY(c, rocprim::segmented_radix_sort_pairs(nullptr, need, a_in, a_out, aq_in, aq_out,
                                                               (size_t)NA, (unsigned)Ub, aseg, aseg + 1, 0,
                                                               c->n_orb, s));
                HIP_TRY(c, tmp.ensure(need));
                HIP_TRY(c, rocprim::segmented_radix_sort_pairs(tmp.p, need, a_in, a_out, aq_in, aq_out,
                                                               (size_t)NA, (unsigned)Ub, aseg, aseg + 1, 0,
                                                               c->n_orb, s));
            }
        }
        B.akey = a_out;
        B.akq = aq_out;
        B.n_alpha = nak;
    }
    // Sub-blocks are generated on the device, group by group: per group, the
    // windows and sub-blocks one of its tiles needs; the host packs small
    // groups into tiles (greedy, at group granularity) and places every
    // group's sub-blocks and windows by prefix sums.
    std::vector<int32_t> ga((size_t)Ua + 1), gb((size_t)Ub + 1);
    HIP_TRY(c, hipMemcpy(ga.data(), B.ga_off, ga.size() * 4, hipMemcpyDeviceToHost));
    HIP_TRY(c, hipMemcpy(gb.data(), B.gb_off, gb.size() * 4, hipMemcpyDeviceToHost));
    std::vector<TileTask> ta, tb;
    int64_t n_wall = 0, n_sb = 0;
    DevBufs gbufs(c);
    struct PartGen {
        int part;
        int32_t U;
        const int32_t* d_off;
        int32_t* d_wt = nullptr;   // windows of one tile of group g
        int32_t* d_sbt = nullptr;  // sub-blocks of one tile of group g
        int64_t* d_wbase = nullptr;
        int64_t* d_sbbase = nullptr;
        int32_t* d_rb = nullptr;
    };
    PartGen parts[2] = {{0, Ua, B.ga_off}, {1, Ub, B.gb_off}};
    for (PartGen& P : parts) {
        if (P.part == 1 && P.U == 0) continue;
        HIP_TRY(c, gbufs.alloc(&P.d_wt, (size_t)P.U + 1));
        HIP_TRY(c, gbufs.alloc(&P.d_sbt, (size_t)P.U + 1));
        HIP_TRY(c, gbufs.alloc(&P.d_wbase, (size_t)P.U + 1));
        HIP_TRY(c, gbufs.alloc(&P.d_sbbase, (size_t)P.U + 1));
        HIP_TRY(c, gbufs.alloc(&P.d_rb, (size_t)P.U + 1));
        hipLaunchKernelGGL(k_group_counts, grid_for(P.U, 256), 256, 0, s, P.part, P.U, P.d_off, B.nbr_off, B.nbr, dbl,
                           P.d_wt, P.d_sbt);
    }
    HIP_TRY(c, hipGetLastError());
    for (PartGen& P : parts) {
        if (P.part == 1 && P.U == 0) continue;
        const std::vector<int32_t>& off = P.part == 0 ? ga : gb;
        std::vector<int32_t> wt((size_t)P.U), sbt((size_t)P.U), rb((size_t)P.U, 0);
        std::vector<int64_t> wbase((size_t)P.U), sbbase((size_t)P.U);
        HIP_TRY(c, hipMemcpy(wt.data(), P.d_wt, wt.size() * 4, hipMemcpyDeviceToHost));
        HIP_TRY(c, hipMemcpy(sbt.data(), P.d_sbt, sbt.size() * 4, hipMemcpyDeviceToHost));
        std::vector<TileTask>& tasks = P.part == 0 ? ta : tb;
        // A group larger than a tile is split into tiles of its own; smaller
        // consecutive groups share a tile of at most kTileRows rows and about
        // kPackWindows windows (their windows are consecutive, and entries
        // carry rows relative to the tile).
        constexpr int64_t kPackWindows = 256;
        TileTask cur{0, 0, 0, 0, 0};
        bool open = false;
        int64_t cur_w = 0;
        auto close = [&]() {
            if (!open) return;
            cur.w_end = (int32_t)n_wall;
            tasks.push_back(cur);
            open = false;
        };
        for (int32_t g = 0; g < P.U; ++g) {
            const int32_t size = off[g + 1] - off[g];
            wbase[g] = n_wall;
            sbbase[g] = n_sb;
            if (size > kTileRows) {
                close();
                for (int32_t r0 = off[g]; r0 < off[g + 1]; r0 += kTileRows) {
                    TileTask t{r0, std::min(kTileRows, off[g + 1] - r0), (int32_t)n_wall, 0, 0};
                    n_wall += wt[g];
                    n_sb += sbt[g];
                    t.w_end = (int32_t)n_wall;
                    tasks.push_back(t);
                }
                continue;
            }
            if (open && (cur.nr + size > kTileRows || cur_w + wt[g] > kPackWindows)) close();
            if (!open) {
                cur = TileTask{off[g], 0, (int32_t)n_wall, 0, 0};
                open = true;
                cur_w = 0;
            }
            rb[g] = cur.nr;
            cur.nr += size;
            cur_w += wt[g];
            n_wall += wt[g];
            n_sb += sbt[g];
            if (n_wall >= (int64_t)INT32_MAX) break;
        }
        close();
        if (n_wall >= (int64_t)INT32_MAX) break;
        HIP_TRY(c, hipMemcpy(P.d_wbase, wbase.data(), wbase.size() * 8, hipMemcpyHostToDevice));
        HIP_TRY(c, hipMemcpy(P.d_sbbase, sbbase.data(), sbbase.size() * 8, hipMemcpyHostToDevice));
        HIP_TRY(c, hipMemcpy(P.d_rb, rb.data(), rb.size() * 4, hipMemcpyHostToDevice));
    }
    if (n_wall >= (int64_t)INT32_MAX) return hci_fail(c, HCI_ERR_ARG, "hci_build_h: too many tiles");
    DevBufs wb(c);
    SubBlock* d_sb = nullptr;
    int64_t *d_wc = nullptr, *d_sz = nullptr, *d_woff = nullptr, *d_sum = nullptr, *d_nnz = nullptr;
    int32_t *d_flag = nullptr, *d_nidx = nullptr, *d_wc0 = nullptr, *d_wnc = nullptr;
    HIP_TRY(c, wb.alloc(&d_sb, (size_t)n_sb + 1));
    HIP_TRY(c, wb.alloc(&d_wc, (size_t)n_wall + 1));
    HIP_TRY(c, wb.alloc(&d_sz, (size_t)n_wall + 1));
    HIP_TRY(c, wb.alloc(&d_woff, (size_t)n_wall + 1));
    HIP_TRY(c, wb.alloc(&d_sum, (size_t)n_wall + 1));
    HIP_TRY(c, wb.alloc(&d_flag, (size_t)n_wall + 1));
    HIP_TRY(c, wb.alloc(&d_nidx, (size_t)n_wall + 1));
    HIP_TRY(c, wb.alloc(&d_wc0, (size_t)n_wall + 1));
    HIP_TRY(c, wb.alloc(&d_wnc, (size_t)n_wall + 1));
    HIP_TRY(c, wb.alloc(&d_nnz, 1));
    HIP_TRY(c, hipMemsetAsync(d_nnz, 0, 8, s));
    for (PartGen& P : parts) {
        if (P.part == 1 && P.U == 0) continue;
        hipLaunchKernelGGL(k_gen_subblocks, grid_for(P.U, kGenWaves), 64 * kGenWaves, 0, s, P.part, P.U, P.d_off,
                           B.nbr_off, B.nbr, dbl, P.d_wbase, P.d_sbbase, P.d_rb, P.d_wt, P.d_sbt, d_sb);
    }
    HIP_TRY(c, hipGetLastError());
    H->n_task_a = (int64_t)ta.size();
    H->n_task_b = (int64_t)tb.size();
    HIP_TRY(c, hipMalloc(&H->d_task_a, ta.size() * sizeof(TileTask) + 16));
    HIP_TRY(c, hipMalloc(&H->d_task_b, tb.size() * sizeof(TileTask) + 16));
    HIP_TRY(c, hipMemcpyAsync(H->d_task_a, ta.data(), ta.size() * sizeof(TileTask), hipMemcpyHostToDevice, s));
    HIP_TRY(c, hipMemcpyAsync(H->d_task_b, tb.data(), tb.size() * sizeof(TileTask), hipMemcpyHostToDevice, s));
    const DevInts T = c->ints();
    TileOut none{nullptr, nullptr, nullptr, nullptr, nullptr};
    if (n_sb > 0) {
        hipLaunchKernelGGL(k_win_init, grid_for(n_sb, 256), 256, 0, s, d_sb, n_sb, d_wc0, d_wnc);
        if (T.sok_all)
            hipLaunchKernelGGL((k_tile_connect<S, false, false>), grid_for(n_sb, kWaves), kBlock, 0, s, d_sb, n_sb, B,
                               ub, T, rules, dbl, d_wc, none, nullptr, nullptr);
        else
            hipLaunchKernelGGL((k_tile_connect<S, false, true>), grid_for(n_sb, kWaves), kBlock, 0, s, d_sb, n_sb, B,
                               ub, T, rules, dbl, d_wc, none, nullptr, nullptr);
    }
    // mirror pairs: dense windows whose mirror is another tile's window; the
    // kept halves are stored alone, numbered (their partial-sum slots) and
    // grouped by the tile that receives their transposed sums.  On the
    // 2e6-determinant workload the FP64 product takes 9.1 ms instead of
    // 10.6 ms, the FP32 hi + lo product 7.25 ms instead of 7.6 ms, the FP32
    // product the same 4.7 ms.  Not with FP16 copies: the FP16-value product
    // is bound by its LDS work (the LDS is busy ~87% of its cycles), which the
    // transposed sums raise, so it takes 3.39 ms with pairs against 3.09 ms
    // without, though it reads 19% fewer bytes
    int8_t* d_pairf = nullptr;
    int32_t* d_pslot = nullptr;
    if (n_wall > 0 && n_sb > 0 && o.value_prec != HCI_PREC_F16) {
        int32_t *d_wr0 = nullptr, *d_wnr = nullptr, *d_mi = nullptr, *d_mi2 = nullptr, *d_kept = nullptr;
        int8_t* d_wpart = nullptr;
        uint64_t *d_mk = nullptr, *d_mk2 = nullptr;
        HIP_TRY(c, wb.alloc(&d_pairf, (size_t)n_wall + 1));
        HIP_TRY(c, wb.alloc(&d_pslot, (size_t)n_wall + 1));
        HIP_TRY(c, wb.alloc(&d_kept, (size_t)n_wall + 1));
        HIP_TRY(c, wb.alloc(&d_wr0, (size_t)n_wall));
        HIP_TRY(c, wb.alloc(&d_wnr, (size_t)n_wall));
        HIP_TRY(c, wb.alloc(&d_wpart, (size_t)n_wall));
        HIP_TRY(c, wb.alloc(&d_mk, (size_t)n_wall));
        HIP_TRY(c, wb.alloc(&d_mk2, (size_t)n_wall));
        HIP_TRY(c, wb.alloc(&d_mi, (size_t)n_wall));
        HIP_TRY(c, wb.alloc(&d_mi2, (size_t)n_wall));
        HIP_TRY(c, hipMemsetAsync(d_pairf, 0, (size_t)n_wall + 1, s));
        HIP_TRY(c, hipMemsetAsync(d_wr0, 0, (size_t)n_wall * 4, s));
        HIP_TRY(c, hipMemsetAsync(d_wnr, 0, (size_t)n_wall * 4, s));
        HIP_TRY(c, hipMemsetAsync(d_wpart, 0, (size_t)n_wall, s));
        if (!ta.empty())
            hipLaunchKernelGGL(k_win_rows, grid_for((int64_t)ta.size(), kPairWaves), 64 * kPairWaves, 0, s,
                               H->d_task_a, (int64_t)ta.size(), (int8_t)0, d_wr0, d_wnr, d_wpart);
        if (!tb.empty())
            hipLaunchKernelGGL(k_win_rows, grid_for((int64_t)tb.size(), kPairWaves), 64 * kPairWaves, 0, s,
                               H->d_task_b, (int64_t)tb.size(), (int8_t)1, d_wr0, d_wnr, d_wpart);
        hipLaunchKernelGGL(k_mirror_keys, grid_for(n_wall, 256), 256, 0, s, d_wc, d_wc0, d_wr0, d_wpart, n_wall,
                           d_mk, d_mi);
        size_t need = 0;
        HIP_TRY(c, rocprim::radix_sort_pairs(nullptr, need, d_mk, d_mk2, d_mi, d_mi2, (size_t)n_wall, 0, 64, s));
        HIP_TRY(c, tmp.ensure(need));
        HIP_TRY(c, rocprim::radix_sort_pairs(tmp.p, need, d_mk, d_mk2, d_mi, d_mi2, (size_t)n_wall, 0, 64, s));
        hipLaunchKernelGGL(k_mirror_match, grid_for(n_wall, 256), 256, 0, s, d_mk2, d_mi2, n_wall, d_wc, d_wc0, d_wnc,
                           d_wr0, d_wnr, d_pairf);
        hipLaunchKernelGGL(k_mirror_drop, grid_for(n_wall + 1, 256), 256, 0, s, d_pairf, n_wall, d_wc, d_kept);
        HIP_TRY(c, exclusive_scan(tmp, d_kept, d_pslot, (size_t)n_wall + 1, s));
        int32_t n_pair = 0;
        HIP_TRY(c, hipMemcpyAsync(&n_pair, d_pslot + n_wall, 4, hipMemcpyDeviceToHost, s));
        HIP_TRY(c, hipStreamSynchronize(s));
        if (n_pair >= (1 << 24) - 1) return hci_fail(c, HCI_ERR_ARG, "hci_build_h: too many mirror window pairs");
        H->n_pair = n_pair;
        if (n_pair > 0) {
            uint64_t *d_rk = nullptr, *d_rk2 = nullptr;
            int32_t *d_rs = nullptr, *d_head = nullptr, *d_hpos = nullptr;
            HIP_TRY(c, wb.alloc(&d_rk, (size_t)n_pair));
            HIP_TRY(c, wb.alloc(&d_rk2, (size_t)n_pair));
            HIP_TRY(c, wb.alloc(&d_rs, (size_t)n_pair));
            HIP_TRY(c, wb.alloc(&d_head, (size_t)n_pair + 1));
            HIP_TRY(c, wb.alloc(&d_hpos, (size_t)n_pair + 1));
            HIP_TRY(c, hipMalloc(&H->d_pair_slot, (size_t)n_pair * 4 + 16));
            hipLaunchKernelGGL(k_mirror_recv, grid_for(n_wall, 256), 256, 0, s, d_pairf, d_pslot, d_wc0, d_wnc,
                               d_wpart, n_wall, d_rk, d_rs);
            need = 0;
            HIP_TRY(c, rocprim::radix_sort_pairs(nullptr, need, d_rk, d_rk2, d_rs, H->d_pair_slot, (size_t)n_pair, 0,
                                                 64, s));
            HIP_TRY(c, tmp.ensure(need));
            HIP_TRY(c, rocprim::radix_sort_pairs(tmp.p, need, d_rk, d_rk2, d_rs, H->d_pair_slot, (size_t)n_pair, 0,
                                                 64, s));
            hipLaunchKernelGGL(k_mirror_heads, grid_for((int64_t)n_pair + 1, 256), 256, 0, s, d_rk2, (int64_t)n_pair,
                               d_head);
            HIP_TRY(c, exclusive_scan(tmp, d_head, d_hpos, (size_t)n_pair + 1, s));
            int32_t n_seg = 0;
            HIP_TRY(c, hipMemcpyAsync(&n_seg, d_hpos + n_pair, 4, hipMemcpyDeviceToHost, s));
            HIP_TRY(c, hipStreamSynchronize(s));
            H->n_pair_seg = n_seg;
            HIP_TRY(c, hipMalloc(&H->d_pair_seg, (size_t)n_seg * sizeof(PairSeg) + 16));
            HIP_TRY(c, hipMalloc(&H->d_pair_part, (size_t)n_pair * kWinCols * 8 + 16));
            hipLaunchKernelGGL(k_mirror_segs, grid_for(n_pair, 256), 256, 0, s, d_rk2, d_head, d_hpos,
                               (int64_t)n_pair, H->d_pair_seg);
            hipLaunchKernelGGL(k_mirror_seg_end, grid_for(n_seg, 256), 256, 0, s, H->d_pair_seg, n_seg);
            // the segments are in key order, part first: count the alpha-major ones
            std::vector<PairSeg> hs((size_t)n_seg);
            HIP_TRY(c, hipMemcpyAsync(hs.data(), H->d_pair_seg, (size_t)n_seg * sizeof(PairSeg),
                                      hipMemcpyDeviceToHost, s));
            HIP_TRY(c, hipStreamSynchronize(s));
            int64_t n0 = 0;
            while (n0 < n_seg && hs[(size_t)n0].part == 0) ++n0;
            for (int64_t g = n0; g < n_seg; ++g)
                if (hs[(size_t)g].part == 0) return hci_fail(c, HCI_ERR_ARG, "hci_build_h: mirror segments out of order");
            H->n_pair_seg0 = n0;
        }
        HIP_TRY(c, hipGetLastError());
    }
    // window slots, offsets, the non-empty windows' new indices and the entry total
    hipLaunchKernelGGL(k_win_sizes, grid_for(n_wall + 1, 256), 256, 0, s, d_wc, n_wall, d_sz, d_flag);
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, exclusive_scan(tmp, d_sz, d_woff, (size_t)n_wall + 1, s));
    HIP_TRY(c, exclusive_scan(tmp, d_flag, d_nidx, (size_t)n_wall + 1, s));
    HIP_TRY(c, hipMemsetAsync(d_wc + n_wall, 0, 8, s));
    HIP_TRY(c, exclusive_scan(tmp, d_wc, d_sum, (size_t)n_wall + 1, s));
    int64_t end = 0, placed = 0;
    int32_t n_win = 0;
    HIP_TRY(c, hipMemcpyAsync(&end, d_woff + n_wall, 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipMemcpyAsync(&placed, d_sum + n_wall, 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipMemcpyAsync(&n_win, d_nidx + n_wall, 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipStreamSynchronize(s));
    (void)placed;
    H->nnz_alloc = end;
    H->n_win = n_win;
    const size_t cap = (size_t)end + kWinAlign;
    H->entry_cap = cap;
    // one 8-byte-per-entry value block for either precision (FP32 hi in its
    // first half, lo in its second), so a recycled block serves both
    if (o.value_prec == HCI_PREC_F64) {
        HIP_TRY(c, recycled_take(c, (void**)&H->d_val, cap * 8));
    } else {
        HIP_TRY(c, recycled_take(c, (void**)&H->d_hi, cap * (o.value_prec == HCI_PREC_F16 ? 10 : 8)));
        H->d_lo = H->d_hi + cap;
        if (o.value_prec == HCI_PREC_F16) H->d_h16 = (uint16_t*)(H->d_lo + cap);
    }
    HIP_TRY(c, recycled_take(c, (void**)&H->d_idx16, cap * 2));
    recycled_flush(c);
    TileOut fo{H->d_idx16, H->d_val, H->d_hi, H->d_lo, d_woff};
    HIP_TRY(c, hipMalloc(&H->d_win, (size_t)n_win * sizeof(Window) + 16));
    if (n_sb > 0) {
        if (T.sok_all)
            hipLaunchKernelGGL((k_tile_connect<S, true, false>), grid_for(n_sb, kWaves), kBlock, 0, s, d_sb, n_sb, B,
                               ub, T, rules, dbl, nullptr, fo, (unsigned long long*)d_nnz, d_pairf);
        else
            hipLaunchKernelGGL((k_tile_connect<S, true, true>), grid_for(n_sb, kWaves), kBlock, 0, s, d_sb, n_sb, B,
                               ub, T, rules, dbl, nullptr, fo, (unsigned long long*)d_nnz, d_pairf);
    }
    hipLaunchKernelGGL(k_win_compact, grid_for(n_wall, 256), 256, 0, s, d_wc, d_woff, d_nidx, d_wc0, d_wnc, n_wall,
                       d_pairf, d_pslot, H->d_win);
    if (H->n_pair > 0) {
        // the stored halves sorted by diagonal before the interleave, each
        // wave through its own scratch (at most ~256 MB in all)
        int32_t* d_plist = nullptr;
        int* d_pmax = nullptr;
        HIP_TRY(c, wb.alloc(&d_plist, (size_t)H->n_pair));
        HIP_TRY(c, wb.alloc(&d_pmax, 1));
        HIP_TRY(c, hipMemsetAsync(d_pmax, 0, 4, s));
        hipLaunchKernelGGL(k_pair_list, grid_for(n_wall, 256), 256, 0, s, d_pairf, d_pslot, d_nidx, d_wc, n_wall, d_plist,
                           d_pmax);
        int pmax = 0;
        HIP_TRY(c, hipMemcpyAsync(&pmax, d_pmax, 4, hipMemcpyDeviceToHost, s));
        HIP_TRY(c, hipStreamSynchronize(s));
        const int64_t per_wave = ((int64_t)pmax + 63) / 64 * 64;
        int64_t waves = std::min<int64_t>((256ll << 20) / (per_wave * 10), 2048);
        waves = std::max<int64_t>(std::min<int64_t>(waves, H->n_pair), 1);
        const int64_t blocks = (waves + kDiagWaves - 1) / kDiagWaves;
        uint16_t* d_sidx = nullptr;
        uint64_t* d_sval = nullptr;
        HIP_TRY(c, wb.alloc(&d_sidx, (size_t)(blocks * kDiagWaves * per_wave)));
        HIP_TRY(c, wb.alloc(&d_sval, (size_t)(blocks * kDiagWaves * per_wave)));
        if (H->d_val)
            hipLaunchKernelGGL(k_pair_diag<true>, (unsigned)blocks, 64 * kDiagWaves, 0, s, H->d_win, d_plist,
                               H->n_pair, H->d_idx16, H->d_val, H->d_hi, H->d_lo, d_sidx, d_sval, per_wave);
        else
            hipLaunchKernelGGL(k_pair_diag<false>, (unsigned)blocks, 64 * kDiagWaves, 0, s, H->d_win, d_plist,
                               H->n_pair, H->d_idx16, H->d_val, H->d_hi, H->d_lo, d_sidx, d_sval, per_wave);
        HIP_TRY(c, hipGetLastError());
    }
    if (n_win > 0) {
        const unsigned ilg = std::min<unsigned>(grid_for(n_win, kIlWaves), 4096u);
        if (H->d_val)
            hipLaunchKernelGGL(k_interleave<true>, ilg, 64 * kIlWaves, 0, s, H->d_win, (int64_t)n_win, H->d_idx16,
                               H->d_val, H->d_hi, H->d_lo, H->d_h16);
        else
            hipLaunchKernelGGL(k_interleave<false>, ilg, 64 * kIlWaves, 0, s, H->d_win, (int64_t)n_win, H->d_idx16,
                               H->d_val, H->d_hi, H->d_lo, H->d_h16);
    }
    if (!ta.empty())
        hipLaunchKernelGGL(k_task_remap, grid_for((int64_t)ta.size(), 256), 256, 0, s, H->d_task_a,
                           (int64_t)ta.size(), d_nidx);
    if (!tb.empty())
        hipLaunchKernelGGL(k_task_remap, grid_for((int64_t)tb.size(), 256), 256, 0, s, H->d_task_b,
                           (int64_t)tb.size(), d_nidx);
    HIP_TRY(c, hipGetLastError());
    {
        // every window lies in the range of exactly one tile of one part.  The
        // windows taken one per lane: at most kSmallWin entries, or
        // kSmallWinDense when rows hold many entries (HCI-selected spaces,
        // whose many small groups leave many windows of a few dozen entries)
        const int small_win = (int64_t)cap >= kDenseRowSlots * N ? kSmallWinDense : kSmallWin;
        Window* parted = nullptr;
        HIP_TRY(c, wb.alloc(&parted, (size_t)n_win + 1));
        if (!ta.empty())
            hipLaunchKernelGGL(k_win_partition, grid_for((int64_t)ta.size(), kPartWaves), 64 * kPartWaves, 0, s,
                               H->d_task_a, (int64_t)ta.size(), H->d_win, parted, small_win);
        if (!tb.empty())
            hipLaunchKernelGGL(k_win_partition, grid_for((int64_t)tb.size(), kPartWaves), 64 * kPartWaves, 0, s,
                               H->d_task_b, (int64_t)tb.size(), H->d_win, parted, small_win);
        HIP_TRY(c, hipGetLastError());
        if (n_win > 0)
            HIP_TRY(c, hipMemcpyAsync(H->d_win, parted, (size_t)n_win * sizeof(Window), hipMemcpyDeviceToDevice, s));
    }
    // alpha-major tiles launch heaviest first (longest-processing-time order),
    // so the largest alpha groups do not start last and stretch the product's
    // tail.  Beta-major tiles stay in position order: neighbouring beta groups
    // add into neighbouring y entries of each alpha group, so tiles running
    // together share y's cache lines (FP16 product on the 2e6-determinant
    // workload 3.125 -> 3.054 ms, FP32 product on the 6x6 lattice 0.905 ->
    // 0.878 ms)
    if (ta.size() >= 2) {
        TileTask* d_t = H->d_task_a;
        const int64_t nt = (int64_t)ta.size();
        // keys, indices and the sorted tasks in one allocation
        const size_t nk = ((size_t)nt + 63) / 64 * 64;
        uint64_t* d_key = nullptr;
        HIP_TRY(c, wb.alloc(&d_key, nk * 2 + nk + (nk * sizeof(TileTask) + 7) / 8));
        uint64_t* d_key2 = d_key + nk;
        int32_t* d_ord = (int32_t*)(d_key2 + nk);
        int32_t* d_ord2 = d_ord + nk;
        TileTask* d_sorted = (TileTask*)(d_ord2 + nk);
        hipLaunchKernelGGL(k_task_work, grid_for(nt, 256), 256, 0, s, d_t, nt, H->d_win, d_key, d_ord);
        HIP_TRY(c, sort_pairs(tmp, d_key, d_key2, d_ord, d_ord2, (size_t)nt, 64, s));
        hipLaunchKernelGGL(k_gather_tasks, grid_for(nt, 256), 256, 0, s, d_t, d_ord2, nt, d_sorted);
        HIP_TRY(c, hipGetLastError());
        HIP_TRY(c, hipMemcpyAsync(d_t, d_sorted, (size_t)nt * sizeof(TileTask), hipMemcpyDeviceToDevice, s));
    }
    HIP_TRY(c, hipMalloc(&H->d_bpos, (size_t)N * 4 + 16));
    HIP_TRY(c, hipMemcpyAsync(H->d_bpos, B.sb_pos, (size_t)N * 4, hipMemcpyDeviceToDevice, s));
    HIP_TRY(c, hipMalloc(&H->d_binv, (size_t)N * 4 + 16));
    hipLaunchKernelGGL(k_invert_perm, grid_for(N, 256), 256, 0, s, H->d_bpos, N, H->d_binv);
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipMalloc(&H->d_xb, (size_t)N * 8 + 16));
    HIP_TRY(c, hipMalloc(&H->d_yb, (size_t)N * 8 + 16));
    if (H->d_h16) {
        HIP_TRY(c, hipMalloc(&H->d_x32, (size_t)N * 4 + 16));
        HIP_TRY(c, hipMalloc(&H->d_xb32, (size_t)N * 4 + 16));
    }
    int64_t nnz = 0;  // stored nonzero entries (both triangles)
    HIP_TRY(c, hipMemcpyAsync(&nnz, d_nnz, 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipStreamSynchronize(s));  // the passes read the work buffers freed on return
    H->nnz_s = nnz;
    H->nnz_u = nnz / 2;
    return HCI_OK;
}

template <class S>
int build_impl(hci_ctx* c, const uint64_t* dets_dev, const double* diag_dev, int64_t N,
               const hci_csr* prev, int64_t n_stored, const hci_build_opts& o, const SingleRules& rules,
               hci_csr** out, hci_build_stats* stats) {
    hipStream_t s = c->stream;
    Scratch tmp;
    DevBufs db(c);
    const int n_orb = c->n_orb;
    const bool full = (o.storage == HCI_STORE_FULL);

    S *up = nullptr, *dn = nullptr;
    int32_t* iota = nullptr;
    int* d_err = nullptr;
    BT(db.alloc(&up, N));
    BT(db.alloc(&dn, N));
    BT(db.alloc(&iota, N));
    BT(db.alloc(&d_err, 2));
    BT(hipMemsetAsync(d_err, 0, 8, s));
    hipLaunchKernelGGL(k_load<S>, grid_for(N, 256), 256, 0, s, dets_dev, N, up, dn, iota);
    hipLaunchKernelGGL(k_validate<S>, grid_for(N, 256), 256, 0, s, up, dn, N, n_orb, d_err);
    BT(hipGetLastError());

    int32_t *ra = nullptr, *rb = nullptr;
    S *ua = nullptr, *ub = nullptr;
    int32_t Ua = 0, Ub = 0;
    BT(db.alloc(&ra, N));
    BT(db.alloc(&rb, N));
    int rc = rank_strings<S>(c, tmp, db, up, N, n_orb, iota, ra, &ua, &Ua);
    if (rc) return rc;
    rc = rank_strings<S>(c, tmp, db, dn, N, n_orb, iota, rb, &ub, &Ub);
    if (rc) return rc;
    int herr[2] = {0, 0};
    BT(hipMemcpyAsync(herr, d_err, 8, hipMemcpyDeviceToHost, s));
    BT(hipStreamSynchronize(s));
    if (herr[0] & 1) return hci_fail(c, HCI_ERR_ARG, "hci_build_h: determinants differ in electron counts");
    if (herr[0] & 2) return hci_fail(c, HCI_ERR_ARG, "hci_build_h: determinant bits beyond n_orb");

    // alpha-major (internal) and beta-major orders
    uint64_t *key = nullptr, *skey = nullptr;
    int32_t *sa_idx = nullptr, *sb_idx = nullptr, *sb_pos = nullptr, *pos = nullptr, *ga_off = nullptr,
            *gb_off = nullptr;
    S *sa_dn = nullptr, *sb_up = nullptr;
    BT(db.alloc(&key, N));
    BT(db.alloc(&skey, N));
    BT(db.alloc(&sb_idx, N));
    BT(db.alloc(&sb_pos, N));
    BT(db.alloc(&sa_dn, N));
    BT(db.alloc(&sb_up, N));
    BT(db.alloc(&ga_off, (size_t)Ua + 1));
    BT(db.alloc(&gb_off, (size_t)Ub + 1));
    hci_csr* H = new hci_csr();
    H->ctx = c;
    H->n = N;
    H->prec = o.value_prec;
    H->storage = o.storage;
    auto fail_free = [&](int code) {
        hci_csr_free(H);
        return code;
    };
#define BTH(expr)                                                                                   \
    do {                                                                                            \
        hipError_t _e = (expr);                                                                     \
        if (_e != hipSuccess)                                                                       \
            return fail_free(hci_fail(c, _e == hipErrorOutOfMemory ? HCI_ERR_OOM : HCI_ERR_HIP,     \
                                      std::string("hci_build_h: ") + #expr + ": " + hipGetErrorString(_e))); \
    } while (0)
    BTH(hipMalloc(&H->d_perm, (size_t)N * 4 + 16));
    BTH(hipMalloc(&H->d_pos, (size_t)N * 4 + 16));
    sa_idx = H->d_perm;
    pos = H->d_pos;
    const int32_t N32 = (int32_t)N;
    BTH(hipMemcpy(ga_off + Ua, &N32, 4, hipMemcpyHostToDevice));
    BTH(hipMemcpy(gb_off + Ub, &N32, 4, hipMemcpyHostToDevice));
    const int key_bits_a = 32 + bits_for(Ua);
    const int key_bits_b = 32 + bits_for(Ub);
    hipLaunchKernelGGL(k_pair_keys, grid_for(N, 256), 256, 0, s, ra, rb, N, key, iota);
    BTH(sort_pairs(tmp, key, skey, iota, sa_idx, (size_t)N, key_bits_a, s));
    hipLaunchKernelGGL(k_sorted_arrays<S>, grid_for(N, 256), 256, 0, s, skey, sa_idx, N, dn, sa_dn,
                       ga_off, d_err + 1);
    hipLaunchKernelGGL(k_pair_keys, grid_for(N, 256), 256, 0, s, rb, ra, N, key, iota);
    BTH(sort_pairs(tmp, key, skey, iota, sb_idx, (size_t)N, key_bits_b, s));
    hipLaunchKernelGGL(k_sorted_arrays<S>, grid_for(N, 256), 256, 0, s, skey, sb_idx, N, up, sb_up,
                       gb_off, d_err + 1);
    hipLaunchKernelGGL(k_inverse_perm, grid_for(N, 256), 256, 0, s, sa_idx, N, pos);
    hipLaunchKernelGGL(k_gather_i32, grid_for(N, 256), 256, 0, s, sb_idx, pos, N, sb_pos);
    BTH(hipGetLastError());
    BTH(hipMemcpyAsync(herr, d_err, 8, hipMemcpyDeviceToHost, s));
    BTH(hipStreamSynchronize(s));
    if (herr[1]) return fail_free(hci_fail(c, HCI_ERR_ARG, "hci_build_h: duplicate determinants"));
    db.release(key);
    db.release(skey);

    // single-connected alpha-group pairs via remove-one keys
    int n_alpha = 0;
    {
        uint64_t w[4];
        BTH(hipMemcpy(w, dets_dev, 32, hipMemcpyDeviceToHost));
        n_alpha = __builtin_popcountll(w[0]) + __builtin_popcountll(w[1]);
    }
    const int64_t E = (int64_t)Ua * n_alpha;
    int32_t *nbr_off = nullptr, *nbr = nullptr;
    int32_t n_pairs = 0;
    BTH(db.alloc(&nbr_off, (size_t)Ua + 1));
    BTH(hipMemsetAsync(nbr_off, 0, ((size_t)Ua + 1) * 4, s));
    if (E > 0) {
        S *k1 = nullptr, *k1s = nullptr;
        int32_t *v1 = nullptr, *v1s = nullptr, *flag = nullptr, *scan = nullptr, *run_start = nullptr,
                *cnt = nullptr, *cursor = nullptr, *nbr_u = nullptr;
        BTH(db.alloc(&k1, E));
        BTH(db.alloc(&k1s, E));
        BTH(db.alloc(&v1, E));
        BTH(db.alloc(&v1s, E));
        BTH(db.alloc(&flag, E));
        BTH(db.alloc(&scan, E));
        BTH(db.alloc(&run_start, E + 1));
        BTH(db.alloc(&cnt, (size_t)Ua + 1));
        BTH(db.alloc(&cursor, (size_t)Ua));
        hipLaunchKernelGGL(k_r1_keys<S>, grid_for(Ua, 256), 256, 0, s, ua, Ua, n_alpha, k1, v1);
        BTH(sort_pairs(tmp, k1, k1s, v1, v1s, (size_t)E, n_orb, s));
        hipLaunchKernelGGL(k_heads<S>, grid_for(E, 256), 256, 0, s, k1s, E, flag);
        BTH(inclusive_scan(tmp, flag, scan, (size_t)E, s));
        int32_t R = 0;
        BTH(hipMemcpyAsync(&R, scan + E - 1, 4, hipMemcpyDeviceToHost, s));
        BTH(hipStreamSynchronize(s));
        const int32_t E32 = (int32_t)E;
        BTH(hipMemcpy(run_start + R, &E32, 4, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_run_starts, grid_for(E, 256), 256, 0, s, flag, scan, E, run_start);
        BTH(hipMemsetAsync(cnt, 0, ((size_t)Ua + 1) * 4, s));
        BTH(hipMemsetAsync(cursor, 0, (size_t)Ua * 4, s));
        hipLaunchKernelGGL(k_nbr_count, grid_for(E, 256), 256, 0, s, v1s, scan, run_start, E, cnt);
        BTH(exclusive_scan(tmp, cnt, nbr_off, (size_t)Ua + 1, s));
        BTH(hipMemcpyAsync(&n_pairs, nbr_off + Ua, 4, hipMemcpyDeviceToHost, s));
        BTH(hipStreamSynchronize(s));
        BTH(db.alloc(&nbr_u, (size_t)n_pairs + 1));
        BTH(db.alloc(&nbr, (size_t)n_pairs + 1));
        hipLaunchKernelGGL(k_nbr_fill, grid_for(E, 256), 256, 0, s, v1s, scan, run_start, E, nbr_off,
                           cursor, nbr_u);
        BTH(hipGetLastError());
        // ascending neighbour lists: a deterministic enumeration order
        if (n_pairs > 0) {
            size_t need = 0;
            const int gbits = bits_for(Ua);
            BTH(rocprim::segmented_radix_sort_keys(nullptr, need, nbr_u, nbr, (size_t)n_pairs, (unsigned)Ua,
                                                   nbr_off, nbr_off + 1, 0, gbits, s));
            BTH(tmp.ensure(need));
            BTH(rocprim::segmented_radix_sort_keys(tmp.p, need, nbr_u, nbr, (size_t)n_pairs, (unsigned)Ua,
                                                   nbr_off, nbr_off + 1, 0, gbits, s));
        }
        BTH(hipStreamSynchronize(s));
        for (void* q : {(void*)k1, (void*)k1s, (void*)v1, (void*)v1s, (void*)flag, (void*)scan,
                        (void*)run_start, (void*)cnt, (void*)cursor, (void*)nbr_u})
            db.release(q);
    } else {
        BTH(db.alloc(&nbr, 1));
    }

    BuildView<S> B;
    B.N = N32;
    B.ga_of = ra;
    B.gb_of = rb;
    B.ua = ua;
    B.ga_off = ga_off;
    B.sa_dn = sa_dn;
    B.sa_idx = sa_idx;
    B.gb_off = gb_off;
    B.sb_up = sb_up;
    B.sb_idx = sb_idx;
    B.sb_pos = sb_pos;
    B.nbr_off = nbr_off;
    B.nbr = nbr;
    B.bkey = nullptr;
    B.bkq = nullptr;
    B.n_beta = 0;
    B.akey = nullptr;
    B.akq = nullptr;
    B.n_alpha = 0;
    const DevInts T = c->ints();

    if (o.storage == HCI_STORE_TILED) {
        int trc = build_tiled<S>(c, tmp, B, ub, Ua, Ub, n_pairs, o, rules, H);
        if (trc) return fail_free(trc);
        BTH(hipMalloc(&H->d_diag, (size_t)N * 8 + 16));
        hipLaunchKernelGGL(k_gather_f64, grid_for(N, 256), 256, 0, s, sa_idx, diag_dev, N, H->d_diag);
        BTH(hipMalloc(&H->d_words, (size_t)N * 32 + 16));
        hipLaunchKernelGGL(k_gather_words, grid_for(4 * N, 256), 256, 0, s, sa_idx, dets_dev, N, H->d_words);
        BTH(hipStreamSynchronize(s));
        if (stats) {
            stats->nnz_u = H->nnz_u;
            stats->n_alpha = Ua;
            stats->n_beta = Ub;
            stats->alpha_pairs = n_pairs;
            stats->mirror_pairs = H->n_pair;
            stats->stored_entries = H->nnz_alloc;
            const int64_t vb = o.value_prec == HCI_PREC_F16 ? 10 : 8;
            stats->matrix_bytes = (int64_t)H->entry_cap * (2 + vb);
        }
        *out = H;
        return HCI_OK;
    }

    // previous build (incremental form): per-row counts of its renumbered entries
    int64_t* prev_cnt = nullptr;
    const int32_t n_prev = prev ? (int32_t)prev->n : 0;
    if (prev) {
        BTH(db.alloc(&prev_cnt, (size_t)N));
        BTH(hipMemsetAsync(prev_cnt, 0, (size_t)N * 8, s));
        if (n_prev > 0) {
            if (full)
                hipLaunchKernelGGL(k_prev_count_full, grid_for(n_prev, 256), 256, 0, s, prev->d_row_ptr,
                                   prev->d_perm, n_prev, pos, prev_cnt);
            else
                hipLaunchKernelGGL(k_prev_count_upper, grid_for(n_prev, 4), 256, 0, s, prev->d_row_ptr,
                                   prev->d_col, prev->d_perm, n_prev, pos, prev_cnt);
        }
    }

    // count pass
    int64_t* row_cnt = nullptr;
    BTH(db.alloc(&row_cnt, (size_t)N + 1));
    BTH(hipMemsetAsync(row_cnt, 0, ((size_t)N + 1) * 8, s));
    const unsigned grid = grid_for(N, kWaves);
    FillOut none{nullptr, nullptr, nullptr, nullptr};
    if (full)
        hipLaunchKernelGGL((k_connect<S, false, true>), grid, kBlock, 0, s, B, T, (int32_t)n_stored,
                           o.single_rule, row_cnt, nullptr, prev_cnt, none);
    else
        hipLaunchKernelGGL((k_connect<S, false, false>), grid, kBlock, 0, s, B, T, (int32_t)n_stored,
                           o.single_rule, row_cnt, nullptr, prev_cnt, none);
    BTH(hipGetLastError());
    BTH(hipMalloc(&H->d_row_ptr, ((size_t)N + 1) * 8));
    int64_t* row_ptr = H->d_row_ptr;
    BTH(exclusive_scan(tmp, row_cnt, row_ptr, (size_t)N + 1, s));
    int64_t nnz = 0;
    BTH(hipMemcpyAsync(&nnz, row_ptr + N, 8, hipMemcpyDeviceToHost, s));
    BTH(hipStreamSynchronize(s));
    db.release(row_cnt);
    H->nnz_s = nnz;
    H->nnz_alloc = nnz;
    H->nnz_u = full ? nnz / 2 : nnz;

    // fill: straight into the final arrays, except an incremental UPPER build,
    // whose renumbered rows need a sort to become deterministic
    const bool sort_rows = prev && !full;
    BTH(hipMalloc(&H->d_col, (size_t)nnz * 4 + 16));
    FillOut fo{H->d_col, nullptr, nullptr, nullptr};
    int32_t* col_tmp = nullptr;
    double* val_tmp = nullptr;
    if (sort_rows) {
        BTH(db.alloc(&col_tmp, (size_t)nnz + 1));
        BTH(db.alloc(&val_tmp, (size_t)nnz + 1));
        fo = FillOut{col_tmp, val_tmp, nullptr, nullptr};
    } else if (o.value_prec == HCI_PREC_F64) {
        BTH(hipMalloc(&H->d_val, (size_t)nnz * 8 + 16));
        fo.val = H->d_val;
    } else {
        BTH(hipMalloc(&H->d_hi, (size_t)nnz * 4 + 16));
        BTH(hipMalloc(&H->d_lo, (size_t)nnz * 4 + 16));
        fo.hi = H->d_hi;
        fo.lo = H->d_lo;
    }
    if (prev && n_prev > 0) {
        int64_t* cursor = nullptr;
        if (!full) {
            BTH(db.alloc(&cursor, (size_t)N));
            BTH(hipMemsetAsync(cursor, 0, (size_t)N * 8, s));
        }
        hipLaunchKernelGGL(k_prev_fill, grid_for(n_prev, 4), 256, 0, s, prev->d_row_ptr, prev->d_col,
                           prev->d_val, prev->d_hi, prev->d_lo, prev->d_perm, n_prev, pos, row_ptr, cursor,
                           full, fo);
    }
    if (full)
        hipLaunchKernelGGL((k_connect<S, true, true>), grid, kBlock, 0, s, B, T, (int32_t)n_stored,
                           o.single_rule, nullptr, row_ptr, prev_cnt, fo);
    else
        hipLaunchKernelGGL((k_connect<S, true, false>), grid, kBlock, 0, s, B, T, (int32_t)n_stored,
                           o.single_rule, nullptr, row_ptr, prev_cnt, fo);
    BTH(hipGetLastError());
    if (sort_rows && nnz > 0) {
        double* val_s = nullptr;
        BTH(db.alloc(&val_s, (size_t)nnz + 1));
        size_t need = 0;
        const int cbits = bits_for(N);
        BTH(rocprim::segmented_radix_sort_pairs(nullptr, need, col_tmp, H->d_col, val_tmp, val_s, (size_t)nnz,
                                                (unsigned)N, row_ptr, row_ptr + 1, 0, cbits, s));
        BTH(tmp.ensure(need));
        BTH(rocprim::segmented_radix_sort_pairs(tmp.p, need, col_tmp, H->d_col, val_tmp, val_s, (size_t)nnz,
                                                (unsigned)N, row_ptr, row_ptr + 1, 0, cbits, s));
        if (o.value_prec == HCI_PREC_F64) {
            BTH(hipMalloc(&H->d_val, (size_t)nnz * 8 + 16));
            BTH(hipMemcpyAsync(H->d_val, val_s, (size_t)nnz * 8, hipMemcpyDeviceToDevice, s));
        } else {
            BTH(hipMalloc(&H->d_hi, (size_t)nnz * 4 + 16));
            BTH(hipMalloc(&H->d_lo, (size_t)nnz * 4 + 16));
            hipLaunchKernelGGL(k_split, grid_for(nnz, 256), 256, 0, s, val_s, nnz, H->d_hi, H->d_lo);
        }
        BTH(hipStreamSynchronize(s));
    } else if (sort_rows) {
        if (o.value_prec == HCI_PREC_F64)
            BTH(hipMalloc(&H->d_val, 16));
        else {
            BTH(hipMalloc(&H->d_hi, 16));
            BTH(hipMalloc(&H->d_lo, 16));
        }
    }
    BTH(hipMalloc(&H->d_diag, (size_t)N * 8 + 16));
    hipLaunchKernelGGL(k_gather_f64, grid_for(N, 256), 256, 0, s, sa_idx, diag_dev, N, H->d_diag);
    BTH(hipMalloc(&H->d_words, (size_t)N * 32 + 16));
    hipLaunchKernelGGL(k_gather_words, grid_for(4 * N, 256), 256, 0, s, sa_idx, dets_dev, N, H->d_words);
    int prc = spmv_prepare(H);
    if (prc) return fail_free(prc);
    BTH(hipStreamSynchronize(s));
#undef BTH
    if (stats) {
        stats->nnz_u = H->nnz_u;
        stats->n_alpha = Ua;
        stats->n_beta = Ub;
        stats->alpha_pairs = n_pairs;
        stats->mirror_pairs = 0;
        stats->stored_entries = H->nnz_s;
        stats->matrix_bytes = H->nnz_s * (4 + 8) + (int64_t)(N + 1) * 8;
    }
    *out = H;
    return HCI_OK;
}

#undef BT

}  // namespace

extern "C" int hci_build_h(hci_ctx* c, const uint64_t* dets, const double* diag, int64_t n,
                           const hci_csr* prev, int64_t n_stored, const hci_build_opts* opts,
                           hci_csr** out, hci_build_stats* stats) {
    if (!c) return HCI_ERR_ARG;
    if (!out || !opts) return hci_fail(c, HCI_ERR_ARG, "hci_build_h: null out/opts");
    *out = nullptr;
    if (!c->d_two) return hci_fail(c, HCI_ERR_STATE, "hci_build_h: integrals not set");
    if (n < 1 || n >= (int64_t)INT32_MAX || !dets || !diag)
        return hci_fail(c, HCI_ERR_ARG, "hci_build_h: need 1 <= n < 2^31 and dets/diag");
    if (opts->value_prec != HCI_PREC_F64 && opts->value_prec != HCI_PREC_F32 && opts->value_prec != HCI_PREC_F16)
        return hci_fail(c, HCI_ERR_ARG, "hci_build_h: value_prec must be 64, 32 or 16");
    if (opts->value_prec == HCI_PREC_F16 && opts->storage != HCI_STORE_TILED)
        return hci_fail(c, HCI_ERR_ARG, "hci_build_h: value_prec 16 needs HCI_STORE_TILED");
    if (opts->single_rule != HCI_SINGLES_ABS_GT_1E12 && opts->single_rule != HCI_SINGLES_NONZERO)
        return hci_fail(c, HCI_ERR_ARG, "hci_build_h: bad single_rule");
    if (opts->storage != HCI_STORE_FULL && opts->storage != HCI_STORE_UPPER && opts->storage != HCI_STORE_TILED)
        return hci_fail(c, HCI_ERR_ARG, "hci_build_h: bad storage");
    if (prev) {
        if (prev->ctx != c) return hci_fail(c, HCI_ERR_ARG, "hci_build_h: prev from another context");
        if (n_stored != prev->n || n < prev->n)
            return hci_fail(c, HCI_ERR_ARG, "hci_build_h: incremental build needs n_stored == prev n <= n");
        if (prev->storage != opts->storage)
            return hci_fail(c, HCI_ERR_ARG, "hci_build_h: prev has another storage kind");
    } else if (n_stored != 0) {
        return hci_fail(c, HCI_ERR_ARG, "hci_build_h: n_stored must be 0 without prev");
    }
    // single-storage rule of every pair: the previous builds' segments, then
    // this build's rule up to n (adjacent equal rules merged)
    std::vector<std::pair<int64_t, int>> segs;
    if (prev) segs = prev->rule_segs;
    if (!segs.empty() && segs.back().second == opts->single_rule)
        segs.back().first = n;
    else
        segs.emplace_back(n, opts->single_rule);
    if ((int)segs.size() > kMaxRuleSegs)
        return hci_fail(c, HCI_ERR_ARG, "hci_build_h: more than 8 alternations of the single rule across builds");
    SingleRules rules{};
    rules.n = (int32_t)segs.size();
    for (size_t k = 0; k < segs.size(); ++k) {
        rules.end[k] = (int32_t)segs[k].first;
        rules.rule[k] = segs[k].second;
    }
    // a tiled matrix is rebuilt whole: the pairs among the previous
    // determinants are found again and keep the rule they were stored under
    const bool tiled_inc = prev && opts->storage == HCI_STORE_TILED;
    auto t0 = std::chrono::steady_clock::now();
    HIP_TRY(c, hipSetDevice(c->device));
    if (opts->storage != HCI_STORE_TILED) recycled_flush(c);
    const uint64_t* d_dets = dets;
    const double* d_diag = diag;
    uint64_t* tmp_dets = nullptr;
    double* tmp_diag = nullptr;
    if (opts->mem == HCI_MEM_HOST) {
        HIP_TRY(c, hipMalloc(&tmp_dets, (size_t)n * 32));
        HIP_TRY(c, hipMalloc(&tmp_diag, (size_t)n * 8));
        HIP_TRY(c, hipMemcpyAsync(tmp_dets, dets, (size_t)n * 32, hipMemcpyHostToDevice, c->stream));
        HIP_TRY(c, hipMemcpyAsync(tmp_diag, diag, (size_t)n * 8, hipMemcpyHostToDevice, c->stream));
        d_dets = tmp_dets;
        d_diag = tmp_diag;
    }
    const hci_csr* pv = tiled_inc ? nullptr : prev;
    const int64_t ns = tiled_inc ? 0 : n_stored;
    int rc = c->n_orb <= 64
                 ? build_impl<uint64_t>(c, d_dets, d_diag, n, pv, ns, *opts, rules, out, stats)
                 : build_impl<u128>(c, d_dets, d_diag, n, pv, ns, *opts, rules, out, stats);
    if (rc == HCI_OK && *out) (*out)->rule_segs = segs;
    hipStreamSynchronize(c->stream);
    hipFree(tmp_dets);
    hipFree(tmp_diag);
    if (stats)
        stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return rc;
}
