// Heat-bath selection of new determinants on the GPU.
//
// Lists (reference src/excite/init.rs:40-249), built once per integrals on the
// host with the reference's stable sorts and summation order:
//   opposite-spin doubles, for every ordered valence pair (p, q): targets
//     (r, s), r != p, s != q, sorted by |(pr|qs)| descending;
//   same-spin doubles, for every valence pair p < q: targets r < s outside
//     {p, q}, sorted by |(pr|qs) - (ps|qr)| descending;
//   singles, for every valence p: targets r != p sorted by the bound
//     max(|h_pr + smallest partial sums|, |h_pr + largest partial sums|) over
//     the n_up - 1 same-spin terms (pr|qq) - (pq|qr) and the n_dn
//     opposite-spin terms (pr|qq).
// Core orbitals are the n_core lowest by h_pp (src/ham/read_ints.rs:161-184).
//
// Selection (src/wf/mod.rs:1467-1631): one thread per determinant D with
// local eps = eps / |c_D| walks, in the reference's order, its opposite-spin
// pairs (i up, j dn ascending), its same-spin pairs (up then dn) and, when
// asked, its singles, each list until the stored magnitude drops below local
// eps; valid targets become candidates, singles only if the exact element
// (elem_single, the same device function as every other element) exceeds local
// eps.  A count pass sizes the candidate array; the fill pass writes the
// candidates in discovery order (determinant-major).  Duplicates and members
// of the current space are removed by stable radix sorts on the determinant
// words: the first occurrence of each new determinant survives, and the
// survivors are returned in discovery order, which is the reference's order
// of new determinants.
//
// Deterministic PT2 (dtm_pt, src/pt/mod.rs:81-122, candidates from
// dets_excites_and_excited_dets, src/excite/iterator.rs:33-73): the same
// lists walked with the reference's test |stored magnitude| * |c_i| >= eps
// (singles then need |H_ai c_i| >= eps), targets inside the space dropped,
// contributions H_ai c_i summed per external determinant in discovery order
// (stable sort by determinant words, then one pass per run), and
// e2 = sum_a (sum_i H_ai c_i)^2 / (e0 - H_aa).  When the candidates exceed a
// budget, external determinants are split into parts by a hash of their words
// and every part walks all source determinants keeping only its own, so each
// external determinant is summed whole, in discovery order, within one part.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <string>
#include <type_traits>
#include <rocprim/rocprim.hpp>

#include "hci_internal.h"

struct SelectLists {
    int n_core = 0, n_up = 0, n_dn = 0, nv = 0;
    int64_t Lo = 0, Ls = 0, Lg = 0;  // list lengths: opposite doubles, same doubles, singles
    double max_opp = 0.0, max_same = 0.0, max_sing = 0.0, eps_start = 0.0;
    uint64_t vmask[2] = {0, 0};
    double *d_oh = nullptr, *d_sh = nullptr, *d_gh = nullptr;
    // sum of the squared magnitudes after each entry of its list
    // (sum_remaining_h_squared, src/excite/init.rs:282-309)
    double *d_or = nullptr, *d_sr = nullptr, *d_gr = nullptr;
    // sum of the magnitudes after each entry (sum_remaining_abs_h)
    double *d_oa = nullptr, *d_sa = nullptr, *d_ga = nullptr;
    uint16_t *d_ot = nullptr, *d_st = nullptr, *d_gt = nullptr;  // targets r | s << 8
    int32_t *d_vpos = nullptr, *d_spos = nullptr;
    ~SelectLists() {
        hipFree(d_oh);
        hipFree(d_sh);
        hipFree(d_gh);
        hipFree(d_or);
        hipFree(d_sr);
        hipFree(d_gr);
        hipFree(d_oa);
        hipFree(d_sa);
        hipFree(d_ga);
        hipFree(d_ot);
        hipFree(d_st);
        hipFree(d_gt);
        hipFree(d_vpos);
        hipFree(d_spos);
    }
};

void select_free(hci_ctx* c) {
    if (c && c->sel) {
        delete c->sel;
        c->sel = nullptr;
    }
}

namespace {

inline size_t sc2(size_t i, size_t j) { return i > j ? (i * (i - 1)) / 2 + j : (j * (j - 1)) / 2 + i; }

struct DevSel {
    int nv;
    int64_t Lo, Ls, Lg;
    double max_opp, max_same, max_sing;
    const double *oh, *sh, *gh;
    const uint16_t *ot, *st, *gt;
    const int32_t *vpos, *spos;
    const double *orem, *srem, *grem;
    const double *oabs, *sabs, *gabs;
};

DevSel dev_sel(const SelectLists& L) {
    return DevSel{L.nv,   L.Lo,   L.Ls,   L.Lg,     L.max_opp, L.max_same, L.max_sing, L.d_oh, L.d_sh,
                  L.d_gh, L.d_ot, L.d_st, L.d_gt,   L.d_vpos,  L.d_spos,   L.d_or,     L.d_sr, L.d_gr,
                  L.d_oa, L.d_sa, L.d_ga};
}

struct NoBreak {
    __device__ void operator()(const double*, int64_t) const {}
};

template <class S>
__device__ __forceinline__ S bit(int p) {
    return ((S)1) << p;
}

// Excitation code of a candidate: bits 0-1 kind (0 opposite-spin double p up,
// q dn -> r up, s dn; 1 same-spin double p < q -> r < s; 2 single p -> r),
// bit 2 alpha (kinds 1, 2), then p, q, r, s in 7 bits each from bit 3.
__device__ __forceinline__ uint32_t exc_code(int kind, int alpha, int p, int q, int r, int s) {
    return (uint32_t)kind | ((uint32_t)alpha << 2) | ((uint32_t)p << 3) | ((uint32_t)q << 10) | ((uint32_t)r << 17) |
           ((uint32_t)s << 24);
}

// Walks one determinant's excitations in the reference's order; emit(up, dn,
// code) receives every candidate.  brk(rem, e) is told where a list was cut
// (rem = that list's sum_remaining_h_squared array, e = the first entry below
// local).  A single is kept when |H| > local (find_new_dets), or with
// singles_ge when |H| >= local (approx_matmul_external_dtm_only).
template <class S, class Emit, class Brk = NoBreak>
__device__ void walk(const DevSel& L, const DevInts& T, S vmask, S up, S dn, double local, bool singles, Emit emit,
                     Brk brk = Brk(), bool singles_ge = false) {
    if (L.max_opp >= local) {
        for (S a = vmask & up; a; a &= a - (S)1) {
            const int i = ctz(a);
            for (S b = vmask & dn; b; b &= b - (S)1) {
                const int j = ctz(b);
                const int64_t base = ((int64_t)L.vpos[i] * L.nv + L.vpos[j]) * L.Lo;
                for (int64_t e = 0; e < L.Lo; ++e) {
                    if (L.oh[base + e] < local) {
                        brk(L.orem, base + e);
                        break;
                    }
                    const uint16_t t = L.ot[base + e];
                    const int r = t & 0xFF, s = t >> 8;
                    if (((up >> r) & (S)1) || ((dn >> s) & (S)1)) continue;
                    emit((up ^ bit<S>(i)) | bit<S>(r), (dn ^ bit<S>(j)) | bit<S>(s), exc_code(0, 0, i, j, r, s));
                }
            }
        }
    }
    if (L.max_same >= local && L.Ls > 0) {
        for (int spin = 0; spin < 2; ++spin) {
            const S cfg = spin == 0 ? up : dn;
            for (S a = vmask & cfg; a; a &= a - (S)1) {
                const int i = ctz(a);
                for (S b = a & (a - (S)1); b; b &= b - (S)1) {
                    const int j = ctz(b);
                    const int64_t base = (int64_t)L.spos[L.vpos[i] * L.nv + L.vpos[j]] * L.Ls;
                    for (int64_t e = 0; e < L.Ls; ++e) {
                        if (L.sh[base + e] < local) {
                            brk(L.srem, base + e);
                            break;
                        }
                        const uint16_t t = L.st[base + e];
                        const int r = t & 0xFF, s = t >> 8;
                        if (((cfg >> r) & (S)1) || ((cfg >> s) & (S)1)) continue;
                        const S nw = (cfg ^ bit<S>(i) ^ bit<S>(j)) | bit<S>(r) | bit<S>(s);
                        const uint32_t code = exc_code(1, spin == 0, i, j, r, s);
                        if (spin == 0)
                            emit(nw, dn, code);
                        else
                            emit(up, nw, code);
                    }
                }
            }
        }
    }
    if (singles && L.max_sing >= local) {
        for (int spin = 0; spin < 2; ++spin) {
            const S cfg = spin == 0 ? up : dn, other = spin == 0 ? dn : up;
            for (S a = vmask & cfg; a; a &= a - (S)1) {
                const int i = ctz(a);
                const int64_t base = (int64_t)L.vpos[i] * L.Lg;
                for (int64_t e = 0; e < L.Lg; ++e) {
                    if (L.gh[base + e] < local) {
                        brk(L.grem, base + e);
                        break;
                    }
                    const int r = L.gt[base + e] & 0xFF;
                    if ((cfg >> r) & (S)1) continue;
                    const S nw = (cfg ^ bit<S>(i)) | bit<S>(r);
                    // ham_sing(D, new) (src/ham/mod.rs:135-168)
                    const double hs = fabs(elem_single<S>(T, cfg, nw, other));
                    if (singles_ge ? hs >= local : hs > local) {
                        const uint32_t code = exc_code(2, spin == 0, i, 0, r, 0);
                        if (spin == 0)
                            emit(nw, dn, code);
                        else
                            emit(up, nw, code);
                    }
                }
            }
        }
    }
}

// dtm_pt's candidates of one determinant with coefficient c, in the order of
// valence_elecs_and_epairs (src/utils/bits.rs:76-85): opposite-spin pairs,
// same-spin pairs (up, then dn), single electrons (up, then dn); each list
// taken while |stored magnitude| * |c| >= eps (src/excite/iterator.rs:44-46).
// emit(up', dn', H_ai * c) for valid targets (doubles always, singles when
// |H_ai * c| >= eps, src/pt/mod.rs:97-102).
template <class S, class Emit>
__device__ void walk_pt(const DevSel& L, const DevInts& T, S vmask, S up, S dn, double c, double eps, Emit emit) {
    const double ac = fabs(c);
    if (L.max_opp * ac >= eps) {
        for (S a = vmask & up; a; a &= a - (S)1) {
            const int i = ctz(a);
            for (S b = vmask & dn; b; b &= b - (S)1) {
                const int j = ctz(b);
                const int64_t base = ((int64_t)L.vpos[i] * L.nv + L.vpos[j]) * L.Lo;
                for (int64_t e = 0; e < L.Lo; ++e) {
                    if (!(L.oh[base + e] * ac >= eps)) break;
                    const uint16_t t = L.ot[base + e];
                    const int r = t & 0xFF, s = t >> 8;
                    if (((up >> r) & (S)1) || ((dn >> s) & (S)1)) continue;
                    const S nu = (up ^ bit<S>(i)) | bit<S>(r), nd = (dn ^ bit<S>(j)) | bit<S>(s);
                    emit(nu, nd, elem_opp_double<S>(T, up, nu, dn, nd) * c);
                }
            }
        }
    }
    if (L.max_same * ac >= eps && L.Ls > 0) {
        for (int spin = 0; spin < 2; ++spin) {
            const S cfg = spin == 0 ? up : dn;
            for (S a = vmask & cfg; a; a &= a - (S)1) {
                const int i = ctz(a);
                for (S b = a & (a - (S)1); b; b &= b - (S)1) {
                    const int j = ctz(b);
                    const int64_t base = (int64_t)L.spos[L.vpos[i] * L.nv + L.vpos[j]] * L.Ls;
                    for (int64_t e = 0; e < L.Ls; ++e) {
                        if (!(L.sh[base + e] * ac >= eps)) break;
                        const uint16_t t = L.st[base + e];
                        const int r = t & 0xFF, s = t >> 8;
                        if (((cfg >> r) & (S)1) || ((cfg >> s) & (S)1)) continue;
                        const S nw = (cfg ^ bit<S>(i) ^ bit<S>(j)) | bit<S>(r) | bit<S>(s);
                        const double h = elem_same_double<S>(T, cfg, nw) * c;
                        if (spin == 0)
                            emit(nw, dn, h);
                        else
                            emit(up, nw, h);
                    }
                }
            }
        }
    }
    if (L.max_sing * ac >= eps) {
        for (int spin = 0; spin < 2; ++spin) {
            const S cfg = spin == 0 ? up : dn, other = spin == 0 ? dn : up;
            for (S a = vmask & cfg; a; a &= a - (S)1) {
                const int i = ctz(a);
                const int64_t base = (int64_t)L.vpos[i] * L.Lg;
                for (int64_t e = 0; e < L.Lg; ++e) {
                    if (!(L.gh[base + e] * ac >= eps)) break;
                    const int r = L.gt[base + e] & 0xFF;
                    if ((cfg >> r) & (S)1) continue;
                    const S nw = (cfg ^ bit<S>(i)) | bit<S>(r);
                    const double h = elem_single<S>(T, cfg, nw, other) * c;
                    if (!(fabs(h) >= eps)) continue;
                    if (spin == 0)
                        emit(nw, dn, h);
                    else
                        emit(up, nw, h);
                }
            }
        }
    }
}

template <class S>
__device__ __forceinline__ void store_string(uint64_t* w, S x);
template <>
__device__ __forceinline__ void store_string<uint64_t>(uint64_t* w, uint64_t x) {
    w[0] = x;
    w[1] = 0;
}
template <>
__device__ __forceinline__ void store_string<u128>(uint64_t* w, u128 x) {
    w[0] = (uint64_t)x;
    w[1] = (uint64_t)(x >> 64);
}

template <class S, bool FILL>
__global__ void k_select(DevSel L, DevInts T, uint64_t vm0, uint64_t vm1, const uint64_t* __restrict__ dets,
                         const double* __restrict__ coeffs, int64_t n, double eps, int singles, int64_t* counts,
                         const int64_t* __restrict__ offs, uint64_t* out, int64_t* src, uint32_t* codes) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const S up = load_string<S>(dets + 4 * k), dn = load_string<S>(dets + 4 * k + 2);
    S vmask;
    if constexpr (sizeof(S) > 8)
        vmask = ((S)vm1 << 64) | (S)vm0;
    else
        vmask = (S)vm0;
    const double c = fabs(coeffs[k]);
    const double local = c > 0.0 ? eps / c : INFINITY;
    int64_t cnt = 0;
    uint64_t* o = FILL ? out + 4 * offs[k] : nullptr;
    walk<S>(L, T, vmask, up, dn, local, singles != 0, [&](S nu, S nd, uint32_t code) {
        if (FILL) {
            store_string<S>(o + 4 * cnt, nu);
            store_string<S>(o + 4 * cnt + 2, nd);
            if (src) {
                src[offs[k] + cnt] = k;
                codes[offs[k] + cnt] = code;
            }
        }
        ++cnt;
    });
    if (!FILL) counts[k] = cnt;
}

// part of an external determinant: a 64-bit mix of its strings, mod n_parts
template <class S>
__device__ __forceinline__ uint32_t part_of(S up, S dn, uint32_t n_parts) {
    uint64_t h = (uint64_t)up * 0x9E3779B97F4A7C15ull ^ (uint64_t)dn;
    if constexpr (sizeof(S) > 8) h ^= (uint64_t)(up >> 64) * 0xC2B2AE3D27D4EB4Full ^ (uint64_t)(dn >> 64) * 0x165667B19E3779F9ull;
    h ^= h >> 31;
    h *= 0xBF58476D1CE4E5B9ull;
    h ^= h >> 29;
    return (uint32_t)(h % n_parts);
}

// Det::new_diag_opp / new_diag_same / new_diag_sing (src/wf/det.rs:271-420):
// the candidate's diagonal from its parent's, in the reference's order of
// operations (direct(i,r,i,r) = (ii|rr) = J3, direct_plus_exchange(i,r,i,r) =
// (ii|rr) - (ir|ri) = D3, the same IEEE values the reference reads).
template <class S>
__device__ double new_diag(const DevInts& T, S up, S dn, double diag, uint32_t code) {
    const int n = T.n;
    const int kind = code & 3, alpha = (code >> 2) & 1;
    const int p = (code >> 3) & 127, q = (code >> 10) & 127, r = (code >> 17) & 127, s = (code >> 24) & 127;
    auto J = [&](int i, int k) { return T.J3[(size_t)(i * n + i) * n + k]; };
    auto Dx = [&](int i, int k) { return T.D3[(size_t)(i * n + i) * n + k]; };
    auto h = [&](int i) { return T.h1[i * n + i]; };
    if (kind == 2) {
        double out = diag + h(r) - h(p);
        const S same = alpha ? up : dn, other = alpha ? dn : up;
        for (S m = same; m; m &= m - (S)1) {
            const int i = ctz(m);
            if (i == p) continue;
            out += Dx(i, r) - Dx(i, p);
        }
        for (S m = other; m; m &= m - (S)1) {
            const int i = ctz(m);
            out += J(i, r) - J(i, p);
        }
        return out;
    }
    double out = diag + h(r) + h(s) - h(p) - h(q);
    if (kind == 0) {
        out += J(r, s) - J(p, q);
        for (S m = up; m; m &= m - (S)1) {
            const int i = ctz(m);
            if (i == p) continue;
            out += Dx(i, r) - Dx(i, p);
            out += J(i, s) - J(i, q);
        }
        for (S m = dn; m; m &= m - (S)1) {
            const int i = ctz(m);
            if (i == q) continue;
            out += Dx(i, s) - Dx(i, q);
            out += J(i, r) - J(i, p);
        }
        return out;
    }
    out += Dx(r, s) - Dx(p, q);
    const S same = alpha ? up : dn, other = alpha ? dn : up;
    for (S m = same; m; m &= m - (S)1) {
        const int i = ctz(m);
        if (i == p || i == q) continue;
        out += Dx(i, r) + Dx(i, s) - Dx(i, p) - Dx(i, q);
    }
    for (S m = other; m; m &= m - (S)1) {
        const int i = ctz(m);
        out += J(i, r) + J(i, s) - J(i, p) - J(i, q);
    }
    return out;
}

// diagonals of the surviving candidates (rank: candidate index per survivor)
template <class S>
__global__ void k_new_diag(DevInts T, const uint64_t* __restrict__ dets, const double* __restrict__ diag,
                           const int64_t* __restrict__ src, const uint32_t* __restrict__ codes,
                           const int64_t* __restrict__ rank, int64_t k, double* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= k) return;
    const int64_t c = rank[i], j = src[c];
    out[i] = new_diag<S>(T, load_string<S>(dets + 4 * j), load_string<S>(dets + 4 * j + 2), diag[j], codes[c]);
}

template <class S, bool FILL>
__global__ void k_pt_walk(DevSel L, DevInts T, uint64_t vm0, uint64_t vm1, const uint64_t* __restrict__ dets,
                          const double* __restrict__ coeffs, int64_t n, double eps, uint32_t n_parts, uint32_t part,
                          int64_t* counts, const int64_t* __restrict__ offs, uint64_t* words, double* vals) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const S up = load_string<S>(dets + 4 * k), dn = load_string<S>(dets + 4 * k + 2);
    S vmask;
    if constexpr (sizeof(S) > 8)
        vmask = ((S)vm1 << 64) | (S)vm0;
    else
        vmask = (S)vm0;
    int64_t cnt = 0;
    const int64_t o = FILL ? offs[k] : 0;
    walk_pt<S>(L, T, vmask, up, dn, coeffs[k], eps, [&](S nu, S nd, double h) {
        if (n_parts > 1 && part_of<S>(nu, nd, n_parts) != part) return;
        if (FILL) {
            store_string<S>(words + 4 * (o + cnt), nu);
            store_string<S>(words + 4 * (o + cnt) + 2, nd);
            vals[o + cnt] = h;
        }
        ++cnt;
    });
    if (!FILL) counts[k] = cnt;
}

// approx_matmul_external_dtm_only (src/wf/mod.rs:186-339), one thread per
// source determinant k with local = eps / |c_k|: the selection walk (singles
// kept when |H| >= local), each candidate with H_ak c_k, its source and its
// excitation code; the count pass also writes c_k^2 times the sum of
// sum_remaining_h_squared over the lists it cut (out_sum_remaining).
template <class S, bool FILL>
__global__ void k_ext_walk(DevSel L, DevInts T, uint64_t vm0, uint64_t vm1, const uint64_t* __restrict__ dets,
                           const double* __restrict__ coeffs, int64_t n, double eps, int64_t* counts,
                           double* sum_rem, const int64_t* __restrict__ offs, uint64_t* words, double* vals,
                           int64_t* src, uint32_t* codes) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const S up = load_string<S>(dets + 4 * k), dn = load_string<S>(dets + 4 * k + 2);
    S vmask;
    if constexpr (sizeof(S) > 8)
        vmask = ((S)vm1 << 64) | (S)vm0;
    else
        vmask = (S)vm0;
    const double c = coeffs[k];
    const double local = eps / fabs(c);  // +inf for c = 0: nothing passes
    int64_t cnt = 0;
    double rem = 0.0;
    const int64_t o = FILL ? offs[k] : 0;
    walk<S>(
        L, T, vmask, up, dn, local, true,
        [&](S nu, S nd, uint32_t code) {
            if (FILL) {
                const int kind = code & 3, alpha = (code >> 2) & 1;
                double h;
                if (kind == 0)
                    h = elem_opp_double<S>(T, up, nu, dn, nd);
                else if (kind == 1)
                    h = alpha ? elem_same_double<S>(T, up, nu) : elem_same_double<S>(T, dn, nd);
                else
                    h = alpha ? elem_single<S>(T, up, nu, dn) : elem_single<S>(T, dn, nd, up);
                store_string<S>(words + 4 * (o + cnt), nu);
                store_string<S>(words + 4 * (o + cnt) + 2, nd);
                vals[o + cnt] = h * c;
                src[o + cnt] = k;
                codes[o + cnt] = code;
            }
            ++cnt;
        },
        [&](const double* r, int64_t e) {
            if (!FILL) rem += r[e];
        },
        true);
    if (!FILL) {
        counts[k] = cnt;
        if (sum_rem) sum_rem[k] = c * c * rem;
    }
}

// kept run u (kpos[u] its rank among them): its first discovery, the
// candidate index of its head in the stable word order
__global__ void k_run_first(const int64_t* cidx, const int64_t* start, const int64_t* keep, const int64_t* kpos,
                            int64_t runs, int64_t* first, int64_t* run_id) {
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= runs || !keep[u]) return;
    first[kpos[u]] = cidx[start[u]];
    run_id[kpos[u]] = u;
}

// output i (in first-discovery order): the run's determinant, its summed
// coefficient and the diagonal its first discoverer gives it (new_diag)
template <class S>
__global__ void k_ext_out(DevInts T, const uint64_t* __restrict__ dets, const double* __restrict__ diag,
                          const int64_t* __restrict__ src, const uint32_t* __restrict__ codes,
                          const int64_t* __restrict__ first, const int64_t* __restrict__ run_id,
                          const uint64_t* __restrict__ rw, const double* __restrict__ rv, int64_t k,
                          uint64_t* out_w, double* out_v, double* out_d) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= k) return;
    const int64_t u = run_id[i], f = first[i], j = src[f];
    for (int q = 0; q < 4; ++q) out_w[4 * i + q] = rw[4 * u + q];
    out_v[i] = rv[u];
    if (out_d)
        out_d[i] = new_diag<S>(T, load_string<S>(dets + 4 * j), load_string<S>(dets + 4 * j + 2), diag[j], codes[f]);
}

__global__ void k_iota64(int64_t* v, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = i;
}

__global__ void k_word(const uint64_t* w, const int64_t* idx, int64_t n, int word, uint64_t* key) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) key[i] = w[4 * idx[i] + word];
}

// lexicographic on words 3, 2, 1, 0 (the order the LSD passes below produce)
__device__ __forceinline__ int cmp_det(const uint64_t* a, const uint64_t* b) {
    for (int w = 3; w >= 0; --w)
        if (a[w] != b[w]) return a[w] < b[w] ? -1 : 1;
    return 0;
}

// keep[i] = 1 iff sorted candidate i is the first of its run and absent from the
// sorted current space
__global__ void k_keep(const uint64_t* cand, const int64_t* cidx, int64_t m, const uint64_t* cur, const int64_t* didx,
                       int64_t n, int64_t* keep) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const uint64_t* a = cand + 4 * cidx[i];
    int64_t k = 1;
    if (i > 0 && cmp_det(a, cand + 4 * cidx[i - 1]) == 0) k = 0;
    if (k) {
        int64_t lo = 0, hi = n;
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (cmp_det(cur + 4 * didx[mid], a) < 0)
                lo = mid + 1;
            else
                hi = mid;
        }
        if (lo < n && cmp_det(cur + 4 * didx[lo], a) == 0) k = 0;
    }
    keep[i] = k;
}

__global__ void k_compact(const int64_t* keep, const int64_t* pos, const int64_t* cidx, int64_t m, int64_t* rank) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m && keep[i]) rank[pos[i]] = cidx[i];
}

__global__ void k_gather_dets(const uint64_t* cand, const int64_t* rank, int64_t k, uint64_t* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= k) return;
    for (int w = 0; w < 4; ++w) out[4 * i + w] = cand[4 * rank[i] + w];
}

// head[i] = 1 iff sorted entry i starts a run of equal determinants
__global__ void k_run_heads(const uint64_t* w, const int64_t* idx, int64_t m, int64_t* head) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) head[i] = (i == 0 || cmp_det(w + 4 * idx[i], w + 4 * idx[i - 1]) != 0) ? 1 : 0;
}

__global__ void k_run_starts(const int64_t* head, const int64_t* pos, int64_t m, int64_t* start) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m && head[i]) start[pos[i]] = i;
    if (i == m) start[pos[m]] = m;
}

// One thread per run: the run's contributions summed in sorted (= discovery)
// order; a determinant of the space (sorted by sidx) gets keep = 0.
__global__ void k_run_sum(const uint64_t* w, const double* v, const int64_t* idx, const int64_t* start, int64_t runs,
                          const uint64_t* space, const int64_t* sidx, int64_t n, uint64_t* out_w, double* out_v,
                          int64_t* keep) {
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= runs) return;
    const int64_t a = start[u], b = start[u + 1];
    double sum = 0.0;
    for (int64_t i = a; i < b; ++i) sum += v[idx[i]];
    const uint64_t* d = w + 4 * idx[a];
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (cmp_det(space + 4 * sidx[mid], d) < 0)
            lo = mid + 1;
        else
            hi = mid;
    }
    keep[u] = (lo < n && cmp_det(space + 4 * sidx[lo], d) == 0) ? 0 : 1;
    for (int q = 0; q < 4; ++q) out_w[4 * u + q] = d[q];
    out_v[u] = sum;
}

// term_a = num_a^2 / (e0 - H_aa) for kept runs; fixed-order partial sums, one per block
constexpr int kPtBlocks = 1024;
template <class S>
__global__ void __launch_bounds__(256)
    k_pt_terms(DevInts T, const uint64_t* w, const double* num, const int64_t* keep, int64_t m, double e0,
               double* partial) {
    __shared__ double sh[256];
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        if (!keep[i]) continue;
        const S up = load_string<S>(w + 4 * i), dn = load_string<S>(w + 4 * i + 2);
        acc += num[i] * num[i] / (e0 - elem_diag<S>(T, up, dn));
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}


// membership of (up, dn) in the sorted space (didx: the space's indices in
// cmp_det order)
template <class S>
__device__ __forceinline__ bool in_space(S up, S dn, const uint64_t* space, const int64_t* didx, int64_t n) {
    uint64_t a[4];
    store_string<S>(a, up);
    store_string<S>(a + 2, dn);
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (cmp_det(space + 4 * didx[mid], a) < 0)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo < n && cmp_det(space + 4 * didx[lo], a) == 0;
}

// The deterministic stage of new_semistoch_enpt2_dtm_diag_singles
// (src/semistoch/mod.rs:262-318) for one determinant with coefficient c:
// every list of its valence pairs and electrons, in the order of
// valence_elecs_and_epairs, read through excites_from_det_and_orbs (targets
// already occupied or inside the space are skipped first,
// src/excite/iterator.rs:113-130).  The first remaining entry with
// |stored magnitude| * |c| < eps ends the list: sample(kind, alpha, p, q,
// sum_remaining_abs_h, sum_remaining_h_squared) of that entry.  Each entry
// before it: emit(up', dn', H_ai c, code, in_psi), in_psi false for a single
// with |H_ai c| < eps.  Targets inside the space above the cut are emitted too
// (the caller's sorted pass drops them); only the cut test needs membership.
template <class S, class Emit, class Sample>
__device__ void walk_sd(const DevSel& L, const DevInts& T, S vmask, S up, S dn, double c, double eps,
                        const uint64_t* space, const int64_t* didx, int64_t n, Emit emit, Sample sample) {
    const double ac = fabs(c);
    for (S a = vmask & up; a; a &= a - (S)1) {
        const int i = ctz(a);
        for (S b = vmask & dn; b; b &= b - (S)1) {
            const int j = ctz(b);
            const int64_t base = ((int64_t)L.vpos[i] * L.nv + L.vpos[j]) * L.Lo;
            for (int64_t e = 0; e < L.Lo; ++e) {
                const uint16_t t = L.ot[base + e];
                const int r = t & 0xFF, s = t >> 8;
                if (((up >> r) & (S)1) || ((dn >> s) & (S)1)) continue;
                const S nu = (up ^ bit<S>(i)) | bit<S>(r), nd = (dn ^ bit<S>(j)) | bit<S>(s);
                if (L.oh[base + e] * ac < eps) {
                    if (in_space<S>(nu, nd, space, didx, n)) continue;
                    sample(0, -1, i, j, L.oabs[base + e], L.orem[base + e]);
                    break;
                }
                emit(nu, nd, elem_opp_double<S>(T, up, nu, dn, nd) * c, exc_code(0, 0, i, j, r, s), true);
            }
        }
    }
    if (L.Ls > 0) {
        for (int spin = 0; spin < 2; ++spin) {
            const S cfg = spin == 0 ? up : dn;
            for (S a = vmask & cfg; a; a &= a - (S)1) {
                const int i = ctz(a);
                for (S b = a & (a - (S)1); b; b &= b - (S)1) {
                    const int j = ctz(b);
                    const int64_t base = (int64_t)L.spos[L.vpos[i] * L.nv + L.vpos[j]] * L.Ls;
                    for (int64_t e = 0; e < L.Ls; ++e) {
                        const uint16_t t = L.st[base + e];
                        const int r = t & 0xFF, s = t >> 8;
                        if (((cfg >> r) & (S)1) || ((cfg >> s) & (S)1)) continue;
                        const S nw = (cfg ^ bit<S>(i) ^ bit<S>(j)) | bit<S>(r) | bit<S>(s);
                        const S nu = spin == 0 ? nw : up, nd = spin == 0 ? dn : nw;
                        if (L.sh[base + e] * ac < eps) {
                            if (in_space<S>(nu, nd, space, didx, n)) continue;
                            sample(1, spin == 0, i, j, L.sabs[base + e], L.srem[base + e]);
                            break;
                        }
                        emit(nu, nd, elem_same_double<S>(T, cfg, nw) * c, exc_code(1, spin == 0, i, j, r, s), true);
                    }
                }
            }
        }
    }
    for (int spin = 0; spin < 2; ++spin) {
        const S cfg = spin == 0 ? up : dn, other = spin == 0 ? dn : up;
        for (S a = vmask & cfg; a; a &= a - (S)1) {
            const int i = ctz(a);
            const int64_t base = (int64_t)L.vpos[i] * L.Lg;
            for (int64_t e = 0; e < L.Lg; ++e) {
                const int r = L.gt[base + e] & 0xFF;
                if ((cfg >> r) & (S)1) continue;
                const S nw = (cfg ^ bit<S>(i)) | bit<S>(r);
                const S nu = spin == 0 ? nw : up, nd = spin == 0 ? dn : nw;
                if (L.gh[base + e] * ac < eps) {
                    if (in_space<S>(nu, nd, space, didx, n)) continue;
                    sample(2, spin == 0, i, -1, L.gabs[base + e], L.grem[base + e]);
                    break;
                }
                const double h = elem_single<S>(T, cfg, nw, other) * c;
                emit(nu, nd, h, exc_code(2, spin == 0, i, 0, r, 0), fabs(h) >= eps);
            }
        }
    }
}

// candidate code bit 31: the candidate joins h_psi
constexpr uint32_t kInPsi = 1u << 31;

template <class S, bool FILL>
__global__ void k_sd_walk(DevSel L, DevInts T, uint64_t vm0, uint64_t vm1, const uint64_t* __restrict__ dets,
                          const double* __restrict__ coeffs, int64_t n, double eps, const int64_t* __restrict__ didx,
                          int64_t* counts, int64_t* scounts, const int64_t* __restrict__ offs,
                          const int64_t* __restrict__ soffs, uint64_t* words, double* vals, int64_t* src,
                          uint32_t* codes, hci_det_orb_sample* samples) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const S up = load_string<S>(dets + 4 * k), dn = load_string<S>(dets + 4 * k + 2);
    S vmask;
    if constexpr (sizeof(S) > 8)
        vmask = ((S)vm1 << 64) | (S)vm0;
    else
        vmask = (S)vm0;
    const double c = coeffs[k];
    int64_t cnt = 0, scnt = 0;
    const int64_t o = FILL ? offs[k] : 0, so = FILL && samples ? soffs[k] : 0;
    walk_sd<S>(
        L, T, vmask, up, dn, c, eps, dets, didx, n,
        [&](S nu, S nd, double h, uint32_t code, bool psi) {
            if (FILL) {
                store_string<S>(words + 4 * (o + cnt), nu);
                store_string<S>(words + 4 * (o + cnt) + 2, nd);
                vals[o + cnt] = h;
                src[o + cnt] = k;
                codes[o + cnt] = code | (psi ? kInPsi : 0u);
            }
            ++cnt;
        },
        [&](int kind, int alpha, int p, int q, double rabs, double rsq) {
            if (FILL && samples) {
                hci_det_orb_sample r;
                r.det = k;
                r.kind = kind;
                r.is_alpha = alpha;
                r.p = p;
                r.q = q;
                r.sum_abs_h = rabs;
                r.sum_h_squared = rsq;
                r.sum_abs_hc = fabs(c) * rabs;
                r.sum_hc_squared = c * c * rsq;
                samples[so + scnt] = r;
            }
            ++scnt;
        });
    if (!FILL) {
        counts[k] = cnt;
        scounts[k] = scnt;
    }
}

// One thread per run of equal targets (sorted, stable: discovery order within
// the run).  A target outside the space: its h_psi candidates summed in
// order, the first of them recorded (it creates the entry), and every
// candidate's |H_ai c_i| raised into rel_probs of its source; keep = outside
// the space and at least one h_psi candidate.
__global__ void k_sd_run(const uint64_t* w, const double* v, const int64_t* src, const uint32_t* codes,
                         const int64_t* idx, const int64_t* start, int64_t runs, const uint64_t* space,
                         const int64_t* sidx, int64_t n, double* num, int64_t* first, int64_t* keep, double* rel) {
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= runs) return;
    const int64_t a = start[u], b = start[u + 1];
    const uint64_t* d = w + 4 * idx[a];
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (cmp_det(space + 4 * sidx[mid], d) < 0)
            lo = mid + 1;
        else
            hi = mid;
    }
    const bool ext = !(lo < n && cmp_det(space + 4 * sidx[lo], d) == 0);
    double sum = 0.0;
    int64_t f = -1;
    if (ext) {
        for (int64_t i = a; i < b; ++i) {
            const int64_t ci = idx[i];
            const double h = v[ci];
            atomicMax((unsigned long long*)(rel + src[ci]), (unsigned long long)__double_as_longlong(fabs(h)));
            if (codes[ci] & kInPsi) {
                if (f < 0) f = ci;
                sum += h;
            }
        }
    }
    keep[u] = (ext && f >= 0) ? 1 : 0;
    num[u] = sum;
    first[u] = f;
}

__global__ void k_sd_first(const int64_t* first_run, const int64_t* keep, const int64_t* kpos, int64_t runs,
                           int64_t* first, int64_t* run_id) {
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= runs || !keep[u]) return;
    first[kpos[u]] = first_run[u];
    run_id[kpos[u]] = u;
}

// e2 terms in first-discovery order: num^2 / (e0 - H_aa), H_aa from the first
// discoverer's diagonal (Det::new_diag); fixed-order partial sums per block
template <class S>
__global__ void __launch_bounds__(256)
    k_sd_terms(DevInts T, const uint64_t* __restrict__ dets, const double* __restrict__ diag,
               const int64_t* __restrict__ src, const uint32_t* __restrict__ codes, const int64_t* __restrict__ first,
               const int64_t* __restrict__ run_id, const double* __restrict__ num, int64_t k, double e0,
               double* partial) {
    __shared__ double sh[256];
    double acc = 0.0;
    const int64_t per = (k + gridDim.x - 1) / gridDim.x;
    const int64_t lo = (int64_t)blockIdx.x * per, hi = min(k, lo + per);
    // each block sums a contiguous range, thread by thread in order of its stride
    for (int64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
        const int64_t f = first[i], j = src[f];
        const double d = new_diag<S>(T, load_string<S>(dets + 4 * j), load_string<S>(dets + 4 * j + 2), diag[j],
                                     codes[f] & ~kInPsi);
        const double x = num[run_id[i]];
        acc += x * x / (e0 - d);
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// The screened-sampler variants of the external product, one determinant
// with coefficient c and local = eps / |c|; lists in the order of
// valence_elecs_and_epairs.  emit(up', dn', code) receives every deterministic
// candidate (targets inside the space too: the sorted pass drops them);
// sample(kind, alpha, p, q, sum_remaining_abs_h, sum_remaining_h_squared,
// scaled) receives the DetOrbSample of a list (scaled: the singles' 1e-6
// factor).  Per variant:
//  * SEPARATE (approx_matmul_external_separate_doubles_and_singles,
//    src/wf/mod.rs:341-533): a kind's lists are walked when its maximum is
//    >= local; the first entry below local ends a list with a record, valid or
//    not; every valid entry before it is a candidate (singles with their
//    exact element, unscreened);
//  * SKIP_SINGLES (:535-711) and the doubles of SEMISTOCH_SINGLES (:714-1033):
//    every doubles list, invalid entries skipped first; entries >= local are
//    candidates, and the first valid entry below local whose target is outside
//    the space ends the list with a record (entries inside it are passed over);
//  * SEMISTOCH_SINGLES singles (when max_sing >= local): entries >= local
//    with a valid external target are candidates when |H_ai| >= local and
//    otherwise give the list's first record; below local, a list with a
//    record ends, and one without ends at its first valid external entry with a
//    record scaled by 1e-6;
//  * NO_SINGLES (:1035-1193): doubles as SEPARATE without the maxima; every
//    valence electron gives a record from its singles list's first entry.
template <class S, int V, class Emit, class Sample>
__device__ void walk_var(const DevSel& L, const DevInts& T, S vmask, S up, S dn, double local, const uint64_t* space,
                         const int64_t* didx, int64_t n, Emit emit, Sample sample) {
    // SEPARATE and NO_SINGLES cut a list at its first entry below local,
    // valid or not; SKIP_SINGLES and SEMISTOCH_SINGLES at the first valid
    // external one
    constexpr bool kCutFirst = V == HCI_EXT_SEPARATE || V == HCI_EXT_NO_SINGLES;
    if (V != HCI_EXT_SEPARATE || L.max_opp >= local) {
        for (S a = vmask & up; a; a &= a - (S)1) {
            const int i = ctz(a);
            for (S b = vmask & dn; b; b &= b - (S)1) {
                const int j = ctz(b);
                const int64_t base = ((int64_t)L.vpos[i] * L.nv + L.vpos[j]) * L.Lo;
                for (int64_t e = 0; e < L.Lo; ++e) {
                    const double h = L.oh[base + e];
                    if (kCutFirst && h < local) {
                        sample(0, -1, i, j, L.oabs[base + e], L.orem[base + e], false);
                        break;
                    }
                    const uint16_t t = L.ot[base + e];
                    const int r = t & 0xFF, s = t >> 8;
                    if (((up >> r) & (S)1) || ((dn >> s) & (S)1)) continue;
                    const S nu = (up ^ bit<S>(i)) | bit<S>(r), nd = (dn ^ bit<S>(j)) | bit<S>(s);
                    if (!kCutFirst && h < local) {
                        if (in_space<S>(nu, nd, space, didx, n)) continue;
                        sample(0, -1, i, j, L.oabs[base + e], L.orem[base + e], false);
                        break;
                    }
                    emit(nu, nd, exc_code(0, 0, i, j, r, s));
                }
            }
        }
    }
    if ((V != HCI_EXT_SEPARATE || L.max_same >= local) && L.Ls > 0) {
        for (int spin = 0; spin < 2; ++spin) {
            const S cfg = spin == 0 ? up : dn;
            for (S a = vmask & cfg; a; a &= a - (S)1) {
                const int i = ctz(a);
                for (S b = a & (a - (S)1); b; b &= b - (S)1) {
                    const int j = ctz(b);
                    const int64_t base = (int64_t)L.spos[L.vpos[i] * L.nv + L.vpos[j]] * L.Ls;
                    for (int64_t e = 0; e < L.Ls; ++e) {
                        const double h = L.sh[base + e];
                        if (kCutFirst && h < local) {
                            sample(1, spin == 0, i, j, L.sabs[base + e], L.srem[base + e], false);
                            break;
                        }
                        const uint16_t t = L.st[base + e];
                        const int r = t & 0xFF, s = t >> 8;
                        if (((cfg >> r) & (S)1) || ((cfg >> s) & (S)1)) continue;
                        const S nw = (cfg ^ bit<S>(i) ^ bit<S>(j)) | bit<S>(r) | bit<S>(s);
                        const S nu = spin == 0 ? nw : up, nd = spin == 0 ? dn : nw;
                        if (!kCutFirst && h < local) {
                            if (in_space<S>(nu, nd, space, didx, n)) continue;
                            sample(1, spin == 0, i, j, L.sabs[base + e], L.srem[base + e], false);
                            break;
                        }
                        emit(nu, nd, exc_code(1, spin == 0, i, j, r, s));
                    }
                }
            }
        }
    }
    if (V == HCI_EXT_SKIP_SINGLES) return;
    if (V != HCI_EXT_NO_SINGLES && L.max_sing < local) return;
    for (int spin = 0; spin < 2; ++spin) {
        const S cfg = spin == 0 ? up : dn, other = spin == 0 ? dn : up;
        for (S a = vmask & cfg; a; a &= a - (S)1) {
            const int i = ctz(a);
            const int64_t base = (int64_t)L.vpos[i] * L.Lg;
            if (V == HCI_EXT_NO_SINGLES) {
                sample(2, spin == 0, i, -1, L.gabs[base], L.grem[base], false);
                continue;
            }
            bool stored = false;
            for (int64_t e = 0; e < L.Lg; ++e) {
                const double h = L.gh[base + e];
                if (V == HCI_EXT_SEPARATE && h < local) {
                    sample(2, spin == 0, i, -1, L.gabs[base + e], L.grem[base + e], false);
                    break;
                }
                if (V == HCI_EXT_SEMISTOCH_SINGLES && h < local && stored) break;
                const int r = L.gt[base + e] & 0xFF;
                if ((cfg >> r) & (S)1) continue;
                const S nw = (cfg ^ bit<S>(i)) | bit<S>(r);
                const S nu = spin == 0 ? nw : up, nd = spin == 0 ? dn : nw;
                const uint32_t code = exc_code(2, spin == 0, i, 0, r, 0);
                if (V == HCI_EXT_SEPARATE) {
                    emit(nu, nd, code);
                    continue;
                }
                if (h >= local) {
                    // ham_sing(D, new) (src/ham/mod.rs:135-168)
                    if (fabs(elem_single<S>(T, cfg, nw, other)) >= local) {
                        emit(nu, nd, code);
                    } else if (!stored && !in_space<S>(nu, nd, space, didx, n)) {
                        sample(2, spin == 0, i, -1, L.gabs[base + e], L.grem[base + e], false);
                        stored = true;
                    }
                } else if (!in_space<S>(nu, nd, space, didx, n)) {
                    sample(2, spin == 0, i, -1, L.gabs[base + e], L.grem[base + e], true);
                    break;
                }
            }
        }
    }
}

template <class S, int V, bool FILL>
__global__ void k_extv_walk(DevSel L, DevInts T, uint64_t vm0, uint64_t vm1, const uint64_t* __restrict__ dets,
                            const double* __restrict__ coeffs, int64_t n, double eps, const int64_t* __restrict__ didx,
                            int64_t* counts, int64_t* scounts, const int64_t* __restrict__ offs,
                            const int64_t* __restrict__ soffs, uint64_t* words, double* vals, int64_t* src,
                            uint32_t* codes, hci_det_orb_sample* samples) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const S up = load_string<S>(dets + 4 * k), dn = load_string<S>(dets + 4 * k + 2);
    S vmask;
    if constexpr (sizeof(S) > 8)
        vmask = ((S)vm1 << 64) | (S)vm0;
    else
        vmask = (S)vm0;
    const double c = coeffs[k];
    const double local = eps / fabs(c);  // +inf for c = 0: every magnitude is below it
    int64_t cnt = 0, scnt = 0;
    const int64_t o = FILL ? offs[k] : 0, so = FILL ? soffs[k] : 0;
    walk_var<S, V>(
        L, T, vmask, up, dn, local, dets, didx, n,
        [&](S nu, S nd, uint32_t code) {
            if (FILL) {
                const int kind = code & 3, alpha = (code >> 2) & 1;
                double h;
                if (kind == 0)
                    h = elem_opp_double<S>(T, up, nu, dn, nd);
                else if (kind == 1)
                    h = alpha ? elem_same_double<S>(T, up, nu) : elem_same_double<S>(T, dn, nd);
                else
                    h = alpha ? elem_single<S>(T, up, nu, dn) : elem_single<S>(T, dn, nd, up);
                store_string<S>(words + 4 * (o + cnt), nu);
                store_string<S>(words + 4 * (o + cnt) + 2, nd);
                vals[o + cnt] = h * c;
                src[o + cnt] = k;
                codes[o + cnt] = code;
            }
            ++cnt;
        },
        [&](int kind, int alpha, int p, int q, double rabs, double rsq, bool scaled) {
            if (FILL) {
                // single_sample_factor (src/wf/mod.rs:895), applied last as the reference does
                constexpr double f = 1e-6;
                hci_det_orb_sample r;
                r.det = k;
                r.kind = kind;
                r.is_alpha = alpha;
                r.p = p;
                r.q = q;
                r.sum_abs_h = scaled ? rabs * f : rabs;
                r.sum_h_squared = scaled ? rsq * f : rsq;
                r.sum_abs_hc = scaled ? fabs(c) * rabs * f : fabs(c) * rabs;
                r.sum_hc_squared = scaled ? c * c * rsq * f : c * c * rsq;
                samples[so + scnt] = r;
            }
            ++scnt;
        });
    if (!FILL) {
        counts[k] = cnt;
        scounts[k] = scnt;
    }
}

struct Tmp {
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
    ~Tmp() { hipFree(p); }
};

struct Bufs {
    std::vector<void*> v;
    template <class T>
    hipError_t alloc(T** p, size_t count) {
        hipError_t e = hipMalloc((void**)p, count * sizeof(T) + 16);
        if (e == hipSuccess) v.push_back(*p);
        return e;
    }
    ~Bufs() {
        for (void* p : v) hipFree(p);
    }
};

// indices 0..n-1 of the 4-word determinants w, stably sorted by words 0..3
// (least significant first), into *idx (device, n entries)
hipError_t sort_dets(Tmp& tmp, Bufs& b, const uint64_t* w, int64_t n, int n_words, int64_t* idx, hipStream_t s) {
    int64_t *i2 = nullptr;
    uint64_t *k1 = nullptr, *k2 = nullptr;
    hipError_t e;
    if ((e = b.alloc(&i2, (size_t)n)) != hipSuccess) return e;
    if ((e = b.alloc(&k1, (size_t)n)) != hipSuccess) return e;
    if ((e = b.alloc(&k2, (size_t)n)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_iota64, grid_for(n, 256), 256, 0, s, idx, n);
    // least significant word first: the order of cmp_det (words 3, 2, 1, 0);
    // the high words are zero when every orbital fits one word
    for (int wd = 0; wd < 4; ++wd) {
        if (n_words == 2 && (wd == 1 || wd == 3)) continue;
        hipLaunchKernelGGL(k_word, grid_for(n, 256), 256, 0, s, w, idx, n, wd, k1);
        size_t need = 0;
        if ((e = rocprim::radix_sort_pairs(nullptr, need, k1, k2, idx, i2, (size_t)n, 0, 64, s)) != hipSuccess) return e;
        if ((e = tmp.ensure(need)) != hipSuccess) return e;
        if ((e = rocprim::radix_sort_pairs(tmp.p, need, k1, k2, idx, i2, (size_t)n, 0, 64, s)) != hipSuccess) return e;
        if ((e = hipMemcpyAsync(idx, i2, (size_t)n * 8, hipMemcpyDeviceToDevice, s)) != hipSuccess) return e;
    }
    return hipSuccess;
}

// host list construction ------------------------------------------------------

struct Ent {
    double a;
    uint16_t t;
};

}  // namespace

extern "C" {

int hci_excite_init(hci_ctx* c, int n_core, int n_up, int n_dn, double eps_var, double* eps_start) {
    if (!c) return HCI_ERR_ARG;
    if (c->h_two.empty()) return hci_fail(c, HCI_ERR_STATE, "hci_excite_init: integrals not set");
    const int n = c->n_orb;
    if (n_core < 0 || n_up < 1 || n_dn < 1 || n_up > n || n_dn > n || n_core > std::min(n_up, n_dn))
        return hci_fail(c, HCI_ERR_ARG, "hci_excite_init: bad electron or core counts");
    HIP_TRY(c, hipSetDevice(c->device));
    auto h1 = [&](int p, int q) { return c->h_one[sc2((size_t)p + 1, (size_t)q + 1)]; };
    auto eri = [&](int p, int q, int r, int s) {
        return c->h_two[sc2(sc2((size_t)p + 1, (size_t)q + 1), sc2((size_t)r + 1, (size_t)s + 1))];
    };
    SelectLists* L = new SelectLists();
    select_free(c);
    c->sel = L;
    std::vector<int> order(n);
    for (int p = 0; p < n; ++p) order[p] = p;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return h1(a, a) < h1(b, b); });
    std::vector<int> val(order.begin() + n_core, order.end());
    const int nv = (int)val.size();
    std::vector<int32_t> vpos(n, -1), spos((size_t)nv * nv, -1);
    for (int a = 0; a < nv; ++a) {
        vpos[val[a]] = a;
        if (val[a] < 64)
            L->vmask[0] |= 1ull << val[a];
        else
            L->vmask[1] |= 1ull << (val[a] - 64);
    }
    L->n_core = n_core;
    L->n_up = n_up;
    L->n_dn = n_dn;
    L->nv = nv;
    L->Lo = (int64_t)(nv - 1) * (nv - 1);
    L->Ls = nv >= 4 ? (int64_t)(nv - 2) * (nv - 3) / 2 : 0;
    L->Lg = nv - 1;
    // opposite-spin doubles: direct(p, q, r, s) = (pr|qs)
    std::vector<double> oh((size_t)nv * nv * L->Lo), sh, gh((size_t)nv * L->Lg);
    std::vector<uint16_t> ot(oh.size()), st, gt(gh.size());
    std::vector<Ent> v;
    for (int p : val)
        for (int q : val) {
            v.clear();
            for (int r : val) {
                if (r == p) continue;
                for (int s : val) {
                    if (s == q) continue;
                    const double a = std::fabs(eri(p, r, q, s));
                    L->max_opp = std::max(L->max_opp, a);
                    v.push_back({a, (uint16_t)(r | (s << 8))});
                }
            }
            std::stable_sort(v.begin(), v.end(), [](const Ent& x, const Ent& y) { return x.a > y.a; });
            const size_t base = ((size_t)vpos[p] * nv + vpos[q]) * L->Lo;
            for (size_t e = 0; e < v.size(); ++e) {
                oh[base + e] = v[e].a;
                ot[base + e] = v[e].t;
            }
        }
    // same-spin doubles: direct_plus_exchange(p, q, r, s) = (pr|qs) - (ps|qr), p < q, r < s
    int n_same = 0;
    for (int p : val)
        for (int q : val)
            if (p < q) spos[(size_t)vpos[p] * nv + vpos[q]] = n_same++;
    sh.assign((size_t)n_same * L->Ls, 0.0);
    st.assign(sh.size(), 0);
    for (int p : val)
        for (int q : val) {
            if (p >= q) continue;
            v.clear();
            for (int r : val) {
                if (r == p || r == q) continue;
                for (int s : val) {
                    if (r >= s || s == p || s == q) continue;
                    const double a = std::fabs(eri(p, r, q, s) - eri(p, s, q, r));
                    L->max_same = std::max(L->max_same, a);
                    v.push_back({a, (uint16_t)(r | (s << 8))});
                }
            }
            std::stable_sort(v.begin(), v.end(), [](const Ent& x, const Ent& y) { return x.a > y.a; });
            const size_t base = (size_t)spos[(size_t)vpos[p] * nv + vpos[q]] * L->Ls;
            for (size_t e = 0; e < v.size(); ++e) {
                sh[base + e] = v[e].a;
                st[base + e] = v[e].t;
            }
        }
    // singles: the bound of src/excite/init.rs:181-231
    std::vector<double> vs, vo;
    for (int p : val) {
        v.clear();
        for (int r : val) {
            if (r == p) continue;
            vs.clear();
            vo.clear();
            for (int q = 0; q < n; ++q) {
                if (q == p || q == r) continue;
                vs.push_back(eri(p, r, q, q) - eri(p, q, q, r));
                vo.push_back(eri(p, r, q, q));
            }
            std::sort(vs.begin(), vs.end());
            std::sort(vo.begin(), vo.end());
            double s1 = 0.0, s2 = 0.0, s3 = 0.0, s4 = 0.0;
            for (int i = 0; i < n_up - 1; ++i) s1 += vs[i];
            for (int i = 0; i < n_dn; ++i) s2 += vo[i];
            for (size_t i = vs.size() - (size_t)(n_up - 1); i < vs.size(); ++i) s3 += vs[i];
            for (size_t i = vs.size() - (size_t)n_dn; i < vo.size(); ++i) s4 += vo[i];
            const double lo = h1(p, r) + s1 + s2, hi = h1(p, r) + s3 + s4;
            v.push_back({std::fabs(lo) > std::fabs(hi) ? std::fabs(lo) : std::fabs(hi), (uint16_t)r});
        }
        std::stable_sort(v.begin(), v.end(), [](const Ent& x, const Ent& y) { return x.a > y.a; });
        const size_t base = (size_t)vpos[p] * L->Lg;
        for (size_t e = 0; e < v.size(); ++e) {
            gh[base + e] = v[e].a;
            gt[base + e] = v[e].t;
        }
        if (!v.empty()) L->max_sing = std::max(L->max_sing, v[0].a);
    }
    // sum_remaining_h_squared of every list entry (compute_sum_remaining,
    // src/excite/init.rs:288-309): 0 for the last entry, then backwards
    // rem[e] = rem[e + 1] + a[e + 1] * a[e + 1] (no fused multiply-add, as in Rust)
    auto suffix = [](const std::vector<double>& a, int64_t len, std::vector<double>& rem) {
        rem.assign(a.size(), 0.0);
        if (len <= 0) return;
        for (size_t base = 0; base < a.size(); base += (size_t)len) {
#pragma clang fp contract(off)
            rem[base + len - 1] = 0.0;
            for (int64_t e = len - 2; e >= 0; --e) {
                const double sq = a[base + e + 1] * a[base + e + 1];
                rem[base + e] = rem[base + e + 1] + sq;
            }
        }
    };
    std::vector<double> orem, srem, grem;
    suffix(oh, L->Lo, orem);
    suffix(sh, L->Ls, srem);
    suffix(gh, L->Lg, grem);
    // sum_remaining_abs_h: rem[e] = rem[e + 1] + a[e + 1] (same file, :288-309)
    auto suffix_abs = [](const std::vector<double>& a, int64_t len, std::vector<double>& rem) {
        rem.assign(a.size(), 0.0);
        if (len <= 0) return;
        for (size_t base = 0; base < a.size(); base += (size_t)len) {
            rem[base + len - 1] = 0.0;
            for (int64_t e = len - 2; e >= 0; --e) rem[base + e] = rem[base + e + 1] + a[base + e + 1];
        }
    };
    std::vector<double> oabs, sabs, gabs;
    suffix_abs(oh, L->Lo, oabs);
    suffix_abs(sh, L->Ls, sabs);
    suffix_abs(gh, L->Lg, gabs);
    // initial epsilon from Hartree-Fock (src/wf/eps.rs:44-171, z_sym = 1)
    {
        auto occ = [](int p, int ne) { return p < ne; };  // Hartree-Fock: the lowest orbitals by index
        double max_sym = eps_var, max_asym = eps_var;
        for (int i = 0; i < n_up; ++i) {
            if (vpos[i] < 0) continue;
            for (int j = 0; j < n_dn; ++j) {
                if (vpos[j] < 0) continue;
                bool fs = false, fa = false;
                const size_t base = ((size_t)vpos[i] * nv + vpos[j]) * L->Lo;
                for (int64_t e = 0; e < L->Lo; ++e) {
                    const int r = ot[base + e] & 0xFF, s = ot[base + e] >> 8;
                    if (occ(r, n_up) || occ(s, n_dn)) continue;
                    const double a = oh[base + e];
                    if (i == j && r == s) {
                        fs = true;
                        max_sym = std::max(max_sym, a);
                    } else {
                        fa = true;
                        max_asym = std::max(max_asym, a);
                    }
                    if (fs && fa) break;
                }
            }
        }
        for (int spin = 0; spin < 2; ++spin) {
            const int ne = spin == 0 ? n_up : n_dn;
            for (int i = 0; i < ne; ++i) {
                if (vpos[i] < 0) continue;
                for (int j = i + 1; j < ne; ++j) {
                    if (vpos[j] < 0 || L->Ls == 0) continue;
                    const size_t base = (size_t)spos[(size_t)vpos[i] * nv + vpos[j]] * L->Ls;
                    for (int64_t e = 0; e < L->Ls; ++e) {
                        const int r = st[base + e] & 0xFF, s = st[base + e] >> 8;
                        if (occ(r, ne) || occ(s, ne)) continue;
                        max_asym = std::max(max_asym, sh[base + e]);
                        break;
                    }
                }
            }
        }
        L->eps_start = (max_sym > max_asym ? max_sym : max_asym) - 1e-9;
    }
    auto up = [&](auto** d, const auto& h) -> int {
        using T = typename std::remove_reference<decltype(h)>::type::value_type;
        HIP_TRY(c, hipMalloc((void**)d, h.size() * sizeof(T) + 16));
        if (!h.empty()) HIP_TRY(c, hipMemcpy(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
        return HCI_OK;
    };
    int rc;
    if ((rc = up(&L->d_oh, oh)) || (rc = up(&L->d_ot, ot)) || (rc = up(&L->d_sh, sh)) || (rc = up(&L->d_st, st)) ||
        (rc = up(&L->d_gh, gh)) || (rc = up(&L->d_gt, gt)) || (rc = up(&L->d_vpos, vpos)) ||
        (rc = up(&L->d_spos, spos)) || (rc = up(&L->d_or, orem)) || (rc = up(&L->d_sr, srem)) ||
        (rc = up(&L->d_gr, grem)) || (rc = up(&L->d_oa, oabs)) || (rc = up(&L->d_sa, sabs)) ||
        (rc = up(&L->d_ga, gabs)))
        return rc;
    if (eps_start) *eps_start = L->eps_start;
    return HCI_OK;
}

int hci_excite_info(const hci_ctx* c, double* max_opp, double* max_same, double* max_sing) {
    if (!c || !c->sel) return HCI_ERR_STATE;
    if (max_opp) *max_opp = c->sel->max_opp;
    if (max_same) *max_same = c->sel->max_same;
    if (max_sing) *max_sing = c->sel->max_sing;
    return HCI_OK;
}

int hci_select(hci_ctx* c, const uint64_t* dets, const double* coeffs, const double* diag, int64_t n, double eps,
               int singles, uint64_t* out, double* out_diag, int64_t cap, int64_t* count, int mem) {
    if (!c) return HCI_ERR_ARG;
    if (!c->sel) return hci_fail(c, HCI_ERR_STATE, "hci_select: call hci_excite_init first");
    if (n < 1 || !dets || !coeffs || !count || cap < 0 || (cap > 0 && !out) || !(eps > 0.0) || (out_diag && !diag))
        return hci_fail(c, HCI_ERR_ARG, "hci_select: bad arguments");
    HIP_TRY(c, hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const SelectLists& L = *c->sel;
    const DevSel D = dev_sel(L);
    const DevInts T = c->ints();
    Bufs b;
    Tmp tmp;
    const uint64_t* d_dets = dets;
    const double* d_coef = coeffs;
    const double* d_diag = diag;
    if (mem == HCI_MEM_HOST) {
        uint64_t* w = nullptr;
        double* q = nullptr;
        HIP_TRY(c, b.alloc(&w, (size_t)n * 4));
        HIP_TRY(c, b.alloc(&q, (size_t)n));
        HIP_TRY(c, hipMemcpyAsync(w, dets, (size_t)n * 32, hipMemcpyHostToDevice, s));
        HIP_TRY(c, hipMemcpyAsync(q, coeffs, (size_t)n * 8, hipMemcpyHostToDevice, s));
        d_dets = w;
        d_coef = q;
        if (out_diag) {
            double* g = nullptr;
            HIP_TRY(c, b.alloc(&g, (size_t)n));
            HIP_TRY(c, hipMemcpyAsync(g, diag, (size_t)n * 8, hipMemcpyHostToDevice, s));
            d_diag = g;
        }
    }
    const bool wide = c->n_orb > 64;
    int64_t *cnt = nullptr, *off = nullptr;
    HIP_TRY(c, b.alloc(&cnt, (size_t)n + 1));
    HIP_TRY(c, b.alloc(&off, (size_t)n + 1));
    if (wide)
        hipLaunchKernelGGL((k_select<u128, false>), grid_for(n, 256), 256, 0, s, D, T, L.vmask[0], L.vmask[1], d_dets,
                           d_coef, n, eps, singles, cnt, nullptr, nullptr, nullptr, nullptr);
    else
        hipLaunchKernelGGL((k_select<uint64_t, false>), grid_for(n, 256), 256, 0, s, D, T, L.vmask[0], L.vmask[1],
                           d_dets, d_coef, n, eps, singles, cnt, nullptr, nullptr, nullptr, nullptr);
    HIP_TRY(c, hipMemsetAsync(cnt + n, 0, 8, s));
    size_t need = 0;
    HIP_TRY(c, rocprim::exclusive_scan(nullptr, need, cnt, off, (int64_t)0, (size_t)n + 1, rocprim::plus<int64_t>(), s));
    HIP_TRY(c, tmp.ensure(need));
    HIP_TRY(c, rocprim::exclusive_scan(tmp.p, need, cnt, off, (int64_t)0, (size_t)n + 1, rocprim::plus<int64_t>(), s));
    int64_t m = 0;
    HIP_TRY(c, hipMemcpyAsync(&m, off + n, 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipStreamSynchronize(s));
    if (m == 0) {
        *count = 0;
        return HCI_OK;
    }
    uint64_t* cand = nullptr;
    int64_t* csrc = nullptr;
    uint32_t* ccode = nullptr;
    HIP_TRY(c, b.alloc(&cand, (size_t)m * 4));
    if (out_diag) {
        HIP_TRY(c, b.alloc(&csrc, (size_t)m));
        HIP_TRY(c, b.alloc(&ccode, (size_t)m));
    }
    if (wide)
        hipLaunchKernelGGL((k_select<u128, true>), grid_for(n, 256), 256, 0, s, D, T, L.vmask[0], L.vmask[1], d_dets,
                           d_coef, n, eps, singles, nullptr, off, cand, csrc, ccode);
    else
        hipLaunchKernelGGL((k_select<uint64_t, true>), grid_for(n, 256), 256, 0, s, D, T, L.vmask[0], L.vmask[1],
                           d_dets, d_coef, n, eps, singles, nullptr, off, cand, csrc, ccode);
    HIP_TRY(c, hipGetLastError());
    // sort candidates and the current space; keep first occurrences not in the space
    const int nw = wide ? 4 : 2;
    int64_t *cidx = nullptr, *didx = nullptr, *keep = nullptr, *pos = nullptr, *rank = nullptr, *rank2 = nullptr;
    HIP_TRY(c, b.alloc(&cidx, (size_t)m));
    HIP_TRY(c, b.alloc(&didx, (size_t)n));
    HIP_TRY(c, sort_dets(tmp, b, cand, m, nw, cidx, s));
    HIP_TRY(c, sort_dets(tmp, b, d_dets, n, nw, didx, s));
    HIP_TRY(c, b.alloc(&keep, (size_t)m + 1));
    HIP_TRY(c, b.alloc(&pos, (size_t)m + 1));
    hipLaunchKernelGGL(k_keep, grid_for(m, 256), 256, 0, s, cand, cidx, m, d_dets, didx, n, keep);
    HIP_TRY(c, hipMemsetAsync(keep + m, 0, 8, s));
    need = 0;
    HIP_TRY(c, rocprim::exclusive_scan(nullptr, need, keep, pos, (int64_t)0, (size_t)m + 1, rocprim::plus<int64_t>(), s));
    HIP_TRY(c, tmp.ensure(need));
    HIP_TRY(c, rocprim::exclusive_scan(tmp.p, need, keep, pos, (int64_t)0, (size_t)m + 1, rocprim::plus<int64_t>(), s));
    int64_t k = 0;
    HIP_TRY(c, hipMemcpyAsync(&k, pos + m, 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipStreamSynchronize(s));
    *count = k;
    if (k > cap) return hci_fail(c, HCI_ERR_ARG, "hci_select: output capacity too small (count holds the need)");
    if (k == 0) return HCI_OK;
    HIP_TRY(c, b.alloc(&rank, (size_t)k));
    HIP_TRY(c, b.alloc(&rank2, (size_t)k));
    hipLaunchKernelGGL(k_compact, grid_for(m, 256), 256, 0, s, keep, pos, cidx, m, rank);
    // discovery order
    need = 0;
    HIP_TRY(c, rocprim::radix_sort_keys(nullptr, need, rank, rank2, (size_t)k, 0, 64, s));
    HIP_TRY(c, tmp.ensure(need));
    HIP_TRY(c, rocprim::radix_sort_keys(tmp.p, need, rank, rank2, (size_t)k, 0, 64, s));
    uint64_t* res = nullptr;
    HIP_TRY(c, b.alloc(&res, (size_t)k * 4));
    hipLaunchKernelGGL(k_gather_dets, grid_for(k, 256), 256, 0, s, cand, rank2, k, res);
    if (out_diag) {
        double* dg = nullptr;
        HIP_TRY(c, b.alloc(&dg, (size_t)k));
        if (wide)
            hipLaunchKernelGGL(k_new_diag<u128>, grid_for(k, 256), 256, 0, s, T, d_dets, d_diag, csrc, ccode, rank2, k,
                               dg);
        else
            hipLaunchKernelGGL(k_new_diag<uint64_t>, grid_for(k, 256), 256, 0, s, T, d_dets, d_diag, csrc, ccode,
                               rank2, k, dg);
        HIP_TRY(c, hipMemcpyAsync(out_diag, dg, (size_t)k * 8,
                                  mem == HCI_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
    }
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipMemcpyAsync(out, res, (size_t)k * 32,
                              mem == HCI_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipStreamSynchronize(s));
    return HCI_OK;
}

int hci_pt2_dtm(hci_ctx* c, const uint64_t* dets, const double* coeffs, int64_t n, double e0, double eps,
                int64_t batch, double* e2, int64_t* n_ext, int mem) {
    if (!c) return HCI_ERR_ARG;
    if (!c->sel) return hci_fail(c, HCI_ERR_STATE, "hci_pt2_dtm: call hci_excite_init first");
    if (n < 1 || !dets || !coeffs || !e2 || !(eps > 0.0) || batch < 0)
        return hci_fail(c, HCI_ERR_ARG, "hci_pt2_dtm: bad arguments");
    HIP_TRY(c, hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const SelectLists& L = *c->sel;
    const DevSel D = dev_sel(L);
    const DevInts T = c->ints();
    const bool wide = c->n_orb > 64;
    const int nw = wide ? 4 : 2;
    Bufs b;
    Tmp tmp;
    const uint64_t* d_dets = dets;
    const double* d_coef = coeffs;
    if (mem == HCI_MEM_HOST) {
        uint64_t* w = nullptr;
        double* q = nullptr;
        HIP_TRY(c, b.alloc(&w, (size_t)n * 4));
        HIP_TRY(c, b.alloc(&q, (size_t)n));
        HIP_TRY(c, hipMemcpyAsync(w, dets, (size_t)n * 32, hipMemcpyHostToDevice, s));
        HIP_TRY(c, hipMemcpyAsync(q, coeffs, (size_t)n * 8, hipMemcpyHostToDevice, s));
        d_dets = w;
        d_coef = q;
    }
    auto scan = [&](const int64_t* in, int64_t* out, int64_t len) -> hipError_t {
        size_t need = 0;
        hipError_t e = rocprim::exclusive_scan(nullptr, need, in, out, (int64_t)0, (size_t)len, rocprim::plus<int64_t>(), s);
        if (e != hipSuccess) return e;
        if ((e = tmp.ensure(need)) != hipSuccess) return e;
        return rocprim::exclusive_scan(tmp.p, need, in, out, (int64_t)0, (size_t)len, rocprim::plus<int64_t>(), s);
    };
    auto walk = [&](bool fill, uint32_t n_parts, uint32_t part, int64_t* cnt, const int64_t* off, uint64_t* cw,
                    double* cv) {
        if (wide) {
            if (fill)
                hipLaunchKernelGGL((k_pt_walk<u128, true>), grid_for(n, 256), 256, 0, s, D, T, L.vmask[0], L.vmask[1],
                                   d_dets, d_coef, n, eps, n_parts, part, cnt, off, cw, cv);
            else
                hipLaunchKernelGGL((k_pt_walk<u128, false>), grid_for(n, 256), 256, 0, s, D, T, L.vmask[0], L.vmask[1],
                                   d_dets, d_coef, n, eps, n_parts, part, cnt, off, cw, cv);
        } else {
            if (fill)
                hipLaunchKernelGGL((k_pt_walk<uint64_t, true>), grid_for(n, 256), 256, 0, s, D, T, L.vmask[0],
                                   L.vmask[1], d_dets, d_coef, n, eps, n_parts, part, cnt, off, cw, cv);
            else
                hipLaunchKernelGGL((k_pt_walk<uint64_t, false>), grid_for(n, 256), 256, 0, s, D, T, L.vmask[0],
                                   L.vmask[1], d_dets, d_coef, n, eps, n_parts, part, cnt, off, cw, cv);
        }
    };
    int64_t *cnt = nullptr, *off = nullptr, *didx = nullptr;
    double* partial = nullptr;
    HIP_TRY(c, b.alloc(&cnt, (size_t)n + 1));
    HIP_TRY(c, b.alloc(&off, (size_t)n + 1));
    HIP_TRY(c, b.alloc(&didx, (size_t)n));
    HIP_TRY(c, b.alloc(&partial, kPtBlocks));
    HIP_TRY(c, sort_dets(tmp, b, d_dets, n, nw, didx, s));
    // all candidates first: their number sets the number of parts
    walk(false, 1, 0, cnt, nullptr, nullptr, nullptr);
    HIP_TRY(c, hipMemsetAsync(cnt + n, 0, 8, s));
    HIP_TRY(c, scan(cnt, off, n + 1));
    int64_t total = 0;
    HIP_TRY(c, hipMemcpyAsync(&total, off + n, 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipStreamSynchronize(s));
    const int64_t budget = batch > 0 ? batch : (int64_t)1 << 28;
    const int64_t parts64 = (total + budget - 1) / budget;
    if (parts64 > (int64_t)1 << 20) return hci_fail(c, HCI_ERR_ARG, "hci_pt2_dtm: budget too small for the candidates");
    const uint32_t n_parts = (uint32_t)std::max<int64_t>(parts64, 1);
    double sum = 0.0;
    int64_t n_out = 0;
    std::vector<double> hp(kPtBlocks);
    for (uint32_t part = 0; part < n_parts; ++part) {
        int64_t m = total;
        if (n_parts > 1) {
            walk(false, n_parts, part, cnt, nullptr, nullptr, nullptr);
            HIP_TRY(c, hipMemsetAsync(cnt + n, 0, 8, s));
            HIP_TRY(c, scan(cnt, off, n + 1));
            HIP_TRY(c, hipMemcpyAsync(&m, off + n, 8, hipMemcpyDeviceToHost, s));
            HIP_TRY(c, hipStreamSynchronize(s));
        }
        if (m == 0) continue;
        Bufs bb;
        uint64_t *cw = nullptr, *rw = nullptr;
        double *cv = nullptr, *rv = nullptr;
        int64_t *cidx = nullptr, *head = nullptr, *hpos = nullptr, *start = nullptr, *keep = nullptr, *kpos = nullptr;
        HIP_TRY(c, bb.alloc(&cw, (size_t)m * 4));
        HIP_TRY(c, bb.alloc(&cv, (size_t)m));
        walk(true, n_parts, part, nullptr, off, cw, cv);
        // stable sort by words: each external determinant's contributions stay
        // in discovery order
        HIP_TRY(c, bb.alloc(&cidx, (size_t)m));
        HIP_TRY(c, sort_dets(tmp, bb, cw, m, nw, cidx, s));
        HIP_TRY(c, bb.alloc(&head, (size_t)m + 1));
        HIP_TRY(c, bb.alloc(&hpos, (size_t)m + 1));
        hipLaunchKernelGGL(k_run_heads, grid_for(m, 256), 256, 0, s, cw, cidx, m, head);
        HIP_TRY(c, hipMemsetAsync(head + m, 0, 8, s));
        HIP_TRY(c, scan(head, hpos, m + 1));
        int64_t runs = 0;
        HIP_TRY(c, hipMemcpyAsync(&runs, hpos + m, 8, hipMemcpyDeviceToHost, s));
        HIP_TRY(c, hipStreamSynchronize(s));
        HIP_TRY(c, bb.alloc(&start, (size_t)runs + 1));
        HIP_TRY(c, bb.alloc(&rw, (size_t)runs * 4));
        HIP_TRY(c, bb.alloc(&rv, (size_t)runs));
        HIP_TRY(c, bb.alloc(&keep, (size_t)runs + 1));
        HIP_TRY(c, bb.alloc(&kpos, (size_t)runs + 1));
        hipLaunchKernelGGL(k_run_starts, grid_for(m + 1, 256), 256, 0, s, head, hpos, m, start);
        hipLaunchKernelGGL(k_run_sum, grid_for(runs, 256), 256, 0, s, cw, cv, cidx, start, runs, d_dets, didx, n, rw, rv,
                           keep);
        HIP_TRY(c, hipMemsetAsync(keep + runs, 0, 8, s));
        HIP_TRY(c, scan(keep, kpos, runs + 1));
        if (wide)
            hipLaunchKernelGGL(k_pt_terms<u128>, kPtBlocks, 256, 0, s, T, rw, rv, keep, runs, e0, partial);
        else
            hipLaunchKernelGGL(k_pt_terms<uint64_t>, kPtBlocks, 256, 0, s, T, rw, rv, keep, runs, e0, partial);
        HIP_TRY(c, hipGetLastError());
        int64_t kept = 0;
        HIP_TRY(c, hipMemcpyAsync(&kept, kpos + runs, 8, hipMemcpyDeviceToHost, s));
        HIP_TRY(c, hipMemcpyAsync(hp.data(), partial, kPtBlocks * 8, hipMemcpyDeviceToHost, s));
        HIP_TRY(c, hipStreamSynchronize(s));
        for (double v : hp) sum += v;
        n_out += kept;
    }
    *e2 = sum;
    if (n_ext) *n_ext = n_out;
    return HCI_OK;
}

}  // extern "C"

namespace {

// The external product of approx_matmul_external_dtm_only (variant 0) or of a
// screened-sampler variant (HCI_EXT_*): walk (count, then fill), stable sort
// of the candidates by their words, run sums in discovery order, members of the
// space dropped, outputs in first-discovery order.
int ext_product(hci_ctx* c, const char* fn, int variant, const uint64_t* dets, const double* coeffs,
                const double* diag, int64_t n, double eps, uint64_t* out, double* out_coeffs, double* out_diag,
                int64_t cap, int64_t* count, double* sum_remaining, hci_det_orb_sample* samples, int64_t scap,
                int64_t* n_samples, int mem) {
    const std::string name(fn);
    if (!c) return HCI_ERR_ARG;
    if (!c->sel) return hci_fail(c, HCI_ERR_STATE, name + ": call hci_excite_init first");
    if (n < 1 || !dets || !coeffs || !count || cap < 0 || (cap > 0 && (!out || !out_coeffs)) || !(eps > 0.0) ||
        (out_diag && !diag) || scap < 0 || (scap > 0 && !samples))
        return hci_fail(c, HCI_ERR_ARG, name + ": bad arguments");
    if (variant == HCI_EXT_NO_SINGLES && c->sel->Lg == 0)
        return hci_fail(c, HCI_ERR_ARG,
                        name + ": no_singles reads the first entry of every valence electron's singles list, "
                               "and these lists are empty");
    HIP_TRY(c, hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const SelectLists& L = *c->sel;
    const DevSel D = dev_sel(L);
    const DevInts T = c->ints();
    const bool wide = c->n_orb > 64;
    const int nw = wide ? 4 : 2;
    const hipMemcpyKind out_kind = mem == HCI_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    Bufs b;
    Tmp tmp;
    const uint64_t* d_dets = dets;
    const double* d_coef = coeffs;
    const double* d_diag = diag;
    if (mem == HCI_MEM_HOST) {
        uint64_t* w = nullptr;
        double* q = nullptr;
        HIP_TRY(c, b.alloc(&w, (size_t)n * 4));
        HIP_TRY(c, b.alloc(&q, (size_t)n));
        HIP_TRY(c, hipMemcpyAsync(w, dets, (size_t)n * 32, hipMemcpyHostToDevice, s));
        HIP_TRY(c, hipMemcpyAsync(q, coeffs, (size_t)n * 8, hipMemcpyHostToDevice, s));
        d_dets = w;
        d_coef = q;
        if (out_diag) {
            double* g = nullptr;
            HIP_TRY(c, b.alloc(&g, (size_t)n));
            HIP_TRY(c, hipMemcpyAsync(g, diag, (size_t)n * 8, hipMemcpyHostToDevice, s));
            d_diag = g;
        }
    }
    auto scan = [&](const int64_t* in, int64_t* o, int64_t len) -> hipError_t {
        size_t need = 0;
        hipError_t e = rocprim::exclusive_scan(nullptr, need, in, o, (int64_t)0, (size_t)len, rocprim::plus<int64_t>(), s);
        if (e != hipSuccess) return e;
        if ((e = tmp.ensure(need)) != hipSuccess) return e;
        return rocprim::exclusive_scan(tmp.p, need, in, o, (int64_t)0, (size_t)len, rocprim::plus<int64_t>(), s);
    };
    int64_t *cnt = nullptr, *off = nullptr, *scnt = nullptr, *soff = nullptr, *didx = nullptr;
    double* srem = nullptr;
    HIP_TRY(c, b.alloc(&cnt, (size_t)n + 1));
    HIP_TRY(c, b.alloc(&off, (size_t)n + 1));
    HIP_TRY(c, b.alloc(&didx, (size_t)n));
    if (variant != 0) {
        HIP_TRY(c, b.alloc(&scnt, (size_t)n + 1));
        HIP_TRY(c, b.alloc(&soff, (size_t)n + 1));
    }
    if (sum_remaining) HIP_TRY(c, b.alloc(&srem, (size_t)n));
    HIP_TRY(c, sort_dets(tmp, b, d_dets, n, nw, didx, s));
    // one launch of the walk: variant 0 through k_ext_walk, the others through
    // k_extv_walk<S, variant, FILL>
    auto launch = [&](bool fill, uint64_t* cw, double* cv, int64_t* csrc, uint32_t* ccode,
                      hci_det_orb_sample* smp) -> hipError_t {
        const dim3 g = grid_for(n, 256);
        auto run = [&](auto tag_s, auto tag_v, auto tag_f) {
            using S = decltype(tag_s);
            constexpr int V = decltype(tag_v)::value;
            constexpr bool F = decltype(tag_f)::value;
            if constexpr (V == 0)
                hipLaunchKernelGGL((k_ext_walk<S, F>), g, 256, 0, s, D, T, L.vmask[0], L.vmask[1], d_dets, d_coef, n,
                                   eps, F ? nullptr : cnt, F ? nullptr : srem, F ? off : nullptr, cw, cv, csrc, ccode);
            else
                hipLaunchKernelGGL((k_extv_walk<S, V, F>), g, 256, 0, s, D, T, L.vmask[0], L.vmask[1], d_dets, d_coef,
                                   n, eps, didx, F ? nullptr : cnt, F ? nullptr : scnt, F ? off : nullptr,
                                   F ? soff : nullptr, cw, cv, csrc, ccode, smp);
        };
        auto by_v = [&](auto tag_s, auto tag_f) {
            switch (variant) {
                case HCI_EXT_SEPARATE: run(tag_s, std::integral_constant<int, HCI_EXT_SEPARATE>(), tag_f); break;
                case HCI_EXT_SKIP_SINGLES: run(tag_s, std::integral_constant<int, HCI_EXT_SKIP_SINGLES>(), tag_f); break;
                case HCI_EXT_SEMISTOCH_SINGLES:
                    run(tag_s, std::integral_constant<int, HCI_EXT_SEMISTOCH_SINGLES>(), tag_f);
                    break;
                case HCI_EXT_NO_SINGLES: run(tag_s, std::integral_constant<int, HCI_EXT_NO_SINGLES>(), tag_f); break;
                default: run(tag_s, std::integral_constant<int, 0>(), tag_f); break;
            }
        };
        if (wide)
            fill ? by_v(u128(), std::true_type()) : by_v(u128(), std::false_type());
        else
            fill ? by_v(uint64_t(), std::true_type()) : by_v(uint64_t(), std::false_type());
        return hipGetLastError();
    };
    HIP_TRY(c, launch(false, nullptr, nullptr, nullptr, nullptr, nullptr));
    HIP_TRY(c, hipMemsetAsync(cnt + n, 0, 8, s));
    HIP_TRY(c, scan(cnt, off, n + 1));
    if (sum_remaining) HIP_TRY(c, hipMemcpyAsync(sum_remaining, srem, (size_t)n * 8, out_kind, s));
    int64_t m = 0, ns = 0;
    HIP_TRY(c, hipMemcpyAsync(&m, off + n, 8, hipMemcpyDeviceToHost, s));
    if (variant != 0) {
        HIP_TRY(c, hipMemsetAsync(scnt + n, 0, 8, s));
        HIP_TRY(c, scan(scnt, soff, n + 1));
        HIP_TRY(c, hipMemcpyAsync(&ns, soff + n, 8, hipMemcpyDeviceToHost, s));
    }
    HIP_TRY(c, hipStreamSynchronize(s));
    if (n_samples) *n_samples = ns;
    if (samples && ns > scap) return hci_fail(c, HCI_ERR_ARG, name + ": sample capacity too small (n_samples holds the need)");
    uint64_t *cw = nullptr, *rw = nullptr;
    double *cv = nullptr, *rv = nullptr;
    int64_t *csrc = nullptr, *cidx = nullptr, *head = nullptr, *hpos = nullptr, *start = nullptr, *keep = nullptr,
            *kpos = nullptr;
    uint32_t* ccode = nullptr;
    hci_det_orb_sample* d_smp = nullptr;
    HIP_TRY(c, b.alloc(&cw, (size_t)m * 4));
    HIP_TRY(c, b.alloc(&cv, (size_t)m));
    HIP_TRY(c, b.alloc(&csrc, (size_t)m));
    HIP_TRY(c, b.alloc(&ccode, (size_t)m));
    if (ns > 0) HIP_TRY(c, b.alloc(&d_smp, (size_t)ns));
    if (m > 0 || ns > 0) HIP_TRY(c, launch(true, cw, cv, csrc, ccode, d_smp));
    // records are copied once the determinant capacity has been checked
    auto copy_samples = [&]() -> hipError_t {
        if (!samples || ns == 0) return hipSuccess;
        return hipMemcpyAsync(samples, d_smp, (size_t)ns * sizeof(hci_det_orb_sample), out_kind, s);
    };
    if (m == 0) {
        *count = 0;
        HIP_TRY(c, copy_samples());
        HIP_TRY(c, hipStreamSynchronize(s));
        return HCI_OK;
    }
    // stable sort by words: each target's contributions stay in discovery
    // order (add_det_with_coeff, src/wf/mod.rs:144-165), summed per run; runs
    // of determinants in the space are dropped
    HIP_TRY(c, b.alloc(&cidx, (size_t)m));
    HIP_TRY(c, sort_dets(tmp, b, cw, m, nw, cidx, s));
    HIP_TRY(c, b.alloc(&head, (size_t)m + 1));
    HIP_TRY(c, b.alloc(&hpos, (size_t)m + 1));
    hipLaunchKernelGGL(k_run_heads, grid_for(m, 256), 256, 0, s, cw, cidx, m, head);
    HIP_TRY(c, hipMemsetAsync(head + m, 0, 8, s));
    HIP_TRY(c, scan(head, hpos, m + 1));
    int64_t runs = 0;
    HIP_TRY(c, hipMemcpyAsync(&runs, hpos + m, 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipStreamSynchronize(s));
    HIP_TRY(c, b.alloc(&start, (size_t)runs + 1));
    HIP_TRY(c, b.alloc(&rw, (size_t)runs * 4));
    HIP_TRY(c, b.alloc(&rv, (size_t)runs));
    HIP_TRY(c, b.alloc(&keep, (size_t)runs + 1));
    HIP_TRY(c, b.alloc(&kpos, (size_t)runs + 1));
    hipLaunchKernelGGL(k_run_starts, grid_for(m + 1, 256), 256, 0, s, head, hpos, m, start);
    hipLaunchKernelGGL(k_run_sum, grid_for(runs, 256), 256, 0, s, cw, cv, cidx, start, runs, d_dets, didx, n, rw, rv,
                       keep);
    HIP_TRY(c, hipMemsetAsync(keep + runs, 0, 8, s));
    HIP_TRY(c, scan(keep, kpos, runs + 1));
    int64_t k = 0;
    HIP_TRY(c, hipMemcpyAsync(&k, kpos + runs, 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipStreamSynchronize(s));
    *count = k;
    if (k > cap) return hci_fail(c, HCI_ERR_ARG, name + ": output capacity too small (count holds the need)");
    HIP_TRY(c, copy_samples());
    if (k == 0) {
        HIP_TRY(c, hipStreamSynchronize(s));
        return HCI_OK;
    }
    // first-discovery order: kept runs sorted by their head's candidate index
    int64_t *first = nullptr, *rid = nullptr, *first2 = nullptr, *rid2 = nullptr;
    HIP_TRY(c, b.alloc(&first, (size_t)k));
    HIP_TRY(c, b.alloc(&rid, (size_t)k));
    HIP_TRY(c, b.alloc(&first2, (size_t)k));
    HIP_TRY(c, b.alloc(&rid2, (size_t)k));
    hipLaunchKernelGGL(k_run_first, grid_for(runs, 256), 256, 0, s, cidx, start, keep, kpos, runs, first, rid);
    size_t need = 0;
    HIP_TRY(c, rocprim::radix_sort_pairs(nullptr, need, first, first2, rid, rid2, (size_t)k, 0, 64, s));
    HIP_TRY(c, tmp.ensure(need));
    HIP_TRY(c, rocprim::radix_sort_pairs(tmp.p, need, first, first2, rid, rid2, (size_t)k, 0, 64, s));
    uint64_t* ow = nullptr;
    double *ov = nullptr, *od = nullptr;
    HIP_TRY(c, b.alloc(&ow, (size_t)k * 4));
    HIP_TRY(c, b.alloc(&ov, (size_t)k));
    if (out_diag) HIP_TRY(c, b.alloc(&od, (size_t)k));
    if (wide)
        hipLaunchKernelGGL(k_ext_out<u128>, grid_for(k, 256), 256, 0, s, T, d_dets, d_diag, csrc, ccode, first2, rid2,
                           rw, rv, k, ow, ov, od);
    else
        hipLaunchKernelGGL(k_ext_out<uint64_t>, grid_for(k, 256), 256, 0, s, T, d_dets, d_diag, csrc, ccode, first2,
                           rid2, rw, rv, k, ow, ov, od);
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipMemcpyAsync(out, ow, (size_t)k * 32, out_kind, s));
    HIP_TRY(c, hipMemcpyAsync(out_coeffs, ov, (size_t)k * 8, out_kind, s));
    if (out_diag) HIP_TRY(c, hipMemcpyAsync(out_diag, od, (size_t)k * 8, out_kind, s));
    HIP_TRY(c, hipStreamSynchronize(s));
    return HCI_OK;
}

}  // namespace

extern "C" {

int hci_matmul_external(hci_ctx* c, const uint64_t* dets, const double* coeffs, const double* diag, int64_t n,
                        double eps, uint64_t* out, double* out_coeffs, double* out_diag, int64_t cap, int64_t* count,
                        double* sum_remaining, int mem) {
    return ext_product(c, "hci_matmul_external", 0, dets, coeffs, diag, n, eps, out, out_coeffs, out_diag, cap, count,
                       sum_remaining, nullptr, 0, nullptr, mem);
}

int hci_matmul_external_sampled(hci_ctx* c, int variant, const uint64_t* dets, const double* coeffs,
                                const double* diag, int64_t n, double eps, uint64_t* out, double* out_coeffs,
                                double* out_diag, int64_t cap, int64_t* count, hci_det_orb_sample* samples,
                                int64_t scap, int64_t* n_samples, int mem) {
    if (c && variant != HCI_EXT_SEPARATE && variant != HCI_EXT_SKIP_SINGLES && variant != HCI_EXT_SEMISTOCH_SINGLES &&
        variant != HCI_EXT_NO_SINGLES)
        return hci_fail(c, HCI_ERR_ARG, "hci_matmul_external_sampled: unknown variant");
    return ext_product(c, "hci_matmul_external_sampled", variant, dets, coeffs, diag, n, eps, out, out_coeffs,
                       out_diag, cap, count, nullptr, samples, scap, n_samples, mem);
}

int hci_semistoch_dtm(hci_ctx* c, const uint64_t* dets, const double* coeffs, const double* diag, int64_t n,
                      double e0, double eps, double* e2, int64_t* n_ext, double* rel_probs,
                      hci_det_orb_sample* samples, int64_t cap, int64_t* n_samples, int mem) {
    if (!c) return HCI_ERR_ARG;
    if (!c->sel) return hci_fail(c, HCI_ERR_STATE, "hci_semistoch_dtm: call hci_excite_init first");
    if (n < 1 || !dets || !coeffs || !diag || !e2 || !(eps > 0.0) || cap < 0 || (cap > 0 && !samples))
        return hci_fail(c, HCI_ERR_ARG, "hci_semistoch_dtm: bad arguments");
    HIP_TRY(c, hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const SelectLists& L = *c->sel;
    const DevSel D = dev_sel(L);
    const DevInts T = c->ints();
    const bool wide = c->n_orb > 64;
    const int nw = wide ? 4 : 2;
    const hipMemcpyKind out_kind = mem == HCI_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    Bufs b;
    Tmp tmp;
    const uint64_t* d_dets = dets;
    const double* d_coef = coeffs;
    const double* d_diag = diag;
    if (mem == HCI_MEM_HOST) {
        uint64_t* w = nullptr;
        double *q = nullptr, *g = nullptr;
        HIP_TRY(c, b.alloc(&w, (size_t)n * 4));
        HIP_TRY(c, b.alloc(&q, (size_t)n));
        HIP_TRY(c, b.alloc(&g, (size_t)n));
        HIP_TRY(c, hipMemcpyAsync(w, dets, (size_t)n * 32, hipMemcpyHostToDevice, s));
        HIP_TRY(c, hipMemcpyAsync(q, coeffs, (size_t)n * 8, hipMemcpyHostToDevice, s));
        HIP_TRY(c, hipMemcpyAsync(g, diag, (size_t)n * 8, hipMemcpyHostToDevice, s));
        d_dets = w;
        d_coef = q;
        d_diag = g;
    }
    auto scan = [&](const int64_t* in, int64_t* o, int64_t len) -> hipError_t {
        size_t need = 0;
        hipError_t e = rocprim::exclusive_scan(nullptr, need, in, o, (int64_t)0, (size_t)len, rocprim::plus<int64_t>(), s);
        if (e != hipSuccess) return e;
        if ((e = tmp.ensure(need)) != hipSuccess) return e;
        return rocprim::exclusive_scan(tmp.p, need, in, o, (int64_t)0, (size_t)len, rocprim::plus<int64_t>(), s);
    };
    int64_t *cnt = nullptr, *off = nullptr, *scnt = nullptr, *soff = nullptr, *didx = nullptr;
    double* rel = nullptr;
    HIP_TRY(c, b.alloc(&cnt, (size_t)n + 1));
    HIP_TRY(c, b.alloc(&off, (size_t)n + 1));
    HIP_TRY(c, b.alloc(&scnt, (size_t)n + 1));
    HIP_TRY(c, b.alloc(&soff, (size_t)n + 1));
    HIP_TRY(c, b.alloc(&didx, (size_t)n));
    HIP_TRY(c, b.alloc(&rel, (size_t)n));
    HIP_TRY(c, hipMemsetAsync(rel, 0, (size_t)n * 8, s));
    HIP_TRY(c, sort_dets(tmp, b, d_dets, n, nw, didx, s));
#define HCI_SD_WALK(FILL, ...)                                                                                   \
    do {                                                                                                          \
        if (wide)                                                                                                 \
            hipLaunchKernelGGL((k_sd_walk<u128, FILL>), grid_for(n, 256), 256, 0, s, D, T, L.vmask[0], L.vmask[1], \
                               d_dets, d_coef, n, eps, didx, __VA_ARGS__);                                        \
        else                                                                                                      \
            hipLaunchKernelGGL((k_sd_walk<uint64_t, FILL>), grid_for(n, 256), 256, 0, s, D, T, L.vmask[0],        \
                               L.vmask[1], d_dets, d_coef, n, eps, didx, __VA_ARGS__);                            \
    } while (0)
    HCI_SD_WALK(false, cnt, scnt, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
    HIP_TRY(c, hipMemsetAsync(cnt + n, 0, 8, s));
    HIP_TRY(c, hipMemsetAsync(scnt + n, 0, 8, s));
    HIP_TRY(c, scan(cnt, off, n + 1));
    HIP_TRY(c, scan(scnt, soff, n + 1));
    int64_t m = 0, ns = 0;
    HIP_TRY(c, hipMemcpyAsync(&m, off + n, 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipMemcpyAsync(&ns, soff + n, 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipStreamSynchronize(s));
    if (n_samples) *n_samples = ns;
    if (samples && ns > cap)
        return hci_fail(c, HCI_ERR_ARG, "hci_semistoch_dtm: sample capacity too small (n_samples holds the need)");
    uint64_t* cw = nullptr;
    double* cv = nullptr;
    int64_t* csrc = nullptr;
    uint32_t* ccode = nullptr;
    hci_det_orb_sample* d_smp = nullptr;
    HIP_TRY(c, b.alloc(&cw, (size_t)m * 4 + 4));
    HIP_TRY(c, b.alloc(&cv, (size_t)m + 1));
    HIP_TRY(c, b.alloc(&csrc, (size_t)m + 1));
    HIP_TRY(c, b.alloc(&ccode, (size_t)m + 1));
    if (samples && ns > 0) {
        if (mem == HCI_MEM_DEVICE)
            d_smp = samples;
        else
            HIP_TRY(c, b.alloc(&d_smp, (size_t)ns));
    }
    HCI_SD_WALK(true, nullptr, nullptr, off, soff, cw, cv, csrc, ccode, d_smp);
#undef HCI_SD_WALK
    HIP_TRY(c, hipGetLastError());
    double e_sum = 0.0;
    int64_t k = 0;
    if (m > 0) {
        int64_t *cidx = nullptr, *head = nullptr, *hpos = nullptr, *start = nullptr, *keep = nullptr, *kpos = nullptr,
                *frun = nullptr;
        double* rv = nullptr;
        HIP_TRY(c, b.alloc(&cidx, (size_t)m));
        HIP_TRY(c, sort_dets(tmp, b, cw, m, nw, cidx, s));
        HIP_TRY(c, b.alloc(&head, (size_t)m + 1));
        HIP_TRY(c, b.alloc(&hpos, (size_t)m + 1));
        hipLaunchKernelGGL(k_run_heads, grid_for(m, 256), 256, 0, s, cw, cidx, m, head);
        HIP_TRY(c, hipMemsetAsync(head + m, 0, 8, s));
        HIP_TRY(c, scan(head, hpos, m + 1));
        int64_t runs = 0;
        HIP_TRY(c, hipMemcpyAsync(&runs, hpos + m, 8, hipMemcpyDeviceToHost, s));
        HIP_TRY(c, hipStreamSynchronize(s));
        HIP_TRY(c, b.alloc(&start, (size_t)runs + 1));
        HIP_TRY(c, b.alloc(&rv, (size_t)runs));
        HIP_TRY(c, b.alloc(&frun, (size_t)runs));
        HIP_TRY(c, b.alloc(&keep, (size_t)runs + 1));
        HIP_TRY(c, b.alloc(&kpos, (size_t)runs + 1));
        hipLaunchKernelGGL(k_run_starts, grid_for(m + 1, 256), 256, 0, s, head, hpos, m, start);
        hipLaunchKernelGGL(k_sd_run, grid_for(runs, 256), 256, 0, s, cw, cv, csrc, ccode, cidx, start, runs, d_dets,
                           didx, n, rv, frun, keep, rel);
        HIP_TRY(c, hipMemsetAsync(keep + runs, 0, 8, s));
        HIP_TRY(c, scan(keep, kpos, runs + 1));
        HIP_TRY(c, hipMemcpyAsync(&k, kpos + runs, 8, hipMemcpyDeviceToHost, s));
        HIP_TRY(c, hipStreamSynchronize(s));
        if (k > 0) {
            // h_psi in first-discovery order (the order Wf::push gives it)
            int64_t *first = nullptr, *rid = nullptr, *first2 = nullptr, *rid2 = nullptr;
            HIP_TRY(c, b.alloc(&first, (size_t)k));
            HIP_TRY(c, b.alloc(&rid, (size_t)k));
            HIP_TRY(c, b.alloc(&first2, (size_t)k));
            HIP_TRY(c, b.alloc(&rid2, (size_t)k));
            hipLaunchKernelGGL(k_sd_first, grid_for(runs, 256), 256, 0, s, frun, keep, kpos, runs, first, rid);
            size_t need = 0;
            HIP_TRY(c, rocprim::radix_sort_pairs(nullptr, need, first, first2, rid, rid2, (size_t)k, 0, 64, s));
            HIP_TRY(c, tmp.ensure(need));
            HIP_TRY(c, rocprim::radix_sort_pairs(tmp.p, need, first, first2, rid, rid2, (size_t)k, 0, 64, s));
            double* part = nullptr;
            HIP_TRY(c, b.alloc(&part, (size_t)kPtBlocks));
            if (wide)
                hipLaunchKernelGGL(k_sd_terms<u128>, kPtBlocks, 256, 0, s, T, d_dets, d_diag, csrc, ccode, first2,
                                   rid2, rv, k, e0, part);
            else
                hipLaunchKernelGGL(k_sd_terms<uint64_t>, kPtBlocks, 256, 0, s, T, d_dets, d_diag, csrc, ccode, first2,
                                   rid2, rv, k, e0, part);
            HIP_TRY(c, hipGetLastError());
            std::vector<double> hp(kPtBlocks);
            HIP_TRY(c, hipMemcpyAsync(hp.data(), part, (size_t)kPtBlocks * 8, hipMemcpyDeviceToHost, s));
            HIP_TRY(c, hipStreamSynchronize(s));
            for (double x : hp) e_sum += x;
        }
    }
    if (rel_probs) HIP_TRY(c, hipMemcpyAsync(rel_probs, rel, (size_t)n * 8, out_kind, s));
    if (samples && ns > 0 && mem == HCI_MEM_HOST)
        HIP_TRY(c, hipMemcpyAsync(samples, d_smp, (size_t)ns * sizeof(hci_det_orb_sample), hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipStreamSynchronize(s));
    *e2 = e_sum;
    if (n_ext) *n_ext = k;
    return HCI_OK;
}

}  // extern "C"
