// CPU oracle for the HCI variational hot path -- TEST INFRASTRUCTURE ONLY.
//
// A plain C++ restatement of the reference (aaholmes/risq, Rust) used solely as
// the result to compare against by tests/, __graft_entry__.smoke() and the
// CPU-baseline leg of bench.py.  The product library never links or calls it.  Every function cites
// the reference file:line it follows.  Compiled with -ffp-contract=off: Rust
// never fuses a*b+c, so neither may the restatement.
//
// Parity pins (what checks each function independently of this file):
//   parity/permute/permute_2 : the reference's parity known answers
//                              (src/utils/bits.rs:198-207) and brute-force
//                              second quantisation (tests/test_oracle_*.py)
//   combine_2/combine_4      : closed form + 8-fold symmetry invariant
//   ham_diag/ham_sing/ham_doub: brute-force second-quantised <Dj|H|Di> on tiny
//                              systems (tests/test_oracle_elements.py)
//   fast-lookup pattern      : brute-force all-pairs degree check
//   SpMV                     : dense H @ x
//   Davidson                 : dense eigh of the same H (numpy), 1e-8 Ha
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <unordered_map>
#include <unordered_set>
#include <utility>
#include <vector>

typedef unsigned __int128 u128;

namespace {

struct U128Hash {
    size_t operator()(const u128& v) const {
        uint64_t lo = (uint64_t)v, hi = (uint64_t)(v >> 64);
        uint64_t h = lo * 0x9E3779B97F4A7C15ull ^ (hi + 0x632BE59BD9B4E019ull + (lo << 6) + (lo >> 2));
        h ^= h >> 31;
        return (size_t)h;
    }
};

struct Config {
    u128 up, dn;
    bool operator==(const Config& o) const { return up == o.up && dn == o.dn; }
};
struct ConfigHash {
    size_t operator()(const Config& c) const {
        U128Hash h;
        return h(c.up) * 31u + h(c.dn);
    }
};

inline u128 mk(uint64_t lo, uint64_t hi) { return ((u128)hi << 64) | lo; }

inline Config load_det(const uint64_t* d) { return Config{mk(d[0], d[1]), mk(d[2], d[3])}; }

inline int count_ones(u128 n) {
    return __builtin_popcountll((uint64_t)n) + __builtin_popcountll((uint64_t)(n >> 64));
}
inline int trailing_zeros(u128 n) {
    uint64_t lo = (uint64_t)n;
    if (lo) return __builtin_ctzll(lo);
    uint64_t hi = (uint64_t)(n >> 64);
    if (hi) return 64 + __builtin_ctzll(hi);
    return 128;
}

// src/utils/bits.rs:14-25 -- ascending set-bit indices
template <class F>
inline void for_bits(u128 n, F f) {
    while (n) {
        int b = trailing_zeros(n);
        n &= ~((u128)1 << b);
        f(b);
    }
}

// src/utils/bits.rs:171-180 -- XOR-fold parity, +1 even / -1 odd
inline int parity(u128 n) {
    n ^= n >> 64;
    n ^= n >> 32;
    n ^= n >> 16;
    n ^= n >> 8;
    n ^= n >> 4;
    n ^= n >> 2;
    n ^= n >> 1;
    return 1 - 2 * (int)(n & 1);
}

// src/utils/ints.rs:76-82
inline int permute(u128 d1, u128 d2) {
    if (d1 > d2) return parity(d1 & (d1 - d2));
    return parity(d2 & (d2 - d1));
}

// src/utils/ints.rs:92-97
inline int permute_2(u128 d1, u128 d2, const int v[4]) {
    u128 m = ((((u128)1) << v[0]) - 1) ^ ((((u128)1) << v[1]) - 1) ^ ((((u128)1) << v[2]) - 1) ^
             ((((u128)1) << v[3]) - 1);
    return parity(d1 & d2 & m);
}

// src/utils/ints.rs:38-46 (1-based, sign ignored)
inline size_t combine_2(int p, int q) {
    int i = p < 0 ? -p : p;
    int j = q < 0 ? -q : q;
    if (i > j) return (size_t)((i * (i - 1)) / 2 + j);
    return (size_t)((j * (j - 1)) / 2 + i);
}
// src/utils/ints.rs:49-55
inline size_t combine_2_usize(size_t i, size_t j) {
    if (i > j) return (i * (i - 1)) / 2 + j;
    return (j * (j - 1)) / 2 + i;
}
// src/utils/ints.rs:66-68
inline size_t combine_4(int p, int q, int r, int s) {
    return combine_2_usize(combine_2(p, q), combine_2(r, s));
}

}  // namespace

struct or_ham {
    int n_orb;
    double nuc;
    std::vector<double> one_body;  // combine_2 layout
    std::vector<double> two_body;  // combine_4 layout

    // src/ham/mod.rs:49-55
    double get_int(int p, int q, int r, int s) const { return two_body[combine_4(p, q, r, s)]; }
    // src/ham/mod.rs:65-67
    double one(int p, int q) const { return one_body[combine_2(p + 1, q + 1)]; }
    // src/ham/mod.rs:72-74
    double direct(int p, int q, int r, int s) const { return get_int(p + 1, r + 1, q + 1, s + 1); }
    // src/ham/mod.rs:80-82
    double dpe(int p, int q, int r, int s) const {
        return get_int(p + 1, r + 1, q + 1, s + 1) - get_int(p + 1, s + 1, q + 1, r + 1);
    }

    // src/ham/mod.rs:92-124
    double ham_diag(const Config& d) const {
        double diag = nuc;
        for_bits(d.up, [&](int i) { diag += one(i, i); });
        for_bits(d.dn, [&](int i) { diag += one(i, i); });
        for_bits(d.up, [&](int i) { for_bits(d.dn, [&](int j) { diag += direct(i, j, i, j); }); });
        // bit_pairs (src/utils/bits.rs:31-52): (i,j), i<j, i outer ascending
        auto pairs = [&](u128 s) {
            std::vector<int> b;
            for_bits(s, [&](int x) { b.push_back(x); });
            for (size_t a = 0; a < b.size(); ++a)
                for (size_t c = a + 1; c < b.size(); ++c) diag += dpe(b[a], b[c], b[a], b[c]);
        };
        pairs(d.up);
        pairs(d.dn);
        return diag;
    }

    // src/ham/mod.rs:135-168
    double ham_sing(const Config& d1, const Config& d2) const {
        double out;
        if (d1.dn == d2.dn) {
            int i = trailing_zeros(d1.up & ~d2.up);
            int j = trailing_zeros(d2.up & ~d1.up);
            out = one(i, j);
            for_bits(d1.up, [&](int k) { out += dpe(i, k, j, k); });
            for_bits(d1.dn, [&](int k) { out += direct(i, k, j, k); });
            out *= (double)permute(d1.up, d2.up);
        } else {
            int i = trailing_zeros(d1.dn & ~d2.dn);
            int j = trailing_zeros(d2.dn & ~d1.dn);
            out = one(i, j);
            for_bits(d1.dn, [&](int k) { out += dpe(i, k, j, k); });
            for_bits(d1.up, [&](int k) { out += direct(i, k, j, k); });
            out *= (double)permute(d1.dn, d2.dn);
        }
        return out;
    }

    // src/ham/mod.rs:179-233
    double ham_doub(const Config& d1, const Config& d2) const {
        if (d1.dn == d2.dn || d1.up == d2.up) {
            u128 a = (d1.dn == d2.dn) ? d1.up : d1.dn;
            u128 b = (d1.dn == d2.dn) ? d2.up : d2.dn;
            int ind[4] = {0, 0, 0, 0};
            int n = 0;
            for_bits(a & ~b, [&](int x) { ind[n++] = x; });
            for_bits(b & ~a, [&](int x) { ind[n++] = x; });
            return (double)permute_2(a, b, ind) * dpe(ind[0], ind[1], ind[2], ind[3]);
        }
        int i1[2] = {0, 0}, i2[2] = {0, 0};
        int n = 0;
        for_bits(d1.up & ~d2.up, [&](int x) { i1[n++] = x; });
        for_bits(d2.up & ~d1.up, [&](int x) { i1[n++] = x; });
        n = 0;
        for_bits(d1.dn & ~d2.dn, [&](int x) { i2[n++] = x; });
        for_bits(d2.dn & ~d1.dn, [&](int x) { i2[n++] = x; });
        return (double)(permute(d1.up, d2.up) * permute(d1.dn, d2.dn)) *
               direct(i1[0], i2[0], i1[1], i2[1]);
    }

    // src/ham/mod.rs:256-303
    double off_diag_no_excite(const Config& d1, const Config& d2) const {
        bool su = d1.up == d2.up, sd = d1.dn == d2.dn;
        if (su && sd) return 0.0;
        if (!su && !sd) {
            if (count_ones(d1.up & ~d2.up) > 1) return 0.0;
            if (count_ones(d1.dn & ~d2.dn) > 1) return 0.0;
            return ham_doub(d1, d2);
        }
        u128 rem = su ? (d1.dn & ~d2.dn) : (d1.up & ~d2.up);
        int n = count_ones(rem);
        if (n > 2) return 0.0;
        if (n == 1) return ham_sing(d1, d2);
        return ham_doub(d1, d2);
    }
};

// src/var/sparse.rs:24-36
struct or_sparse {
    int64_t n = 0;
    std::vector<double> diag;
    std::vector<size_t> nnz;
    std::vector<std::vector<std::pair<int64_t, double>>> off_diag;

    // src/var/sparse.rs:44-56
    void sort_remove_duplicates() {
        for (int64_t i = 0; i < n; ++i) {
            auto& v = off_diag[i];
            std::stable_sort(v.begin() + (ptrdiff_t)nnz[i], v.end(),
                             [](const std::pair<int64_t, double>& a,
                                const std::pair<int64_t, double>& b) { return a.first < b.first; });
            // dedup_by_key: drop consecutive entries with an equal column, keep the first
            size_t w = 0;
            for (size_t r = 0; r < v.size(); ++r) {
                if (w > 0 && v[w - 1].first == v[r].first) continue;
                v[w++] = v[r];
            }
            v.resize(w);
            nnz[i] = v.size();
        }
    }
};

namespace {

// add_el / add_off_diag_elem, src/var/off_diag.rs:26-53
inline void add_el(or_sparse* h, const or_ham* ham, const std::vector<Config>& dets, int64_t i,
                   int64_t j) {
    if (i == j) return;
    double e = ham->off_diag_no_excite(dets[i], dets[j]);
    if (e != 0.0) {
        if (i < j)
            h->off_diag[i].push_back({j, e});
        else
            h->off_diag[j].push_back({i, e});
    }
}

// are_strings_double_connected, src/var/fast_ham.rs:321-336
inline bool double_connected(u128 a, u128 b) {
    if (count_ones(a ^ b) != 4) return false;
    return count_ones(a) == count_ones(b);
}

// build_alpha_singles / build_beta_singles, src/var/fast_ham.rs:105-173
std::unordered_map<u128, std::vector<u128>, U128Hash> build_singles(const std::vector<u128>& uniq) {
    std::unordered_map<u128, std::vector<u128>, U128Hash> r1map, out;
    for (u128 s : uniq)
        for_bits(s, [&](int o) { r1map[s & ~((u128)1 << o)].push_back(s); });
    for (u128 s : uniq) {
        std::vector<u128> conn;
        for_bits(s, [&](int o) {
            auto it = r1map.find(s & ~((u128)1 << o));
            if (it != r1map.end())
                for (u128 c : it->second)
                    if (c != s) conn.push_back(c);
        });
        out[s] = std::move(conn);
    }
    return out;
}

}  // namespace

extern "C" {

int or_parity(uint64_t lo, uint64_t hi) { return parity(mk(lo, hi)); }
int or_permute(const uint64_t* d1, const uint64_t* d2) {
    return permute(mk(d1[0], d1[1]), mk(d2[0], d2[1]));
}
int or_permute_2(const uint64_t* d1, const uint64_t* d2, const int32_t* v) {
    int vv[4] = {v[0], v[1], v[2], v[3]};
    return permute_2(mk(d1[0], d1[1]), mk(d2[0], d2[1]), vv);
}
uint64_t or_combine_2(int32_t p, int32_t q) { return combine_2(p, q); }
uint64_t or_combine_4(int32_t p, int32_t q, int32_t r, int32_t s) { return combine_4(p, q, r, s); }

or_ham* or_ham_create(int32_t n_orb, double e_core, const double* one_body, size_t n1,
                      const double* two_body, size_t n2) {
    or_ham* h = new or_ham();
    h->n_orb = n_orb;
    h->nuc = e_core;
    h->one_body.assign(one_body, one_body + n1);
    h->two_body.assign(two_body, two_body + n2);
    return h;
}
void or_ham_free(or_ham* h) { delete h; }

void or_ham_diag(const or_ham* h, const uint64_t* dets, int64_t n, double* out) {
    for (int64_t i = 0; i < n; ++i) out[i] = h->ham_diag(load_det(dets + 4 * i));
}

void or_off_diag(const or_ham* h, const uint64_t* det1, const uint64_t* det2, int64_t m,
                 double* value, int8_t* degree) {
    for (int64_t k = 0; k < m; ++k) {
        Config a = load_det(det1 + 4 * k), b = load_det(det2 + 4 * k);
        int deg = count_ones(a.up & ~b.up) + count_ones(a.dn & ~b.dn);
        if (degree) degree[k] = (int8_t)(deg > 2 ? 3 : deg);
        value[k] = h->off_diag_no_excite(a, b);
    }
}

or_sparse* or_sparse_create(int64_t n, const double* diag) {
    or_sparse* s = new or_sparse();
    s->n = n;
    s->diag.assign(diag, diag + n);
    s->nnz.assign((size_t)n, 0);
    s->off_diag.resize((size_t)n);
    return s;
}

// gen_sparse_ham_fast_lookup, src/var/fast_ham.rs:527-575, with
// VariationalSpace::new (:58-101) and find_all_connections_original_pattern (:341-366)
int or_build_fast_lookup(or_sparse* h, const or_ham* ham, const uint64_t* dets_w, int64_t n,
                         const double* diag, int64_t n_stored) {
    if (n < h->n || n_stored > h->n) return -1;
    // expand_sparse_ham_rows, src/wf/mod.rs:1664-1683
    for (int64_t i = h->n; i < n; ++i) {
        h->diag.push_back(diag[i]);
        h->nnz.push_back(0);
        h->off_diag.emplace_back();
    }
    h->n = n;

    std::vector<Config> dets((size_t)n);
    for (int64_t i = 0; i < n; ++i) dets[i] = load_det(dets_w + 4 * i);

    // group determinants by alpha string, det indices ascending (:64-83)
    std::vector<u128> alpha_strings;
    std::unordered_map<u128, size_t, U128Hash> alpha_to_index;
    std::vector<std::vector<int64_t>> alpha_to_dets;
    for (int64_t i = 0; i < n; ++i) {
        auto it = alpha_to_index.find(dets[i].up);
        if (it == alpha_to_index.end()) {
            alpha_to_index[dets[i].up] = alpha_strings.size();
            alpha_strings.push_back(dets[i].up);
            alpha_to_dets.emplace_back();
            alpha_to_dets.back().push_back(i);
        } else {
            alpha_to_dets[it->second].push_back(i);
        }
    }
    auto alpha_singles = build_singles(alpha_strings);
    std::vector<u128> uniq_beta;
    {
        std::unordered_set<u128, U128Hash> seen;
        for (auto& d : dets)
            if (seen.insert(d.dn).second) uniq_beta.push_back(d.dn);
    }
    auto beta_singles = build_singles(uniq_beta);
    auto beta_single_connected = [&](u128 b1, u128 b2) {  // :310-316
        auto it = beta_singles.find(b1);
        if (it == beta_singles.end()) return false;
        return std::find(it->second.begin(), it->second.end(), b2) != it->second.end();
    };

    // new alpha strings (:347-352)
    std::vector<u128> new_alpha;
    {
        std::unordered_set<u128, U128Hash> seen;
        for (int64_t i = n_stored; i < n; ++i)
            if (seen.insert(dets[i].up).second) new_alpha.push_back(dets[i].up);
    }

    // find_single_excitations, :473-517
    for (int64_t src = n_stored; src < n; ++src) {
        const Config& s = dets[src];
        for (int64_t tgt = 0; tgt < n; ++tgt) {
            if (src == tgt) continue;
            const Config& t = dets[tgt];
            int ud = count_ones(s.up ^ t.up), dd = count_ones(s.dn ^ t.dn);
            if (!((ud == 2 && dd == 0) || (ud == 0 && dd == 2))) continue;
            if (src < tgt) {
                double e = ham->off_diag_no_excite(s, t);
                if (std::fabs(e) > 1e-12) h->off_diag[src].push_back({tgt, e});
            } else {
                double e = ham->off_diag_no_excite(t, s);
                if (std::fabs(e) > 1e-12) h->off_diag[tgt].push_back({src, e});
            }
        }
    }

    // find_opposite_spin_for_alpha_string, :369-407
    for (u128 a : new_alpha) {
        auto sit = alpha_singles.find(a);
        if (sit == alpha_singles.end()) continue;
        const auto& src_dets = alpha_to_dets[alpha_to_index[a]];
        for (u128 b : sit->second) {
            auto bit = alpha_to_index.find(b);
            if (bit == alpha_to_index.end()) continue;
            const auto& con_dets = alpha_to_dets[bit->second];
            for (int64_t i : src_dets)
                for (int64_t j : con_dets)
                    if (beta_single_connected(dets[i].dn, dets[j].dn))
                        if (i >= n_stored || j >= n_stored) add_el(h, ham, dets, i, j);
        }
    }

    // find_same_spin_for_alpha_string, :410-470
    for (u128 a : alpha_strings) {
        const auto& ad = alpha_to_dets[alpha_to_index[a]];
        bool has_new = false;
        for (int64_t d : ad) has_new |= (d >= n_stored);
        if (!has_new) continue;
        for (size_t x = 0; x < ad.size(); ++x)
            for (size_t y = x + 1; y < ad.size(); ++y) {
                int64_t d1 = ad[x], d2 = ad[y];
                if (d1 >= n_stored || d2 >= n_stored)
                    if (double_connected(dets[d1].dn, dets[d2].dn)) add_el(h, ham, dets, d1, d2);
            }
        std::vector<int64_t> new_dets;
        for (int64_t d : ad)
            if (d >= n_stored) new_dets.push_back(d);
        for (u128 o : alpha_strings) {
            if (o == a || !double_connected(a, o)) continue;
            const auto& od = alpha_to_dets[alpha_to_index[o]];
            for (int64_t nd : new_dets)
                for (int64_t x : od)
                    if (dets[nd].dn == dets[x].dn) add_el(h, ham, dets, nd, x);
        }
    }

    h->sort_remove_duplicates();
    return 0;
}

// gen_sparse_ham_fast with opp_algo = same_algo = 1 (the reference's
// defaults, src/config.rs:98-99), the build sparse_optimize uses while the
// space holds at most 100 determinants (src/var/davidson.rs:55-61):
// src/var/ham_gen.rs:60-292 (set-up), :374-433 (opposite-spin algo 1),
// :621-637 (same-spin algo 1), with add_el / add_el_and_spin_flipped
// (src/var/off_diag.rs:26-102) and remove_1e (src/var/utils.rs:10-12).
// Needs a space closed under spin flip: the reference panics when the partner
// of a determinant is missing (off_diag.rs:88-98); here that returns -2.
int or_build_small(or_sparse* h, const or_ham* ham, const uint64_t* dets_w, int64_t n, const double* diag,
                   int64_t n_stored) {
    if (n < h->n || n_stored > h->n) return -1;
    std::vector<Config> dets((size_t)n);
    std::unordered_map<Config, int64_t, ConfigHash> inds;
    for (int64_t i = 0; i < n; ++i) {
        dets[i] = load_det(dets_w + 4 * i);
        inds[dets[i]] = i;
    }
    // 1. up string -> [(index, dn)] for all and for new determinants (:73-90)
    typedef std::vector<std::pair<int64_t, u128>> DetList;
    std::unordered_map<u128, DetList, U128Hash> unique_up_dict, new_unique_up_dict;
    std::vector<u128> up_order, new_up_order;  // first-seen order of the keys
    for (int64_t i = 0; i < n; ++i) {
        if (unique_up_dict.find(dets[i].up) == unique_up_dict.end()) up_order.push_back(dets[i].up);
        unique_up_dict[dets[i].up].push_back({i, dets[i].dn});
        if (i >= n_stored) {
            if (new_unique_up_dict.find(dets[i].up) == new_unique_up_dict.end()) new_up_order.push_back(dets[i].up);
            new_unique_up_dict[dets[i].up].push_back({i, dets[i].dn});
        }
    }
    // 2. unique ups sorted by decreasing number of determinants (:97-118)
    auto by_count = [&](std::vector<u128>& v, std::unordered_map<u128, DetList, U128Hash>& m) {
        std::stable_sort(v.begin(), v.end(), [&](u128 a, u128 b) { return m[a].size() > m[b].size(); });
    };
    by_count(up_order, unique_up_dict);
    by_count(new_up_order, new_unique_up_dict);
    // 4. up_singles: every new unique up -> the unique ups one electron away (:130-157)
    std::unordered_map<u128, std::vector<size_t>, U128Hash> up_single_excite_constructor;
    for (size_t ind = 0; ind < up_order.size(); ++ind)
        for_bits(up_order[ind], [&](int o) {
            up_single_excite_constructor[up_order[ind] & ~((u128)1 << o)].push_back(ind);
        });
    std::unordered_map<u128, std::vector<u128>, U128Hash> up_singles;
    for (u128 u : new_up_order)
        for_bits(u, [&](int o) {
            auto it = up_single_excite_constructor.find(u & ~((u128)1 << o));
            if (it == up_single_excite_constructor.end()) return;
            for (size_t ci : it->second) {
                if (u == up_order[ci]) continue;  // the diagonal
                up_singles[u].push_back(up_order[ci]);
            }
        });
    // new_sparse_ham / expand_sparse_ham_rows (src/wf/mod.rs:1646-1683)
    for (int64_t i = h->n; i < n; ++i) {
        h->diag.push_back(diag[i]);
        h->nnz.push_back(0);
        h->off_diag.emplace_back();
    }
    h->n = n;
    auto add = [&](int64_t i, int64_t j, double e) {  // add_off_diag_elem, off_diag.rs:45-53
        if (i < j)
            h->off_diag[i].push_back({j, e});
        else
            h->off_diag[j].push_back({i, e});
    };
    // opposite-spin algo 1 (:374-433): the dn strings of the new determinants
    // of this up string, looked up as up strings (the space is closed under
    // spin flip, so every dn string is also an up string)
    for (u128 u : new_up_order) {
        auto ups = up_singles.find(u);
        if (ups == up_singles.end()) continue;
        std::unordered_map<u128, std::vector<int64_t>, U128Hash> dn_candidates;
        for (auto& dn : new_unique_up_dict[u]) {
            auto dns = up_singles.find(dn.second);
            if (dns == up_singles.end()) continue;
            for (u128 dn2 : dns->second) dn_candidates[dn2].push_back(dn.first);
        }
        for (u128 up : ups->second)
            for (auto& dn : unique_up_dict[up]) {
                auto cand = dn_candidates.find(dn.second);
                if (cand == dn_candidates.end()) continue;
                const int64_t ind1 = inds[Config{up, dn.second}];
                for (int64_t c : cand->second)
                    if (c >= n_stored) add_el(h, ham, dets, ind1, c);
            }
    }
    // same-spin algo 1 (:621-637): pairs of one up string, at least one new,
    // each stored with its spin-flipped pair (add_el_and_spin_flipped)
    for (u128 u : up_order) {
        auto nd = new_unique_up_dict.find(u);
        if (nd == new_unique_up_dict.end()) continue;
        for (auto& a : unique_up_dict[u])
            for (auto& b : nd->second) {
                if (!(b.first > a.first)) continue;
                const double e = ham->off_diag_no_excite(dets[a.first], dets[b.first]);
                if (e == 0.0) continue;
                add(a.first, b.first, e);
                auto fa = inds.find(Config{dets[a.first].dn, dets[a.first].up});
                auto fb = inds.find(Config{dets[b.first].dn, dets[b.first].up});
                if (fa == inds.end() || fb == inds.end()) return -2;
                if (fa->second != fb->second) add(fa->second, fb->second, e);
            }
    }
    h->sort_remove_duplicates();
    return 0;
}

int64_t or_sparse_n(const or_sparse* s) { return s->n; }
int64_t or_sparse_nnz(const or_sparse* s) {
    int64_t t = 0;
    for (auto& r : s->off_diag) t += (int64_t)r.size();
    return t;
}
void or_sparse_export(const or_sparse* s, int64_t* row_ptr, int32_t* col, double* val,
                      double* diag) {
    int64_t p = 0;
    for (int64_t i = 0; i < s->n; ++i) {
        row_ptr[i] = p;
        for (auto& e : s->off_diag[i]) {
            col[p] = (int32_t)e.first;
            val[p] = e.second;
            ++p;
        }
        if (diag) diag[i] = s->diag[i];
    }
    row_ptr[s->n] = p;
}
void or_sparse_free(or_sparse* s) { delete s; }

// SparseMatUpperTri::matrix_vector_prod, src/var/sparse.rs:64-78
void or_spmv(const or_sparse* s, const double* x, double* y) {
    for (int64_t i = 0; i < s->n; ++i) y[i] = 0.0;
    for (int64_t r = 0; r < s->n; ++r) {
        y[r] += s->diag[r] * x[r];
        for (auto& e : s->off_diag[r]) {
            y[r] += e.second * x[e.first];
            y[e.first] += e.second * x[r];
        }
    }
}

void or_spmv_csr(int64_t n, const int64_t* row_ptr, const int32_t* col, const double* val,
                 const double* diag, const double* x, double* y) {
    for (int64_t i = 0; i < n; ++i) y[i] = 0.0;
    for (int64_t r = 0; r < n; ++r) {
        y[r] += diag[r] * x[r];
        double xr = x[r];
        for (int64_t p = row_ptr[r]; p < row_ptr[r + 1]; ++p) {
            y[r] += val[p] * x[col[p]];
            y[col[p]] += val[p] * xr;
        }
    }
}

// Cyclic Jacobi rotations; stands in for nalgebra 0.26.2 SymmetricEigen +
// sort_eigenpairs (src/var/eigenvalues/utils.rs:48-79).
void or_sym_eigen(int k, const double* a_in, double* vals, double* vecs) {
    std::vector<double> a(a_in, a_in + (size_t)k * k), v((size_t)k * k, 0.0);
    for (int i = 0; i < k; ++i) v[(size_t)i * k + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0;
        for (int p = 0; p < k; ++p)
            for (int q = p + 1; q < k; ++q) off += a[(size_t)p * k + q] * a[(size_t)p * k + q];
        if (off < 1e-300) break;
        for (int p = 0; p < k; ++p)
            for (int q = p + 1; q < k; ++q) {
                double apq = a[(size_t)p * k + q];
                if (apq == 0.0) continue;
                double app = a[(size_t)p * k + p], aqq = a[(size_t)q * k + q];
                double theta = (aqq - app) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int r = 0; r < k; ++r) {
                    double arp = a[(size_t)r * k + p], arq = a[(size_t)r * k + q];
                    a[(size_t)r * k + p] = c * arp - s * arq;
                    a[(size_t)r * k + q] = s * arp + c * arq;
                }
                for (int r = 0; r < k; ++r) {
                    double apr = a[(size_t)p * k + r], aqr = a[(size_t)q * k + r];
                    a[(size_t)p * k + r] = c * apr - s * aqr;
                    a[(size_t)q * k + r] = s * apr + c * aqr;
                }
                for (int r = 0; r < k; ++r) {
                    double vrp = v[(size_t)r * k + p], vrq = v[(size_t)r * k + q];
                    v[(size_t)r * k + p] = c * vrp - s * vrq;
                    v[(size_t)r * k + q] = s * vrp + c * vrq;
                }
            }
    }
    std::vector<int> idx(k);
    for (int i = 0; i < k; ++i) idx[i] = i;
    std::sort(idx.begin(), idx.end(),
              [&](int x, int y) { return a[(size_t)x * k + x] < a[(size_t)y * k + y]; });
    for (int c = 0; c < k; ++c) {
        vals[c] = a[(size_t)idx[c] * k + idx[c]];
        for (int r = 0; r < k; ++r) vecs[(size_t)r * k + c] = v[(size_t)r * k + idx[c]];
    }
}

namespace {

typedef void (*spmv_fn)(const void* ctx, const double* x, double* y);

// Davidson::new for nvalues=1, DPR, SpectrumTarget::Lowest
// (src/var/eigenvalues/algorithms/davidson.rs:53-277, 301-338, 391-408, 496-500;
//  MGS::orthonormalize src/var/eigenvalues/modified_gram_schmidt.rs:22-42)
int davidson_core(int64_t N, const double* d, spmv_fn spmv, const void* ctx, const double* init,
                  double tol, double etol, int max_iters, double* e0, double* coeffs, int* iters,
                  int* n_sigma) {
    const int nvalues = 1;
    const int init_dim = init ? 1 : 2;
    const int max_sub = (nvalues + 10 < N) ? nvalues + 10 : (int)N;
    int nsig = 0;
    std::vector<std::vector<double>> V((size_t)max_sub, std::vector<double>((size_t)N, 0.0));
    std::vector<std::vector<double>> W;
    for (int c = 0; c < max_sub; ++c) V[c][c] = 1.0;  // identity(N, max_sub)
    if (init) std::copy(init, init + N, V[0].begin());
    // The reference takes basis.columns(0, init_dim) of an N x max_sub basis
    // and panics when N = 1 without a start vector; the subspace is clamped
    // to N here (a one-determinant space returns H_00 and c = [1]).
    int dim_sub = init_dim < N ? init_dim : (int)N;
    auto dot = [&](const std::vector<double>& a, const std::vector<double>& b) {
        double s = 0.0;
        for (int64_t i = 0; i < N; ++i) s += a[i] * b[i];
        return s;
    };
    auto sigma = [&](const std::vector<double>& v) {
        std::vector<double> w((size_t)N);
        spmv(ctx, v.data(), w.data());
        ++nsig;
        return w;
    };
    for (int c = 0; c < dim_sub; ++c) W.push_back(sigma(V[c]));
    std::vector<double> proj((size_t)dim_sub * dim_sub);
    for (int r = 0; r < dim_sub; ++r)
        for (int c = 0; c < dim_sub; ++c) proj[(size_t)r * dim_sub + c] = dot(V[r], W[c]);
    double last_eig = 0.0;
    int rc = 1;
    std::vector<double> ritz((size_t)N), res((size_t)N);
    int it;
    for (it = 0; it < max_iters; ++it) {
        std::vector<double> vals(dim_sub), vecs((size_t)dim_sub * dim_sub);
        or_sym_eigen(dim_sub, proj.data(), vals.data(), vecs.data());
        double lam = vals[0];
        // Ritz vector 0 and residual 0 (only column 0 is ever used for nvalues=1)
        for (int64_t i = 0; i < N; ++i) {
            double x = 0.0, w = 0.0;
            for (int c = 0; c < dim_sub; ++c) {
                x += V[c][i] * vecs[(size_t)c * dim_sub];
                w += W[c][i] * vecs[(size_t)c * dim_sub];
            }
            ritz[i] = x;
            res[i] = w - x * lam;
        }
        double err = std::sqrt(dot(res, res));
        auto finish = [&]() {
            *e0 = lam;
            std::copy(ritz.begin(), ritz.end(), coeffs);
            rc = 0;
        };
        if (N <= 2 || (err < tol && std::fabs(lam - last_eig) < etol)) {
            finish();
            break;
        }
        last_eig = lam;
        if (dim_sub + 1 <= max_sub) {
            std::vector<double>& t = V[dim_sub];
            for (int64_t i = 0; i < N; ++i) t[i] = res[i] / (lam - d[i]);
            if (std::sqrt(dot(t, t)) < 1e-9) {
                finish();
                break;
            }
            for (int j = 0; j < dim_sub; ++j) {
                double m = dot(V[j], t) / dot(V[j], V[j]);
                for (int64_t i = 0; i < N; ++i) t[i] = t[i] - V[j][i] * m;
            }
            double nrm = std::sqrt(dot(t, t));
            if (nrm < 1e-9) {
                finish();
                break;
            }
            for (int64_t i = 0; i < N; ++i) t[i] /= nrm;
            if ((int)W.size() > dim_sub)
                W[dim_sub] = sigma(t);
            else
                W.push_back(sigma(t));
            int nd = dim_sub + 1;
            std::vector<double> np((size_t)nd * nd);
            for (int r = 0; r < dim_sub; ++r)
                for (int c = 0; c < dim_sub; ++c) np[(size_t)r * nd + c] = proj[(size_t)r * dim_sub + c];
            for (int r = 0; r < nd; ++r) {
                double b = dot(V[r], W[dim_sub]);
                np[(size_t)r * nd + dim_sub] = b;
                np[(size_t)dim_sub * nd + r] = b;
            }
            proj.swap(np);
            dim_sub = nd;
        } else {
            dim_sub = 1;
            for (auto& col : V) std::fill(col.begin(), col.end(), 0.0);
            V[0] = ritz;
            W.clear();
            W.push_back(sigma(V[0]));
            proj.assign(1, dot(V[0], W[0]));
        }
    }
    *iters = it;
    *n_sigma = nsig;
    return rc;
}

struct CsrCtx {
    int64_t n;
    const int64_t* rp;
    const int32_t* col;
    const double* val;
    const double* diag;
};
void csr_spmv_cb(const void* c, const double* x, double* y) {
    const CsrCtx* k = (const CsrCtx*)c;
    or_spmv_csr(k->n, k->rp, k->col, k->val, k->diag, x, y);
}
void sparse_spmv_cb(const void* c, const double* x, double* y) {
    or_spmv((const or_sparse*)c, x, y);
}

}  // namespace

int or_davidson(const or_sparse* s, const double* init, double tol, double etol, int max_iters,
                double* e0, double* coeffs, int* iters, int* n_sigma) {
    return davidson_core(s->n, s->diag.data(), sparse_spmv_cb, s, init, tol, etol, max_iters, e0,
                         coeffs, iters, n_sigma);
}

int or_davidson_csr(int64_t n, const int64_t* row_ptr, const int32_t* col, const double* val,
                    const double* diag, const double* init, double tol, double etol,
                    int max_iters, double* e0, double* coeffs, int* iters, int* n_sigma) {
    CsrCtx c{n, row_ptr, col, val, diag};
    return davidson_core(n, diag, csr_spmv_cb, &c, init, tol, etol, max_iters, e0, coeffs, iters,
                         n_sigma);
}

}  // extern "C"
