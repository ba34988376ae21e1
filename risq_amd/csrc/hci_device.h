// Device-side determinant bit operations and Slater-Condon elements.
//
// A spin string is S = uint64_t (n_orb <= 64) or u128 (n_orb <= 128).  Element
// functions keep the reference's operand order and summation order
// (src/ham/mod.rs:135-233) so their FP64 results are bit-identical to it: every
// addend is the same IEEE value (table entries, or one subtraction of two table
// entries done once on the host), added in the same order; sign factors are
// exact multiplications by +-1.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef unsigned __int128 u128;

__device__ __forceinline__ int popc(uint64_t x) { return __popcll(x); }
__device__ __forceinline__ int popc(u128 x) {
    return __popcll((uint64_t)x) + __popcll((uint64_t)(x >> 64));
}
// x != 0
__device__ __forceinline__ int ctz(uint64_t x) { return __builtin_ctzll(x); }
__device__ __forceinline__ int ctz(u128 x) {
    uint64_t lo = (uint64_t)x;
    return lo ? __builtin_ctzll(lo) : 64 + __builtin_ctzll((uint64_t)(x >> 64));
}

// parity of the set-bit count, +1 even / -1 odd (src/utils/bits.rs:171-180)
template <class S>
__device__ __forceinline__ int sgn_parity(S x) {
    return 1 - 2 * (popc(x) & 1);
}
// src/utils/ints.rs:76-82
template <class S>
__device__ __forceinline__ int permute(S a, S b) {
    return a > b ? sgn_parity(a & (a - b)) : sgn_parity(b & (b - a));
}
template <class S>
__device__ __forceinline__ S lowmask(int v) {
    return (((S)1) << v) - (S)1;
}

// Integral tables in device memory.
//   two : (pq|rs) in the reference's combine_4 layout (1-based)
//   h1  : h[p*n+q] dense
//   J3  : J3[(i*n+j)*n+k] = (ij|kk)             = Ham::direct(i,k,j,k)
//   D3  : D3[(i*n+j)*n+k] = (ij|kk) - (ik|kj)   = Ham::direct_plus_exchange(i,k,j,k)
struct DevInts {
    int n;
    double nuc;
    const double* two;
    const double* h1;
    const double* J3;
    const double* D3;
    // sok[i * n + j] = 0: the single excitation i -> j has a zero element for
    // every occupation (h_ij and every (ij|kk), (ij|kk) - (ik|kj) are zero)
    const uint8_t* sok;
    int sok_all;  // every single can be nonzero (sok is then not read)
};

// combine_2 on positive 1-based indices (src/utils/ints.rs:38-46)
__device__ __forceinline__ int c2(int i, int j) {
    return i > j ? ((i * (i - 1)) >> 1) + j : ((j * (j - 1)) >> 1) + i;
}
// Ham::get_int (src/ham/mod.rs:49-55), 1-based
__device__ __forceinline__ double get_int(const DevInts& T, int p, int q, int r, int s) {
    return T.two[c2(c2(p, q), c2(r, s))];
}

// Ham::ham_sing for a single excitation in spin sigma (src/ham/mod.rs:135-168):
// d1s/d2s are the sigma strings of det1/det2, d1o the other-spin string of det1.
template <class S>
__device__ __forceinline__ double elem_single(const DevInts& T, S d1s, S d2s, S d1o) {
    const int n = T.n;
    const int i = ctz(d1s & ~d2s);
    const int j = ctz(d2s & ~d1s);
    const double* D = T.D3 + (size_t)(i * n + j) * n;
    const double* J = T.J3 + (size_t)(i * n + j) * n;
    double out = T.h1[i * n + j];
    S m = d1s;
    while (m) {
        out += D[ctz(m)];
        m &= m - (S)1;
    }
    m = d1o;
    while (m) {
        out += J[ctz(m)];
        m &= m - (S)1;
    }
    return out * (double)permute(d1s, d2s);
}

// false when the single excitation from -> to (strings of the changed spin)
// has a zero element whatever the other electrons: ham_sing then sums only
// zero terms, so neither storage rule ever keeps the pair
template <class S>
__device__ __forceinline__ bool single_possible(const DevInts& T, S from, S to) {
    return T.sok_all || T.sok[ctz(from & ~to) * T.n + ctz(to & ~from)] != 0;
}

// Ham::ham_doub, same-spin branch (src/ham/mod.rs:180-207): a/b are the
// changed-spin strings of det1/det2.
template <class S>
__device__ __forceinline__ double elem_same_double(const DevInts& T, S a, S b) {
    S rem = a & ~b, add = b & ~a;
    const int p = ctz(rem);
    rem &= rem - (S)1;
    const int q = ctz(rem);
    const int r = ctz(add);
    add &= add - (S)1;
    const int s = ctz(add);
    const int sg = sgn_parity(a & b & (lowmask<S>(p) ^ lowmask<S>(q) ^ lowmask<S>(r) ^ lowmask<S>(s)));
    const double v = get_int(T, p + 1, r + 1, q + 1, s + 1) - get_int(T, p + 1, s + 1, q + 1, r + 1);
    return (double)sg * v;
}

// Ham::ham_doub, opposite-spin branch (src/ham/mod.rs:208-232)
template <class S>
__device__ __forceinline__ double elem_opp_double(const DevInts& T, S u1, S u2, S d1, S d2) {
    const int p = ctz(u1 & ~u2);
    const int r = ctz(u2 & ~u1);
    const int q = ctz(d1 & ~d2);
    const int s = ctz(d2 & ~d1);
    const int sg = permute(u1, u2) * permute(d1, d2);
    return (double)sg * get_int(T, p + 1, r + 1, q + 1, s + 1);
}

// Ham::ham_off_diag_no_excite (src/ham/mod.rs:256-303); *deg = excitation
// degree (3 = more than double)
template <class S>
__device__ __forceinline__ double elem_generic(const DevInts& T, S u1, S d1, S u2, S d2, int* deg) {
    const int ru = popc(u1 & ~u2), rd = popc(d1 & ~d2);
    *deg = (ru + rd) > 2 ? 3 : (ru + rd);
    const bool su = (u1 == u2), sd = (d1 == d2);
    if (su && sd) return 0.0;
    if (!su && !sd) {
        if (ru > 1 || rd > 1) return 0.0;
        return elem_opp_double(T, u1, u2, d1, d2);
    }
    if (su) {
        if (rd > 2) return 0.0;
        return rd == 1 ? elem_single(T, d1, d2, u1) : elem_same_double(T, d1, d2);
    }
    if (ru > 2) return 0.0;
    return ru == 1 ? elem_single(T, u1, u2, d1) : elem_same_double(T, u1, u2);
}

// Ham::ham_diag (src/ham/mod.rs:92-124)
template <class S>
__device__ __forceinline__ double elem_diag(const DevInts& T, S up, S dn) {
    const int n = T.n;
    double d = T.nuc;
    S m = up;
    while (m) {
        const int k = ctz(m);
        d += T.h1[k * n + k];
        m &= m - (S)1;
    }
    m = dn;
    while (m) {
        const int k = ctz(m);
        d += T.h1[k * n + k];
        m &= m - (S)1;
    }
    m = up;
    while (m) {
        const int i = ctz(m);
        m &= m - (S)1;
        S w = dn;
        while (w) {
            d += T.J3[(size_t)(i * n + i) * n + ctz(w)];  // (ii|jj)
            w &= w - (S)1;
        }
    }
    for (int spin = 0; spin < 2; ++spin) {
        S a = spin ? dn : up;
        while (a) {
            const int i = ctz(a);
            a &= a - (S)1;
            S b = a;
            while (b) {
                d += T.D3[(size_t)(i * n + i) * n + ctz(b)];  // (ii|jj) - (ij|ji)
                b &= b - (S)1;
            }
        }
    }
    return d;
}

template <class S>
__device__ __forceinline__ S load_string(const uint64_t* w);
template <>
__device__ __forceinline__ uint64_t load_string<uint64_t>(const uint64_t* w) {
    return w[0];
}
template <>
__device__ __forceinline__ u128 load_string<u128>(const uint64_t* w) {
    return ((u128)w[1] << 64) | (u128)w[0];
}
