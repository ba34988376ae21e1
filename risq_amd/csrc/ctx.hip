// Context, integral tables, per-determinant and per-pair element entry points.
#include <algorithm>
#include <cstring>
#include <thread>

#include "hci_internal.h"
#include "mlp.h"

int hci_fail(hci_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return code;
}

namespace {

template <class S>
__global__ void __launch_bounds__(256) k_ham_diag(DevInts T, const uint64_t* dets, int64_t n,
                                                  double* out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t* w = dets + 4 * i;
    out[i] = elem_diag<S>(T, load_string<S>(w), load_string<S>(w + 2));
}

template <class S>
__global__ void __launch_bounds__(256) k_elements(DevInts T, const uint64_t* a, const uint64_t* b,
                                                  int64_t m, double* val, int8_t* deg) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    int d = 0;
    const uint64_t* x = a + 4 * k;
    const uint64_t* y = b + 4 * k;
    val[k] = elem_generic<S>(T, load_string<S>(x), load_string<S>(x + 2), load_string<S>(y),
                             load_string<S>(y + 2), &d);
    if (deg) deg[k] = (int8_t)d;
}

// combine_2 / combine_4 (src/utils/ints.rs:38-68), 1-based
inline size_t hc2(size_t i, size_t j) { return i > j ? (i * (i - 1)) / 2 + j : (j * (j - 1)) / 2 + i; }
inline size_t hc4(size_t p, size_t q, size_t r, size_t s) { return hc2(hc2(p, q), hc2(r, s)); }

// Bit check of the word layout against n_orb: words above n_orb must be zero.
bool words_fit(const uint64_t* w, int64_t n, int n_orb) {
    uint64_t lo_mask = n_orb >= 64 ? ~0ull : ((1ull << n_orb) - 1);
    uint64_t hi_mask = n_orb <= 64 ? 0ull : (n_orb >= 128 ? ~0ull : ((1ull << (n_orb - 64)) - 1));
    for (int64_t i = 0; i < n; ++i)
        for (int s = 0; s < 2; ++s)
            if ((w[4 * i + 2 * s] & ~lo_mask) || (w[4 * i + 2 * s + 1] & ~hi_mask)) return false;
    return true;
}

}  // namespace

hipError_t recycled_take(hci_ctx* c, void** p, size_t bytes) {
    int best = -1;
    for (int i = 0; i < (int)c->recycled.size(); ++i)
        if (c->recycled[i].bytes >= bytes && (best < 0 || c->recycled[i].bytes < c->recycled[best].bytes)) best = i;
    if (best >= 0) {
        *p = c->recycled[best].p;
        c->recycled.erase(c->recycled.begin() + best);
        return hipSuccess;
    }
    recycled_flush(c);  // nothing fits: give the memory back before allocating
    hipError_t e = hipMalloc(p, bytes);
    if (e == hipErrorOutOfMemory && !c->scratch_cache.empty()) {
        (void)hipGetLastError();
        scratch_flush(c);  // the kept build work buffers
        e = hipMalloc(p, bytes);
    }
    if (e == hipErrorOutOfMemory && c->work_dev) {
        (void)hipGetLastError();
        work_release(c);  // the Davidson work vectors are cheaper to allocate again
        e = hipMalloc(p, bytes);
    }
    return e;
}

void recycled_flush(hci_ctx* c) {
    for (auto& r : c->recycled) hipFree(r.p);
    c->recycled.clear();
}

hipError_t scratch_take(hci_ctx* c, void** p, size_t bytes) {
    int best = -1;
    for (int i = 0; i < (int)c->scratch_cache.size(); ++i) {
        const size_t b = c->scratch_cache[i].bytes;
        if (b >= bytes && b - bytes <= bytes / 4 && (best < 0 || b < c->scratch_cache[best].bytes)) best = i;
    }
    if (best >= 0) {
        *p = c->scratch_cache[best].p;
        c->scratch_cached -= c->scratch_cache[best].bytes;
        c->scratch_cache.erase(c->scratch_cache.begin() + best);
        return hipSuccess;
    }
    hipError_t e = hipMalloc(p, bytes);
    if (e == hipErrorOutOfMemory && !c->scratch_cache.empty()) {
        (void)hipGetLastError();
        scratch_flush(c);
        e = hipMalloc(p, bytes);
    }
    return e;
}

void scratch_give(hci_ctx* c, void* p, size_t bytes) {
    if (!p) return;
    if (bytes <= kScratchCacheBuf && c->scratch_cached + bytes <= kScratchCacheMax) {
        c->scratch_cache.push_back({p, bytes});
        c->scratch_cached += bytes;
    } else {
        hipFree(p);
    }
}

void scratch_flush(hci_ctx* c) {
    for (auto& r : c->scratch_cache) hipFree(r.p);
    c->scratch_cache.clear();
    c->scratch_cached = 0;
}

extern "C" {

int hci_create(int device, hci_ctx** out) {
    if (!out) return HCI_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return HCI_ERR_HIP;
    if (device < 0 || device >= ndev) return HCI_ERR_ARG;
    hci_ctx* c = new hci_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess) {
        delete c;
        return HCI_ERR_HIP;
    }
    if (sigma_scratch_init(c) != HCI_OK) {
        hipStreamDestroy(c->stream);
        delete c;
        return HCI_ERR_HIP;
    }
    *out = c;
    return HCI_OK;
}

void hci_destroy(hci_ctx* c) {
    if (!c) return;
    hipSetDevice(c->device);
    recycled_flush(c);
    scratch_flush(c);
    work_release(c);
    mlp_free_state(c);
    select_free(c);
    hipFree(c->d_two);
    hipFree(c->d_h1);
    hipFree(c->d_J3);
    hipFree(c->d_D3);
    hipFree(c->d_sok);
    for (auto& e : c->events)
        if (e) hipEventDestroy(e);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

int hci_dev_alloc(hci_ctx* c, size_t bytes, void** ptr) {
    if (!c || !ptr) return HCI_ERR_ARG;
    *ptr = nullptr;
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipMalloc(ptr, bytes ? bytes : 16));
    return HCI_OK;
}

int hci_dev_free(hci_ctx* c, void* ptr) {
    if (!c) return HCI_ERR_ARG;
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipFree(ptr));
    return HCI_OK;
}

int hci_memcpy(hci_ctx* c, void* dst, const void* src, size_t bytes, int dst_mem, int src_mem) {
    if (!c || (bytes && (!dst || !src))) return HCI_ERR_ARG;
    HIP_TRY(c, hipSetDevice(c->device));
    hipMemcpyKind k = dst_mem == HCI_MEM_DEVICE
                          ? (src_mem == HCI_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice)
                          : (src_mem == HCI_MEM_DEVICE ? hipMemcpyDeviceToHost : hipMemcpyHostToHost);
    HIP_TRY(c, hipMemcpyAsync(dst, src, bytes, k, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return HCI_OK;
}

int hci_event_record(hci_ctx* c, int slot) {
    if (!c || slot < 0 || slot >= 16) return HCI_ERR_ARG;
    HIP_TRY(c, hipSetDevice(c->device));
    if (!c->events[slot]) HIP_TRY(c, hipEventCreate(&c->events[slot]));
    HIP_TRY(c, hipEventRecord(c->events[slot], c->stream));
    return HCI_OK;
}

int hci_event_elapsed_ms(hci_ctx* c, int a, int b, float* ms) {
    if (!c || !ms || a < 0 || a >= 16 || b < 0 || b >= 16 || !c->events[a] || !c->events[b])
        return c ? hci_fail(c, HCI_ERR_ARG, "hci_event_elapsed_ms: unrecorded slot") : HCI_ERR_ARG;
    HIP_TRY(c, hipEventSynchronize(c->events[b]));
    HIP_TRY(c, hipEventElapsedTime(ms, c->events[a], c->events[b]));
    return HCI_OK;
}

const char* hci_last_error(const hci_ctx* c) { return c ? c->err.c_str() : "null context"; }

int hci_device_sync(hci_ctx* c) {
    if (!c) return HCI_ERR_ARG;
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return HCI_OK;
}

int hci_set_integrals(hci_ctx* c, int n_orb, double e_core, const double* one_body, size_t n1,
                      const double* two_body, size_t n2) {
    if (!c) return HCI_ERR_ARG;
    if (n_orb < 1 || n_orb > 128 || !one_body || !two_body)
        return hci_fail(c, HCI_ERR_ARG, "hci_set_integrals: need 1 <= n_orb <= 128 and tables");
    const size_t N = (size_t)n_orb;
    const size_t need1 = hc2(N + 1, N + 1), need2 = hc4(N + 1, N + 1, N + 1, N + 1);
    if (n1 < need1 || n2 < need2)
        return hci_fail(c, HCI_ERR_ARG, "hci_set_integrals: tables shorter than the combine_2/"
                                        "combine_4 layout for n_orb");
    HIP_TRY(c, hipSetDevice(c->device));
    std::vector<double> h1(N * N), J3(N * N * N), D3(N * N * N);
    for (size_t p = 0; p < N; ++p)
        for (size_t q = 0; q < N; ++q) h1[p * N + q] = one_body[hc2(p + 1, q + 1)];
    for (size_t i = 0; i < N; ++i)
        for (size_t j = 0; j < N; ++j)
            for (size_t k = 0; k < N; ++k) {
                // Ham::direct(i,k,j,k) = (ij|kk); direct_plus_exchange(i,k,j,k) = (ij|kk) - (ik|kj)
                double ijkk = two_body[hc4(i + 1, j + 1, k + 1, k + 1)];
                double ikkj = two_body[hc4(i + 1, k + 1, k + 1, j + 1)];
                J3[(i * N + j) * N + k] = ijkk;
                D3[(i * N + j) * N + k] = ijkk - ikkj;
            }
    // singles whose element is zero for every occupation (on a lattice: every
    // hop that is not a bond); symmetric, as h, J3 and D3 are
    std::vector<uint8_t> sok(N * N, 0);
    for (size_t i = 0; i < N; ++i)
        for (size_t j = 0; j < N; ++j) {
            bool nz = h1[i * N + j] != 0.0;
            for (size_t k = 0; k < N && !nz; ++k)
                nz = J3[(i * N + j) * N + k] != 0.0 || D3[(i * N + j) * N + k] != 0.0;
            if (nz) sok[i * N + j] = sok[j * N + i] = 1;
        }
    int all = 1;
    for (size_t i = 0; i < N; ++i)
        for (size_t j = 0; j < N; ++j)
            if (i != j && !sok[i * N + j]) all = 0;
    hipFree(c->d_two);
    hipFree(c->d_h1);
    hipFree(c->d_J3);
    hipFree(c->d_D3);
    hipFree(c->d_sok);
    c->d_two = c->d_h1 = c->d_J3 = c->d_D3 = nullptr;
    c->d_sok = nullptr;
    HIP_TRY(c, hipMalloc(&c->d_two, need2 * sizeof(double)));
    HIP_TRY(c, hipMalloc(&c->d_h1, N * N * sizeof(double)));
    HIP_TRY(c, hipMalloc(&c->d_J3, N * N * N * sizeof(double)));
    HIP_TRY(c, hipMalloc(&c->d_D3, N * N * N * sizeof(double)));
    HIP_TRY(c, hipMemcpy(c->d_two, two_body, need2 * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->d_h1, h1.data(), N * N * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->d_J3, J3.data(), N * N * N * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->d_D3, D3.data(), N * N * N * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(c, hipMalloc(&c->d_sok, N * N));
    HIP_TRY(c, hipMemcpy(c->d_sok, sok.data(), N * N, hipMemcpyHostToDevice));
    c->sok_all = all;
    // double excitations only read (pr|qs) with p != r and q != s
    bool vanish = true;
    for (size_t i = 1; i <= N && vanish; ++i)
        for (size_t j = i + 1; j <= N && vanish; ++j)
            for (size_t k = 1; k <= N && vanish; ++k)
                for (size_t l = k + 1; l <= N; ++l)
                    if (two_body[hc4(i, j, k, l)] != 0.0) {
                        vanish = false;
                        break;
                    }
    c->n_orb = n_orb;
    c->e_core = e_core;
    c->n_two = need2;
    c->doubles_vanish = vanish;
    c->h_one.assign(one_body, one_body + need1);
    c->h_two.assign(two_body, two_body + need2);
    select_free(c);  // lists of the previous integrals
    return HCI_OK;
}

int hci_ham_diag(hci_ctx* c, const uint64_t* dets, int64_t n, double* diag, int mem) {
    if (!c) return HCI_ERR_ARG;
    if (!c->d_two) return hci_fail(c, HCI_ERR_STATE, "hci_ham_diag: integrals not set");
    if (n < 0 || (n > 0 && (!dets || !diag))) return hci_fail(c, HCI_ERR_ARG, "hci_ham_diag: bad arguments");
    if (n == 0) return HCI_OK;
    if (mem == HCI_MEM_HOST && !words_fit(dets, n, c->n_orb))
        return hci_fail(c, HCI_ERR_ARG, "hci_ham_diag: determinant bits beyond n_orb");
    HIP_TRY(c, hipSetDevice(c->device));
    const uint64_t* d_dets = dets;
    double* d_out = diag;
    uint64_t* tmp_d = nullptr;
    double* tmp_o = nullptr;
    if (mem == HCI_MEM_HOST) {
        HIP_TRY(c, hipMalloc(&tmp_d, (size_t)n * 32));
        HIP_TRY(c, hipMalloc(&tmp_o, (size_t)n * 8));
        HIP_TRY(c, hipMemcpyAsync(tmp_d, dets, (size_t)n * 32, hipMemcpyHostToDevice, c->stream));
        d_dets = tmp_d;
        d_out = tmp_o;
    }
    if (c->n_orb <= 64)
        hipLaunchKernelGGL(k_ham_diag<uint64_t>, grid_for(n, 256), 256, 0, c->stream, c->ints(), d_dets, n, d_out);
    else
        hipLaunchKernelGGL(k_ham_diag<u128>, grid_for(n, 256), 256, 0, c->stream, c->ints(), d_dets, n, d_out);
    HIP_TRY(c, hipGetLastError());
    if (mem == HCI_MEM_HOST) {
        HIP_TRY(c, hipMemcpyAsync(diag, tmp_o, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        hipFree(tmp_d);
        hipFree(tmp_o);
    }
    return HCI_OK;
}

int hci_elements(hci_ctx* c, const uint64_t* det1, const uint64_t* det2, int64_t m, double* value,
                 int8_t* degree, int mem) {
    if (!c) return HCI_ERR_ARG;
    if (!c->d_two) return hci_fail(c, HCI_ERR_STATE, "hci_elements: integrals not set");
    if (m < 0 || (m > 0 && (!det1 || !det2 || !value)))
        return hci_fail(c, HCI_ERR_ARG, "hci_elements: bad arguments");
    if (m == 0) return HCI_OK;
    if (mem == HCI_MEM_HOST && (!words_fit(det1, m, c->n_orb) || !words_fit(det2, m, c->n_orb)))
        return hci_fail(c, HCI_ERR_ARG, "hci_elements: determinant bits beyond n_orb");
    HIP_TRY(c, hipSetDevice(c->device));
    const uint64_t *a = det1, *b = det2;
    double* v = value;
    int8_t* dg = degree;
    uint64_t *ta = nullptr, *tb = nullptr;
    double* tv = nullptr;
    int8_t* td = nullptr;
    if (mem == HCI_MEM_HOST) {
        HIP_TRY(c, hipMalloc(&ta, (size_t)m * 32));
        HIP_TRY(c, hipMalloc(&tb, (size_t)m * 32));
        HIP_TRY(c, hipMalloc(&tv, (size_t)m * 8));
        if (degree) HIP_TRY(c, hipMalloc(&td, (size_t)m));
        HIP_TRY(c, hipMemcpyAsync(ta, det1, (size_t)m * 32, hipMemcpyHostToDevice, c->stream));
        HIP_TRY(c, hipMemcpyAsync(tb, det2, (size_t)m * 32, hipMemcpyHostToDevice, c->stream));
        a = ta;
        b = tb;
        v = tv;
        dg = td;
    }
    if (c->n_orb <= 64)
        hipLaunchKernelGGL(k_elements<uint64_t>, grid_for(m, 256), 256, 0, c->stream, c->ints(), a, b, m, v, dg);
    else
        hipLaunchKernelGGL(k_elements<u128>, grid_for(m, 256), 256, 0, c->stream, c->ints(), a, b, m, v, dg);
    HIP_TRY(c, hipGetLastError());
    if (mem == HCI_MEM_HOST) {
        HIP_TRY(c, hipMemcpyAsync(value, tv, (size_t)m * 8, hipMemcpyDeviceToHost, c->stream));
        if (degree) HIP_TRY(c, hipMemcpyAsync(degree, td, (size_t)m, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        hipFree(ta);
        hipFree(tb);
        hipFree(tv);
        hipFree(td);
    }
    return HCI_OK;
}

int hci_csr_info(const hci_csr* h, int64_t* n, int64_t* nnz_u, int* value_prec) {
    if (!h) return HCI_ERR_ARG;
    if (n) *n = h->n;
    if (nnz_u) *nnz_u = h->nnz_u;
    if (value_prec) *value_prec = h->prec;
    return HCI_OK;
}

}  // extern "C"

namespace {
__global__ void k_gather_perm(const int32_t* perm, const int32_t* idx, int64_t n, int32_t* out) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) out[k] = perm[idx[k]];
}

__global__ void k_widen(const float* hi, const float* lo, double* out, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (double)hi[i] + (double)lo[i];
}

// internal arrays -> host (values widened to double)
int fetch_internal(const hci_csr* h, std::vector<int64_t>& rp, std::vector<int32_t>& col,
                   std::vector<double>& val, std::vector<int32_t>& perm) {
    hci_ctx* c = h->ctx;
    rp.resize((size_t)h->n + 1);
    col.resize((size_t)h->nnz_alloc);
    val.resize((size_t)h->nnz_alloc);
    perm.resize((size_t)h->n);
    if (h->d_row_ptr) HIP_TRY(c, hipMemcpy(rp.data(), h->d_row_ptr, rp.size() * 8, hipMemcpyDeviceToHost));
    HIP_TRY(c, hipMemcpy(perm.data(), h->d_perm, perm.size() * 4, hipMemcpyDeviceToHost));
    if (h->nnz_alloc == 0) return HCI_OK;
    if (h->d_idx16) {
        std::vector<uint16_t> u(col.size());
        HIP_TRY(c, hipMemcpy(u.data(), h->d_idx16, u.size() * 2, hipMemcpyDeviceToHost));
        std::copy(u.begin(), u.end(), col.begin());
    } else {
        HIP_TRY(c, hipMemcpy(col.data(), h->d_col, col.size() * 4, hipMemcpyDeviceToHost));
    }
    if (h->prec == HCI_PREC_F64) {
        HIP_TRY(c, hipMemcpy(val.data(), h->d_val, val.size() * 8, hipMemcpyDeviceToHost));
    } else {
        double* tmp = nullptr;
        HIP_TRY(c, hipMalloc(&tmp, val.size() * 8));
        hipLaunchKernelGGL(k_widen, grid_for(h->nnz_alloc, 256), 256, 0, c->stream, h->d_hi, h->d_lo, tmp,
                           h->nnz_alloc);
        hipError_t e = hipStreamSynchronize(c->stream);
        if (e == hipSuccess) e = hipMemcpy(val.data(), tmp, val.size() * 8, hipMemcpyDeviceToHost);
        hipFree(tmp);
        HIP_TRY(c, e);
    }
    return HCI_OK;
}
// FULL storage: internal row p holds both triangles of caller row perm[p],
// so the caller's upper row perm[p] is exactly row p's entries with
// perm[col] > perm[p]: rows are counted, placed and sorted on host threads,
// each thread a contiguous range of internal rows (no two write one row)
int download_full(const hci_csr* h, const std::vector<int64_t>& rp, const std::vector<int32_t>& ic,
                  const std::vector<double>& iv, const std::vector<int32_t>& perm, int64_t* row_ptr, int32_t* col,
                  double* val, double* diag) {
    hci_ctx* c = h->ctx;
    const int64_t n = h->n;
    const int nt = (int)std::max<unsigned>(1, std::min<unsigned>(32, std::thread::hardware_concurrency()));
    auto parallel = [&](auto f) {
        std::vector<std::thread> th;
        for (int t = 0; t < nt; ++t) th.emplace_back([&, t]() { f(n * t / nt, n * (t + 1) / nt); });
        for (auto& x : th) x.join();
    };
    std::vector<int64_t> cnt((size_t)n + 1, 0);
    parallel([&](int64_t p0, int64_t p1) {
        for (int64_t p = p0; p < p1; ++p) {
            const int32_t i = perm[p];
            int64_t k = 0;
            for (int64_t e = rp[p]; e < rp[p + 1]; ++e) k += perm[ic[e]] > i;
            cnt[(size_t)i + 1] = k;
        }
    });
    for (int64_t i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
    if (row_ptr) std::copy(cnt.begin(), cnt.end(), row_ptr);
    parallel([&](int64_t p0, int64_t p1) {
        std::vector<std::pair<int32_t, double>> row;
        for (int64_t p = p0; p < p1; ++p) {
            const int32_t i = perm[p];
            row.clear();
            for (int64_t e = rp[p]; e < rp[p + 1]; ++e) {
                const int32_t j = perm[ic[e]];
                if (j > i) row.emplace_back(j, iv[e]);
            }
            std::sort(row.begin(), row.end(),
                      [](const std::pair<int32_t, double>& a, const std::pair<int32_t, double>& b) {
                          return a.first < b.first;
                      });
            int64_t o = cnt[i];
            for (const auto& q : row) {
                if (col) col[o] = q.first;
                if (val) val[o] = q.second;
                ++o;
            }
        }
    });
    if (diag && n) {
        std::vector<double> d((size_t)n);
        HIP_TRY(c, hipMemcpy(d.data(), h->d_diag, (size_t)n * 8, hipMemcpyDeviceToHost));
        for (int64_t p = 0; p < n; ++p) diag[perm[p]] = d[p];
    }
    return HCI_OK;
}
}  // namespace

extern "C" {

// Caller-order upper triangle: row i holds (j, H_ij) for j > i, ascending j
// (the reference's SparseMatUpperTri after sort_remove_duplicates).
int hci_csr_download(const hci_csr* h, int64_t* row_ptr, int32_t* col, double* val, double* diag) {
    if (!h) return HCI_ERR_ARG;
    hci_ctx* c = h->ctx;
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    std::vector<int64_t> rp;
    std::vector<int32_t> ic, perm, bpos;
    std::vector<double> iv;
    std::vector<Window> wins;
    std::vector<TileTask> tasks;
    int rc = fetch_internal(h, rp, ic, iv, perm);
    if (rc) return rc;
    const int64_t n = h->n;
    if (h->storage == HCI_STORE_TILED) {
        wins.resize((size_t)h->n_win);
        tasks.resize((size_t)(h->n_task_a + h->n_task_b));
        bpos.resize((size_t)n);
        if (h->n_win) HIP_TRY(c, hipMemcpy(wins.data(), h->d_win, wins.size() * sizeof(Window), hipMemcpyDeviceToHost));
        if (h->n_task_a)
            HIP_TRY(c, hipMemcpy(tasks.data(), h->d_task_a, (size_t)h->n_task_a * sizeof(TileTask), hipMemcpyDeviceToHost));
        if (h->n_task_b)
            HIP_TRY(c, hipMemcpy(tasks.data() + h->n_task_a, h->d_task_b, (size_t)h->n_task_b * sizeof(TileTask),
                                 hipMemcpyDeviceToHost));
        HIP_TRY(c, hipMemcpy(bpos.data(), h->d_bpos, (size_t)n * 4, hipMemcpyDeviceToHost));
    }
    if (h->storage == HCI_STORE_FULL) return download_full(h, rp, ic, iv, perm, row_ptr, col, val, diag);
    std::vector<int64_t> cnt((size_t)n + 1, 0);
    auto each = [&](auto f) {
        if (h->storage == HCI_STORE_TILED) {
            for (size_t t = 0; t < tasks.size(); ++t) {
                const bool beta_major = (int64_t)t >= h->n_task_a;
                for (int32_t k = tasks[t].w_begin; k < tasks[t].w_end; ++k) {
                    const Window& b = wins[(size_t)k];
                    // the stored half of a mirror pair holds both triangles' pairs
                    const bool pair = b.sp() >= 0;
                    for (int64_t j = 0; j < b.cnt(); ++j) {
                        const int64_t e = b.o() + win_slot(j, b.cnt());
                        const uint32_t u = (uint32_t)ic[(size_t)e];
                        int32_t p = tasks[t].r0 + (int32_t)(u >> 8), q = b.c0 + (int32_t)(u & 0xFFu);
                        if (beta_major) {
                            p = bpos[p];
                            q = bpos[q];
                        }
                        const int32_t i = perm[p], j2 = perm[q];
                        if (iv[(size_t)e] == 0.0) continue;
                        if (pair)
                            f(std::min(i, j2), std::max(i, j2), iv[(size_t)e]);
                        else if (i < j2)
                            f(i, j2, iv[(size_t)e]);
                    }
                }
            }
            return;
        }
        for (int64_t p = 0; p < n; ++p)
            for (int64_t e = rp[p]; e < rp[p + 1]; ++e) {
                int32_t i = perm[p], j = perm[ic[e]];
                if (h->storage == HCI_STORE_FULL) {
                    if (i < j) f(i, j, iv[e]);
                } else {
                    f(std::min(i, j), std::max(i, j), iv[e]);
                }
            }
    };
    each([&](int32_t i, int32_t, double) { ++cnt[(size_t)i + 1]; });
    for (int64_t i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
    std::vector<std::pair<int32_t, double>> ent((size_t)cnt[n]);
    std::vector<int64_t> cur(cnt.begin(), cnt.end() - 1);
    each([&](int32_t i, int32_t j, double v) { ent[(size_t)cur[i]++] = {j, v}; });
    for (int64_t i = 0; i < n; ++i)
        std::sort(ent.begin() + cnt[i], ent.begin() + cnt[i + 1],
                  [](const std::pair<int32_t, double>& a, const std::pair<int32_t, double>& b) {
                      return a.first < b.first;
                  });
    if (row_ptr) std::copy(cnt.begin(), cnt.end(), row_ptr);
    for (size_t e = 0; e < ent.size(); ++e) {
        if (col) col[e] = ent[e].first;
        if (val) val[e] = ent[e].second;
    }
    if (diag && n) {
        std::vector<double> d((size_t)n);
        HIP_TRY(c, hipMemcpy(d.data(), h->d_diag, (size_t)n * 8, hipMemcpyDeviceToHost));
        for (int64_t p = 0; p < n; ++p) diag[perm[p]] = d[p];
    }
    return HCI_OK;
}

int hci_csr_row(const hci_csr* h, int64_t row, int32_t* cols, double* vals, int64_t cap, int64_t* count) {
    if (!h || !count) return HCI_ERR_ARG;
    hci_ctx* c = h->ctx;
    if (h->storage != HCI_STORE_FULL) return hci_fail(c, HCI_ERR_ARG, "hci_csr_row: needs HCI_STORE_FULL");
    if (row < 0 || row >= h->n) return hci_fail(c, HCI_ERR_ARG, "hci_csr_row: row out of range");
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    int32_t p = 0;
    HIP_TRY(c, hipMemcpy(&p, h->d_pos + row, 4, hipMemcpyDeviceToHost));
    int64_t rr[2];
    HIP_TRY(c, hipMemcpy(rr, h->d_row_ptr + p, 16, hipMemcpyDeviceToHost));
    const int64_t len = rr[1] - rr[0];
    std::vector<int32_t> q((size_t)len);
    std::vector<double> v((size_t)len);
    if (len) {
        HIP_TRY(c, hipMemcpy(q.data(), h->d_col + rr[0], (size_t)len * 4, hipMemcpyDeviceToHost));
        if (h->prec == HCI_PREC_F64) {
            HIP_TRY(c, hipMemcpy(v.data(), h->d_val + rr[0], (size_t)len * 8, hipMemcpyDeviceToHost));
        } else {
            std::vector<float> a((size_t)len), b((size_t)len);
            HIP_TRY(c, hipMemcpy(a.data(), h->d_hi + rr[0], (size_t)len * 4, hipMemcpyDeviceToHost));
            HIP_TRY(c, hipMemcpy(b.data(), h->d_lo + rr[0], (size_t)len * 4, hipMemcpyDeviceToHost));
            for (int64_t k = 0; k < len; ++k) v[k] = (double)a[k] + (double)b[k];
        }
    }
    std::vector<int32_t> jj((size_t)len);
    if (len) {
        int32_t* tmp = nullptr;
        HIP_TRY(c, hipMalloc(&tmp, (size_t)len * 4));
        hipLaunchKernelGGL(k_gather_perm, grid_for(len, 256), 256, 0, c->stream, h->d_perm, h->d_col + rr[0], len, tmp);
        hipError_t e = hipStreamSynchronize(c->stream);
        if (e == hipSuccess) e = hipMemcpy(jj.data(), tmp, (size_t)len * 4, hipMemcpyDeviceToHost);
        hipFree(tmp);
        HIP_TRY(c, e);
    }
    std::vector<std::pair<int32_t, double>> ent;
    for (int64_t k = 0; k < len; ++k)
        if (jj[k] > row) ent.push_back({jj[k], v[k]});
    std::sort(ent.begin(), ent.end(), [](const std::pair<int32_t, double>& a, const std::pair<int32_t, double>& b) {
        return a.first < b.first;
    });
    *count = (int64_t)ent.size();
    for (int64_t k = 0; k < (int64_t)ent.size() && k < cap; ++k) {
        if (cols) cols[k] = ent[k].first;
        if (vals) vals[k] = ent[k].second;
    }
    return HCI_OK;
}

int hci_csr_recycle(hci_csr* h) {
    if (!h) return HCI_ERR_ARG;
    hci_ctx* c = h->ctx;
    if (h->storage == HCI_STORE_TILED && h->entry_cap > 0) {
        HIP_TRY(c, hipSetDevice(c->device));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        recycled_flush(c);
        const size_t cap = h->entry_cap;
        c->recycled.push_back({h->d_idx16, cap * 2});
        c->recycled.push_back({h->d_val ? (void*)h->d_val : (void*)h->d_hi, cap * (h->d_h16 ? 10 : 8)});
        h->d_idx16 = nullptr;
        h->d_val = nullptr;
        h->d_hi = nullptr;
        h->d_lo = nullptr;
        h->d_h16 = nullptr;
    }
    hci_csr_free(h);
    return HCI_OK;
}

void hci_csr_free(hci_csr* h) {
    if (!h) return;
    hipSetDevice(h->ctx->device);
    hipStreamSynchronize(h->ctx->stream);
    hipFree(h->d_row_ptr);
    hipFree(h->d_col);
    hipFree(h->d_val);
    hipFree(h->d_hi);
    if (h->storage != HCI_STORE_TILED) hipFree(h->d_lo);  // TILED: lo lives in hi's block
    hipFree(h->d_diag);
    hipFree(h->d_chunk_row);
    hipFree(h->d_item_row);
    hipFree(h->d_rownorm);
    hipFree(h->d_pair_part);
    hipFree(h->d_pair_slot);
    hipFree(h->d_pair_seg);
    hipFree(h->d_perm);
    hipFree(h->d_pos);
    hipFree(h->d_words);
    hipFree(h->d_idx16);
    hipFree(h->d_win);
    hipFree(h->d_task_a);
    hipFree(h->d_task_b);
    hipFree(h->d_bpos);
    hipFree(h->d_binv);
    hipFree(h->d_xb);
    hipFree(h->d_yb);
    hipFree(h->d_x32);
    hipFree(h->d_xb32);
    delete h;
}

}  // extern "C"
