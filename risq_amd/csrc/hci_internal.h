// Internal host-side structures of the HCI library.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>
#include <vector>

#include "../../include/hci_gpu.h"
#include "hci_device.h"

struct MlpState;     // learned preconditioner (mlp.hip)
struct SelectLists;  // heat-bath excitation lists (select.hip)

// TILED storage.  The build enumerates sub-blocks: the pairs between nr <= 256
// rows starting at r0 and nc <= 4096 columns starting at c0 (both in the
// part's order: alpha-major for kinds 0/1/3, beta-major for kind 2).
//   kind 0: rows and columns in the same alpha group (beta singles/doubles)
//   kind 1: columns in a single-connected alpha group (alpha singles with equal
//           beta, opposite-spin doubles)
//   kind 2: rows and columns in the same beta group, beta-major (alpha doubles)
//   kind 3: as kind 1, but the build finds the partners through sorted
//           "beta minus one electron" keys instead of scanning the group
// Storage is by window: the entries of one sub-block whose columns fall in one
// 256-wide slice, each a 16-bit (local_row << 8 | local_col) index plus its
// value.  A window of cnt entries starts at a multiple of 8 and occupies
// ceil(cnt / 8) groups of 8 slots.  Each full chunk of 512 entries is stored
// lane-major: entry 64 q + l of the chunk sits in group l, slot q, so lane l of
// a wave reads its 8 entries as one 16-byte index vector and in each of the 8
// steps the wave handles 64 consecutive entries.  The last m < 512 entries use
// R = ceil(m / 8) groups: entry j sits in group j mod R, slot j div R.
constexpr int kTileRows = 256;
constexpr int kTileCols = 4096;  // build sub-block width
constexpr int kWinCols = 256;
constexpr int kWinPerSb = kTileCols / kWinCols;
constexpr int kWinAlign = 8;
constexpr int kTileWaves = 4;
constexpr int kTileThreads = 64 * kTileWaves;
struct SubBlock {
    int32_t r0, nr, c0, nc;
    int32_t g_row, g_col;
    int32_t kind, win_begin;  // first of its ceil(nc / 256) windows
    int32_t rb;               // r0 - the first row of its tile (tiles pack small groups)
};
struct Window {
    int64_t off_sp;   // first slot (multiple of kWinAlign) | (pair slot + 1) << 40
    int32_t c0;       // first column, in the part's order
    uint32_t cnt_nc;  // entries | columns << 20
    __host__ __device__ int cnt() const { return (int)(cnt_nc & 0xFFFFFu); }
    __host__ __device__ int nc() const { return (int)(cnt_nc >> 20); }
    __host__ __device__ int64_t o() const { return off_sp & ((1ll << 40) - 1); }
    // the window's slot among the stored halves of mirror pairs, -1 if none
    __host__ __device__ int32_t sp() const { return (int32_t)(off_sp >> 40) - 1; }
};
// Mirror pairs.  A dense window whose rows are one tile's whole row range and
// whose columns are another tile's whole row range has a mirror (the same
// entries transposed) in that other tile.  Only the half above the diagonal
// block is stored; its product also forms the transposed sums (one per
// column), which go to a partial-sum buffer, slot sp x kWinCols, and are added
// into the receiving tile's rows by a reduction after the product.  Windows
// with at least kPairMinEntries entries are paired (builds without FP16
// copies).  The stored half's entries are sorted by diagonal ((column - row)
// mod 256, stable) before they move to win_slot's slots: the 64 consecutive
// entries of one step of the product then lie on one or a few diagonals, on
// distinct rows and distinct columns, so neither the row nor the column
// accumulator sees many lanes on one address.
constexpr int kPairMinEntries = 2048;
// receiving tile of a run of slots: rows [r0, r0 + nr) of part `part`
// receive the slots' partial sums, slots s_begin .. s_end of the sorted list
struct PairSeg {
    int32_t part, r0, nr, s_begin, s_end;
};
constexpr int kWinChunk = 64 * kWinAlign;
__host__ __device__ inline int64_t win_slot(int64_t j, int64_t cnt) {
    const int64_t full = cnt / kWinChunk * kWinChunk;
    if (j < full) return (j & ~(int64_t)(kWinChunk - 1)) + ((j & 63) << 3) + ((j >> 6) & 7);
    const int64_t t = j - full, R = (cnt - full + kWinAlign - 1) / kWinAlign;
    return full + (t % R) * kWinAlign + t / R;
}
// one workgroup: rows [r0, r0+nr) and windows [w_begin, w_end); windows of
// more than kSmallWin entries come first and are taken one per wavefront,
// those from w_split on are taken one per lane
constexpr int kSmallWin = 8;
// ... or kSmallWinDense entries in matrices of at least kDenseRowSlots
// stored slots per row
constexpr int kSmallWinDense = 32;
constexpr int64_t kDenseRowSlots = 64;
struct TileTask {
    int32_t r0, nr, w_begin, w_end, w_split;
};
// Which single-storage rule (HCI_SINGLES_*) a pair follows: the rule of the
// first segment whose end exceeds the pair's larger caller index.  A fresh
// build has one segment; an incremental build appends (n, its rule) to the
// previous matrix's segments, as the reference's builds add the pairs with a
// new index under the rule of that build (src/var/davidson.rs:55-61).
constexpr int kMaxRuleSegs = 8;
struct SingleRules {
    int32_t n;
    int32_t end[kMaxRuleSegs];
    int32_t rule[kMaxRuleSegs];
};

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
    uint8_t* d_sok = nullptr;  // DevInts::sok
    int sok_all = 0;
    size_t n_two = 0;
    bool doubles_vanish = false;  // every (pq|rs) with p != q and r != s is zero (Hubbard)
    std::vector<double> h_one, h_two;  // host copies of the packed tables (selection lists)
    MlpState* mlp = nullptr;
    SelectLists* sel = nullptr;
    hipEvent_t events[16] = {};
    // entry arrays of a recycled matrix, offered to the next TILED build
    // (device memory handed back and allocated again is cleared by the GPU
    // kernel driver before reuse: about 4 s for 200 GB)
    struct Recycled {
        void* p;
        size_t bytes;
    };
    std::vector<Recycled> recycled;
    // the builds' work buffers of at most kScratchCacheBuf bytes, kept for the
    // next build (up to kScratchCacheMax bytes in all): a build allocates and
    // frees ~140 of them, and every hipFree waits for the device
    std::vector<Recycled> scratch_cache;
    size_t scratch_cached = 0;
    // Davidson and sigma work vectors, kept between calls (davidson.hip):
    // one device block and one pinned host block, grown on demand
    void* work_dev = nullptr;
    size_t work_dev_bytes = 0;
    void* work_host = nullptr;
    size_t work_host_bytes = 0;
    // device address of the fixed scratch of hci_sigma_noalloc (spmv.hip)
    double* scratch = nullptr;

    DevInts ints() const {
        DevInts t;
        t.n = n_orb;
        t.nuc = e_core;
        t.two = d_two;
        t.h1 = d_h1;
        t.J3 = d_J3;
        t.D3 = d_D3;
        t.sok = d_sok;
        t.sok_all = sok_all;
        return t;
    }
};

// Device-resident Hamiltonian.  Rows/columns are numbered in the internal
// alpha-major order p (determinants sorted by alpha string, then beta string);
// perm[p] is the caller's index of internal row p and pos is its inverse.
// storage FULL holds both triangles (every connected pair twice), UPPER holds
// p < q only.  nnz_u counts the reference's stored upper-triangle pairs.
struct hci_csr {
    hci_ctx* ctx = nullptr;
    int64_t n = 0;
    int64_t nnz_u = 0;             // reference count (upper pairs)
    int n_up = -1, n_dn = -1;      // electrons per spin, read from d_words on first use
    int64_t nnz_s = 0;             // stored entries (2*nnz_u for FULL)
    int64_t nnz_alloc = 0;         // length of the entry arrays (nnz_s plus TILED alignment gaps)
    int prec = HCI_PREC_F64;
    int storage = HCI_STORE_FULL;
    int32_t* d_perm = nullptr;     // internal -> caller index
    int32_t* d_pos = nullptr;      // caller index -> internal
    int64_t* d_row_ptr = nullptr;  // n+1
    int32_t* d_col = nullptr;      // nnz_s, internal column index
    double* d_val = nullptr;       // nnz_s (HCI_PREC_F64)
    float* d_hi = nullptr;         // nnz_s (HCI_PREC_F32)
    float* d_lo = nullptr;         // nnz_s (HCI_PREC_F32)
    uint16_t* d_h16 = nullptr;     // FP16 copies (HCI_PREC_F16, TILED; inside the value block)
    double* d_diag = nullptr;      // n, internal order
    uint64_t* d_words = nullptr;   // 4n: determinant words in internal order
    // nnz-balanced SpMV work split: row containing the first entry of every chunk
    // (UPPER storage)
    int32_t* d_chunk_row = nullptr;
    int64_t n_chunks = 0;
    // row-aligned work items (FULL storage): item k covers rows
    // [item_row[k], item_row[k+1]); a long row is an item of its own
    int32_t* d_item_row = nullptr;
    int64_t n_items = 0;
    // per-row absolute sums (learned-preconditioner feature), lazily built
    double* d_rownorm = nullptr;
    // TILED storage
    size_t entry_cap = 0;          // allocated length of d_idx16 / d_val / d_hi / d_lo
    uint16_t* d_idx16 = nullptr;   // nnz_alloc packed (local_row << 8 | local_col)
    Window* d_win = nullptr;       // non-empty windows, tile by tile
    int64_t n_win = 0;
    TileTask* d_task_a = nullptr;  // alpha-major tiles (kinds 0, 1, 3)
    int64_t n_task_a = 0;
    TileTask* d_task_b = nullptr;  // beta-major tiles (kind 2)
    int64_t n_task_b = 0;
    int32_t* d_bpos = nullptr;     // beta-major position -> internal position
    int32_t* d_binv = nullptr;     // internal position -> beta-major position
    double* d_xb = nullptr;        // x in beta-major order
    double* d_yb = nullptr;        // the beta-major part's row sums (beta-major order)
    float* d_x32 = nullptr;        // FP32 copies of x for FP16-value products (HCI_PREC_F16)
    float* d_xb32 = nullptr;
    // mirror pairs (TILED): partial sums (n_pair x kWinCols doubles), the
    // slots in receiving order and the receiving segments (the alpha-major
    // part's first: n_pair_seg0 of them)
    double* d_pair_part = nullptr;
    int32_t* d_pair_slot = nullptr;
    PairSeg* d_pair_seg = nullptr;
    int64_t n_pair = 0, n_pair_seg = 0, n_pair_seg0 = 0;
    // single-storage rule segments of the pairs (see SingleRules)
    std::vector<std::pair<int64_t, int>> rule_segs;
};

// recycled entry arrays (ctx.hip): recycled_take hands out a kept buffer of at
// least `bytes` or allocates; recycled_flush frees what is left
hipError_t recycled_take(hci_ctx* c, void** p, size_t bytes);
void recycled_flush(hci_ctx* c);
// build work buffers (ctx.hip): scratch_take hands out a kept buffer of
// bytes .. 5/4 bytes or allocates (on the context's stream, whose order makes
// a kept buffer free for reuse); scratch_give keeps or frees one;
// scratch_flush frees them all
constexpr size_t kScratchCacheBuf = (size_t)128 << 20;
constexpr size_t kScratchCacheMax = (size_t)1 << 30;
hipError_t scratch_take(hci_ctx* c, void** p, size_t bytes);
void scratch_give(hci_ctx* c, void* p, size_t bytes);
void scratch_flush(hci_ctx* c);
// the context's Davidson / sigma work block (davidson.hip): work_reserve
// returns a device block of at least dev_bytes and a pinned host block of at
// least host_bytes, reallocating only when they grow; work_release frees both
// (builds call it before they allocate when an allocation fails)
int work_reserve(hci_ctx* c, size_t dev_bytes, size_t host_bytes);
void work_release(hci_ctx* c);

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
constexpr int kShortRows = 1024;  // FULL SpMV: rows are packed into items of at most this many entries
int spmv_prepare(hci_csr* h);      // builds d_chunk_row
// y = H x on device vectors, on the context stream; mode = HCI_SIGMA_*
// y = H x on device vectors in internal order, on the context stream
int spmv_device(const hci_csr* h, const double* x, double* y, int mode);
int csr_rownorm(hci_csr* h);       // fills d_rownorm[p] = sum_q |H_pq| (off-diagonal)
int sigma_scratch_init(hci_ctx* c);  // device address of hci_sigma_noalloc's scratch
// caller order <-> internal order
void permute_to_internal(const hci_csr* h, const double* x_caller, double* x_int);
void permute_to_caller(const hci_csr* h, const double* x_int, double* x_caller);

// heat-bath selection (select.hip)
void select_free(hci_ctx* c);

// launch helpers
inline unsigned grid_for(int64_t n, int block) {
    int64_t g = (n + block - 1) / block;
    return (unsigned)(g < 1 ? 1 : g);
}
