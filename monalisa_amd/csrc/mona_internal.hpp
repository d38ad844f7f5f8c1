// Internal helpers shared by the HIP translation units of libmona.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mona.h"

namespace mona {

// ---------------------------------------------------------------- errors
void set_error(const std::string& msg);
const char* last_error();

struct Error {
    int code;
    std::string msg;
};

#define MONA_HIP(expr)                                                              \
    do {                                                                            \
        hipError_t _e = (expr);                                                     \
        if (_e != hipSuccess)                                                       \
            throw ::mona::Error{MONA_ERR_HIP, std::string(#expr) + ": " +          \
                                                  hipGetErrorString(_e)};           \
    } while (0)

#define MONA_CHECK(cond, code, msg)                                                 \
    do {                                                                            \
        if (!(cond)) throw ::mona::Error{(code), (msg)};                            \
    } while (0)

// Runs f(); converts exceptions to a status code and records the message.
template <class F>
int guarded(F&& f) {
    try {
        f();
        return MONA_OK;
    } catch (const Error& e) {
        set_error(e.msg);
        return e.code;
    } catch (const std::bad_alloc&) {
        set_error("host allocation failed");
        return MONA_ERR_OOM;
    } catch (const std::exception& e) {
        set_error(e.what());
        return MONA_ERR_INTERNAL;
    }
}

// The reference's void entry points have no status; a failure there prints
// the reason and aborts the process rather than returning garbage.
void fatal(const char* where, int code);

// ---------------------------------------------------------------- memory
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    explicit DevBuf(size_t n) { alloc(n); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t n) {
        release();
        if (n == 0) return;
        hipError_t e = hipMalloc(&p, n);
        if (e != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            throw Error{MONA_ERR_OOM, "hipMalloc of " + std::to_string(n) + " bytes failed"};
        }
        bytes = n;
    }
    void ensure(size_t n) { if (n > bytes) alloc(n); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

// Per-process context for the synchronous host-pointer entry points:
// one stream and grow-only staging buffers, guarded by a mutex.
struct HostCallCtx {
    std::mutex mu;
    hipStream_t stream = nullptr;
    DevBuf slot[12];
    void init();
};
HostCallCtx& host_ctx();

// ---------------------------------------------------------------- launch
inline unsigned blocks_for(int64_t n, int threads) {
    int64_t b = (n + threads - 1) / threads;
    if (b < 1) b = 1;
    if (b > 0x7fffffff) b = 0x7fffffff;
    return (unsigned)b;
}

// Device scans (scan.hip).  All run on `s`, need no host synchronisation.
void scan_i32_to_i32_inclusive(const int* in, int* out, int64_t n, hipStream_t s);
void scan_i32_to_i64_exclusive(const int* in, int64_t* out, int64_t n, hipStream_t s);  // out has n+1
void scan_i32_to_i64_inclusive(const int* in, int64_t* out, int64_t n, hipStream_t s);
void fill_f32(float* p, float v, int64_t n, hipStream_t s);
void fill_i32(int* p, int v, int64_t n, hipStream_t s);

// ---------------------------------------------------------------- launch order
// Workgroup b runs on XCD b % 8 (round-robin dispatch over the 8 XCDs, each
// with its own L2).  With G > 1 the logical block of workgroup b is chosen so
// that each XCD receives runs of G consecutive logical blocks: neighbouring
// blocks (gridding rows that gather the same points, FFT tiles on the same
// rows) then share one L2, while the work still spreads over every XCD.
// Blocks past the last full run of 8 G keep their index.
__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t nblocks, int G) {
    if (G <= 1) return b;
    const int64_t span = 8LL * G;
    const int64_t full = nblocks / span * span;
    if (b >= full) return b;
    const int64_t xcd = b & 7, k = b >> 3;
    return ((k / G) * 8 + xcd) * G + (k % G);
}

// ---------------------------------------------------------------- storage types
// A complex value stored as two bfloat16 (real in the low half), the
// reduced-precision storage of the mixed-precision solver.  Arithmetic is
// always float: ld_c widens exactly, st_c rounds to nearest even.
struct bf16x2 {
    uint32_t v;
};
__device__ __forceinline__ float2 ld_c(const float2* p) { return *p; }
__device__ __forceinline__ float2 ld_c(const bf16x2* p) {
    const uint32_t u = p->v;
    return make_float2(__uint_as_float(u << 16), __uint_as_float(u & 0xffff0000u));
}
__device__ __forceinline__ uint32_t f32_to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (u >> 16) | 0x40u;  // NaN stays NaN
    u += 0x7fffu + ((u >> 16) & 1u);
    return u >> 16;
}
__device__ __forceinline__ void st_c(float2* p, float2 v) { *p = v; }
__device__ __forceinline__ void st_c(bf16x2* p, float2 v) { p->v = f32_to_bf16(v.x) | (f32_to_bf16(v.y) << 16); }

// ---------------------------------------------------------------- sparse
// One stored entry, column and value side by side (one 8-byte load).
struct SpEntry {
    int col;
    float val;
};

// Device-resident, decoded form of a cpp_prepared one-block bmSparseMat.
// The trajectory a gridding matrix was built from (bmTraj2SparseMat
// arguments), kept by the library's own builders so that operator plans can
// grid without the stored entries.
struct TrajInfo {
    int imDim = 0, nPt = 0, ktype = 0, nWin = 0, nParam = 0;
    int N[3] = {1, 1, 1};
    double dK[3] = {1, 1, 1};
    double kp[3] = {0, 0, 0};
    std::vector<double> t;  // [nPt][imDim], physical k units
};

// Per-point data of the matrix-free gridding of a trajectory, in the
// builder's float64 semantics (bmTraj2SparseMat.m:58-230): the base cell
// (0-based, wrapped, nearest grid point for odd nWin), the offset of the
// point from it (t - floor(t + 1/2), in [-1/2, 1/2)), and 1 / sum of the
// point's nWin^imDim kernel weights taken without the kernel's constant
// factor -- I0(alpha sqrt(1 - d^2/tau^2)) for Kaiser, exp(-d^2/(2 sigma^2))
// for Gauss -- (its Gu row normaliser for those weights; 0 for a deleted
// point).  Odd nWin only.
struct TrajPoints {
    std::vector<int> base;      // [nPt][3]
    std::vector<double> rest;   // [nPt][3]
    std::vector<double> inv_sum;
};
void traj_points(const TrajInfo& T, TrajPoints* out);
std::shared_ptr<const TrajInfo> make_traj_info(int imDim, int nPt, const double* t, const int* N_u,
                                               const double* dK_u, int kernelType, int nWin,
                                               const double* kp, int nParam);

// One segment of the range form: entries [eb, eb + len) of one long row,
// summed into partial number part.
struct SpSeg {
    int64_t eb;
    int len;
    int part;
};

struct SparseDev {
    int r_size = 0, l_size = 0, l_nJump = 0;
    int64_t nnz = 0;
    bool squeezed = false;
    bool covers_all = false;  // every output row is a stored row (the product writes all of w)
    int max_row = 0;          // longest stored row
    DevBuf col;               // int32[nnz]   absolute 0-based column
    DevBuf val;               // float[nnz]
    DevBuf rowptr;            // int64[l_nJump+1]
    DevBuf rowidx;            // int32[l_nJump] output row (only if squeezed)
    DevBuf cv;                // SpEntry[nnz] (column, value) interleaved (operator matrices)
    // rows longer than kLongRow entries, split into chunks for the operator
    // products (fixed-order partial sums; see spmv_c2_rowmajor)
    int n_long = 0;
    int64_t n_chunk = 0;
    DevBuf long_row;          // int32[n_long]   stored-row index
    DevBuf long_first;        // int32[n_long+1] first chunk of each long row
    DevBuf chunk_begin;       // int64[2*n_chunk] [begin, end) entry offsets of each chunk
    mutable DevBuf partial;   // float2[n_chunk * n_vec] scratch
    // Staged form of the short rows (operator products over n_vec ==
    // group_vec vectors): runs of consecutive short rows whose entries
    // reference at most a group's capacity of distinct columns.  A group's
    // columns (ascending) are loaded into LDS once and its rows are summed
    // from there; slot[k] is entry k's position in its group's column list.
    int n_group = 0, group_vec = 0;
    int group_cap = 0, group_ecap = 0, group_rcap = 0;  // most columns, entries, rows of a group
    DevBuf group;             // int4[n_group]: first row, end row, first column in group_col, column count
    DevBuf group_e;           // int64[2 n_group]: first and end entry of each group
    DevBuf group_col;         // int32[n_group_col]: group g's columns at g * group_cap (padded with its last)
    int64_t n_group_col = 0;
    DevBuf slot;              // uint16[nnz] (entries of long rows unused)
    // Range form of the long rows (built with the groups; the operator
    // products then use it instead of the chunks): the columns the long rows
    // reference are cut into ranges of at most group_cap consecutive
    // columns; within a range every long row that has entries there owns one
    // segment (its entries are ascending in column, so a segment is a
    // contiguous run of entries).  A block stages a range's vectors (one
    // contiguous run of v) and sums each segment into its partial; each
    // long row then adds its partials in range order (spmv_c2_combine with
    // range_first in place of long_first).
    int n_range = 0, range_vec = 0, range_cap = 0;  // ranges for products over range_vec vectors, columns per range
    int64_t n_seg = 0, n_rpart = 0;
    DevBuf range;             // int4[n_range]: first column, column count, first segment, segment count
    DevBuf seg;               // SpSeg[n_seg], grouped by range
    DevBuf range_first;       // int32[n_long+1]: first partial of each long row
    // host copy of the reference encoding (kept for export; decoded onto the
    // device on first use)
    std::vector<int> h_r_jump, h_r_nJump, h_l_jump;
    std::vector<float> h_m_val;
    bool on_device = false;
    std::mutex mu;
    std::shared_ptr<const TrajInfo> traj;  // set by the trajectory builders
};

// Validates a jump-encoded matrix like bmSparseMat.check and keeps the host
// copy; no device work.
void sparse_set_host(SparseDev& S, int r_size, const int* r_jump, const int* r_nJump,
                     const float* m_val, int l_size, const int* l_jump, int l_nJump, int64_t nnz);
void sparse_set_host(SparseDev& S, int r_size, std::vector<int>&& r_jump, std::vector<int>&& r_nJump,
                     std::vector<float>&& m_val, int l_size, std::vector<int>&& l_jump);
// Uploads and decodes on the current device (idempotent, thread-safe).
void sparse_upload(SparseDev& S, hipStream_t s);
// Device matrix from decoded host arrays (absolute columns, per-row counts,
// optional output rows); no host copy of the jump encoding is kept.
void sparse_from_decoded(SparseDev& S, int r_size, int l_size, const std::vector<int>& col,
                         const std::vector<float>& val, const std::vector<int>& r_nJump,
                         const std::vector<int>* rowidx, hipStream_t s);
void build_chunks(SparseDev& S, const int* r_nJump, hipStream_t s);
// Builds S's staged groups for products over n_vec vectors from the decoded
// host arrays (per-row counts, columns); no groups when n_vec is too large
// for a group to hold the longest short row.  stage_kb / entries / rows:
// LDS budget of a group's vectors (0 = no groups), most entries and rows of
// a group; negative = the defaults (MONA_STAGE_KB, MONA_GROUP_ENTRIES,
// MONA_GROUP_ROWS or 16, 1024, 256).
void build_groups(SparseDev& S, const std::vector<int>& r_nJump, const std::vector<int>& col, int n_vec,
                  hipStream_t s, int stage_kb = -1, int entries = -1, int rows = -1);
// Builds the range form of S's long rows for products over n_vec <= 256
// vectors (a 16 KB stage of consecutive columns, MONA_RANGE_KB overrides it;
// 0 = no ranges) from the decoded host arrays; none when a long row's
// columns are not ascending.
void build_ranges(SparseDev& S, const std::vector<int>& r_nJump, const std::vector<int>& col, int n_vec,
                  hipStream_t s);
// The same from S's own device arrays (rebuilds the groups of an uploaded
// operator matrix with other sizes).
void rebuild_groups(SparseDev& S, int n_vec, hipStream_t s, int stage_kb, int entries, int rows);

// Rows with more entries than this are summed in chunks by the operator path.
constexpr int kLongRow = 64;

// w = S v for complex vectors stored interleaved (float2) with `n_vec`
// vectors, entry (row, k) at base[row * n_vec + k] (vector index fastest).
// Rows of w not stored in S are set to 0.  Rows of up to kLongRow entries
// are summed sequentially in stored order without FMA (bit-identical to the
// reference CPU kernel); longer rows are summed as consecutive chunks of
// kLongRow entries whose partial sums are then added in chunk order
// (deterministic; differs from the sequential sum only by rounding).
// A leading part of a matrix: its first `rows` stored rows (with their
// long-row chunks) writing output rows [0, l_size).
struct SpView {
    int rows = 0;
    int64_t l_size = 0;
    int n_long = 0;
    int64_t n_chunk = 0;
    int n_group = 0;
    int n_range = 0;
    int64_t n_rpart = 0;
};
// zero_rest = false leaves rows of w that S does not store untouched (for a
// consumer that never reads them).
// staged = false: the short rows gather from global memory even when S has
// staged groups (same results bit for bit).
void spmv_c2_rowmajor(const SparseDev& S, const float2* v, float2* w, int n_vec,
                      const float* yscale, hipStream_t s, const SpView* view = nullptr, bool zero_rest = true,
                      bool staged = true);
// The same product with v and / or w stored as bf16x2 (sums in float).
void spmv_c2_rowmajor_st(const SparseDev& S, const void* v, bool v_bf16, void* w, bool w_bf16, int n_vec,
                         const float* yscale, hipStream_t s, const SpView* view = nullptr,
                         bool zero_rest = true, bool staged = true);
// Block-diagonal stacking on the device: part f's columns are shifted by
// col_off[f] and its output rows by row_off[f] (out has r_size columns and
// l_size output rows; its stored rows are the parts' in order).  views[f]
// (n + 1 entries) is the leading part made of parts 0 .. f-1.
void sparse_block_diag(SparseDev& out, const std::vector<const SparseDev*>& parts,
                       const std::vector<int64_t>& col_off, const std::vector<int64_t>& row_off, int r_size,
                       int l_size, std::vector<SpView>* views, hipStream_t s);

}  // namespace mona

struct mona_sparse_s {
    mona::SparseDev S;
};
