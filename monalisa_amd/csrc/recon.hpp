// Reconstruction plan and operator declarations (recon.hip, admm.hip).
#pragma once
#include <memory>
#include <vector>

#include "gridmf.hpp"
#include "mona_internal.hpp"

namespace mona {

// Geometry passed by value to the prologue / epilogue kernels.
struct Geom {
    int N[3], n[3], off[3];
    int64_t Nu, nu;
    int nCh;
};

// Device-resident CG scalars (FP64).
struct CgScalars {
    double sqn, sqn_new, den, a, b;
    double den2;  // second term of the ADMM denominator
    int done, pad;
};

// A frame's gridding matrices.  The operators run on copies whose points
// (rows of Gu, columns of Gut) are in tile order -- points sorted by the 8^3
// cell tile that holds their first neighbour, ties kept in input order -- so
// that consecutive work-items touch neighbouring grid cells and k-space
// points.  Every row keeps its entries in the caller's order, so the sums
// are unchanged.  perm[s] is the caller's index of sorted point s.
struct Frame {
    const SparseDev* Gu = nullptr;   // caller's matrices
    const SparseDev* Gut = nullptr;
    std::shared_ptr<SparseDev> Gu_s, Gut_s;
    std::shared_ptr<DevBuf> perm;    // int32[P]
    // bit c: grid cell c is read by the forward gridding (fmask) / written
    // by the adjoint gridding (amask)
    std::shared_ptr<DevBuf> fmask, amask;
    bool masked = false;             // sparse frame: Gut_s stores touched cells only, passes masked
    // dense frame with row intervals: per column of the last-axis pass the
    // rows [lo, hi) spanning its touched cells; Gut_s stores the cells
    // inside them only and the last-axis passes skip the rest
    std::shared_ptr<DevBuf> ival;
    // 3-D: per x, the rows [lo, hi) of the middle-axis pass spanning the
    // columns with cells (one copy per pruned z row); columns outside that
    // span are marked (-1, -1) in ival and skipped by both passes
    std::shared_ptr<DevBuf> ivaly;
    bool ivals = false;
    int64_t P = 0;
    // matrix-free gridding data (frames whose Gu came with its trajectory)
    std::shared_ptr<MfPoints> mf;
};

// Records HIP events around a region when kernel timing is enabled.
class TimedRegion {
  public:
    TimedRegion(int kind, hipStream_t s);
    ~TimedRegion();
  private:
    int kind_;
    hipStream_t s_;
    hipEvent_t a_ = nullptr, b_ = nullptr;
    bool on_ = false;
};

struct Work {
    DevBuf hy, k1, k2, i1, i2, i3, i4, sc, partial, hist;
    DevBuf ys, hy0;             // k-space data and weights in tile order
    DevBuf kpub;                // staging of the public operator calls
    DevBuf z1, z2, z3, z4, z5, z6;  // ADMM image-space buffers (z1, z2: one block per term)
};
// the reduction scratch of compute_hy (the partial sums' buffer, free then)
inline float* hy_scratch(Work& w) {
    w.partial.ensure(1025 * sizeof(float));
    return w.partial.as<float>();
}

struct Plan {
    int imDim = 0, nCh = 0, nWin = 3, ktype_ = 0;
    int N[3] = {1, 1, 1}, n[3] = {1, 1, 1};
    double dK[3] = {1, 1, 1};
    double kparam[3] = {0, 0, 0};
    int64_t Nu = 1, nu = 1;
    double HX = 1.0;
    bool coils_set = false;
    DevBuf K, KFC, KFCc, grid;
    DevBuf fft_a, fft_b;  // pruned-pass intermediates
    std::vector<Frame> frames;
    // per-frame deformations (Tu, Tut) of the solvers' two difference terms;
    // frames without one use the identity
    std::vector<std::pair<const SparseDev*, const SparseDev*>> deform[2];
    Work work;

    void init(int imDim, const int* N, const int* n, const double* dK, int nCh, int ktype, int nWin,
              const double* kp, int nParam);
    void set_coils(const float2* C, hipStream_t s);
    Geom geom() const;
    const Frame& frame(int f) const;
    void ensure_grid();
    void launch_check();
    float kmax() const { return (float)kparam[ktype_ == MONA_KERNEL_GAUSS ? 1 : 2]; }

    // FFT stages of the operators over nb frames stored back to back
    // (images [nb][nu], grids [nb][Nu][nCh]); scratch a/b sized for nb.
    // (bf: grid-side arrays stored as bf16x2)
    // 3-D pruned transforms with y-major grid and intermediate layouts (the
    // gridding matrices of set_frame then address cells y-major)
    bool grid_yz() const;
    void forward_fft_pruned(const float2* x, float2* grid, int nb, DevBuf& a, DevBuf& b, hipStream_t s,
                            bool bf = false, const uint32_t* out_mask = nullptr, const int2* out_ival = nullptr,
                            const int2* mid_ival = nullptr);
    void adjoint_fft_pruned(const float2* grid, float2* x, bool combine, float scale, int nb, DevBuf& a, DevBuf& b,
                            hipStream_t s, bool bf = false, const uint32_t* in_mask = nullptr,
                            const int2* in_ival = nullptr, const int2* mid_ival = nullptr);
    // the last-axis pass can be confined (register kernel, fused x pass)
    bool limits_usable() const;
    // frame f (f < 0: every frame of the plan) runs with cell masks / row intervals
    bool masked_grid(int f) const;
    bool ival_grid(int f) const;
    // the grid-side confinement of frame f for the forward / adjoint passes
    const uint32_t* fwd_mask(int f) const;
    const uint32_t* adj_mask(int f) const;
    const int2* ivals_of(int f) const;
    const int2* mid_ivals_of(int f) const;
    int64_t fft_scratch_elems() const;
    int64_t mid_rs() const;  // y row stride of the 3-D y-major middle intermediate
    void ensure_fft_scratch();

    // Batched form over all frames: one block-diagonal gridding launch and
    // frame-batched FFT passes per operator application.
    std::shared_ptr<SparseDev> Gu_b, Gut_b;
    std::vector<int64_t> koff_b;       // k-space offset (float2) of each frame
    DevBuf grid_b, fft_ab, fft_bb;
    DevBuf mask_fb, mask_ab;           // stacked cell masks of the frames
    DevBuf ival_b, ivaly_b;            // stacked row intervals of the frames
    bool use_batch = false;            // teva_chain uses the batched operators
    bool use_graph = false;            // ADMM iterations replayed from a HIP graph
    int fft_tile = 0;                  // plain FFT pass tile (0 = default)
    // matrix-free gridding: used for the frames that have its data (all
    // frames of a batch); off = the stored matrices everywhere
    bool use_mf = true;
    bool use_staged = true;            // adjoint gridding from LDS-staged groups (Gut_s groups)
    // group sizes of the staged adjoint gridding (-1: the defaults); set by
    // the autotuner, used by set_frame
    int grp_stage_kb = -1, grp_entries = -1, grp_rows = -1;
    void regroup(hipStream_t s);       // every frame's Gut_s with the current sizes
    bool mf_kernel_ok = false;
    MfKernel mfk;
    MfGeo mf_geo() const;
    bool frame_mf(int f) const;        // frame f grids matrix-free
    bool batch_mf(int nA) const;       // frames 0..nA-1 all do
    MfPoints mf_b;                     // frames' points back to back
    DevBuf mf_b_frame;                 // uint16 frame of each point
    int fft_lines = 0;                 // register FFT kernel lines per tile (16/32/64; 0 default, -1 off)
    std::vector<SpView> view_u, view_t;  // leading frames 0..nA-1 of Gu_b / Gut_b
    void build_batch(hipStream_t s);
    // over frames 0..nA-1 (nA <= 0: all)
    void forward_batch(const float2* x, float2* y, hipStream_t s, int nA = 0);
    void adjoint_batch(const float2* y, float2* x, const float* hy, float scale, hipStream_t s, int nA = 0);

    void set_frame(int f, const SparseDev* Gu, const SparseDev* Gut, hipStream_t s);

    // Witness snapshots of the solvers (bmWitnessInfo.m:92-190): after outer
    // iteration c (0-based) a solver copies frame 0's witness image (the
    // whole image in 1-D / 2-D, its middle z slice in 3-D) into
    // wit_slices + c * witness_elems() while c < wit_slots, and the whole x
    // into wit_full + k * x_elems when c + 1 == wit_full_ind[k].
    float2* wit_slices = nullptr;
    int wit_slots = 0;
    float2* wit_full = nullptr;
    std::vector<int> wit_full_ind;
    int64_t witness_elems() const;
    void witness(int c, const float2* x, int64_t x_elems, hipStream_t s);

    // y = M x  (x [nu], y [P][nCh] in the frame's tile order)
    void forward(int f, const float2* x, float2* y, hipStream_t s);
    // x = scale * sum_c KFC_conj .* crop(cifft(Gut (hy .* y)))   (hy may be
    // NULL; y and hy in tile order)
    void adjoint(int f, const float2* y, float2* x, bool combine, const float* hy, float scale,
                 hipStream_t s);
    // The same operators with k-space in the caller's point order.
    void forward_public(int f, const float2* x, float2* y, hipStream_t s);
    void adjoint_public(int f, const float2* y, float2* x, bool combine, hipStream_t s);
    // tile-order copies of caller-order k-space data / point weights
    void to_tile_order(int f, const float2* src, float2* dst, hipStream_t s);
    void weights_to_tile_order(int f, const float* src, float* dst, hipStream_t s);
    void sensa(int f, float2* x, const float2* y, const float* ve, float ve_max, int nIter, int nCGD,
               double* residual, hipStream_t s);
    // all frames at once (independent systems, batched operators)
    void sensa_batch(int nA, float2* x, const float2* y, const long long* y_off, const float* ve,
                     const long long* ve_off, const float* ve_max, int nIter, int nCGD, double* residual,
                     hipStream_t s);
    // device memory held by the plan (its own buffers and frame matrices)
    size_t device_bytes() const;
    // mixed precision: grid-side arrays and k-space in bf16x2 storage
    void check_bf16() const;
    void forward_bf16(int f, const float2* x, bf16x2* y, hipStream_t s);
    void adjoint_bf16(int f, const bf16x2* y, float2* x, const float* hy, float scale, hipStream_t s);
    void sensa_ir(int f, float2* x, const float2* y, const float* ve, float ve_max, int nInner, int nOuter,
                  double tol, bool bf16, double* rel_out, int* n_outer, int* n_steps, hipStream_t s);
};

// vector helpers (recon.hip)
void vsub(const float2* a, const float2* b, float2* out, int64_t n, hipStream_t s);
void vcopy(const float2* a, float2* out, int64_t n, hipStream_t s);
void vaxpy(float2* y, const float2* x, const double* a, float sign, int64_t n, hipStream_t s);
void vxpby(float2* p, const float2* d, const double* b, int64_t n, hipStream_t s);
void sqnorm(const float2* a, int64_t n, const float* w, int nvec, double cw, double* out, bool acc,
            double* partial, hipStream_t s);
void dotre(const float2* a, const float2* b, int64_t n, double cw, double* out, double* partial, hipStream_t s);
// HY = min(ve, vmax) (bmSensa.m:29-32); vmax <= 0 takes the maximum of ve,
// reduced on the device in scratch (at least 1025 floats: hy_scratch)
void compute_hy(const float* ve, float vmax, float* hy, int64_t n, float* scratch, hipStream_t s);
// a = sqn / den; the sticky done flag is raised when sqn < eps (bmSensa.m:54
// uses 10 * eps('single'))
void cg_alpha(CgScalars* sc, hipStream_t s, double eps = 10.0 * 1.1920928955078125e-07);
void cg_beta(CgScalars* sc, hipStream_t s);
void cg_reset(CgScalars* sc, hipStream_t s);
void dstore(double* dst, const double* src, hipStream_t s);

}  // namespace mona

struct mona_plan_s {
    mona::Plan P;
};
