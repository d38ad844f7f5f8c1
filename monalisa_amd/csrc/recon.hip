// Reconstruction plan, deapodisation, the forward/adjoint operators and the
// CG-SENSE / ADMM-TV solvers.
//
//   forward  y = Gu * cfft(zerofill(x .* KFC))          bmShanna_MATLAB.m:6-56
//   adjoint  x = sum_c KFC_conj .* crop(cifft(Gut * y))  bmNakatsha_MATLAB.m:6-60
//
// All vectors stay in HBM; the CG scalars live in a small device struct and
// are computed in FP64 by single-workgroup kernels, so a solver iteration
// never waits on the host.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "fft.hpp"
#include "mona_internal.hpp"
#include "recon.hpp"

namespace mona {

// ================================================================ timing
namespace {
struct TimingRec {
    int kind;
    hipEvent_t a, b;
};
std::mutex g_tmu;
bool g_timing = false;
std::vector<TimingRec> g_trecs;
double g_tms[4] = {0, 0, 0, 0};
long long g_tn[4] = {0, 0, 0, 0};

void timing_flush_locked() {
    for (auto& r : g_trecs) {
        float ms = 0;
        if (hipEventSynchronize(r.b) == hipSuccess && hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            g_tms[r.kind] += ms;
            g_tn[r.kind] += 1;
        }
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    g_trecs.clear();
}
}  // namespace

TimedRegion::TimedRegion(int kind, hipStream_t s) : kind_(kind), s_(s) {
    if (!g_timing) return;
    hipStreamCaptureStatus cst = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cst) != hipSuccess || cst != hipStreamCaptureStatusNone) return;
    if (hipEventCreate(&a_) != hipSuccess || hipEventCreate(&b_) != hipSuccess) return;
    (void)hipEventRecord(a_, s_);
    on_ = true;
}
TimedRegion::~TimedRegion() {
    if (!on_) return;
    (void)hipEventRecord(b_, s_);
    std::lock_guard<std::mutex> lk(g_tmu);
    g_trecs.push_back({kind_, a_, b_});
    if (g_trecs.size() > 4096) timing_flush_locked();
}

// ================================================================ kernels

__device__ __forceinline__ float2 cmulf(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Deapodisation on the image grid n_u (bmK.m:29-151 followed by the crop of
// bmKF.m:16): K(k) = | sum_x w(|x|) cos(2 pi sum_d x_d k_d / N_d) | / sum_x w
// over the window points x (half-integer spacing, per-axis box |x_d| <= h),
// then min(1/K, K_max).  The maximum of |Re DFT| over the grid is its value
// at k = 0, sum_x w, because every w >= 0.
__global__ void deapod_kernel(float* K, int imDim, int n0, int n1, int n2, int N0, int N1, int N2,
                              const double* wpts, int npts, double wsum, double Kmax) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t tot = (int64_t)n0 * n1 * n2;
    if (i >= tot) return;
    int ix = (int)(i % n0);
    int iy = (int)((i / n0) % n1);
    int iz = (int)(i / ((int64_t)n0 * n1));
    double k0 = ix - n0 / 2, k1 = iy - n1 / 2, k2 = iz - n2 / 2;
    double acc = 0.0;
    for (int p = 0; p < npts; p++) {
        const double* q = wpts + 4 * p;
        double ph = q[0] * k0 / N0;
        if (imDim > 1) ph += q[1] * k1 / N1;
        if (imDim > 2) ph += q[2] * k2 / N2;
        acc += q[3] * cospi(2.0 * ph);
    }
    double v = fabs(acc) / wsum;
    double r = 1.0 / v;
    K[i] = (float)(r < Kmax ? r : Kmax);
}

// KFC = (K .* F) .* C and KFC_conj = (K .* F_conj) .* conj(C) in single
// precision (bmKF.m:33 with F = single(1/prod(N_u)/prod(dK_u)); bmKF_conj.m:36).
__global__ void kfc_kernel(const float* K, const float2* C, float2* KFC, float2* KFCc, int64_t nu,
                           int nCh, float F, float Fc) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nu * nCh) return;
    int64_t i = t / nCh;
    float2 c = C[t];
    float kf = K[i] * F, kfc = K[i] * Fc;
    KFC[t] = make_float2(kf * c.x, kf * c.y);
    KFCc[t] = make_float2(kfc * c.x, kfc * -c.y);
}

// Forward prologue: grid[g][c] = x[i] * KFC[i][c] inside the centred n_u
// window (bmImZeroFill.m:41-51), 0 elsewhere.
__global__ void prologue_kernel(const float2* __restrict__ x, const float2* __restrict__ KFC,
                                float2* __restrict__ grid, Geom g) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t tot = g.Nu * g.nCh;
    if (t >= tot) return;
    int c = (int)(t % g.nCh);
    int64_t cell = t / g.nCh;
    int gx = (int)(cell % g.N[0]);
    int gy = (int)((cell / g.N[0]) % g.N[1]);
    int gz = (int)(cell / ((int64_t)g.N[0] * g.N[1]));
    int ix = gx - g.off[0], iy = gy - g.off[1], iz = gz - g.off[2];
    float2 v = make_float2(0.f, 0.f);
    if (ix >= 0 && ix < g.n[0] && iy >= 0 && iy < g.n[1] && iz >= 0 && iz < g.n[2]) {
        int64_t i = ix + (int64_t)g.n[0] * (iy + (int64_t)g.n[1] * iz);
        v = cmulf(x[i], KFC[i * g.nCh + c]);
    }
    grid[t] = v;
}

// Adjoint epilogue: x[i] = scale * sum_c KFC_conj[i][c] * grid[crop(i)][c]
// (bmImCrope, then .*KFC_conj and sum over coils, bmNakatsha_MATLAB.m:46-55);
// with combine == 0 the per-coil products are written instead.
__global__ void epilogue_kernel(const float2* __restrict__ grid, const float2* __restrict__ KFCc,
                                float2* __restrict__ x, Geom g, int combine, float scale) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.nu) return;
    int ix = (int)(i % g.n[0]);
    int iy = (int)((i / g.n[0]) % g.n[1]);
    int iz = (int)(i / ((int64_t)g.n[0] * g.n[1]));
    int64_t cell = (ix + g.off[0]) + (int64_t)g.N[0] * ((iy + g.off[1]) + (int64_t)g.N[1] * (iz + g.off[2]));
    const float2* gp = grid + cell * g.nCh;
    const float2* kp = KFCc + i * g.nCh;
    if (combine) {
        float2 acc = make_float2(0.f, 0.f);
        for (int c = 0; c < g.nCh; c++) {
            float2 p = cmulf(gp[c], kp[c]);
            acc.x += p.x;
            acc.y += p.y;
        }
        x[i] = make_float2(acc.x * scale, acc.y * scale);
    } else {
        for (int c = 0; c < g.nCh; c++) {
            float2 p = cmulf(gp[c], kp[c]);
            x[i * g.nCh + c] = make_float2(p.x * scale, p.y * scale);
        }
    }
}

// ---------------------------------------------------------------- vectors
__global__ void k_sub(const float2* a, const float2* b, float2* out, int64_t n) {  // out = a - b
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) { float2 x = a[i], y = b[i]; out[i] = make_float2(x.x - y.x, x.y - y.y); }
}
__global__ void k_copy(const float2* a, float2* out, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = a[i];
}
// y += coef * x where coef = sign * scalar (device double, rounded to float)
__global__ void k_axpy_dev(float2* y, const float2* x, const double* scal, float sign, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        float a = sign * (float)(*scal);
        float2 xv = x[i], yv = y[i];
        yv.x = yv.x + a * xv.x;
        yv.y = yv.y + a * xv.y;
        y[i] = yv;
    }
}
// p = d + b * p
__global__ void k_xpby_dev(float2* p, const float2* d, const double* scal, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        float b = (float)(*scal);
        float2 pv = p[i], dv = d[i];
        p[i] = make_float2(dv.x + b * pv.x, dv.y + b * pv.y);
    }
}
// The same updates on two complex values per work-item (16-byte accesses,
// every pointer 16-byte aligned); the last value of an odd n is done alone
// by the last work-item.  Each value gets the arithmetic of the one-value
// kernels above.
__global__ void k_sub2(const float2* a, const float2* b, float2* out, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (2 * i + 1 < n) {
        const float4 x = reinterpret_cast<const float4*>(a)[i], y = reinterpret_cast<const float4*>(b)[i];
        reinterpret_cast<float4*>(out)[i] = make_float4(x.x - y.x, x.y - y.y, x.z - y.z, x.w - y.w);
    } else if (2 * i < n) {
        const float2 x = a[2 * i], y = b[2 * i];
        out[2 * i] = make_float2(x.x - y.x, x.y - y.y);
    }
}
__global__ void k_copy2(const float2* a, float2* out, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (2 * i + 1 < n) reinterpret_cast<float4*>(out)[i] = reinterpret_cast<const float4*>(a)[i];
    else if (2 * i < n) out[2 * i] = a[2 * i];
}
__global__ void k_axpy_dev2(float2* y, const float2* x, const double* scal, float sign, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const float a = sign * (float)(*scal);
    if (2 * i + 1 < n) {
        const float4 xv = reinterpret_cast<const float4*>(x)[i];
        float4 yv = reinterpret_cast<float4*>(y)[i];
        yv.x = yv.x + a * xv.x;
        yv.y = yv.y + a * xv.y;
        yv.z = yv.z + a * xv.z;
        yv.w = yv.w + a * xv.w;
        reinterpret_cast<float4*>(y)[i] = yv;
    } else if (2 * i < n) {
        const float2 xv = x[2 * i];
        float2 yv = y[2 * i];
        yv.x = yv.x + a * xv.x;
        yv.y = yv.y + a * xv.y;
        y[2 * i] = yv;
    }
}
__global__ void k_xpby_dev2(float2* p, const float2* d, const double* scal, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const float b = (float)(*scal);
    if (2 * i + 1 < n) {
        const float4 pv = reinterpret_cast<const float4*>(p)[i], dv = reinterpret_cast<const float4*>(d)[i];
        reinterpret_cast<float4*>(p)[i] = make_float4(dv.x + b * pv.x, dv.y + b * pv.y, dv.z + b * pv.z,
                                                      dv.w + b * pv.w);
    } else if (2 * i < n) {
        const float2 pv = p[2 * i], dv = d[2 * i];
        p[2 * i] = make_float2(dv.x + b * pv.x, dv.y + b * pv.y);
    }
}
bool al16(const void* a, const void* b = nullptr, const void* c = nullptr) {
    return ((uintptr_t)a | (uintptr_t)b | (uintptr_t)c) % 16 == 0;
}
// dst[s][c] = src[perm[s]][c]  (caller order -> tile order)
__global__ void k_gather_pts(const float2* __restrict__ src, const int* __restrict__ perm, float2* __restrict__ dst,
                             int64_t P, int nCh) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= P * nCh) return;
    int64_t s = t / nCh;
    int c = (int)(t - s * nCh);
    dst[t] = src[(int64_t)perm[s] * nCh + c];
}
// dst[perm[s]][c] = src[s][c]  (tile order -> caller order)
__global__ void k_scatter_pts(const float2* __restrict__ src, const int* __restrict__ perm, float2* __restrict__ dst,
                              int64_t P, int nCh) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= P * nCh) return;
    int64_t s = t / nCh;
    int c = (int)(t - s * nCh);
    dst[(int64_t)perm[s] * nCh + c] = src[t];
}
// bf16x2 tile order <-> float2 caller order
__global__ void k_scatter_pts_bf(const bf16x2* __restrict__ src, const int* __restrict__ perm,
                                 float2* __restrict__ dst, int64_t P, int nCh) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= P * nCh) return;
    int64_t s = t / nCh;
    int c = (int)(t - s * nCh);
    dst[(int64_t)perm[s] * nCh + c] = ld_c(src + t);
}
__global__ void k_gather_pts_bf(const float2* __restrict__ src, const int* __restrict__ perm,
                                bf16x2* __restrict__ dst, int64_t P, int nCh) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= P * nCh) return;
    int64_t s = t / nCh;
    int c = (int)(t - s * nCh);
    st_c(dst + t, src[(int64_t)perm[s] * nCh + c]);
}
__global__ void k_gather_w(const float* __restrict__ src, const int* __restrict__ perm, float* __restrict__ dst,
                           int64_t P) {
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s < P) dst[s] = src[perm[s]];
}

__global__ void k_hy(const float* ve, float vmax, float* hy, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) hy[i] = fminf(ve[i], vmax);
}
// the same with the bound read from device memory
__global__ void k_hy_dev(const float* ve, const float* vmax, float* hy, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) hy[i] = fminf(ve[i], *vmax);
}

// FP64 weighted squared norm sum_i w(i) |a_i|^2 (w = per-point weight for
// [n/nvec][nvec] data, or a constant).  Fixed launch shape and fixed
// reduction tree: deterministic.
constexpr int kRedThreads = 256;
constexpr int kRedBlocks = 1024;

// Each thread sums the elements i0, i0 + G, i0 + 2G, ... (G = gridDim *
// kRedThreads) in that order; the loads of four consecutive terms are issued
// before their adds, which keeps the order.  Weight rows use 32-bit
// division (n < 2^31 is checked on the host).
template <bool W>
__device__ __forceinline__ double sqn_term(const float2* a, const float* w, int64_t i, unsigned nvec) {
    float2 v = a[i];
    double m = (double)v.x * v.x + (double)v.y * v.y;
    if constexpr (W) m *= (double)w[(unsigned)i / nvec];
    return m;
}

template <bool W>
__global__ void __launch_bounds__(kRedThreads)
k_sqnorm_partial(const float2* a, int64_t n, const float* w, int nvec, double cw, double* partial) {
    __shared__ double sh[kRedThreads];
    double s = 0.0;
    const int64_t G = (int64_t)gridDim.x * kRedThreads;
    int64_t i = (int64_t)blockIdx.x * kRedThreads + threadIdx.x;
    const unsigned nv = (unsigned)nvec;
    for (; i + 3 * G < n; i += 4 * G) {
        float2 v0 = a[i], v1 = a[i + G], v2 = a[i + 2 * G], v3 = a[i + 3 * G];
        double m0 = (double)v0.x * v0.x + (double)v0.y * v0.y;
        double m1 = (double)v1.x * v1.x + (double)v1.y * v1.y;
        double m2 = (double)v2.x * v2.x + (double)v2.y * v2.y;
        double m3 = (double)v3.x * v3.x + (double)v3.y * v3.y;
        if constexpr (W) {
            m0 *= (double)w[(unsigned)i / nv];
            m1 *= (double)w[(unsigned)(i + G) / nv];
            m2 *= (double)w[(unsigned)(i + 2 * G) / nv];
            m3 *= (double)w[(unsigned)(i + 3 * G) / nv];
        }
        s += m0;
        s += m1;
        s += m2;
        s += m3;
    }
    for (; i < n; i += G) s += sqn_term<W>(a, w, i, nv);
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = kRedThreads / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0] * cw;
}

// partial[block] = cw * sum_i Re(conj(a_i) b_i), FP64 accumulation in a
// fixed order (deterministic; loads issued four terms ahead)
__global__ void __launch_bounds__(kRedThreads)
k_dotre_partial(const float2* a, const float2* b, int64_t n, double cw, double* partial) {
    __shared__ double sh[kRedThreads];
    double s = 0.0;
    const int64_t G = (int64_t)gridDim.x * kRedThreads;
    int64_t i = (int64_t)blockIdx.x * kRedThreads + threadIdx.x;
    for (; i + 3 * G < n; i += 4 * G) {
        float2 u0 = a[i], u1 = a[i + G], u2 = a[i + 2 * G], u3 = a[i + 3 * G];
        float2 v0 = b[i], v1 = b[i + G], v2 = b[i + 2 * G], v3 = b[i + 3 * G];
        s += (double)u0.x * v0.x + (double)u0.y * v0.y;
        s += (double)u1.x * v1.x + (double)u1.y * v1.y;
        s += (double)u2.x * v2.x + (double)u2.y * v2.y;
        s += (double)u3.x * v3.x + (double)u3.y * v3.y;
    }
    for (; i < n; i += G) {
        float2 u = a[i], v = b[i];
        s += (double)u.x * v.x + (double)u.y * v.y;
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = kRedThreads / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0] * cw;
}

// out[slot] (= or +=) sum of partials
__global__ void __launch_bounds__(kRedThreads)
k_reduce_final(const double* partial, int np, double* out, int accumulate) {
    __shared__ double sh[kRedThreads];
    double s = 0.0;
    for (int i = threadIdx.x; i < np; i += kRedThreads) s += partial[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = kRedThreads / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = accumulate ? (*out + sh[0]) : sh[0];
}

// Maximum of v[0, n) (NaN values ignored, -inf for none): partial maxima
// per block, then one block folds them into out.  A maximum is exact, so the
// launch shape does not change the result.
__global__ void __launch_bounds__(kRedThreads) k_max_partial(const float* v, int64_t n, float* partial) {
    __shared__ float sh[kRedThreads];
    float m = -INFINITY;
    const int64_t G = (int64_t)gridDim.x * kRedThreads;
    for (int64_t i = (int64_t)blockIdx.x * kRedThreads + threadIdx.x; i < n; i += G) m = fmaxf(m, v[i]);
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int o = kRedThreads / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] = fmaxf(sh[threadIdx.x], sh[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}
__global__ void __launch_bounds__(kRedThreads) k_max_final(const float* partial, int nb, float* out) {
    __shared__ float sh[kRedThreads];
    float m = -INFINITY;
    for (int i = threadIdx.x; i < nb; i += kRedThreads) m = fmaxf(m, partial[i]);
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int o = kRedThreads / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] = fmaxf(sh[threadIdx.x], sh[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

// CG scalar updates (bmSensa.m:54-75):  at the top of an iteration the
// sticky `done` flag is raised when sqn < eps; a = done ? 0 : sqn / den.
__global__ void k_cg_alpha(CgScalars* s, double eps) {
    if (!s->done && s->sqn < eps) s->done = 1;
    s->a = s->done ? 0.0 : s->sqn / s->den;
}
__global__ void k_cg_beta(CgScalars* s) {
    s->b = s->done ? 0.0 : s->sqn_new / s->sqn;
    if (!s->done) s->sqn = s->sqn_new;
}
__global__ void k_cg_reset(CgScalars* s) {
    s->done = 0;
    s->a = s->b = 0.0;
}
__global__ void k_store(double* dst, const double* src) { *dst = *src; }

// ================================================================ host side

void Plan::launch_check() { MONA_HIP(hipGetLastError()); }

Geom Plan::geom() const {
    Geom g{};
    for (int d = 0; d < 3; d++) {
        g.N[d] = N[d];
        g.n[d] = n[d];
        g.off[d] = N[d] / 2 - n[d] / 2;
    }
    g.Nu = Nu;
    g.nu = nu;
    g.nCh = nCh;
    return g;
}

void Plan::init(int imDim_, const int* N_, const int* n_, const double* dK_, int nCh_, int ktype,
                int nWin_, const double* kp, int nParam) {
    MONA_CHECK(imDim_ >= 1 && imDim_ <= 3, MONA_ERR_ARG, "imDim must be 1, 2 or 3");
    MONA_CHECK(nCh_ >= 1, MONA_ERR_ARG, "nCh must be >= 1");
    MONA_CHECK((ktype == MONA_KERNEL_GAUSS && nParam == 2) || (ktype == MONA_KERNEL_KAISER && nParam == 3),
               MONA_ERR_ARG, "Wrong list of gridding kernel parameters.");
    MONA_CHECK(nWin_ >= 1 && nWin_ <= 9, MONA_ERR_ARG, "nWin must be in [1, 9]");
    imDim = imDim_;
    nCh = nCh_;
    ktype_ = ktype;
    nWin = nWin_;
    for (int i = 0; i < 3; i++) kparam[i] = i < nParam ? (double)(float)kp[i] : 0.0;
    Nu = 1;
    nu = 1;
    HX = 1.0;
    for (int d = 0; d < 3; d++) {
        if (d < imDim) {
            N[d] = N_[d];
            n[d] = n_ ? n_[d] : N_[d];
            dK[d] = (double)(float)dK_[d];
            MONA_CHECK(N[d] % 2 == 0, MONA_ERR_ARG, "N_u must have all components even for the Fourier transform. ");
            MONA_CHECK(n[d] % 2 == 0, MONA_ERR_ARG, "n_u must have all components even for the Fourier transform. ");
            MONA_CHECK(n[d] >= 2 && n[d] <= N[d], MONA_ERR_ARG, "n_u must be in [2, N_u]");
            MONA_CHECK(fft_size_supported(N[d]), MONA_ERR_ARG, "N_u component is not a supported FFT length");
            MONA_CHECK(dK[d] > 0, MONA_ERR_ARG, "dK_u must be > 0");
            // HX = prod(dX_u), dX_u = single((1/single(dK_u))/single(N_u))  (bmSensa.m:26-27)
            float dx = (1.0f / (float)dK[d]) / (float)N[d];
            HX = (double)((float)HX * dx);
        } else {
            N[d] = 1;
            n[d] = 1;
            dK[d] = 1.0;
        }
        Nu *= N[d];
        nu *= n[d];
    }
    // matrix-free gridding (nWin 3, 2-D / 3-D grids of at most 1024 cells
    // an axis, a kernel whose series fits kMfTerms); MONA_MF=0 turns it off
    mf_kernel_ok = nWin == 3 && (imDim == 2 || imDim == 3) && N[0] <= 1024 && N[1] <= 1024 && N[2] <= 1024 &&
                   mf_kernel_params(ktype, nParam, kp, &mfk);
    static const bool mf_env = [] {
        const char* e = std::getenv("MONA_MF");
        return !(e && std::atoi(e) == 0);
    }();
    use_mf = mf_env;
}

MfGeo Plan::mf_geo() const {
    MfGeo g;
    g.N0 = N[0]; g.N1 = N[1]; g.N2 = N[2];
    if (imDim == 2) {
        g.sy = N[0]; g.sz = 0;
    } else if (grid_yz()) {
        g.sy = N[0] * N[2]; g.sz = N[0];     // [N_y][N_z][N_x]
    } else {
        g.sy = N[0]; g.sz = N[0] * N[1];     // [N_z][N_y][N_x]
    }
    return g;
}

bool Plan::frame_mf(int f) const { return use_mf && frame(f).mf != nullptr; }

bool Plan::batch_mf(int nA) const {
    if (!use_mf || !mf_b.P) return false;
    for (int f = 0; f < nA; f++)
        if (!frame(f).mf) return false;
    return true;
}

void Plan::ensure_grid() {
    grid.ensure((size_t)Nu * nCh * sizeof(float2));
}

static inline unsigned nb256(int64_t n) { return blocks_for(n, 256); }

void Plan::set_coils(const float2* C, hipStream_t s) {
    // window points of bmK (half-integer grid, per-axis box |x| <= nWin/2)
    int h = nWin / 2;
    std::vector<double> pts;
    double wsum = 0.0;
    double I0a = 1.0;
    auto i0 = [](double x) {
        double q = 0.25 * x * x, term = 1.0, sum = 1.0;
        for (int k = 1; k < 500; k++) {
            term *= q / ((double)k * k);
            sum += term;
            if (term < sum * 1e-17) break;
        }
        return sum;
    };
    if (ktype_ == MONA_KERNEL_KAISER) I0a = i0(kparam[1]);
    auto weight = [&](double d) {
        if (ktype_ == MONA_KERNEL_GAUSS) {
            double r = d / kparam[0];
            return std::exp(-0.5 * r * r) / (std::sqrt(2.0 * M_PI) * kparam[0]);
        }
        double r = d / kparam[0];
        double w = 1.0 - r * r;
        if (w < 0) w = 0;
        return i0(kparam[1] * std::sqrt(w)) / I0a;
    };
    int span = 4 * h + 1;  // x in {-h, -h+1/2, ..., h}
    int sy = imDim > 1 ? span : 1, sz = imDim > 2 ? span : 1;
    for (int c = 0; c < sz; c++)
        for (int b = 0; b < sy; b++)
            for (int a = 0; a < span; a++) {
                double x = -h + 0.5 * a, y = imDim > 1 ? -h + 0.5 * b : 0.0, z = imDim > 2 ? -h + 0.5 * c : 0.0;
                if (imDim == 1 && !(x >= h && x <= h)) continue;  // bmK.m:88 (1-D mask x < h | x > h)
                double w = weight(std::sqrt(x * x + y * y + z * z));
                pts.insert(pts.end(), {x, y, z, w});
                wsum += w;
            }
    int npts = (int)(pts.size() / 4);
    DevBuf dpts(pts.size() * sizeof(double));
    MONA_HIP(hipMemcpyAsync(dpts.p, pts.data(), pts.size() * sizeof(double), hipMemcpyHostToDevice, s));
    K.ensure((size_t)nu * sizeof(float));
    deapod_kernel<<<nb256(nu), 256, 0, s>>>(K.as<float>(), imDim, n[0], n[1], n[2], N[0], N[1], N[2],
                                            dpts.as<double>(), npts, wsum, kparam[imDim == 0 ? 0 : (ktype_ == MONA_KERNEL_GAUSS ? 1 : 2)]);
    launch_check();
    float F = 1.0f, Fc = 1.0f;
    {
        float pN = 1.0f, pdK = 1.0f;
        for (int d = 0; d < imDim; d++) { pN *= (float)N[d]; pdK *= (float)dK[d]; }
        F = (float)(1.0 / (double)pN / (double)pdK);
        Fc = (float)(1.0 / (double)pdK);
    }
    KFC.ensure((size_t)nu * nCh * sizeof(float2));
    KFCc.ensure((size_t)nu * nCh * sizeof(float2));
    kfc_kernel<<<nb256(nu * nCh), 256, 0, s>>>(K.as<float>(), C, KFC.as<float2>(), KFCc.as<float2>(), nu,
                                                nCh, F, Fc);
    launch_check();
    MONA_HIP(hipStreamSynchronize(s));  // dpts goes out of scope
    coils_set = true;
}

const Frame& Plan::frame(int f) const {
    MONA_CHECK(f >= 0 && f < (int)frames.size() && frames[f].Gu_s && frames[f].Gut_s, MONA_ERR_STATE,
               "frame " + std::to_string(f) + " has no gridding matrices");
    return frames[f];
}

// Tile-order copies of a frame's matrices (host decode of the reference
// encoding, stable sort of the points by 8^3 cell tile, upload).
// Row intervals for dense frames (MONA_GRID_INTERVALS=0 turns them off).
// y-major layouts of the 3-D pruned transforms (MONA_FFT_YZ=0: the plain
// [z][y][x] order).  Each frame's grid is stored [N_y][N_z][N_x][nCh]
// (the gridding matrices address cells in that order) and the middle
// intermediate [N_y][n_z][N_x nCh], so that every z pass reads and writes
// rows N_x nCh values apart; only the y passes keep one large-stride side,
// the forward one on its writes and the adjoint one on its reads.
bool yz_interleave() {
    static const bool on = [] {
        const char* e = std::getenv("MONA_FFT_YZ");
        return e ? std::atoi(e) != 0 : true;
    }();
    return on;
}

bool Plan::grid_yz() const { return imDim == 3 && yz_interleave() && fft_pass_fusable(N[0], nCh); }

// Order of the 8^3 cell tiles in the gridding structures (points of Gu_s,
// rows of Gut_s): 0 = the grid's storage order (raster over tiles), 1 =
// Morton order (bits of the three tile coordinates interleaved), so that
// runs of consecutive tiles are compact in all three axes and neighbouring
// tiles' shared cells meet in one L2.  MONA_TILE_ORDER overrides it.
static int tile_order() {
    static const int o = [] {
        const char* e = std::getenv("MONA_TILE_ORDER");
        return e ? std::atoi(e) : 0;
    }();
    return o;
}

// Order of the cells of an 8^3 tile in the adjoint matrix: cubes of c cells
// a side in raster order (slowest axis first), raster order inside a cube;
// c = 8 is the tile's plain raster order, c = 0 the Morton order (cubes of
// 2 in cubes of 4 in the tile).  Cubes of 2 (the default) keep a group of
// consecutive rows compact in all three axes, so it stages fewer samples
// than a slab of whole x-rows would.  MONA_CELL_CUBE overrides it.
static int cell_cube() {
    static const int c = [] {
        const char* e = std::getenv("MONA_CELL_CUBE");
        const int v = e ? std::atoi(e) : 2;
        return (v == 0 || v == 4 || v == 8) ? v : 2;
    }();
    return c;
}

// (slowest, middle, x) offsets of the 512 cells of a tile in that order
static std::vector<int3> tile_cell_order(int c) {
    std::vector<int3> o;
    o.reserve(512);
    if (c == 0) {
        for (int m = 0; m < 512; m++) {
            int a = 0, b = 0, x = 0;
            for (int l = 0; l < 3; l++) {
                a |= ((m >> (3 * l + 2)) & 1) << l;
                b |= ((m >> (3 * l + 1)) & 1) << l;
                x |= ((m >> (3 * l)) & 1) << l;
            }
            o.push_back(make_int3(a, b, x));
        }
        return o;
    }
    for (int a1 = 0; a1 < 8; a1 += c)
        for (int b1 = 0; b1 < 8; b1 += c)
            for (int x1 = 0; x1 < 8; x1 += c)
                for (int a = a1; a < a1 + c; a++)
                    for (int b = b1; b < b1 + c; b++)
                        for (int x = x1; x < x1 + c; x++) o.push_back(make_int3(a, b, x));
    return o;
}

// Morton code of tile coordinates (a slowest, then b, then x), 10 bits each
static int64_t morton3(int64_t a, int64_t b, int64_t x) {
    int64_t m = 0;
    for (int i = 0; i < 10; i++) {
        m |= ((x >> i) & 1) << (3 * i);
        m |= ((b >> i) & 1) << (3 * i + 1);
        m |= ((a >> i) & 1) << (3 * i + 2);
    }
    return m;
}

bool grid_intervals() {
    static const bool on = [] {
        const char* e = std::getenv("MONA_GRID_INTERVALS");
        return e ? std::atoi(e) != 0 : true;
    }();
    return on;
}

void Plan::set_frame(int f, const SparseDev* Gu, const SparseDev* Gut, hipStream_t s) {
    MONA_CHECK(f >= 0, MONA_ERR_ARG, "negative frame index");
    MONA_CHECK(!Gu->squeezed && Gu->l_nJump == Gu->l_size, MONA_ERR_ARG,
               "Gu must have one stored row per point (not squeezed)");
    MONA_CHECK((int64_t)Gu->h_r_jump.size() == Gu->nnz && (int64_t)Gut->h_r_jump.size() == Gut->nnz,
               MONA_ERR_STATE, "gridding matrices without their host encoding");
    // a frame with the same matrices shares their tile-order copies
    for (const Frame& o : frames)
        if (o.Gu == Gu && o.Gut == Gut && o.Gu_s) {
            Frame copy = o;
            if ((int)frames.size() <= f) frames.resize(f + 1);
            frames[f] = copy;
            Gu_b.reset();
            Gut_b.reset();
            koff_b.clear();
            return;
        }
    const int64_t P = Gu->l_size;
    // decode Gu's columns (running deltas, first absolute)
    std::vector<int64_t> rowptr(P + 1, 0);
    for (int64_t p = 0; p < P; p++) rowptr[p + 1] = rowptr[p] + Gu->h_r_nJump[p];
    std::vector<int> col(Gu->nnz);
    MONA_CHECK(Gu->r_size == Nu, MONA_ERR_ARG, "Gu must have one column per grid cell (r_size == prod(N_u))");
    {
        int64_t c = 0;
        for (int64_t k = 0; k < Gu->nnz; k++) {
            c += Gu->h_r_jump[k];
            MONA_CHECK(c >= 0 && c < Nu, MONA_ERR_ARG, "Gu column outside the grid");
            col[k] = (int)c;
        }
    }
    // tile key of each point: the tile of its first (lowest) cell, tiles
    // with y slowest, then z, then x (the storage order of the y-major 3-D
    // grids).  The point order is the same for every grid layout, so the
    // long rows' range sums, which follow it, are too.
    const bool yzg = grid_yz();
    const int tx = (N[0] + 7) / 8, tz = (N[2] + 7) / 8;
    const bool morton = tile_order() == 1;
    // order key of the tile of cell (x, y, z)
    auto tile_key = [&](int64_t x, int64_t y, int64_t z) -> int64_t {
        if (morton) return morton3(y >> 3, z >> 3, x >> 3);
        return ((y >> 3) * tz + (z >> 3)) * tx + (x >> 3);
    };
    std::vector<int64_t> key(P);
    // A Gu built from a trajectory of this plan's geometry and kernel grids
    // matrix-free: each point keeps its base cell, offsets and row
    // normaliser.
    TrajPoints tp;
    bool mf = false;
    if (mf_kernel_ok && Gu->traj) {
        const TrajInfo& T = *Gu->traj;
        mf = T.imDim == imDim && T.nPt == P && T.nWin == nWin && T.ktype == ktype_;
        for (int d = 0; d < imDim && mf; d++)
            mf = T.N[d] == N[d] && (double)(float)T.dK[d] == dK[d];
        for (int i = 0; i < T.nParam && mf; i++) mf = (double)(float)T.kp[i] == kparam[i];
        if (mf) traj_points(T, &tp);
    }
    // key: the tile of the point's lowest cell (its first column), then the
    // cell's raster position inside the tile (neighbouring points gather the
    // same cells); the order depends on Gu alone, whether or not the
    // trajectory came with it
    for (int64_t p = 0; p < P; p++) {
        if (rowptr[p + 1] == rowptr[p]) { key[p] = INT64_MAX; continue; }
        const int64_t cell = col[rowptr[p]];
        const int64_t x = cell % N[0], y = (cell / N[0]) % N[1], z = cell / ((int64_t)N[0] * N[1]);
        key[p] = tile_key(x, y, z) * 512 + ((y & 7) * 8 + (z & 7)) * 8 + (x & 7);
    }
    std::vector<int> perm(P), inv(P);
    for (int64_t p = 0; p < P; p++) perm[p] = (int)p;
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return key[a] < key[b]; });
    for (int64_t s2 = 0; s2 < P; s2++) inv[perm[s2]] = (int)s2;
    // Gu in tile order
    std::vector<int> scol(Gu->nnz), snj(P);
    std::vector<float> sval(Gu->nnz);
    {
        int64_t o = 0;
        for (int64_t s2 = 0; s2 < P; s2++) {
            int64_t p = perm[s2];
            int n = Gu->h_r_nJump[p];
            snj[s2] = n;
            for (int64_t k = rowptr[p]; k < rowptr[p + 1]; k++, o++) {
                scol[o] = col[k];
                sval[o] = Gu->h_m_val[k];
            }
        }
    }
    // cell index in the grid's storage order (y-major for the 3-D pruned
    // transforms)
    auto stored_cell = [&](int64_t c) {
        if (!yzg) return (int)c;
        const int64_t x = c % N[0], y = (c / N[0]) % N[1], z = c / ((int64_t)N[0] * N[1]);
        return (int)(x + (int64_t)N[0] * (z + (int64_t)N[2] * y));
    };
    if (yzg)
        for (auto& c : scol) c = stored_cell(c);
    auto Gu_s = std::make_shared<SparseDev>();
    sparse_from_decoded(*Gu_s, Gu->r_size, Gu->l_size, scol, sval, snj, nullptr, s);
    // Gut with relabelled point columns (entries of each cell keep their order)
    std::vector<int> tcol(Gut->nnz);
    {
        int64_t c = 0;
        for (int64_t k = 0; k < Gut->nnz; k++) {
            c += Gut->h_r_jump[k];
            MONA_CHECK(c >= 0 && c < P, MONA_ERR_ARG, "Gut column outside [0, nPt)");
            tcol[k] = inv[c];
        }
    }
    // Gut with its rows in 8^3-tile order (raster order inside a tile):
    // neighbouring rows gather the same points; each row keeps its entries
    // in order.  A frame that touches a large part of the grid stores one
    // row per cell (empty ones too): the product then writes the whole grid
    // and needs no zero fill.  A sparse frame stores only its touched cells
    // and the FFT passes along the last axis skip the others through cell
    // masks (skipping cells of a densely touched grid leaves partly
    // written cache lines all over it, which costs more than it saves).
    const int64_t G = Gut->l_size;
    MONA_CHECK(G == Nu, MONA_ERR_ARG, "Gut rows must be the grid cells");
    std::vector<int64_t> cell_row(G, -1), trp(Gut->l_nJump + 1, 0);
    {
        int64_t r = 0;
        for (int i = 0; i < Gut->l_nJump; i++) {
            r = Gut->squeezed ? r + Gut->h_l_jump[i] : i;
            MONA_CHECK(r >= 0 && r < G, MONA_ERR_ARG, "Gut row outside the grid");
            cell_row[r] = i;
            trp[i + 1] = trp[i] + Gut->h_r_nJump[i];
        }
    }
    int64_t touched = 0;
    for (int64_t c = 0; c < G; c++) touched += cell_row[c] >= 0 && trp[cell_row[c] + 1] > trp[cell_row[c]];
    const bool sparse_frame = touched < (G * 3) / 10;
    // row intervals of the last-axis pass columns (dense frames): rows
    // [lo, hi) of column c span every cell of c that Gu reads or Gut writes
    const int64_t NLa = N[imDim - 1], colsa = G / NLa;
    const bool use_ival = !sparse_frame && imDim >= 2 && grid_intervals();
    // sparse 3-D frames keep their cell masks and also carry the intervals:
    // the columns outside the y span of their x are then skipped by every
    // pass (dead columns), the rows inside stay masked
    const bool sparse_ival = sparse_frame && imDim == 3 && grid_intervals();
    std::vector<int2> ival;
    if (use_ival || sparse_ival) {
        ival.assign(colsa, make_int2((int)NLa, 0));
        auto widen = [&](int64_t cell) {
            const int64_t z = cell / colsa, c = cell - z * colsa;
            ival[c].x = std::min(ival[c].x, (int)z);
            ival[c].y = std::max(ival[c].y, (int)z + 1);
        };
        for (int64_t k = 0; k < Gu->nnz; k++) widen(col[k]);
        for (int64_t c = 0; c < G; c++)
            if (cell_row[c] >= 0 && trp[cell_row[c] + 1] > trp[cell_row[c]]) widen(c);
        for (auto& q : ival)
            if (q.y <= q.x) q = make_int2(0, 0);
    }
    // 3-D: the span over y of the non-empty columns of each x; columns
    // outside it are dead (-1, -1): no pass loads or stores them
    std::vector<int2> ivaly;
    if ((use_ival || sparse_ival) && imDim == 3) {
        std::vector<int2> span(N[0], make_int2(N[1], 0));
        for (int y = 0; y < N[1]; y++)
            for (int x = 0; x < N[0]; x++)
                if (ival[(int64_t)y * N[0] + x].y > 0) {
                    span[x].x = std::min(span[x].x, y);
                    span[x].y = std::max(span[x].y, y + 1);
                }
        for (auto& q : span)
            if (q.y <= q.x) q = make_int2(0, 0);
        for (int y = 0; y < N[1]; y++)
            for (int x = 0; x < N[0]; x++)
                if (y < span[x].x || y >= span[x].y) ival[(int64_t)y * N[0] + x] = make_int2(-1, -1);
        ivaly.resize((size_t)n[2] * N[0]);
        for (int z = 0; z < n[2]; z++)
            for (int x = 0; x < N[0]; x++) ivaly[(size_t)z * N[0] + x] = span[x];
    }
    auto inside = [&](int64_t cell) {
        const int64_t z = cell / colsa, c = cell - z * colsa;
        return z >= ival[c].x && z < ival[c].y;
    };
    std::vector<int> dcol(Gut->nnz), dnj(G), drow(G);
    std::vector<float> dval(Gut->nnz);
    {
        int64_t o = 0, q = 0;
        // slowest axis sa, then sb, then x (storage order of the grid)
        const int sa = yzg ? 1 : 2, sb = yzg ? 2 : 1;
        int c3[3] = {0, 0, 0};
        // the tiles in the order of tile_key
        std::vector<std::pair<int64_t, int3>> tiles;
        for (int a0 = 0; a0 < N[sa]; a0 += 8)
            for (int b0 = 0; b0 < N[sb]; b0 += 8)
                for (int x0 = 0; x0 < N[0]; x0 += 8) {
                    c3[0] = x0;
                    c3[sa] = a0;
                    c3[sb] = b0;
                    tiles.push_back({tile_key(c3[0], c3[1], c3[2]), make_int3(a0, b0, x0)});
                }
        std::stable_sort(tiles.begin(), tiles.end(),
                         [](const std::pair<int64_t, int3>& u, const std::pair<int64_t, int3>& v) {
                             return u.first < v.first;
                         });
        // the cells of a tile in the order of cell_cube()
        const std::vector<int3> in_tile = tile_cell_order(cell_cube());
        for (const auto& tl : tiles) {
            const int a0 = tl.second.x, b0 = tl.second.y, x0 = tl.second.z;
            for (const int3& d : in_tile) {
                const int a = a0 + d.x, b = b0 + d.y, x = x0 + d.z;
                if (a >= N[sa] || b >= N[sb] || x >= N[0]) continue;
                c3[0] = x;
                c3[sa] = a;
                c3[sb] = b;
                const int64_t cell = ((int64_t)c3[2] * N[1] + c3[1]) * N[0] + c3[0];
                const int64_t r = cell_row[cell];
                const int n = r < 0 ? 0 : (int)(trp[r + 1] - trp[r]);
                if (n == 0 && (sparse_frame || (use_ival && !inside(cell)))) continue;
                drow[q] = (int)cell;
                dnj[q] = n;
                for (int64_t k = n ? trp[r] : 0; k < (n ? trp[r + 1] : 0); k++, o++) {
                    dcol[o] = tcol[k];
                    dval[o] = Gut->h_m_val[k];
                }
                q++;
            }
        }
        dnj.resize(q);
        drow.resize(q);
    }
    // rows longer than kLongRow (the k-space centre): entries by ascending
    // point, so that a row's entries inside a range of consecutive points
    // are one run (the range form of the long rows)
    {
        int64_t o = 0;
        std::vector<std::pair<int, float>> ent;
        for (size_t r = 0; r < dnj.size(); o += dnj[r], r++) {
            const int nr = dnj[r];
            if (nr <= kLongRow) continue;
            ent.resize(nr);
            for (int k = 0; k < nr; k++) ent[k] = {dcol[o + k], dval[o + k]};
            std::sort(ent.begin(), ent.end(),
                      [](const std::pair<int, float>& a, const std::pair<int, float>& b) { return a.first < b.first; });
            for (int k = 0; k < nr; k++) {
                dcol[o + k] = ent[k].first;
                dval[o + k] = ent[k].second;
            }
        }
    }
    if (yzg)
        for (auto& c : drow) c = stored_cell(c);
    auto Gut_s = std::make_shared<SparseDev>();
    sparse_from_decoded(*Gut_s, Gut->r_size, Gut->l_size, dcol, dval, dnj, &drow, s);
    Gut_s->covers_all = !sparse_frame && !use_ival;
    build_groups(*Gut_s, dnj, dcol, nCh, s, grp_stage_kb, grp_entries, grp_rows);
    build_ranges(*Gut_s, dnj, dcol, nCh, s);
    // cell masks: cells the forward gridding reads (Gu columns) and cells the
    // adjoint gridding writes (Gut rows with entries); equal for a
    // transposed pair.  Layout of the pass along the last axis: for each
    // column (all other coordinates, x fastest) ceil(N_last / 32) words,
    // bit z of the column's cell at last-axis coordinate z.
    const int64_t NL = N[imDim - 1], cols = G / NL, mwords = (NL + 31) / 32;
    const int64_t nw = cols * mwords;
    std::vector<uint32_t> fmask(nw, 0u), amask(nw, 0u);
    auto set_bit = [&](std::vector<uint32_t>& m, int64_t cell) {
        const int64_t z = cell / cols, c = cell - z * cols;
        m[c * mwords + (z >> 5)] |= 1u << (z & 31);
    };
    for (int64_t k = 0; k < Gu->nnz; k++) set_bit(fmask, col[k]);
    for (int64_t c = 0; c < G; c++)
        if (cell_row[c] >= 0 && trp[cell_row[c] + 1] > trp[cell_row[c]]) set_bit(amask, c);
    auto dfm = std::make_shared<DevBuf>((size_t)nw * sizeof(uint32_t));
    MONA_HIP(hipMemcpyAsync(dfm->p, fmask.data(), (size_t)nw * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    std::shared_ptr<DevBuf> dam = dfm;
    if (amask != fmask) {
        dam = std::make_shared<DevBuf>((size_t)nw * sizeof(uint32_t));
        MONA_HIP(hipMemcpyAsync(dam->p, amask.data(), (size_t)nw * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    }
    std::shared_ptr<DevBuf> div, divy;
    if (use_ival || sparse_ival) {
        div = std::make_shared<DevBuf>((size_t)colsa * sizeof(int2));
        MONA_HIP(hipMemcpyAsync(div->p, ival.data(), (size_t)colsa * sizeof(int2), hipMemcpyHostToDevice, s));
    }
    if (!ivaly.empty()) {
        divy = std::make_shared<DevBuf>(ivaly.size() * sizeof(int2));
        MONA_HIP(hipMemcpyAsync(divy->p, ivaly.data(), ivaly.size() * sizeof(int2), hipMemcpyHostToDevice, s));
    }
    auto dperm = std::make_shared<DevBuf>((size_t)std::max<int64_t>(P, 1) * sizeof(int));
    MONA_HIP(hipMemcpyAsync(dperm->p, perm.data(), (size_t)P * sizeof(int), hipMemcpyHostToDevice, s));
    std::shared_ptr<MfPoints> dmf;
    std::vector<uint32_t> hb;
    std::vector<float4> hr;
    if (mf) {
        hb.assign(P, 0u);
        hr.assign(P, make_float4(0.f, 0.f, 0.f, 0.f));
        for (int64_t s2 = 0; s2 < P; s2++) {
            const int64_t p = perm[s2];
            if (rowptr[p + 1] == rowptr[p]) continue;   // deleted point: zeros
            const int* b = &tp.base[p * 3];
            const double* r = &tp.rest[p * 3];
            hb[s2] = (uint32_t)b[0] | ((uint32_t)b[1] << 10) | ((uint32_t)(imDim == 3 ? b[2] : 0) << 20);
            hr[s2] = make_float4((float)r[0], (float)r[1], imDim == 3 ? (float)r[2] : 0.f, (float)tp.inv_sum[p]);
        }
        dmf = std::make_shared<MfPoints>();
        dmf->P = (int)P;
        dmf->base.alloc((size_t)std::max<int64_t>(P, 1) * sizeof(uint32_t));
        dmf->rec.alloc((size_t)std::max<int64_t>(P, 1) * sizeof(float4));
        MONA_HIP(hipMemcpyAsync(dmf->base.p, hb.data(), (size_t)P * sizeof(uint32_t), hipMemcpyHostToDevice, s));
        MONA_HIP(hipMemcpyAsync(dmf->rec.p, hr.data(), (size_t)P * sizeof(float4), hipMemcpyHostToDevice, s));
    }
    MONA_HIP(hipStreamSynchronize(s));
    if ((int)frames.size() <= f) frames.resize(f + 1);
    Frame& fr = frames[f];
    fr.Gu = Gu;
    fr.Gut = Gut;
    fr.Gu_s = Gu_s;
    fr.Gut_s = Gut_s;
    fr.perm = dperm;
    fr.P = P;
    fr.fmask = dfm;
    fr.amask = dam;
    fr.masked = sparse_frame;
    fr.ival = div;
    fr.ivaly = divy;
    fr.ivals = use_ival;
    fr.mf = dmf;
    Gu_b.reset();  // the batched matrices are rebuilt on next use
    Gut_b.reset();
    koff_b.clear();
}

// The pruned operators confine the grid to the frame's cells: the forward
// pass along the last axis stores only the cells the gridding reads, the
// adjoint pass loads only the cells the adjoint gridding writes (the others
// are read as zero), so the untouched cells are never written, filled or
// read.  Needs the register kernel on the last axis.
bool Plan::limits_usable() const {
    return imDim >= 2 && fft_pass_fusable(N[0], nCh) && fft_plain_pass_is_reg(N[imDim - 1], fft_lines);
}

bool Plan::masked_grid(int f) const {
    const bool frames_sparse = f >= 0 ? frame(f).masked : [&] {
        for (const Frame& fr : frames)
            if (!fr.masked) return false;
        return !frames.empty();
    }();
    return frames_sparse && limits_usable();
}

bool Plan::ival_grid(int f) const {
    const bool all = f >= 0 ? frame(f).ivals : [&] {
        for (const Frame& fr : frames)
            if (!fr.ivals) return false;
        return !frames.empty();
    }();
    return all && limits_usable();
}

const uint32_t* Plan::fwd_mask(int f) const {
    if (f < 0) return masked_grid(-1) && mask_fb.p ? mask_fb.as<uint32_t>() : nullptr;
    return masked_grid(f) ? frame(f).fmask->as<uint32_t>() : nullptr;
}
const uint32_t* Plan::adj_mask(int f) const {
    if (f < 0) return masked_grid(-1) && mask_ab.p ? mask_ab.as<uint32_t>() : nullptr;
    return masked_grid(f) ? frame(f).amask->as<uint32_t>() : nullptr;
}
// Intervals of the passes: dense frames (ival_grid), and sparse 3-D frames
// that carry them beside their masks (dead columns only).
const int2* Plan::ivals_of(int f) const {
    if (f < 0) return (ival_grid(-1) || masked_grid(-1)) && ival_b.p ? ival_b.as<int2>() : nullptr;
    return (ival_grid(f) || masked_grid(f)) && frame(f).ival ? frame(f).ival->as<int2>() : nullptr;
}
// (with ivals_of: the two go together, dead columns rely on both)
const int2* Plan::mid_ivals_of(int f) const {
    if (f < 0) return (ival_grid(-1) || masked_grid(-1)) && ival_b.p && ivaly_b.p ? ivaly_b.as<int2>() : nullptr;
    return (ival_grid(f) || masked_grid(f)) && frame(f).ival && frame(f).ivaly ? frame(f).ivaly->as<int2>()
                                                                               : nullptr;
}

void Plan::forward(int f, const float2* x, float2* y, hipStream_t s) {
    MONA_CHECK(coils_set, MONA_ERR_STATE, "coil maps not set");
    const Frame& fr = frame(f);
    ensure_grid();
    float2* gr = grid.as<float2>();
    {
        TimedRegion t(2, s);
        if (fft_pass_fusable(N[0], nCh)) {
            ensure_fft_scratch();
            forward_fft_pruned(x, gr, 1, fft_a, fft_b, s, false, fwd_mask(f), ivals_of(f), mid_ivals_of(f));
        } else {
            Geom g = geom();
            prologue_kernel<<<nb256(Nu * nCh), 256, 0, s>>>(x, KFC.as<float2>(), gr, g);
            launch_check();
            cfft_nd(gr, imDim, N, nCh, false, s);
        }
    }
    {
        TimedRegion t(0, s);
        if (frame_mf(f))
            u2n_mf(*fr.mf, nullptr, fr.mf->P, gr, false, Nu, y, false, nCh, imDim, mf_geo(), mfk, s);
        else
            spmv_c2_rowmajor(*fr.Gu_s, gr, y, nCh, nullptr, s);
    }
}

// Scratch for the pruned passes: the largest intermediate is
// prod(N_d, d <= a) * prod(n_d, d > a) * nCh over the axes a < imDim - 1.
int64_t Plan::fft_scratch_elems() const {
    int64_t mx = 1;
    for (int a = 0; a + 1 < imDim; a++) {
        int64_t v = nCh;
        for (int d = 0; d < imDim; d++) v *= (d <= a) ? N[d] : n[d];
        mx = std::max(mx, v);
    }
    if (grid_yz()) mx = std::max(mx, (int64_t)N[1] * mid_rs());
    return mx;
}

// Row stride along y of the y-major middle intermediate [N_y][n_z][N_x nCh]:
// its n_z N_x nCh values plus a pad (MONA_FFT_MID_PAD complex values,
// default 32 = 256 bytes), so that the y pass's rows are not an exact
// multiple of a large power of two apart.
int64_t Plan::mid_rs() const {
    static const int64_t pad = [] {
        const char* e = std::getenv("MONA_FFT_MID_PAD");
        return e ? std::max<int64_t>(0, std::atoll(e)) : (int64_t)32;
    }();
    return (int64_t)n[2] * N[0] * nCh + pad;
}

void Plan::ensure_fft_scratch() {
    int64_t mx = fft_scratch_elems();
    fft_a.ensure((size_t)mx * sizeof(float2));
    fft_b.ensure((size_t)mx * sizeof(float2));
}

// Forward: x .* KFC zero-filled to N_u and transformed, one pass per axis;
// pass a reads the compact [nb][prod(n_d, d>a)][n_a][S_a] array, S_a =
// nCh*prod(N_d, d<a), and writes [nb][prod(n_d, d>a)][N_a][S_a].
void Plan::forward_fft_pruned(const float2* x, float2* gr, int nb, DevBuf& ba, DevBuf& bb, hipStream_t s,
                              bool bf, const uint32_t* out_mask, const int2* out_ival, const int2* mid_ival) {
    const float2* src = nullptr;
    float2* bufs[2] = {ba.as<float2>(), bb.as<float2>()};
    const bool yz = grid_yz();
    int64_t S = nCh;
    for (int a = 0; a < imDim; a++) {
        int64_t outer = nb;
        for (int d = a + 1; d < imDim; d++) outer *= n[d];
        if (yz && a == 2) {
            S /= N[1];       // lines (frame, y, x, c): outer frame * N_y + y, inner x c
            outer = (int64_t)nb * N[1];
        }
        AxisPass p;
        p.tile = fft_tile;
        p.lines = fft_lines;
        p.kfc_period = nu;
        p.N = N[a];
        p.S = S;
        p.outer = outer;
        p.in_off = N[a] / 2 - n[a] / 2;
        p.in_cnt = n[a];
        p.in = src;
        p.in_os = (int64_t)n[a] * S;
        p.out = (a == imDim - 1) ? gr : bufs[a & 1];
        p.out_os = (int64_t)N[a] * S;
        p.out_off = 0;
        p.out_cnt = N[a];
        p.inverse = false;
        // bf16 storage for the grid only: the pruned intermediates stay
        // float2 (the passes are not bandwidth-bound; bf16 intermediates
        // made them slower)
        p.bf_in = false;
        p.bf_out = bf && a == imDim - 1;
        if (a == imDim - 1 && (out_mask || out_ival)) {
            p.out_mask = out_mask;
            p.out_ival = out_ival;
            p.mask_nch = nCh;
        }
        if (a == 1 && imDim == 3 && mid_ival) {
            p.out_ival = mid_ival;
            p.mask_nch = nCh;
        }
        if (yz && a == 1) {  // out per frame [N_y][n_z][S] (y rows mid_rs apart); line o = frame * n_z + z
            p.out_on = n[2];
            p.out_os2 = (int64_t)N[1] * mid_rs();
            p.out_os = S;
            p.out_rs = mid_rs();
        }
        if (yz && a == 2) {  // in per frame [N_y][n_z][S], out the grid [N_y][N_z][S]
            p.in_os = mid_rs();
            p.in_rs = S;
            p.out_os = (int64_t)N[2] * S;
            p.out_rs = S;
        }
        if (a == 0) {
            p.mode = 1;
            p.img = x;
            p.kfc = KFC.as<float2>();
        }
        fft_pass(p, s);
        src = p.out;
        S *= N[a];
    }
}

// Adjoint: the grid transformed back one axis at a time from the last,
// each pass keeping only the central n_a rows; the pass along x ends with
// .* KFC_conj (and the coil sum when combine).
void Plan::adjoint_fft_pruned(const float2* gr, float2* x, bool combine, float scale, int nb, DevBuf& ba,
                              DevBuf& bb, hipStream_t s, bool bf, const uint32_t* in_mask, const int2* in_ival,
                              const int2* mid_ival) {
    const float2* src = gr;
    float2* bufs[2] = {ba.as<float2>(), bb.as<float2>()};
    const bool yz = grid_yz();
    for (int a = imDim - 1; a >= 0; a--) {
        int64_t S = nCh, outer = nb;
        for (int d = 0; d < a; d++) S *= N[d];
        for (int d = a + 1; d < imDim; d++) outer *= n[d];
        if (yz && a == 2) {
            S /= N[1];
            outer = (int64_t)nb * N[1];
        }
        AxisPass p;
        p.tile = fft_tile;
        p.lines = fft_lines;
        p.kfc_period = nu;
        p.N = N[a];
        p.S = S;
        p.outer = outer;
        p.in = src;
        p.in_os = (int64_t)N[a] * S;
        p.in_off = 0;
        p.in_cnt = N[a];
        p.out_off = N[a] / 2 - n[a] / 2;
        p.out_cnt = n[a];
        p.out_os = (int64_t)n[a] * S;
        p.inverse = true;
        p.scale = 1.0f / (float)N[a];
        p.bf_in = bf && a == imDim - 1;
        p.bf_out = false;
        if (a == imDim - 1 && (in_mask || in_ival)) {
            p.in_mask = in_mask;
            p.in_ival = in_ival;
            p.mask_nch = nCh;
        }
        if (a == 1 && imDim == 3 && mid_ival) {
            p.in_ival = mid_ival;
            p.mask_nch = nCh;
        }
        if (yz && a == 2) {  // in the grid [N_y][N_z][S], out per frame [N_y][n_z][S] (y rows mid_rs apart)
            p.in_os = (int64_t)N[2] * S;
            p.in_rs = S;
            p.out_os = mid_rs();
            p.out_rs = S;
        }
        if (yz && a == 1) {  // in per frame [N_y][n_z][S]; line o = frame * n_z + z
            p.in_on = n[2];
            p.in_os2 = (int64_t)N[1] * mid_rs();
            p.in_os = S;
            p.in_rs = mid_rs();
        }
        if (a == 0) {
            p.mode = combine ? 2 : 3;
            p.kfc = KFCc.as<float2>();
            p.epi_scale = scale;
            p.img_out = x;
            p.out = x;
        } else {
            p.out = bufs[a & 1];
        }
        fft_pass(p, s);
        src = p.out;
    }
}

void Plan::adjoint(int f, const float2* y, float2* x, bool combine, const float* hy, float scale,
                   hipStream_t s) {
    MONA_CHECK(coils_set, MONA_ERR_STATE, "coil maps not set");
    const Frame& fr = frame(f);
    ensure_grid();
    Geom g = geom();
    float2* gr = grid.as<float2>();
    const bool limited = masked_grid(f) || ival_grid(f);
    {
        TimedRegion t(1, s);
        spmv_c2_rowmajor(*fr.Gut_s, y, gr, nCh, hy, s, nullptr, !limited, use_staged);
    }
    {
        TimedRegion t(2, s);
        if (fft_pass_fusable(N[0], nCh)) {
            ensure_fft_scratch();
            adjoint_fft_pruned(gr, x, combine, scale, 1, fft_a, fft_b, s, false, adj_mask(f), ivals_of(f),
                               mid_ivals_of(f));
        } else {
            cfft_nd(gr, imDim, N, nCh, true, s);
            epilogue_kernel<<<nb256(nu), 256, 0, s>>>(gr, KFCc.as<float2>(), x, g, combine ? 1 : 0, scale);
            launch_check();
        }
    }
}

void Plan::to_tile_order(int f, const float2* src, float2* dst, hipStream_t s) {
    const Frame& fr = frame(f);
    k_gather_pts<<<nb256(fr.P * nCh), 256, 0, s>>>(src, fr.perm->as<int>(), dst, fr.P, nCh);
    launch_check();
}

void Plan::weights_to_tile_order(int f, const float* src, float* dst, hipStream_t s) {
    const Frame& fr = frame(f);
    k_gather_w<<<nb256(fr.P), 256, 0, s>>>(src, fr.perm->as<int>(), dst, fr.P);
    launch_check();
}

void Plan::forward_public(int f, const float2* x, float2* y, hipStream_t s) {
    const Frame& fr = frame(f);
    work.kpub.ensure((size_t)fr.P * nCh * sizeof(float2));
    forward(f, x, work.kpub.as<float2>(), s);
    k_scatter_pts<<<nb256(fr.P * nCh), 256, 0, s>>>(work.kpub.as<float2>(), fr.perm->as<int>(), y, fr.P, nCh);
    launch_check();
}

void Plan::adjoint_public(int f, const float2* y, float2* x, bool combine, hipStream_t s) {
    const Frame& fr = frame(f);
    work.kpub.ensure((size_t)fr.P * nCh * sizeof(float2));
    to_tile_order(f, y, work.kpub.as<float2>(), s);
    adjoint(f, work.kpub.as<float2>(), x, combine, nullptr, 1.0f, s);
}

// ---------------------------------------------------------------- batched form
// Block-diagonal stacking of every frame's tile-order matrices, on the
// device: frame f's grid occupies columns [f*Nu, (f+1)*Nu) of Gu_b and
// output rows of Gut_b, its points follow those of frames 0..f-1.  The
// views of the leading frames let a batch use only frames 0..nA-1.
void Plan::build_batch(hipStream_t s) {
    const int nFr = (int)frames.size();
    MONA_CHECK(nFr >= 1, MONA_ERR_STATE, "the plan has no frames");
    MONA_CHECK((int64_t)nFr * Nu < 0x7fffffff, MONA_ERR_ARG, "nFrames * prod(N_u) too large for one launch");
    std::vector<const SparseDev*> us, ts;
    std::vector<int64_t> ucol, urow, tcol, trow;
    koff_b.assign(nFr + 1, 0);
    int64_t poff = 0;
    for (int f = 0; f < nFr; f++) {
        const Frame& fr = frame(f);
        us.push_back(fr.Gu_s.get());
        ts.push_back(fr.Gut_s.get());
        ucol.push_back((int64_t)f * Nu);
        urow.push_back(poff);
        tcol.push_back(poff);
        trow.push_back((int64_t)f * Nu);
        poff += fr.P;
        koff_b[f + 1] = koff_b[f] + fr.P * nCh;
    }
    // stacked cell masks of the frames (frame f's columns after those of
    // frames 0..f-1)
    mask_fb.release();
    mask_ab.release();
    {
        const size_t wb = frame(0).fmask->bytes;
        mask_fb.alloc(wb * nFr);
        mask_ab.alloc(wb * nFr);
        for (int f = 0; f < nFr; f++) {
            MONA_HIP(hipMemcpyAsync(static_cast<char*>(mask_fb.p) + f * wb, frame(f).fmask->p, wb,
                                    hipMemcpyDeviceToDevice, s));
            MONA_HIP(hipMemcpyAsync(static_cast<char*>(mask_ab.p) + f * wb, frame(f).amask->p, wb,
                                    hipMemcpyDeviceToDevice, s));
        }
    }
    // stacked row intervals (when every frame has them)
    ival_b.release();
    ivaly_b.release();
    {
        bool all = true;
        for (int f = 0; f < nFr; f++) all = all && frame(f).ival && (frame(f).ivaly != nullptr) == (frame(0).ivaly != nullptr);
        if (all) {
            const size_t ib = frame(0).ival->bytes;
            ival_b.alloc(ib * nFr);
            for (int f = 0; f < nFr; f++)
                MONA_HIP(hipMemcpyAsync(static_cast<char*>(ival_b.p) + f * ib, frame(f).ival->p, ib,
                                        hipMemcpyDeviceToDevice, s));
            if (frame(0).ivaly) {
                const size_t yb = frame(0).ivaly->bytes;
                ivaly_b.alloc(yb * nFr);
                for (int f = 0; f < nFr; f++)
                    MONA_HIP(hipMemcpyAsync(static_cast<char*>(ivaly_b.p) + f * yb, frame(f).ivaly->p, yb,
                                            hipMemcpyDeviceToDevice, s));
            }
        }
    }
    // matrix-free points of the frames back to back, with their frame
    mf_b.P = 0;
    mf_b.base.release();
    mf_b.rec.release();
    mf_b_frame.release();
    bool all_mf = nFr <= 65535;
    for (int f = 0; f < nFr; f++) all_mf = all_mf && frame(f).mf;
    if (all_mf) {
        mf_b.base.alloc((size_t)std::max<int64_t>(poff, 1) * sizeof(uint32_t));
        mf_b.rec.alloc((size_t)std::max<int64_t>(poff, 1) * sizeof(float4));
        mf_b_frame.alloc((size_t)std::max<int64_t>(poff, 1) * sizeof(uint16_t));
        std::vector<uint16_t> fo((size_t)poff);
        int64_t o = 0;
        for (int f = 0; f < nFr; f++) {
            const MfPoints& m = *frame(f).mf;
            MONA_HIP(hipMemcpyAsync(mf_b.base.as<uint32_t>() + o, m.base.p, (size_t)m.P * sizeof(uint32_t),
                                    hipMemcpyDeviceToDevice, s));
            MONA_HIP(hipMemcpyAsync(mf_b.rec.as<float4>() + o, m.rec.p, (size_t)m.P * sizeof(float4),
                                    hipMemcpyDeviceToDevice, s));
            std::fill(fo.begin() + o, fo.begin() + o + m.P, (uint16_t)f);
            o += m.P;
        }
        MONA_HIP(hipMemcpyAsync(mf_b_frame.p, fo.data(), (size_t)poff * sizeof(uint16_t), hipMemcpyHostToDevice, s));
        MONA_HIP(hipStreamSynchronize(s));
        mf_b.P = (int)poff;
    }
    auto gu = std::make_shared<SparseDev>();
    sparse_block_diag(*gu, us, ucol, urow, (int)(nFr * Nu), (int)poff, &view_u, s);
    auto gut = std::make_shared<SparseDev>();
    sparse_block_diag(*gut, ts, tcol, trow, (int)poff, (int)(nFr * Nu), &view_t, s);
    Gu_b = gu;
    Gut_b = gut;
}

void Plan::forward_batch(const float2* x, float2* y, hipStream_t s, int nA) {
    MONA_CHECK(coils_set && Gu_b, MONA_ERR_STATE, "batched operators not built");
    const int nFr = (int)frames.size();
    if (nA <= 0) nA = nFr;
    MONA_CHECK(nA <= nFr, MONA_ERR_ARG, "more batch frames than the plan has");
    MONA_CHECK(fft_pass_fusable(N[0], nCh), MONA_ERR_ARG, "batched form needs N_x * nCh <= 8192");
    grid_b.ensure((size_t)nFr * Nu * nCh * sizeof(float2));
    fft_ab.ensure((size_t)nFr * fft_scratch_elems() * sizeof(float2));
    fft_bb.ensure((size_t)nFr * fft_scratch_elems() * sizeof(float2));
    {
        TimedRegion t(2, s);
        forward_fft_pruned(x, grid_b.as<float2>(), nA, fft_ab, fft_bb, s, false, fwd_mask(-1), ivals_of(-1),
                           mid_ivals_of(-1));
    }
    {
        TimedRegion t(0, s);
        if (batch_mf(nA))
            u2n_mf(mf_b, mf_b_frame.as<uint16_t>(), (int)(koff_b[nA] / nCh), grid_b.as<float2>(), false, Nu, y,
                   false, nCh, imDim, mf_geo(), mfk, s);
        else
            spmv_c2_rowmajor(*Gu_b, grid_b.as<float2>(), y, nCh, nullptr, s, &view_u[nA]);
    }
}

void Plan::adjoint_batch(const float2* y, float2* x, const float* hy, float scale, hipStream_t s, int nA) {
    MONA_CHECK(coils_set && Gut_b, MONA_ERR_STATE, "batched operators not built");
    const int nFr = (int)frames.size();
    if (nA <= 0) nA = nFr;
    MONA_CHECK(nA <= nFr, MONA_ERR_ARG, "more batch frames than the plan has");
    MONA_CHECK(fft_pass_fusable(N[0], nCh), MONA_ERR_ARG, "batched form needs N_x * nCh <= 8192");
    grid_b.ensure((size_t)nFr * Nu * nCh * sizeof(float2));
    fft_ab.ensure((size_t)nFr * fft_scratch_elems() * sizeof(float2));
    fft_bb.ensure((size_t)nFr * fft_scratch_elems() * sizeof(float2));
    const bool limited = fwd_mask(-1) || ivals_of(-1);
    {
        TimedRegion t(1, s);
        spmv_c2_rowmajor(*Gut_b, y, grid_b.as<float2>(), nCh, hy, s, &view_t[nA], !limited, use_staged);
    }
    {
        TimedRegion t(2, s);
        adjoint_fft_pruned(grid_b.as<float2>(), x, true, scale, nA, fft_ab, fft_bb, s, false, adj_mask(-1),
                           ivals_of(-1), mid_ivals_of(-1));
    }
}

// ---------------------------------------------------------------- vector helpers
void vsub(const float2* a, const float2* b, float2* out, int64_t n, hipStream_t s) {
    if (al16(a, b, out)) k_sub2<<<nb256((n + 1) / 2), 256, 0, s>>>(a, b, out, n);
    else k_sub<<<nb256(n), 256, 0, s>>>(a, b, out, n);
    MONA_HIP(hipGetLastError());
}
void vcopy(const float2* a, float2* out, int64_t n, hipStream_t s) {
    if (al16(a, out)) k_copy2<<<nb256((n + 1) / 2), 256, 0, s>>>(a, out, n);
    else k_copy<<<nb256(n), 256, 0, s>>>(a, out, n);
    MONA_HIP(hipGetLastError());
}
void vaxpy(float2* y, const float2* x, const double* a, float sign, int64_t n, hipStream_t s) {
    if (al16(y, x)) k_axpy_dev2<<<nb256((n + 1) / 2), 256, 0, s>>>(y, x, a, sign, n);
    else k_axpy_dev<<<nb256(n), 256, 0, s>>>(y, x, a, sign, n);
    MONA_HIP(hipGetLastError());
}
void vxpby(float2* p, const float2* d, const double* b, int64_t n, hipStream_t s) {
    if (al16(p, d)) k_xpby_dev2<<<nb256((n + 1) / 2), 256, 0, s>>>(p, d, b, n);
    else k_xpby_dev<<<nb256(n), 256, 0, s>>>(p, d, b, n);
    MONA_HIP(hipGetLastError());
}
void sqnorm(const float2* a, int64_t n, const float* w, int nvec, double cw, double* out, bool acc,
            double* partial, hipStream_t s) {
    int nb = (int)std::min<int64_t>(kRedBlocks, std::max<int64_t>(1, (n + kRedThreads - 1) / kRedThreads));
    MONA_CHECK(!w || n < 2147483647LL, MONA_ERR_ARG, "weighted norm over 2^31 or more elements");
    if (w) k_sqnorm_partial<true><<<nb, kRedThreads, 0, s>>>(a, n, w, nvec, cw, partial);
    else k_sqnorm_partial<false><<<nb, kRedThreads, 0, s>>>(a, n, w, nvec, cw, partial);
    MONA_HIP(hipGetLastError());
    k_reduce_final<<<1, kRedThreads, 0, s>>>(partial, nb, out, acc ? 1 : 0);
    MONA_HIP(hipGetLastError());
}
void compute_hy(const float* ve, float vmax, float* hy, int64_t n, float* scratch, hipStream_t s) {
    if (vmax > 0) {
        k_hy<<<nb256(n), 256, 0, s>>>(ve, vmax, hy, n);
    } else {
        // bmSensa.m:29-30: no bound given, the maximum of ve (on the device,
        // so the solve keeps running ahead of the host)
        const int nb = (int)std::min<int64_t>(kRedBlocks, std::max<int64_t>(1, (n + kRedThreads - 1) / kRedThreads));
        k_max_partial<<<nb, kRedThreads, 0, s>>>(ve, n, scratch);
        MONA_HIP(hipGetLastError());
        k_max_final<<<1, kRedThreads, 0, s>>>(scratch, nb, scratch + kRedBlocks);
        MONA_HIP(hipGetLastError());
        k_hy_dev<<<nb256(n), 256, 0, s>>>(ve, scratch + kRedBlocks, hy, n);
    }
    MONA_HIP(hipGetLastError());
}
void cg_alpha(CgScalars* sc, hipStream_t s, double eps) {
    k_cg_alpha<<<1, 1, 0, s>>>(sc, eps);
    MONA_HIP(hipGetLastError());
}
void cg_beta(CgScalars* sc, hipStream_t s) {
    k_cg_beta<<<1, 1, 0, s>>>(sc);
    MONA_HIP(hipGetLastError());
}
void cg_reset(CgScalars* sc, hipStream_t s) {
    k_cg_reset<<<1, 1, 0, s>>>(sc);
    MONA_HIP(hipGetLastError());
}
void dstore(double* dst, const double* src, hipStream_t s) {
    k_store<<<1, 1, 0, s>>>(dst, src);
    MONA_HIP(hipGetLastError());
}

// max of a float array on the host side (ve_max default = max(ve))
// ================================================================ witness
int64_t Plan::witness_elems() const { return imDim == 3 ? (int64_t)n[0] * n[1] : nu; }

void Plan::witness(int c, const float2* x, int64_t x_elems, hipStream_t s) {
    if (wit_slices && c < wit_slots) {
        // the image is x-fastest: the middle z slice (z = n_z / 2, the
        // reference's fix(n_u(end)/2) + 1) is one contiguous block
        const int64_t off = imDim == 3 ? (int64_t)(n[2] / 2) * n[0] * n[1] : 0;
        vcopy(x + off, wit_slices + (int64_t)c * witness_elems(), witness_elems(), s);
    }
    if (wit_full)
        for (size_t k = 0; k < wit_full_ind.size(); k++)
            if (wit_full_ind[k] == c + 1) vcopy(x, wit_full + (int64_t)k * x_elems, x_elems, s);
}

// ================================================================ CG-SENSE
void Plan::sensa(int f, float2* x, const float2* y, const float* ve, float ve_max, int nIter, int nCGD,
                 double* residual, hipStream_t s) {
    MONA_CHECK(nIter >= 0 && nCGD >= 0, MONA_ERR_ARG, "nIter and nCGD must be >= 0");
    const Frame& fr = frame(f);
    int64_t P = fr.Gu->l_size;
    int64_t ny = P * nCh;
    MONA_CHECK(fr.Gut->r_size == P, MONA_ERR_ARG, "Gut does not match Gu");
    Work& w = work;
    w.hy.ensure((size_t)P * sizeof(float));
    w.k1.ensure((size_t)ny * sizeof(float2));
    w.k2.ensure((size_t)ny * sizeof(float2));
    w.i1.ensure((size_t)nu * sizeof(float2));
    w.i2.ensure((size_t)nu * sizeof(float2));
    w.sc.ensure(sizeof(CgScalars));
    w.partial.ensure(kRedBlocks * sizeof(double));
    w.hist.ensure((size_t)std::max(nIter, 1) * sizeof(double));
    float* hy = w.hy.as<float>();
    float2* res = w.k1.as<float2>();
    float2* Mp = w.k2.as<float2>();
    float2* dag = w.i1.as<float2>();
    float2* p = w.i2.as<float2>();
    CgScalars* sc = w.sc.as<CgScalars>();
    double* part = w.partial.as<double>();
    // k-space data and weights in the frame's tile order for the whole solve
    w.hy0.ensure((size_t)P * sizeof(float));
    w.ys.ensure((size_t)ny * sizeof(float2));
    compute_hy(ve, ve_max, w.hy0.as<float>(), P, hy_scratch(w), s);
    weights_to_tile_order(f, w.hy0.as<float>(), hy, s);
    to_tile_order(f, y, w.ys.as<float2>(), s);
    y = w.ys.as<float2>();
    const float invHX = (float)(1.0 / HX);

    for (int it = 0; it < nIter; it++) {
        cg_reset(sc, s);
        forward(f, x, Mp, s);
        vsub(y, Mp, res, ny, s);                             // res = y - M x
        adjoint(f, res, dag, true, hy, invHX, s);            // dag = (1/HX) M^H (HY .* res)
        sqnorm(dag, nu, nullptr, 1, HX, &sc->sqn, false, part, s);
        vcopy(dag, p, nu, s);
        for (int i = 1; i <= nCGD; i++) {
            forward(f, p, Mp, s);                            // Mp = M p
            sqnorm(Mp, ny, hy, nCh, 1.0, &sc->den, false, part, s);
            cg_alpha(sc, s);                                 // done check, a
            vaxpy(x, p, &sc->a, 1.0f, nu, s);                // x += a p
            if (i == nCGD) break;
            vaxpy(res, Mp, &sc->a, -1.0f, ny, s);            // res -= a Mp
            adjoint(f, res, dag, true, hy, invHX, s);
            sqnorm(dag, nu, nullptr, 1, HX, &sc->sqn_new, false, part, s);
            cg_beta(sc, s);
            vxpby(p, dag, &sc->b, nu, s);                    // p = dag + b p
        }
        // monitoring residual <r, HY r>, r = y - M x   (bmSensa.m:82-85)
        forward(f, x, Mp, s);
        vsub(y, Mp, res, ny, s);
        sqnorm(res, ny, hy, nCh, 1.0, w.hist.as<double>() + it, false, part, s);
        witness(it, x, nu, s);
    }
    if (residual && nIter > 0) {
        MONA_HIP(hipMemcpyAsync(residual, w.hist.p, nIter * sizeof(double), hipMemcpyDeviceToHost, s));
        MONA_HIP(hipStreamSynchronize(s));
    }
}

// ================================================================ memory
namespace {
size_t sparse_bytes(const SparseDev& S) {
    return S.col.bytes + S.val.bytes + S.rowptr.bytes + S.rowidx.bytes + S.cv.bytes + S.long_row.bytes +
           S.long_first.bytes + S.chunk_begin.bytes + S.partial.bytes + S.group.bytes + S.group_col.bytes + S.group_e.bytes +
           S.slot.bytes + S.range.bytes + S.seg.bytes + S.range_first.bytes;
}
}  // namespace

size_t Plan::device_bytes() const {
    size_t b = K.bytes + KFC.bytes + KFCc.bytes + grid.bytes + fft_a.bytes + fft_b.bytes + grid_b.bytes +
               fft_ab.bytes + fft_bb.bytes + mask_fb.bytes + mask_ab.bytes + ival_b.bytes + ivaly_b.bytes;
    const Work& w = work;
    for (const DevBuf* d : {&w.hy, &w.k1, &w.k2, &w.i1, &w.i2, &w.i3, &w.i4, &w.sc, &w.partial, &w.hist, &w.ys,
                            &w.hy0, &w.kpub, &w.z1, &w.z2, &w.z3, &w.z4, &w.z5})
        b += d->bytes;
    std::vector<const void*> seen;
    auto once = [&](const void* q) {
        if (!q || std::find(seen.begin(), seen.end(), q) != seen.end()) return false;
        seen.push_back(q);
        return true;
    };
    for (const Frame& fr : frames) {
        if (once(fr.Gu_s.get())) b += sparse_bytes(*fr.Gu_s);
        if (once(fr.Gut_s.get())) b += sparse_bytes(*fr.Gut_s);
        if (once(fr.perm.get())) b += fr.perm->bytes;
        if (once(fr.fmask.get())) b += fr.fmask->bytes;
        if (once(fr.amask.get())) b += fr.amask->bytes;
        if (fr.mf && once(fr.mf.get())) b += fr.mf->base.bytes + fr.mf->rec.bytes;
    }
    b += mf_b.base.bytes + mf_b.rec.bytes + mf_b_frame.bytes;
    if (Gu_b) b += sparse_bytes(*Gu_b);
    if (Gut_b) b += sparse_bytes(*Gut_b);
    return b;
}

// ================================================================ batched CG-SENSE
// bmSensa over every frame of the plan at once: frame f solves its own
// system (own data, weights and CG scalars); the operators run as one
// block-diagonal gridding launch and frame-batched FFT passes over all
// frames, the reductions and vector updates frame by frame with the same
// launches as sensa(), so each frame's result equals sensa() on that frame
// alone.  x [nFr][nu]; y / ve packed per frame at y_off / ve_off (caller
// point order); residual [nFr][nIter].
void Plan::sensa_batch(int nA, float2* x, const float2* y, const long long* y_off, const float* ve,
                       const long long* ve_off, const float* ve_max_in, int nIter, int nCGD, double* residual,
                       hipStream_t s) {
    MONA_CHECK(nIter >= 0 && nCGD >= 0, MONA_ERR_ARG, "nIter and nCGD must be >= 0");
    MONA_CHECK(!frames.empty(), MONA_ERR_STATE, "the plan has no frames");
    if (nA <= 0) nA = (int)frames.size();
    MONA_CHECK(nA <= (int)frames.size(), MONA_ERR_ARG, "more batch frames than the plan has");
    if (!Gu_b || koff_b.size() != frames.size() + 1) build_batch(s);
    const int nFr = nA;  // the leading frames solved in this call
    std::vector<int64_t> Pf(nFr), koff(nFr + 1, 0);
    for (int f = 0; f < nFr; f++) {
        Pf[f] = frame(f).Gu->l_size;
        koff[f + 1] = koff[f] + Pf[f] * nCh;
    }
    const int64_t ny = koff[nFr], nx = (int64_t)nFr * nu;
    Work& w = work;
    w.hy.ensure((size_t)(ny / nCh) * sizeof(float));
    w.hy0.ensure((size_t)(ny / nCh) * sizeof(float));
    w.ys.ensure((size_t)ny * sizeof(float2));
    w.k1.ensure((size_t)ny * sizeof(float2));
    w.k2.ensure((size_t)ny * sizeof(float2));
    w.i1.ensure((size_t)nx * sizeof(float2));
    w.i2.ensure((size_t)nx * sizeof(float2));
    w.sc.ensure((size_t)nFr * sizeof(CgScalars));
    w.partial.ensure(kRedBlocks * sizeof(double));
    w.hist.ensure((size_t)std::max(nIter, 1) * nFr * sizeof(double));
    float* hy = w.hy.as<float>();
    float2* ys = w.ys.as<float2>();
    float2* res = w.k1.as<float2>();
    float2* Mp = w.k2.as<float2>();
    float2* dag = w.i1.as<float2>();
    float2* p = w.i2.as<float2>();
    CgScalars* sc = w.sc.as<CgScalars>();
    double* part = w.partial.as<double>();
    double* hist = w.hist.as<double>();
    for (int f = 0; f < nFr; f++) {
        float vmax = ve_max_in ? ve_max_in[f] : 0.0f;
        compute_hy(ve + ve_off[f], vmax, w.hy0.as<float>() + koff[f] / nCh, Pf[f], hy_scratch(w), s);
        weights_to_tile_order(f, w.hy0.as<float>() + koff[f] / nCh, hy + koff[f] / nCh, s);
        to_tile_order(f, y + y_off[f], ys + koff[f], s);
    }
    const float invHX = (float)(1.0 / HX);
    auto K = [&](float2* a, int f) { return a + koff[f]; };
    auto X = [&](float2* a, int f) { return a + (int64_t)f * nu; };
    for (int it = 0; it < nIter; it++) {
        for (int f = 0; f < nFr; f++) cg_reset(sc + f, s);
        forward_batch(x, Mp, s, nA);
        for (int f = 0; f < nFr; f++) vsub(ys + koff[f], K(Mp, f), K(res, f), koff[f + 1] - koff[f], s);
        adjoint_batch(res, dag, hy, invHX, s, nA);
        for (int f = 0; f < nFr; f++) {
            sqnorm(X(dag, f), nu, nullptr, 1, HX, &sc[f].sqn, false, part, s);
            vcopy(X(dag, f), X(p, f), nu, s);
        }
        for (int i = 1; i <= nCGD; i++) {
            forward_batch(p, Mp, s, nA);
            for (int f = 0; f < nFr; f++) {
                sqnorm(K(Mp, f), koff[f + 1] - koff[f], hy + koff[f] / nCh, nCh, 1.0, &sc[f].den, false, part, s);
                cg_alpha(sc + f, s);
                vaxpy(X(x, f), X(p, f), &sc[f].a, 1.0f, nu, s);
            }
            if (i == nCGD) break;
            for (int f = 0; f < nFr; f++) vaxpy(K(res, f), K(Mp, f), &sc[f].a, -1.0f, koff[f + 1] - koff[f], s);
            adjoint_batch(res, dag, hy, invHX, s, nA);
            for (int f = 0; f < nFr; f++) {
                sqnorm(X(dag, f), nu, nullptr, 1, HX, &sc[f].sqn_new, false, part, s);
                cg_beta(sc + f, s);
                vxpby(X(p, f), X(dag, f), &sc[f].b, nu, s);
            }
        }
        forward_batch(x, Mp, s, nA);
        for (int f = 0; f < nFr; f++) {
            vsub(ys + koff[f], K(Mp, f), K(res, f), koff[f + 1] - koff[f], s);
            sqnorm(K(res, f), koff[f + 1] - koff[f], hy + koff[f] / nCh, nCh, 1.0, hist + (int64_t)f * nIter + it,
                   false, part, s);
        }
    }
    if (residual && nIter > 0) {
        MONA_HIP(hipMemcpyAsync(residual, hist, (size_t)nIter * nFr * sizeof(double), hipMemcpyDeviceToHost, s));
        MONA_HIP(hipStreamSynchronize(s));
    }
}

// ================================================================ mixed precision
// Operators with the grid-side arrays (FFT intermediates, oversampled grid)
// and the k-space vector stored as bf16x2; gridding and FFT arithmetic in
// float.  k-space in the frame's tile order.
void Plan::check_bf16() const {
    MONA_CHECK(nCh <= 64, MONA_ERR_ARG, "bf16 storage needs nCh <= 64");
    for (int d = 0; d < imDim; d++)
        MONA_CHECK(fft_reg_supported(N[d]), MONA_ERR_ARG,
                   "bf16 storage needs N_u = 16 * {2,3,4,5,6,8,10,12,16} in every dimension");
}

void Plan::forward_bf16(int f, const float2* x, bf16x2* y, hipStream_t s) {
    MONA_CHECK(coils_set, MONA_ERR_STATE, "coil maps not set");
    check_bf16();
    const Frame& fr = frame(f);
    ensure_grid();
    ensure_fft_scratch();
    {
        TimedRegion t(2, s);
        forward_fft_pruned(x, grid.as<float2>(), 1, fft_a, fft_b, s, true, fwd_mask(f), ivals_of(f),
                           mid_ivals_of(f));
    }
    {
        TimedRegion t(0, s);
        if (frame_mf(f))
            u2n_mf(*fr.mf, nullptr, fr.mf->P, grid.p, true, Nu, y, true, nCh, imDim, mf_geo(), mfk, s);
        else
            spmv_c2_rowmajor_st(*fr.Gu_s, grid.p, true, y, true, nCh, nullptr, s);
    }
}

void Plan::adjoint_bf16(int f, const bf16x2* y, float2* x, const float* hy, float scale, hipStream_t s) {
    MONA_CHECK(coils_set, MONA_ERR_STATE, "coil maps not set");
    check_bf16();
    const Frame& fr = frame(f);
    ensure_grid();
    ensure_fft_scratch();
    const bool limited = masked_grid(f) || ival_grid(f);
    {
        TimedRegion t(1, s);
        spmv_c2_rowmajor_st(*fr.Gut_s, y, true, grid.p, true, nCh, hy, s, nullptr, !limited, use_staged);
    }
    {
        TimedRegion t(2, s);
        adjoint_fft_pruned(grid.as<float2>(), x, true, scale, 1, fft_a, fft_b, s, true, adj_mask(f), ivals_of(f),
                           mid_ivals_of(f));
    }
}

void dotre(const float2* a, const float2* b, int64_t n, double cw, double* out, double* partial, hipStream_t s) {
    int nb = (int)std::min<int64_t>(kRedBlocks, std::max<int64_t>(1, (n + kRedThreads - 1) / kRedThreads));
    k_dotre_partial<<<nb, kRedThreads, 0, s>>>(a, b, n, cw, partial);
    MONA_HIP(hipGetLastError());
    k_reduce_final<<<1, kRedThreads, 0, s>>>(partial, nb, out, 0);
    MONA_HIP(hipGetLastError());
}

// CG-SENSE by iterative refinement with reliable updates.  CG on the normal
// equations of bmSensa.m:6-96, (1/HX) M^H HY M x = (1/HX) M^H HY y, with the
// same HX / HY inner products, in image space.  Its products A p use the
// operators with bf16x2 storage (bf16) or float storage; every nInner steps
// the recursively updated residual is replaced by the true residual
//   g_k = (1/HX) M^H HY (y - M x_k)     (float operators)
// and CG continues with its search direction (no restart), so the
// reduced-precision products steer the Krylov iteration while the float
// residual sets where it converges.  rel = ||g||_X / ||b||_X with
// b = (1/HX) M^H HY y (the residual of x = 0).  After every step the
// recursive residual's norm is read; when it falls below tol the true
// residual is computed at once (at most once per period) and the solve stops
// if that is below tol too; it also stops after nOuter periodic
// replacements.  Scalars in FP64.
void Plan::sensa_ir(int f, float2* x, const float2* y, const float* ve, float ve_max, int nInner, int nOuter,
                    double tol, bool bf16, double* rel_out, int* n_outer, int* n_steps, hipStream_t s) {
    MONA_CHECK(nInner >= 1 && nOuter >= 0, MONA_ERR_ARG, "nInner must be >= 1 and nOuter >= 0");
    if (bf16) check_bf16();
    const Frame& fr = frame(f);
    int64_t P = fr.Gu->l_size;
    int64_t ny = P * nCh;
    Work& w = work;
    w.hy.ensure((size_t)P * sizeof(float));
    w.k1.ensure((size_t)ny * sizeof(float2));
    w.k2.ensure((size_t)ny * sizeof(float2));
    w.i1.ensure((size_t)nu * sizeof(float2));
    w.i2.ensure((size_t)nu * sizeof(float2));
    w.i4.ensure((size_t)nu * sizeof(float2));
    w.sc.ensure(sizeof(CgScalars));
    w.partial.ensure(kRedBlocks * sizeof(double));
    w.hist.ensure(sizeof(double));
    float* hy = w.hy.as<float>();
    float2* res = w.k1.as<float2>();
    float2* r = w.i1.as<float2>();
    float2* p = w.i2.as<float2>();
    float2* q = w.i4.as<float2>();
    CgScalars* sc = w.sc.as<CgScalars>();
    double* part = w.partial.as<double>();
    double* hist = w.hist.as<double>();
    w.hy0.ensure((size_t)P * sizeof(float));
    w.ys.ensure((size_t)ny * sizeof(float2));
    compute_hy(ve, ve_max, w.hy0.as<float>(), P, hy_scratch(w), s);
    weights_to_tile_order(f, w.hy0.as<float>(), hy, s);
    to_tile_order(f, y, w.ys.as<float2>(), s);
    y = w.ys.as<float2>();
    const float invHX = (float)(1.0 / HX);
    // true residual r = (1/HX) M^H HY (y - M x) into r, its squared norm into dst
    auto true_residual = [&](double* dst) {
        forward(f, x, res, s);
        vsub(y, res, res, ny, s);
        adjoint(f, res, r, true, hy, invHX, s);
        sqnorm(r, nu, nullptr, 1, HX, dst, false, part, s);
    };
    double bb = 0.0;
    adjoint(f, y, r, true, hy, invHX, s);                    // b = (1/HX) M^H HY y
    sqnorm(r, nu, nullptr, 1, HX, hist, false, part, s);
    MONA_HIP(hipMemcpyAsync(&bb, hist, sizeof(double), hipMemcpyDeviceToHost, s));
    cg_reset(sc, s);
    true_residual(&sc->sqn);
    vcopy(r, p, nu, s);
    // host copy of a device FP64 value; the stream is synchronised
    auto read = [&](const double* d) {
        double h = 0.0;
        MONA_HIP(hipMemcpyAsync(&h, d, sizeof(double), hipMemcpyDeviceToHost, s));
        MONA_HIP(hipStreamSynchronize(s));
        return h;
    };
    auto rel_of = [&](double sq) { return bb > 0 ? std::sqrt(sq / bb) : 0.0; };
    int k = 0;  // replacements recorded
    double rel = rel_of(read(&sc->sqn));
    if (rel_out) rel_out[0] = rel;
    int periodic = 0, in_period = 0, steps = 0;
    bool armed = true;  // an early replacement may be tried (once per period)
    while (rel >= tol && periodic < nOuter && bb > 0) {
        in_period++;
        steps++;
        if (bf16) {
            bf16x2* Mp = reinterpret_cast<bf16x2*>(w.k2.p);
            forward_bf16(f, p, Mp, s);
            adjoint_bf16(f, Mp, q, hy, invHX, s);
        } else {
            float2* Mp = w.k2.as<float2>();
            forward(f, p, Mp, s);
            adjoint(f, Mp, q, true, hy, invHX, s);
        }
        dotre(p, q, nu, HX, &sc->den, part, s);              // <p, A p>_X
        cg_alpha(sc, s, 2.2250738585072014e-308);            // relative solve: stop only at 0
        vaxpy(x, p, &sc->a, 1.0f, nu, s);                    // x += a p
        bool replaced = false;
        if (in_period < nInner) {
            vaxpy(r, q, &sc->a, -1.0f, nu, s);               // r -= a A p
            sqnorm(r, nu, nullptr, 1, HX, &sc->sqn_new, false, part, s);
            if (armed && rel_of(read(&sc->sqn_new)) < tol) {  // converged by the estimate:
                true_residual(&sc->sqn_new);                  // confirm with the true one
                armed = false;
                replaced = true;
            }
        } else {
            true_residual(&sc->sqn_new);                      // periodic reliable update
            periodic++;
            in_period = 0;
            armed = true;
            replaced = true;
        }
        if (replaced) {
            rel = rel_of(read(&sc->sqn_new));
            k++;
            if (rel_out) rel_out[k] = rel;
        }
        cg_beta(sc, s);
        vxpby(p, r, &sc->b, nu, s);                          // p = r + b p
    }
    if (n_outer) *n_outer = k;
    if (n_steps) *n_steps = steps;
}

}  // namespace mona

using namespace mona;

extern "C" {

int mona_plan_create(int imDim, const int* N_u, const int* n_u, const double* dK_u, int nCh,
                     int kernelType, int nWin, const double* kernelParam, int nParam, mona_plan* out) {
    return guarded([&] {
        MONA_CHECK(out && N_u && dK_u && kernelParam, MONA_ERR_ARG, "NULL argument");
        *out = nullptr;
        auto* h = new mona_plan_s;
        try {
            h->P.init(imDim, N_u, n_u, dK_u, nCh, kernelType, nWin, kernelParam, nParam);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

int mona_plan_destroy(mona_plan P) {
    return guarded([&] { delete P; });
}

int mona_plan_set_coils(mona_plan P, const void* C, hipStream_t stream) {
    return guarded([&] {
        MONA_CHECK(P && C, MONA_ERR_ARG, "NULL argument");
        P->P.set_coils(static_cast<const float2*>(C), stream);
    });
}

int mona_plan_set_frame(mona_plan P, int frame, mona_sparse Gu, mona_sparse Gut) {
    return guarded([&] {
        MONA_CHECK(P && Gu && Gut && frame >= 0, MONA_ERR_ARG, "bad argument");
        const SparseDev& u = Gu->S;
        const SparseDev& ut = Gut->S;
        MONA_CHECK(u.r_size == P->P.Nu, MONA_ERR_ARG, "Gu has r_size != prod(N_u)");
        MONA_CHECK(ut.l_size == P->P.Nu, MONA_ERR_ARG, "Gut has l_size != prod(N_u)");
        MONA_CHECK(ut.r_size == u.l_size, MONA_ERR_ARG, "Gut is not the transpose shape of Gu");
        HostCallCtx& C = host_ctx();
        {
            std::lock_guard<std::mutex> lk(C.mu);
            C.init();
        }
        P->P.set_frame(frame, &Gu->S, &Gut->S, C.stream);
    });
}

int mona_plan_get_K(mona_plan P, float* K, hipStream_t stream) {
    return guarded([&] {
        MONA_CHECK(P && K && P->P.coils_set, MONA_ERR_STATE, "coil maps not set");
        MONA_HIP(hipMemcpyAsync(K, P->P.K.p, P->P.nu * sizeof(float), hipMemcpyDeviceToDevice, stream));
    });
}

int mona_plan_get_KFC(mona_plan P, void* KFC, void* KFCc, hipStream_t stream) {
    return guarded([&] {
        MONA_CHECK(P && P->P.coils_set, MONA_ERR_STATE, "coil maps not set");
        size_t b = (size_t)P->P.nu * P->P.nCh * sizeof(float2);
        if (KFC) MONA_HIP(hipMemcpyAsync(KFC, P->P.KFC.p, b, hipMemcpyDeviceToDevice, stream));
        if (KFCc) MONA_HIP(hipMemcpyAsync(KFCc, P->P.KFCc.p, b, hipMemcpyDeviceToDevice, stream));
    });
}

int mona_shanna(mona_plan P, int frame, const void* x, void* y, hipStream_t stream) {
    return guarded([&] {
        MONA_CHECK(P && x && y, MONA_ERR_ARG, "NULL argument");
        P->P.forward_public(frame, static_cast<const float2*>(x), static_cast<float2*>(y), stream);
    });
}

int mona_nakatsha(mona_plan P, int frame, const void* y, void* x, int C_flag, hipStream_t stream) {
    return guarded([&] {
        MONA_CHECK(P && x && y, MONA_ERR_ARG, "NULL argument");
        P->P.adjoint_public(frame, static_cast<const float2*>(y), static_cast<float2*>(x), C_flag != 0, stream);
    });
}

int mona_sensa(mona_plan P, int frame, void* x, const void* y, const float* ve, float ve_max, int nIter,
               int nCGD, double* residual, hipStream_t stream) {
    return guarded([&] {
        MONA_CHECK(P && x && y && ve, MONA_ERR_ARG, "NULL argument");
        P->P.sensa(frame, static_cast<float2*>(x), static_cast<const float2*>(y), ve, ve_max, nIter, nCGD,
                   residual, stream);
    });
}

int mona_sensa_batch(mona_plan P, int nFrames, void* x, const void* y, const long long* y_off, const float* ve,
                     const long long* ve_off, const float* ve_max, int nIter, int nCGD, double* residual,
                     hipStream_t stream) {
    return guarded([&] {
        MONA_CHECK(P && x && y && y_off && ve && ve_off, MONA_ERR_ARG, "NULL argument");
        P->P.sensa_batch(nFrames, static_cast<float2*>(x), static_cast<const float2*>(y), y_off, ve, ve_off, ve_max,
                         nIter, nCGD, residual, stream);
    });
}

int mona_plan_memory(mona_plan P, long long* bytes) {
    return guarded([&] {
        MONA_CHECK(P && bytes, MONA_ERR_ARG, "NULL argument");
        *bytes = (long long)P->P.device_bytes();
    });
}

int mona_sensa_ir(mona_plan P, int frame, void* x, const void* y, const float* ve, float ve_max, int nInner,
                  int nOuter, double tol, int bf16, double* rel_res, int* n_outer, int* n_steps,
                  hipStream_t stream) {
    return guarded([&] {
        MONA_CHECK(P && x && y && ve, MONA_ERR_ARG, "NULL argument");
        P->P.sensa_ir(frame, static_cast<float2*>(x), static_cast<const float2*>(y), ve, ve_max, nInner, nOuter,
                      tol, bf16 != 0, rel_res, n_outer, n_steps, stream);
    });
}

int mona_shanna_bf16(mona_plan P, int frame, const void* x, void* y, hipStream_t stream) {
    return guarded([&] {
        MONA_CHECK(P && x && y, MONA_ERR_ARG, "NULL argument");
        Plan& p = P->P;
        const Frame& fr = p.frame(frame);
        p.work.kpub.ensure((size_t)fr.P * p.nCh * sizeof(float2));
        bf16x2* t = reinterpret_cast<bf16x2*>(p.work.kpub.p);
        p.forward_bf16(frame, static_cast<const float2*>(x), t, stream);
        k_scatter_pts_bf<<<nb256(fr.P * p.nCh), 256, 0, stream>>>(t, fr.perm->as<int>(), static_cast<float2*>(y),
                                                                   fr.P, p.nCh);
        MONA_HIP(hipGetLastError());
    });
}

int mona_nakatsha_bf16(mona_plan P, int frame, const void* y, void* x, hipStream_t stream) {
    return guarded([&] {
        MONA_CHECK(P && x && y, MONA_ERR_ARG, "NULL argument");
        Plan& p = P->P;
        const Frame& fr = p.frame(frame);
        p.work.kpub.ensure((size_t)fr.P * p.nCh * sizeof(float2));
        bf16x2* t = reinterpret_cast<bf16x2*>(p.work.kpub.p);
        k_gather_pts_bf<<<nb256(fr.P * p.nCh), 256, 0, stream>>>(static_cast<const float2*>(y), fr.perm->as<int>(),
                                                                  t, fr.P, p.nCh);
        MONA_HIP(hipGetLastError());
        p.adjoint_bf16(frame, t, static_cast<float2*>(x), nullptr, 1.0f, stream);
    });
}

int mona_plan_set_witness(mona_plan P, void* slices, int n_slots, void* full, const int* full_ind, int n_full) {
    return guarded([&] {
        MONA_CHECK(P, MONA_ERR_ARG, "NULL plan");
        MONA_CHECK(n_slots >= 0 && n_full >= 0, MONA_ERR_ARG, "n_slots and n_full must be >= 0");
        MONA_CHECK(!slices || n_slots > 0, MONA_ERR_ARG, "a witness slice buffer needs n_slots >= 1");
        MONA_CHECK(!full || (full_ind && n_full > 0), MONA_ERR_ARG, "a full-image buffer needs its iteration list");
        Plan& p = P->P;
        p.wit_slices = static_cast<float2*>(slices);
        p.wit_slots = slices ? n_slots : 0;
        p.wit_full = static_cast<float2*>(full);
        p.wit_full_ind.assign(full ? full_ind : nullptr, full ? full_ind + n_full : nullptr);
    });
}

int mona_plan_witness_elems(mona_plan P, long long* elems) {
    return guarded([&] {
        MONA_CHECK(P && elems, MONA_ERR_ARG, "NULL argument");
        *elems = (long long)P->P.witness_elems();
    });
}

int mona_plan_set_option(mona_plan P, int option, long long value) {
    return guarded([&] {
        MONA_CHECK(P, MONA_ERR_ARG, "NULL plan");
        switch (option) {
            case MONA_OPT_BATCH_FRAMES: P->P.use_batch = value != 0; break;
            case MONA_OPT_GRAPH: P->P.use_graph = value != 0; break;
            case MONA_OPT_FFT_TILE:
                MONA_CHECK(value == 0 || value == 2048 || value == 4096 || value == 8192, MONA_ERR_ARG,
                           "FFT tile must be 0, 2048, 4096 or 8192");
                P->P.fft_tile = (int)value;
                break;
            case MONA_OPT_FFT_LINES:
                MONA_CHECK(value == -1 || value == 0 || value == 16 || value == 32 || value == 64, MONA_ERR_ARG,
                           "FFT lines per tile must be -1, 0, 16, 32 or 64");
                P->P.fft_lines = (int)value;
                break;
            case MONA_OPT_MATRIX_FREE: P->P.use_mf = value != 0; break;
            case MONA_OPT_STAGED_ADJOINT: P->P.use_staged = value != 0; break;
            default: MONA_CHECK(false, MONA_ERR_ARG, "unknown option");
        }
    });
}

// Times one forward + adjoint of frame `frame` for each FFT pass variant
// (register kernel with 16 / 32 / 64 lines per tile; LDS Stockham kernel
// with tiles 2048 / 4096 / 8192) and keeps the fastest; two timed
// repetitions after one warm-up each.  *chosen = lines per tile of the
// winner, or minus its Stockham tile.
int mona_plan_autotune(mona_plan P, int frame, hipStream_t stream, long long* chosen) {
    return guarded([&] {
        MONA_CHECK(P, MONA_ERR_ARG, "NULL plan");
        Plan& p = P->P;
        const Frame& fr = p.frame(frame);
        DevBuf xi((size_t)p.nu * sizeof(float2)), xo((size_t)p.nu * sizeof(float2)),
            yk((size_t)fr.P * p.nCh * sizeof(float2));
        fill_f32(xi.as<float>(), 1.0f, p.nu * 2, stream);
        hipEvent_t a, b;
        MONA_HIP(hipEventCreate(&a));
        MONA_HIP(hipEventCreate(&b));
        struct Variant { int tile, lines; };
        const Variant vars[] = {{4096, 16}, {4096, 32}, {4096, 64}, {2048, -1}, {4096, -1}, {8192, -1}};
        float best = 1e30f;
        Variant win = {4096, 16};
        try {
            for (const Variant& v : vars) {
                p.fft_tile = v.tile;
                p.fft_lines = v.lines;
                p.forward(frame, xi.as<float2>(), yk.as<float2>(), stream);
                p.adjoint(frame, yk.as<float2>(), xo.as<float2>(), true, nullptr, 1.0f, stream);
                MONA_HIP(hipEventRecord(a, stream));
                for (int r = 0; r < 2; r++) {
                    p.forward(frame, xi.as<float2>(), yk.as<float2>(), stream);
                    p.adjoint(frame, yk.as<float2>(), xo.as<float2>(), true, nullptr, 1.0f, stream);
                }
                MONA_HIP(hipEventRecord(b, stream));
                MONA_HIP(hipEventSynchronize(b));
                float ms = 0;
                MONA_HIP(hipEventElapsedTime(&ms, a, b));
                if (ms < best) { best = ms; win = v; }
            }
        } catch (...) {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
            throw;
        }
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        p.fft_tile = win.tile;
        p.fft_lines = win.lines;
        if (chosen) *chosen = win.lines > 0 ? win.lines : -(long long)win.tile;
    });
}

void Plan::regroup(hipStream_t s) {
    std::vector<const SparseDev*> done;
    for (Frame& fr : frames) {
        if (!fr.Gut_s || std::find(done.begin(), done.end(), fr.Gut_s.get()) != done.end()) continue;
        rebuild_groups(*fr.Gut_s, nCh, s, grp_stage_kb, grp_entries, grp_rows);
        done.push_back(fr.Gut_s.get());
    }
    Gu_b.reset();  // the batched matrices stack the new groups on next use
    Gut_b.reset();
    koff_b.clear();
}

// Times the weighted adjoint gridding of frame `frame` for several group
// sizes of its staged form (LDS budget of the k-space rows, rows per group)
// and keeps the fastest for every frame; two timed repetitions after one
// warm-up each.
int mona_plan_autotune_groups(mona_plan P, int frame, hipStream_t stream, int* stage_kb, int* rows) {
    return guarded([&] {
        MONA_CHECK(P, MONA_ERR_ARG, "NULL plan");
        Plan& p = P->P;
        const Frame& fr = p.frame(frame);
        DevBuf xo((size_t)p.nu * sizeof(float2)), yk((size_t)fr.P * p.nCh * sizeof(float2)),
            hy((size_t)std::max<int64_t>(fr.P, 1) * sizeof(float));
        fill_f32(yk.as<float>(), 1.0f, fr.P * p.nCh * 2, stream);
        fill_f32(hy.as<float>(), 1.0f, fr.P, stream);
        hipEvent_t a, b;
        MONA_HIP(hipEventCreate(&a));
        MONA_HIP(hipEventCreate(&b));
        struct Variant { int kb, rows; };
        const Variant vars[] = {{16, 256}, {12, 256}, {24, 256}, {16, 128}, {16, 512}};
        float best = 1e30f;
        Variant win = vars[0];
        const bool staged = p.use_staged;
        p.use_staged = true;
        try {
            for (const Variant& v : vars) {
                rebuild_groups(*fr.Gut_s, p.nCh, stream, v.kb, p.grp_entries, v.rows);
                if (fr.Gut_s->n_group == 0) continue;
                p.adjoint(frame, yk.as<float2>(), xo.as<float2>(), true, hy.as<float>(), 1.0f, stream);
                MONA_HIP(hipEventRecord(a, stream));
                for (int r = 0; r < 2; r++)
                    p.adjoint(frame, yk.as<float2>(), xo.as<float2>(), true, hy.as<float>(), 1.0f, stream);
                MONA_HIP(hipEventRecord(b, stream));
                MONA_HIP(hipEventSynchronize(b));
                float ms = 0;
                MONA_HIP(hipEventElapsedTime(&ms, a, b));
                if (ms < best) { best = ms; win = v; }
            }
        } catch (...) {
            p.use_staged = staged;
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
            throw;
        }
        p.use_staged = staged;
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        p.grp_stage_kb = win.kb;
        p.grp_rows = win.rows;
        p.regroup(stream);
        if (stage_kb) *stage_kb = win.kb;
        if (rows) *rows = win.rows;
    });
}

int mona_timing_enable(int on) {
    return guarded([&] {
        std::lock_guard<std::mutex> lk(g_tmu);
        if (!on) timing_flush_locked();
        g_timing = on != 0;
        if (on) {
            for (int k = 0; k < 4; k++) { g_tms[k] = 0; g_tn[k] = 0; }
        }
    });
}

int mona_timing_read(int kind, long long* launches, double* total_ms) {
    return guarded([&] {
        MONA_CHECK(kind >= 0 && kind < 4, MONA_ERR_ARG, "kind must be 0..3");
        std::lock_guard<std::mutex> lk(g_tmu);
        timing_flush_locked();
        if (launches) *launches = g_tn[kind];
        if (total_ms) *total_ms = g_tms[kind];
    });
}

}  // extern "C"
