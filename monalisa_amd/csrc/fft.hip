// Centred FFT on the oversampled grid: one LDS-staged Stockham pass per
// axis, mixed radix 8/5/4/3/2, batched over everything else.
//
// The centring of bmShanna_MATLAB.m:46-50 / bmNakatsha_MATLAB.m:36-40,
// fftshift(fft(ifftshift(x))) along an axis of even length N, equals
//   Y[k] = (-1)^(k + N/2) * sum_n (-1)^n x[n] e^(-+2 pi i n k / N)
// so a pass multiplies by (-1)^n when it loads a line and by
// (-1)^(k + N/2) (and 1/N for the inverse) when it stores it; no shift
// copies are made.
//
// Pruning and fusion.  A pass reads only the window of rows that can be
// non-zero (the zero-filled image occupies the centre n of N) and writes
// only the window of rows that is kept (the adjoint crops to the centre), so
// the forward operator's three passes move 1/8 -> 1/4 -> 1/2 -> 1 of the grid
// instead of 1 -> 1 three times.  The forward pass along x takes its input
// straight from x .* KFC (bmShanna_MATLAB.m:34-40) and the adjoint pass
// along x ends with .* KFC_conj and the coil sum (bmNakatsha_MATLAB.m:46-55).
//
// Tile.  A workgroup owns L lines = Bo outer indices x W consecutive inner
// positions, all N rows, stored transposed in LDS as [n][line]: the global
// rows are runs of W float2, the LDS writes of a wave are contiguous, and
// the butterfly reads/writes of a wave touch consecutive lines (no bank
// conflicts).  Each work-item keeps its butterflies' inputs in registers
// across one barrier per stage (in-place Stockham autosort).
#include <cstdlib>
#include <map>
#include <memory>

#include "fft.hpp"
#include "mona_internal.hpp"

namespace mona {

namespace {

constexpr int kMaxStages = 16;

struct PassArgs {
    int N, nstage;
    int radix[kMaxStages];
    int64_t S, outer;
    int W, Bo;
    int64_t ntile_w;
    const float2* in;
    int64_t in_os;
    int in_off, in_cnt;
    float2* out;
    int64_t out_os;
    int out_off, out_cnt;
    int inverse, presign, postsign;
    float scale, epi_scale;
    const float2* img;   // prologue: image, [outer][in_cnt]
    const float2* kfc;   // prologue: KFC [outer*in_cnt][S]; epilogue: KFC_conj [outer*out_cnt][S]
    int64_t kfc_period;  // KFC index taken modulo this many pixels (frames share the maps)
    float2* img_out;     // epilogue (combine): [outer][out_cnt]
    int xs;              // register kernel: LDS stride between the N1 exchange planes
    const uint32_t* in_mask;   // register kernel, mode 0: cell masks (see AxisPass)
    const uint32_t* out_mask;
    int mask_nch;
    int64_t mask_cpl;          // mask columns per line block: S / mask_nch
    const int2* in_ival;       // register kernel, mode 0: row intervals per column (see AxisPass)
    const int2* out_ival;
    int64_t in_rs, out_rs;     // row strides of in / out (S unless the layout interleaves lines)
    int64_t in_on, in_os2;     // two-level line offsets (line_off); in_on = 0: o * in_os
    int64_t out_on, out_os2;
    // register kernel: ceil(2^20 / L) and ceil(2^20 / W) (exact quotients of
    // work-item indices below 1024), ceil(2^32 / mask_nch) (exact for inner
    // positions below 2^32 / mask_nch; 0 when mask_nch is 1)
    uint32_t lmag, wmag, nchmag;
    // register kernel: ceil(2^32 / d) for the divisors ntile_w, in_on and
    // out_on (exact for the quotients the launch takes; 0: divide)
    uint32_t twmag, inmag, outmag;
    int xg;  // register kernel: runs of consecutive tiles per XCD (1 = round-robin)
};

// Offset of line o: o * os, or with on > 0 (o / on) * os2 + (o % on) * os
// (lines of several frames whose per-frame layout interleaves rows).
__device__ __forceinline__ int64_t line_off(int64_t o, int64_t os, int64_t on, int64_t os2) {
    if (on <= 0) return o * os;
    const int64_t hi = o / on;
    return hi * os2 + (o - hi * on) * os;
}
// The same with o / on as a multiply-high when mag = ceil(2^32 / on) is
// exact for every line of the launch (o on < 2^32, checked on the host)
__device__ __forceinline__ int64_t line_off_m(int64_t o, int64_t os, int64_t on, int64_t os2, uint32_t mag) {
    if (on <= 0) return o * os;
    const int64_t hi = mag ? (int64_t)__umulhi((uint32_t)o, mag) : o / on;
    return hi * os2 + (o - hi * on) * os;
}
// n / d for the host-computed magic of d (mag = 0: divide)
__device__ __forceinline__ int64_t div_m(int64_t n, int64_t d, uint32_t mag) {
    return mag ? (int64_t)__umulhi((uint32_t)n, mag) : n / d;
}

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// multiply by sigma*i, sigma = -1 (forward) or +1 (inverse)
__device__ __forceinline__ float2 mul_si(float2 a, float sigma) { return make_float2(-sigma * a.y, sigma * a.x); }

template <int R> struct Dft;

template <> struct Dft<2> {
    __device__ static void run(float2* a, float) {
        float2 t = a[0];
        a[0] = cadd(t, a[1]);
        a[1] = csub(t, a[1]);
    }
};

template <> struct Dft<4> {
    __device__ static void run(float2* a, float sg) {
        float2 t0 = cadd(a[0], a[2]), t1 = csub(a[0], a[2]);
        float2 t2 = cadd(a[1], a[3]), t3 = mul_si(csub(a[1], a[3]), sg);
        a[0] = cadd(t0, t2);
        a[2] = csub(t0, t2);
        a[1] = cadd(t1, t3);
        a[3] = csub(t1, t3);
    }
};

template <> struct Dft<3> {
    __device__ static void run(float2* a, float sg) {
        const float h = 0.86602540378443864676f;  // sin(2 pi / 3)
        float2 t1 = cadd(a[1], a[2]), t2 = csub(a[1], a[2]);
        float2 y0 = cadd(a[0], t1);
        float2 m1 = make_float2(a[0].x - 0.5f * t1.x, a[0].y - 0.5f * t1.y);
        float2 m2 = mul_si(make_float2(h * t2.x, h * t2.y), sg);
        a[0] = y0;
        a[1] = cadd(m1, m2);
        a[2] = csub(m1, m2);
    }
};

template <> struct Dft<5> {
    __device__ static void run(float2* a, float sg) {
        const float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
        const float s1 = 0.95105651629515357212f, s2 = 0.58778525229247312917f;
        float2 t1 = cadd(a[1], a[4]), t2 = cadd(a[2], a[3]);
        float2 t3 = csub(a[1], a[4]), t4 = csub(a[2], a[3]);
        float2 y0 = make_float2(a[0].x + t1.x + t2.x, a[0].y + t1.y + t2.y);
        float2 b1 = make_float2(a[0].x + c1 * t1.x + c2 * t2.x, a[0].y + c1 * t1.y + c2 * t2.y);
        float2 b2 = make_float2(a[0].x + c2 * t1.x + c1 * t2.x, a[0].y + c2 * t1.y + c1 * t2.y);
        float2 d1 = mul_si(make_float2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y), sg);
        float2 d2 = mul_si(make_float2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y), sg);
        a[0] = y0;
        a[1] = cadd(b1, d1);
        a[4] = csub(b1, d1);
        a[2] = cadd(b2, d2);
        a[3] = csub(b2, d2);
    }
};

template <> struct Dft<8> {
    __device__ static void run(float2* a, float sg) {
        const float r = 0.70710678118654752440f;
        float2 e[4] = {a[0], a[2], a[4], a[6]};
        float2 o[4] = {a[1], a[3], a[5], a[7]};
        Dft<4>::run(e, sg);
        Dft<4>::run(o, sg);
        // w8^k, w8 = exp(sg * i * pi / 4)
        float2 w1 = make_float2(r, sg * r), w3 = make_float2(-r, sg * r);
        float2 o1 = cmul(o[1], w1), o2 = mul_si(o[2], sg), o3 = cmul(o[3], w3);
        a[0] = cadd(e[0], o[0]);
        a[4] = csub(e[0], o[0]);
        a[1] = cadd(e[1], o1);
        a[5] = csub(e[1], o1);
        a[2] = cadd(e[2], o2);
        a[6] = csub(e[2], o2);
        a[3] = cadd(e[3], o3);
        a[7] = csub(e[3], o3);
    }
};

// Radix 16 as 4 x 4: n = 4 n1 + n2, k = k1 + 4 k2.
template <> struct Dft<16> {
    __device__ static void run(float2* a, float sg) {
        // cos / sin of 2 pi m / 16, m = 0..9
        const float c[10] = {1.0f, 0.92387953251128675613f, 0.70710678118654752440f, 0.38268343236508977173f,
                             0.0f, -0.38268343236508977173f, -0.70710678118654752440f, -0.92387953251128675613f,
                             -1.0f, -0.92387953251128675613f};
        const float sn[10] = {0.0f, 0.38268343236508977173f, 0.70710678118654752440f, 0.92387953251128675613f,
                              1.0f, 0.92387953251128675613f, 0.70710678118654752440f, 0.38268343236508977173f,
                              0.0f, -0.38268343236508977173f};
        float2 b[4][4];
#pragma unroll
        for (int n2 = 0; n2 < 4; n2++) {
            float2 t[4] = {a[n2], a[4 + n2], a[8 + n2], a[12 + n2]};
            Dft<4>::run(t, sg);
#pragma unroll
            for (int k1 = 0; k1 < 4; k1++) {
                int m = n2 * k1;  // w16^m, m <= 9
                float2 w = make_float2(c[m], sg * sn[m]);
                b[n2][k1] = (m == 0) ? t[k1] : cmul(t[k1], w);
            }
        }
#pragma unroll
        for (int k1 = 0; k1 < 4; k1++) {
            float2 t[4] = {b[0][k1], b[1][k1], b[2][k1], b[3][k1]};
            Dft<4>::run(t, sg);
#pragma unroll
            for (int k2 = 0; k2 < 4; k2++) a[k1 + 4 * k2] = t[k2];
        }
    }
    // The same DFT when only a[4..11] can be non-zero (a line whose input is
    // the central half): each first-layer radix-4 butterfly then has inputs
    // (0, p, q, 0), giving (q + p, -q + s p, q - p, -q - s p) with s p = p
    // times sg i, in half the additions.  a[0..3] and a[12..15] are not
    // read.  Results agree with run() on such input to float rounding (the
    // compiler may fuse the surrounding products into multiply-adds
    // differently).
    __device__ static void run_mid(float2* a, float sg) {
        const float c[10] = {1.0f, 0.92387953251128675613f, 0.70710678118654752440f, 0.38268343236508977173f,
                             0.0f, -0.38268343236508977173f, -0.70710678118654752440f, -0.92387953251128675613f,
                             -1.0f, -0.92387953251128675613f};
        const float sn[10] = {0.0f, 0.38268343236508977173f, 0.70710678118654752440f, 0.92387953251128675613f,
                              1.0f, 0.92387953251128675613f, 0.70710678118654752440f, 0.38268343236508977173f,
                              0.0f, -0.38268343236508977173f};
        float2 b[4][4];
#pragma unroll
        for (int n2 = 0; n2 < 4; n2++) {
            const float2 p = a[4 + n2], q = a[8 + n2];
            const float2 sp = mul_si(p, sg);
            float2 t[4];
            t[0] = cadd(q, p);
            t[1] = make_float2(sp.x - q.x, sp.y - q.y);
            t[2] = csub(q, p);
            t[3] = make_float2(-q.x - sp.x, -q.y - sp.y);
#pragma unroll
            for (int k1 = 0; k1 < 4; k1++) {
                int m = n2 * k1;
                float2 w = make_float2(c[m], sg * sn[m]);
                b[n2][k1] = (m == 0) ? t[k1] : cmul(t[k1], w);
            }
        }
#pragma unroll
        for (int k1 = 0; k1 < 4; k1++) {
            float2 t[4] = {b[0][k1], b[1][k1], b[2][k1], b[3][k1]};
            Dft<4>::run(t, sg);
#pragma unroll
            for (int k2 = 0; k2 < 4; k2++) a[k1 + 4 * k2] = t[k2];
        }
    }
};

// Composite register DFTs R = A * B (Cooley-Tukey, n = B n1 + n2,
// k = k1 + A k2) with twiddles exp(sg 2 pi i m / R) folded to constants.
constexpr double kPi = 3.14159265358979323846;
constexpr double ce_sin(double x) {  // |x| <= pi, Taylor series to x^27
    double term = x, sum = x;
    for (int k = 1; k < 14; k++) {
        term *= -x * x / ((2.0 * k) * (2.0 * k + 1.0));
        sum += term;
    }
    return sum;
}
constexpr double ce_cos(double x) {
    double term = 1.0, sum = 1.0;
    for (int k = 1; k < 14; k++) {
        term *= -x * x / ((2.0 * k - 1.0) * (2.0 * k));
        sum += term;
    }
    return sum;
}
template <int R> struct TwConst {
    float c[R], s[R];
    constexpr TwConst() : c(), s() {
        for (int m = 0; m < R; m++) {
            double x = 2.0 * kPi * m / R;
            if (x > kPi) x -= 2.0 * kPi;
            c[m] = (float)ce_cos(x);
            s[m] = (float)ce_sin(x);
        }
    }
};

template <int A, int B> struct DftCT {
    __device__ static void run(float2* a, float sg) {
        constexpr TwConst<A * B> W{};
        float2 y[B][A];
#pragma unroll
        for (int n2 = 0; n2 < B; n2++) {
            float2 t[A];
#pragma unroll
            for (int n1 = 0; n1 < A; n1++) t[n1] = a[B * n1 + n2];
            Dft<A>::run(t, sg);
#pragma unroll
            for (int k1 = 0; k1 < A; k1++) {
                const int m = n2 * k1;
                y[n2][k1] = (m == 0) ? t[k1] : cmul(t[k1], make_float2(W.c[m], sg * W.s[m]));
            }
        }
#pragma unroll
        for (int k1 = 0; k1 < A; k1++) {
            float2 t[B];
#pragma unroll
            for (int n2 = 0; n2 < B; n2++) t[n2] = y[n2][k1];
            Dft<B>::run(t, sg);
#pragma unroll
            for (int k2 = 0; k2 < B; k2++) a[k1 + A * k2] = t[k2];
        }
    }
};
template <> struct Dft<6> : DftCT<2, 3> {};
template <> struct Dft<10> : DftCT<2, 5> {};
template <> struct Dft<12> : DftCT<4, 3> {};

// One Stockham stage of radix R over L lines of length N held transposed
// in LDS ([n][line]).  j in [0, N/R): inputs at rows j + r*N/R, twiddle
// w_(Ns R)^(j%Ns * r), outputs at rows (j/Ns)*Ns*R + j%Ns + r*Ns.
template <int TILE, int NT, int R>
__device__ void stage(float2* lds, int L, int N, int Ns, const float2* __restrict__ tw, float sg) {
    constexpr int MAXT = (TILE / NT + R - 1) / R;
    float2 v[MAXT][R];
    const int nb = N / R;
    const int ntask = L * nb;
    const int tstride = N / (Ns * R);
#pragma unroll
    for (int i = 0; i < MAXT; i++) {
        int t = threadIdx.x + i * NT;
        if (t < ntask) {
            int j = t / L, line = t - j * L;
#pragma unroll
            for (int r = 0; r < R; r++) v[i][r] = lds[(j + r * nb) * L + line];
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < MAXT; i++) {
        int t = threadIdx.x + i * NT;
        if (t < ntask) {
            int j = t / L, line = t - j * L;
            int k = j % Ns;
            if (Ns > 1) {
#pragma unroll
                for (int r = 1; r < R; r++) {
                    float2 w = tw[k * r * tstride];
                    w.y *= -sg;  // table holds exp(-2 pi i t / N); conj for the inverse
                    v[i][r] = cmul(v[i][r], w);
                }
            }
            Dft<R>::run(v[i], sg);
            int base = (j / Ns) * Ns * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) lds[(base + r * Ns) * L + line] = v[i][r];
        }
    }
    __syncthreads();
}

// MODE 0: plain pass; 1: forward prologue (input = img .* KFC);
// 2: adjoint epilogue with coil sum; 3: adjoint epilogue per coil.
template <int TILE, int NT, int MODE>
__global__ void __launch_bounds__(NT) fft_pass_kernel(PassArgs a, const float2* __restrict__ twg) {
    // tiles up to 32 KiB also stage the twiddle table in LDS (in front of
    // the tile); the 64 KiB fused tiles read it from global memory
    constexpr bool TWL = TILE <= 4096;
    extern __shared__ float2 smem[];
    const int N = a.N;
    float2* lds = smem + (TWL ? N : 0);
    const float2* tw = twg;
    if (TWL) {
        for (int i = threadIdx.x; i < N; i += NT) smem[i] = twg[i];
        tw = smem;  // visible after the barrier that ends the load phase
    }
    const int64_t tile_w = blockIdx.x % a.ntile_w;
    const int64_t to = blockIdx.x / a.ntile_w;
    const int64_t w0 = tile_w * a.W;
    const int Wt = (int)((a.S - w0) < a.W ? (a.S - w0) : a.W);
    const int64_t o0 = to * a.Bo;
    const int Bt = (int)((a.outer - o0) < a.Bo ? (a.outer - o0) : a.Bo);
    const int L = Bt * Wt;

    // rows outside the input window are zero
    const int nzero = (N - a.in_cnt) * L;
    for (int e = threadIdx.x; e < nzero; e += NT) {
        int row = e / L, line = e - row * L;
        int n = row < a.in_off ? row : row + a.in_cnt;
        lds[n * L + line] = make_float2(0.f, 0.f);
    }
    // load the input window (coalesced along w), pre-sign (-1)^n.  All of a
    // work-item's loads are issued before any LDS write so that they are in
    // flight together.
    const int nin = a.in_cnt * L;
    constexpr int EPT = TILE / NT;
    if (MODE == 0 && Bt == 1 && (NT % Wt) == 0) {
        // one outer index, Wt divides the workgroup: w is fixed per work-item
        const int w = threadIdx.x % Wt;
        const int rstep = NT / Wt;
        const int r0 = threadIdx.x / Wt;
        const float2* src = a.in + line_off(o0, a.in_os, a.in_on, a.in_os2) + w0 + w;
        float2 v[EPT];
#pragma unroll
        for (int i = 0; i < EPT; i++) {
            int r = r0 + i * rstep;
            if (r < a.in_cnt) v[i] = src[(int64_t)r * a.in_rs];
        }
#pragma unroll
        for (int i = 0; i < EPT; i++) {
            int r = r0 + i * rstep;
            if (r < a.in_cnt) {
                int n = a.in_off + r;
                float2 x = v[i];
                if (a.presign && (n & 1)) { x.x = -x.x; x.y = -x.y; }
                lds[n * L + w] = x;
            }
        }
    } else {
        float2 v[EPT];
#pragma unroll
        for (int i = 0; i < EPT; i++) {
            int e = threadIdx.x + i * NT;
            if (e < nin) {
                int w = e % Wt;
                int rest = e / Wt;
                int r = rest % a.in_cnt;
                int ob = rest / a.in_cnt;
                int64_t o = o0 + ob;
                if (MODE == 1) {
                    int64_t ii = o * a.in_cnt + r;
                    v[i] = cmul(a.img[ii], a.kfc[(ii % a.kfc_period) * a.S + w0 + w]);
                } else {
                    v[i] = a.in[line_off(o, a.in_os, a.in_on, a.in_os2) + (int64_t)r * a.in_rs + w0 + w];
                }
            }
        }
#pragma unroll
        for (int i = 0; i < EPT; i++) {
            int e = threadIdx.x + i * NT;
            if (e < nin) {
                int w = e % Wt;
                int rest = e / Wt;
                int r = rest % a.in_cnt;
                int ob = rest / a.in_cnt;
                int n = a.in_off + r;
                float2 x = v[i];
                if (a.presign && (n & 1)) { x.x = -x.x; x.y = -x.y; }
                lds[n * L + ob * Wt + w] = x;
            }
        }
    }
    __syncthreads();

    const float sg = a.inverse ? 1.0f : -1.0f;
    int Ns = 1;
    for (int s = 0; s < a.nstage; s++) {
        int R = a.radix[s];
        switch (R) {
            case 16: stage<TILE, NT, 16>(lds, L, N, Ns, tw, sg); break;
            case 8: stage<TILE, NT, 8>(lds, L, N, Ns, tw, sg); break;
            case 5: stage<TILE, NT, 5>(lds, L, N, Ns, tw, sg); break;
            case 4: stage<TILE, NT, 4>(lds, L, N, Ns, tw, sg); break;
            case 3: stage<TILE, NT, 3>(lds, L, N, Ns, tw, sg); break;
            default: stage<TILE, NT, 2>(lds, L, N, Ns, tw, sg); break;
        }
        Ns *= R;
    }

    // store the output window, post-sign (-1)^(k + N/2), scale
    const int half = N / 2;
    if (MODE == 2) {
        // one work-item per (outer, kept row): sum over the coils of the line
        const int nout = a.out_cnt * Bt;
        for (int e = threadIdx.x; e < nout; e += NT) {
            int kl = e % a.out_cnt, ob = e / a.out_cnt;
            int k = a.out_off + kl;
            float sc = a.scale;
            if (a.postsign && ((k + half) & 1)) sc = -sc;
            int64_t i = (o0 + ob) * a.out_cnt + kl;
            const float2* kc = a.kfc + (i % a.kfc_period) * a.S;
            float2 acc = make_float2(0.f, 0.f);
            for (int c = 0; c < Wt; c++) {
                float2 v = lds[k * L + ob * Wt + c];
                v = make_float2(v.x * sc, v.y * sc);
                float2 p = cmul(v, kc[c]);
                acc.x += p.x;
                acc.y += p.y;
            }
            a.img_out[i] = make_float2(acc.x * a.epi_scale, acc.y * a.epi_scale);
        }
        return;
    }
    const int nout = a.out_cnt * L;
    for (int e = threadIdx.x; e < nout; e += NT) {
        int w = e % Wt;
        int rest = e / Wt;
        int kl = rest % a.out_cnt;
        int ob = rest / a.out_cnt;
        int k = a.out_off + kl;
        float sc = a.scale;
        if (a.postsign && ((k + half) & 1)) sc = -sc;
        float2 v = lds[k * L + ob * Wt + w];
        v = make_float2(v.x * sc, v.y * sc);
        if (MODE == 3) {
            int64_t i = (o0 + ob) * a.out_cnt + kl;
            float2 p = cmul(v, a.kfc[(i % a.kfc_period) * a.S + w0 + w]);
            a.out[i * a.S + w0 + w] = make_float2(p.x * a.epi_scale, p.y * a.epi_scale);
        } else {
            a.out[line_off(o0 + ob, a.out_os, a.out_on, a.out_os2) + (int64_t)kl * a.out_rs + w0 + w] = v;
        }
    }
}

// Bits b in [0, NB) such that row j + STRIDE b lies in [lo, hi).
template <int STRIDE, int NB>
__device__ __forceinline__ uint32_t row_bits(int lo, int hi, int j) {
    const int b0 = lo > j ? (lo - j + STRIDE - 1) / STRIDE : 0;
    const int b1 = hi > j ? min((hi - j + STRIDE - 1) / STRIDE, NB) : 0;
    return b1 > b0 ? ((1u << b1) - 1u) & ~((1u << b0) - 1u) : 0u;
}

// Register-resident pass for N = 16 * N2 (N2 <= 16): the four-step split
// n = n2 + N2 n1, k = k1 + 16 k2.  A tile holds L = Bo * W lines, 16
// work-items per line.  Work-item (line, j): step 1 (j < N2) loads rows
// j + N2 n1 of its line straight into registers (rows outside the input
// window are zero), runs the 16-point DFT over n1 and multiplies by
// w_N^(j k1); one LDS exchange ([k1][n2][line], planes padded so that the
// reads of a wave spread over all banks); step 2 runs the N2-point DFT over
// n2 for k1 = j and stores rows j + 16 k2 straight from registers.  A wave's
// loads and stores cover runs of consecutive lines (inner positions) in
// consecutive rows.  MODE as for the LDS kernel; the coil sum of MODE 2 goes
// through a second LDS pass and adds the coils in order.
// SPLIT: the exchange goes through LDS in two halves of eight planes
// (k1 < 8, then k1 >= 8), so a tile needs half the exchange space and more
// tiles fit on a CU.
// DIR: 0 direction from a.inverse; 1 forward, 2 inverse as a compile-time
// constant (the plain passes), which folds the signs of the DFT butterflies
// and of the twiddles into the arithmetic.
// WIN: 1 when the input window is exactly the central half, rows [N/4,
// 3N/4) (the zero-filled forward passes): a work-item's rows j + N2 n1 are
// then inside only for n1 in [4, 12), the others are zero at compile time
// and step 1 runs the half-input DFT.  2 when the output window is exactly
// the central half (the cropping adjoint passes, N2 a multiple of 4): rows
// j + 16 k2 are kept only for k2 in [N2/4, 3 N2/4), and the outputs of the
// other k2 are never computed.  0: windows from the arguments alone.
template <int N2, int MODE, class TI, class TO, bool SPLIT = false, int DIR = 0, int WIN = 0>
__global__ void __launch_bounds__(1024) fft_reg_kernel(PassArgs a, const float2* __restrict__ twg) {
    static_assert(WIN != 2 || N2 % 4 == 0, "central output window needs N2 % 4 == 0");
    // compile-time window of the rows of step 1 (n1) and step 2 (k2)
    constexpr auto in_win = [](int n1) { return WIN != 1 || (n1 >= 4 && n1 < 12); };
    constexpr auto out_win = [](int k2) { return WIN != 2 || (k2 >= N2 / 4 && k2 < 3 * N2 / 4); };
    const TI* __restrict__ in = reinterpret_cast<const TI*>(a.in);
    TO* __restrict__ out = reinterpret_cast<TO*>(a.out);
    constexpr int N1 = 16;
    constexpr int N = N1 * N2;
    extern __shared__ float2 smem[];
    float2* tw = smem;      // N twiddles exp(-2 pi i t / N)
    float2* ex = smem + N;  // exchange planes
    const int W = a.W, L = a.W * a.Bo, XS = a.xs;
    const int t = threadIdx.x;
    const int j = (int)(__umul24((unsigned)t, a.lmag) >> 20);  // t / L
    const int line = t - j * L;
    const int ob = (int)(__umul24((unsigned)line, a.wmag) >> 20), w = line - ob * W;
    const int64_t bid = xcd_block(blockIdx.x, gridDim.x, a.xg);
    const int64_t to = div_m(bid, a.ntile_w, a.twmag), tile_w = bid - to * a.ntile_w;
    const int64_t o = to * a.Bo + ob;
    const int64_t ws = tile_w * W + w;
    const bool live = o < a.outer && ws < a.S;
    const int64_t ibase = line_off_m(o, a.in_os, a.in_on, a.in_os2, a.inmag);
    const int64_t obase = line_off_m(o, a.out_os, a.out_on, a.out_os2, a.outmag);
    const float sg = DIR == 1 ? -1.0f : DIR == 2 ? 1.0f : (a.inverse ? 1.0f : -1.0f);
    // column of the masks and intervals: o * mask_cpl + ws / mask_nch
    const int64_t mcol =
        o * a.mask_cpl + (a.nchmag ? (int64_t)__umulhi((uint32_t)ws, a.nchmag) : ws);

    // interval passes: this work-item's column keeps rows [ilo, ihi); a dead
    // column (-1, -1) is neither loaded nor stored.  The interval is loaded
    // here, in flight with the mask words below.
    int2 civ = make_int2(0, N);
    {
        const int2* iv = MODE == 0 ? (a.in_ival ? a.in_ival : a.out_ival) : nullptr;
        if (iv && live) civ = iv[mcol];
    }
    // masked passes: the bits of this work-item's column, reduced at once to
    // the bits of the rows it loads (j + N2 n1, bit n1 of lbits) and stores
    // (j + N1 k2, bit k2 of sbits), with compile-time word indices only (an
    // indexed mask array went to scratch memory, whose loads wait for every
    // load in flight)
    constexpr int MW = (N + 31) / 32;
    uint32_t lbits = ~0u, sbits = ~0u;
    const uint32_t* mask = MODE == 0 ? (a.in_mask ? a.in_mask : a.out_mask) : nullptr;
    if (mask && live) {
        // a column's words are MW * 4 bytes at a multiple of that: vector loads
        const uint32_t* mp = mask + mcol * MW;
        uint32_t mw[MW];
        if constexpr (MW % 4 == 0) {
#pragma unroll
            for (int i = 0; i < MW / 4; i++) {
                const uint4 q = reinterpret_cast<const uint4*>(mp)[i];
                mw[4 * i] = q.x;
                mw[4 * i + 1] = q.y;
                mw[4 * i + 2] = q.z;
                mw[4 * i + 3] = q.w;
            }
        } else if constexpr (MW % 2 == 0) {
#pragma unroll
            for (int i = 0; i < MW / 2; i++) {
                const uint2 q = reinterpret_cast<const uint2*>(mp)[i];
                mw[2 * i] = q.x;
                mw[2 * i + 1] = q.y;
            }
        } else {
#pragma unroll
            for (int i = 0; i < MW; i++) mw[i] = mp[i];
        }
        // row j + N2 n1 (j < N2 <= 16; rows of j >= N2 are never loaded)
        // lies in word (N2 n1) / 32 or the next one, at most word MW - 1;
        // row j + 16 k2 (j < 16) in word k2 / 2.  Only the side that has
        // the mask takes its bits.
        if (a.in_mask) {
            lbits = 0u;
#pragma unroll
            for (int n1 = 0; n1 < N1; n1++) {
                if (!in_win(n1)) continue;
                const int row = j + N2 * n1;
                const int w0 = (N2 * n1) >> 5, w1 = (N2 * n1 + N2 - 1) >> 5;
                const uint32_t wv = (w1 != w0 && row >= 32 * w1) ? mw[w1] : mw[w0];
                lbits |= ((wv >> (row & 31)) & 1u) << n1;
            }
        }
        if (a.out_mask) {
            sbits = 0u;
#pragma unroll
            for (int k2 = 0; k2 < N2; k2++) {
                if (!out_win(k2)) continue;
                const int row = j + N1 * k2;
                sbits |= ((mw[k2 >> 1] >> (row & 31)) & 1u) << k2;
            }
        }
    }
    const int ilo = civ.x, ihi = civ.y;
    const bool dead = civ.x < 0;
    // Rows this work-item loads (j + N2 n1, bit n1 of lmask) and stores
    // (j + N1 k2, bit k2 of smask): the input / output window, the column's
    // interval and its cell mask, intersected once here so that the row
    // loops below only test a bit.
    uint32_t lmask = 0u, smask = 0u;
    if (live && !dead) {
        int lo = a.in_off, hi = a.in_off + a.in_cnt;
        if (MODE == 0 && a.in_ival) {
            lo = max(lo, ilo);
            hi = min(hi, ihi);
        }
        if (j < N2) lmask = row_bits<N2, N1>(lo, hi, j);
        if (MODE == 0 && a.in_mask) lmask &= lbits;
        lo = a.out_off;
        hi = a.out_off + a.out_cnt;
        if (MODE == 0 && a.out_ival) {
            lo = max(lo, ilo);
            hi = min(hi, ihi);
        }
        if (j < N1) smask = row_bits<N1, N2>(lo, hi, j);
        if (MODE == 0 && a.out_mask) smask &= sbits;
    }
    // coil-map pixel of row r of this line: kin + r (MODE 1) / kout + r
    // (MODE 2, 3).  The maps repeat every kfc_period pixels, a whole number
    // of lines (checked on the host), so kin + r stays below kfc_period.
    const int64_t kin = MODE == 1 ? (o * a.in_cnt) % a.kfc_period : 0;
    const int64_t kout = MODE >= 2 ? (o * a.out_cnt) % a.kfc_period : 0;
    float2 v[N1];
    if (j < N2) {
        // element offsets of row j - in_off (r = 0 at n1 = 0 would be row
        // in_off); row j + N2 n1 is n1 steps of N2 rows further
        const int64_t r0 = j - a.in_off;
        const int64_t ioff = ibase + ws + r0 * a.in_rs, istep = (int64_t)N2 * a.in_rs;
        const int64_t koff = (kin + r0) * a.S + ws, kstep = (int64_t)N2 * a.S;
        const int64_t goff = o * a.in_cnt + r0;
#pragma unroll
        for (int n1 = 0; n1 < N1; n1++) {
            float2 x = make_float2(0.f, 0.f);
            if (in_win(n1) && ((lmask >> n1) & 1u)) {
                if (MODE == 1) x = cmul(a.img[goff + N2 * n1], a.kfc[koff + n1 * kstep]);
                else x = ld_c(in + ioff + n1 * istep);
            }
            v[n1] = x;
        }
    }
    // MODE 2: the KFC_conj values of this work-item's kept output rows are
    // loaded now, in flight with the input, instead of after the transform
    float2 kf[MODE == 2 ? N2 : 1];
    if constexpr (MODE == 2) {
        const int64_t koff = (kout + j - a.out_off) * a.S + ws, kstep = (int64_t)N1 * a.S;
#pragma unroll
        for (int k2 = 0; k2 < N2; k2++) {
            kf[k2] = make_float2(0.f, 0.f);
            if (out_win(k2) && ((smask >> k2) & 1u)) kf[k2] = a.kfc[koff + k2 * kstep];
        }
    }
    // plain passes: a tile none of whose work-items stores a row (every line
    // dead, out of range or with an empty interval) leaves after the first
    // barrier.  Each wave records whether any of its work-items stores; all
    // work-items read the same records, so the whole tile leaves together.
    __shared__ uint32_t wstore[16];
    if constexpr (MODE == 0) {
        const int any = __any(smask != 0u);
        if ((t & 63) == 0) wstore[t >> 6] = any ? 1u : 0u;
    }
    for (int i = t; i < N; i += blockDim.x) tw[i] = twg[i];
    __syncthreads();
    if constexpr (MODE == 0) {
        uint32_t st = 0u;
        const int nw = ((int)blockDim.x + 63) >> 6;
        for (int i = 0; i < nw; i++) st |= wstore[i];
        if (st == 0u) return;
    }
    if (j < N2) {
        // a work-item that loaded no row holds zeros, which it passes on
        // untransformed (a wave all of whose work-items did so skips the
        // arithmetic)
        if (lmask != 0u) {
            if (a.presign) {
#pragma unroll
                for (int n1 = 0; n1 < N1; n1++)
                    if (in_win(n1) && ((j + N2 * n1) & 1)) v[n1] = make_float2(-v[n1].x, -v[n1].y);
            }
            if constexpr (WIN == 1) Dft<N1>::run_mid(v, sg);
            else Dft<N1>::run(v, sg);
#pragma unroll
            for (int k1 = 1; k1 < N1; k1++) {
                float2 wv = tw[j * k1];
                wv.y *= -sg;  // table holds exp(-2 pi i t / N); conj for the inverse
                v[k1] = cmul(v[k1], wv);
            }
        }
        if constexpr (!SPLIT) {
#pragma unroll
            for (int k1 = 0; k1 < N1; k1++) ex[k1 * XS + j * L + line] = v[k1];
        } else {
#pragma unroll
            for (int k1 = 0; k1 < N1 / 2; k1++) ex[k1 * XS + j * L + line] = v[k1];
        }
    }
    __syncthreads();
    float2 u[N2];
    if constexpr (!SPLIT) {
#pragma unroll
        for (int n2 = 0; n2 < N2; n2++) u[n2] = ex[j * XS + n2 * L + line];
    } else {
        // work-item j reads plane k1 = j: the first half now, the second
        // after the planes k1 >= 8 replace them
        if (j < N1 / 2) {
#pragma unroll
            for (int n2 = 0; n2 < N2; n2++) u[n2] = ex[j * XS + n2 * L + line];
        }
        __syncthreads();
        if (j < N2) {
#pragma unroll
            for (int k1 = N1 / 2; k1 < N1; k1++) ex[(k1 - N1 / 2) * XS + j * L + line] = v[k1];
        }
        __syncthreads();
        if (j >= N1 / 2) {
#pragma unroll
            for (int n2 = 0; n2 < N2; n2++) u[n2] = ex[(j - N1 / 2) * XS + n2 * L + line];
        }
    }
    // only work-items that store a row need step 2's outputs
    if (smask != 0u) Dft<N2>::run(u, sg);
    const int half = N / 2;
    if (MODE == 2) {
        // products with KFC_conj into LDS as [ob][kept row][coil] (odd
        // stride), then one work-item per (ob, kept row) adds the coils
        const int Wp = W | 1;
        __syncthreads();
        if (live) {
            const float sc = (a.postsign && ((j + half) & 1)) ? -a.scale : a.scale;
#pragma unroll
            for (int k2 = 0; k2 < N2; k2++) {
                if (out_win(k2) && ((smask >> k2) & 1u)) {
                    const int kl = j + N1 * k2 - a.out_off;
                    const float2 y = make_float2(u[k2].x * sc, u[k2].y * sc);
                    ex[(ob * a.out_cnt + kl) * Wp + w] = cmul(y, kf[k2]);
                }
            }
        }
        __syncthreads();
        const int nred = a.Bo * a.out_cnt;
        for (int e = t; e < nred; e += blockDim.x) {
            const int obb = e / a.out_cnt, kl = e - obb * a.out_cnt;
            const int64_t oo = to * a.Bo + obb;
            if (oo < a.outer) {
                float2 acc = make_float2(0.f, 0.f);
                for (int c = 0; c < W; c++) {
                    const float2 p = ex[(obb * a.out_cnt + kl) * Wp + c];
                    acc.x += p.x;
                    acc.y += p.y;
                }
                a.img_out[oo * a.out_cnt + kl] = make_float2(acc.x * a.epi_scale, acc.y * a.epi_scale);
            }
        }
        return;
    }
    if (!live || dead) return;
    // postsign: (k + N/2) is odd for every k = j + 16 k2 of this work-item
    // or for none (16 is even)
    const float sc = (a.postsign && ((j + half) & 1)) ? -a.scale : a.scale;
    const int64_t r0 = j - a.out_off;
    const int64_t ooff = obase + ws + r0 * a.out_rs, ostep = (int64_t)N1 * a.out_rs;
    const int64_t koff = (kout + r0) * a.S + ws, kstep = (int64_t)N1 * a.S;
    const int64_t goff = (o * a.out_cnt + r0) * a.S + ws;
#pragma unroll
    for (int k2 = 0; k2 < N2; k2++) {
        if (out_win(k2) && ((smask >> k2) & 1u)) {
            const float2 y = make_float2(u[k2].x * sc, u[k2].y * sc);
            if (MODE == 3) {
                const float2 p = cmul(y, a.kfc[koff + k2 * kstep]);
                st_c(out + goff + k2 * kstep, make_float2(p.x * a.epi_scale, p.y * a.epi_scale));
            } else {
                st_c(out + ooff + k2 * ostep, y);
            }
        }
    }
}

struct TwiddleCache {
    std::mutex mu;
    std::map<int, std::unique_ptr<DevBuf>> tables;
    const float2* get(int N) {
        std::lock_guard<std::mutex> lk(mu);
        auto it = tables.find(N);
        if (it != tables.end()) return it->second->as<float2>();
        std::vector<float2> h(N);
        for (int t = 0; t < N; t++) {
            double ang = -2.0 * M_PI * (double)t / (double)N;
            h[t] = make_float2((float)std::cos(ang), (float)std::sin(ang));
        }
        auto buf = std::make_unique<DevBuf>((size_t)N * sizeof(float2));
        MONA_HIP(hipMemcpy(buf->p, h.data(), (size_t)N * sizeof(float2), hipMemcpyHostToDevice));
        const float2* p = buf->as<float2>();
        tables[N] = std::move(buf);
        return p;
    }
};

TwiddleCache& twiddles() {
    static TwiddleCache c;
    return c;
}

// Tile of the fused (prologue / epilogue) passes: a whole [N][nCh] line
// group must fit, so it is the largest one (64 KiB, 512 work-items).
constexpr int kFusedTile = 8192;
// Tile of the plain passes: 2048, 4096 or 8192 complex values
// (MONA_FFT_PLAIN_TILE; default 4096 = 32 KiB, 256 work-items, several
// workgroups per CU).
int plain_tile() {
    static int t = [] {
        const char* e = std::getenv("MONA_FFT_PLAIN_TILE");
        int v = e ? std::atoi(e) : 4096;
        return (v == 2048 || v == 8192) ? v : 4096;
    }();
    return t;
}

template <int TILE, int NT>
void launch(const PassArgs& a, int mode, unsigned nblocks, size_t lds, const float2* tw, hipStream_t s) {
    switch (mode) {
        case 1: fft_pass_kernel<TILE, NT, 1><<<nblocks, NT, lds, s>>>(a, tw); break;
        case 2: fft_pass_kernel<TILE, NT, 2><<<nblocks, NT, lds, s>>>(a, tw); break;
        case 3: fft_pass_kernel<TILE, NT, 3><<<nblocks, NT, lds, s>>>(a, tw); break;
        default: fft_pass_kernel<TILE, NT, 0><<<nblocks, NT, lds, s>>>(a, tw); break;
    }
}

// Two-half exchange of the plain register passes (MONA_FFT_SPLIT=0 turns
// it off).
bool split_exchange() {
    static const bool on = [] {
        const char* e = std::getenv("MONA_FFT_SPLIT");
        return e ? std::atoi(e) != 0 : true;
    }();
    return on;
}

// Register kernel: lines per tile on strided axes (16, 32 or 64;
// MONA_FFT_LINES, default 32); -1 disables the register kernel.
int reg_lines_default() {
    static int t = [] {
        const char* e = std::getenv("MONA_FFT_LINES");
        int v = e ? std::atoi(e) : 32;
        return (v == -1 || v == 16 || v == 64) ? v : 32;
    }();
    return t;
}

// N = 16 * N2 with a register DFT of length N2.
int reg_n2(int N) {
    if (N % 16) return 0;
    const int n2 = N / 16;
    switch (n2) {
        case 2: case 3: case 4: case 5: case 6: case 8: case 10: case 12: case 16: return n2;
        default: return 0;
    }
}

template <int N2, int MODE, class TI, class TO, bool SPLIT = false, int DIR = 0, int WIN = 0>
void launch_reg_dir(const PassArgs& a, unsigned nblocks, int nt, size_t lds, const float2* tw, hipStream_t s) {
    auto kern = fft_reg_kernel<N2, MODE, TI, TO, SPLIT, DIR, WIN>;
    // the kernel's static LDS (16 words of wave records) counts too
    if (lds + 16 * sizeof(uint32_t) > 65536) {
        static std::once_flag once;
        std::call_once(once, [&] {
            (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                      hipFuncAttributeMaxDynamicSharedMemorySize,
                                      160 * 1024 - (int)(16 * sizeof(uint32_t)));
        });
    }
    kern<<<nblocks, nt, lds, s>>>(a, tw);
}

// plain passes with the direction as a template argument; passes whose
// input (forward, x .* KFC) or output (adjoint, coil sum) window is exactly
// the central half with that window at compile time (N2 a multiple of 4)
// (MONA_FFT_WIN=0 keeps the windows at run time)
bool fft_win_on() {
    static const bool on = [] {
        const char* e = std::getenv("MONA_FFT_WIN");
        return e ? std::atoi(e) != 0 : true;
    }();
    return on;
}

template <int N2, int MODE, class TI, class TO, bool SPLIT = false>
void launch_reg_mode(const PassArgs& a, unsigned nblocks, int nt, size_t lds, const float2* tw, hipStream_t s) {
    constexpr int N = 16 * N2;
    constexpr bool wins = N2 % 4 == 0;
    const bool on = wins && fft_win_on();
    const bool cin = on && a.in_off == N / 4 && a.in_cnt == N / 2;
    const bool cout = on && a.out_off == N / 4 && a.out_cnt == N / 2;
    if constexpr (MODE == 0) {
        if (a.inverse) {
            if constexpr (wins)
                if (cout) return launch_reg_dir<N2, 0, TI, TO, SPLIT, 2, 2>(a, nblocks, nt, lds, tw, s);
            launch_reg_dir<N2, 0, TI, TO, SPLIT, 2>(a, nblocks, nt, lds, tw, s);
        } else {
            if constexpr (wins)
                if (cin) return launch_reg_dir<N2, 0, TI, TO, SPLIT, 1, 1>(a, nblocks, nt, lds, tw, s);
            launch_reg_dir<N2, 0, TI, TO, SPLIT, 1>(a, nblocks, nt, lds, tw, s);
        }
    } else if constexpr (MODE == 1) {
        if constexpr (wins)
            if (cin) return launch_reg_dir<N2, 1, TI, TO, SPLIT, 0, 1>(a, nblocks, nt, lds, tw, s);
        launch_reg_dir<N2, 1, TI, TO, SPLIT, 0>(a, nblocks, nt, lds, tw, s);
    } else if constexpr (MODE == 2) {
        if constexpr (wins)
            if (cout) return launch_reg_dir<N2, 2, TI, TO, SPLIT, 0, 2>(a, nblocks, nt, lds, tw, s);
        launch_reg_dir<N2, 2, TI, TO, SPLIT, 0>(a, nblocks, nt, lds, tw, s);
    } else {
        launch_reg_dir<N2, MODE, TI, TO, SPLIT, 0>(a, nblocks, nt, lds, tw, s);
    }
}

// Storage variants: float2 throughout, or bf16x2 for the grid-side arrays
// (MODE 1 writes, MODE 0 reads and writes, MODE 2 reads bf16x2).
template <int N2, bool SPLIT>
void launch_reg_split(const PassArgs& a, int mode, bool bf_in, bool bf_out, unsigned nblocks, int nt, size_t lds,
                      const float2* tw, hipStream_t s) {
    if (bf_in || bf_out) {
        switch (mode) {
            case 1:
                MONA_CHECK(bf_out, MONA_ERR_INTERNAL, "bad bf16 pass");
                launch_reg_mode<N2, 1, float2, bf16x2, SPLIT>(a, nblocks, nt, lds, tw, s);
                break;
            case 2:
                MONA_CHECK(bf_in, MONA_ERR_INTERNAL, "bad bf16 pass");
                launch_reg_mode<N2, 2, bf16x2, float2, SPLIT>(a, nblocks, nt, lds, tw, s);
                break;
            case 0:
                if (bf_in && bf_out) launch_reg_mode<N2, 0, bf16x2, bf16x2, SPLIT>(a, nblocks, nt, lds, tw, s);
                else if (bf_in) launch_reg_mode<N2, 0, bf16x2, float2, SPLIT>(a, nblocks, nt, lds, tw, s);
                else launch_reg_mode<N2, 0, float2, bf16x2, SPLIT>(a, nblocks, nt, lds, tw, s);
                break;
            default: MONA_CHECK(false, MONA_ERR_INTERNAL, "bf16 storage has no per-coil epilogue");
        }
        return;
    }
    switch (mode) {
        case 1: launch_reg_mode<N2, 1, float2, float2, SPLIT>(a, nblocks, nt, lds, tw, s); break;
        case 2: launch_reg_mode<N2, 2, float2, float2, SPLIT>(a, nblocks, nt, lds, tw, s); break;
        case 3: launch_reg_mode<N2, 3, float2, float2, SPLIT>(a, nblocks, nt, lds, tw, s); break;
        default: launch_reg_mode<N2, 0, float2, float2, SPLIT>(a, nblocks, nt, lds, tw, s); break;
    }
}

template <int N2>
void launch_reg(const PassArgs& a, int mode, bool bf_in, bool bf_out, unsigned nblocks, int nt, size_t lds,
                const float2* tw, hipStream_t s, bool split) {
    if (split) launch_reg_split<N2, true>(a, mode, bf_in, bf_out, nblocks, nt, lds, tw, s);
    else launch_reg_split<N2, false>(a, mode, bf_in, bf_out, nblocks, nt, lds, tw, s);
}

}  // namespace

bool fft_size_supported(int N) {
    if (N < 2 || N > 2048) return false;
    int n = N;
    for (int f : {2, 3, 5})
        while (n % f == 0) n /= f;
    return n == 1;
}

bool fft_reg_supported(int N) { return reg_n2(N) != 0; }

bool fft_plain_pass_is_reg(int N, int lines) {
    const int l = lines != 0 ? lines : reg_lines_default();
    return reg_n2(N) != 0 && l > 0;
}

bool fft_pass_fusable(int N, int64_t S) { return (int64_t)N * S <= kFusedTile || (reg_n2(N) && S <= 64); }

namespace {
void fft_pass_reg(const AxisPass& p, int n2, int lines, hipStream_t s) {
    const int N = p.N;
    PassArgs a{};
    a.N = N;
    a.S = p.S;
    a.in_rs = p.in_rs > 0 ? p.in_rs : p.S;
    a.out_rs = p.out_rs > 0 ? p.out_rs : p.S;
    a.in_on = p.in_on;
    a.in_os2 = p.in_os2;
    a.out_on = p.out_on;
    a.out_os2 = p.out_os2;
    a.outer = p.outer;
    if (p.S <= 64) {
        // whole inner dimension in a tile, >= 16 lines where the outer
        // dimension allows
        a.W = (int)p.S;
        int bo = (int)std::max<int64_t>(1, (16 + p.S - 1) / p.S);
        while (bo > 1 && bo * p.S > 64) bo--;
        a.Bo = (int)std::min<int64_t>(bo, p.outer);
    } else {
        a.W = lines;
        a.Bo = 1;
    }
    const int L = a.W * a.Bo;
    // exchange plane stride: n2 * L rounded so that consecutive planes
    // start 128 bytes apart modulo the 256-byte bank row
    int xs = n2 * L;
    xs += ((16 - xs % 32) + 32) % 32;
    a.xs = xs;
    a.ntile_w = (p.S + a.W - 1) / a.W;
    a.in = p.in;
    a.in_os = p.in_os;
    a.in_off = p.in_off;
    a.in_cnt = p.in_cnt;
    a.out = p.out;
    a.out_os = p.out_os;
    a.out_off = p.out_off;
    a.out_cnt = p.out_cnt;
    a.inverse = p.inverse;
    a.presign = p.presign;
    a.postsign = p.postsign;
    a.scale = p.scale;
    a.epi_scale = p.epi_scale;
    a.img = p.img;
    a.kfc = p.kfc;
    a.kfc_period = p.kfc_period > 0 ? p.kfc_period : INT64_MAX;
    // the register kernel adds the row to a line's first coil-map pixel
    // without wrapping: a period must hold whole lines
    MONA_CHECK(p.kfc_period <= 0 || (p.mode == 1 ? p.kfc_period % p.in_cnt : p.kfc_period % p.out_cnt) == 0,
               MONA_ERR_INTERNAL, "coil-map period is not a whole number of lines");
    a.img_out = p.img_out;
    a.in_mask = p.in_mask;
    a.out_mask = p.out_mask;
    a.mask_nch = p.mask_nch;
    a.mask_cpl = p.S / std::max(p.mask_nch, 1);
    a.in_ival = p.in_ival;
    a.out_ival = p.out_ival;
    const bool limited = p.in_mask || p.out_mask || p.in_ival || p.out_ival;
    MONA_CHECK(!limited || (p.mode == 0 && p.mask_nch >= 1 && p.S % p.mask_nch == 0), MONA_ERR_INTERNAL,
               "cell masks / intervals need a plain pass and S divisible by mask_nch");
    // exact quotients in the kernel: t / L and line / W for t < 1024
    // (2^20 scale), inner position / mask_nch below 2^32 / mask_nch
    a.lmag = (uint32_t)(((1u << 20) + L - 1) / L);
    a.wmag = (uint32_t)(((1u << 20) + a.W - 1) / a.W);
    a.nchmag = 0;
    if (limited && p.mask_nch > 1) {
        MONA_CHECK((uint64_t)p.S * (uint64_t)p.mask_nch < (1ull << 32), MONA_ERR_INTERNAL,
                   "mask columns exceed the 32-bit quotient");
        a.nchmag = (uint32_t)(((1ull << 32) + p.mask_nch - 1) / p.mask_nch);
    }
    const int64_t ntile_o = (p.outer + a.Bo - 1) / a.Bo;
    const int64_t nblocks = a.ntile_w * ntile_o;
    MONA_CHECK(nblocks < 0x7fffffff, MONA_ERR_ARG, "FFT grid too large");
    // multiply-high quotients where exact: n / d with n d < 2^32 for every
    // numerator of the launch (blocks, or lines up to ntile_o Bo)
    auto magic = [](int64_t d, int64_t nmax) -> uint32_t {
        if (d <= 1 || nmax * d >= ((int64_t)1 << 32)) return 0u;
        return (uint32_t)((((int64_t)1 << 32) + d - 1) / d);
    };
    a.twmag = magic(a.ntile_w, nblocks);
    // plain passes: runs of 4 consecutive tiles per XCD (neighbouring row
    // segments share an L2; C1 z passes 2-3% faster, the fused x passes
    // not), MONA_XCD_GROUP_FFT overrides it
    static const int fft_xg = [] {
        const char* e = std::getenv("MONA_XCD_GROUP_FFT");
        return e ? std::max(1, std::atoi(e)) : 0;
    }();
    a.xg = fft_xg ? fft_xg : (p.mode == 0 ? 4 : 1);
    a.inmag = magic(a.in_on, ntile_o * a.Bo);
    a.outmag = magic(a.out_on, ntile_o * a.Bo);
    const int nt = 16 * L;
    MONA_CHECK(nt <= 1024 && (p.mode != 2 || a.W == p.S), MONA_ERR_INTERNAL, "bad register FFT tile");
    const size_t coil_sum = p.mode == 2 ? (size_t)a.Bo * p.out_cnt * (a.W | 1) : 0;
    size_t ex = std::max<size_t>((size_t)16 * xs, coil_sum);
    // passes whose exchange needs more than 20 KiB (fewer than eight tiles
    // on a CU) exchange in two halves (MONA_FFT_SPLIT_KB sets the bound);
    // the coil-sum epilogue (mode 2) exchanges in one piece: its two extra
    // barriers cost more than the added tiles per CU gain
    static const size_t split_bytes = [] {
        const char* e = std::getenv("MONA_FFT_SPLIT_KB");
        return (size_t)(e ? std::atoi(e) : 20) * 1024;
    }();
    const bool split = split_exchange() && p.mode != 2 && ((size_t)N + ex) * sizeof(float2) > split_bytes;
    if (split) ex = std::max<size_t>((size_t)8 * xs, coil_sum);
    size_t lds = ((size_t)N + ex) * sizeof(float2);
    MONA_CHECK(lds + 16 * sizeof(uint32_t) <= 160 * 1024, MONA_ERR_INTERNAL, "register FFT tile exceeds the LDS");
    const float2* tw = twiddles().get(N);
    switch (n2) {
        case 2: launch_reg<2>(a, p.mode, p.bf_in, p.bf_out, (unsigned)nblocks, nt, lds, tw, s, split); break;
        case 3: launch_reg<3>(a, p.mode, p.bf_in, p.bf_out, (unsigned)nblocks, nt, lds, tw, s, split); break;
        case 4: launch_reg<4>(a, p.mode, p.bf_in, p.bf_out, (unsigned)nblocks, nt, lds, tw, s, split); break;
        case 5: launch_reg<5>(a, p.mode, p.bf_in, p.bf_out, (unsigned)nblocks, nt, lds, tw, s, split); break;
        case 6: launch_reg<6>(a, p.mode, p.bf_in, p.bf_out, (unsigned)nblocks, nt, lds, tw, s, split); break;
        case 8: launch_reg<8>(a, p.mode, p.bf_in, p.bf_out, (unsigned)nblocks, nt, lds, tw, s, split); break;
        case 10: launch_reg<10>(a, p.mode, p.bf_in, p.bf_out, (unsigned)nblocks, nt, lds, tw, s, split); break;
        case 12: launch_reg<12>(a, p.mode, p.bf_in, p.bf_out, (unsigned)nblocks, nt, lds, tw, s, split); break;
        default: launch_reg<16>(a, p.mode, p.bf_in, p.bf_out, (unsigned)nblocks, nt, lds, tw, s, split); break;
    }
    MONA_HIP(hipGetLastError());
}
}  // namespace

void fft_pass(const AxisPass& p, hipStream_t s) {
    const int N = p.N;
    MONA_CHECK(fft_size_supported(N), MONA_ERR_ARG,
               "FFT length " + std::to_string(N) + " is not 2^a 3^b 5^c <= 2048");
    MONA_CHECK(!(p.presign || p.postsign) || N % 2 == 0, MONA_ERR_ARG,
               "N_u must have all components even for the Fourier transform.");
    MONA_CHECK(p.in_off >= 0 && p.in_cnt >= 1 && p.in_off + p.in_cnt <= N, MONA_ERR_INTERNAL, "bad input window");
    MONA_CHECK(p.out_off >= 0 && p.out_cnt >= 1 && p.out_off + p.out_cnt <= N, MONA_ERR_INTERNAL,
               "bad output window");
    MONA_CHECK(p.S >= 1 && p.outer >= 1, MONA_ERR_INTERNAL, "empty FFT pass");
    const int TILE = p.mode != 0 ? kFusedTile
                     : (p.tile == 2048 || p.tile == 4096 || p.tile == 8192) ? p.tile : plain_tile();
    const int lines = p.lines != 0 ? p.lines : reg_lines_default();
    const int n2 = reg_n2(N);
    const bool old_fits = (int64_t)N * p.S <= kFusedTile;
    if (p.bf_in || p.bf_out) {
        MONA_CHECK(n2 && (p.mode == 0 || p.S <= 64), MONA_ERR_ARG,
                   "bf16 storage needs FFT lengths 16 * {2,3,4,5,6,8,10,12,16} and <= 64 coils");
        // 4-byte elements: a pass writing bf16 takes 32 lines (128-byte row
        // segments; 64 lines: 27% slower at C1, 16 lines: 26% slower),
        // a pass reading bf16 16 lines (32: 12% slower, 64: 49% slower)
        fft_pass_reg(p, n2, p.lines > 0 ? p.lines : (p.bf_in ? 16 : 32), s);
        return;
    }
    if (n2 && ((p.mode == 0 && lines > 0) || (p.mode != 0 && p.S <= 64 && (lines > 0 || !old_fits)))) {
        fft_pass_reg(p, n2, lines > 0 ? lines : 16, s);
        return;
    }
    MONA_CHECK(!(p.in_mask || p.out_mask || p.in_ival || p.out_ival), MONA_ERR_INTERNAL,
               "cell masks and intervals need the register FFT kernel");
    MONA_CHECK(p.mode == 0 || old_fits, MONA_ERR_INTERNAL, "fused FFT pass does not fit a tile");
    PassArgs a{};
    a.N = N;
    int n = N, ns = 0;
    while (n % 16 == 0) { a.radix[ns++] = 16; n /= 16; }
    while (n % 8 == 0 && n > 4) { a.radix[ns++] = 8; n /= 8; }
    while (n % 4 == 0) { a.radix[ns++] = 4; n /= 4; }
    while (n % 2 == 0) { a.radix[ns++] = 2; n /= 2; }
    while (n % 3 == 0) { a.radix[ns++] = 3; n /= 3; }
    while (n % 5 == 0) { a.radix[ns++] = 5; n /= 5; }
    a.nstage = ns;
    a.S = p.S;
    a.in_rs = p.in_rs > 0 ? p.in_rs : p.S;
    a.out_rs = p.out_rs > 0 ? p.out_rs : p.S;
    a.in_on = p.in_on;
    a.in_os2 = p.in_os2;
    a.out_on = p.out_on;
    a.out_os2 = p.out_os2;
    a.outer = p.outer;
    const int Lmax = TILE / N;
    if (p.S <= Lmax) {
        a.W = (int)p.S;
        a.Bo = Lmax / (int)p.S;
        if (p.mode == 2) a.Bo = 1;
    } else {
        MONA_CHECK(p.mode == 0, MONA_ERR_INTERNAL, "fused FFT pass needs the whole inner dimension in one tile");
        int w = 1;
        while (w * 2 <= Lmax && w * 2 <= 64) w *= 2;
        a.W = w;
        a.Bo = 1;
    }
    if (a.Bo > a.outer) a.Bo = (int)a.outer;
    a.ntile_w = (p.S + a.W - 1) / a.W;
    a.in = p.in;
    a.in_os = p.in_os;
    a.in_off = p.in_off;
    a.in_cnt = p.in_cnt;
    a.out = p.out;
    a.out_os = p.out_os;
    a.out_off = p.out_off;
    a.out_cnt = p.out_cnt;
    a.inverse = p.inverse;
    a.presign = p.presign;
    a.postsign = p.postsign;
    a.scale = p.scale;
    a.epi_scale = p.epi_scale;
    a.img = p.img;
    a.kfc = p.kfc;
    a.kfc_period = p.kfc_period > 0 ? p.kfc_period : INT64_MAX;
    a.img_out = p.img_out;
    int64_t ntile_o = (p.outer + a.Bo - 1) / a.Bo;
    int64_t nblocks = a.ntile_w * ntile_o;
    MONA_CHECK(nblocks < 0x7fffffff, MONA_ERR_ARG, "FFT grid too large");
    size_t lds = (size_t)a.W * a.Bo * N * sizeof(float2);
    if (TILE <= 4096) lds += (size_t)N * sizeof(float2);  // twiddles staged in LDS
    MONA_CHECK(lds <= (size_t)(TILE + (TILE <= 4096 ? N : 0)) * sizeof(float2), MONA_ERR_INTERNAL,
               "FFT tile too large");
    const float2* tw = twiddles().get(N);
    if (TILE == 2048) launch<2048, 256>(a, p.mode, (unsigned)nblocks, lds, tw, s);
    else if (TILE == 4096) launch<4096, 256>(a, p.mode, (unsigned)nblocks, lds, tw, s);
    else launch<8192, 512>(a, p.mode, (unsigned)nblocks, lds, tw, s);
    MONA_HIP(hipGetLastError());
}

void fft_axis(const float2* in, float2* out, int64_t outer, int N, int64_t S, bool inverse,
              bool presign, bool postsign, float scale, hipStream_t s) {
    AxisPass p;
    p.N = N;
    p.S = S;
    p.outer = outer;
    p.in = in;
    p.in_os = (int64_t)N * S;
    p.in_off = 0;
    p.in_cnt = N;
    p.out = out;
    p.out_os = (int64_t)N * S;
    p.out_off = 0;
    p.out_cnt = N;
    p.inverse = inverse;
    p.presign = presign;
    p.postsign = postsign;
    p.scale = scale;
    fft_pass(p, s);
}

void cfft_nd(float2* data, int imDim, const int* N, int64_t nBatch, bool inverse, hipStream_t s) {
    int64_t inner = nBatch;
    int64_t total = nBatch;
    for (int d = 0; d < imDim; d++) total *= N[d];
    for (int d = 0; d < imDim; d++) {
        int64_t outer = total / (inner * N[d]);
        fft_axis(data, data, outer, N[d], inner, inverse, true, true, inverse ? 1.0f / (float)N[d] : 1.0f, s);
        inner *= N[d];
    }
}

}  // namespace mona

extern "C" int mona_cfft(int imDim, const int* N, int nBatch, void* data, int inverse, hipStream_t stream) {
    return mona::guarded([&] {
        MONA_CHECK(imDim >= 1 && imDim <= 3 && N && data && nBatch >= 1, MONA_ERR_ARG, "bad argument");
        mona::cfft_nd(static_cast<float2*>(data), imDim, N, nBatch, inverse != 0, stream);
    });
}
