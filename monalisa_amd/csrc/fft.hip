// Centred FFT on the oversampled grid: one LDS-staged Stockham pass per
// axis, mixed radix 8/5/4/3/2, batched over everything else.
//
// The centring of bmShanna_MATLAB.m:46-50 / bmNakatsha_MATLAB.m:36-40,
// fftshift(fft(ifftshift(x))) along an axis of even length N, equals
//   Y[k] = (-1)^(k + N/2) * sum_n (-1)^n x[n] e^(-+2 pi i n k / N)
// so each pass multiplies by (-1)^n when it loads a line and by
// (-1)^(k + N/2) (times 1/N for the inverse) when it stores it; no shift
// copies are made.
//
// A workgroup owns a tile of L lines ((L+1)*N <= 4096 complex values, 32 KiB
// of LDS with the twiddle table): W consecutive positions of the contiguous inner dimension times Bo
// outer indices, so every global load/store row is a run of W float2.  The
// Stockham stages run in LDS with each work-item holding its butterflies'
// inputs in registers across one barrier (in-place autosort).
#include <map>
#include <memory>

#include "mona_internal.hpp"
#include "fft.hpp"

namespace mona {

namespace {

constexpr int kFftThreads = 256;
constexpr int kFftMaxElems = 4096;  // twiddles + tile, per workgroup (32 KiB)
constexpr int kMaxStages = 16;

struct FftArgs {
    int N;
    int nstage;
    int radix[kMaxStages];
    int64_t S;        // inner stride (float2 elements)
    int64_t outer;    // number of outer indices
    int W;            // tile width along the inner dimension
    int Bo;           // outer indices per tile
    int64_t ntile_w;  // tiles along the inner dimension
    int inverse;
    int presign;
    int postsign;
    float scale;
};

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

// One Stockham stage of radix R over L lines of length N held in LDS with
// leading dimension LD.  j in [0, N/R): inputs at j + r*N/R, twiddle
// w_(Ns R)^(j%Ns * r), outputs at (j/Ns)*Ns*R + j%Ns + r*Ns.
template <int R>
__device__ void stage(float2* lds, int L, int N, int LD, int Ns, const float2* tw, float sg) {
    constexpr int MAXT = (kFftMaxElems / kFftThreads + R - 1) / R;
    float2 v[MAXT][R];
    const int nb = N / R;
    const int ntask = L * nb;
    const int tstride = N / (Ns * R);
#pragma unroll
    for (int i = 0; i < MAXT; i++) {
        int t = threadIdx.x + i * kFftThreads;
        if (t < ntask) {
            int line = t / nb, j = t - line * nb;
            const float2* src = lds + line * LD + j;
#pragma unroll
            for (int r = 0; r < R; r++) v[i][r] = src[r * nb];
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < MAXT; i++) {
        int t = threadIdx.x + i * kFftThreads;
        if (t < ntask) {
            int line = t / nb, j = t - line * nb;
            int k = j % Ns;
            if (Ns > 1) {
#pragma unroll
                for (int r = 1; r < R; r++) {
                    float2 w = tw[k * r * tstride];
                    w.y *= -sg;  // table holds exp(-2 pi i t / N); conj for inverse
                    v[i][r] = cmul(v[i][r], w);
                }
            }
            Dft<R>::run(v[i], sg);
            float2* dst = lds + line * LD + (j / Ns) * Ns * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) dst[r * Ns] = v[i][r];
        }
    }
    __syncthreads();
}

__global__ void __launch_bounds__(kFftThreads)
fft_axis_kernel(const float2* in, float2* out, FftArgs a,  // in == out allowed (no restrict)
                const float2* __restrict__ twg) {
    extern __shared__ float2 smem[];
    const int N = a.N;
    float2* tw = smem;            // N twiddles
    float2* lds = smem + N;       // tile
    for (int i = threadIdx.x; i < N; i += kFftThreads) tw[i] = twg[i];

    int64_t tw_idx = blockIdx.x % a.ntile_w;
    int64_t to = blockIdx.x / a.ntile_w;
    int64_t w0 = tw_idx * a.W;
    int Wt = (int)((a.S - w0) < a.W ? (a.S - w0) : a.W);
    int64_t o0 = to * a.Bo;
    int Bt = (int)((a.outer - o0) < a.Bo ? (a.outer - o0) : a.Bo);
    const int L = Bt * Wt;
    const int LD = N + 1;  // odd pitch: the line-strided load/store phases avoid bank conflicts
    const int64_t lineN = (int64_t)N * a.S;

    // load (coalesced along w), pre-sign (-1)^n
    const int tot = L * N;
    for (int e = threadIdx.x; e < tot; e += kFftThreads) {
        int w = e % Wt;
        int rest = e / Wt;
        int n = rest % N;
        int ob = rest / N;
        int64_t g = (o0 + ob) * lineN + (int64_t)n * a.S + w0 + w;
        float2 v = in[g];
        if (a.presign && (n & 1)) { v.x = -v.x; v.y = -v.y; }
        lds[(ob * Wt + w) * LD + n] = v;
    }
    __syncthreads();

    const float sg = a.inverse ? 1.0f : -1.0f;
    int Ns = 1;
    for (int s = 0; s < a.nstage; s++) {
        int R = a.radix[s];
        switch (R) {
            case 8: stage<8>(lds, L, N, LD, Ns, tw, sg); break;
            case 5: stage<5>(lds, L, N, LD, Ns, tw, sg); break;
            case 4: stage<4>(lds, L, N, LD, Ns, tw, sg); break;
            case 3: stage<3>(lds, L, N, LD, Ns, tw, sg); break;
            default: stage<2>(lds, L, N, LD, Ns, tw, sg); break;
        }
        Ns *= R;
    }

    // store, post-sign (-1)^(k + N/2), scale
    const int half = N / 2;
    for (int e = threadIdx.x; e < tot; e += kFftThreads) {
        int w = e % Wt;
        int rest = e / Wt;
        int k = rest % N;
        int ob = rest / N;
        int64_t g = (o0 + ob) * lineN + (int64_t)k * a.S + w0 + w;
        float2 v = lds[(ob * Wt + w) * LD + k];
        float sc = a.scale;
        if (a.postsign && ((k + half) & 1)) sc = -sc;
        out[g] = make_float2(v.x * sc, v.y * sc);
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

}  // namespace

bool fft_size_supported(int N) {
    if (N < 2 || 2 * N + 1 > kFftMaxElems) return false;
    int n = N;
    for (int f : {2, 3, 5})
        while (n % f == 0) n /= f;
    return n == 1;
}

void fft_axis(const float2* in, float2* out, int64_t outer, int N, int64_t S, bool inverse,
              bool presign, bool postsign, float scale, hipStream_t s) {
    MONA_CHECK(fft_size_supported(N), MONA_ERR_ARG,
               "FFT length " + std::to_string(N) + " is not 2^a 3^b 5^c <= 2048");
    MONA_CHECK(!(presign || postsign) || N % 2 == 0, MONA_ERR_ARG,
               "N_u must have all components even for the Fourier transform.");
    FftArgs a{};
    a.N = N;
    int n = N, ns = 0;
    while (n % 8 == 0 && n > 4) { a.radix[ns++] = 8; n /= 8; }
    while (n % 4 == 0) { a.radix[ns++] = 4; n /= 4; }
    while (n % 2 == 0) { a.radix[ns++] = 2; n /= 2; }
    while (n % 3 == 0) { a.radix[ns++] = 3; n /= 3; }
    while (n % 5 == 0) { a.radix[ns++] = 5; n /= 5; }
    a.nstage = ns;
    a.S = S;
    a.outer = outer;
    int Lmax = (kFftMaxElems - N) / (N + 1);
    if (Lmax < 1) Lmax = 1;
    if (S >= 16 || S >= Lmax) {
        a.W = (int)(S < Lmax ? S : Lmax);
        a.Bo = 1;
    } else {
        a.W = (int)S;
        a.Bo = Lmax / (int)S;
        if (a.Bo < 1) a.Bo = 1;
    }
    a.ntile_w = (S + a.W - 1) / a.W;
    a.inverse = inverse;
    a.presign = presign;
    a.postsign = postsign;
    a.scale = scale;
    int64_t ntile_o = (outer + a.Bo - 1) / a.Bo;
    int64_t nblocks = a.ntile_w * ntile_o;
    MONA_CHECK(nblocks < 0x7fffffff, MONA_ERR_ARG, "FFT grid too large");
    size_t lds = (size_t)(N + (size_t)a.W * a.Bo * (N + 1)) * sizeof(float2);
    MONA_CHECK(lds <= (size_t)kFftMaxElems * sizeof(float2), MONA_ERR_INTERNAL, "FFT tile too large");
    const float2* tw = twiddles().get(N);
    fft_axis_kernel<<<(unsigned)nblocks, kFftThreads, lds, s>>>(in, out, a, tw);
    MONA_HIP(hipGetLastError());
}

void cfft_nd(float2* data, int imDim, const int* N, int64_t nBatch, bool inverse, hipStream_t s) {
    int64_t inner = nBatch;
    int64_t total = nBatch;
    for (int d = 0; d < imDim; d++) total *= N[d];
    for (int d = 0; d < imDim; d++) {
        int64_t outer = total / (inner * N[d]);
        fft_axis(data, data, outer, N[d], inner, inverse, true, true,
                 inverse ? 1.0f / (float)N[d] : 1.0f, s);
        inner *= N[d];
    }
}

}  // namespace mona

extern "C" int mona_cfft(int imDim, const int* N, int nBatch, void* data, int inverse,
                         hipStream_t stream) {
    return mona::guarded([&] {
        MONA_CHECK(imDim >= 1 && imDim <= 3 && N && data && nBatch >= 1, MONA_ERR_ARG, "bad argument");
        mona::cfft_nd(static_cast<float2*>(data), imDim, N, nBatch, inverse != 0, stream);
    });
}
