// Matrix-free gridding: the forward gridding y = Gu * grid computed from
// the trajectory instead of the stored entries of Gu.
//
// A point's Gu row is w_j / sum_k w_k over its nWin^imDim neighbour cells
// (bmTraj2SparseMat.m:166-281, odd nWin: cells round(t) + {-h..h} per axis,
// wrapped), with w the radial kernel of the distance d:
//   Kaiser   I0(alpha sqrt(max(1 - d^2/tau^2, 0))) / I0(alpha),
//   Gauss    exp(-d^2 / (2 sigma^2)) / (sqrt(2 pi) sigma).
// The constant factors cancel in the normalisation, and I0(alpha sqrt(v))
// = sum_k (alpha^2 v / 4)^k / (k!)^2 is a power series in v = 1 - d^2/tau^2,
// so the Kaiser weight needs no square root or Bessel call: a fixed-length
// Horner evaluation of the series (24 terms, relative truncation below
// 1e-10 for alpha <= 12, checked when the plan is made).  1/sum_k w_k comes
// per point from the host (float64).  Products are float32 with FMA; the
// result differs from the stored-matrix product by rounding only (about
// 1e-7 relative), well inside the 1e-5 gridding tolerance.
//
// Work layout: a block takes consecutive points of the frame's tile order
// (points sorted by the tile, then the raster position, of their base
// cell).  It first computes every (point, neighbour) weight and cell index
// into LDS, then each point's neighbour cells are gathered by nCh/CPL lanes,
// CPL coils each (16-byte loads of float2 pairs, or four bf16 coils).
// Consecutive blocks run in groups of 16 per XCD so that neighbouring points
// share an L2.
#include "gridmf.hpp"

namespace mona {

namespace {

__device__ __forceinline__ float mf_weight(float s, const MfKernel& k) {
    if (k.kind == MONA_KERNEL_KAISER) {
        const float v = fmaxf(1.0f - s * k.inv, 0.0f);
        float r = k.c[kMfTerms - 1];
#pragma unroll
        for (int i = kMfTerms - 2; i >= 0; i--) r = fmaf(r, v, k.c[i]);
        return r;
    }
    return exp2f(-s * k.inv);
}

__device__ __forceinline__ int wrap_axis(int a, int N) { return a < 0 ? a + N : (a >= N ? a - N : a); }

// CPL coils of one cell: float2 storage, CPL = 1 or 2
template <int CPL>
__device__ __forceinline__ void ld_cells(const float2* p, float* re, float* im) {
    if constexpr (CPL == 2) {
        const float4 a = *reinterpret_cast<const float4*>(p);
        re[0] = a.x; im[0] = a.y; re[1] = a.z; im[1] = a.w;
    } else {
        const float2 a = *p;
        re[0] = a.x; im[0] = a.y;
    }
}
// bf16x2 storage, CPL = 1, 2 or 4
template <int CPL>
__device__ __forceinline__ void ld_cells(const bf16x2* p, float* re, float* im) {
    uint32_t u[4];
    if constexpr (CPL == 4) {
        const uint4 a = *reinterpret_cast<const uint4*>(p);
        u[0] = a.x; u[1] = a.y; u[2] = a.z; u[3] = a.w;
    } else if constexpr (CPL == 2) {
        const uint2 a = *reinterpret_cast<const uint2*>(p);
        u[0] = a.x; u[1] = a.y;
    } else {
        u[0] = p->v;
    }
#pragma unroll
    for (int c = 0; c < CPL; c++) {
        const bf16x2 b{u[c]};
        const float2 f = ld_c(&b);
        re[c] = f.x;
        im[c] = f.y;
    }
}
template <int CPL>
__device__ __forceinline__ void st_cells(float2* p, const float* re, const float* im) {
    if constexpr (CPL == 2) *reinterpret_cast<float4*>(p) = make_float4(re[0], im[0], re[1], im[1]);
    else if constexpr (CPL == 4) {
        reinterpret_cast<float4*>(p)[0] = make_float4(re[0], im[0], re[1], im[1]);
        reinterpret_cast<float4*>(p)[1] = make_float4(re[2], im[2], re[3], im[3]);
    } else *p = make_float2(re[0], im[0]);
}
template <int CPL>
__device__ __forceinline__ void st_cells(bf16x2* p, const float* re, const float* im) {
#pragma unroll
    for (int c = 0; c < CPL; c++) st_c(p + c, make_float2(re[c], im[c]));
}

// Points of a block: kMfBlock threads, ppb = kMfBlock / (nCh / CPL) points.
template <int D, int CPL, class TG, class TY>
__global__ void __launch_bounds__(kMfBlock)
u2n_mf_kernel(const uint32_t* __restrict__ base, const float4* __restrict__ rec,
              const uint16_t* __restrict__ frame_of, int P, const TG* __restrict__ grid, int64_t frame_cells,
              TY* __restrict__ y, int nCh, MfGeo g, MfKernel k, int ppb, int xg) {
    constexpr int NB = D == 3 ? 27 : 9;
    __shared__ float sw[kMfMaxPts * NB];
    __shared__ int soff[kMfMaxPts * NB];
    __shared__ int sfr[kMfMaxPts];
    const int64_t blk = xcd_block(blockIdx.x, gridDim.x, xg);
    const int64_t p0 = blk * ppb;
    for (int t = threadIdx.x; t < ppb * NB; t += kMfBlock) {
        const int i = t / NB, j = t - i * NB;
        const int64_t p = p0 + i;
        float w = 0.0f;
        int off = 0;
        if (p < P) {
            const uint32_t b = base[p];
            const float4 r = rec[p];
            const int cx = j % 3 - 1, cy = (j / 3) % 3 - 1, cz = D == 3 ? j / 9 - 1 : 0;
            const float dx = r.x - cx, dy = r.y - cy, dz = r.z - cz;
            w = mf_weight(dx * dx + dy * dy + dz * dz, k) * r.w;
            const int x = wrap_axis((int)(b & 1023u) + cx, g.N0);
            const int yy = wrap_axis((int)((b >> 10) & 1023u) + cy, g.N1);
            const int z = D == 3 ? wrap_axis((int)(b >> 20) + cz, g.N2) : 0;
            off = x + g.sy * yy + g.sz * z;
            // a deleted point (no Gu row) reads nothing: -1 marks it
            if (j == 0) sfr[i] = r.w == 0.0f ? -1 : (frame_of ? (int)frame_of[p] : 0);
        } else if (j == 0) {
            sfr[i] = -1;
        }
        sw[t] = w;
        soff[t] = off;
    }
    __syncthreads();
    const int tpp = nCh / CPL;
    const int i = threadIdx.x / tpp, q = threadIdx.x - i * tpp;
    const int64_t p = p0 + i;
    if (i >= ppb || p >= P) return;
    float ar[CPL], ai[CPL];
#pragma unroll
    for (int c = 0; c < CPL; c++) ar[c] = ai[c] = 0.0f;
    if (sfr[i] < 0) {
        st_cells<CPL>(y + p * nCh + q * CPL, ar, ai);
        return;
    }
    const TG* gb = grid + (int64_t)sfr[i] * frame_cells * nCh + q * CPL;
    // nine neighbour loads in flight (independent addresses), then their sums
#pragma unroll
    for (int j0 = 0; j0 < NB; j0 += 9) {
        float vr[9][CPL], vi[9][CPL];
#pragma unroll
        for (int j = 0; j < 9; j++) ld_cells<CPL>(gb + (int64_t)soff[i * NB + j0 + j] * nCh, vr[j], vi[j]);
#pragma unroll
        for (int j = 0; j < 9; j++) {
            const float w = sw[i * NB + j0 + j];
#pragma unroll
            for (int c = 0; c < CPL; c++) {
                ar[c] = fmaf(w, vr[j][c], ar[c]);
                ai[c] = fmaf(w, vi[j][c], ai[c]);
            }
        }
    }
    st_cells<CPL>(y + p * nCh + q * CPL, ar, ai);
}

// x-run form of u2n_mf_kernel: the nCh / CPL lanes of a (point, x offset)
// pair gather that offset's nine (3-D) or three (2-D) cells, so the lanes
// of one point read its three x-neighbours of a (y, z) row together, one
// contiguous run of 3 nCh values (fewer cache lines per load instruction
// than one cell at a time).  The three x partial sums of a point are added
// in x order by its first lane.  Steps: each wave takes ppw points a step.
constexpr int kXrMaxPts = 96;
template <int D, int CPL, class TG, class TY>
__global__ void __launch_bounds__(kMfBlock)
u2n_mf_xrun(const uint32_t* __restrict__ base, const float4* __restrict__ rec, const uint16_t* __restrict__ frame_of,
            int P, const TG* __restrict__ grid, int64_t frame_cells, TY* __restrict__ y, int nCh, MfGeo g,
            MfKernel k, int ppw, int steps, int xg) {
    constexpr int NB = D == 3 ? 27 : 9;
    constexpr int NR = NB / 3;
    __shared__ float sw[kXrMaxPts * NB];
    __shared__ int soff[kXrMaxPts * NB];
    __shared__ int sfr[kXrMaxPts];
    const int ppb = (kMfBlock / 64) * ppw * steps;
    const int64_t p0 = xcd_block(blockIdx.x, gridDim.x, xg) * ppb;
    for (int t = threadIdx.x; t < ppb * NB; t += kMfBlock) {
        const int i = t / NB, j = t - i * NB;
        const int64_t p = p0 + i;
        float w = 0.0f;
        int off = 0;
        if (p < P) {
            const uint32_t b = base[p];
            const float4 r = rec[p];
            const int cx = j % 3 - 1, cy = (j / 3) % 3 - 1, cz = D == 3 ? j / 9 - 1 : 0;
            const float dx = r.x - cx, dy = r.y - cy, dz = r.z - cz;
            w = mf_weight(dx * dx + dy * dy + dz * dz, k) * r.w;
            const int x = wrap_axis((int)(b & 1023u) + cx, g.N0);
            const int yy = wrap_axis((int)((b >> 10) & 1023u) + cy, g.N1);
            const int z = D == 3 ? wrap_axis((int)(b >> 20) + cz, g.N2) : 0;
            off = x + g.sy * yy + g.sz * z;
            // a deleted point (no Gu row) reads nothing: -1 marks it
            if (j == 0) sfr[i] = r.w == 0.0f ? -1 : (frame_of ? (int)frame_of[p] : 0);
        } else if (j == 0) {
            sfr[i] = -1;
        }
        sw[t] = w;
        soff[t] = off;
    }
    __syncthreads();
    const int tpp = nCh / CPL, lpp = 3 * tpp;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int iw = lane / lpp, rl = lane - iw * lpp, dx = rl / tpp, q = rl - dx * tpp;
    for (int st = 0; st < steps; st++) {
        const int i = (st * (kMfBlock / 64) + wv) * ppw + iw;
        const int64_t p = p0 + i;
        const bool on = iw < ppw && p < P;
        float ar[CPL], ai[CPL];
#pragma unroll
        for (int c = 0; c < CPL; c++) ar[c] = ai[c] = 0.0f;
        const int fr = on ? sfr[i] : -1;
        if (fr >= 0) {
            const TG* gb = grid + (int64_t)fr * frame_cells * nCh + q * CPL;
            float vr[NR][CPL], vi[NR][CPL];
#pragma unroll
            for (int jr = 0; jr < NR; jr++) ld_cells<CPL>(gb + (int64_t)soff[i * NB + jr * 3 + dx] * nCh, vr[jr], vi[jr]);
#pragma unroll
            for (int jr = 0; jr < NR; jr++) {
                const float w = sw[i * NB + jr * 3 + dx];
#pragma unroll
                for (int c = 0; c < CPL; c++) {
                    ar[c] = fmaf(w, vr[jr][c], ar[c]);
                    ai[c] = fmaf(w, vi[jr][c], ai[c]);
                }
            }
        }
        // the x offsets 0 and +1 of the point to its first lanes (same wave)
        float br[CPL], bi[CPL], cr[CPL], ci[CPL];
#pragma unroll
        for (int c = 0; c < CPL; c++) {
            br[c] = __shfl(ar[c], lane + tpp, 64);
            bi[c] = __shfl(ai[c], lane + tpp, 64);
            cr[c] = __shfl(ar[c], lane + 2 * tpp, 64);
            ci[c] = __shfl(ai[c], lane + 2 * tpp, 64);
        }
        if (on && dx == 0) {
#pragma unroll
            for (int c = 0; c < CPL; c++) {
                ar[c] = (ar[c] + br[c]) + cr[c];
                ai[c] = (ai[c] + bi[c]) + ci[c];
            }
            st_cells<CPL>(y + p * nCh + q * CPL, ar, ai);
        }
    }
}

template <int D, class TG, class TY>
void launch_u2n(int cpl, const uint32_t* base, const float4* rec, const uint16_t* frame_of, int P, const TG* grid,
                int64_t frame_cells, TY* y, int nCh, const MfGeo& g, const MfKernel& k, hipStream_t s) {
    const int tpp = nCh / cpl;
    // runs of xg consecutive blocks per XCD (MONA_XCD_GROUP_MF, default 16)
    static const int xg = [] {
        const char* e = std::getenv("MONA_XCD_GROUP_MF");
        return e ? std::max(1, std::atoi(e)) : 16;
    }();
    // x-run form when a point's three x offsets fit one wave (MONA_MF_XRUN=0
    // turns it off)
    static const bool xrun = [] {
        const char* e = std::getenv("MONA_MF_XRUN");
        return !(e && std::atoi(e) == 0);
    }();
    if (xrun && 3 * tpp <= 64) {
        const int ppw = 64 / (3 * tpp);
        const int steps = std::max(1, 32 / ((kMfBlock / 64) * ppw));
        const int ppb = (kMfBlock / 64) * ppw * steps;
        MONA_CHECK(ppb <= kXrMaxPts, MONA_ERR_INTERNAL, "x-run block too large");
        const unsigned nb = (unsigned)((P + ppb - 1) / ppb);
        switch (cpl) {
            case 4:
                if constexpr (std::is_same<TG, bf16x2>::value)
                    u2n_mf_xrun<D, 4, TG, TY><<<nb, kMfBlock, 0, s>>>(base, rec, frame_of, P, grid, frame_cells, y,
                                                                     nCh, g, k, ppw, steps, xg);
                else
                    MONA_CHECK(false, MONA_ERR_INTERNAL, "four coils per lane is a bf16 layout");
                break;
            case 2:
                u2n_mf_xrun<D, 2, TG, TY><<<nb, kMfBlock, 0, s>>>(base, rec, frame_of, P, grid, frame_cells, y, nCh,
                                                                 g, k, ppw, steps, xg);
                break;
            default:
                u2n_mf_xrun<D, 1, TG, TY><<<nb, kMfBlock, 0, s>>>(base, rec, frame_of, P, grid, frame_cells, y, nCh,
                                                                 g, k, ppw, steps, xg);
        }
        MONA_HIP(hipGetLastError());
        return;
    }
    const int ppb = std::min(kMfMaxPts, kMfBlock / tpp);
    MONA_CHECK(ppb >= 1, MONA_ERR_ARG, "too many coils for the matrix-free gridding");
    const unsigned nb = (unsigned)((P + ppb - 1) / ppb);
    switch (cpl) {
        case 4:
            if constexpr (std::is_same<TG, bf16x2>::value)
                u2n_mf_kernel<D, 4, TG, TY><<<nb, kMfBlock, 0, s>>>(base, rec, frame_of, P, grid, frame_cells, y, nCh,
                                                                   g, k, ppb, xg);
            else
                MONA_CHECK(false, MONA_ERR_INTERNAL, "four coils per lane is a bf16 layout");
            break;
        case 2:
            u2n_mf_kernel<D, 2, TG, TY><<<nb, kMfBlock, 0, s>>>(base, rec, frame_of, P, grid, frame_cells, y, nCh, g,
                                                               k, ppb, xg);
            break;
        default:
            u2n_mf_kernel<D, 1, TG, TY><<<nb, kMfBlock, 0, s>>>(base, rec, frame_of, P, grid, frame_cells, y, nCh, g,
                                                               k, ppb, xg);
    }
    MONA_HIP(hipGetLastError());
}

template <class TG, class TY>
void u2n_mf_typed(const MfPoints& pts, const uint16_t* frame_of, int P, const TG* grid, int64_t frame_cells, TY* y,
                  int nCh, int imDim, const MfGeo& g, const MfKernel& k, hipStream_t s) {
    if (P <= 0) return;
    int cpl = nCh % 2 ? 1 : 2;
    if (std::is_same<TG, bf16x2>::value && nCh % 4 == 0) cpl = 4;
    const uint32_t* b = pts.base.as<uint32_t>();
    const float4* r = pts.rec.as<float4>();
    if (imDim == 3) launch_u2n<3>(cpl, b, r, frame_of, P, grid, frame_cells, y, nCh, g, k, s);
    else launch_u2n<2>(cpl, b, r, frame_of, P, grid, frame_cells, y, nCh, g, k, s);
}

}  // namespace

bool mf_kernel_params(int ktype, int nParam, const double* kp, MfKernel* out) {
    out->kind = ktype;
    for (int i = 0; i < kMfTerms; i++) out->c[i] = 0.0f;
    if (ktype == MONA_KERNEL_KAISER) {
        if (nParam < 3) return false;
        const double tau = (double)(float)kp[0], alpha = (double)(float)kp[1];
        if (!(tau > 0) || !(alpha >= 0)) return false;
        out->inv = (float)(1.0 / (tau * tau));
        // c_k = (alpha^2 / 4)^k / (k!)^2; the terms beyond kMfTerms must be
        // negligible next to the series' value at v = 1 (I0(alpha))
        const double u = alpha * alpha / 4.0;
        double term = 1.0, sum = 0.0;
        for (int i = 0; i < kMfTerms; i++) {
            out->c[i] = (float)term;
            sum += term;
            term *= u / ((double)(i + 1) * (double)(i + 1));
        }
        return term < 1e-10 * sum;
    }
    if (nParam < 2) return false;
    const double sigma = (double)(float)kp[0];
    if (!(sigma > 0)) return false;
    out->inv = (float)(0.5 / (sigma * sigma) * 1.4426950408889634);  // exp(-x) = 2^(-x log2 e)
    return true;
}

void u2n_mf(const MfPoints& pts, const uint16_t* frame_of, int P, const void* grid, bool grid_bf16,
            int64_t frame_cells, void* y, bool y_bf16, int nCh, int imDim, const MfGeo& g, const MfKernel& k,
            hipStream_t s) {
    MONA_CHECK(imDim == 2 || imDim == 3, MONA_ERR_INTERNAL, "matrix-free gridding is 2-D or 3-D");
    if (grid_bf16 && y_bf16)
        u2n_mf_typed(pts, frame_of, P, static_cast<const bf16x2*>(grid), frame_cells, static_cast<bf16x2*>(y), nCh,
                     imDim, g, k, s);
    else if (grid_bf16)
        u2n_mf_typed(pts, frame_of, P, static_cast<const bf16x2*>(grid), frame_cells, static_cast<float2*>(y), nCh,
                     imDim, g, k, s);
    else if (y_bf16)
        u2n_mf_typed(pts, frame_of, P, static_cast<const float2*>(grid), frame_cells, static_cast<bf16x2*>(y), nCh,
                     imDim, g, k, s);
    else
        u2n_mf_typed(pts, frame_of, P, static_cast<const float2*>(grid), frame_cells, static_cast<float2*>(y), nCh,
                     imDim, g, k, s);
}

}  // namespace mona
