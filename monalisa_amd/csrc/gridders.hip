// Direct normalised n2u gridders and the in-box pair count.
//
//   bmGridder_n2u_leight{1,2,3}_mex  Gaussian weights on the fly, clipped box
//                                    (bmGridder_n2u_leight3_mex.cpp:116-358)
//   bmGridder_n2u_mex                precomputed weights, running index jumps
//                                    (bmGridder_n2u_mex.cpp:72-201)
//   bmGriddingMatrix_secret_length*  count of in-box pairs with d^2 < (nWin/2)^2
//                                    (bmGriddingMatrix_secret_length3_mex.cpp:70-222)
//
// The reference scatters point by point.  Here every (point, neighbour)
// entry is generated in the reference's loop order, entries are sorted
// stably by cell (radix sort, integer only), and one work-item per
// (cell, coil) gathers its segment: the weight sum and the weighted data sum
// see the same terms in the same order as the reference's serial scatter,
// with no atomics.  Products and sums are rounded separately (no FMA).
#include <algorithm>
#include <cmath>
#include <vector>

#include "mona_internal.hpp"

#pragma clang fp contract(off)

namespace mona {

void radix_sort_pairs(unsigned* keys, int* vals, int64_t n, int bits, hipStream_t s);

namespace {

struct BoxArgs {
    int imDim, nWin, nh, nnb;  // nnb = nWin^imDim
    int N[3];
    float t_shift, nhf;
    float g1, g2;
    float dmax;
    unsigned sentinel;
};

// Entry e = p * nnb + ((iz * nWin) + iy) * nWin + ix in the reference's
// z-outer / x-inner order.  Out-of-box neighbours get key = sentinel.
template <bool WEIGHT, bool COUNT>
__global__ void k_box_entries(BoxArgs a, const float* __restrict__ tx, const float* __restrict__ ty,
                              const float* __restrict__ tz, const float* __restrict__ Dn, int64_t nent,
                              unsigned* __restrict__ keys, int* __restrict__ vals, float* __restrict__ wts,
                              unsigned long long* __restrict__ count) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int inbox = 0;
    if (e < nent) {
        int64_t p = e / a.nnb;
        int loc = (int)(e - p * a.nnb);
        int ix = loc % a.nWin;
        int iy = a.imDim > 1 ? (loc / a.nWin) % a.nWin : 0;
        int iz = a.imDim > 2 ? loc / (a.nWin * a.nWin) : 0;
        float t[3] = {tx[p], a.imDim > 1 ? ty[p] : 1.0f, a.imDim > 2 ? tz[p] : 1.0f};
        int c[3];
        float r[3];
        for (int d = 0; d < 3; d++) {
            float cf = truncf(t[d] + a.t_shift);  // modff integer part
            r[d] = t[d] - cf;
            c[d] = (int)cf - 1;
        }
        int idx[3] = {ix, iy, iz};
        int n[3];
        float u[3];
        bool ok = true;
        for (int d = 0; d < 3; d++) {
            n[d] = c[d] - a.nh + idx[d];
            u[d] = -a.nhf + (float)idx[d];
            if (d < a.imDim && (n[d] < 0 || n[d] > a.N[d] - 1)) ok = false;
        }
        float dsq = 0.0f;
        if (ok) {
            float dx = r[0] - u[0];
            dsq = dx * dx;
            if (a.imDim > 1) { float dy = r[1] - u[1]; dsq = dsq + dy * dy; }
            if (a.imDim > 2) { float dz = r[2] - u[2]; dsq = dsq + dz * dz; }
        }
        if (COUNT) {
            inbox = (ok && dsq < a.dmax) ? 1 : 0;
        } else {
            unsigned cell = a.sentinel;
            float w = 0.0f;
            if (ok) {
                cell = (unsigned)(n[0] + (int64_t)a.N[0] * ((a.imDim > 1 ? n[1] : 0) +
                                                            (int64_t)a.N[1] * (a.imDim > 2 ? n[2] : 0)));
                if (WEIGHT) {
                    float ex = expf(a.g2 * dsq);
                    float gw = a.g1 * ex;
                    w = gw * Dn[p];
                }
            }
            keys[e] = cell;
            vals[e] = (int)e;
            if (WEIGHT) wts[e] = w;
        }
    }
    if (COUNT) {
        // workgroup reduction of the in-box count, then one integer atomic
        __shared__ int sh[256];
        sh[threadIdx.x] = inbox;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
            __syncthreads();
        }
        if (threadIdx.x == 0 && sh[0]) atomicAdd(count, (unsigned long long)sh[0]);
    }
}

// Precomputed-weight entries: cell = inclusive prefix of the running jumps.
__global__ void k_jump_entries(const int64_t* __restrict__ cellsum, int64_t nent, int64_t Nu,
                               unsigned* __restrict__ keys, int* __restrict__ vals, int* __restrict__ bad) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nent) return;
    int64_t c = cellsum[e];
    if (c < 0 || c >= Nu) { *bad = 1; c = Nu; }
    keys[e] = (unsigned)c;
    vals[e] = (int)e;
}

__global__ void k_count_keys(const unsigned* __restrict__ keys, int64_t n, int* __restrict__ cnt) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) atomicAdd(&cnt[keys[i]], 1);
}

// b[cell][k] = (sum_e w_e a[p_e][k]) / max(sum_e w_e, 1 if < 1e-5), terms in
// sorted (= reference) order.
__global__ void k_gather_norm(const int64_t* __restrict__ start, const int* __restrict__ sorted_e,
                              const float* __restrict__ wts, int nnb, int64_t Nu, int nCh,
                              const float* __restrict__ a_re, const float* __restrict__ a_im,
                              float* __restrict__ b_re, float* __restrict__ b_im) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= Nu * nCh) return;
    int64_t cell = t / nCh;
    int k = (int)(t - cell * nCh);
    float nrm = 0.0f, ar = 0.0f, ai = 0.0f;
    for (int64_t j = start[cell], e = start[cell + 1]; j < e; j++) {
        int en = sorted_e[j];
        int64_t p = en / nnb;
        float w = wts[en];
        nrm = nrm + w;
        float pr = w * a_re[p * nCh + k], pi = w * a_im[p * nCh + k];
        ar = ar + pr;
        ai = ai + pi;
    }
    if (nrm < 1e-5f) nrm = 1.0f;
    b_re[t] = ar / nrm;
    b_im[t] = ai / nrm;
}

int bits_for(int64_t v) {
    int b = 1;
    while (((int64_t)1 << b) <= v) b++;
    return b;
}

BoxArgs box_args(int imDim, int Nx, int Ny, int Nz, int nWin, float kp1) {
    BoxArgs a{};
    a.imDim = imDim;
    a.nWin = nWin;
    if (nWin % 2) { a.t_shift = 0.5f; a.nh = (nWin - 1) / 2; }
    else { a.t_shift = 0.0f; a.nh = nWin / 2 - 1; }
    a.nhf = (float)a.nh;
    a.N[0] = Nx; a.N[1] = Ny; a.N[2] = Nz;
    a.nnb = nWin;
    for (int d = 1; d < imDim; d++) a.nnb *= nWin;
    float sigma = kp1;
    float sigma_square = sigma * sigma;
    a.g1 = 1 / ((float)std::sqrt(2 * 3.141592) * sigma);  // leight3_mex.cpp:191 constants
    a.g2 = -1 / (2 * sigma_square);
    a.dmax = ((float)nWin / 2.0f) * ((float)nWin / 2.0f);
    return a;
}

// Sort entries by cell, build segment starts, gather.
void normalised_gather(unsigned* keys, int* vals, const float* wts, int64_t nent, int nnb, int64_t Nu, int nCh,
                       const float* a_re, const float* a_im, float* b_re, float* b_im, hipStream_t s) {
    radix_sort_pairs(keys, vals, nent, bits_for(Nu), s);  // sentinel Nu sorts last
    DevBuf cnt((size_t)(Nu + 1) * sizeof(int)), start((size_t)(Nu + 2) * sizeof(int64_t));
    fill_i32(cnt.as<int>(), 0, Nu + 1, s);
    k_count_keys<<<blocks_for(nent, 256), 256, 0, s>>>(keys, nent, cnt.as<int>());
    MONA_HIP(hipGetLastError());
    scan_i32_to_i64_exclusive(cnt.as<int>(), start.as<int64_t>(), Nu + 1, s);
    k_gather_norm<<<blocks_for(Nu * nCh, 256), 256, 0, s>>>(start.as<int64_t>(), vals, wts, nnb, Nu, nCh, a_re,
                                                             a_im, b_re, b_im);
    MONA_HIP(hipGetLastError());
    MONA_HIP(hipStreamSynchronize(s));
}

void leight(int imDim, float* aReal, float* aImag, float* bReal, float* bImag, float* tx, float* ty, float* tz,
            float* Dn, int nCh, int nPt, int Nx, int Ny, int Nz, int nWin, float kp1) {
    MONA_CHECK(nCh >= 1 && nPt >= 0 && Nx >= 1 && Ny >= 1 && Nz >= 1 && nWin >= 1 && nWin <= 9, MONA_ERR_ARG,
               "bad gridder arguments");
    HostCallCtx& C = host_ctx();
    std::lock_guard<std::mutex> lk(C.mu);
    C.init();
    hipStream_t s = C.stream;
    BoxArgs a = box_args(imDim, Nx, Ny, Nz, nWin, kp1);
    int64_t Nu = (int64_t)Nx * Ny * Nz;
    MONA_CHECK(Nu < 0x7fffffff, MONA_ERR_ARG, "grid too large");
    a.sentinel = (unsigned)Nu;
    int64_t nent = (int64_t)nPt * a.nnb;
    MONA_CHECK(nent < 0x7fffffff, MONA_ERR_ARG, "too many (point, neighbour) pairs");
    size_t pb = (size_t)std::max(nPt, 1) * sizeof(float);
    DevBuf dtx(pb), dty(pb), dtz(pb), ddn(pb);
    DevBuf dar((size_t)std::max(nPt, 1) * nCh * sizeof(float)), dai((size_t)std::max(nPt, 1) * nCh * sizeof(float));
    DevBuf dbr((size_t)Nu * nCh * sizeof(float)), dbi((size_t)Nu * nCh * sizeof(float));
    DevBuf keys((size_t)std::max<int64_t>(nent, 1) * 4), vals((size_t)std::max<int64_t>(nent, 1) * 4),
        wts((size_t)std::max<int64_t>(nent, 1) * 4);
    if (nPt > 0) {
        MONA_HIP(hipMemcpyAsync(dtx.p, tx, (size_t)nPt * 4, hipMemcpyHostToDevice, s));
        if (imDim > 1) MONA_HIP(hipMemcpyAsync(dty.p, ty, (size_t)nPt * 4, hipMemcpyHostToDevice, s));
        if (imDim > 2) MONA_HIP(hipMemcpyAsync(dtz.p, tz, (size_t)nPt * 4, hipMemcpyHostToDevice, s));
        MONA_HIP(hipMemcpyAsync(ddn.p, Dn, (size_t)nPt * 4, hipMemcpyHostToDevice, s));
        MONA_HIP(hipMemcpyAsync(dar.p, aReal, (size_t)nPt * nCh * 4, hipMemcpyHostToDevice, s));
        MONA_HIP(hipMemcpyAsync(dai.p, aImag, (size_t)nPt * nCh * 4, hipMemcpyHostToDevice, s));
        k_box_entries<true, false><<<blocks_for(nent, 256), 256, 0, s>>>(
            a, dtx.as<float>(), dty.as<float>(), dtz.as<float>(), ddn.as<float>(), nent, keys.as<unsigned>(),
            vals.as<int>(), wts.as<float>(), nullptr);
        MONA_HIP(hipGetLastError());
    }
    normalised_gather(keys.as<unsigned>(), vals.as<int>(), wts.as<float>(), nent, a.nnb, Nu, nCh, dar.as<float>(),
                      dai.as<float>(), dbr.as<float>(), dbi.as<float>(), s);
    MONA_HIP(hipMemcpyAsync(bReal, dbr.p, (size_t)Nu * nCh * 4, hipMemcpyDeviceToHost, s));
    MONA_HIP(hipMemcpyAsync(bImag, dbi.p, (size_t)Nu * nCh * 4, hipMemcpyDeviceToHost, s));
    MONA_HIP(hipStreamSynchronize(s));
}

long long secret_length(int imDim, float* tx, float* ty, float* tz, int nPt, int Nx, int Ny, int Nz, int nWin) {
    MONA_CHECK(nPt >= 0 && Nx >= 1 && Ny >= 1 && Nz >= 1 && nWin >= 1 && nWin <= 9, MONA_ERR_ARG, "bad arguments");
    HostCallCtx& C = host_ctx();
    std::lock_guard<std::mutex> lk(C.mu);
    C.init();
    hipStream_t s = C.stream;
    BoxArgs a = box_args(imDim, Nx, Ny, Nz, nWin, 1.0f);
    int64_t nent = (int64_t)nPt * a.nnb;
    if (nent == 0) return 0;
    size_t pb = (size_t)nPt * sizeof(float);
    DevBuf dtx(pb), dty(pb), dtz(pb), cnt(sizeof(unsigned long long));
    MONA_HIP(hipMemcpyAsync(dtx.p, tx, pb, hipMemcpyHostToDevice, s));
    if (imDim > 1) MONA_HIP(hipMemcpyAsync(dty.p, ty, pb, hipMemcpyHostToDevice, s));
    if (imDim > 2) MONA_HIP(hipMemcpyAsync(dtz.p, tz, pb, hipMemcpyHostToDevice, s));
    MONA_HIP(hipMemsetAsync(cnt.p, 0, sizeof(unsigned long long), s));
    k_box_entries<false, true><<<blocks_for(nent, 256), 256, 0, s>>>(a, dtx.as<float>(), dty.as<float>(),
                                                                      dtz.as<float>(), nullptr, nent, nullptr,
                                                                      nullptr, nullptr,
                                                                      cnt.as<unsigned long long>());
    MONA_HIP(hipGetLastError());
    unsigned long long h = 0;
    MONA_HIP(hipMemcpyAsync(&h, cnt.p, sizeof(h), hipMemcpyDeviceToHost, s));
    MONA_HIP(hipStreamSynchronize(s));
    return (long long)h;
}

}  // namespace
}  // namespace mona

using namespace mona;

extern "C" {

void bmGridder_n2u_leight1_mex(float* aReal, float* aImag, float* bReal, float* bImag, float* tx, float* Dn,
                               int nCh, int nPt, int Nx, int nWin, float kp1, float kp2) {
    (void)kp2;
    int st = guarded([&] { leight(1, aReal, aImag, bReal, bImag, tx, nullptr, nullptr, Dn, nCh, nPt, Nx, 1, 1, nWin, kp1); });
    if (st) fatal("bmGridder_n2u_leight1_mex", st);
}

void bmGridder_n2u_leight2_mex(float* aReal, float* aImag, float* bReal, float* bImag, float* tx, float* ty,
                               float* Dn, int nCh, int nPt, int Nx, int Ny, int nWin, float kp1, float kp2) {
    (void)kp2;
    int st = guarded([&] { leight(2, aReal, aImag, bReal, bImag, tx, ty, nullptr, Dn, nCh, nPt, Nx, Ny, 1, nWin, kp1); });
    if (st) fatal("bmGridder_n2u_leight2_mex", st);
}

void bmGridder_n2u_leight3_mex(float* aReal, float* aImag, float* bReal, float* bImag, float* tx, float* ty,
                               float* tz, float* Dn, int nCh, int nPt, int Nx, int Ny, int Nz, int nWin, float kp1,
                               float kp2) {
    (void)kp2;
    int st = guarded([&] { leight(3, aReal, aImag, bReal, bImag, tx, ty, tz, Dn, nCh, nPt, Nx, Ny, Nz, nWin, kp1); });
    if (st) fatal("bmGridder_n2u_leight3_mex", st);
}

void bmGridder_n2u_mex(float* aReal, float* aImag, float* bReal, float* bImag, float* w, int* nJump, int nNb,
                       int nCh, int nPt, int Nu_tot) {
    int st = guarded([&] {
        MONA_CHECK(nNb >= 1 && nCh >= 1 && nPt >= 0 && Nu_tot >= 1, MONA_ERR_ARG, "bad gridder arguments");
        HostCallCtx& C = host_ctx();
        std::lock_guard<std::mutex> lk(C.mu);
        C.init();
        hipStream_t s = C.stream;
        int64_t nent = (int64_t)nPt * nNb;
        MONA_CHECK(nent < 0x7fffffff, MONA_ERR_ARG, "too many (point, neighbour) pairs");
        int64_t Nu = Nu_tot;
        DevBuf dar((size_t)std::max(nPt, 1) * nCh * 4), dai((size_t)std::max(nPt, 1) * nCh * 4);
        DevBuf dbr((size_t)Nu * nCh * 4), dbi((size_t)Nu * nCh * 4);
        DevBuf dw((size_t)std::max<int64_t>(nent, 1) * 4), dj((size_t)std::max<int64_t>(nent, 1) * 4);
        DevBuf csum((size_t)std::max<int64_t>(nent, 1) * 8);
        DevBuf keys((size_t)std::max<int64_t>(nent, 1) * 4), vals((size_t)std::max<int64_t>(nent, 1) * 4);
        DevBuf bad(sizeof(int));
        MONA_HIP(hipMemsetAsync(bad.p, 0, sizeof(int), s));
        if (nent > 0) {
            MONA_HIP(hipMemcpyAsync(dar.p, aReal, (size_t)nPt * nCh * 4, hipMemcpyHostToDevice, s));
            MONA_HIP(hipMemcpyAsync(dai.p, aImag, (size_t)nPt * nCh * 4, hipMemcpyHostToDevice, s));
            MONA_HIP(hipMemcpyAsync(dw.p, w, (size_t)nent * 4, hipMemcpyHostToDevice, s));
            MONA_HIP(hipMemcpyAsync(dj.p, nJump, (size_t)nent * 4, hipMemcpyHostToDevice, s));
            scan_i32_to_i64_inclusive(dj.as<int>(), csum.as<int64_t>(), nent, s);
            k_jump_entries<<<blocks_for(nent, 256), 256, 0, s>>>(csum.as<int64_t>(), nent, Nu, keys.as<unsigned>(),
                                                                 vals.as<int>(), bad.as<int>());
            MONA_HIP(hipGetLastError());
            int hb = 0;
            MONA_HIP(hipMemcpyAsync(&hb, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
            MONA_HIP(hipStreamSynchronize(s));
            MONA_CHECK(!hb, MONA_ERR_ARG, "nJump decodes to a cell outside [0, Nu_tot)");
        }
        normalised_gather(keys.as<unsigned>(), vals.as<int>(), dw.as<float>(), nent, nNb, Nu, nCh, dar.as<float>(),
                          dai.as<float>(), dbr.as<float>(), dbi.as<float>(), s);
        MONA_HIP(hipMemcpyAsync(bReal, dbr.p, (size_t)Nu * nCh * 4, hipMemcpyDeviceToHost, s));
        MONA_HIP(hipMemcpyAsync(bImag, dbi.p, (size_t)Nu * nCh * 4, hipMemcpyDeviceToHost, s));
        MONA_HIP(hipStreamSynchronize(s));
    });
    if (st) fatal("bmGridder_n2u_mex", st);
}

void bmGriddingMatrix_secret_length1_mex(float* tx, int nCh, int nPt, int Nx, int nWin, long long* out) {
    (void)nCh;
    int st = guarded([&] { *out = secret_length(1, tx, nullptr, nullptr, nPt, Nx, 1, 1, nWin); });
    if (st) fatal("bmGriddingMatrix_secret_length1_mex", st);
}

void bmGriddingMatrix_secret_length2_mex(float* tx, float* ty, int nCh, int nPt, int Nx, int Ny, int nWin,
                                         long long* out) {
    (void)nCh;
    int st = guarded([&] { *out = secret_length(2, tx, ty, nullptr, nPt, Nx, Ny, 1, nWin); });
    if (st) fatal("bmGriddingMatrix_secret_length2_mex", st);
}

void bmGriddingMatrix_secret_length3_mex(float* tx, float* ty, float* tz, int nCh, int nPt, int Nx, int Ny, int Nz,
                                         int nWin, long long* out) {
    (void)nCh;
    int st = guarded([&] { *out = secret_length(3, tx, ty, tz, nPt, Nx, Ny, Nz, nWin); });
    if (st) fatal("bmGriddingMatrix_secret_length3_mex", st);
}

}  // extern "C"
