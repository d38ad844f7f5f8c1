// Trajectory -> gridding matrices Gn, Gu, Gut built on the device, the same
// construction as structure.cpp (src/gridding123/m/bmTraj2SparseMat.m:58-297
// + the conversions of bmSparseMat_matlabSparse2bmSparseMat.m:6-40 and
// bmSparseMat.m:62-188):
//
//   * one work-item per sample computes, in float64 with the same operation
//     sequence as the host builder, the grid position t/dK + fix(N/2+1),
//     the deletion test, floor(t + 0.5) or floor(t), the wrapped neighbour
//     cells and their kernel weights, and keeps its entries sorted by cell
//     (equal cells merged, as sparse() sums them);
//   * Gu rows (samples) are normalised by their ascending-column sums;
//   * Gut / Gn rows (cells) come from a stable radix sort of the entries by
//     cell, so each cell keeps its samples in ascending order; Gn rows are
//     normalised by their sums in that order;
//   * the jump encodings (running deltas, first entry absolute) and the row
//     squeeze are elementwise kernels and scans.
//
// The integer arrays equal the host builder's exactly; the weights go
// through the device's exp / sqrt / Bessel series, so m_val may differ from
// the host's in the last float bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "mona_internal.hpp"

#pragma clang fp contract(off)

namespace mona {

void radix_sort_pairs(unsigned* keys, int* vals, int64_t n, int bits, hipStream_t s);

namespace {

struct TrajArgs {
    int imDim, nw, nNb, ktype;
    int r[9];
    int64_t Ni[3];
    double N[3], dK[3], shift[3], fshift, p[3], I0a;
};

__device__ double dev_i0(double x) {
    // power series sum_k ((x/2)^k / k!)^2
    double q = 0.25 * x * x, term = 1.0, sum = 1.0;
    for (int k = 1; k < 500; k++) {
        term *= q / ((double)k * (double)k);
        sum += term;
        if (term < sum * 1e-17) break;
    }
    return sum;
}

__device__ double dev_weight(const TrajArgs& a, double d) {
    if (a.ktype == MONA_KERNEL_GAUSS) {
        const double s = a.p[0];
        const double r = d / s;
        return exp(-0.5 * r * r) / (sqrt(2.0 * 3.14159265358979323846) * s);
    }
    const double r = d / a.p[0];
    double w = 1.0 - r * r;
    if (w < 0) w = 0;
    w = a.p[1] * sqrt(w);
    return dev_i0(w) / a.I0a;
}

// Entries of sample p, sorted by cell, duplicates merged.  cols / w hold
// nNb slots per sample.
__global__ void k_traj_entries(TrajArgs a, const double* __restrict__ t, const double* __restrict__ ve, int nPt,
                               int* __restrict__ cnt, int* __restrict__ cols, double* __restrict__ w,
                               double* __restrict__ Dn) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nPt) return;
    double tf[3], tr[3];
    bool del = false;
    for (int d = 0; d < a.imDim; d++) {
        const double tg = t[(int64_t)p * a.imDim + d] / a.dK[d] + a.shift[d];
        if (tg < 1 || tg > a.N[d]) del = true;
        tf[d] = floor(tg + a.fshift);
        tr[d] = tg - tf[d];
    }
    if (ve) {
        double dn = ve[p];
        for (int d = 0; d < a.imDim; d++) dn = dn / a.dK[d];
        Dn[p] = dn;
    }
    if (del) {
        cnt[p] = 0;
        return;
    }
    int* cp = cols + (int64_t)p * a.nNb;
    double* wp = w + (int64_t)p * a.nNb;
    int m = 0;
    for (int j = 0; j < a.nNb; j++) {
        const int c[3] = {a.r[j % a.nw], a.imDim > 1 ? a.r[(j / a.nw) % a.nw] : 0,
                          a.imDim > 2 ? a.r[j / (a.nw * a.nw)] : 0};
        double sq = 0.0;
        int64_t lin = 0, stride = 1;
        for (int d = 0; d < a.imDim; d++) {
            const double dd = tr[d] - c[d];
            sq = sq + dd * dd;
            const double n = tf[d] + c[d];
            double mm = fmod(n - 1, a.N[d]);
            if (mm < 0) mm += a.N[d];
            lin += (int64_t)mm * stride;
            stride *= a.Ni[d];
        }
        const double wt = dev_weight(a, sqrt(sq));
        // insert in cell order; an equal cell adds to the entry already there
        int q = m;
        while (q > 0 && cp[q - 1] > (int)lin) q--;
        if (q > 0 && cp[q - 1] == (int)lin) {
            wp[q - 1] += wt;
            continue;
        }
        for (int k = m; k > q; k--) {
            cp[k] = cp[k - 1];
            wp[k] = wp[k - 1];
        }
        cp[q] = (int)lin;
        wp[q] = wt;
        m++;
    }
    cnt[p] = m;
}

// Gu rows: m = (w * Du) * (1 / sum_row(w * Du)), columns ascending.
__global__ void k_gu_rows(const int* __restrict__ cnt, const int* __restrict__ cols, const double* __restrict__ w,
                          int nNb, double Du, const int64_t* __restrict__ rowptr, int nPt, int* __restrict__ gcol,
                          double* __restrict__ gval, float* __restrict__ mval) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nPt) return;
    const int n = cnt[p];
    const int* cp = cols + (int64_t)p * nNb;
    const double* wp = w + (int64_t)p * nNb;
    double s = 0.0;
    for (int j = 0; j < n; j++) s = s + wp[j] * Du;
    const double inv = 1.0 / s;
    const int64_t o = rowptr[p];
    for (int j = 0; j < n; j++) {
        const double v = (wp[j] * Du) * inv;
        gcol[o + j] = cp[j];
        gval[o + j] = v;
        mval[o + j] = (float)v;
    }
}

// raw weights and their sample of every entry (for Gn)
__global__ void k_entry_meta(const int* __restrict__ cnt, const double* __restrict__ w, int nNb,
                             const int64_t* __restrict__ rowptr, int nPt, int* __restrict__ erow,
                             double* __restrict__ wraw) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nPt) return;
    const int64_t o = rowptr[p];
    for (int j = 0; j < cnt[p]; j++) {
        erow[o + j] = p;
        wraw[o + j] = w[(int64_t)p * nNb + j];
    }
}

// jump encoding: running deltas, the first entry absolute
__global__ void k_jumps(const int* __restrict__ v, int64_t n, int* __restrict__ out) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) out[k] = v[k] - (k ? v[k - 1] : 0);
}

__global__ void k_iota(int* v, int64_t n) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) v[k] = (int)k;
}

__global__ void k_cell_count(const unsigned* __restrict__ keys, int64_t n, int* __restrict__ cnt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) atomicAdd(&cnt[keys[i]], 1);
}

__global__ void k_flags(const int* __restrict__ cnt, int64_t n, int* __restrict__ f) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) f[i] = cnt[i] > 0;
}

// squeezed rows: row q = q-th cell with entries
__global__ void k_rows(const int* __restrict__ cnt, const int64_t* __restrict__ pos, int64_t Nu,
                       int* __restrict__ r_nJump, int* __restrict__ rowcell) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < Nu && cnt[c] > 0) {
        r_nJump[pos[c]] = cnt[c];
        rowcell[pos[c]] = (int)c;
    }
}

// transposed entries: sample and weight of the sorted entry i
__global__ void k_transpose(const int* __restrict__ sorted_e, const int* __restrict__ erow,
                            const double* __restrict__ gval, int64_t n, int* __restrict__ tcol,
                            float* __restrict__ tval) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int e = sorted_e[i];
    tcol[i] = erow[e];
    tval[i] = (float)gval[e];
}

// Gn rows (cells): w_raw * Dn over the row's samples in ascending order,
// normalised by their sum in that order
__global__ void k_gn_rows(const int* __restrict__ sorted_e, const int* __restrict__ erow,
                          const double* __restrict__ wraw, const double* __restrict__ Dn,
                          const int64_t* __restrict__ rstart, int nrows, float* __restrict__ mval) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nrows) return;
    const int64_t b = rstart[q], e = rstart[q + 1];
    double s = 0.0;
    for (int64_t i = b; i < e; i++) {
        const int k = sorted_e[i];
        s = s + wraw[k] * Dn[erow[k]];
    }
    const double inv = 1.0 / s;
    for (int64_t i = b; i < e; i++) {
        const int k = sorted_e[i];
        mval[i] = (float)((wraw[k] * Dn[erow[k]]) * inv);
    }
}

int bits_for(int64_t n) {
    int b = 1;
    while (((int64_t)1 << b) <= n) b++;
    return b;
}

template <class T>
std::vector<T> d2h(const DevBuf& b, int64_t n, hipStream_t s) {
    std::vector<T> h(n);
    if (n > 0) MONA_HIP(hipMemcpyAsync(h.data(), b.p, (size_t)n * sizeof(T), hipMemcpyDeviceToHost, s));
    return h;
}

}  // namespace

struct DeviceSparseOut {
    int r_size = 0, l_size = 0;
    std::vector<int> r_jump, r_nJump, l_jump;
    std::vector<float> m_val;
};

void traj_to_sparse_device(int imDim, int nPt, const double* t, const double* ve, const int* N_in,
                           const double* dK_in, int kernelType, int nWin, const double* kp, int nParam,
                           bool want_gn, bool want_gu, DeviceSparseOut* Gn, DeviceSparseOut* Gu,
                           DeviceSparseOut* Gut, hipStream_t s) {
    MONA_CHECK(imDim >= 1 && imDim <= 3, MONA_ERR_ARG, "imDim must be 1, 2 or 3");
    MONA_CHECK(nPt >= 1, MONA_ERR_ARG, "nPt must be >= 1");
    MONA_CHECK(nWin >= 1 && nWin <= 9, MONA_ERR_ARG, "nWin must be in [1, 9]");
    MONA_CHECK(kernelType == MONA_KERNEL_GAUSS || kernelType == MONA_KERNEL_KAISER, MONA_ERR_ARG,
               "unknown kernel type");
    MONA_CHECK((kernelType == MONA_KERNEL_GAUSS && nParam == 2) ||
                   (kernelType == MONA_KERNEL_KAISER && nParam == 3),
               MONA_ERR_ARG, "Wrong list of gridding kernel parameters.");
    MONA_CHECK(!want_gn || ve, MONA_ERR_ARG, "Gn needs the volume elements ve");
    TrajArgs a{};
    a.imDim = imDim;
    a.ktype = kernelType;
    for (int i = 0; i < 3; i++) a.p[i] = i < nParam ? (double)(float)kp[i] : 0.0;
    // I0(p1) on the host series (the device's equals it: same operations)
    {
        double q = 0.25 * a.p[1] * a.p[1], term = 1.0, sum = 1.0;
        for (int k = 1; k < 500; k++) {
            term *= q / ((double)k * (double)k);
            sum += term;
            if (term < sum * 1e-17) break;
        }
        a.I0a = kernelType == MONA_KERNEL_KAISER ? sum : 1.0;
    }
    int64_t Nu = 1;
    double Du = 1.0;
    for (int d = 0; d < 3; d++) {
        a.N[d] = 1;
        a.dK[d] = 1;
        a.Ni[d] = 1;
        a.shift[d] = 0;
    }
    for (int d = 0; d < imDim; d++) {
        MONA_CHECK(N_in[d] >= 1, MONA_ERR_ARG, "N_u must be >= 1");
        a.N[d] = (double)(float)(double)N_in[d];
        a.Ni[d] = N_in[d];
        a.dK[d] = (double)(float)dK_in[d];
        MONA_CHECK(a.dK[d] > 0, MONA_ERR_ARG, "dK_u must be > 0");
        Nu *= a.Ni[d];
        a.shift[d] = std::trunc(a.N[d] / 2 + 1);
        Du = Du * ((1.0 / a.dK[d]) / a.N[d]);
    }
    MONA_CHECK(Nu < 2147483647LL, MONA_ERR_ARG, "prod(N_u) must be < intmax('int32')");
    int nw = 0;
    if (nWin % 2 == 0) {
        for (int v = -nWin / 2 - 1; v <= nWin / 2; v++) a.r[nw++] = v;
        a.fshift = 0.0;
    } else {
        for (int v = -(nWin / 2); v <= nWin / 2; v++) a.r[nw++] = v;
        a.fshift = 0.5;
    }
    a.nw = nw;
    a.nNb = nw;
    for (int d = 1; d < imDim; d++) a.nNb *= nw;

    DevBuf dt((size_t)nPt * imDim * sizeof(double)), dve(ve ? (size_t)nPt * sizeof(double) : 0);
    MONA_HIP(hipMemcpyAsync(dt.p, t, (size_t)nPt * imDim * sizeof(double), hipMemcpyHostToDevice, s));
    if (ve) MONA_HIP(hipMemcpyAsync(dve.p, ve, (size_t)nPt * sizeof(double), hipMemcpyHostToDevice, s));
    DevBuf cnt((size_t)nPt * sizeof(int)), cols((size_t)nPt * a.nNb * sizeof(int)),
        w((size_t)nPt * a.nNb * sizeof(double)), Dn((size_t)nPt * sizeof(double));
    k_traj_entries<<<blocks_for(nPt, 128), 128, 0, s>>>(a, dt.as<double>(), ve ? dve.as<double>() : nullptr, nPt,
                                                        cnt.as<int>(), cols.as<int>(), w.as<double>(),
                                                        Dn.as<double>());
    MONA_HIP(hipGetLastError());
    DevBuf rowptr((size_t)(nPt + 1) * sizeof(int64_t));
    scan_i32_to_i64_exclusive(cnt.as<int>(), rowptr.as<int64_t>(), nPt, s);
    int64_t nnz = 0;
    MONA_HIP(hipMemcpyAsync(&nnz, rowptr.as<int64_t>() + nPt, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    MONA_HIP(hipStreamSynchronize(s));
    MONA_CHECK(nnz > 0, MONA_ERR_ARG, "every trajectory point lies outside the grid");
    MONA_CHECK(nnz < 0x7fffffff, MONA_ERR_ARG, "too many matrix entries for int32 entry indices");

    // Gu
    DevBuf gcol((size_t)nnz * sizeof(int)), gval((size_t)nnz * sizeof(double)), mval((size_t)nnz * sizeof(float));
    k_gu_rows<<<blocks_for(nPt, 256), 256, 0, s>>>(cnt.as<int>(), cols.as<int>(), w.as<double>(), a.nNb, Du,
                                                   rowptr.as<int64_t>(), nPt, gcol.as<int>(), gval.as<double>(),
                                                   mval.as<float>());
    MONA_HIP(hipGetLastError());
    DevBuf erow((size_t)nnz * sizeof(int)), wraw(want_gn ? (size_t)nnz * sizeof(double) : 0);
    {
        DevBuf wr((size_t)nnz * sizeof(double));
        k_entry_meta<<<blocks_for(nPt, 256), 256, 0, s>>>(cnt.as<int>(), w.as<double>(), a.nNb,
                                                          rowptr.as<int64_t>(), nPt, erow.as<int>(),
                                                          wr.as<double>());
        MONA_HIP(hipGetLastError());
        if (want_gn) std::swap(wraw, wr);
        MONA_HIP(hipStreamSynchronize(s));
    }
    if (want_gu) {
        DevBuf jump((size_t)nnz * sizeof(int));
        k_jumps<<<blocks_for(nnz, 256), 256, 0, s>>>(gcol.as<int>(), nnz, jump.as<int>());
        MONA_HIP(hipGetLastError());
        Gu->r_size = (int)Nu;
        Gu->l_size = nPt;
        Gu->r_jump = d2h<int>(jump, nnz, s);
        Gu->r_nJump = d2h<int>(cnt, nPt, s);
        Gu->m_val = d2h<float>(mval, nnz, s);
        Gu->l_jump.clear();
        MONA_HIP(hipStreamSynchronize(s));
    }

    // cell-major order of the entries (stable: samples ascending per cell)
    DevBuf keys((size_t)nnz * sizeof(unsigned)), order((size_t)nnz * sizeof(int));
    MONA_HIP(hipMemcpyAsync(keys.p, gcol.p, (size_t)nnz * sizeof(int), hipMemcpyDeviceToDevice, s));
    k_iota<<<blocks_for(nnz, 256), 256, 0, s>>>(order.as<int>(), nnz);
    MONA_HIP(hipGetLastError());
    radix_sort_pairs(keys.as<unsigned>(), order.as<int>(), nnz, bits_for(Nu), s);
    DevBuf ccount((size_t)Nu * sizeof(int)), flags((size_t)Nu * sizeof(int)), pos((size_t)(Nu + 1) * sizeof(int64_t));
    fill_i32(ccount.as<int>(), 0, Nu, s);
    k_cell_count<<<blocks_for(nnz, 256), 256, 0, s>>>(keys.as<unsigned>(), nnz, ccount.as<int>());
    k_flags<<<blocks_for(Nu, 256), 256, 0, s>>>(ccount.as<int>(), Nu, flags.as<int>());
    MONA_HIP(hipGetLastError());
    scan_i32_to_i64_exclusive(flags.as<int>(), pos.as<int64_t>(), Nu, s);
    int64_t nrows = 0;
    MONA_HIP(hipMemcpyAsync(&nrows, pos.as<int64_t>() + Nu, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    MONA_HIP(hipStreamSynchronize(s));
    DevBuf rnj((size_t)nrows * sizeof(int)), rcell((size_t)nrows * sizeof(int)), ljump((size_t)nrows * sizeof(int));
    k_rows<<<blocks_for(Nu, 256), 256, 0, s>>>(ccount.as<int>(), pos.as<int64_t>(), Nu, rnj.as<int>(),
                                               rcell.as<int>());
    k_jumps<<<blocks_for(nrows, 256), 256, 0, s>>>(rcell.as<int>(), nrows, ljump.as<int>());
    DevBuf tcol((size_t)nnz * sizeof(int)), tval((size_t)nnz * sizeof(float)), tjump((size_t)nnz * sizeof(int));
    k_transpose<<<blocks_for(nnz, 256), 256, 0, s>>>(order.as<int>(), erow.as<int>(), gval.as<double>(), nnz,
                                                     tcol.as<int>(), tval.as<float>());
    k_jumps<<<blocks_for(nnz, 256), 256, 0, s>>>(tcol.as<int>(), nnz, tjump.as<int>());
    MONA_HIP(hipGetLastError());
    auto fill_rows = [&](DeviceSparseOut* o, const DevBuf& mv) {
        o->r_size = nPt;
        o->l_size = (int)Nu;
        o->r_jump = d2h<int>(tjump, nnz, s);
        o->r_nJump = d2h<int>(rnj, nrows, s);
        o->l_jump = d2h<int>(ljump, nrows, s);
        o->m_val = d2h<float>(mv, nnz, s);
    };
    if (want_gu) fill_rows(Gut, tval);
    if (want_gn) {
        DevBuf rstart((size_t)(nrows + 1) * sizeof(int64_t)), gnv((size_t)nnz * sizeof(float));
        scan_i32_to_i64_exclusive(rnj.as<int>(), rstart.as<int64_t>(), nrows, s);
        k_gn_rows<<<blocks_for(nrows, 256), 256, 0, s>>>(order.as<int>(), erow.as<int>(), wraw.as<double>(),
                                                          Dn.as<double>(), rstart.as<int64_t>(), (int)nrows,
                                                          gnv.as<float>());
        MONA_HIP(hipGetLastError());
        fill_rows(Gn, gnv);
        MONA_HIP(hipStreamSynchronize(s));  // gnv is released here
    }
    MONA_HIP(hipStreamSynchronize(s));
}

}  // namespace mona

using namespace mona;

extern "C" int mona_traj2sparsemat_device(int imDim, int nPt, const double* t, const double* ve,
                                          const int* N_u, const double* dK_u, int kernelType, int nWin,
                                          const double* kernelParam, int nParam, int which, mona_sparse* Gn,
                                          mona_sparse* Gu, mona_sparse* Gut, hipStream_t stream) {
    return guarded([&] {
        const bool want_gn = which & 1, want_gu = which & 2;
        MONA_CHECK(want_gn || want_gu, MONA_ERR_ARG, "nothing requested");
        MONA_CHECK(t && N_u && dK_u && kernelParam, MONA_ERR_ARG, "NULL input");
        MONA_CHECK(!want_gn || Gn, MONA_ERR_ARG, "Gn output is NULL");
        MONA_CHECK(!want_gu || (Gu && Gut), MONA_ERR_ARG, "Gu/Gut output is NULL");
        DeviceSparseOut hn, hu, hut;
        traj_to_sparse_device(imDim, nPt, t, ve, N_u, dK_u, kernelType, nWin, kernelParam, nParam, want_gn,
                              want_gu, &hn, &hu, &hut, stream);
        auto make = [&](DeviceSparseOut& h) {
            auto* sp = new mona_sparse_s;
            try {
                sparse_set_host(sp->S, h.r_size, std::move(h.r_jump), std::move(h.r_nJump), std::move(h.m_val),
                                h.l_size, std::move(h.l_jump));
            } catch (...) {
                delete sp;
                throw;
            }
            return sp;
        };
        if (want_gn) *Gn = make(hn);
        if (want_gu) {
            *Gu = make(hu);
            try {
                *Gut = make(hut);
            } catch (...) {
                delete *Gu;
                *Gu = nullptr;
                throw;
            }
            auto T = make_traj_info(imDim, nPt, t, N_u, dK_u, kernelType, nWin, kernelParam, nParam);
            (*Gu)->S.traj = T;
            (*Gut)->S.traj = T;
        }
    });
}
