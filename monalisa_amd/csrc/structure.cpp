// Trajectory -> gridding matrices Gn, Gu, Gut in the bmSparseMat jump
// encoding, following src/gridding123/m/bmTraj2SparseMat.m:58-297 with
// MATLAB's float64 semantics for every index computation:
//   t/dK_u + fix(N/2+1); delete points with t < 1 or t > N; floor(t + 0.5)
//   (odd window) or floor(t) (even window, offsets -nWin/2-1 .. nWin/2);
//   periodic wrap mod(n-1, N)+1; x-fastest linear cell index;
//   weights normpdf / besseli-Kaiser of the Euclidean distance;
//   rows normalised by diag(1./sum(A,2)) (sum in ascending column order);
// and the conversion of bmSparseMat_matlabSparse2bmSparseMat.m:6-40 +
// bmSparseMat.m:62-188 (ascending columns per row, row squeeze, one block,
// running deltas with an absolute 0-based first entry).
//
// Setup code run once per trajectory on the host cores; the products that
// use the matrices run on the GPU.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "mona_internal.hpp"

namespace mona {

namespace {

double bessel_i0(double x) {
    // power series sum_k ((x/2)^k / k!)^2
    double q = 0.25 * x * x, term = 1.0, sum = 1.0;
    for (int k = 1; k < 500; k++) {
        term *= q / ((double)k * (double)k);
        sum += term;
        if (term < sum * 1e-17) break;
    }
    return sum;
}

struct KernelSpec {
    int type;  // 0 gauss, 1 kaiser
    double p[3];
    double I0a;
    double weight(double d) const {
        if (type == MONA_KERNEL_GAUSS) {
            double s = p[0];
            double r = d / s;
            return std::exp(-0.5 * r * r) / (std::sqrt(2.0 * M_PI) * s);
        }
        double r = d / p[0];
        double w = 1.0 - r * r;
        if (w < 0) w = 0;
        w = p[1] * std::sqrt(w);
        return bessel_i0(w) / I0a;
    }
};

double as_single(double v) { return (double)(float)v; }

}  // namespace

struct HostSparse {
    int r_size = 0, l_size = 0, l_nJump = 0;
    std::vector<int> r_jump, r_nJump, l_jump;  // l_jump empty when not squeezed
    std::vector<float> m_val;
};

// Gu (rows = samples, ascending columns, m = wts * (1 / row sum)), Gut = Gu'
// (rows = cells, ascending samples) and Gn (rows = cells, wraw * Dn
// normalised by the row sum in ascending sample order) from the per-sample
// entries (cnt[p] entries at p * nNb, columns ascending).  squeeze: Gut and
// Gn store only the cells with entries (bmSparseMat.l_squeeze), else every
// cell.
void assemble(int nPt, int64_t Nu, int nNb, const std::vector<int64_t>& cnt, const std::vector<int>& cols,
              const std::vector<double>& wts, const std::vector<double>& wraw, const std::vector<double>& Dn,
              bool want_gn, bool want_gu, bool squeeze, HostSparse* Gn, HostSparse* Gu, HostSparse* Gut) {
    if (want_gu) {
        // Gu rows: m = (w*Du) * (1/sum_row)
        std::vector<int64_t> rowptr(nPt + 1, 0);
        for (int p = 0; p < nPt; p++) rowptr[p + 1] = rowptr[p] + cnt[p];
        int64_t nnz = rowptr[nPt];
        MONA_CHECK(nnz > 0, MONA_ERR_ARG, "every point lies outside the grid");
        Gu->r_size = (int)Nu;
        Gu->l_size = nPt;
        Gu->l_nJump = nPt;
        Gu->r_nJump.assign(nPt, 0);
        Gu->r_jump.assign(nnz, 0);
        Gu->m_val.assign(nnz, 0.0f);
        Gu->l_jump.clear();
        std::vector<int> gcol(nnz);
        std::vector<double> gval(nnz);
#pragma omp parallel for schedule(static)
        for (int p = 0; p < nPt; p++) {
            int64_t o = rowptr[p];
            Gu->r_nJump[p] = (int)cnt[p];
            double s = 0.0;
            for (int j = 0; j < cnt[p]; j++) s = s + wts[(size_t)p * nNb + j];
            double inv = 1.0 / s;
            for (int j = 0; j < cnt[p]; j++) {
                double v = wts[(size_t)p * nNb + j] * inv;
                gcol[o + j] = cols[(size_t)p * nNb + j];
                gval[o + j] = v;
                Gu->m_val[o + j] = (float)v;
            }
        }
        int prev = 0;
        for (int64_t k = 0; k < nnz; k++) {
            Gu->r_jump[k] = gcol[k] - prev;
            prev = gcol[k];
        }
        // Gut = Gu' (rows = cells), entries in ascending point order
        std::vector<int64_t> ccount(Nu + 1, 0);
        for (int64_t k = 0; k < nnz; k++) ccount[gcol[k] + 1]++;
        for (int64_t c = 0; c < Nu; c++) ccount[c + 1] += ccount[c];
        std::vector<int> tcol(nnz);
        std::vector<float> tval(nnz);
        {
            std::vector<int64_t> cur(ccount.begin(), ccount.end() - 1);
            for (int p = 0; p < nPt; p++)
                for (int64_t k = rowptr[p]; k < rowptr[p + 1]; k++) {
                    int64_t pos = cur[gcol[k]]++;
                    tcol[pos] = p;
                    tval[pos] = (float)gval[k];
                }
        }
        Gut->r_size = nPt;
        Gut->l_size = (int)Nu;
        Gut->r_nJump.clear();
        Gut->l_jump.clear();
        int prow = 0;
        for (int64_t c = 0; c < Nu; c++) {
            int64_t n = ccount[c + 1] - ccount[c];
            if (n == 0 && squeeze) continue;
            Gut->r_nJump.push_back((int)n);
            if (squeeze) Gut->l_jump.push_back((int)(c - prow));
            prow = (int)c;
        }
        Gut->l_nJump = (int)Gut->r_nJump.size();
        Gut->r_jump.resize(nnz);
        int pc = 0;
        for (int64_t k = 0; k < nnz; k++) {
            Gut->r_jump[k] = tcol[k] - pc;
            pc = tcol[k];
        }
        Gut->m_val.swap(tval);
    }

    if (want_gn) {
        // Gn = rownorm(sparse(pt, cell, w.*Dn)') : rows = cells, ascending points
        std::vector<int64_t> rowptr(nPt + 1, 0);
        for (int p = 0; p < nPt; p++) rowptr[p + 1] = rowptr[p] + cnt[p];
        int64_t nnz = rowptr[nPt];
        MONA_CHECK(nnz > 0, MONA_ERR_ARG, "every point lies outside the grid");
        std::vector<int64_t> ccount(Nu + 1, 0);
        for (int p = 0; p < nPt; p++)
            for (int j = 0; j < cnt[p]; j++) ccount[cols[(size_t)p * nNb + j] + 1]++;
        for (int64_t c = 0; c < Nu; c++) ccount[c + 1] += ccount[c];
        std::vector<int> tcol(nnz);
        std::vector<double> tval(nnz);
        {
            std::vector<int64_t> cur(ccount.begin(), ccount.end() - 1);
            for (int p = 0; p < nPt; p++)
                for (int j = 0; j < cnt[p]; j++) {
                    int64_t pos = cur[cols[(size_t)p * nNb + j]]++;
                    tcol[pos] = p;
                    tval[pos] = wraw[(size_t)p * nNb + j] * Dn[p];
                }
        }
        Gn->r_size = nPt;
        Gn->l_size = (int)Nu;
        Gn->r_nJump.clear();
        Gn->l_jump.clear();
        Gn->m_val.resize(nnz);
        int prow = 0;
        for (int64_t c = 0; c < Nu; c++) {
            int64_t b = ccount[c], e = ccount[c + 1];
            if (e == b && squeeze) continue;
            double s = 0.0;
            for (int64_t k = b; k < e; k++) s = s + tval[k];
            double inv = 1.0 / s;
            for (int64_t k = b; k < e; k++) Gn->m_val[k] = (float)(tval[k] * inv);
            Gn->r_nJump.push_back((int)(e - b));
            if (squeeze) Gn->l_jump.push_back((int)(c - prow));
            prow = (int)c;
        }
        Gn->l_nJump = (int)Gn->r_nJump.size();
        Gn->r_jump.resize(nnz);
        int pc = 0;
        for (int64_t k = 0; k < nnz; k++) {
            Gn->r_jump[k] = tcol[k] - pc;
            pc = tcol[k];
        }
    }
}

// Builds the requested matrices.  t is [nPt][imDim].
void traj_to_sparse(int imDim, int nPt, const double* t, const double* ve, const int* N_in,
                    const double* dK_in, int kernelType, int nWin, const double* kp, int nParam,
                    bool want_gn, bool want_gu, HostSparse* Gn, HostSparse* Gu, HostSparse* Gut) {
    MONA_CHECK(imDim >= 1 && imDim <= 3, MONA_ERR_ARG, "imDim must be 1, 2 or 3");
    MONA_CHECK(nPt >= 1, MONA_ERR_ARG, "nPt must be >= 1");
    MONA_CHECK(nWin >= 1 && nWin <= 9, MONA_ERR_ARG, "nWin must be in [1, 9]");
    MONA_CHECK(kernelType == MONA_KERNEL_GAUSS || kernelType == MONA_KERNEL_KAISER, MONA_ERR_ARG,
               "unknown kernel type");
    MONA_CHECK((kernelType == MONA_KERNEL_GAUSS && nParam == 2) ||
                   (kernelType == MONA_KERNEL_KAISER && nParam == 3),
               MONA_ERR_ARG, "Wrong list of gridding kernel parameters.");
    MONA_CHECK(!want_gn || ve, MONA_ERR_ARG, "Gn needs the volume elements ve");
    KernelSpec K;
    K.type = kernelType;
    for (int i = 0; i < 3; i++) K.p[i] = i < nParam ? as_single(kp[i]) : 0.0;
    K.I0a = kernelType == MONA_KERNEL_KAISER ? bessel_i0(K.p[1]) : 1.0;

    double N[3] = {1, 1, 1}, dK[3] = {1, 1, 1};
    int64_t Ni[3] = {1, 1, 1};
    int64_t Nu = 1;
    double Du = 1.0;
    double shift[3] = {0, 0, 0};
    for (int d = 0; d < imDim; d++) {
        MONA_CHECK(N_in[d] >= 1, MONA_ERR_ARG, "N_u must be >= 1");
        N[d] = as_single((double)N_in[d]);
        Ni[d] = N_in[d];
        dK[d] = as_single(dK_in[d]);
        MONA_CHECK(dK[d] > 0, MONA_ERR_ARG, "dK_u must be > 0");
        Nu *= Ni[d];
        shift[d] = std::trunc(N[d] / 2 + 1);
        Du = Du * ((1.0 / dK[d]) / N[d]);
    }
    MONA_CHECK(Nu < 2147483647LL, MONA_ERR_ARG, "prod(N_u) must be < intmax('int32')");

    // neighbour offsets (bmTraj2SparseMat.m:134-164), x fastest
    std::vector<int> r;
    double fshift;
    if (nWin % 2 == 0) {
        for (int v = -nWin / 2 - 1; v <= nWin / 2; v++) r.push_back(v);
        fshift = 0.0;
    } else {
        for (int v = -(nWin / 2); v <= nWin / 2; v++) r.push_back(v);
        fshift = 0.5;
    }
    int nw = (int)r.size();
    int nNb = nw;
    for (int d = 1; d < imDim; d++) nNb *= nw;

    // per-point entries, rows of Gu (columns sorted ascending)
    std::vector<int64_t> cnt(nPt);
    std::vector<int> cols((size_t)nPt * nNb);
    std::vector<double> wts((size_t)nPt * nNb);    // w * Du, normalised later
    std::vector<double> wraw((size_t)nPt * nNb);   // w (for Gn)
    std::vector<unsigned char> keep(nPt);
    std::vector<double> Dn(nPt, 0.0);

#pragma omp parallel for schedule(static)
    for (int p = 0; p < nPt; p++) {
        double tg[3], tf[3], tr[3];
        bool del = false;
        for (int d = 0; d < imDim; d++) {
            tg[d] = t[(int64_t)p * imDim + d] / dK[d] + shift[d];
            if (tg[d] < 1 || tg[d] > N[d]) del = true;
            tf[d] = std::floor(tg[d] + fshift);
            tr[d] = tg[d] - tf[d];
        }
        if (ve) {
            double dn = ve[p];
            for (int d = 0; d < imDim; d++) dn = dn / dK[d];
            Dn[p] = dn;
        }
        keep[p] = !del;
        if (del) { cnt[p] = 0; continue; }
        std::pair<int, double> e[1000];
        for (int j = 0; j < nNb; j++) {
            int c[3] = {r[j % nw], imDim > 1 ? r[(j / nw) % nw] : 0, imDim > 2 ? r[j / (nw * nw)] : 0};
            double sq = 0.0;
            int64_t lin = 0, stride = 1;
            for (int d = 0; d < imDim; d++) {
                double dd = tr[d] - c[d];
                sq = sq + dd * dd;
                double n = tf[d] + c[d];
                double m = std::fmod(n - 1, N[d]);
                if (m < 0) m += N[d];
                lin += (int64_t)m * stride;
                stride *= Ni[d];
            }
            double w = K.weight(std::sqrt(sq));
            e[j] = {(int)lin, w};
        }
        for (int j = 1; j < nNb; j++)  // stable: duplicates summed in neighbour order
            for (int i = j; i > 0 && e[i - 1].first > e[i].first; i--) std::swap(e[i - 1], e[i]);
        // merge duplicate cells (sparse() sums them)
        int m = 0;
        for (int j = 0; j < nNb; j++) {
            if (m > 0 && e[m - 1].first == e[j].first) e[m - 1].second += e[j].second;
            else e[m++] = e[j];
        }
        cnt[p] = m;
        int* cp = &cols[(size_t)p * nNb];
        double* wp = &wts[(size_t)p * nNb];
        double* rp = &wraw[(size_t)p * nNb];
        for (int j = 0; j < m; j++) {
            cp[j] = e[j].first;
            rp[j] = e[j].second;
            wp[j] = e[j].second * Du;
        }
    }

    assemble(nPt, Nu, nNb, cnt, cols, wts, wraw, Dn, want_gn, want_gu, true, Gn, Gu, Gut);
}

// Deformation field -> interpolation matrices
// (src/gridding123/m/bmImDeformField2SparseMat.m:14-147 with
// bmImDeformField2sparseMat_ind_weight.m:6-150): sample p is pixel p
// (1-based grid coordinates, x fastest) moved by v[p]; it is deleted when
// it leaves [1, N] in some dimension (unless torus); its neighbours are the
// 2^imDim cells floor(t) + c, c in {0, 1}^imDim, wrapped mod N, weighted
// by the bump exp(-1 / (1 - d^2)) of the Euclidean distance d (0 for
// d >= 1); zero weights are not stored (sparse() ignores them).  Gu rows
// are normalised by their sums, Gut = Gu', Gn = rownorm((w .* Dn)'); Gut
// and Gn keep every cell as a row (no squeeze).  v is [nPt][imDim].
void deform_to_sparse(int imDim, const int* N_in, const double* v, const double* Dn_in, bool torus, bool want_gn,
                      bool want_gu, HostSparse* Gn, HostSparse* Gu, HostSparse* Gut) {
    MONA_CHECK(imDim >= 1 && imDim <= 3, MONA_ERR_ARG, "imDim must be 1, 2 or 3");
    MONA_CHECK(!want_gn || Dn_in, MONA_ERR_ARG, "Gn needs Dn");
    int64_t Ni[3] = {1, 1, 1}, Nu = 1;
    double N[3] = {1, 1, 1};
    for (int d = 0; d < imDim; d++) {
        MONA_CHECK(N_in[d] >= 1, MONA_ERR_ARG, "N_u must be >= 1");
        Ni[d] = N_in[d];
        N[d] = (double)N_in[d];
        Nu *= Ni[d];
    }
    MONA_CHECK(Nu < 2147483647LL, MONA_ERR_ARG, "prod(N_u) must be < intmax('int32')");
    const int nPt = (int)Nu;
    const int nNb = 1 << imDim;
    std::vector<int64_t> cnt(nPt);
    std::vector<int> cols((size_t)nPt * nNb);
    std::vector<double> wts((size_t)nPt * nNb);
    std::vector<double> Dn(nPt, 1.0);
    if (Dn_in)
        for (int p = 0; p < nPt; p++) Dn[p] = Dn_in[p];
#pragma omp parallel for schedule(static)
    for (int p = 0; p < nPt; p++) {
        double tf[3], tr[3];
        bool del = false;
        int64_t q = p;
        for (int d = 0; d < imDim; d++) {
            const double x = (double)(q % Ni[d] + 1);  // 1-based pixel coordinate
            q /= Ni[d];
            const double t = x + v[(int64_t)p * imDim + d];
            if (t < 1 || t > N[d]) del = true;
            tf[d] = std::floor(t);
            tr[d] = t - tf[d];
        }
        if (del && !torus) {
            cnt[p] = 0;
            continue;
        }
        std::pair<int, double> e[8];
        for (int j = 0; j < nNb; j++) {
            double sq = 0.0;
            int64_t lin = 0, stride = 1;
            for (int d = 0; d < imDim; d++) {
                const int c = (j >> d) & 1;
                const double dd = tr[d] - c;
                sq = sq + dd * dd;
                const double n = tf[d] + c;
                double m = std::fmod(n - 1, N[d]);
                if (m < 0) m += N[d];
                lin += (int64_t)m * stride;
                stride *= Ni[d];
            }
            const double dist = std::sqrt(sq);
            double w = std::exp(-1.0 / (1.0 - dist * dist));
            if (std::isinf(w)) w = 0.0;
            w = w * (double)(std::fabs(dist) < 1);
            e[j] = {(int)lin, w};
        }
        std::sort(e, e + nNb, [](const std::pair<int, double>& a, const std::pair<int, double>& b) {
            return a.first < b.first;
        });
        int m = 0;
        for (int j = 0; j < nNb; j++) {
            if (m > 0 && e[m - 1].first == e[j].first) e[m - 1].second += e[j].second;
            else e[m++] = e[j];
        }
        int k = 0;
        for (int j = 0; j < m; j++)
            if (e[j].second != 0.0) {
                cols[(size_t)p * nNb + k] = e[j].first;
                wts[(size_t)p * nNb + k] = e[j].second;
                k++;
            }
        cnt[p] = k;
    }
    assemble(nPt, Nu, nNb, cnt, cols, wts, wts, Dn, want_gn, want_gu, false, Gn, Gu, Gut);
}

void traj_points(const TrajInfo& T, TrajPoints* out) {
    MONA_CHECK(T.nWin % 2 == 1, MONA_ERR_ARG, "matrix-free gridding needs an odd window width");
    KernelSpec K;
    K.type = T.ktype;
    for (int i = 0; i < 3; i++) K.p[i] = i < T.nParam ? as_single(T.kp[i]) : 0.0;
    K.I0a = T.ktype == MONA_KERNEL_KAISER ? bessel_i0(K.p[1]) : 1.0;
    const int D = T.imDim, h = T.nWin / 2, nw = T.nWin;
    double N[3] = {1, 1, 1}, dK[3] = {1, 1, 1}, shift[3] = {0, 0, 0};
    for (int d = 0; d < D; d++) {
        N[d] = as_single((double)T.N[d]);
        dK[d] = as_single(T.dK[d]);
        shift[d] = std::trunc(N[d] / 2 + 1);
    }
    int nNb = 1;
    for (int d = 0; d < D; d++) nNb *= nw;
    out->base.assign((size_t)T.nPt * 3, 0);
    out->rest.assign((size_t)T.nPt * 3, 0.0);
    out->inv_sum.assign(T.nPt, 0.0);
#pragma omp parallel for schedule(static)
    for (int p = 0; p < T.nPt; p++) {
        double tg[3], tf[3], tr[3];
        bool del = false;
        for (int d = 0; d < D; d++) {
            tg[d] = T.t[(int64_t)p * D + d] / dK[d] + shift[d];
            if (tg[d] < 1 || tg[d] > N[d]) del = true;
            tf[d] = std::floor(tg[d] + 0.5);
            tr[d] = tg[d] - tf[d];
        }
        if (del) continue;
        double sum = 0.0;
        for (int j = 0; j < nNb; j++) {
            int c[3] = {j % nw - h, D > 1 ? (j / nw) % nw - h : 0, D > 2 ? j / (nw * nw) - h : 0};
            double sq = 0.0;
            for (int d = 0; d < D; d++) {
                double dd = tr[d] - c[d];
                sq = sq + dd * dd;
            }
            sum += K.weight(std::sqrt(sq));
        }
        // the device evaluates the kernel without its constant factor
        // (I0(alpha) for Kaiser, sqrt(2 pi) sigma for Gauss)
        sum *= T.ktype == MONA_KERNEL_KAISER ? K.I0a : std::sqrt(2.0 * M_PI) * K.p[0];
        for (int d = 0; d < D; d++) {
            double m = std::fmod(tf[d] - 1, N[d]);
            if (m < 0) m += N[d];
            out->base[(size_t)p * 3 + d] = (int)m;
            out->rest[(size_t)p * 3 + d] = tr[d];
        }
        out->inv_sum[p] = sum > 0 ? 1.0 / sum : 0.0;
    }
}

std::shared_ptr<const TrajInfo> make_traj_info(int imDim, int nPt, const double* t, const int* N_u,
                                               const double* dK_u, int kernelType, int nWin,
                                               const double* kp, int nParam) {
    auto T = std::make_shared<TrajInfo>();
    T->imDim = imDim; T->nPt = nPt; T->ktype = kernelType; T->nWin = nWin; T->nParam = nParam;
    for (int d = 0; d < imDim; d++) { T->N[d] = N_u[d]; T->dK[d] = dK_u[d]; }
    for (int i = 0; i < nParam && i < 3; i++) T->kp[i] = kp[i];
    T->t.assign(t, t + (size_t)nPt * imDim);
    return T;
}

}  // namespace mona

using namespace mona;

extern "C" int mona_deform2sparsemat(int imDim, const int* N_u, const double* v, const double* Dn, int torus,
                                     int which, mona_sparse* Gn, mona_sparse* Gu, mona_sparse* Gut) {
    return guarded([&] {
        bool want_gn = which & 1, want_gu = which & 2;
        MONA_CHECK(want_gn || want_gu, MONA_ERR_ARG, "nothing requested");
        MONA_CHECK(N_u && v, MONA_ERR_ARG, "NULL input");
        MONA_CHECK(!want_gn || Gn, MONA_ERR_ARG, "Gn output is NULL");
        MONA_CHECK(!want_gu || (Gu && Gut), MONA_ERR_ARG, "Gu/Gut output is NULL");
        HostSparse hn, hu, hut;
        deform_to_sparse(imDim, N_u, v, Dn, torus != 0, want_gn, want_gu, &hn, &hu, &hut);
        auto make = [&](HostSparse& h) {
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
        }
    });
}

extern "C" int mona_traj2sparsemat(int imDim, int nPt, const double* t, const double* ve,
                                   const int* N_u, const double* dK_u, int kernelType, int nWin,
                                   const double* kernelParam, int nParam, int which,
                                   mona_sparse* Gn, mona_sparse* Gu, mona_sparse* Gut) {
    return guarded([&] {
        bool want_gn = which & 1, want_gu = which & 2;
        MONA_CHECK(want_gn || want_gu, MONA_ERR_ARG, "nothing requested");
        MONA_CHECK(t && N_u && dK_u && kernelParam, MONA_ERR_ARG, "NULL input");
        MONA_CHECK(!want_gn || Gn, MONA_ERR_ARG, "Gn output is NULL");
        MONA_CHECK(!want_gu || (Gu && Gut), MONA_ERR_ARG, "Gu/Gut output is NULL");
        HostSparse hn, hu, hut;
        traj_to_sparse(imDim, nPt, t, ve, N_u, dK_u, kernelType, nWin, kernelParam, nParam, want_gn,
                       want_gu, &hn, &hu, &hut);
        auto make = [&](HostSparse& h) {
            auto* s = new mona_sparse_s;
            try {
                sparse_set_host(s->S, h.r_size, std::move(h.r_jump), std::move(h.r_nJump),
                                std::move(h.m_val), h.l_size, std::move(h.l_jump));
            } catch (...) {
                delete s;
                throw;
            }
            return s;
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
            // the trajectory travels with Gu / Gut: plans can grid from it
            auto T = make_traj_info(imDim, nPt, t, N_u, dK_u, kernelType, nWin, kernelParam, nParam);
            (*Gu)->S.traj = T;
            (*Gut)->S.traj = T;
        }
    });
}
