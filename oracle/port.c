/*
 * ORACLE -- test infrastructure only.  Nothing in the product library links,
 * loads or calls this file; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, as the checker and as the timed CPU
 * baseline.
 *
 * Plain-C restatement of the reference's native (C++ MEX) kernels on the
 * reconstruction hot path, with the same argument meaning, memory layout,
 * loop order and float32 arithmetic (one rounded multiply, then one rounded
 * add, per term -- compiled with -ffp-contract=off so no FMA is formed).
 *
 * Pinned by: tests/test_oracle_port.py (bit-exact against the reference's own
 * bmSparseMat_cC_oBlock_omp.cpp built into oracle/_ref, against dense float64
 * products on small matrices, and against brute-force per-cell sums for the
 * gridders).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------
 * Jump-encoded CSR product, complex, column layout, one block, OpenMP.
 * Follows src/sparseMat/cpp/bmSparseMat_cC_oBlock_omp.cpp:25-177
 * (= src/sparseMat/mex/bmSparseMat_vec/bmSparseMat_cC_oBlock_omp/
 *  bmSparseMat_cC_oBlock_omp_mex.cpp:106-260).
 * v_* are [n_vec][r_size] planes, w_* are [n_vec][l_size] planes.
 * The column cursor is NOT reset between rows: r_jump is a running delta
 * whose first entry is the absolute 0-based column.  l_jump (may be NULL)
 * is the running delta of the output row, applied before each stored row.
 * ---------------------------------------------------------------------- */
static void plane_product(int l_size, const int *r_jump, const int *r_nJump,
                          const float *m_val, const int *l_jump, int l_nJump,
                          const float *v, float *w)
{
    const float *vrun = v;
    float *wrun = w;
    const int *rj = r_jump;
    const float *mv = m_val;
    for (int l = 0; l < l_size; l++) w[l] = 0.0f;
    for (int i = 0; i < l_nJump; i++) {
        int n = r_nJump[i];
        if (l_jump) wrun += l_jump[i];
        for (int j = 0; j < n; j++) {
            vrun += *rj++;
            float prod = (*mv++) * (*vrun);
            *wrun = *wrun + prod;
        }
        if (!l_jump) wrun++;
    }
}

void port_cC_oBlock_omp(int r_size, const int *r_jump, const int *r_nJump,
                        const float *m_val, int l_size, const int *l_jump,
                        int l_nJump, const float *v_re, const float *v_im,
                        int n_vec, float *w_re, float *w_im)
{
    /* thread rule of ...cC_oBlock_omp.cpp:37-44 */
    int nthr = 1;
#ifdef _OPENMP
    nthr = omp_get_max_threads();
    if (2 * n_vec < nthr) nthr = 2 * n_vec;
#endif
    (void)nthr;
#pragma omp parallel for num_threads(nthr) schedule(static)
    for (int k = 0; k < 2 * n_vec; k++) {
        int kh = k / 2;
        const float *v = (k % 2 == 0 ? v_re : v_im) + (int64_t)r_size * kh;
        float *w = (k % 2 == 0 ? w_re : w_im) + (int64_t)l_size * kh;
        plane_product(l_size, r_jump, r_nJump, m_val, l_jump, l_nJump, v, w);
    }
}

/* Serial complex column variant:
 * src/sparseMat/mex/bmSparseMat_vec/bmSparseMat_cC_oBlock/bmSparseMat_cC_oBlock_mex.cpp:82-193
 * (re and im of one vector in one sweep; per-plane arithmetic identical). */
void port_cC_oBlock(int r_size, const int *r_jump, const int *r_nJump,
                    const float *m_val, int l_size, const int *l_jump,
                    int l_nJump, const float *v_re, const float *v_im,
                    int n_vec, float *w_re, float *w_im)
{
    for (int k = 0; k < n_vec; k++) {
        plane_product(l_size, r_jump, r_nJump, m_val, l_jump, l_nJump,
                      v_re + (int64_t)r_size * k, w_re + (int64_t)l_size * k);
        plane_product(l_size, r_jump, r_nJump, m_val, l_jump, l_nJump,
                      v_im + (int64_t)r_size * k, w_im + (int64_t)l_size * k);
    }
}

/* Real column variant (rC_oBlock, ...rC_oBlock_mex.cpp:77-174;
 * rC_oBlock_omp ...rC_oBlock_omp_mex.cpp:79-152). */
void port_rC_oBlock(int r_size, const int *r_jump, const int *r_nJump,
                    const float *m_val, int l_size, const int *l_jump,
                    int l_nJump, const float *v, int n_vec, float *w)
{
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n_vec; k++)
        plane_product(l_size, r_jump, r_nJump, m_val, l_jump, l_nJump,
                      v + (int64_t)r_size * k, w + (int64_t)l_size * k);
}

/* Row layout variants: V is [r_size][n_vec] (vector index fastest), W is
 * [l_size][n_vec].  Per row a temp_sum[n_vec] accumulates v*m in entry
 * order, then is copied to the output row.  Follows
 * src/sparseMat/mex/bmSparseMat_vec/bmSparseMat_cR_oBlock/bmSparseMat_cR_oBlock_mex.cpp:80-223
 * and bmSparseMat_rR_oBlock_mex.cpp:77-194 (real). */
static void row_product(int l_size, const int *r_jump, const int *r_nJump,
                        const float *m_val, const int *l_jump, int l_nJump,
                        const float *v, int n_vec, float *w)
{
    float *tmp = (float *)malloc(sizeof(float) * (size_t)(n_vec > 0 ? n_vec : 1));
    const float *vp = v;
    float *wp = w;
    const int *rj = r_jump;
    const float *mv = m_val;
    if (l_jump) memset(w, 0, sizeof(float) * (size_t)l_size * (size_t)n_vec);
    for (int i = 0; i < l_nJump; i++) {
        for (int k = 0; k < n_vec; k++) tmp[k] = 0.0f;
        int n = r_nJump[i];
        for (int j = 0; j < n; j++) {
            float m = *mv++;
            vp += (int64_t)(*rj++) * n_vec;
            for (int k = 0; k < n_vec; k++) {
                float prod = vp[k] * m;
                tmp[k] = tmp[k] + prod;
            }
        }
        if (l_jump) {
            wp += (int64_t)l_jump[i] * n_vec;
            for (int k = 0; k < n_vec; k++) wp[k] = tmp[k];
        } else {
            for (int k = 0; k < n_vec; k++) wp[k] = tmp[k];
            wp += n_vec;
        }
    }
    free(tmp);
}

void port_cR_oBlock(int r_size, const int *r_jump, const int *r_nJump,
                    const float *m_val, int l_size, const int *l_jump,
                    int l_nJump, const float *v_re, const float *v_im,
                    int n_vec, float *w_re, float *w_im)
{
    (void)r_size;
    row_product(l_size, r_jump, r_nJump, m_val, l_jump, l_nJump, v_re, n_vec, w_re);
    row_product(l_size, r_jump, r_nJump, m_val, l_jump, l_nJump, v_im, n_vec, w_im);
}

void port_rR_oBlock(int r_size, const int *r_jump, const int *r_nJump,
                    const float *m_val, int l_size, const int *l_jump,
                    int l_nJump, const float *v, int n_vec, float *w)
{
    (void)r_size;
    row_product(l_size, r_jump, r_nJump, m_val, l_jump, l_nJump, v, n_vec, w);
}

/* Row-id expansion: src/sparseMat/mex/bmSparseMat_help_function/
 * bmSparseMat_r_nJump2index/bmSparseMat_r_nJump2index_mex.cpp:70-89. */
void port_r_nJump2index(int l_size, const int *r_nJump, int *ind)
{
    int *run = ind;
    for (int i = 0; i < l_size; i++)
        for (int j = 0; j < r_nJump[i]; j++) *run++ = i;
}

/* ------------------------------------------------------------------------
 * Direct normalised n2u gridding with on-the-fly Gaussian weights, points
 * outside the grid box clipped (not wrapped).  One function for 1/2/3
 * dimensions; follows src/gridding123/mex/bmGridder_n2u_leight3/
 * bmGridder_n2u_leight3_mex.cpp:116-358 (leight1/leight2 are the same with
 * fewer axes).  a is [nPt][nCh] (coil fastest), b is [Nu][nCh].
 * t[d] are 1-based grid coordinates.  N[d] = 1 for unused axes.
 * ---------------------------------------------------------------------- */
void port_gridder_n2u_leight(int imDim, const float *a_re, const float *a_im,
                             float *b_re, float *b_im, const float *tx,
                             const float *ty, const float *tz, const float *Dn,
                             int nCh, int nPt, int Nx, int Ny, int Nz, int nWin,
                             float kp1, float kp2)
{
    (void)kp2;
    int64_t Nu = (int64_t)Nx * Ny * Nz;
    float *norm = (float *)calloc((size_t)Nu, sizeof(float));
    memset(b_re, 0, sizeof(float) * (size_t)(Nu * nCh));
    memset(b_im, 0, sizeof(float) * (size_t)(Nu * nCh));
    float t_shift;
    int nh;
    if (nWin % 2) { t_shift = 0.5f; nh = (nWin - 1) / 2; }
    else { t_shift = 0.0f; nh = nWin / 2 - 1; }
    float nhf = (float)nh;
    float sigma = kp1;
    float g1 = 1 / ((float)sqrt(2 * 3.141592) * sigma);
    float g2 = -1 / (2 * (sigma * sigma));
    int wz = imDim > 2 ? nWin : 1, wy = imDim > 1 ? nWin : 1;
    for (int i = 0; i < nPt; i++) {
        float t[3] = {tx[i], imDim > 1 ? ty[i] : 1.0f, imDim > 2 ? tz[i] : 1.0f};
        float r[3];
        int c[3];
        for (int d = 0; d < 3; d++) {
            float cf;
            modff(t[d] + t_shift, &cf);
            r[d] = t[d] - cf;
            c[d] = (int)cf - 1;
        }
        float dn = Dn[i];
        int nz = imDim > 2 ? c[2] - nh : 0;
        float uz = -nhf;
        for (int iz = 0; iz < wz; iz++, nz++, uz += 1.0f) {
            if (imDim > 2 && (nz < 0 || nz > Nz - 1)) continue;
            int ny = imDim > 1 ? c[1] - nh : 0;
            float uy = -nhf;
            for (int iy = 0; iy < wy; iy++, ny++, uy += 1.0f) {
                if (imDim > 1 && (ny < 0 || ny > Ny - 1)) continue;
                int nx = c[0] - nh;
                float ux = -nhf;
                for (int ix = 0; ix < nWin; ix++, nx++, ux += 1.0f) {
                    if (nx < 0 || nx > Nx - 1) continue;
                    float dx = r[0] - ux, dsq = dx * dx;
                    if (imDim > 1) { float dy = r[1] - uy; dsq = dsq + dy * dy; }
                    if (imDim > 2) { float dz = r[2] - uz; dsq = dsq + dz * dz; }
                    float w = g1 * expf(g2 * dsq) * dn;
                    int64_t cell = (int64_t)nx + (int64_t)ny * Nx + (int64_t)nz * Nx * Ny;
                    norm[cell] = norm[cell] + w;
                    const float *ar = a_re + (int64_t)i * nCh, *ai = a_im + (int64_t)i * nCh;
                    float *br = b_re + cell * nCh, *bi = b_im + cell * nCh;
                    for (int k = 0; k < nCh; k++) {
                        float pr = w * ar[k], pi = w * ai[k];
                        br[k] = br[k] + pr;
                        bi[k] = bi[k] + pi;
                    }
                }
            }
        }
    }
    for (int64_t q = 0; q < Nu; q++) {
        float nv = norm[q];
        if (nv < 1e-5f) nv = 1.0f;
        for (int k = 0; k < nCh; k++) {
            b_re[q * nCh + k] /= nv;
            b_im[q * nCh + k] /= nv;
        }
    }
    free(norm);
}

/* Normalised n2u with precomputed weights w[nPt][nNb] and running index
 * jumps nJump[nPt*nNb] (first absolute, 0-based).  Follows
 * src/gridding123/mex/bmGridder_n2u/bmGridder_n2u_mex.cpp:72-201. */
void port_gridder_n2u(const float *a_re, const float *a_im, float *b_re,
                      float *b_im, const float *w, const int *nJump, int nNb,
                      int nCh, int nPt, int Nu_tot)
{
    float *norm = (float *)calloc((size_t)Nu_tot, sizeof(float));
    memset(b_re, 0, sizeof(float) * (size_t)Nu_tot * (size_t)nCh);
    memset(b_im, 0, sizeof(float) * (size_t)Nu_tot * (size_t)nCh);
    int64_t cell = 0;
    for (int i = 0; i < nPt; i++) {
        const float *ar = a_re + (int64_t)i * nCh, *ai = a_im + (int64_t)i * nCh;
        for (int j = 0; j < nNb; j++) {
            float wt = w[(int64_t)i * nNb + j];
            cell += nJump[(int64_t)i * nNb + j];
            norm[cell] = norm[cell] + wt;
            float *br = b_re + cell * nCh, *bi = b_im + cell * nCh;
            for (int k = 0; k < nCh; k++) {
                float pr = wt * ar[k], pi = wt * ai[k];
                br[k] = br[k] + pr;
                bi[k] = bi[k] + pi;
            }
        }
    }
    for (int64_t q = 0; q < Nu_tot; q++) {
        float nv = norm[q];
        if (nv < 1e-5f) nv = 1.0f;
        for (int k = 0; k < nCh; k++) {
            b_re[q * nCh + k] /= nv;
            b_im[q * nCh + k] /= nv;
        }
    }
    free(norm);
}

/* Count of (point, neighbour) pairs inside the box with d^2 < (nWin/2)^2.
 * Follows src/gridding123/mex/bmGriddingMatrix_secret_length3/
 * bmGriddingMatrix_secret_length3_mex.cpp:70-222 (1/2-D analogues). */
long long port_secret_length(int imDim, const float *tx, const float *ty,
                             const float *tz, int nPt, int Nx, int Ny, int Nz,
                             int nWin)
{
    long long count = 0;
    float dmax = ((float)nWin / 2.0f) * ((float)nWin / 2.0f);
    float t_shift;
    int nh;
    if (nWin % 2) { t_shift = 0.5f; nh = (nWin - 1) / 2; }
    else { t_shift = 0.0f; nh = nWin / 2 - 1; }
    float nhf = (float)nh;
    int wz = imDim > 2 ? nWin : 1, wy = imDim > 1 ? nWin : 1;
    for (int i = 0; i < nPt; i++) {
        float t[3] = {tx[i], imDim > 1 ? ty[i] : 1.0f, imDim > 2 ? tz[i] : 1.0f};
        float r[3];
        int c[3];
        for (int d = 0; d < 3; d++) {
            float cf;
            modff(t[d] + t_shift, &cf);
            r[d] = t[d] - cf;
            c[d] = (int)cf - 1;
        }
        int nz = imDim > 2 ? c[2] - nh : 0;
        float uz = -nhf;
        for (int iz = 0; iz < wz; iz++, nz++, uz += 1.0f) {
            if (imDim > 2 && (nz < 0 || nz > Nz - 1)) continue;
            int ny = imDim > 1 ? c[1] - nh : 0;
            float uy = -nhf;
            for (int iy = 0; iy < wy; iy++, ny++, uy += 1.0f) {
                if (imDim > 1 && (ny < 0 || ny > Ny - 1)) continue;
                int nx = c[0] - nh;
                float ux = -nhf;
                for (int ix = 0; ix < nWin; ix++, nx++, ux += 1.0f) {
                    if (nx < 0 || nx > Nx - 1) continue;
                    float dx = r[0] - ux, dsq = dx * dx;
                    if (imDim > 1) { float dy = r[1] - uy; dsq = dsq + dy * dy; }
                    if (imDim > 2) { float dz = r[2] - uz; dsq = dsq + dz * dz; }
                    if (dsq >= dmax) continue;
                    count++;
                }
            }
        }
    }
    return count;
}
