// ADMM total-variation solvers and the gridded reconstruction.
//
//   bmSteva               spatial TV, one frame   (bmSteva.m:6-138)
//   Morphosia solvers     temporal TV (Teva) or quadratic (Sensitiva) penalties over
//                         frame chains and sheets, motion-compensated by deformation
//                         matrices Tu/Tut (imDim_morphosia/nonCartesian/*.m)
//   bmNasha               n2u with Gn + centred ifft + deapodisation (bmNasha.m:6-91)
//   bmMathilda            the same tail after a direct gridder (bmMathilda.m:14-125)
//
// The finite differences, their transposes and the soft-threshold prox are
// fused elementwise kernels; inner CG scalars are FP64 on the device.
#include <algorithm>
#include <cmath>

#include "fft.hpp"
#include "recon.hpp"

namespace mona {

namespace {

inline unsigned nb(int64_t n) { return blocks_for(n, 256); }

struct Dims {
    int n[3];
    float inv_dx[3];
    int imDim;
    int64_t nu;
};

__device__ __forceinline__ void coords(int64_t i, const Dims& d, int c[3]) {
    c[0] = (int)(i % d.n[0]);
    c[1] = (int)((i / d.n[0]) % d.n[1]);
    c[2] = (int)(i / ((int64_t)d.n[0] * d.n[1]));
}

__device__ __forceinline__ int64_t stride_of(const Dims& d, int a) {
    return a == 0 ? 1 : (a == 1 ? (int64_t)d.n[0] : (int64_t)d.n[0] * d.n[1]);
}

// g[a][i] = (x[i] - x[i - e_a (cyclic)]) / dX_a, 0 on the first slice of
// axis a (bmBackGradient_n.m:36-44).  With zu != NULL writes
// out = (z - u) - g instead (the ADMM residual res_z).
__global__ void k_grad(const float2* __restrict__ x, float2* __restrict__ out, const float2* __restrict__ z,
                       const float2* __restrict__ u, Dims d) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.nu) return;
    int c[3];
    coords(i, d, c);
    float2 xi = x[i];
    for (int a = 0; a < d.imDim; a++) {
        float2 g = make_float2(0.f, 0.f);
        if (c[a] != 0) {
            float2 xm = x[i - stride_of(d, a)];
            g = make_float2((xi.x - xm.x) * d.inv_dx[a], (xi.y - xm.y) * d.inv_dx[a]);
        }
        int64_t o = (int64_t)a * d.nu + i;
        if (z) {
            float2 zz = z[o], uu = u[o];
            out[o] = make_float2((zz.x - uu.x) - g.x, (zz.y - uu.y) - g.y);
        } else {
            out[o] = g;
        }
    }
}

// out[i] (+)= coef * sum_a (w_a[i] * [c_a != 0] - w_a[i + e_a (cyclic)]) / dX_a
// (bmBackGradient_nT.m:34-44 summed over axes by bmBackGradientT.m).
__global__ void k_gradT(const float2* __restrict__ w, float2* __restrict__ out, Dims d, float coef, int accumulate) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.nu) return;
    int c[3];
    coords(i, d, c);
    float2 s = make_float2(0.f, 0.f);
    for (int a = 0; a < d.imDim; a++) {
        const float2* wa = w + (int64_t)a * d.nu;
        float2 gi = c[a] != 0 ? wa[i] : make_float2(0.f, 0.f);
        int64_t j = (c[a] == d.n[a] - 1) ? i - (int64_t)(d.n[a] - 1) * stride_of(d, a) : i + stride_of(d, a);
        float2 wn = wa[j];
        s.x += (gi.x - wn.x) * d.inv_dx[a];
        s.y += (gi.y - wn.y) * d.inv_dx[a];
    }
    s.x *= coef;
    s.y *= coef;
    if (accumulate) {
        float2 o = out[i];
        s.x += o.x;
        s.y += o.y;
    }
    out[i] = s;
}

__device__ __forceinline__ float soft(float v, float s) {
    return (v > s) ? (v - s) : ((v < -s) ? (v + s) : 0.0f);
}

// w = Fx + u ; z = prox_{s}(w) (re and im separately, bmProx_oneNorm.m:30-41) ; u = w - z
__global__ void k_prox_update(const float2* __restrict__ Fx, float2* __restrict__ z, float2* __restrict__ u,
                              int64_t n, float s) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float2 f = Fx[i], uu = u[i];
    float2 w = make_float2(uu.x + f.x, uu.y + f.y);
    float2 zz = make_float2(soft(w.x, s), soft(w.y, s));
    z[i] = zz;
    u[i] = make_float2(w.x - zz.x, w.y - zz.y);
}

// Frame f = i + n1*j of an n1 x n2 cell array (MATLAB order) moved by d
// along axis 0 (i) or 1 (j), cyclically.
__device__ __forceinline__ int frame_shift(int f, int n1, int n2, int axis, int d) {
    int i = f % n1, j = f / n1;
    if (axis == 0) i = (i + d + n1) % n1;
    else j = (j + d + n2) % n2;
    return i + n1 * j;
}

struct FrameGrid {
    int nFr, n1, n2;
    int64_t nu;
};

// Difference term of the Morphosia solvers: g = scale * (xa_f - x_shift(f))
// with xa_f = Tu_f x_f (private_F / private_F1 / private_F2 of
// bmTevaMorphosia_chain.m:266-272, bmTevaDuoMorphosia_chain.m:313-329,
// bmTevaMorphosia_sheet.m:322-344; xa = x for identity deformations).
// mode 0: out = g; 1: out = (z - u) - g (ADMM residual); 2: out = 0 - g.
__global__ void k_tdiff(const float2* __restrict__ xa, const float2* __restrict__ x, float2* __restrict__ out,
                        const float2* __restrict__ z, const float2* __restrict__ u, FrameGrid G, int axis, int dir,
                        float scale, int mode) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)G.nFr * G.nu) return;
    int f = (int)(t / G.nu);
    int64_t i = t - (int64_t)f * G.nu;
    int fm = frame_shift(f, G.n1, G.n2, axis, dir);
    float2 a = xa[t], b = x[(int64_t)fm * G.nu + i];
    float2 g = make_float2(a.x - b.x, a.y - b.y);
    if (scale != 1.0f) g = make_float2(scale * g.x, scale * g.y);
    if (mode == 1) {
        float2 zz = z[t], uu = u[t];
        out[t] = make_float2((zz.x - uu.x) - g.x, (zz.y - uu.y) - g.y);
    } else if (mode == 2) {
        out[t] = make_float2(0.0f - g.x, 0.0f - g.y);
    } else {
        out[t] = g;
    }
}

// out_f += coef * (za_f - z_shift(f, -dir)), za_f = Tut_f z_f (private_dagF
// of the same files: bmTevaMorphosia_chain.m:285-291)
__global__ void k_tdiffT_acc(const float2* __restrict__ za, const float2* __restrict__ z, float2* __restrict__ out,
                             FrameGrid G, int axis, int dir, float coef) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)G.nFr * G.nu) return;
    int f = (int)(t / G.nu);
    int64_t i = t - (int64_t)f * G.nu;
    int fp = frame_shift(f, G.n1, G.n2, axis, -dir);
    float2 a = za[t], b = z[(int64_t)fp * G.nu + i];
    float2 o = out[t];
    out[t] = make_float2(o.x + coef * (a.x - b.x), o.y + coef * (a.y - b.y));
}

// out[i] = sum_f a[f][i], frames added in order (bmSensaMorphosia_chain.m:63-70)
__global__ void k_frame_sum(const float2* __restrict__ a, float2* __restrict__ out, int nFr, int64_t nu) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nu) return;
    float2 s = a[i];
    for (int f = 1; f < nFr; f++) {
        float2 v = a[(int64_t)f * nu + i];
        s.x += v.x;
        s.y += v.y;
    }
    out[i] = s;
}

constexpr int kRT = 256;
constexpr int kRB = 1024;

// FP64 sum of |Re| + |Im|
__global__ void __launch_bounds__(kRT) k_abs1_partial(const float2* a, int64_t n, double* partial) {
    __shared__ double sh[kRT];
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kRT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kRT) {
        float2 v = a[i];
        s += fabs((double)v.x) + fabs((double)v.y);
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = kRT / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

__global__ void __launch_bounds__(kRT) k_sum_partial(const double* partial, int np, double* out, double scale) {
    __shared__ double sh[kRT];
    double s = 0.0;
    for (int i = threadIdx.x; i < np; i += kRT) s += partial[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = kRT / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0] * scale;
}

void abs1(const float2* a, int64_t n, double scale, double* out, double* partial, hipStream_t s) {
    int nbk = (int)std::min<int64_t>(kRB, std::max<int64_t>(1, (n + kRT - 1) / kRT));
    k_abs1_partial<<<nbk, kRT, 0, s>>>(a, n, partial);
    MONA_HIP(hipGetLastError());
    k_sum_partial<<<1, kRT, 0, s>>>(partial, nbk, out, scale);
    MONA_HIP(hipGetLastError());
}

// regul = sum of the per-term values h[3..3+nT); objective = 0.5 * dafi + 0.5 * delta * regul
__global__ void k_objective(double* h, double delta, int nT) {
    double r = h[3];
    for (int t = 1; t < nT; t++) r += h[3 + t];
    h[2] = r;
    h[0] = 0.5 * h[1] + 0.5 * delta * h[2];
}

// hist[n*ctr .. n*ctr+n-1] = slot[0..n-1]; ctr += 1
__global__ void k_hist_commit(double* hist, const double* slot, int* ctr, int n) {
    int c = *ctr;
    for (int i = 0; i < n; i++) hist[(int64_t)n * c + i] = slot[i];
    *ctr = c + 1;
}

// Gridded-recon epilogue (bmNasha.m:66-87): per-coil image
// v_c = grid[crop(i)][c] * F * K[i]; combined sum(conj(C) v)/sum|C|^2.
__global__ void k_nasha_epilogue(const float2* __restrict__ grid, const float* __restrict__ K,
                                 const float2* __restrict__ C, float2* __restrict__ x, Geom g, float F) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.nu) return;
    int ix = (int)(i % g.n[0]);
    int iy = (int)((i / g.n[0]) % g.n[1]);
    int iz = (int)(i / ((int64_t)g.n[0] * g.n[1]));
    int64_t cell = (ix + g.off[0]) + (int64_t)g.N[0] * ((iy + g.off[1]) + (int64_t)g.N[1] * (iz + g.off[2]));
    const float2* gp = grid + cell * g.nCh;
    float k = K[i];
    if (C) {
        float2 num = make_float2(0.f, 0.f);
        float den = 0.f;
        for (int c = 0; c < g.nCh; c++) {
            float2 v = gp[c];
            v = make_float2(v.x * F * k, v.y * F * k);
            float2 cc = C[i * g.nCh + c];
            num.x += cc.x * v.x + cc.y * v.y;     // conj(C) * v
            num.y += cc.x * v.y - cc.y * v.x;
            den += cc.x * cc.x + cc.y * cc.y;
        }
        x[i] = make_float2(num.x / den, num.y / den);
    } else {
        for (int c = 0; c < g.nCh; c++) {
            float2 v = gp[c];
            x[i * g.nCh + c] = make_float2(v.x * F * k, v.y * F * k);
        }
    }
}

std::vector<float> schedule(const float* v, int nIter) {
    // linspace(v0, v1, nIter) in single (private_init_delta_rho, bmSteva.m:140-157)
    std::vector<float> out(std::max(nIter, 1));
    double a = std::fabs((double)v[0]), b = std::fabs((double)v[1]);
    for (int i = 0; i < nIter; i++)
        out[i] = (float)(nIter == 1 ? b : a + (b - a) * (double)i / (double)(nIter - 1));
    return out;
}

}  // namespace

// ================================================================ bmSteva
void steva(Plan& P, int f, float2* x, float2* z, float2* u, bool zu_given, const float2* y, const float* ve,
           float ve_max, const float* delta_in, const float* rho_in, int nIter, int nCGD, double* hist_out,
           hipStream_t s) {
    MONA_CHECK(nIter >= 0 && nCGD >= 0, MONA_ERR_ARG, "nIter and nCGD must be >= 0");
    const Frame& fr = P.frame(f);
    int64_t Pn = fr.Gu->l_size, ny = Pn * P.nCh, nu = P.nu;
    int D = P.imDim;
    std::vector<float> delta = schedule(delta_in, nIter), rho = schedule(rho_in, nIter);
    Dims d{};
    for (int a = 0; a < 3; a++) {
        d.n[a] = P.n[a];
        d.inv_dx[a] = a < D ? 1.0f / ((1.0f / (float)P.dK[a]) / (float)P.N[a]) : 1.0f;
    }
    d.imDim = D;
    d.nu = nu;
    Work& w = P.work;
    w.hy.ensure((size_t)Pn * sizeof(float));
    w.k1.ensure((size_t)ny * sizeof(float2));
    w.k2.ensure((size_t)ny * sizeof(float2));
    w.i1.ensure((size_t)nu * sizeof(float2));
    w.i2.ensure((size_t)nu * sizeof(float2));
    w.z1.ensure((size_t)D * nu * sizeof(float2));
    w.z2.ensure((size_t)D * nu * sizeof(float2));
    w.sc.ensure(sizeof(CgScalars));
    w.partial.ensure(kRB * sizeof(double));
    w.hist.ensure((size_t)std::max(nIter, 1) * 2 * sizeof(double));
    float* hy = w.hy.as<float>();
    float2 *res_y = w.k1.as<float2>(), *Mp = w.k2.as<float2>();
    float2 *dag = w.i1.as<float2>(), *p = w.i2.as<float2>();
    float2 *res_z = w.z1.as<float2>(), *Fp = w.z2.as<float2>();
    CgScalars* sc = w.sc.as<CgScalars>();
    double* part = w.partial.as<double>();
    double* hist = w.hist.as<double>();
    // data and weights in the frame's tile order for the whole solve
    w.hy0.ensure((size_t)Pn * sizeof(float));
    w.ys.ensure((size_t)ny * sizeof(float2));
    compute_hy(ve, ve_max, w.hy0.as<float>(), Pn, hy_scratch(w), s);
    P.weights_to_tile_order(f, w.hy0.as<float>(), hy, s);
    P.to_tile_order(f, y, w.ys.as<float2>(), s);
    y = w.ys.as<float2>();
    const float invHX = (float)(1.0 / P.HX);
    const float HZ = (float)P.HX;
    if (!zu_given) {
        k_grad<<<nb(nu), 256, 0, s>>>(x, z, nullptr, nullptr, d);
        MONA_HIP(hipGetLastError());
        fill_f32(reinterpret_cast<float*>(u), 0.0f, (int64_t)D * nu * 2, s);
    }
    for (int c = 0; c < nIter; c++) {
        float rc = rho[c];
        cg_reset(sc, s);
        P.forward(f, x, Mp, s);
        vsub(y, Mp, res_y, ny, s);
        k_grad<<<nb(nu), 256, 0, s>>>(x, res_z, z, u, d);                  // res_z = (z-u) - Fx
        MONA_HIP(hipGetLastError());
        P.adjoint(f, res_y, dag, true, hy, invHX, s);                      // dagM
        k_gradT<<<nb(nu), 256, 0, s>>>(res_z, dag, d, rc * invHX * HZ, 1);  // + rho/HX F^T(HZ res_z)
        MONA_HIP(hipGetLastError());
        vcopy(dag, p, nu, s);
        sqnorm(dag, nu, nullptr, 1, P.HX, &sc->sqn, false, part, s);
        for (int i = 1; i <= nCGD; i++) {
            P.forward(f, p, Mp, s);
            k_grad<<<nb(nu), 256, 0, s>>>(p, Fp, nullptr, nullptr, d);
            MONA_HIP(hipGetLastError());
            sqnorm(Mp, ny, hy, P.nCh, 1.0, &sc->den, false, part, s);
            sqnorm(Fp, (int64_t)D * nu, nullptr, 1, (double)rc * HZ, &sc->den, true, part, s);
            cg_alpha(sc, s);
            vaxpy(x, p, &sc->a, 1.0f, nu, s);
            if (i == nCGD) break;
            vaxpy(res_y, Mp, &sc->a, -1.0f, ny, s);
            vaxpy(res_z, Fp, &sc->a, -1.0f, (int64_t)D * nu, s);
            P.adjoint(f, res_y, dag, true, hy, invHX, s);
            k_gradT<<<nb(nu), 256, 0, s>>>(res_z, dag, d, rc * invHX * HZ, 1);
            MONA_HIP(hipGetLastError());
            sqnorm(dag, nu, nullptr, 1, P.HX, &sc->sqn_new, false, part, s);
            cg_beta(sc, s);
            vxpby(p, dag, &sc->b, nu, s);
        }
        // z, u update
        k_grad<<<nb(nu), 256, 0, s>>>(x, Fp, nullptr, nullptr, d);
        MONA_HIP(hipGetLastError());
        k_prox_update<<<nb((int64_t)D * nu), 256, 0, s>>>(Fp, z, u, (int64_t)D * nu, delta[c] / rc);
        MONA_HIP(hipGetLastError());
        // monitoring: R = <r, HY r>, TV = HZ * sum |Re| + |Im| of grad x
        P.forward(f, x, Mp, s);
        vsub(y, Mp, res_y, ny, s);
        sqnorm(res_y, ny, hy, P.nCh, 1.0, hist + 2 * c, false, part, s);
        abs1(Fp, (int64_t)D * nu, (double)HZ, hist + 2 * c + 1, part, s);
        P.witness(c, x, nu, s);
    }
    if (hist_out && nIter > 0) {
        MONA_HIP(hipMemcpyAsync(hist_out, hist, (size_t)nIter * 2 * sizeof(double), hipMemcpyDeviceToHost, s));
        MONA_HIP(hipStreamSynchronize(s));
    }
}

// ================================================================ bmTevaMorphosia_chain
// ================================================================ Morphosia solvers
// One difference term: F_t x_f = scale (Tu_{set,f} x_f - x_shift(f, axis, dir)).
struct MorphTerm {
    int axis, dir;
    float scale;
    int set;
};

struct MorphSpec {
    int nT;
    MorphTerm t[2];
    bool l1;                  // ADMM with soft threshold (Teva) or quadratic penalty (Sensitiva)
    bool first_dagF_scaled;   // bmSensitivaMorphosia_sheet.m:118-119 leaves delta off
};

MorphSpec morph_spec(int solver) {
    MorphSpec m{};
    m.l1 = true;
    m.first_dagF_scaled = true;
    switch (solver) {
        case MONA_MORPH_TEVA_CHAIN:  // bmTevaMorphosia_chain.m:266-291
            m.nT = 1;
            m.t[0] = {0, -1, 1.0f, 0};
            break;
        case MONA_MORPH_TEVA_DUO_CHAIN:  // bmTevaDuoMorphosia_chain.m:313-359
            m.nT = 2;
            m.t[0] = {0, -1, 0.5f, 0};
            m.t[1] = {0, +1, 0.5f, 1};
            break;
        case MONA_MORPH_TEVA_SHEET:  // bmTevaMorphosia_sheet.m:322-379
            m.nT = 2;
            m.t[0] = {0, -1, 1.0f, 0};
            m.t[1] = {1, -1, 1.0f, 1};
            break;
        case MONA_MORPH_SENSITIVA_CHAIN:  // bmSensitivaMorphosia_chain.m:251-277
            m.nT = 1;
            m.t[0] = {0, -1, 1.0f, 0};
            m.l1 = false;
            break;
        case MONA_MORPH_SENSITIVA_DUO_CHAIN:  // bmSensitivaDuoMorphosia_chain.m:150-183
            m.nT = 2;
            m.t[0] = {0, -1, 0.5f, 0};
            m.t[1] = {0, +1, 0.5f, 1};
            m.l1 = false;
            break;
        case MONA_MORPH_SENSITIVA_SHEET:  // bmSensitivaMorphosia_sheet.m
            m.nT = 2;
            m.t[0] = {0, -1, 1.0f, 0};
            m.t[1] = {1, -1, 1.0f, 1};
            m.l1 = false;
            m.first_dagF_scaled = false;
            break;
        default:
            MONA_CHECK(false, MONA_ERR_ARG, "unknown Morphosia solver");
    }
    return m;
}

// The Morphosia solvers over the plan's frames, regul_mode 'normal'
// (src/fourier123/solver_function/imDim_morphosia/nonCartesian/: Teva = ADMM
// with z_t = F_t x split off and soft-thresholded, Sensitiva = quadratic
// penalty delta*HZ*||F_t x||^2 minimised by nCGD CG steps per iteration).
// Frames f = i + nFr1*j of an nFr1 x nFr2 cell array; z, u hold nT blocks of
// nFr images.  hist: nIter x {objective, dafi, regul}.
void morphosia(Plan& P, int solver, int nFr1, float2* x, float2* z, float2* u, bool zu_given, const float2* y,
               const long long* y_off, const float* ve, const long long* ve_off, float ve_max,
               const float* delta_in, const float* rho_in, int nIter, int nCGD, double* hist_out, hipStream_t s) {
    const MorphSpec spec = morph_spec(solver);
    int nFr = (int)P.frames.size();
    MONA_CHECK(nFr >= 1, MONA_ERR_STATE, "the plan has no frames");
    MONA_CHECK(ve_max > 0, MONA_ERR_ARG, "ve_max must be > 0");
    MONA_CHECK(nFr1 >= 1 && nFr % nFr1 == 0, MONA_ERR_ARG, "nFr1 must divide the number of frames");
    MONA_CHECK(!spec.l1 || (z && u && rho_in), MONA_ERR_ARG, "the ADMM solvers need z, u and rho");
    const FrameGrid FG{nFr, nFr1, nFr / nFr1, P.nu};
    int64_t nu = P.nu, nx = (int64_t)nFr * nu;
    std::vector<int64_t> Pf(nFr), koff(nFr + 1, 0);
    for (int f = 0; f < nFr; f++) {
        Pf[f] = P.frame(f).Gu->l_size;
        koff[f + 1] = koff[f] + Pf[f] * P.nCh;
    }
    int64_t ny = koff[nFr];
    std::vector<float> delta = schedule(delta_in, nIter);
    std::vector<float> rho = spec.l1 ? schedule(rho_in, nIter) : delta;  // weight of the penalty terms
    Work& w = P.work;
    w.hy.ensure((size_t)(ny / P.nCh) * sizeof(float));
    w.k1.ensure((size_t)ny * sizeof(float2));
    w.k2.ensure((size_t)ny * sizeof(float2));
    w.z1.ensure((size_t)nx * spec.nT * sizeof(float2));
    w.z2.ensure((size_t)nx * spec.nT * sizeof(float2));
    w.z3.ensure((size_t)nx * sizeof(float2));
    w.z4.ensure((size_t)nx * sizeof(float2));
    w.sc.ensure(sizeof(CgScalars));
    w.partial.ensure(kRB * sizeof(double));
    w.hist.ensure((size_t)std::max(nIter, 1) * 3 * sizeof(double));
    float* hy = w.hy.as<float>();
    float2 *res_y = w.k1.as<float2>(), *Mp = w.k2.as<float2>();
    float2 *dag = w.z3.as<float2>(), *p = w.z4.as<float2>();
    float2* res_z[2] = {w.z1.as<float2>(), w.z1.as<float2>() + nx};  // per term
    float2* Fp[2] = {w.z2.as<float2>(), w.z2.as<float2>() + nx};
    float2* zt[2] = {z, z ? z + nx : nullptr};
    float2* ut[2] = {u, u ? u + nx : nullptr};
    CgScalars* sc = w.sc.as<CgScalars>();
    double* part = w.partial.as<double>();
    double* hist = w.hist.as<double>();
    const float invHX = (float)(1.0 / P.HX);
    const float HZ = (float)P.HX;
    // HY and the data per frame, packed like the k-space vectors and in each
    // frame's tile order
    w.hy0.ensure((size_t)(ny / P.nCh) * sizeof(float));
    w.ys.ensure((size_t)ny * sizeof(float2));
    for (int f = 0; f < nFr; f++) {
        compute_hy(ve + ve_off[f], ve_max, w.hy0.as<float>() + koff[f] / P.nCh, Pf[f], hy_scratch(w), s);
        P.weights_to_tile_order(f, w.hy0.as<float>() + koff[f] / P.nCh, hy + koff[f] / P.nCh, s);
        P.to_tile_order(f, y + y_off[f], w.ys.as<float2>() + koff[f], s);
    }
    const float2* ysorted = w.ys.as<float2>();
    auto yF = [&](int f) { return ysorted + koff[f]; };
    // deformations: Tu_f x_f and Tut_f z_f go to a scratch image set, frames
    // without a deformation are copied
    auto T_of = [&](int set, int f, bool transposed) -> const SparseDev* {
        if (f >= (int)P.deform[set].size()) return nullptr;
        return transposed ? P.deform[set][f].second : P.deform[set][f].first;
    };
    bool any_deform = false;
    for (int t = 0; t < spec.nT; t++)
        for (int f = 0; f < nFr; f++) any_deform = any_deform || T_of(spec.t[t].set, f, false);
    if (any_deform) w.z6.ensure((size_t)nx * sizeof(float2));
    float2* tx = any_deform ? w.z6.as<float2>() : nullptr;
    auto warp = [&](int set, const float2* v, bool transposed, hipStream_t s) -> const float2* {
        bool mine = false;
        for (int f = 0; f < nFr; f++) mine = mine || T_of(set, f, false);
        if (!mine) return v;
        for (int f = 0; f < nFr; f++) {
            const SparseDev* T = T_of(set, f, transposed);
            if (T) spmv_c2_rowmajor(*T, v + (int64_t)f * nu, tx + (int64_t)f * nu, 1, nullptr, s);
            else vcopy(v + (int64_t)f * nu, tx + (int64_t)f * nu, nu, s);
        }
        return tx;
    };
    auto Fop = [&](int t, const float2* v, float2* out, int mode, float scale, hipStream_t s) {
        const MorphTerm& m = spec.t[t];
        const float2* va = warp(m.set, v, false, s);
        k_tdiff<<<nb(nx), 256, 0, s>>>(va, v, out, mode == 1 ? zt[t] : nullptr, mode == 1 ? ut[t] : nullptr, FG,
                                       m.axis, m.dir, scale, mode);
        MONA_HIP(hipGetLastError());
    };
    auto FtAcc = [&](int t, const float2* zz, float2* out, float coef, hipStream_t s) {
        const MorphTerm& m = spec.t[t];
        const float2* za = warp(m.set, zz, true, s);
        k_tdiffT_acc<<<nb(nx), 256, 0, s>>>(za, zz, out, FG, m.axis, m.dir, coef * m.scale);
        MONA_HIP(hipGetLastError());
    };
    // operators: one launch per gridding op over all frames (batched form), or
    // frame by frame
    if (P.use_batch && (!P.Gu_b || (int)P.koff_b.size() != nFr + 1)) P.build_batch(s);
    hipStream_t cs = s;  // stream the iteration body is issued on
    auto M = [&](const float2* xx, float2* out) {
        if (P.use_batch) { P.forward_batch(xx, out, cs); return; }
        for (int f = 0; f < nFr; f++) P.forward(f, xx + (int64_t)f * nu, out + koff[f], cs);
    };
    auto Mh = [&](const float2* r, float2* out) {
        if (P.use_batch) { P.adjoint_batch(r, out, hy, invHX, cs); return; }
        for (int f = 0; f < nFr; f++)
            P.adjoint(f, r + koff[f], out + (int64_t)f * nu, true, hy + koff[f] / P.nCh, invHX, cs);
    };
    if (spec.l1 && !zu_given) {
        for (int t = 0; t < spec.nT; t++) Fop(t, x, zt[t], 0, spec.t[t].scale, s);
        fill_f32(reinterpret_cast<float*>(u), 0.0f, nx * 2 * spec.nT, s);
    }
    // monitoring values go to a fixed slot and are committed to hist[3*iter]
    // by a kernel that reads and advances a device counter, so the body has
    // no per-iteration pointers and can be replayed from a graph
    w.z5.ensure(6 * sizeof(double) + sizeof(int));
    double* hslot = w.z5.as<double>();
    int* hctr = reinterpret_cast<int*>(hslot + 6);
    fill_i32(hctr, 0, 1, s);
    auto body = [&](float dl, float rc) {
        hipStream_t s = cs;
        const float wz = rc * invHX * HZ;   // penalty weight times 1/HX times HZ
        cg_reset(sc, s);
        M(x, Mp);
        for (int f = 0; f < nFr; f++) vsub(yF(f), Mp + koff[f], res_y + koff[f], koff[f + 1] - koff[f], s);
        for (int t = 0; t < spec.nT; t++) Fop(t, x, res_z[t], spec.l1 ? 1 : 2, spec.t[t].scale, s);
        Mh(res_y, dag);
        for (int t = 0; t < spec.nT; t++) FtAcc(t, res_z[t], dag, spec.first_dagF_scaled ? wz : invHX * HZ, s);
        vcopy(dag, p, nx, s);
        sqnorm(dag, nx, nullptr, 1, P.HX, &sc->sqn, false, part, s);
        for (int j = 1; j <= nCGD; j++) {
            M(p, Mp);
            for (int t = 0; t < spec.nT; t++) Fop(t, p, Fp[t], 0, spec.t[t].scale, s);
            sqnorm(Mp, ny, hy, P.nCh, 1.0, &sc->den, false, part, s);
            for (int t = 0; t < spec.nT; t++) sqnorm(Fp[t], nx, nullptr, 1, (double)rc * HZ, &sc->den, true, part, s);
            cg_alpha(sc, s);
            vaxpy(x, p, &sc->a, 1.0f, nx, s);
            if (j == nCGD) break;
            vaxpy(res_y, Mp, &sc->a, -1.0f, ny, s);
            for (int t = 0; t < spec.nT; t++) vaxpy(res_z[t], Fp[t], &sc->a, -1.0f, nx, s);
            Mh(res_y, dag);
            for (int t = 0; t < spec.nT; t++) FtAcc(t, res_z[t], dag, wz, s);
            sqnorm(dag, nx, nullptr, 1, P.HX, &sc->sqn_new, false, part, s);
            cg_beta(sc, s);
            vxpby(p, dag, &sc->b, nx, s);
        }
        if (spec.l1)
            for (int t = 0; t < spec.nT; t++) {
                Fop(t, x, Fp[t], 0, spec.t[t].scale, s);
                k_prox_update<<<nb(nx), 256, 0, s>>>(Fp[t], zt[t], ut[t], nx, dl / rc);
                MONA_HIP(hipGetLastError());
            }
        // monitoring: dafi, regul of the unscaled differences (private_dafi_regul), objective
        M(x, Mp);
        for (int f = 0; f < nFr; f++) vsub(yF(f), Mp + koff[f], res_y + koff[f], koff[f + 1] - koff[f], s);
        sqnorm(res_y, ny, hy, P.nCh, 1.0, hslot + 1, false, part, s);
        for (int t = 0; t < spec.nT; t++) {
            Fop(t, x, Fp[0], 0, 1.0f, s);
            if (spec.l1) abs1(Fp[0], nx, (double)HZ, hslot + 3 + t, part, s);
            else sqnorm(Fp[0], nx, nullptr, 1, (double)HZ, hslot + 3 + t, false, part, s);
        }
        k_objective<<<1, 1, 0, s>>>(hslot, (double)dl, spec.nT);
        MONA_HIP(hipGetLastError());
        k_hist_commit<<<1, 1, 0, s>>>(hist, hslot, hctr, 3);
        MONA_HIP(hipGetLastError());
    };
    bool constant = true;
    for (int c = 1; c < nIter; c++) constant = constant && delta[c] == delta[0] && rho[c] == rho[0];
    // witness snapshots differ per iteration: those runs stay eager
    const bool watched = P.wit_slices || P.wit_full;
    if (P.use_graph && constant && nIter > 1 && !watched) {
        // iteration 0 eagerly (allocates every work buffer), then one
        // captured iteration replayed nIter - 1 times, on a private stream
        // ordered after / before the caller's stream
        struct GraphRun {  // releases everything on every exit path
            hipStream_t gs = nullptr;
            hipEvent_t e0 = nullptr, e1 = nullptr;
            hipGraph_t graph = nullptr;
            hipGraphExec_t exec = nullptr;
            ~GraphRun() {
                if (gs) {
                    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
                    if (hipStreamIsCapturing(gs, &st) == hipSuccess && st != hipStreamCaptureStatusNone) {
                        hipGraph_t g = nullptr;
                        (void)hipStreamEndCapture(gs, &g);
                        if (g) (void)hipGraphDestroy(g);
                    }
                    (void)hipStreamSynchronize(gs);
                }
                if (exec) (void)hipGraphExecDestroy(exec);
                if (graph) (void)hipGraphDestroy(graph);
                if (e0) (void)hipEventDestroy(e0);
                if (e1) (void)hipEventDestroy(e1);
                if (gs) (void)hipStreamDestroy(gs);
            }
        } g;
        MONA_HIP(hipStreamCreateWithFlags(&g.gs, hipStreamNonBlocking));
        MONA_HIP(hipEventCreateWithFlags(&g.e0, hipEventDisableTiming));
        MONA_HIP(hipEventCreateWithFlags(&g.e1, hipEventDisableTiming));
        MONA_HIP(hipEventRecord(g.e0, s));
        MONA_HIP(hipStreamWaitEvent(g.gs, g.e0, 0));
        cs = g.gs;
        struct RestoreStream {
            hipStream_t& r;
            hipStream_t v;
            ~RestoreStream() { r = v; }
        } restore{cs, s};
        body(delta[0], rho[0]);
        MONA_HIP(hipStreamBeginCapture(g.gs, hipStreamCaptureModeThreadLocal));
        body(delta[0], rho[0]);
        MONA_HIP(hipStreamEndCapture(g.gs, &g.graph));
        MONA_HIP(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
        for (int c = 1; c < nIter; c++) MONA_HIP(hipGraphLaunch(g.exec, g.gs));
        MONA_HIP(hipEventRecord(g.e1, g.gs));
        MONA_HIP(hipStreamWaitEvent(s, g.e1, 0));
    } else {
        for (int c = 0; c < nIter; c++) {
            body(delta[c], rho[c]);
            P.witness(c, x, nx, s);
        }
    }
    if (hist_out && nIter > 0) {
        MONA_HIP(hipMemcpyAsync(hist_out, hist, (size_t)nIter * 3 * sizeof(double), hipMemcpyDeviceToHost, s));
        MONA_HIP(hipStreamSynchronize(s));
    }
}

// ================================================================ bmSensaMorphosia_chain
// One image x seen through every frame's deformation: minimises
// sum_f ||M_f Tu_f x - y_f||^2_HY by nCGD CG steps per iteration
// (bmSensaMorphosia_chain.m:6-104; deformation set 0).  hist[c] = R, the
// HY-weighted residual after iteration c.
void sensa_morphosia(Plan& P, float2* x, const float2* y, const long long* y_off, const float* ve,
                     const long long* ve_off, float ve_max, int nIter, int nCGD, double* hist_out, hipStream_t s) {
    int nFr = (int)P.frames.size();
    MONA_CHECK(nFr >= 1, MONA_ERR_STATE, "the plan has no frames");
    MONA_CHECK(ve_max > 0, MONA_ERR_ARG, "ve_max must be > 0");
    int64_t nu = P.nu, nx = (int64_t)nFr * nu;
    std::vector<int64_t> Pf(nFr), koff(nFr + 1, 0);
    for (int f = 0; f < nFr; f++) {
        Pf[f] = P.frame(f).Gu->l_size;
        koff[f + 1] = koff[f] + Pf[f] * P.nCh;
    }
    int64_t ny = koff[nFr];
    Work& w = P.work;
    w.hy.ensure((size_t)(ny / P.nCh) * sizeof(float));
    w.hy0.ensure((size_t)(ny / P.nCh) * sizeof(float));
    w.ys.ensure((size_t)ny * sizeof(float2));
    w.k1.ensure((size_t)ny * sizeof(float2));
    w.k2.ensure((size_t)ny * sizeof(float2));
    w.z1.ensure((size_t)nx * sizeof(float2));
    w.z2.ensure((size_t)nx * sizeof(float2));
    w.z3.ensure((size_t)nu * sizeof(float2));
    w.z4.ensure((size_t)nu * sizeof(float2));
    w.z6.ensure((size_t)nx * sizeof(float2));
    w.sc.ensure(sizeof(CgScalars));
    w.partial.ensure(kRB * sizeof(double));
    w.hist.ensure((size_t)std::max(nIter, 1) * sizeof(double));
    float* hy = w.hy.as<float>();
    float2 *res = w.k1.as<float2>(), *Ap = w.k2.as<float2>();
    float2 *tx = w.z1.as<float2>(), *img = w.z2.as<float2>(), *tmp = w.z6.as<float2>();
    float2 *dag = w.z3.as<float2>(), *p = w.z4.as<float2>();
    CgScalars* sc = w.sc.as<CgScalars>();
    double* part = w.partial.as<double>();
    double* hist = w.hist.as<double>();
    const float invHX = (float)(1.0 / P.HX);
    for (int f = 0; f < nFr; f++) {
        compute_hy(ve + ve_off[f], ve_max, w.hy0.as<float>() + koff[f] / P.nCh, Pf[f], hy_scratch(w), s);
        P.weights_to_tile_order(f, w.hy0.as<float>() + koff[f] / P.nCh, hy + koff[f] / P.nCh, s);
        P.to_tile_order(f, y + y_off[f], w.ys.as<float2>() + koff[f], s);
    }
    const float2* ys = w.ys.as<float2>();
    if (P.use_batch && (!P.Gu_b || (int)P.koff_b.size() != nFr + 1)) P.build_batch(s);
    auto T_of = [&](int f, bool transposed) -> const SparseDev* {
        if (f >= (int)P.deform[0].size()) return nullptr;
        return transposed ? P.deform[0][f].second : P.deform[0][f].first;
    };
    // A v: Tu_f v for every frame, then the forward operator of each frame
    auto A = [&](const float2* v, float2* out) {
        for (int f = 0; f < nFr; f++) {
            const SparseDev* T = T_of(f, false);
            if (T) spmv_c2_rowmajor(*T, v, tx + (int64_t)f * nu, 1, nullptr, s);
            else vcopy(v, tx + (int64_t)f * nu, nu, s);
        }
        if (P.use_batch) P.forward_batch(tx, out, s);
        else
            for (int f = 0; f < nFr; f++) P.forward(f, tx + (int64_t)f * nu, out + koff[f], s);
    };
    // dag = sum_f (1/HX) Tut_f Nakatsha(HY r_f)
    auto Ah = [&](const float2* r, float2* out) {
        if (P.use_batch) P.adjoint_batch(r, img, hy, invHX, s);
        else
            for (int f = 0; f < nFr; f++)
                P.adjoint(f, r + koff[f], img + (int64_t)f * nu, true, hy + koff[f] / P.nCh, invHX, s);
        const float2* src = img;
        bool any = false;
        for (int f = 0; f < nFr; f++) any = any || T_of(f, true);
        if (any) {
            for (int f = 0; f < nFr; f++) {
                const SparseDev* T = T_of(f, true);
                if (T) spmv_c2_rowmajor(*T, img + (int64_t)f * nu, tmp + (int64_t)f * nu, 1, nullptr, s);
                else vcopy(img + (int64_t)f * nu, tmp + (int64_t)f * nu, nu, s);
            }
            src = tmp;
        }
        k_frame_sum<<<nb(nu), 256, 0, s>>>(src, out, nFr, nu);
        MONA_HIP(hipGetLastError());
    };
    for (int c = 0; c < nIter; c++) {
        cg_reset(sc, s);
        A(x, Ap);
        vsub(ys, Ap, res, ny, s);
        Ah(res, dag);
        sqnorm(dag, nu, nullptr, 1, P.HX, &sc->sqn, false, part, s);
        vcopy(dag, p, nu, s);
        for (int j = 1; j <= nCGD; j++) {
            A(p, Ap);
            sqnorm(Ap, ny, hy, P.nCh, 1.0, &sc->den, false, part, s);
            cg_alpha(sc, s);
            vaxpy(x, p, &sc->a, 1.0f, nu, s);
            if (j == nCGD) break;
            vaxpy(res, Ap, &sc->a, -1.0f, ny, s);
            Ah(res, dag);
            sqnorm(dag, nu, nullptr, 1, P.HX, &sc->sqn_new, false, part, s);
            cg_beta(sc, s);
            vxpby(p, dag, &sc->b, nu, s);
        }
        A(x, Ap);
        vsub(ys, Ap, res, ny, s);
        sqnorm(res, ny, hy, P.nCh, 1.0, hist + c, false, part, s);
        P.witness(c, x, nu, s);
    }
    if (hist_out && nIter > 0) {
        MONA_HIP(hipMemcpyAsync(hist_out, hist, (size_t)nIter * sizeof(double), hipMemcpyDeviceToHost, s));
        MONA_HIP(hipStreamSynchronize(s));
    }
}

// ================================================================ bmNasha / bmMathilda
// Image from the gridded data in P.grid: centred inverse FFT times
// prod(N_u) prod(dK_u) (bmIDF*), times K, crop to n_u, coil combine with C
// (bmCoilSense_pinv) or per-coil images when C is NULL (bmNasha.m:66-87,
// bmMathilda.m:94-123).
void image_from_grid(Plan& P, const float2* C, float2* x, hipStream_t s) {
    float2* gr = P.grid.as<float2>();
    {
        TimedRegion t(2, s);
        cfft_nd(gr, P.imDim, P.N, P.nCh, true, s);
    }
    float pN = 1.0f, pdK = 1.0f;
    for (int a = 0; a < P.imDim; a++) { pN *= (float)P.N[a]; pdK *= (float)P.dK[a]; }
    float F = pN * pdK;  // bmIDF3.m:18 F = single(prod(N_u)*prod(dK_u))
    Geom g = P.geom();
    k_nasha_epilogue<<<nb(P.nu), 256, 0, s>>>(gr, P.K.as<float>(), C, x, g, F);
    MONA_HIP(hipGetLastError());
}

void nasha(Plan& P, const SparseDev& Gn, const float2* y, const float2* C, float2* x, hipStream_t s) {
    MONA_CHECK(P.coils_set, MONA_ERR_STATE, "coil maps not set (the deapodisation is computed there)");
    MONA_CHECK(Gn.l_size == P.Nu, MONA_ERR_ARG, "Gn has l_size != prod(N_u)");
    P.ensure_grid();
    {
        TimedRegion t(1, s);
        spmv_c2_rowmajor(Gn, y, P.grid.as<float2>(), P.nCh, nullptr, s);
    }
    image_from_grid(P, C, x, s);
}

// bmMathilda (bmMathilda.m:14-125) after its gridder: grid [Nu][nCh]
// (device, the normalised n2u of bmGridder_n2u[_leight]) to the image.
void mathilda(Plan& P, const float2* grid, const float2* C, float2* x, hipStream_t s) {
    MONA_CHECK(P.coils_set, MONA_ERR_STATE, "coil maps not set (the deapodisation is computed there)");
    P.ensure_grid();
    MONA_HIP(hipMemcpyAsync(P.grid.p, grid, (size_t)P.Nu * P.nCh * sizeof(float2), hipMemcpyDeviceToDevice, s));
    image_from_grid(P, C, x, s);
}

}  // namespace mona

using namespace mona;

extern "C" {

int mona_steva(mona_plan P, int frame, void* x, void* z, void* u, int zu_given, const void* y, const float* ve,
               float ve_max, const float* delta, const float* rho, int nIter, int nCGD, double* hist,
               hipStream_t stream) {
    return guarded([&] {
        MONA_CHECK(P && x && z && u && y && ve && delta && rho, MONA_ERR_ARG, "NULL argument");
        steva(P->P, frame, static_cast<float2*>(x), static_cast<float2*>(z), static_cast<float2*>(u),
              zu_given != 0, static_cast<const float2*>(y), ve, ve_max, delta, rho, nIter, nCGD, hist, stream);
    });
}

int mona_teva_chain(mona_plan P, void* x, void* z, void* u, int zu_given, const void* y, const long long* y_off,
                    const float* ve, const long long* ve_off, float ve_max, const float* delta, const float* rho,
                    int nIter, int nCGD, double* hist, hipStream_t stream) {
    return guarded([&] {
        MONA_CHECK(P && x && z && u && y && y_off && ve && ve_off && delta && rho, MONA_ERR_ARG, "NULL argument");
        morphosia(P->P, MONA_MORPH_TEVA_CHAIN, (int)P->P.frames.size(), static_cast<float2*>(x),
                  static_cast<float2*>(z), static_cast<float2*>(u), zu_given != 0, static_cast<const float2*>(y),
                  y_off, ve, ve_off, ve_max, delta, rho, nIter, nCGD, hist, stream);
    });
}

int mona_morphosia(mona_plan P, int solver, int nFr1, void* x, void* z, void* u, int zu_given, const void* y,
                   const long long* y_off, const float* ve, const long long* ve_off, float ve_max,
                   const float* delta, const float* rho, int nIter, int nCGD, double* hist, hipStream_t stream) {
    return guarded([&] {
        MONA_CHECK(P && x && y && y_off && ve && ve_off && delta, MONA_ERR_ARG, "NULL argument");
        morphosia(P->P, solver, nFr1, static_cast<float2*>(x), static_cast<float2*>(z), static_cast<float2*>(u),
                  zu_given != 0, static_cast<const float2*>(y), y_off, ve, ve_off, ve_max, delta, rho, nIter, nCGD,
                  hist, stream);
    });
}

int mona_sensa_morphosia(mona_plan P, void* x, const void* y, const long long* y_off, const float* ve,
                         const long long* ve_off, float ve_max, int nIter, int nCGD, double* hist,
                         hipStream_t stream) {
    return guarded([&] {
        MONA_CHECK(P && x && y && y_off && ve && ve_off, MONA_ERR_ARG, "NULL argument");
        sensa_morphosia(P->P, static_cast<float2*>(x), static_cast<const float2*>(y), y_off, ve, ve_off, ve_max,
                        nIter, nCGD, hist, stream);
    });
}

int mona_plan_set_deform(mona_plan P, int set, int frame, mona_sparse Tu, mona_sparse Tut) {
    return guarded([&] {
        MONA_CHECK(P && frame >= 0 && (set == 0 || set == 1), MONA_ERR_ARG, "bad argument");
        MONA_CHECK((Tu == nullptr) == (Tut == nullptr), MONA_ERR_ARG, "Tu and Tut are given together");
        Plan& pl = P->P;
        if (Tu) {
            MONA_CHECK(Tu->S.r_size == pl.nu && Tu->S.l_size == pl.nu, MONA_ERR_ARG,
                       "Tu must be prod(n_u) x prod(n_u)");
            MONA_CHECK(Tut->S.r_size == pl.nu && Tut->S.l_size == pl.nu, MONA_ERR_ARG,
                       "Tut must be prod(n_u) x prod(n_u)");
            HostCallCtx& hc = host_ctx();
            {
                std::lock_guard<std::mutex> lk(hc.mu);
                hc.init();
            }
            sparse_upload(Tu->S, hc.stream);
            sparse_upload(Tut->S, hc.stream);
        }
        auto& d = pl.deform[set];
        if ((int)d.size() <= frame) d.resize(frame + 1, {nullptr, nullptr});
        d[frame] = {Tu ? &Tu->S : nullptr, Tut ? &Tut->S : nullptr};
    });
}

int mona_mathilda(mona_plan P, const void* grid, const void* C, void* x, hipStream_t stream) {
    return guarded([&] {
        MONA_CHECK(P && grid && x, MONA_ERR_ARG, "NULL argument");
        mathilda(P->P, static_cast<const float2*>(grid), static_cast<const float2*>(C), static_cast<float2*>(x),
                 stream);
    });
}

int mona_nasha(mona_plan P, mona_sparse Gn, const void* y, const void* C, void* x, hipStream_t stream) {
    return guarded([&] {
        MONA_CHECK(P && Gn && y && x, MONA_ERR_ARG, "NULL argument");
        HostCallCtx& hc = host_ctx();
        {
            std::lock_guard<std::mutex> lk(hc.mu);
            hc.init();
        }
        sparse_upload(Gn->S, hc.stream);
        nasha(P->P, Gn->S, static_cast<const float2*>(y), static_cast<const float2*>(C), static_cast<float2*>(x),
              stream);
    });
}

}  // extern "C"
