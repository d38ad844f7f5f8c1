/*
 * libmona -- MI355X (gfx950) implementation of Monalisa's non-Cartesian
 * reconstruction hot path behind a C ABI.
 *
 * Two groups of entry points:
 *
 *  1. Drop-in replacements for the reference's native (C++ MEX) inner
 *     functions.  Same names, same argument meaning and layout, host
 *     pointers, synchronous, outputs zero-filled by the callee, no status
 *     (the reference returns void); on a failure they print the reason to
 *     stderr and abort, and mona_last_error() holds the message.
 *
 *  2. A device-resident handle API (mona_*) for the operators and solvers:
 *     device pointers, an explicit hipStream_t (0 = null stream), an int
 *     status (MONA_OK or an error code; no exceptions cross the ABI).
 *
 * Complex data on the device API is interleaved float2 (re, im).  Layouts:
 *   image  x   [prod(n_u)]            x fastest (the reference's column order)
 *   k-space y  [nPt][nCh]             coil index fastest
 *   grid       [prod(N_u)][nCh]       coil index fastest
 *   coil maps  [prod(n_u)][nCh]       coil index fastest
 */
#ifndef MONA_H
#define MONA_H

#include <hip/hip_runtime.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MONA_OK 0
#define MONA_ERR_ARG 1       /* invalid argument / failed structural check */
#define MONA_ERR_HIP 2       /* HIP runtime error */
#define MONA_ERR_OOM 3       /* device or host allocation failed */
#define MONA_ERR_STATE 4     /* object not in the required state */
#define MONA_ERR_INTERNAL 5

#define MONA_KERNEL_GAUSS 0
#define MONA_KERNEL_KAISER 1

/* Message of the last failure in this process ("" if none). */
const char* mona_last_error(void);
/* Number of visible HIP devices (0 when none). */
int mona_device_count(void);
/* Library version string. */
const char* mona_version(void);

/* ======================================================================
 * 1. Drop-in entry points (host pointers, synchronous, void)
 * ====================================================================== */

/* W = S*V, complex, column layout: v_re/v_im are n_vec planes of r_size
 * floats, w_re/w_im are n_vec planes of l_size floats.  S is a cpp_prepared
 * one-block bmSparseMat: r_jump running column deltas (first absolute,
 * 0-based), r_nJump entries per stored row, l_jump running output-row
 * deltas or NULL when rows are not squeezed.
 * Replaces src/sparseMat/cpp/bmSparseMat_cC_oBlock_omp.h:9-14 (library
 * symbol) and the inner function of
 * src/sparseMat/mex/bmSparseMat_vec/bmSparseMat_cC_oBlock_omp/bmSparseMat_cC_oBlock_omp_mex.cpp:106-117.
 * The reference's header declares the library symbol with C++ linkage; the
 * library also defines that mangled twin
 * (_Z25bmSparseMat_cC_oBlock_ompiPiS_PfiS_iS0_S0_iS0_S0_, cxx_linkage.cpp),
 * so C++ code built against that header links unchanged. */
void bmSparseMat_cC_oBlock_omp(int r_size, int* r_jump, int* r_nJump, float* m_val,
                               int l_size, int* l_jump, int l_nJump,
                               float* v_real, float* v_imag, int n_vec,
                               float* w_real, float* w_imag);

/* Serial variant, same semantics.
 * Replaces .../bmSparseMat_cC_oBlock/bmSparseMat_cC_oBlock_mex.cpp:82-193. */
void bmSparseMat_cC_oBlock(int r_size, int* r_jump, int* r_nJump, float* m_val,
                           int l_size, int* l_jump, int l_nJump,
                           float* v_real, float* v_imag, int n_vec,
                           float* w_real, float* w_imag);

/* Variant that requires exactly 2*n_vec threads in the reference (planes are
 * skipped when fewer are granted); here every plane is always computed.
 * Replaces .../bmSparseMat_cC_oBlock_2_omp/bmSparseMat_cC_oBlock_2_omp_mex.cpp:83-201. */
void bmSparseMat_cC_oBlock_2_omp(int r_size, int* r_jump, int* r_nJump, float* m_val,
                                 int l_size, int* l_jump, int l_nJump,
                                 float* v_real, float* v_imag, int n_vec,
                                 float* w_real, float* w_imag);

/* Complex row layout: v is [r_size][n_vec] (vector index fastest), w is
 * [l_size][n_vec].
 * Replaces .../bmSparseMat_cR_oBlock/bmSparseMat_cR_oBlock_mex.cpp:80-223. */
void bmSparseMat_cR_oBlock(int r_size, int* r_jump, int* r_nJump, float* m_val,
                           int l_size, int* l_jump, int l_nJump,
                           float* v_real, float* v_imag, int n_vec,
                           float* w_real, float* w_imag);

/* Real column layout.
 * Replaces .../bmSparseMat_rC_oBlock/bmSparseMat_rC_oBlock_mex.cpp:77-174
 * and .../bmSparseMat_rC_oBlock_omp/bmSparseMat_rC_oBlock_omp_mex.cpp:79-152. */
void bmSparseMat_rC_oBlock(int r_size, int* r_jump, int* r_nJump, float* m_val,
                           int l_size, int* l_jump, int l_nJump,
                           float* v, int n_vec, float* w);

/* Real row layout.
 * Replaces .../bmSparseMat_rR_oBlock/bmSparseMat_rR_oBlock_mex.cpp:77-194. */
void bmSparseMat_rR_oBlock(int r_size, int* r_jump, int* r_nJump, float* m_val,
                           int l_size, int* l_jump, int l_nJump,
                           float* v, int n_vec, float* w);

/* Multi-block row-layout products: the block partition only changes how
 * the reference splits the work over threads; entries are still a single
 * running delta sequence, except that the first entry of every block (at
 * m_block_start[b]) is absolute 0-based, and, for a squeezed matrix
 * (l_jump != NULL, the reference's l_sparsity branch), so is the first
 * output row of every block (l_jump[l_block_start[b]]).  Rows not stored
 * are zero in w.
 * Replace .../bmSparseMat_cR_nBlock_omp/bmSparseMat_cR_nBlock_omp_mex.cpp:97-483
 * and .../bmSparseMat_rR_nBlock_omp/bmSparseMat_rR_nBlock_omp_mex.cpp:97-436. */
void bmSparseMat_cR_nBlock_omp(int r_size, int* r_jump, int* r_nJump, float* m_val,
                               int l_size, int* l_jump, int l_nJump,
                               int nBlock, int* block_length, int* l_block_start,
                               long long* m_block_start,
                               float* v_real, float* v_imag, int n_vec,
                               float* w_real, float* w_imag);
void bmSparseMat_rR_nBlock_omp(int r_size, int* r_jump, int* r_nJump, float* m_val,
                               int l_size, int* l_jump, int l_nJump,
                               int nBlock, int* block_length, int* l_block_start,
                               long long* m_block_start,
                               float* v, int n_vec, float* w);

/* Row-id expansion: ind[sum(r_nJump)] = 0-based row of each entry.
 * Replaces .../bmSparseMat_r_nJump2index/bmSparseMat_r_nJump2index_mex.cpp:70-89. */
void bmSparseMat_r_nJump2index_mex(int l_size, int* r_nJump, int* ind);

/* Normalised n2u gridding, Gaussian weights computed on the fly, clipped
 * box.  a_* are [nPt][nCh], b_* are [Nu][nCh] (coil fastest); t* are
 * 1-based grid coordinates.
 * Replace src/gridding123/mex/bmGridder_n2u_leight{1,2,3}/bmGridder_n2u_leight{1,2,3}_mex.cpp
 * (leight3: :116-358). */
void bmGridder_n2u_leight1_mex(float* aReal, float* aImag, float* bReal, float* bImag,
                               float* tx, float* Dn, int nCh, int nPt, int Nx,
                               int nWin, float kernelParam_1, float kernelParam_2);
void bmGridder_n2u_leight2_mex(float* aReal, float* aImag, float* bReal, float* bImag,
                               float* tx, float* ty, float* Dn, int nCh, int nPt,
                               int Nx, int Ny, int nWin, float kernelParam_1,
                               float kernelParam_2);
void bmGridder_n2u_leight3_mex(float* aReal, float* aImag, float* bReal, float* bImag,
                               float* tx, float* ty, float* tz, float* Dn, int nCh,
                               int nPt, int Nx, int Ny, int Nz, int nWin,
                               float kernelParam_1, float kernelParam_2);

/* Normalised n2u with precomputed weights w[nPt][nNb] and running index
 * jumps nJump[nPt*nNb] (first absolute, 0-based).
 * Replaces src/gridding123/mex/bmGridder_n2u/bmGridder_n2u_mex.cpp:72-201. */
void bmGridder_n2u_mex(float* aReal, float* aImag, float* bReal, float* bImag, float* w,
                       int* nJump, int nNb, int nCh, int nPt, int Nu_tot);

/* Number of in-box (point, neighbour) pairs with d^2 < (nWin/2)^2.
 * Replace src/gridding123/mex/bmGriddingMatrix_secret_length{1,2,3}/..._mex.cpp
 * (secret_length3: :70-222). */
void bmGriddingMatrix_secret_length1_mex(float* tx, int nCh, int nPt, int Nx, int nWin,
                                         long long* secret_length);
void bmGriddingMatrix_secret_length2_mex(float* tx, float* ty, int nCh, int nPt, int Nx,
                                         int Ny, int nWin, long long* secret_length);
void bmGriddingMatrix_secret_length3_mex(float* tx, float* ty, float* tz, int nCh, int nPt,
                                         int Nx, int Ny, int Nz, int nWin,
                                         long long* secret_length);

/* ======================================================================
 * 2. Device-resident handle API
 * ====================================================================== */

typedef struct mona_sparse_s* mona_sparse;
typedef struct mona_plan_s* mona_plan;

/* --- bmSparseMat on the device ---------------------------------------- */

/* Uploads a cpp_prepared one-block matrix given in the reference encoding
 * (host arrays) and decodes it on the device.  Applies the rules of
 * bmSparseMat.check (src/sparseMat/m/bmSparseMat.m:192-393). */
int mona_sparse_create(int r_size, const int* r_jump, const int* r_nJump,
                       const float* m_val, int l_size, const int* l_jump, int l_nJump,
                       long long nnz, mona_sparse* out);
int mona_sparse_destroy(mona_sparse S);
/* nnz, l_nJump, r_size, l_size, squeezed flag (any pointer may be NULL). */
int mona_sparse_info(mona_sparse S, long long* nnz, int* l_nJump, int* r_size,
                     int* l_size, int* squeezed);
/* Copies the reference encoding back to host arrays (l_jump ignored when
 * not squeezed; any pointer may be NULL). */
int mona_sparse_export(mona_sparse S, int* r_jump, int* r_nJump, float* m_val,
                       int* l_jump);
/* w = S v on the device, complex interleaved, vector index fastest:
 * v [r_size][n_vec] float2, w [l_size][n_vec] float2. */
int mona_sparse_apply(mona_sparse S, const void* v, void* w, int n_vec,
                      hipStream_t stream);
/* mona_sparse_apply without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops).  The partial sums of rows
 * longer than 64 entries go to a file-scope device array of 64 MiB per GPU
 * (held until the process ends) instead of a per-matrix buffer, so
 * n_chunk * n_vec * 8 bytes must fit in 64 MiB (n_chunk = sum over rows of
 * more than 64 entries of ceil(entries / 64)); larger calls return
 * MONA_ERR_ARG.  S must already be on the device (mona_sparse_apply or
 * mona_plan_set_frame uploads it), else MONA_ERR_STATE.  Results equal
 * mona_sparse_apply's bit for bit.  All calls on one GPU share the array:
 * calls that overlap on one GPU (other streams or threads) race. */
int mona_sparse_apply_noalloc(mona_sparse S, const void* v, void* w, int n_vec,
                              hipStream_t stream);

/* --- trajectory -> gridding matrices ----------------------------------
 * bmTraj2SparseMat (src/gridding123/m/bmTraj2SparseMat.m:12-314):
 * t [nPt][imDim] (physical k, point-major), ve [nPt] (may be NULL when
 * only Gu/Gut are requested), N_u[imDim], dK_u[imDim].  kernelParam has
 * 2 (gauss: sigma, K_max) or 3 (kaiser: tau, alpha, K_max) entries.
 * which: bit 0 = Gn, bit 1 = Gu and Gut.  Outputs not requested may be NULL. */
int mona_traj2sparsemat(int imDim, int nPt, const double* t, const double* ve,
                        const int* N_u, const double* dK_u, int kernelType, int nWin,
                        const double* kernelParam, int nParam, int which,
                        mona_sparse* Gn, mona_sparse* Gu, mona_sparse* Gut);
/* The same construction on the GPU (same host arguments; the work runs on
 * `stream` and the call returns when the matrices are ready).  The integer
 * arrays equal mona_traj2sparsemat's exactly; m_val may differ in the last
 * float bit (device exp / sqrt). */
int mona_traj2sparsemat_device(int imDim, int nPt, const double* t, const double* ve,
                               const int* N_u, const double* dK_u, int kernelType, int nWin,
                               const double* kernelParam, int nParam, int which,
                               mona_sparse* Gn, mona_sparse* Gu, mona_sparse* Gut, hipStream_t stream);
/* Interpolation matrices of a deformation field
 * (bmImDeformField2SparseMat.m:14-147, bmImDeformField2sparseMat_ind_weight.m:6-150):
 * v [prod(N_u)][imDim] displaces pixel p (1-based coordinates) to t_p;
 * bump weights exp(-1/(1-d^2)) on the 2^imDim cells around t_p, wrapped
 * mod N_u; points leaving [1, N_u] are deleted unless torus != 0.  Gu rows
 * (pixels) are normalised, Gut = Gu', Gn = rownorm((w .* Dn)') (needs Dn
 * [prod(N_u)]).  No matrix is squeezed.  which: 1 = Gn, 2 = Gu+Gut, 3 = all. */
int mona_deform2sparsemat(int imDim, const int* N_u, const double* v, const double* Dn, int torus,
                          int which, mona_sparse* Gn, mona_sparse* Gu, mona_sparse* Gut);

/* --- reconstruction plan -----------------------------------------------
 * Geometry + gridding kernel + coil maps; owns KFC/KFC_conj (bmKF.m:6-35,
 * bmKF_conj.m:6-38), the oversampled-grid workspace and nFrames gridding
 * matrix pairs (Gu, Gut), which it references but does not own. */
int mona_plan_create(int imDim, const int* N_u, const int* n_u, const double* dK_u,
                     int nCh, int kernelType, int nWin, const double* kernelParam,
                     int nParam, mona_plan* out);
int mona_plan_destroy(mona_plan P);
/* Coil maps C [prod(n_u)][nCh] complex (device pointer).  Computes the
 * deapodisation K (bmK.m:6-156) and KFC, KFC_conj on the device. */
int mona_plan_set_coils(mona_plan P, const void* C, hipStream_t stream);
/* Frame f uses Gu (nPt x prod(N_u)) and Gut (its transpose). */
int mona_plan_set_frame(mona_plan P, int frame, mona_sparse Gu, mona_sparse Gut);
/* Deapodisation K cropped to n_u [prod(n_u)] float (device out). */
int mona_plan_get_K(mona_plan P, float* K, hipStream_t stream);
/* KFC and KFC_conj [prod(n_u)][nCh] float2 (device out, either may be NULL). */
int mona_plan_get_KFC(mona_plan P, void* KFC, void* KFC_conj, hipStream_t stream);

/* Forward model y = Gu * cfft(zerofill(x .* KFC))
 * (bmShanna_MATLAB.m:6-56).  x [prod(n_u)], y [nPt][nCh]. */
int mona_shanna(mona_plan P, int frame, const void* x, void* y, hipStream_t stream);
/* Adjoint x = sum_c KFC_conj .* crop(cifft(Gut * y))  (C_flag = 1) or the
 * per-coil images [prod(n_u)][nCh] (C_flag = 0)  (bmNakatsha_MATLAB.m:6-60). */
int mona_nakatsha(mona_plan P, int frame, const void* y, void* x, int C_flag,
                  hipStream_t stream);

/* Centred n-D DFT over the first imDim axes of a [prod(N)][nBatch] float2
 * array, in place: fftshift(fft(ifftshift(.))) per axis (inverse: ifft,
 * scaled 1/N per axis).  Follows bmShanna_MATLAB.m:46-50 /
 * bmNakatsha_MATLAB.m:36-40. */
int mona_cfft(int imDim, const int* N, int nBatch, void* data, int inverse,
              hipStream_t stream);

/* CG-SENSE (src/fourier123/solver_function/imDim/nonCartesian/bmSensa.m:6-96).
 * x [prod(n_u)] in/out (initial guess), y [nPt][nCh], ve [nPt] (volume
 * elements; HY = min(ve, ve_max); ve_max <= 0 means max(ve)).
 * residual (host, may be NULL) receives nIter values of <r, HY r>.
 * Vectors FP32, inner products and CG scalars FP64. */
int mona_sensa(mona_plan P, int frame, void* x, const void* y, const float* ve,
               float ve_max, int nIter, int nCGD, double* residual, hipStream_t stream);

/* Spatial-TV ADMM (src/fourier123/solver_function/imDim/nonCartesian/bmSteva.m:6-138).
 * z, u [imDim][prod(n_u)] float2 in/out (z_init/u_init flags: 0 = compute
 * z = grad(x), u = 0).  delta/rho: 2 values each (linear schedule start,
 * end over nIter).  hist (host, may be NULL): nIter x {residual, TV}. */
int mona_steva(mona_plan P, int frame, void* x, void* z, void* u, int zu_given,
               const void* y, const float* ve, float ve_max, const float* delta,
               const float* rho, int nIter, int nCGD, double* hist,
               hipStream_t stream);

/* Temporal-TV ADMM over the plan's nFr frames (bmTevaMorphosia_chain.m:6-294,
 * regul_mode 'normal'): F x_f = Tu_f x_f - x_{f-1} with the deformations
 * given by mona_plan_set_deform (identity for frames without one, Tu = []).
 * x, z, u: nFr images each, contiguous [nFr][prod(n_u)] float2.
 * y: frame f at y_off[f] float2 elements into y ([nPt_f][nCh]); ve likewise
 * at ve_off[f] (floats).  hist (host, may be NULL): nIter x {objective,
 * dafi, regul}. */
int mona_teva_chain(mona_plan P, void* x, void* z, void* u, int zu_given,
                    const void* y, const long long* y_off, const float* ve,
                    const long long* ve_off, float ve_max, const float* delta,
                    const float* rho, int nIter, int nCGD, double* hist,
                    hipStream_t stream);

/* Deformation matrices of frame f for the Morphosia solvers, applied as
 * bmImDeform.m / bmImDeformT.m do (K = []): set 0 holds the Tu{f}, Tut{f}
 * (or Tu1, Tu1t) arguments of bmTevaMorphosia_chain.m:6-13,
 * bmTevaDuoMorphosia_chain.m:6-16, bmTevaMorphosia_sheet.m:16-26 and
 * bmSensitivaMorphosia_chain.m / _sheet.m; set 1 holds Tu2, Tu2t.  Both
 * prod(n_u) x prod(n_u), referenced, not owned; Tu = Tut = NULL restores
 * the identity. */
int mona_plan_set_deform(mona_plan P, int set, int frame, mona_sparse Tu, mona_sparse Tut);

/* The Morphosia solvers over the plan's frames, regul_mode 'normal'.
 * Frames f = i + nFr1*j of an nFr1 x nFr2 cell array (y(:) order; chains
 * use nFr1 = nFr).  x [nFr][prod(n_u)] float2; z, u: one block of nFr
 * images per difference term (z1 then z2), NULL for the Sensitiva solvers;
 * rho is unused (may be NULL) for them.  Other arguments as mona_teva_chain.
 *   MONA_MORPH_TEVA_CHAIN      bmTevaMorphosia_chain.m:6-294
 *   MONA_MORPH_TEVA_DUO_CHAIN  bmTevaDuoMorphosia_chain.m:6-359 (0.5 (Tu1 x_i - x_{i-1}),
 *                              0.5 (Tu2 x_i - x_{i+1}))
 *   MONA_MORPH_TEVA_SHEET      bmTevaMorphosia_sheet.m:16-379 (differences along i and j)
 *   MONA_MORPH_SENSITIVA_CHAIN bmSensitivaMorphosia_chain.m:6-280 (delta HZ ||Tu x_i - x_{i-1}||^2)
 *   MONA_MORPH_SENSITIVA_SHEET bmSensitivaMorphosia_sheet.m:6-336 (both axes; the first
 *                              gradient of each iteration carries the penalty without
 *                              delta, as bmSensitivaMorphosia_sheet.m:118-119 does)
 *   MONA_MORPH_SENSITIVA_DUO_CHAIN bmSensitivaDuoMorphosia_chain.m:6-313 (the duo
 *                              differences under the quadratic penalty) */
#define MONA_MORPH_TEVA_CHAIN 0
#define MONA_MORPH_TEVA_DUO_CHAIN 1
#define MONA_MORPH_TEVA_SHEET 2
#define MONA_MORPH_SENSITIVA_CHAIN 3
#define MONA_MORPH_SENSITIVA_SHEET 4
#define MONA_MORPH_SENSITIVA_DUO_CHAIN 5
int mona_morphosia(mona_plan P, int solver, int nFr1, void* x, void* z, void* u, int zu_given,
                   const void* y, const long long* y_off, const float* ve,
                   const long long* ve_off, float ve_max, const float* delta,
                   const float* rho, int nIter, int nCGD, double* hist, hipStream_t stream);

/* Motion-compensated SENSE (bmSensaMorphosia_chain.m:6-104): one image x
 * [prod(n_u)] float2 observed by every frame through its deformation
 * (set 0; identity where none): minimises sum_f ||M_f Tu_f x - y_f||^2_HY
 * by nCGD CG steps per iteration, restarted nIter times.  y, ve as
 * mona_teva_chain; hist (host, may be NULL): nIter residuals R. */
int mona_sensa_morphosia(mona_plan P, void* x, const void* y, const long long* y_off,
                         const float* ve, const long long* ve_off, float ve_max, int nIter,
                         int nCGD, double* hist, hipStream_t stream);

/* Gridded reconstruction (bmNasha.m:6-91): n2u with Gn, centred ifft times
 * prod(N_u)*prod(dK_u), times K, crop to n_u; with C != NULL the coil
 * combine sum(conj(C) x)/sum|C|^2 (bmCoilSense_pinv.m:6-12), output
 * [prod(n_u)]; without C the per-coil images [prod(n_u)][nCh]. */
int mona_nasha(mona_plan P, mona_sparse Gn, const void* y, const void* C, void* x,
               hipStream_t stream);

/* Gridded reconstruction after a direct gridder (bmMathilda.m:14-125): grid
 * [prod(N_u)][nCh] float2 (device; the normalised n2u that
 * bmGridder_n2u[_leight] returns, cells x fastest, coil fastest within a
 * cell) -> centred inverse FFT times prod(N_u) prod(dK_u) (bmIDF*) -> .* K
 * -> crop to n_u -> coil combine with C (bmCoilSense_pinv.m:6-12), or
 * per-coil images [prod(n_u)][nCh] when C is NULL.  Needs the plan's coil
 * maps (the deapodisation K is computed there). */
int mona_mathilda(mona_plan P, const void* grid, const void* C, void* x, hipStream_t stream);

/* bmSensa over frames 0 .. nFrames-1 of the plan at once (nFrames <= 0: all;
 * no reference counterpart: the batched form of the serving loop).  Frame
 * f solves its own system
 * with data y + y_off[f] ([P_f][nCh], caller point order), weights
 * ve + ve_off[f] and ve_max[f] (<= 0: max of its ve; ve_max may be NULL);
 * x is [nFr][prod(n_u)].  The operators run as one block-diagonal gridding
 * launch and frame-batched FFT passes; each frame's result equals
 * mona_sensa on that frame alone.  residual: [nFr][nIter] (may be NULL). */
int mona_sensa_batch(mona_plan P, int nFrames, void* x, const void* y, const long long* y_off, const float* ve,
                     const long long* ve_off, const float* ve_max, int nIter, int nCGD, double* residual,
                     hipStream_t stream);
/* Device memory held by a plan: its buffers and frame matrices. */
int mona_plan_memory(mona_plan P, long long* bytes);

/* Mixed precision (no reference counterpart).  CG-SENSE by iterative
 * refinement with reliable updates: CG on the normal equations of bmSensa's
 * system (same HX / HY inner products) whose products use the operators
 * with the grid-side arrays and k-space vector stored as bf16 (bf16 = 1) or
 * float (bf16 = 0); every nInner steps the recursive residual is replaced
 * by the true one computed in float, g = (1/HX) M^H HY (y - M x), and CG
 * continues.  rel = ||g|| / ||b||, b = (1/HX) M^H HY y (the residual of
 * x = 0).  When the recursive residual falls below tol the true one is
 * computed at once (at most once per nInner steps); the solve stops when a
 * true rel < tol or after nOuter periodic replacements.  x in place (any
 * start); rel_res receives rel at the start and after every replacement
 * (room for 2 * nOuter + 1 doubles), *n_outer the replacements done and
 * *n_steps the CG steps (either may be NULL).
 * bf16 storage needs every N_u = 16 * {2,3,4,5,6,8,10,12,16} and
 * nCh <= 64. */
int mona_sensa_ir(mona_plan P, int frame, void* x, const void* y, const float* ve, float ve_max, int nInner,
                  int nOuter, double tol, int bf16, double* rel_res, int* n_outer, int* n_steps,
                  hipStream_t stream);
/* bmShanna / bmNakatsha (C_flag = 1) with bf16 storage of the grid-side
 * arrays and of the k-space vector; caller-order float k-space in / out. */
int mona_shanna_bf16(mona_plan P, int frame, const void* x, void* y, hipStream_t stream);
int mona_nakatsha_bf16(mona_plan P, int frame, const void* y, void* x, hipStream_t stream);

/* Plan options (no reference counterpart: execution choices of this
 * implementation). */
#define MONA_OPT_BATCH_FRAMES 1 /* temporal-TV solver: all frames in one launch per gridding op
                                   (block-diagonal matrices, frame-batched FFT passes) */
#define MONA_OPT_GRAPH 2        /* temporal-TV solver: ADMM iterations replayed from a HIP graph
                                   (constant delta/rho schedules) */
#define MONA_OPT_FFT_TILE 3     /* plain FFT pass tile: 2048, 4096 or 8192 complex values, 0 = default */
#define MONA_OPT_FFT_LINES 4    /* register-resident FFT passes (N = 16 * N2, N2 <= 16): lines per tile
                                   on strided axes 16, 32 or 64, 0 = default, -1 = LDS kernel only */
#define MONA_OPT_MATRIX_FREE 5  /* forward gridding computed from the trajectory that mona_traj2sparsemat[_device]
                                   keeps with Gu (frames whose Gu has one and whose geometry and kernel match the
                                   plan; nWin 3): 1 = on (default), 0 = the stored entries */
#define MONA_OPT_STAGED_ADJOINT 6 /* adjoint gridding: the k-space rows of groups of neighbouring grid cells
                                     staged in LDS once per group (1 = on, default) or gathered per entry (0);
                                     the results are the same bit for bit */
int mona_plan_set_option(mona_plan P, int option, long long value);
/* Witness snapshots of the solvers (replaces the per-iteration
 * witnessInfo.watch calls of bmSensa.m:85, bmSteva and the Morphosia solvers,
 * bmWitnessInfo.m:92-190).  While set, mona_sensa, mona_steva,
 * mona_teva_chain, mona_morphosia and mona_sensa_morphosia copy, after outer
 * iteration c (0-based), frame 0's witness image -- the whole image in 1-D
 * and 2-D, its middle z slice (z = n_z / 2, 0-based) in 3-D;
 * mona_plan_witness_elems complex values -- into slices + c * elems while
 * c < n_slots, and the whole x (all frames) into full + k * (x elements)
 * after iteration full_ind[k] (1-based, as the reference numbers them).
 * Device buffers owned by the caller; slices = full = NULL turns it off.
 * Runs with a witness set replay no HIP graph. */
int mona_plan_set_witness(mona_plan P, void* slices, int n_slots, void* full, const int* full_ind, int n_full);
int mona_plan_witness_elems(mona_plan P, long long* elems);
/* Autotuner: times forward + adjoint of one frame for each FFT pass variant
 * and keeps the fastest; *chosen (may be NULL) receives the register kernel's
 * lines per tile, or minus the LDS kernel's tile. */
int mona_plan_autotune(mona_plan P, int frame, hipStream_t stream, long long* chosen);
/* Autotuner of the staged adjoint gridding (no reference counterpart): times
 * the weighted adjoint gridding of one frame for several group sizes (LDS
 * budget of the staged k-space rows in KiB, rows per group), keeps the
 * fastest, regroups every frame with it and reports it (either pointer may
 * be NULL).  Results do not depend on the group sizes. */
int mona_plan_autotune_groups(mona_plan P, int frame, hipStream_t stream, int* stage_kb, int* rows);

/* Kernel timing: when enabled, every gridding launch (forward Gu product,
 * adjoint Gut product) records HIP events on its stream; the accumulated
 * launch count and total milliseconds per kind are read with
 * mona_timing_read (kind 0 = u2n, 1 = n2u, 2 = one whole transform, all its FFT
 * passes; kind 3 is reserved and reads zero). */
int mona_timing_enable(int on);
int mona_timing_read(int kind, long long* launches, double* total_ms);

#ifdef __cplusplus
}
#endif
#endif /* MONA_H */
