// ORACLE -- test infrastructure only.  Gives the reference's C++-linkage
// library symbol (src/sparseMat/cpp/bmSparseMat_cC_oBlock_omp.h:9-14) a C name
// so that the tests can call the reference build through ctypes.
#include "bmSparseMat_cC_oBlock_omp.h"
extern "C" void ref_cC_oBlock_omp(int r_size, int *r_jump, int *r_nJump, float *m_val,
                                  int l_size, int *l_jump, int l_nJump, float *v_re,
                                  float *v_im, int n_vec, float *w_re, float *w_im)
{
    bmSparseMat_cC_oBlock_omp(r_size, r_jump, r_nJump, m_val, l_size, l_jump, l_nJump,
                              v_re, v_im, n_vec, w_re, w_im);
}
