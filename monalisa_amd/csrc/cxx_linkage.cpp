// C++-linkage twin of the one library symbol the reference exports with C++
// linkage: src/sparseMat/cpp/bmSparseMat_cC_oBlock_omp.h:9-14 declares
//   extern void bmSparseMat_cC_oBlock_omp(int, int*, int*, float*, int, int*,
//                                         int, float*, float*, int, float*, float*);
// without extern "C", so C++ code compiled against that header references the
// mangled name.  This definition forwards to the C entry point of the same
// name (sparse.hip), reached through an assembler label because one
// translation unit cannot declare both linkages of one name; for that reason
// this file does not include mona.h.

extern "C" void mona_c_entry_cC_oBlock_omp(int r_size, int* r_jump, int* r_nJump, float* m_val, int l_size,
                                           int* l_jump, int l_nJump, float* v_real, float* v_imag, int n_vec,
                                           float* w_real, float* w_imag) __asm__("bmSparseMat_cC_oBlock_omp");

void bmSparseMat_cC_oBlock_omp(int r_size, int* r_jump, int* r_nJump, float* m_val, int l_size, int* l_jump,
                               int l_nJump, float* v_real, float* v_imag, int n_vec, float* w_real, float* w_imag) {
    mona_c_entry_cC_oBlock_omp(r_size, r_jump, r_nJump, m_val, l_size, l_jump, l_nJump, v_real, v_imag, n_vec,
                               w_real, w_imag);
}
