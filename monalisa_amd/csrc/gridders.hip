// Direct n2u gridders (leight / precomputed weights) and the in-box pair
// count: filled in after the sparse-product path.
#include "mona_internal.hpp"

using namespace mona;

extern "C" {

void bmGridder_n2u_leight1_mex(float*, float*, float*, float*, float*, float*, int, int, int, int, float,
                               float) {
    set_error("bmGridder_n2u_leight1_mex: not built yet");
    fatal("bmGridder_n2u_leight1_mex", MONA_ERR_STATE);
}
void bmGridder_n2u_leight2_mex(float*, float*, float*, float*, float*, float*, float*, int, int, int, int,
                               int, float, float) {
    set_error("bmGridder_n2u_leight2_mex: not built yet");
    fatal("bmGridder_n2u_leight2_mex", MONA_ERR_STATE);
}
void bmGridder_n2u_leight3_mex(float*, float*, float*, float*, float*, float*, float*, float*, int, int,
                               int, int, int, int, float, float) {
    set_error("bmGridder_n2u_leight3_mex: not built yet");
    fatal("bmGridder_n2u_leight3_mex", MONA_ERR_STATE);
}
void bmGridder_n2u_mex(float*, float*, float*, float*, float*, int*, int, int, int, int) {
    set_error("bmGridder_n2u_mex: not built yet");
    fatal("bmGridder_n2u_mex", MONA_ERR_STATE);
}
void bmGriddingMatrix_secret_length1_mex(float*, int, int, int, int, long long*) {
    set_error("not built yet");
    fatal("bmGriddingMatrix_secret_length1_mex", MONA_ERR_STATE);
}
void bmGriddingMatrix_secret_length2_mex(float*, float*, int, int, int, int, int, long long*) {
    set_error("not built yet");
    fatal("bmGriddingMatrix_secret_length2_mex", MONA_ERR_STATE);
}
void bmGriddingMatrix_secret_length3_mex(float*, float*, float*, int, int, int, int, int, int, long long*) {
    set_error("not built yet");
    fatal("bmGriddingMatrix_secret_length3_mex", MONA_ERR_STATE);
}

}  // extern "C"
