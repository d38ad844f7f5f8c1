// ADMM total-variation solvers and the gridded reconstruction (filled in
// below the operator layer).
#include "recon.hpp"

using namespace mona;

extern "C" {

int mona_steva(mona_plan, int, void*, void*, void*, int, const void*, const float*, float, const float*,
               const float*, int, int, double*, hipStream_t) {
    return guarded([&] { MONA_CHECK(false, MONA_ERR_STATE, "mona_steva: not built yet"); });
}

int mona_teva_chain(mona_plan, void*, void*, void*, int, const void*, const long long*, const float*,
                    const long long*, float, const float*, const float*, int, int, double*, hipStream_t) {
    return guarded([&] { MONA_CHECK(false, MONA_ERR_STATE, "mona_teva_chain: not built yet"); });
}

int mona_nasha(mona_plan, mona_sparse, const void*, const void*, void*, hipStream_t) {
    return guarded([&] { MONA_CHECK(false, MONA_ERR_STATE, "mona_nasha: not built yet"); });
}

}  // extern "C"
