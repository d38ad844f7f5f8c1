// Matrix-free gridding (gridmf.hip): forward gridding from per-point data
// of the trajectory instead of the stored entries of Gu.
#pragma once
#include <algorithm>
#include <cstdlib>
#include <type_traits>

#include "mona_internal.hpp"

namespace mona {

constexpr int kMfTerms = 24;    // series terms of the Kaiser weight
constexpr int kMfBlock = 256;   // threads per block
constexpr int kMfMaxPts = 128;  // points per block at most

// Cell index of grid coordinates (x, y, z) in storage order: x + sy*y + sz*z.
struct MfGeo {
    int N0 = 1, N1 = 1, N2 = 1;
    int sy = 0, sz = 0;
};

// Kernel of the weights: Kaiser (inv = 1/tau^2, c = series coefficients of
// I0(alpha sqrt(v))) or Gauss (inv = log2(e) / (2 sigma^2)).
struct MfKernel {
    int kind = 0;
    float inv = 0.0f;
    float c[kMfTerms];
};

// Per-point data in the frame's point order: base cell x | y << 10 |
// z << 20 (0-based) and (offset x, offset y, offset z, 1 / sum of weights);
// a deleted point has all four 0.
struct MfPoints {
    int P = 0;
    DevBuf base;  // uint32[P]
    DevBuf rec;   // float4[P]
};

// Fills the kernel parameters from the plan's kernel; false when the series
// cannot reach float accuracy in kMfTerms terms.
bool mf_kernel_params(int ktype, int nParam, const double* kp, MfKernel* out);

// y[p][nCh] = sum over p's neighbour cells of w * grid[cell][nCh] for the
// P points; frame_of (may be NULL) gives each point's frame, whose grid
// starts frame_cells cells after the previous one.
void u2n_mf(const MfPoints& pts, const uint16_t* frame_of, int P, const void* grid, bool grid_bf16,
            int64_t frame_cells, void* y, bool y_bf16, int nCh, int imDim, const MfGeo& g, const MfKernel& k,
            hipStream_t s);

}  // namespace mona
