// Centred Stockham FFT passes (fft.hip).
#pragma once
#include "mona_internal.hpp"

namespace mona {

// True when N = 2^a 3^b 5^c fits one LDS tile.
bool fft_size_supported(int N);

// Transforms every line of length N of a [outer][N][S] float2 array (S
// contiguous) along the middle axis; in == out is allowed.  presign
// multiplies the input by (-1)^n, postsign the output by (-1)^(k+N/2),
// scale multiplies the output.
void fft_axis(const float2* in, float2* out, int64_t outer, int N, int64_t S, bool inverse,
              bool presign, bool postsign, float scale, hipStream_t s);

// Centred n-D DFT of a [prod(N)][nBatch] array (x fastest after the batch
// index), in place; inverse scales by 1/prod(N).
void cfft_nd(float2* data, int imDim, const int* N, int64_t nBatch, bool inverse, hipStream_t s);

}  // namespace mona
