// Centred Stockham FFT passes (fft.hip).
#pragma once
#include "mona_internal.hpp"

namespace mona {

// True when N = 2^a 3^b 5^c <= 2048.
bool fft_size_supported(int N);
// True when N = 16 * N2 with N2 in {2,3,4,5,6,8,10,12,16}: passes of this
// length run register-resident (and support bf16x2 storage).
bool fft_reg_supported(int N);
// True when a plain pass of length N with this lines setting (AxisPass::lines)
// runs on the register kernel (so it can take cell masks).
bool fft_plain_pass_is_reg(int N, int lines);

// One pass along an axis of length N over an [outer][rows][S] array (S
// contiguous).  Input rows in_off .. in_off+in_cnt-1 of the length-N line
// are read from in + o*in_os + r*in_rs (r = row - in_off), the others are
// zero; output rows out_off .. out_off+out_cnt-1 are written to
// out + o*out_os + (row - out_off)*out_rs (row strides default to S).  presign multiplies input row n by
// (-1)^n, postsign output row k by (-1)^(k+N/2), scale multiplies the
// output.  mode 1: the input is img[o*in_cnt + r] * kfc[(o*in_cnt + r)*S + w];
// mode 2: out is replaced by img_out[o*out_cnt + k] = epi_scale *
// sum_w kfc[(o*out_cnt + k)*S + w] * y(k, w); mode 3: out[(o*out_cnt +
// k)*S + w] = epi_scale * kfc[..] * y(k, w).
struct AxisPass {
    int N = 0;
    int64_t S = 1, outer = 1;
    const float2* in = nullptr;
    int64_t in_os = 0;
    int in_off = 0, in_cnt = 0;
    float2* out = nullptr;
    int64_t out_os = 0;
    int out_off = 0, out_cnt = 0;
    bool inverse = false, presign = true, postsign = true;
    float scale = 1.0f, epi_scale = 1.0f;
    int mode = 0;
    const float2* img = nullptr;
    const float2* kfc = nullptr;
    int64_t kfc_period = 0;  // > 0: kfc pixel index taken modulo this (batched frames)
    int tile = 0;            // plain-pass tile (2048 / 4096 / 8192; 0 = default)
    int lines = 0;           // register kernel (N = 16 * N2): lines per tile on strided axes
                             // (16 / 32 / 64), 0 = default, -1 = LDS Stockham kernel only
    float2* img_out = nullptr;
    bool bf_in = false, bf_out = false;  // in / out stored as bf16x2 (register kernel only)
    // Cell masks (register kernel, mode 0), one bit per (line block, row):
    // element (o, row, w) belongs to column c = o * (S / mask_nch) + w /
    // mask_nch, whose bits are the ceil(N / 32) words at c * ceil(N / 32)
    // (bit row).  An input element whose in_mask bit is 0 is read as 0
    // without a load; an output element whose out_mask bit is 0 is not
    // stored.
    const uint32_t* in_mask = nullptr;
    const uint32_t* out_mask = nullptr;
    int mask_nch = 1;
    // Row intervals (register kernel, mode 0), the contiguous form of the
    // masks: column c (as above) holds rows [x, y) of in_ival[c] /
    // out_ival[c]; rows outside are read as 0 without a load / not stored,
    // and a column (-1, -1) is skipped whole (dead).  Given with masks, an
    // element is read / stored only when both allow it.
    const int2* in_ival = nullptr;
    const int2* out_ival = nullptr;
    // Row strides (elements between rows n and n+1 of a line) of in / out;
    // 0 = S.  Mode 0 only: a pass can then read or write lines whose rows
    // are interleaved with other lines' rows ([row][outer][S] instead of
    // [outer][row][S]).
    int64_t in_rs = 0, out_rs = 0;
    // Two-level line offsets (mode 0): with in_on > 0 line o starts at
    // (o / in_on) * in_os2 + (o % in_on) * in_os (frames of a batch whose
    // own layout interleaves rows); likewise out_on / out_os2.
    int64_t in_on = 0, in_os2 = 0, out_on = 0, out_os2 = 0;
};

void fft_pass(const AxisPass& p, hipStream_t s);
// True when a line of N rows times S inner values fits one tile (required
// for the fused prologue / epilogue modes).
bool fft_pass_fusable(int N, int64_t S);

// Transforms every line of length N of a [outer][N][S] float2 array along
// the middle axis; in == out is allowed.
void fft_axis(const float2* in, float2* out, int64_t outer, int N, int64_t S, bool inverse,
              bool presign, bool postsign, float scale, hipStream_t s);

// Centred n-D DFT of a [prod(N)][nBatch] array (x fastest after the batch
// index), in place; inverse scales by 1/prod(N).
void cfft_nd(float2* data, int imDim, const int* N, int64_t nBatch, bool inverse, hipStream_t s);

}  // namespace mona
