// Device prefix sums used to decode the jump encoding of bmSparseMat
// (running deltas -> absolute indices, row counts -> row pointers).
// Three launches: per-tile totals, one-workgroup scan of the totals, per-tile
// scan with the tile's carry.  Integer arithmetic, so the result does not
// depend on the order of the partial sums.
#include "mona_internal.hpp"

namespace mona {
namespace {

constexpr int kThreads = 256;
constexpr int kItems = 16;
constexpr int kTile = kThreads * kItems;

template <class T>
__device__ T block_exclusive_scan(T v, T* lds, T* total) {
    // Hillis-Steele over 256 values in LDS (integer, exact).
    int t = threadIdx.x;
    lds[t] = v;
    __syncthreads();
    for (int off = 1; off < kThreads; off <<= 1) {
        T add = (t >= off) ? lds[t - off] : T(0);
        __syncthreads();
        lds[t] += add;
        __syncthreads();
    }
    T incl = lds[t];
    *total = lds[kThreads - 1];
    __syncthreads();
    return incl - v;
}

template <class Tin, class Tacc>
__global__ void __launch_bounds__(kThreads) tile_totals(const Tin* in, int64_t n, Tacc* totals) {
    __shared__ Tacc lds[kThreads];
    int64_t base = (int64_t)blockIdx.x * kTile + (int64_t)threadIdx.x * kItems;
    Tacc s = 0;
#pragma unroll
    for (int i = 0; i < kItems; i++)
        if (base + i < n) s += (Tacc)in[base + i];
    Tacc tot;
    block_exclusive_scan<Tacc>(s, lds, &tot);
    if (threadIdx.x == 0) totals[blockIdx.x] = tot;
}

// Exclusive scan of `totals` in place by one workgroup, chunk by chunk.
template <class Tacc>
__global__ void __launch_bounds__(kThreads) scan_totals(Tacc* totals, int64_t n) {
    __shared__ Tacc lds[kThreads];
    Tacc carry = 0;
    for (int64_t c0 = 0; c0 < n; c0 += kThreads) {
        int64_t i = c0 + threadIdx.x;
        Tacc v = i < n ? totals[i] : Tacc(0);
        Tacc tot;
        Tacc ex = block_exclusive_scan<Tacc>(v, lds, &tot);
        if (i < n) totals[i] = carry + ex;
        carry += tot;
    }
}

template <class Tin, class Tacc, class Tout, bool INCLUSIVE>
__global__ void __launch_bounds__(kThreads) tile_scan(const Tin* in, int64_t n, const Tacc* offs,
                                                      Tout* out) {
    __shared__ Tacc lds[kThreads];
    int64_t base = (int64_t)blockIdx.x * kTile + (int64_t)threadIdx.x * kItems;
    Tacc v[kItems];
    Tacc s = 0;
#pragma unroll
    for (int i = 0; i < kItems; i++) {
        v[i] = (base + i < n) ? (Tacc)in[base + i] : Tacc(0);
        s += v[i];
    }
    Tacc tot;
    Tacc run = offs[blockIdx.x] + block_exclusive_scan<Tacc>(s, lds, &tot);
#pragma unroll
    for (int i = 0; i < kItems; i++) {
        if (INCLUSIVE) run += v[i];
        if (base + i < n) out[base + i] = (Tout)run;
        if (!INCLUSIVE) run += v[i];
    }
}

template <class T>
__global__ void fill_kernel(T* p, T v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

template <class Tacc>
__global__ void write_scalar(Tacc* dst, const Tacc* totals, int64_t ntiles, const Tacc* last_in_tot) {
    // dst = total sum = offs[last tile] + total of last tile
    *dst = totals[ntiles - 1] + *last_in_tot;
}

template <class Tin, class Tacc, class Tout, bool INCLUSIVE>
void run_scan(const Tin* in, Tout* out, int64_t n, hipStream_t s, Tacc* total_out) {
    if (n <= 0) {
        if (total_out) fill_kernel<Tacc><<<1, 1, 0, s>>>(total_out, Tacc(0), 1);
        return;
    }
    int64_t ntiles = (n + kTile - 1) / kTile;
    DevBuf tot((size_t)(ntiles + 1) * sizeof(Tacc));
    Tacc* t = tot.as<Tacc>();
    tile_totals<Tin, Tacc><<<(unsigned)ntiles, kThreads, 0, s>>>(in, n, t);
    MONA_HIP(hipGetLastError());
    if (total_out) {
        // keep the last tile's own total before the in-place exclusive scan
        MONA_HIP(hipMemcpyAsync(t + ntiles, t + ntiles - 1, sizeof(Tacc), hipMemcpyDeviceToDevice, s));
    }
    scan_totals<Tacc><<<1, kThreads, 0, s>>>(t, ntiles);
    MONA_HIP(hipGetLastError());
    tile_scan<Tin, Tacc, Tout, INCLUSIVE><<<(unsigned)ntiles, kThreads, 0, s>>>(in, n, t, out);
    MONA_HIP(hipGetLastError());
    if (total_out) {
        write_scalar<Tacc><<<1, 1, 0, s>>>(total_out, t, ntiles, t + ntiles);
        MONA_HIP(hipGetLastError());
    }
    // `tot` is released at scope exit: wait for the kernels that use it.
    MONA_HIP(hipStreamSynchronize(s));
}

}  // namespace

void scan_i32_to_i32_inclusive(const int* in, int* out, int64_t n, hipStream_t s) {
    run_scan<int, int64_t, int, true>(in, out, n, s, nullptr);
}

void scan_i32_to_i64_exclusive(const int* in, int64_t* out, int64_t n, hipStream_t s) {
    // out[0..n-1] exclusive prefix, out[n] = total
    run_scan<int, int64_t, int64_t, false>(in, out, n, s, out + n);
}

void scan_i32_to_i64_inclusive(const int* in, int64_t* out, int64_t n, hipStream_t s) {
    run_scan<int, int64_t, int64_t, true>(in, out, n, s, nullptr);
}

void fill_f32(float* p, float v, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    unsigned b = blocks_for(n, 256);
    if (b > 8192) b = 8192;
    fill_kernel<float><<<b, 256, 0, s>>>(p, v, n);
    MONA_HIP(hipGetLastError());
}

void fill_i32(int* p, int v, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    unsigned b = blocks_for(n, 256);
    if (b > 8192) b = 8192;
    fill_kernel<int><<<b, 256, 0, s>>>(p, v, n);
    MONA_HIP(hipGetLastError());
}

}  // namespace mona
