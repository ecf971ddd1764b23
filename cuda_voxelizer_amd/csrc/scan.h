// scan.h -- device-wide exclusive scans (uint64 sum, int64 max): three
// kernels (block reduce, one-workgroup scan of the block sums, block scan).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace voxamd {

constexpr int kScanPer = 8;                  // items per thread
constexpr int kScanTile = 256 * kScanPer;    // items per block

struct SumOp {
    __device__ static uint64_t id() { return 0; }
    __device__ static uint64_t op(uint64_t a, uint64_t b) { return a + b; }
};
struct MaxOp {  // on int64 stored as uint64
    __device__ static uint64_t id() { return (uint64_t)(int64_t)-1; }
    __device__ static uint64_t op(uint64_t a, uint64_t b) { return (int64_t)a > (int64_t)b ? a : b; }
};

template <class Op>
__device__ __forceinline__ uint64_t wave_incl_scan(uint64_t v) {
    const int lane = __lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint64_t o = __shfl_up(v, d);
        if (lane >= d) v = Op::op(o, v);
    }
    return v;
}

// Block-wide exclusive scan of one value per thread (256 threads); returns
// the exclusive prefix, *total gets the block total.
template <class Op>
__device__ __forceinline__ uint64_t block_excl_scan(uint64_t v, uint64_t* total) {
    __shared__ uint64_t s_wave[4];
    const int lane = __lane_id(), wave = threadIdx.x >> 6;
    const uint64_t incl = wave_incl_scan<Op>(v);
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    uint64_t before = Op::id();
    for (int k = 0; k < wave; k++) before = Op::op(before, s_wave[k]);
    uint64_t all = Op::id();
    for (int k = 0; k < 4; k++) all = Op::op(all, s_wave[k]);
    __syncthreads();
    uint64_t excl = __shfl_up(incl, 1);
    if (lane == 0) excl = Op::id();
    *total = all;
    return Op::op(before, excl);
}

template <class Op, class T>
__global__ void __launch_bounds__(256) scan_reduce_kernel(const T* __restrict__ in, uint64_t n,
                                                          uint64_t* __restrict__ sums) {
    const uint64_t i0 = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanPer;
    uint64_t v = Op::id();
#pragma unroll
    for (int k = 0; k < kScanPer; k++)
        if (i0 + k < n) v = Op::op(v, (uint64_t)in[i0 + k]);
    uint64_t total;
    block_excl_scan<Op>(v, &total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// One workgroup: exclusive scan of the block sums, in place; kScanPer
// consecutive sums per thread, so up to kScanTile sums take one block scan.
template <class Op>
__global__ void __launch_bounds__(256) scan_top_kernel(uint64_t* __restrict__ sums, uint64_t nb) {
    uint64_t carry = Op::id();
    for (uint64_t b0 = 0; b0 < nb; b0 += kScanTile) {
        const uint64_t i0 = b0 + (uint64_t)threadIdx.x * kScanPer;
        uint64_t vals[kScanPer];
        uint64_t v = Op::id();
#pragma unroll
        for (int k = 0; k < kScanPer; k++) {
            vals[k] = i0 + k < nb ? sums[i0 + k] : Op::id();
            v = Op::op(v, vals[k]);
        }
        uint64_t total;
        uint64_t run = Op::op(carry, block_excl_scan<Op>(v, &total));
#pragma unroll
        for (int k = 0; k < kScanPer; k++) {
            if (i0 + k < nb) sums[i0 + k] = run;
            run = Op::op(run, vals[k]);
        }
        carry = Op::op(carry, total);
    }
}

template <class Op, class T>
__global__ void __launch_bounds__(256) scan_down_kernel(const T* __restrict__ in, uint64_t n,
                                                        const uint64_t* __restrict__ sums,
                                                        uint64_t* __restrict__ out) {
    const uint64_t i0 = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanPer;
    uint64_t vals[kScanPer];
    uint64_t v = Op::id();
#pragma unroll
    for (int k = 0; k < kScanPer; k++) {
        vals[k] = i0 + k < n ? (uint64_t)in[i0 + k] : Op::id();
        v = Op::op(v, vals[k]);
    }
    uint64_t total;
    uint64_t run = Op::op(sums[blockIdx.x], block_excl_scan<Op>(v, &total));
#pragma unroll
    for (int k = 0; k < kScanPer; k++) {
        if (i0 + k < n) out[i0 + k] = run;
        run = Op::op(run, vals[k]);
    }
}

template <class Op, class T>
inline hipError_t device_scan(const T* in, uint64_t n, uint64_t* sums, uint64_t* out, hipStream_t st) {
    const uint64_t nb = (n + kScanTile - 1) / kScanTile;
    if (nb == 0) return hipSuccess;
    scan_reduce_kernel<Op, T><<<(unsigned)nb, 256, 0, st>>>(in, n, sums);
    scan_top_kernel<Op><<<1, 256, 0, st>>>(sums, nb);
    scan_down_kernel<Op, T><<<(unsigned)nb, 256, 0, st>>>(in, n, sums, out);
    return hipGetLastError();
}

}  // namespace voxamd
