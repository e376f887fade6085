// pzl_sort.hip -- ordering of a batch's hits into PRESTO C scan order.
//
// Replaces the reference's thrust sort of the hit array (sort_search_results,
// accel_kernels.cu:990-1006).  The harmonic sum compacts hits with one atomic
// per wave, so they arrive in an arbitrary order; their keys encode
// (rstep, stage, w, z, r) and are distinct.  This is a least-significant-digit
// radix sort of (key, power) pairs with digits of at most 8 bits:
//
//   k_rs_hist     per tile of RS_TILE keys: a histogram of the pass's digit,
//                 written digit-major ([digit][tile]) so that one exclusive
//                 scan of the whole table gives every (digit, tile) its output
//                 offset;
//   k_scan_*      that exclusive scan (one workgroup for small tables, three
//                 kernels otherwise);
//   k_rs_scatter  per tile, slice by slice of 256 consecutive keys: the rank
//                 of each key among the keys of its digit before it, from wave
//                 ballots on the digit's bits (lanes with the same digit) and
//                 per-wave digit counts in LDS; keys are written to the digit's
//                 running offset, so every pass is stable.
//
// The number of passes is made odd (digits narrowed as needed), so that the
// ordered pairs end in the second buffer pair.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "pzl_kernels.h"

namespace {

constexpr int RS_NT = 256;                  // threads per workgroup (4 waves)
constexpr int RS_SLICES = 8;                // slices of RS_NT keys per tile
constexpr int RS_TILE = RS_NT * RS_SLICES;  // keys per tile
constexpr int RS_BINS = 256;                // digits of at most 8 bits
constexpr int SCAN_NT = 1024;
constexpr int SCAN_PER = 4;                           // entries per thread of the block scan
constexpr int SCAN_BLOCK = SCAN_NT * SCAN_PER;        // entries per block of the block scan
constexpr long long SCAN_SMALL = (long long)SCAN_NT * 64;  // one workgroup up to this length

__device__ __forceinline__ int digit_of(long long key, int shift, int bits)
{
    return (int)(((unsigned long long)key >> shift) & ((1u << bits) - 1u));
}

__global__ __launch_bounds__(RS_NT) void k_rs_hist(const long long *keys, long long n, int shift, int bits,
                                                    unsigned *counts, int ntiles)
{
    __shared__ unsigned hist[RS_BINS];
    const int tid = threadIdx.x;
    hist[tid] = 0;
    __syncthreads();
    const long long base = (long long)blockIdx.x * RS_TILE;
#pragma unroll
    for (int j = 0; j < RS_SLICES; j++) {
        const long long i = base + j * RS_NT + tid;
        if (i < n) atomicAdd(&hist[digit_of(keys[i], shift, bits)], 1u);
    }
    __syncthreads();
    counts[(long long)tid * ntiles + blockIdx.x] = hist[tid];
}

// Exclusive scan of a workgroup's values (one per thread), returning the
// exclusive prefix of this thread and the workgroup total in *total.
template <int NT> __device__ __forceinline__ unsigned block_exclusive_scan(unsigned v, unsigned *total)
{
    __shared__ unsigned wsum[NT / 64];
    const int lane = __lane_id(), wave = threadIdx.x / 64;
    unsigned x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const unsigned y = __shfl_up(x, d);
        if (lane >= d) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    unsigned before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < NT / 64; w++) {
        const unsigned s = wsum[w];
        before += w < wave ? s : 0u;
        all += s;
    }
    __syncthreads();
    *total = all;
    return before + x - v;
}

// The whole table in one workgroup: each thread scans a contiguous run.
__global__ __launch_bounds__(SCAN_NT) void k_scan_small(unsigned *a, long long len)
{
    const long long per = (len + SCAN_NT - 1) / SCAN_NT;
    const long long lo = threadIdx.x * per, hi = std::min(len, lo + per);
    unsigned s = 0;
    for (long long i = lo; i < hi; i++) s += a[i];
    unsigned total;
    unsigned run = block_exclusive_scan<SCAN_NT>(s, &total);
    for (long long i = lo; i < hi; i++) {
        const unsigned v = a[i];
        a[i] = run;
        run += v;
    }
}

// Long tables: per block of SCAN_BLOCK entries an in-place exclusive scan
// and the block's total, then the totals scanned, then added back.
__global__ __launch_bounds__(SCAN_NT) void k_scan_blocks(unsigned *a, long long len, unsigned *totals)
{
    const long long lo = (long long)blockIdx.x * SCAN_BLOCK + threadIdx.x * SCAN_PER;
    unsigned v[SCAN_PER], s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_PER; k++) {
        v[k] = lo + k < len ? a[lo + k] : 0u;
        s += v[k];
    }
    unsigned total;
    unsigned run = block_exclusive_scan<SCAN_NT>(s, &total);
#pragma unroll
    for (int k = 0; k < SCAN_PER; k++) {
        if (lo + k < len) a[lo + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = total;
}

__global__ __launch_bounds__(SCAN_NT) void k_scan_add(unsigned *a, long long len, const unsigned *totals)
{
    const long long lo = (long long)blockIdx.x * SCAN_BLOCK + threadIdx.x * SCAN_PER;
    const unsigned t = totals[blockIdx.x];
#pragma unroll
    for (int k = 0; k < SCAN_PER; k++)
        if (lo + k < len) a[lo + k] += t;
}

__global__ __launch_bounds__(RS_NT) void k_rs_scatter(const long long *keys_in, const float *pows_in,
                                                       long long *keys_out, float *pows_out, long long n,
                                                       int shift, int bits, const unsigned *offsets, int ntiles)
{
    constexpr int NW = RS_NT / 64;
    __shared__ unsigned run[RS_BINS];       // next output position of each digit
    __shared__ unsigned wcnt[NW][RS_BINS];  // this slice's count of each digit per wave
    const int tid = threadIdx.x, lane = __lane_id(), wave = tid / 64;
    run[tid] = offsets[(long long)tid * ntiles + blockIdx.x];
#pragma unroll
    for (int w = 0; w < NW; w++) wcnt[w][tid] = 0;
    const unsigned long long lanemask_lt = (1ull << lane) - 1ull;
    const long long base = (long long)blockIdx.x * RS_TILE;
    __syncthreads();
    for (int j = 0; j < RS_SLICES; j++) {
        const long long i = base + j * RS_NT + tid;
        const bool valid = i < n;
        long long key = 0;
        float pw = 0.f;
        int d = 0;
        if (valid) {
            key = keys_in[i];
            pw = pows_in[i];
            d = digit_of(key, shift, bits);
        }
        // lanes of this wave holding the same digit
        unsigned long long peers = __ballot(valid);
        for (int k = 0; k < bits; k++) {
            const bool bit = (d >> k) & 1;
            const unsigned long long bb = __ballot(bit);
            peers &= bit ? bb : ~bb;
        }
        const unsigned rank = (unsigned)__popcll(peers & lanemask_lt);
        if (valid && rank == 0) wcnt[wave][d] = (unsigned)__popcll(peers);
        __syncthreads();
        if (valid) {
            unsigned pos = run[d] + rank;
            for (int w = 0; w < wave; w++) pos += wcnt[w][d];
            if (pos < (unsigned long long)n) {  // always, when the offsets are the scan of the histograms
                keys_out[pos] = key;
                pows_out[pos] = pw;
            }
        }
        __syncthreads();
        unsigned add = 0;
#pragma unroll
        for (int w = 0; w < NW; w++) {
            add += wcnt[w][tid];
            wcnt[w][tid] = 0;
        }
        run[tid] += add;
        __syncthreads();
    }
}

hipError_t scan(unsigned *a, long long len, unsigned *totals, hipStream_t st)
{
    if (len <= SCAN_SMALL) {
        hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(SCAN_NT), 0, st, a, len);
        return hipGetLastError();
    }
    const long long nb = (len + SCAN_BLOCK - 1) / SCAN_BLOCK;
    if (nb > SCAN_SMALL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_scan_blocks, dim3((unsigned)nb), dim3(SCAN_NT), 0, st, a, len, totals);
    hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(SCAN_NT), 0, st, totals, nb);
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nb), dim3(SCAN_NT), 0, st, a, len, (const unsigned *)totals);
    return hipGetLastError();
}

}  // namespace

// Orders n (key, power) pairs by key on bits [0, end_bit): the pairs start in
// (keys, pows) and end in (keys_alt, pows_alt); both pairs are overwritten.
// temp == nullptr: *temp_bytes gets the scratch n needs.
extern "C" hipError_t pzl_sort_hits(void *temp, size_t *temp_bytes, long long *keys, float *pows,
                                    long long *keys_alt, float *pows_alt, long long n, int end_bit,
                                    hipStream_t st)
{
    const long long ntiles = std::max<long long>((n + RS_TILE - 1) / RS_TILE, 1);
    const long long table = (long long)RS_BINS * ntiles;
    const long long ntot = (table + SCAN_BLOCK - 1) / SCAN_BLOCK;
    const size_t need = sizeof(unsigned) * (size_t)(table + ntot + 64);
    if (!temp) {
        *temp_bytes = need;
        return hipSuccess;
    }
    if (*temp_bytes < need || n < 0 || end_bit < 1 || end_bit > 64 || ntiles > 0x7fffffffLL)
        return hipErrorInvalidValue;
    if (n == 0) return hipSuccess;
    int passes = (end_bit + 7) / 8;
    if (passes % 2 == 0) passes++;  // the ordered pairs end in the alternate buffers
    const int bits = (end_bit + passes - 1) / passes;
    unsigned *counts = (unsigned *)temp, *totals = counts + table;
    long long *kin = keys, *kout = keys_alt;
    float *pin = pows, *pout = pows_alt;
    for (int p = 0, shift = 0; p < passes; p++, shift += bits) {
        // (passes - 1) * bits < end_bit <= 64: every pass has at least one bit
        const int b = std::min(bits, 64 - shift);
        hipLaunchKernelGGL(k_rs_hist, dim3((unsigned)ntiles), dim3(RS_NT), 0, st, (const long long *)kin, n, shift,
                           b, counts, (int)ntiles);
        hipError_t e = scan(counts, table, totals, st);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_rs_scatter, dim3((unsigned)ntiles), dim3(RS_NT), 0, st, (const long long *)kin,
                           (const float *)pin, kout, pout, n, shift, b, (const unsigned *)counts, (int)ntiles);
        std::swap(kin, kout);
        std::swap(pin, pout);
    }
    return hipGetLastError();
}
