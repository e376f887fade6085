// pzl_realfft.hip -- realfft (fastffts.c:163-235) on the GPU: the .dat -> .fft
// producer of the search.  A real series of n floats is an n/2-point complex
// series; its FFT and the Numerical Recipes packing (bin 0 = (DC, Nyquist))
// give the n/2 bins of the real FFT, and the inverse undoes both.
//
//   n/2 = M <= RF_MAXE points: one workgroup transforms the series in LDS.
//   larger M = M1 * M2 (four steps, both factors <= 8192):
//     k_rf_cols   FFTs of length M1 down the columns of the [M1][M2] view,
//                 C adjacent columns per workgroup, times exp(SIGN 2 pi i k1 j2 / M)
//     k_rf_rows   FFTs of length M2 along the rows, C2 rows per workgroup,
//                 written transposed: X[k1 + M1 k2]
//   k_rf_unpack_fwd / k_rf_pack_inv   the packing step of the forward and
//                 inverse transform (one thread per pair of bins il, n/2 - il)
//
// Data stay float (as FFTW's float transforms that the reference calls), with
// twiddles computed in double and rounded once; the packing step computes in
// double from float inputs, as the reference does.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "pzl_accel.h"
#include "pzl_fft.h"
#include "pzl_plan.h"

#define RF_NT 1024
#define RF_MAXE 16448  // float2 per workgroup's LDS buffer (128.5 KiB)
#define RF_U 8         // global loads in flight per thread while staging a tile

namespace {

// Division by an invariant d (1 <= d < 2^31) of n < 2^31 with one multiply-high
// (Granlund and Montgomery, "Division by invariant integers using
// multiplication", PLDI 1994): l = ceil(log2 d), m = floor(2^32 (2^l - d) / d)
// + 1, n / d = (mulhi(m, n) + n) >> l.
struct RfDiv {
    unsigned m;
    int l;
    __device__ explicit RfDiv(int d)
    {
        l = 0;
        while ((1u << l) < (unsigned)d) l++;
        m = (unsigned)((((unsigned long long)1 << 32) * (((unsigned long long)1 << l) - (unsigned)d)) / (unsigned)d + 1);
    }
    __device__ int div(int n) const { return (int)((__umulhi(m, (unsigned)n) + (unsigned)n) >> l); }
};

// One Stockham pass of radix R over nseq sequences of length L in LDS
// (sequence s at buf + s * sstride); twiddle exp(SIGN 2 pi i jm r / (Ns R)) =
// tw[jm r L / (Ns R)] (conjugated for SIGN +1), tw[j] = exp(-2 pi i j / L).
template <int SIGN, int R>
__device__ __forceinline__ void rf_pass(float2 *buf, int L, int nseq, int sstride, int Ns, const float2 *__restrict__ tw, int tid)
{
    constexpr int MAXB = (RF_MAXE / R + RF_NT - 1) / RF_NT;
    const int nbf = L / R, total = nbf * nseq, step = L / (Ns * R);
    const RfDiv dnbf(nbf), dns(Ns);
    float2 v[MAXB][R];
#pragma unroll
    for (int b = 0; b < MAXB; b++) {
        int g = tid + b * RF_NT;
        // opaque to the optimiser: the butterfly indices are recomputed in
        // every pass instead of being hoisted (and spilled) for all passes
        asm volatile("" : "+v"(g));
        if (g < total) {
            const int s = dnbf.div(g), j = g - s * nbf, jm = j - dns.div(j) * Ns;
            const float2 *q = buf + s * sstride;
            v[b][0] = q[j];
#pragma unroll
            for (int r = 1; r < R; r++) {
                float2 t = tw[jm * r * step];
                if (SIGN > 0) t.y = -t.y;
                v[b][r] = cmul(q[j + r * nbf], t);
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < MAXB; b++) {
        int g = tid + b * RF_NT;
        asm volatile("" : "+v"(g));
        if (g < total) {
            const int s = dnbf.div(g), j = g - s * nbf, jq = dns.div(j), jm = j - jq * Ns;
            if constexpr (R == 7 || R == 11) dft_odd<R, SIGN>(v[b]);
            else dftR<R, SIGN>(v[b]);
            float2 *q = buf + s * sstride;
            const int o = jq * Ns * R + jm;
#pragma unroll
            for (int r = 0; r < R; r++) q[o + r * Ns] = v[b][r];
        }
    }
    __syncthreads();
}

// Unnormalised FFTs of nseq sequences of length L = 2^a 3^b 5^c 7^d 11^e in
// LDS: passes of radix 16 while 16 divides what is left, then 8, 4, 2, 3, 5,
// 7, 11.
template <int SIGN>
__device__ __forceinline__ void rf_fft(float2 *buf, int L, int nseq, int sstride, const float2 *__restrict__ tw, int tid)
{
    int Ns = 1;
    while (Ns < L) {
        const int rem = L / Ns;
        if (rem % 16 == 0) {
            rf_pass<SIGN, 16>(buf, L, nseq, sstride, Ns, tw, tid);
            Ns *= 16;
        } else if (rem % 8 == 0) {
            rf_pass<SIGN, 8>(buf, L, nseq, sstride, Ns, tw, tid);
            Ns *= 8;
        } else if (rem % 4 == 0) {
            rf_pass<SIGN, 4>(buf, L, nseq, sstride, Ns, tw, tid);
            Ns *= 4;
        } else if (rem % 2 == 0) {
            rf_pass<SIGN, 2>(buf, L, nseq, sstride, Ns, tw, tid);
            Ns *= 2;
        } else if (rem % 3 == 0) {
            rf_pass<SIGN, 3>(buf, L, nseq, sstride, Ns, tw, tid);
            Ns *= 3;
        } else if (rem % 5 == 0) {
            rf_pass<SIGN, 5>(buf, L, nseq, sstride, Ns, tw, tid);
            Ns *= 5;
        } else if (rem % 7 == 0) {
            rf_pass<SIGN, 7>(buf, L, nseq, sstride, Ns, tw, tid);
            Ns *= 7;
        } else {
            rf_pass<SIGN, 11>(buf, L, nseq, sstride, Ns, tw, tid);
            Ns *= 11;
        }
    }
}

// Workgroups are dispatched to the 8 XCDs in turn; give each XCD a contiguous
// run of column (row) groups so that the groups in flight on one XCD share the
// cache lines of their partial-line accesses in its L2.
__device__ inline long long rf_group(int b, int nb)
{
    return (nb % 8) ? b : (long long)(b % 8) * (nb / 8) + b / 8;
}

// The same pass with L and Ns known at compile time (powers of two): the
// butterfly indices are shifts and masks.  The sequence count stays a
// run-time value.
template <int SIGN, int R, int L, int Ns>
__device__ __forceinline__ void rf_pass_ct(float2 *buf, int nseq, int sstride, const float2 *__restrict__ tw, int tid)
{
    constexpr int MAXB = (RF_MAXE / R + RF_NT - 1) / RF_NT;
    constexpr int nbf = L / R, step = L / (Ns * R);
    const int total = nbf * nseq;
    float2 v[MAXB][R];
#pragma unroll
    for (int b = 0; b < MAXB; b++) {
        const int g = tid + b * RF_NT;
        if (g < total) {
            const int s = g / nbf, j = g % nbf, jm = j % Ns;
            const float2 *q = buf + s * sstride;
            v[b][0] = q[j];
#pragma unroll
            for (int r = 1; r < R; r++) {
                float2 t = tw[jm * r * step];
                if (SIGN > 0) t.y = -t.y;
                v[b][r] = cmul(q[j + r * nbf], t);
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < MAXB; b++) {
        const int g = tid + b * RF_NT;
        if (g < total) {
            const int s = g / nbf, j = g % nbf, jm = j % Ns, jq = j / Ns;
            dftR<R, SIGN>(v[b]);
            float2 *q = buf + s * sstride;
            const int o = jq * Ns * R + jm;
#pragma unroll
            for (int r = 0; r < R; r++) q[o + r * Ns] = v[b][r];
        }
    }
    __syncthreads();
}

template <int SIGN, int L, int Ns = 1>
__device__ __forceinline__ void rf_fft_ct(float2 *buf, int nseq, int sstride, const float2 *__restrict__ tw, int tid)
{
    if constexpr (Ns < L) {
        constexpr int rem = L / Ns;
        constexpr int R = rem % 16 == 0 ? 16 : rem % 8 == 0 ? 8 : rem % 4 == 0 ? 4 : 2;
        rf_pass_ct<SIGN, R, L, Ns>(buf, nseq, sstride, tw, tid);
        rf_fft_ct<SIGN, L, Ns * R>(buf, nseq, sstride, tw, tid);
    }
}

// FFT length L of the four-step kernels: 0 for the run-time pass loop, else a
// power of two compiled into the pass sequence.
template <int SIGN, int L>
__device__ __forceinline__ void rf_fft_any(float2 *buf, int n, int nseq, int sstride, const float2 *__restrict__ tw,
                                           int tid)
{
    if constexpr (L == 0) rf_fft<SIGN>(buf, n, nseq, sstride, tw, tid);
    else rf_fft_ct<SIGN, L>(buf, nseq, sstride, tw, tid);
}

__device__ inline float2 rf_scale(float2 v, double scale)
{
    return float2{(float)((double)v.x * scale), (float)((double)v.y * scale)};
}

// tw[j] = exp(-2 pi i j / L), j < L
__global__ __launch_bounds__(256) void k_rf_twiddles(float2 *__restrict__ tw, int L)
{
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j < L) {
        double s, c;
        sincospi(-2.0 * (double)j / (double)L, &s, &c);
        tw[j] = float2{(float)c, (float)s};
    }
}

// The four-step twiddles exp(-2 pi i p / M), p < M, as the product of
// lo[p % RF_TB] = exp(-2 pi i (p % RF_TB) / M) and hi[p / RF_TB] =
// exp(-2 pi i RF_TB (p / RF_TB) / M), both in double.
#define RF_TB 4096
__global__ __launch_bounds__(256) void k_rf_twiddles_step(double2 *__restrict__ lo, double2 *__restrict__ hi,
                                                          long long M, int nhi)
{
    const int j = blockIdx.x * 256 + threadIdx.x;
    double s, c;
    if (j < RF_TB) {
        sincospi(-2.0 * (double)j / (double)M, &s, &c);
        lo[j] = double2{c, s};
    }
    if (j < nhi) {
        sincospi(-2.0 * (double)((long long)j * RF_TB % M) / (double)M, &s, &c);
        hi[j] = double2{c, s};
    }
}

// The whole series in one workgroup, in place; every value times 'scale'
// (1: none).
template <int SIGN>
__global__ __launch_bounds__(RF_NT) void k_rf_single(float2 *__restrict__ data, int M, const float2 *__restrict__ tw,
                                                     double scale)
{
    extern __shared__ float2 rbuf[];
    const int tid = threadIdx.x;
    data += (long long)blockIdx.y * M;  // series of the batch
    for (int e = tid; e < M; e += RF_NT) rbuf[e] = data[e];
    __syncthreads();
    rf_fft<SIGN>(rbuf, M, 1, M, tw, tid);
    for (int e = tid; e < M; e += RF_NT) data[e] = scale == 1.0 ? rbuf[e] : rf_scale(rbuf[e], scale);
}

// Columns j2b .. j2b + C - 1 of the [M1][M2] view of 'in': FFT of length M1
// down each, times exp(SIGN 2 pi i k1 j2 / M), into 'out' at the same places.
template <int SIGN, int L>
__global__ __launch_bounds__(RF_NT) void k_rf_cols(const float2 *__restrict__ in, float2 *__restrict__ out, int M1,
                                                   int M2, int C, const float2 *__restrict__ tw,
                                                   const double2 *__restrict__ tlo, const double2 *__restrict__ thi)
{
    extern __shared__ float2 rbuf[];
    const int tid = threadIdx.x, sstride = M1 + 1;
    const long long j2b = rf_group(blockIdx.x, gridDim.x) * C;
    const long long M = (long long)M1 * M2;
    in += blockIdx.y * M;  // series of the batch (M float2 each)
    out += blockIdx.y * M;
    const int total = C * M1;
    const RfDiv dc(C);
    // RF_U loads in flight per thread before their LDS stores
    for (int e0 = 0; e0 < total; e0 += RF_U * RF_NT) {
        float2 t[RF_U];
#pragma unroll
        for (int u = 0; u < RF_U; u++) {
            const int e = e0 + u * RF_NT + tid;
            if (e < total) {
                const int j1 = dc.div(e), c = e - j1 * C;
                t[u] = in[(long long)j1 * M2 + j2b + c];
            }
        }
#pragma unroll
        for (int u = 0; u < RF_U; u++) {
            const int e = e0 + u * RF_NT + tid;
            if (e < total) {
                const int j1 = dc.div(e), c = e - j1 * C;
                rbuf[c * sstride + j1] = t[u];
            }
        }
    }
    __syncthreads();
    rf_fft_any<SIGN, L>(rbuf, M1, C, sstride, tw, tid);
#pragma unroll 4
    for (int e = tid; e < total; e += RF_NT) {
        const int k1 = dc.div(e), c = e - k1 * C;
        const long long j2 = j2b + c, p = (long long)k1 * j2;  // < M1 M2 = M
        const double2 a = tlo[p % RF_TB], h = thi[p / RF_TB];
        const double co = a.x * h.x - a.y * h.y, sn = (a.x * h.y + a.y * h.x) * (double)(-SIGN);
        out[(long long)k1 * M2 + j2] = cmul(rbuf[c * sstride + k1], float2{(float)co, (float)sn});
    }
}

// Rows k1b .. k1b + C2 - 1 of the [M1][M2] view of 'in': FFT of length M2
// along each, element k2 of row k1 to out[k1 + M1 k2], times 'scale'.
template <int SIGN, int L>
__global__ __launch_bounds__(RF_NT) void k_rf_rows(const float2 *__restrict__ in, float2 *__restrict__ out, int M1,
                                                   int M2, int C2, const float2 *__restrict__ tw, double scale)
{
    extern __shared__ float2 rbuf[];
    const int tid = threadIdx.x, sstride = M2 + 1;
    const long long k1b = rf_group(blockIdx.x, gridDim.x) * C2;
    in += blockIdx.y * (long long)M1 * M2;  // series of the batch
    out += blockIdx.y * (long long)M1 * M2;
    const int total = C2 * M2;
    const RfDiv dm2(M2), dc2(C2);
    for (int e0 = 0; e0 < total; e0 += RF_U * RF_NT) {
        float2 t[RF_U];
#pragma unroll
        for (int u = 0; u < RF_U; u++) {
            const int e = e0 + u * RF_NT + tid;
            if (e < total) {
                const int c = dm2.div(e), j2 = e - c * M2;
                t[u] = in[(k1b + c) * M2 + j2];
            }
        }
#pragma unroll
        for (int u = 0; u < RF_U; u++) {
            const int e = e0 + u * RF_NT + tid;
            if (e < total) {
                const int c = dm2.div(e), j2 = e - c * M2;
                rbuf[c * sstride + j2] = t[u];
            }
        }
    }
    __syncthreads();
    rf_fft_any<SIGN, L>(rbuf, M2, C2, sstride, tw, tid);
#pragma unroll 4
    for (int e = tid; e < total; e += RF_NT) {
        const int k2 = dc2.div(e), c = e - k2 * C2;
        const float2 v = rbuf[c * sstride + k2];
        out[(long long)k2 * M1 + k1b + c] = scale == 1.0 ? v : rf_scale(v, scale);
    }
}

// The packing step of realfft's forward branch (fastffts.c:191-223) for the
// pair il, ih = n/2 - il (il <= n/4), with w = exp(-2 pi i il / n); il = 0
// sets bin 0 to (DC, Nyquist).  Where il = ih the second write is the one
// that stays, as in the reference.
__global__ __launch_bounds__(256) void k_rf_unpack_fwd(float2 *__restrict__ d, long long n)
{
    const long long il = (long long)blockIdx.x * 256 + threadIdx.x;
    const long long nby2 = n >> 1;
    d += blockIdx.y * nby2;  // series of the batch
    if (il > (n >> 2)) return;
    if (il == 0) {
        const float2 z = d[0];
        const double t = z.x;
        d[0] = float2{(float)(t + z.y), (float)(t - z.y)};
        return;
    }
    const long long ih = nby2 - il;
    const double cc = -0.5;
    double wi, wr;
    sincospi(-2.0 * (double)il / (double)n, &wi, &wr);
    const float2 a = d[il], b = d[ih];
    const double h1r = 0.5 * ((double)a.x + b.x), h1i = 0.5 * ((double)a.y - b.y);
    const double h2r = -cc * ((double)a.y + b.y), h2i = cc * ((double)a.x - b.x);
    const double h2rwr = h2r * wr, h2rwi = h2r * wi, h2iwr = h2i * wr, h2iwi = h2i * wi;
    const float2 lo{(float)(h1r + h2rwr - h2iwi), (float)(h1i + h2iwr + h2rwi)};
    const float2 hi{(float)(h1r - h2rwr + h2iwi), (float)(-h1i + h2iwr + h2rwi)};
    if (il != ih) d[il] = lo;
    d[ih] = hi;
}

// The packing step of the inverse branch (fastffts.c:183-189, 191-223,
// 224-230): bin n/4's imaginary part negated before and after the pair loop
// when n % 4 == 0, w = exp(+2 pi i il / n), bin 0 from (DC, Nyquist).
__global__ __launch_bounds__(256) void k_rf_pack_inv(float2 *__restrict__ d, long long n)
{
    const long long il = (long long)blockIdx.x * 256 + threadIdx.x;
    const long long nby2 = n >> 1;
    d += blockIdx.y * nby2;  // series of the batch
    if (il > (n >> 2)) return;
    if (il == 0) {
        const float2 z = d[0];
        const double t = z.x;
        d[0] = float2{(float)(0.5 * (t + z.y)), (float)(0.5 * (t - z.y))};
        return;
    }
    const long long ih = nby2 - il;
    const bool flip = ((n + 2) % 4) != 0 && il == (n >> 2);  // the bin n/4 = this pair's il (= ih)
    const double cc = 0.5;
    double wi, wr;
    sincospi(2.0 * (double)il / (double)n, &wi, &wr);
    float2 a = d[il], b = d[ih];
    if (flip) {
        a.y = -a.y;
        b.y = -b.y;
    }
    const double h1r = 0.5 * ((double)a.x + b.x), h1i = 0.5 * ((double)a.y - b.y);
    const double h2r = -cc * ((double)a.y + b.y), h2i = cc * ((double)a.x - b.x);
    const double h2rwr = h2r * wr, h2rwi = h2r * wi, h2iwr = h2i * wr, h2iwi = h2i * wi;
    float2 lo{(float)(h1r + h2rwr - h2iwi), (float)(h1i + h2iwr + h2rwi)};
    float2 hi{(float)(h1r - h2rwr + h2iwi), (float)(-h1i + h2iwr + h2rwi)};
    if (flip) hi.y = -hi.y;
    if (il != ih) d[il] = lo;
    d[ih] = hi;
}

bool smooth(long long m)
{
    if (m < 1) return false;
    for (int p : {2, 3, 5, 7, 11})
        while (m % p == 0) m /= p;
    return m == 1;
}

// Columns (or rows) per workgroup: the largest power of two up to 16 that
// divides 'count' and fits C * (L + 1) float2 in the LDS buffer.
int per_block(int count, int L)
{
    int c = 16;
    while (c > 1 && (count % c != 0 || c * (L + 1) > RF_MAXE)) c >>= 1;
    return c;
}

// The split M = M1 * M2 (both smooth, both <= 8192) with the factors closest
// to each other, the column length M1 >= M2; false if none.
bool split(long long M, int *M1, int *M2)
{
    long long best = 0;
    for (long long m2 = 1; m2 <= 8192 && m2 * m2 <= M; m2++) {
        if (M % m2 || !smooth(m2)) continue;
        const long long m1 = M / m2;
        if (m1 > 8192 || !smooth(m1)) continue;
        best = m2;
    }
    if (!best) return false;
    *M2 = (int)best;
    *M1 = (int)(M / best);
    return true;
}

// Workspace per device: one n/2 complex buffer and the two twiddle tables;
// calls on a device are serialised.
struct RfWork {
    std::mutex mu;
    float2 *buf = nullptr;
    long long cap = 0;  // float2
};
std::mutex g_rf_mu;
std::vector<RfWork *> g_rf_work;

RfWork *work_for(int device)
{
    std::lock_guard<std::mutex> lk(g_rf_mu);
    if ((int)g_rf_work.size() <= device) g_rf_work.resize(device + 1, nullptr);
    if (!g_rf_work[device]) g_rf_work[device] = new RfWork();
    return g_rf_work[device];
}

}  // namespace

using namespace pzl;

// The column / row kernels for FFT length L: the compiled pass sequence for the
// powers of two 256 .. 8192, the run-time pass loop otherwise.
template <int SIGN, typename... A>
static hipError_t launch_cols(int L, dim3 grid, size_t lds, hipStream_t st, A... a)
{
    switch (L) {
    case 256: hipLaunchKernelGGL((k_rf_cols<SIGN, 256>), grid, dim3(RF_NT), lds, st, a...); break;
    case 512: hipLaunchKernelGGL((k_rf_cols<SIGN, 512>), grid, dim3(RF_NT), lds, st, a...); break;
    case 1024: hipLaunchKernelGGL((k_rf_cols<SIGN, 1024>), grid, dim3(RF_NT), lds, st, a...); break;
    case 2048: hipLaunchKernelGGL((k_rf_cols<SIGN, 2048>), grid, dim3(RF_NT), lds, st, a...); break;
    case 4096: hipLaunchKernelGGL((k_rf_cols<SIGN, 4096>), grid, dim3(RF_NT), lds, st, a...); break;
    case 8192: hipLaunchKernelGGL((k_rf_cols<SIGN, 8192>), grid, dim3(RF_NT), lds, st, a...); break;
    default: hipLaunchKernelGGL((k_rf_cols<SIGN, 0>), grid, dim3(RF_NT), lds, st, a...); break;
    }
    return hipGetLastError();
}

template <int SIGN, typename... A>
static hipError_t launch_rows(int L, dim3 grid, size_t lds, hipStream_t st, A... a)
{
    switch (L) {
    case 256: hipLaunchKernelGGL((k_rf_rows<SIGN, 256>), grid, dim3(RF_NT), lds, st, a...); break;
    case 512: hipLaunchKernelGGL((k_rf_rows<SIGN, 512>), grid, dim3(RF_NT), lds, st, a...); break;
    case 1024: hipLaunchKernelGGL((k_rf_rows<SIGN, 1024>), grid, dim3(RF_NT), lds, st, a...); break;
    case 2048: hipLaunchKernelGGL((k_rf_rows<SIGN, 2048>), grid, dim3(RF_NT), lds, st, a...); break;
    case 4096: hipLaunchKernelGGL((k_rf_rows<SIGN, 4096>), grid, dim3(RF_NT), lds, st, a...); break;
    case 8192: hipLaunchKernelGGL((k_rf_rows<SIGN, 8192>), grid, dim3(RF_NT), lds, st, a...); break;
    default: hipLaunchKernelGGL((k_rf_rows<SIGN, 0>), grid, dim3(RF_NT), lds, st, a...); break;
    }
    return hipGetLastError();
}

// The workspace (float2 elements) of nb transforms of n floats: the series
// workspace, the two FFT twiddle tables and the four-step tables (double2 = 2
// float2).  Checks the arguments; fills M1, M2 and 'single'.
static int realfft_workspace(long long n, int nb, int isign, int *M1, int *M2, bool *single, long long *need)
{
    if (n < 2 || n % 2) return fail(PZL_ERR_ARG, "realfft: n = %lld must be even and >= 2", n);
    if (isign != -1 && isign != 1) return fail(PZL_ERR_ARG, "realfft: isign must be -1 or +1");
    if (nb < 1 || nb > 65535) return fail(PZL_ERR_ARG, "realfft: %d series (1 to 65535)", nb);
    const long long M = n / 2;
    *M1 = (int)M;
    *M2 = 1;
    *single = M <= RF_MAXE;
    if (!smooth(M) || (!*single && !split(M, M1, M2)))
        return fail(PZL_ERR_UNSUPPORTED, "realfft: n/2 = %lld is not a product of 2, 3, 5, 7, 11 up to 2^26", M);
    const long long nhi = *single ? 0 : (M + RF_TB - 1) / RF_TB;
    *need = (*single ? 0 : M * nb) + *M1 + *M2 + (*single ? 0 : 2 * (RF_TB + nhi) + 1);
    return 0;
}

// The transforms of nb series of n floats each, back to back at d (device
// memory of the current device), enqueued on st with the workspace 'buf' of
// at least realfft_workspace's 'need' elements.
static int realfft_enqueue(float *d_data, long long n, int nb, int isign, hipStream_t st, float2 *buf)
{
    int M1, M2;
    bool single;
    long long need;
    int rc = realfft_workspace(n, nb, isign, &M1, &M2, &single, &need);
    if (rc) return rc;
    (void)need;
    const long long M = n / 2;
    float2 *d = (float2 *)d_data;
    const long long nhi = single ? 0 : (M + RF_TB - 1) / RF_TB;
    float2 *ws = buf, *tw1 = buf + (single ? 0 : M * nb), *tw2 = tw1 + M1;
    const long long toff = (tw2 + M2 - buf + 1) & ~1LL;  // 16-byte aligned (float2 units)
    double2 *tlo = (double2 *)(buf + toff), *thi = tlo + RF_TB;
    hipLaunchKernelGGL(k_rf_twiddles, dim3((M1 + 255) / 256), dim3(256), 0, st, tw1, M1);
    hipLaunchKernelGGL(k_rf_twiddles, dim3((M2 + 255) / 256), dim3(256), 0, st, tw2, M2);
    if (!single) {
        const long long nt = std::max<long long>(RF_TB, nhi);
        hipLaunchKernelGGL(k_rf_twiddles_step, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, st, tlo, thi, M,
                           (int)nhi);
    }
    PZL_HIPCHK(hipGetLastError());
    const int npairs = (int)((n >> 2) + 1);
    if (isign == 1) {
        hipLaunchKernelGGL(k_rf_pack_inv, dim3((npairs + 255) / 256, nb), dim3(256), 0, st, d, n);
        PZL_HIPCHK(hipGetLastError());
    }
    const double scale = isign == 1 ? 2.0 / (double)n : 1.0;
    if (single) {
        const size_t lds = sizeof(float2) * (size_t)M;
        if (isign == -1)
            hipLaunchKernelGGL(k_rf_single<-1>, dim3(1, nb), dim3(RF_NT), lds, st, d, (int)M, tw1, scale);
        else
            hipLaunchKernelGGL(k_rf_single<1>, dim3(1, nb), dim3(RF_NT), lds, st, d, (int)M, tw1, scale);
    } else {
        const int C = per_block(M2, M1), C2 = per_block(M1, M2);
        const size_t lds1 = sizeof(float2) * (size_t)C * (M1 + 1), lds2 = sizeof(float2) * (size_t)C2 * (M2 + 1);
        if (isign == -1) {
            PZL_HIPCHK(launch_cols<-1>(M1, dim3(M2 / C, nb), lds1, st, d, ws, M1, M2, C, tw1, tlo, thi));
            PZL_HIPCHK(launch_rows<-1>(M2, dim3(M1 / C2, nb), lds2, st, ws, d, M1, M2, C2, tw2, scale));
        } else {
            PZL_HIPCHK(launch_cols<1>(M1, dim3(M2 / C, nb), lds1, st, d, ws, M1, M2, C, tw1, tlo, thi));
            PZL_HIPCHK(launch_rows<1>(M2, dim3(M1 / C2, nb), lds2, st, ws, d, M1, M2, C2, tw2, scale));
        }
    }
    PZL_HIPCHK(hipGetLastError());
    if (isign == -1) {
        hipLaunchKernelGGL(k_rf_unpack_fwd, dim3((npairs + 255) / 256, nb), dim3(256), 0, st, d, n);
        PZL_HIPCHK(hipGetLastError());
    }
    return 0;
}

// The same with the device's cached workspace, grown with hipMalloc as needed.
static int realfft_on_device(float *d_data, long long n, int nb, int isign, hipStream_t st, RfWork *w)
{
    int M1, M2;
    bool single;
    long long need;
    int rc = realfft_workspace(n, nb, isign, &M1, &M2, &single, &need);
    if (rc) return rc;
    if (need > w->cap) {
        (void)hipFree(w->buf);
        w->buf = nullptr;
        w->cap = 0;
        PZL_HIPCHK(hipMalloc(&w->buf, sizeof(float2) * need));
        w->cap = need;
    }
    return realfft_enqueue(d_data, n, nb, isign, st, w->buf);
}

// Workspace of pzl_realfft_device_noalloc: a file-scope device array (each GPU
// the library runs on holds its own copy until the process ends), reached
// from the host only through the address hipGetSymbolAddress gives on that
// device.  2^21 + 2^16 float2 (16.5 MiB): one series of up to 2^22 samples,
// or n * nseries up to 2^22 in general.
#define PZL_RF_SCRATCH_ELEMS ((1LL << 21) + (1LL << 16))
static __device__ __align__(256) float2 scratch[PZL_RF_SCRATCH_ELEMS];

namespace {
std::mutex g_rf_scratch_mu;
std::vector<float2 *> g_rf_scratch;  // per device: the address of 'scratch' there

// The device address of 'scratch' on 'device' (which must be current), looked
// up on the first call for that device.
int scratch_on(int device, float2 **out)
{
    std::lock_guard<std::mutex> lk(g_rf_scratch_mu);
    if ((int)g_rf_scratch.size() <= device) g_rf_scratch.resize(device + 1, nullptr);
    if (!g_rf_scratch[device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(PZL_ERR_HIP, "realfft: the scratch array of device %d has no address: %s", device,
                        hipGetErrorString(e));
        g_rf_scratch[device] = (float2 *)ptr;
    }
    *out = g_rf_scratch[device];
    return 0;
}
}  // namespace

extern "C" {

int pzl_realfft_device_noalloc(int device, void *d_data, long long n, int nseries, int isign, void *stream)
{
    if (!d_data) return fail(PZL_ERR_ARG, "realfft: null data");
    int M1, M2;
    bool single;
    long long need;
    int rc = realfft_workspace(n, nseries, isign, &M1, &M2, &single, &need);
    if (rc) return rc;
    if (need > PZL_RF_SCRATCH_ELEMS)
        return fail(PZL_ERR_ARG,
                    "realfft: %d series of %lld samples need %lld bytes of workspace; the allocation-free entry "
                    "point has %lld (n * nseries up to 2^22)",
                    nseries, n, (long long)sizeof(float2) * need, (long long)sizeof(float2) * PZL_RF_SCRATCH_ELEMS);
    PZL_HIPCHK(hipSetDevice(device));
    float2 *ws = nullptr;
    rc = scratch_on(device, &ws);
    if (rc) return rc;
    return realfft_enqueue((float *)d_data, n, nseries, isign, (hipStream_t)stream, ws);
}

int pzl_realfft_device(int device, void *d_data, long long n, int isign)
{
    return pzl_realfft_device_batch(device, d_data, n, 1, isign);
}

int pzl_realfft_device_batch(int device, void *d_data, long long n, int nseries, int isign)
{
    if (!d_data) return fail(PZL_ERR_ARG, "realfft: null data");
    PZL_HIPCHK(hipSetDevice(device));
    RfWork *w = work_for(device);
    std::lock_guard<std::mutex> lk(w->mu);
    int rc = realfft_on_device((float *)d_data, n, nseries, isign, nullptr, w);
    if (rc) return rc;
    PZL_HIPCHK(hipStreamSynchronize(nullptr));
    return 0;
}

int pzl_realfft(float *idata, long long n, int isign)
{
    if (!idata) return fail(PZL_ERR_ARG, "realfft: null data");
    if (n < 2 || n % 2) return fail(PZL_ERR_ARG, "realfft: n = %lld must be even and >= 2", n);
    int device = 0;
    PZL_HIPCHK(hipGetDevice(&device));
    RfWork *w = work_for(device);
    std::lock_guard<std::mutex> lk(w->mu);
    float *d = nullptr;
    PZL_HIPCHK(hipMalloc(&d, sizeof(float) * n));
    hipError_t e = hipMemcpy(d, idata, sizeof(float) * n, hipMemcpyHostToDevice);
    int rc = 0;
    if (e == hipSuccess) rc = realfft_on_device(d, n, 1, isign, nullptr, w);
    if (e == hipSuccess && rc == 0) e = hipMemcpy(idata, d, sizeof(float) * n, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (rc) return rc;
    if (e != hipSuccess) return fail(PZL_ERR_HIP, "realfft: %s", hipGetErrorString(e));
    return 0;
}

}  // extern "C"
