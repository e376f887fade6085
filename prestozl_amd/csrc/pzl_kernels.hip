// pzl_kernels.hip -- the jerk-search hot path as gfx950 kernels.
//
//   k_bank_build   z/w response-kernel bank, built once per search shape and
//                  kept in HBM (accel_utils.c:151-168, responses.c:159-418,
//                  corr_prep.c:92-114)
//   k_segments     per (rstep, harmonic): gather, block-median normalise,
//                  spread + forward FFT; subharmonic r-index table
//                  (accel_utils.c:1104-1220, :1058-1092, median.c:12-69)
//   k_corr<N>      per (kernel row, rstep): product with conj(kernel),
//                  inverse FFT of length N in LDS, |.|^2 / N^2, plane store
//                  (accel_utils.c:1222-1289; accel_kernels.cu:930-988)
//   k_harmsum      harmonic sum in the reference's order, threshold and
//                  wave-aggregated compaction of hits
//                  (accel_utils.c:1356-1377, :1507-1555; accel_kernels.cu:781-862)
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <mutex>


#include "pzl_fft.h"
#include "pzl_kernels.h"
#include "pzl_launch.h"
#include "pzl_math.h"
#include "pzl_response.h"
#include "pzl_stats.h"

// ---------------------------------------------------------------------------
// Kernel bank
// ---------------------------------------------------------------------------

#define BANK_NT 256
#define BANK_MAXV 8  // response points per thread: numkern <= 2048

__device__ static int find_harm(const PzlHarm *harms, int nh, int row_global, int *row)
{
    int acc = 0;
    for (int h = 0; h < nh; h++) {
        if (row_global < acc + harms[h].nrows) {
            *row = row_global - acc;
            return h;
        }
        acc += harms[h].nrows;
    }
    *row = 0;
    return -1;
}

// One workgroup per (harmonic, z, w) kernel.
__global__ __launch_bounds__(BANK_NT) void k_bank_build(const PzlHarm *__restrict__ harms, int nh,
                                                        float2 *__restrict__ bank)
{
    extern __shared__ double2 sbuf[];  // fftlen complex doubles
    const int tid = threadIdx.x;
    int row;
    const int h = find_harm(harms, nh, blockIdx.x, &row);
    if (h < 0) return;
    const PzlHarm H = harms[h];
    const int wi = row / H.numz, zi = row % H.numz;
    const double z = (double)(-H.zmax + zi * PZL_DZ);
    const double w = (double)(-H.wmax + wi * PZL_DW);
    const int khw = pzl_w_resp_halfwidth(z, w);
    const int numkern = 2 * PZL_NUMBETWEEN * khw;
    const int half = numkern / 2;
    const int fftlen = H.fftlen;
    double2 resp[BANK_MAXV];
    pzl_gen_response<BANK_NT, BANK_MAXV>(0.0, PZL_NUMBETWEEN, z, w, numkern, tid, (float *)sbuf, resp);
    // place_complex_kernel (corr_prep.c:92-114) then forward FFT
    for (int k = tid; k < fftlen; k += BANK_NT) sbuf[k] = double2{0.0, 0.0};
    __syncthreads();
#pragma unroll
    for (int i = 0; i < BANK_MAXV; i++) {
        const int m = tid + i * BANK_NT;
        if (m < numkern) {
            const int pos = (m >= half) ? (m - half) : (fftlen - half + m);
            sbuf[pos] = resp[i];
        }
    }
    __syncthreads();
    lds_fft_f64<-1, BANK_NT>(sbuf, fftlen, tid);
    // For power-of-two lengths the bank holds K / fftlen: the correlation's
    // 1 / fftlen^2 power normalisation then comes out of the product, exactly
    // (a power-of-two scale commutes with every rounding on the way).
    const double bscale = (fftlen & (fftlen - 1)) == 0 ? 1.0 / fftlen : 1.0;
    float2 *out = bank + H.bank_off + (long long)row * fftlen;
    for (int k = tid; k < fftlen; k += BANK_NT)
        out[k] = float2{(float)(sbuf[k].x * bscale), (float)(sbuf[k].y * bscale)};
}

// ---------------------------------------------------------------------------
// Segments: gather, normalise, spread + forward FFT, r-index tables
// ---------------------------------------------------------------------------

#define SEG_NT 256
#define SEG_MAXDATA 2048  // segment points: half the longest FFT length (4096)

// The element of rank k (0-based, ascending) of n non-negative floats in LDS,
// by the SEG_NT threads of a workgroup: a radix select on their bit patterns
// (which order non-negative floats as their values), 8 bits at a time from
// the top.  Per digit: a histogram of the keys that match the digits found so
// far, a scan of it, and the digit whose bin holds rank k.  The keys are left
// as they are.
__device__ __forceinline__ float seg_select(const float *keys, int n, unsigned k, int tid)
{
    __shared__ unsigned hist[256];
    __shared__ unsigned wsum[SEG_NT / 64];
    __shared__ unsigned found[2];
    static_assert(SEG_NT == 256, "one histogram bin per thread");
    const int lane = __lane_id(), wave = tid / 64;
    unsigned prefix = 0, mask = 0;
    for (int shift = 24; shift >= 0; shift -= 8) {
        hist[tid] = 0;
        __syncthreads();
        for (int i = tid; i < n; i += SEG_NT) {
            const unsigned u = __float_as_uint(keys[i]);
            if ((u & mask) == prefix) atomicAdd(&hist[(u >> shift) & 255u], 1u);
        }
        __syncthreads();
        // inclusive scan of the 256 bins, one per thread
        const unsigned c = hist[tid];
        unsigned x = c;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const unsigned y = __shfl_up(x, d);
            if (lane >= d) x += y;
        }
        if (lane == 63) wsum[wave] = x;
        __syncthreads();
#pragma unroll
        for (int w = 0; w < SEG_NT / 64; w++) x += w < wave ? wsum[w] : 0u;
        if (x - c <= k && k < x) {  // exactly one bin holds rank k
            found[0] = (unsigned)tid;
            found[1] = k - (x - c);
        }
        __syncthreads();
        prefix |= found[0] << shift;
        mask |= 255u << shift;
        k = found[1];
        __syncthreads();
    }
    return __uint_as_float(prefix);
}

__global__ __launch_bounds__(SEG_NT) void k_segments(PzlBatchArgs a)
{
    extern __shared__ double2 sbuf[];  // fftlen/2 complex doubles, then keys
    const int tid = threadIdx.x;
    const int b = blockIdx.x / a.nh, h = blockIdx.x % a.nh;
    const PzlHarm H = a.harms[h];
    const long long rstep = (a.rstep0_dev ? *a.rstep0_dev : a.rstep0) + b;
    const double fullrlo = a.startr[rstep];
    const double fullrhi = fullrlo + a.rstep_bins - PZL_DR;
    const PzlHarmR hr = pzl_harm_r(H.hf, h == 0, fullrlo, fullrhi, a.corr_uselen);
    const long long lobin = hr.rlo - H.halfwidth;
    const int numdata = H.fftlen / PZL_NUMBETWEEN;
    int n2 = 1;
    while (n2 < numdata) n2 <<= 1;
    float *keys = (float *)(sbuf + numdata);

    if (tid == 0) {
        a.hr_tab[b * a.nh + h].rlo = hr.rlo;
        a.hr_tab[b * a.nh + h].numrs = hr.numrs;
    }
    // get_fourier_amplitudes: zero outside [obs lobin, numbins)
    for (int i = tid; i < n2; i += SEG_NT) {
        float2 x = float2{0.f, 0.f};
        if (i < numdata) {
            const long long idx = lobin - a.file_lobin + i;
            if (idx >= 0 && idx < a.numbins) x = a.fft[idx];
            sbuf[i] = double2{(double)x.x, (double)x.y};
            keys[i] = __fadd_rn(__fmul_rn(x.x, x.x), __fmul_rn(x.y, x.y));
        } else {
            keys[i] = __int_as_float(0x7f800000);  // +inf padding
        }
    }
    __syncthreads();
    double norm;
    if (a.nph > 0.0) {
        norm = 1.0 / sqrt(a.nph);
    } else {
        // the lower median: element (n-1)/2 of the powers in ascending order
        norm = 1.0 / sqrt((double)seg_select(keys, n2, (unsigned)(numdata - 1) / 2, tid) / log(2.0));
    }
    __syncthreads();
    for (int i = tid; i < numdata; i += SEG_NT) {
        const double2 x = sbuf[i];
        sbuf[i] = double2{(double)(float)(x.x * norm), (double)(float)(x.y * norm)};
    }
    __syncthreads();
    // FFT of the fftlen-point spread array == numdata-point FFT, repeated
    // (numdata = fftlen / 2 <= 2048)
    lds_fft_f64<-1, SEG_NT, SEG_MAXDATA>(sbuf, numdata, tid);
    float2 *seg = a.segs + (long long)b * a.seg_stride + H.seg_off;
    for (int k = tid; k < numdata; k += SEG_NT) seg[k] = float2{(float)sbuf[k].x, (float)sbuf[k].y};
    if (h > 0) {
        uint16_t *rinds = a.rinds + ((long long)b * a.nh + h) * a.corr_uselen;
        for (int ii = tid; ii < a.corr_uselen; ii += SEG_NT)
            rinds[ii] = (uint16_t)pzl_rind(H.hf, fullrlo, hr.rlo, ii);
    }
}

// ---------------------------------------------------------------------------
// Correlation: product, inverse FFT, power, plane store
// ---------------------------------------------------------------------------
//
// One workgroup takes G kernel rows of one harmonic and runs them against the
// segments of every rstep of the batch: the rows are read from HBM once per
// batch and stay in registers.  Per (row, rstep) the inverse FFT is a Stockham
// sequence of in-register radix-R DFTs with LDS exchanges between passes
// (ping-pong buffers: one barrier per exchange).  Pass 0 forms the product
// with conj(kernel) straight from the segment; the last pass turns the
// natural-order outputs into powers and stores the wanted r window.
// Twiddles depend only on the thread and the pass, so each workgroup computes
// them once (double-precision sincospi) and keeps them in registers for all of
// its (row, rstep) pairs (the grid is persistent).

// LDS index padding: one element after every 2^S (breaks bank conflicts of
// strided exchanges)
template <int S> __host__ __device__ constexpr int lds_pad(int p) { return p + (p >> S); }

// Word 3 of a raw (stride 0) buffer resource on gfx9-family parts: 32-bit
// data format; with stride 0, an access at byte offset >= num_records is
// dropped (stores) or returns 0 (loads).
#define PZL_BUF_DWORD3 0x00020000

// A buffer resource built from wave-uniform values (read from the first
// lane, so that the compiler keeps it in scalar registers).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t uniform_rsrc(const void *p, int nbytes)
{
    const unsigned long long addr = (unsigned long long)p;
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)addr);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(addr >> 32));
    const int n = __builtin_amdgcn_readfirstlane(nbytes);
    return __builtin_amdgcn_make_buffer_rsrc((void *)(((unsigned long long)hi << 32) | lo), 0, n, PZL_BUF_DWORD3);
}

template <int N, int NT, int G, bool PING, int PADSH, int... Rs> struct CorrPlan {
    static constexpr int NP = sizeof...(Rs);
    static constexpr int RAD[NP] = {Rs...};
    static constexpr int ns(int p)
    {
        int s = 1;
        for (int i = 0; i < p; i++) s *= RAD[i];
        return s;
    }
    static constexpr int nbf(int p) { return G * N / RAD[p]; }     // butterflies of pass p
    static constexpr int per(int p) { return (nbf(p) + NT - 1) / NT; }  // ... per thread
    static constexpr bool full(int p) { return nbf(p) % NT == 0; }
    static constexpr int twoff(int p)
    {
        int o = 0;
        for (int i = 1; i < p; i++) o += per(i) * (RAD[i] - 1);
        return o;
    }
    static constexpr int TWN = twoff(NP);
    static constexpr int PPT = per(0) * RAD[0];  // kernel values held per thread
    static constexpr int BUF = lds_pad<PADSH>(G * N) + 1;
    static constexpr size_t smem() { return (size_t)(PING ? 2 : 1) * BUF * sizeof(float2); }
    static_assert(ns(NP) == N, "radices do not multiply to N");
};

// N: FFT length, G: FFTs per workgroup.  SH = 1: each kernel row is done as
// two N-point FFTs (its even and odd outputs), FFT g holding row g >> 1,
// output parity g & 1.
template <int N, int NT, int G, bool PING, int SH, int PADSH, int... Rs> struct CorrFFT {
    using CP = CorrPlan<N, NT, G, PING, PADSH, Rs...>;
    static constexpr int NP = CP::NP;

    // butterfly bf of pass P: FFT g, index j within that FFT's N / R butterflies
    template <int P> static __device__ __forceinline__ void where(int bf, int &g, int &j)
    {
        constexpr int NBP = N / CP::RAD[P];
        g = bf / NBP;
        j = bf - g * NBP;
    }

    template <int P> static __device__ __forceinline__ void init_tw(float2 *tw, int tid)
    {
        if constexpr (P < NP) {
            constexpr int R = CP::RAD[P], Ns = CP::ns(P);
#pragma unroll
            for (int i = 0; i < CP::per(P); i++) {
                int g, j;
                where<P>(tid + i * NT, g, j);
                const int m = j % Ns;
#pragma unroll
                for (int r = 1; r < R; r++) {
                    double s, c;
                    sincospi(2.0 * (double)(m * r) / (double)(Ns * R), &s, &c);
                    tw[CP::twoff(P) + i * (R - 1) + (r - 1)] = float2{(float)c, (float)s};
                }
            }
            init_tw<P + 1>(tw, tid);
        }
    }

    // passes 1 .. NP-1 of one FFT set; 'src' holds the output of pass P-1
    template <int P>
    static __device__ __forceinline__ void pass(float2 *b0, float2 *b1, const float2 *tw, int tid, float *pbase,
                                                long long rs_stride, int gvalid, int numrs, int offset, float norm)
    {
        constexpr int R = CP::RAD[P], Ns = CP::ns(P), NBP = N / R;
        constexpr bool last = (P == NP - 1);
        const float2 *src = ((P - 1) & 1) && PING ? b1 : b0;
        float2 *dst = (P & 1) && PING ? b1 : b0;
        float2 v[CP::per(P)][R];
#pragma unroll
        for (int i = 0; i < CP::per(P); i++) {
            const int bf = tid + i * NT;
            if (CP::full(P) || bf < CP::nbf(P)) {
                int g, j;
                where<P>(bf, g, j);
                // stride NBP: with NBP a multiple of 16 the padded index is linear in r
                constexpr int PM = 1 << PADSH;
                const float2 *rp = src + lds_pad<PADSH>(g * N + j);
#pragma unroll
                for (int r = 0; r < R; r++) {
                    float2 x = (NBP % PM == 0) ? rp[r * lds_pad<PADSH>(NBP)]
                                               : src[lds_pad<PADSH>(g * N + j + r * NBP)];
                    if (r > 0) x = cmul(x, tw[CP::twoff(P) + i * (R - 1) + (r - 1)]);
                    v[i][r] = x;
                }
            }
        }
        if constexpr (!PING && !last) __syncthreads();
#pragma unroll
        for (int i = 0; i < CP::per(P); i++) {
            const int bf = tid + i * NT;
            if (CP::full(P) || bf < CP::nbf(P)) {
                int g, j;
                where<P>(bf, g, j);
                dftR<R, 1>(v[i]);
                if constexpr (!last) {
                    const int idxD = (j / Ns) * Ns * R + (j % Ns);
                    float2 *wp = dst + lds_pad<PADSH>(g * N + idxD);
#pragma unroll
                    for (int r = 0; r < R; r++) {
                        if constexpr (Ns % (1 << PADSH) == 0) wp[r * lds_pad<PADSH>(Ns)] = v[i][r];
                        else dst[lds_pad<PADSH>(g * N + idxD + r * Ns)] = v[i][r];
                    }
                } else if constexpr (NBP % 64 == 0) {
                    // g is uniform in a wave: a buffer store whose range is the
                    // row's numrs powers drops the outputs outside the r window
                    // (negative offsets wrap to out of range) without branches
                    const int gu = __builtin_amdgcn_readfirstlane(g);
                    const int row = gu >> SH, par = gu & ((1 << SH) - 1);
                    const __amdgpu_buffer_rsrc_t rs =
                        uniform_rsrc(pbase + row * rs_stride, row < gvalid ? numrs * 4 : 0);
#pragma unroll
                    for (int r = 0; r < R; r++) {
                        // natural-order output index (of the row) minus offset
                        const int ro = ((j + r * NBP) << SH) + par - offset;
                        const float2 y = v[i][r];
                        __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint((y.x * y.x + y.y * y.y) * norm), rs,
                                                              ro * 4, 0, 0);
                    }
                } else {
                    const int row = g >> SH, par = g & ((1 << SH) - 1);
                    if (row < gvalid) {
                        float *pl = pbase + row * rs_stride;
#pragma unroll
                        for (int r = 0; r < R; r++) {
                            const int ro = ((j + r * NBP) << SH) + par - offset;
                            if (ro >= 0 && ro < numrs) {
                                const float2 y = v[i][r];
                                pl[ro] = (y.x * y.x + y.y * y.y) * norm;
                            }
                        }
                    }
                }
            }
        }
        if constexpr (!last) {
            __syncthreads();
            pass<P + 1>(b0, b1, tw, tid, pbase, rs_stride, gvalid, numrs, offset, norm);
        }
    }
};

// The twiddles of every thread of a correlation variant, [TWN][NT]: computed
// once per device in double precision and read by each workgroup of k_corr.
template <int N, int NT, int G, bool PING, int SH, int PADSH, int... Rs>
__global__ __launch_bounds__(NT) void k_corr_twiddles(float2 *table)
{
    using F = CorrFFT<(N >> SH), NT, (G << SH), PING, SH, PADSH, Rs...>;
    using CP = typename F::CP;
    float2 tw[CP::TWN > 0 ? CP::TWN : 1];
    F::template init_tw<1>(tw, threadIdx.x);
#pragma unroll
    for (int t = 0; t < CP::TWN; t++) table[t * NT + threadIdx.x] = tw[t];
}

// exp(-2 pi i m / R) for m < R, as compile-time constants
template <int R> struct ExpTab {
    float c[R], s[R];
    constexpr ExpTab() : c(), s()
    {
        for (int m = 0; m < R; m++) {
            c[m] = (float)ct_cos2pi(-m, R);
            s[m] = (float)ct_sin2pi(-m, R);
        }
    }
};

// N: kernel-row (correlation) length, G: kernel rows per workgroup, SH: 1 to
// transform each row as two N/2-point FFTs.  Because the segment spectrum
// repeats with period N/2, y[2m + e] = sum_{k<N/2} P_e[k] exp(2 pi i k m/(N/2))
// with P_0 = pdata (K*[k] + K*[k+N/2]) and P_1 = pdata (K*[k] - K*[k+N/2])
// exp(2 pi i k/N): the first radix-2 stage of the N-point transform done on the
// kernel row once per batch instead of per rstep.
thread_local std::vector<PzlKernelRec> *pzl_recorder = nullptr;

// Work units u0, u0 + ustep, ... (< a.total_items) of one correlation
// variant, by the NT threads of a workgroup; smem holds CP::smem() bytes.
template <int N, int NT, int G, bool PING, int SH, int PADSH, int... Rs>
__device__ __forceinline__ void corr_units(const PzlCorrArgs &a, long long u0, long long ustep, float2 *smem)
{
    constexpr int NF = N >> SH;  // FFT length
    constexpr int GF = G << SH;  // FFTs per workgroup
    using F = CorrFFT<NF, NT, GF, PING, SH, PADSH, Rs...>;
    using CP = typename F::CP;
    constexpr int R0 = CP::RAD[0], NB0 = NF / R0;
    float2 *b0 = smem, *b1 = PING ? smem + CP::BUF : smem;
    const int tid = threadIdx.x;
    // power normalisation 1 / N^2 (accel_utils.c:1234, 1278-1282); for power-of-two N
    // the bank rows carry 1 / N each, so the powers need no further scaling
    constexpr bool POW2 = (N & (N - 1)) == 0;
    const float norm = POW2 ? 1.0f : (float)(1.0 / ((double)N * (double)N));

    // twiddles of this thread, from the variant's table ([TWN][NT], see k_corr_twiddles)
    float2 tw[CP::TWN > 0 ? CP::TWN : 1];
#pragma unroll
    for (int t = 0; t < CP::TWN; t++) tw[t] = a.twtab[t * NT + tid];
    // SH = 1: exp(-2 pi i j / N) of each pass-0 butterfly (the odd half's twist)
    float2 wj[CP::per(0)];
    if constexpr (SH == 1) {
#pragma unroll
        for (int i = 0; i < CP::per(0); i++) {
            int g, j;
            F::template where<0>(tid + i * NT, g, j);
            double sn, cs;
            sincospi(-2.0 * (double)j / (double)N, &sn, &cs);
            wj[i] = float2{(float)cs, (float)sn};
        }
    }

    for (long long unit = u0; unit < a.total_items; unit += ustep) {
        // the unit's rows (per harmonic of the class, ceil(nrows / G) row groups) and
        // its block of rsteps: the blocks of one row group are adjacent units
        const int ru = (int)(unit / a.nblk), blk = (int)(unit - (long long)ru * a.nblk);
        int h = -1, row0 = 0;
        {
            int acc = 0;
            for (int c = 0; c < a.nclass; c++) {
                const int hh = a.class_harms[c];
                const int cnt = (a.harms[hh].nrows + G - 1) / G;
                if (ru < acc + cnt) {
                    h = hh;
                    row0 = (ru - acc) * G;
                    break;
                }
                acc += cnt;
            }
        }
        const int b_lo = __builtin_amdgcn_readfirstlane(blk * a.rb);
        const int b_hi = __builtin_amdgcn_readfirstlane(min(a.nb, b_lo + a.rb));
        h = __builtin_amdgcn_readfirstlane(h);
        row0 = __builtin_amdgcn_readfirstlane(row0);
        const PzlHarm &H = a.harms[h];
        const int nrows = __builtin_amdgcn_readfirstlane(H.nrows);
        // the kernel rows (conjugated in the product below), at the positions pass 0 reads them
        float2 kv[CP::per(0)][R0];
#pragma unroll
        for (int i = 0; i < CP::per(0); i++) {
            const int bf = tid + i * NT;
            int g, j;
            F::template where<0>(bf, g, j);
            if constexpr (NB0 % 64 == 0 && CP::full(0)) {
                // g is uniform in a wave: rows past the harmonic read as zeros
                const int gu = __builtin_amdgcn_readfirstlane(g);
                const int row = gu >> SH;
                const __amdgpu_buffer_rsrc_t rs =
                    uniform_rsrc(a.bank + H.bank_off + (long long)(row0 + row) * N, row0 + row < nrows ? N * 8 : 0);
                auto ld = [&](int k) {
                    return float2{__uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, k * 8, 0, 0)),
                                  __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, k * 8 + 4, 0, 0))};
                };
                if constexpr (SH == 0) {
#pragma unroll
                    for (int r = 0; r < R0; r++) kv[i][r] = ld(j + r * NB0);
                } else {
                    const int par = gu & 1;
                    constexpr ExpTab<2 * R0> et{};  // exp(-2 pi i r / (2 R0))
#pragma unroll
                    for (int r = 0; r < R0; r++) {
                        const int k = j + r * NB0;
                        const float2 k1 = ld(k), k2 = ld(k + NF);
                        if (par == 0) {
                            kv[i][r] = cadd(k1, k2);
                        } else {
                            // (K[k] - K[k+N/2]) exp(-2 pi i k/N), k = j + r N/(2 R0)
                            kv[i][r] = cmul(cmul(csub(k1, k2), wj[i]), float2{et.c[r], et.s[r]});
                        }
                    }
                }
            } else {
                static_assert(SH == 0, "split rows need wave-uniform rows in pass 0");
                const bool ok = (CP::full(0) || bf < CP::nbf(0)) && row0 + g < nrows;
                const float2 *kr = a.bank + H.bank_off + (long long)(row0 + g) * N;
#pragma unroll
                for (int r = 0; r < R0; r++) kv[i][r] = ok ? kr[j + r * NB0] : float2{0.f, 0.f};
            }
        }
        if constexpr (SH == 0 && R0 % 2 == 0) {
            // The two inputs r and r + R0/2 of a pass-0 butterfly share their
            // segment value (the spread segment has period N/2), so the first
            // radix-2 stage of the butterfly's DFT is done on the kernel row,
            // once per unit: K[r] + K[r+R0/2], and (K[r] - K[r+R0/2]) times
            // exp(-2 pi i r / R0) (the conjugate of that stage's twiddle, the
            // product below conjugates the kernel value).
            constexpr int HR = R0 / 2;
            constexpr ExpTab<R0> et{};
#pragma unroll
            for (int i = 0; i < CP::per(0); i++)
#pragma unroll
                for (int r = 0; r < HR; r++) {
                    const float2 k1 = kv[i][r], k2 = kv[i][r + HR];
                    kv[i][r] = cadd(k1, k2);
                    kv[i][r + HR] = cmul(csub(k1, k2), float2{et.c[r], et.s[r]});
                }
        }
        const int offset = 2 * H.halfwidth;
        // segment values of the next rstep are loaded while this one is transformed;
        // without the split, the spread segment's period N/2 lets an even radix
        // read each value once
        constexpr bool HALFPER = SH == 0 && R0 % 2 == 0;
        constexpr int NSV = HALFPER ? R0 / 2 : R0;
        float2 sv[CP::per(0)][NSV];
        auto load_seg = [&](int b) {
            const float2 *seg = a.segs + (long long)b * a.seg_stride + H.seg_off;
#pragma unroll
            for (int i = 0; i < CP::per(0); i++) {
                const int bf = tid + i * NT;
                int g, j;
                F::template where<0>(bf, g, j);
                if (CP::full(0) || bf < CP::nbf(0)) {
#pragma unroll
                    for (int r = 0; r < NSV; r++) {
                        const int k = j + r * NB0;
                        sv[i][r] = seg[k < N / 2 ? k : k - N / 2];
                    }
                }
            }
        };
        load_seg(b_lo);
        for (int b = b_lo; b < b_hi; b++) {
            const int numrs = __builtin_amdgcn_readfirstlane(a.hr_tab[b * a.nh + h].numrs);
            float *pbase = a.planes + (long long)b * a.plane_stride + H.plane_off + (long long)row0 * H.rs_stride;
            float2 cur[CP::per(0)][NSV];
#pragma unroll
            for (int i = 0; i < CP::per(0); i++)
#pragma unroll
                for (int r = 0; r < NSV; r++) cur[i][r] = sv[i][r];
            if (b + 1 < b_hi) load_seg(b + 1);
            // pass 0: P[k] = seg[k mod N/2] * conj(kern[k])
#pragma unroll
            for (int i = 0; i < CP::per(0); i++) {
                const int bf = tid + i * NT;
                if (CP::full(0) || bf < CP::nbf(0)) {
                    int g, j;
                    F::template where<0>(bf, g, j);
                    float2 v[R0];
                    if constexpr (HALFPER) {
                        // outputs 2m and 2m+1 of the R0-point DFT: the R0/2-point
                        // DFTs of seg * conj(the combined kernel values)
                        constexpr int HR = R0 / 2;
                        float2 ev[HR], od[HR];
#pragma unroll
                        for (int r = 0; r < HR; r++) {
                            const float2 pd = cur[i][r];
                            const float2 ka = kv[i][r], kb = kv[i][r + HR];
                            // re = dr*kr + di*ki, im = di*kr - dr*ki (accel_utils.c:1262-1265)
                            ev[r] = float2{pd.x * ka.x + pd.y * ka.y, pd.y * ka.x - pd.x * ka.y};
                            od[r] = float2{pd.x * kb.x + pd.y * kb.y, pd.y * kb.x - pd.x * kb.y};
                        }
                        dftR<HR, 1>(ev);
                        dftR<HR, 1>(od);
#pragma unroll
                        for (int m = 0; m < HR; m++) {
                            v[2 * m] = ev[m];
                            v[2 * m + 1] = od[m];
                        }
                    } else {
#pragma unroll
                        for (int r = 0; r < R0; r++) {
                            const float2 pd = cur[i][r];
                            const float2 kk = kv[i][r];
                            // re = dr*kr + di*ki, im = di*kr - dr*ki (accel_utils.c:1262-1265)
                            v[r] = float2{pd.x * kk.x + pd.y * kk.y, pd.y * kk.x - pd.x * kk.y};
                        }
                        dftR<R0, 1>(v);
                    }
#pragma unroll
                    for (int r = 0; r < R0; r++) b0[lds_pad<PADSH>(g * NF + j * R0 + r)] = v[r];
                }
            }
            __syncthreads();
            F::template pass<1>(b0, b1, tw, tid, pbase, H.rs_stride, nrows - row0, numrs, offset, norm);
            // the next pass 0 overwrites b0: wait for this FFT's readers of b0
            if constexpr (!PING || (CP::NP % 2 == 0)) __syncthreads();
        }
    }
}

template <int N, int NT, int G, bool PING, int SH, int PADSH, int... Rs>
__global__ __launch_bounds__(NT) void k_corr(PzlCorrArgs a)
{
    extern __shared__ float2 smem[];
    corr_units<N, NT, G, PING, SH, PADSH, Rs...>(a, blockIdx.x, gridDim.x, smem);
}

// One correlation variant as a type: its launch geometry, twiddle table and
// work units, for launches of one variant and for grouped launches.
template <int N_, int NT_, int G_, bool PING, int SH, int PADSH, int... Rs> struct CorrV {
    static constexpr int N = N_, NT = NT_, G = G_;
    using CP = CorrPlan<(N >> SH), NT, (G << SH), PING, PADSH, Rs...>;
    static constexpr size_t smem() { return CP::smem(); }
    static __device__ __forceinline__ void unit(const PzlCorrArgs &a, long long u, float2 *sm)
    {
        corr_units<N, NT, G, PING, SH, PADSH, Rs...>(a, u, a.total_items, sm);
    }
    // The variant's twiddles ([TWN][NT], see k_corr_twiddles), built on the
    // current device on first use (never inside a stream capture: the first
    // batch of every plan runs eagerly) and kept until the process ends.
    static hipError_t table(hipStream_t st, const float2 **out)
    {
        static std::mutex mu;
        static float2 *tables[64] = {};
        int dev = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e != hipSuccess) return e;
        if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
        std::lock_guard<std::mutex> lk(mu);
        if (!tables[dev] && pzl_recorder) return hipErrorNotReady;  // built by an enqueued launch first
        if (!tables[dev]) {
            float2 *t = nullptr;
            e = hipMalloc(&t, sizeof(float2) * (size_t)(CP::TWN > 0 ? CP::TWN : 1) * NT);
            if (e != hipSuccess) return e;
            hipLaunchKernelGGL((k_corr_twiddles<N, NT, G, PING, SH, PADSH, Rs...>), dim3(1), dim3(NT), 0, st, t);
            e = hipGetLastError();
            if (e == hipSuccess) e = hipStreamSynchronize(st);
            if (e != hipSuccess) {
                (void)hipFree(t);
                return e;
            }
            tables[dev] = t;
        }
        *out = tables[dev];
        return hipSuccess;
    }
};

// Waves a correlation launch aims for at least (two rounds of 4 waves per SIMD).
constexpr long long kCorrWaves = 2 * 4 * 1024;

// Work units of one class for a variant of NT threads and G rows per
// workgroup: a class with few row groups (the short FFT lengths) splits the
// batch's rsteps over several units, so that the launch still has about
// kCorrWaves waves; its rows are then read once per block of rsteps instead
// of once per batch.  Returns false when the units exceed 2^31 - 1.
static bool corr_units_of(const PzlCorrArgs &a, const int *class_nrows, int NT, int G, PzlCorrArgs &out)
{
    long long rowunits = 0;
    for (int c = 0; c < a.nclass; c++) rowunits += (class_nrows[c] + G - 1) / G;
    out = a;
    out.total_items = 0;
    out.rb = out.nblk = 1;
    if (rowunits < 1 || a.nb < 1) return true;
    const long long waves = rowunits * (NT / 64);
    const int nblk = (int)std::min<long long>(a.nb, std::max<long long>(1, (kCorrWaves + waves - 1) / waves));
    out.rb = (a.nb + nblk - 1) / nblk;
    out.nblk = (a.nb + out.rb - 1) / out.rb;
    out.total_items = rowunits * out.nblk;
    return out.total_items <= 0x7fffffffLL;
}

// ---------------------------------------------------------------------------
// Harmonic sum, threshold, compaction
// ---------------------------------------------------------------------------

#define SUM_ZB 32  // fundamental z rows per workgroup

// Threads per workgroup and fundamental r per workgroup row chunk, for RPT
// consecutive r per thread.
template <int RPT> struct SumShape {
    static constexpr int NT = RPT == 4 ? 128 : 256;
    static constexpr int RCH = NT * RPT;
};

// RPT consecutive floats at byte offset voff + soff of a buffer, one load.
template <int RPT>
__device__ __forceinline__ void load_run(__amdgpu_buffer_rsrc_t rs, int voff, int soff, float (&x)[RPT])
{
    if constexpr (RPT == 1) {
        x[0] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, voff, soff, 0));
    } else if constexpr (RPT == 2) {
        const auto v = __builtin_amdgcn_raw_buffer_load_b64(rs, voff, soff, 0);
        x[0] = __uint_as_float(v[0]);
        x[1] = __uint_as_float(v[1]);
    } else {
        static_assert(RPT == 4, "1, 2 or 4 r per thread");
        const auto v = __builtin_amdgcn_raw_buffer_load_b128(rs, voff, soff, 0);
        x[0] = __uint_as_float(v[0]);
        x[1] = __uint_as_float(v[1]);
        x[2] = __uint_as_float(v[2]);
        x[3] = __uint_as_float(v[3]);
    }
}

// The subharmonic index of fundamental r + i is rint((m + i) j / 2^s) for
// some real m (half-integer steps of r), so its distance from the index of r
// lies within one of i j / 2^s: it is floor(i j / 2^s) or one more, or, when
// i j / 2^s is an integer D, one of D - 1, D, D + 1.  SelRange gives that
// range (clamped to the RPT loaded values) for harmonic j / 2^s.
template <int RPT> struct SelRange {
    static constexpr int lo(int i, int j, int s)
    {
        const int num = i * j, den = 1 << s;
        const int v = num % den == 0 ? num / den - 1 : num / den;
        return v < 0 ? 0 : (v > RPT - 1 ? RPT - 1 : v);
    }
    static constexpr int hi(int i, int j, int s)
    {
        const int num = i * j, den = 1 << s;
        const int v = num % den == 0 ? num / den + 1 : num / den + 1;
        const int c = v > i ? i : v;
        return c > RPT - 1 ? RPT - 1 : c;
    }
    // subharmonic q (stage order: stage s = 1.., odd j ascending) is j / 2^s
    static constexpr int stage_of(int q)
    {
        int s = 1;
        while ((1 << s) - 1 <= q) s++;
        return s;
    }
    static constexpr int j_of(int q) { return 2 * (q - ((1 << (stage_of(q) - 1)) - 1)) + 1; }
    // selection bits of subharmonic q for its r + i' with i' < i
    static constexpr int bits_before_i(int q, int i)
    {
        int n = 0;
        for (int ii = 1; ii < i; ii++) n += hi(ii, j_of(q), stage_of(q)) - lo(ii, j_of(q), stage_of(q));
        return n;
    }
    // selection bits needed by subharmonics q' < q
    static constexpr int bits_before(int q)
    {
        int n = 0;
        for (int qq = 0; qq < q; qq++)
            for (int i = 1; i < RPT; i++) n += hi(i, j_of(qq), stage_of(qq)) - lo(i, j_of(qq), stage_of(qq));
        return n;
    }
};

// f(std::integral_constant<int, I>) for I = B .. E - 1, unrolled at compile time.
template <int B, int E, typename F> __device__ __forceinline__ void static_for(F &&f)
{
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// One workgroup: RCH consecutive r of SUM_ZB consecutive z rows of one
// (rstep, w) of the fundamental, RPT consecutive r per thread.  The
// subharmonic r indices depend only on (rstep, harmonic, r), so each thread
// computes them once (integer round-half-even of m j / 2^s split into a uniform
// 64-bit part and a small per-thread 32-bit part).  A subharmonic's ratio j/2^s
// is below 1, so the RPT indices of a thread's r are non-decreasing with steps
// of 0 or 1: they lie in the RPT floats from the first one, which one load
// fetches.  The thread keeps the first index as a byte offset, and for every
// other r one bit per index that r's map can still take beyond the lowest
// (SelRange): the value is then chosen with bitfield selects, with no
// per-lane condition registers held across the loop.  The byte offsets of
// the subharmonic rows of each fundamental (w, z) come from a per-plan table
// through scalar loads and are the scalar offset of buffer loads over the
// rstep's planes.  For every cell all subharmonic powers are loaded first and
// then added in the reference's order (stage 1..S, odd harmonics ascending).
//
// Workgroup ids are dealt round-robin over the 8 XCDs; each XCD is given a
// contiguous range of the (rstep, r chunk, z block, w) space, w fastest
// (order 0), or of (rstep, z block, w, r chunk), r chunk fastest (order 1):
// neighbouring w rows share subharmonic rows, neighbouring r chunks the cache
// lines at their edges, which then meet in one L2.
// u where bit 'pos' of 'bits' is clear, the bits of x where it is set: a
// sign-extending bitfield extract of the bit, then a bitfield select, in
// registers, so that the compiler keeps the selection in the loop instead of
// turning it into per-lane conditions held in scalar registers across it.
template <int POS> __device__ __forceinline__ unsigned pick_bit_t(unsigned u, float x, unsigned bits)
{
    unsigned m;
    asm("v_bfe_i32 %0, %1, %2, 1\n\tv_bfi_b32 %0, %0, %3, %4"
        : "=&v"(m)
        : "v"(bits), "n"(POS), "v"(__float_as_uint(x)), "v"(u));
    return m;
}

// The row-offset table is written once per plan, before any search: reading
// it through the constant address space lets the compiler use scalar loads.
typedef const __attribute__((address_space(4))) int pzl_const_int;

template <int NSTAGES, int RPT>
__global__ __launch_bounds__(SumShape<RPT>::NT) void k_harmsum(PzlSumArgs a)
{
    constexpr int NT = SumShape<RPT>::NT, RCH = SumShape<RPT>::RCH;
    constexpr int NSUB = (1 << (NSTAGES - 1)) - 1;
    constexpr int NQ = NSUB > 0 ? NSUB : 1;
    constexpr int NBITS = SelRange<RPT>::bits_before(NSUB);
    constexpr int NW = NBITS > 0 ? (NBITS + 31) / 32 : 1;
    const int tid = threadIdx.x;
    const int chunks = (a.corr_uselen + RCH - 1) / RCH;
    const int zblocks = (a.numz + SUM_ZB - 1) / SUM_ZB;
    const long long nblk = (long long)a.nb * a.numw * zblocks * chunks;
    const long long per_xcd = (nblk + 7) / 8;
    const long long L = (long long)(blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
    if (L >= nblk) return;
    int w, zb, chunk, b;
    if (a.order == 0) {
        w = (int)(L % a.numw);
        long long t = L / a.numw;
        zb = (int)(t % zblocks);
        t /= zblocks;
        chunk = (int)(t % chunks);
        b = (int)(t / chunks);
    } else {
        chunk = (int)(L % chunks);
        long long t = L / chunks;
        w = (int)(t % a.numw);
        t /= a.numw;
        zb = (int)(t % zblocks);
        b = (int)(t / zblocks);
    }
    const int z0 = zb * SUM_ZB;
    const int nz = min(SUM_ZB, a.numz - z0);
    const int rbase = chunk * RCH + tid * RPT;
    bool active[RPT];
    int rr[RPT];
#pragma unroll
    for (int i = 0; i < RPT; i++) {
        active[i] = rbase + i < a.corr_uselen;
        rr[i] = min(rbase + i, a.corr_uselen - 1);
    }
    const long long rstep = (a.rstep0_dev ? *a.rstep0_dev : a.rstep0) + b;
    const PzlHarm F = a.harms[0];
    // every plane of this rstep (and the slack behind the set's planes), as
    // one buffer (byte offsets < 2^31)
    const __amdgpu_buffer_rsrc_t planes = uniform_rsrc(a.planes + (long long)b * a.plane_stride,
                                                       (int)(4 * (a.plane_stride + PZL_PLANE_SLACK)));

    int rind[NQ];        // byte offset of the first subharmonic index within its row
    unsigned sel[NW];    // selection bits (SelRange), harmonic by harmonic
#pragma unroll
    for (int k = 0; k < NW; k++) sel[k] = 0;
    if constexpr (NSUB > 0) {
        const long long m0 = a.int_rlo ? (long long)(2.0 * a.startr[rstep]) : 0;
        static_for<0, NSUB>([&](auto QC) {
            constexpr int q = decltype(QC)::value;
            constexpr int s = SelRange<RPT>::stage_of(q), j = SelRange<RPT>::j_of(q);
            const int h = q + 1;
            int sidx[RPT];
            if (a.int_rlo) {
                // rint((m0 + r) j / 2^s) - 2 rlo_h with m0 j split as q0 2^s + rem0
                const long long p0 = m0 * j;
                const long long q0 = p0 >> s;
                const int rem0 = (int)(p0 - (q0 << s));
                const int base_h = (int)(q0 - 2 * a.hr_tab[b * a.nh + h].rlo);
                constexpr int half = (1 << s) >> 1;
#pragma unroll
                for (int i = 0; i < RPT; i++) {
                    const int tt = rem0 + rr[i] * j;
                    int qq = tt >> s;
                    const int rem = tt - (qq << s);
                    // round half to even on the full quotient q0 + qq
                    if (rem > half || (rem == half && ((q0 + qq) & 1))) qq++;
                    sidx[i] = base_h + qq;
                }
            } else {
#pragma unroll
                for (int i = 0; i < RPT; i++)
                    sidx[i] = a.rinds[((long long)b * a.nh + h) * a.corr_uselen + rr[i]];
            }
            rind[q] = 4 * sidx[0];
            static_for<1, RPT>([&](auto IC) {
                constexpr int i = decltype(IC)::value;
                constexpr int lo = SelRange<RPT>::lo(i, j, s), hi = SelRange<RPT>::hi(i, j, s);
                constexpr int b0 = SelRange<RPT>::bits_before(q) + SelRange<RPT>::bits_before_i(q, i);
                const int d = sidx[i] - sidx[0];
                static_for<lo + 1, hi + 1>([&](auto CC) {
                    constexpr int c = decltype(CC)::value;
                    constexpr int bit = b0 + (c - lo - 1);
                    sel[bit >> 5] |= (unsigned)(d >= c) << (bit & 31);
                });
            });
        });
    }

    const unsigned long long lanemask_lt = (1ull << (__lane_id())) - 1ull;
    const int frow = (int)(4 * (F.plane_off + ((long long)w * a.numz + z0) * F.rs_stride));
    const pzl_const_int *ztab = (const pzl_const_int *)(a.zoff_tab + ((long long)w * a.numz + z0) * NQ);
    // A subharmonic's z map repeats rows (z h rounds to the same row for about
    // 1 - h of consecutive z): its values are then still in registers and are
    // not loaded again.
    float x[NQ][RPT];
    int zprev[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) zprev[q] = -1;
    for (int zi = 0; zi < nz; zi++) {
        const int z = z0 + zi;
        float acc[RPT];
        load_run<RPT>(planes, 4 * rbase, frow + zi * 4 * (int)F.rs_stride, acc);
        if constexpr (NSUB > 0) {
            const pzl_const_int *zt = ztab + zi * NQ;
            int off[NQ];
#pragma unroll
            for (int q = 0; q < NSUB; q++) off[q] = zt[q];  // all scalar loads first, then the branches
#pragma unroll
            for (int q = 0; q < NSUB; q++) {
                if (off[q] != zprev[q]) {
                    load_run<RPT>(planes, rind[q], off[q], x[q]);
                    zprev[q] = off[q];
                }
            }
        }
        // threshold of stage s after its harmonics are in: compaction of the
        // wave's hits in scan order (one atomic per wave)
        auto stage_check = [&](int s) {
            bool hit[RPT];
            bool any = false;
#pragma unroll
            for (int i = 0; i < RPT; i++) {
                hit[i] = active[i] && (acc[i] > a.powcut[s]);
                any = any || hit[i];
            }
            if (__ballot(any)) {
                unsigned long long mask[RPT];
                int cnt = 0;
#pragma unroll
                for (int i = 0; i < RPT; i++) {
                    mask[i] = __ballot(hit[i]);
                    cnt += __popcll(mask[i]);
                }
                unsigned long long slot0 = 0;
                if (__lane_id() == 0) slot0 = atomicAdd(a.nhits, (unsigned long long)cnt);
                slot0 = __shfl(slot0, 0);
#pragma unroll
                for (int i = 0; i < RPT; i++) {
                    if (hit[i]) {
                        const unsigned long long slot = slot0 + __popcll(mask[i] & lanemask_lt);
                        if (slot < a.maxhits) {
                            a.hit_keys[slot] = ((((rstep * NSTAGES + s) * a.numw + w) * (long long)a.numz + z) *
                                                    a.corr_uselen +
                                                rbase + i);
                            a.hit_pows[slot] = acc[i];
                        }
                    }
                    slot0 += __popcll(mask[i]);
                }
            }
        };
        stage_check(0);
        static_for<0, NSUB>([&](auto QC) {
            constexpr int q = decltype(QC)::value;
            constexpr int s = SelRange<RPT>::stage_of(q), j = SelRange<RPT>::j_of(q);
            acc[0] += x[q][0];
            static_for<1, RPT>([&](auto IC) {
                constexpr int i = decltype(IC)::value;
                constexpr int lo = SelRange<RPT>::lo(i, j, s), hi = SelRange<RPT>::hi(i, j, s);
                constexpr int b0 = SelRange<RPT>::bits_before(q) + SelRange<RPT>::bits_before_i(q, i);
                unsigned u = __float_as_uint(x[q][lo]);
                static_for<lo + 1, hi + 1>([&](auto CC) {
                    constexpr int c = decltype(CC)::value;
                    constexpr int bit = b0 + (c - lo - 1);
                    u = pick_bit_t<bit & 31>(u, x[q][c], sel[bit >> 5]);
                });
                acc[i] += __uint_as_float(u);
            });
            // the last harmonic of stage s (q = 2^s - 2): its threshold
            if constexpr (((q + 2) & (q + 1)) == 0) stage_check(s);
        });
        if (a.sums_out) {
#pragma unroll
            for (int i = 0; i < RPT; i++)
                if (active[i])
                    a.sums_out[(((long long)b * a.numw + w) * a.numz + z) * a.corr_uselen + rbase + i] = acc[i];
        }
    }
}

// ---------------------------------------------------------------------------
// Hit records: decode, fundamental r and significance of every sorted hit
// (accel_utils.c:1518-1533 with candidate_sigma, characteristics.c:647-668)
// ---------------------------------------------------------------------------

struct HitPrepArgs {
    const long long *keys;
    const float *pows;
    long long n;
    const double *startr;
    int corr_uselen, numz, numw, nstages;
    double numindep[6];
    PzlHostHit *out;
};

__global__ __launch_bounds__(256) void k_hitprep(HitPrepArgs a)
{
    // one hit per thread, no loop: constants of the significance are not kept
    // live across hits, so the kernel stays within the registers a finished
    // correlation wave frees (it runs beside the correlation)
    {
        const long long i = blockIdx.x * 256ll + threadIdx.x;
        if (i >= a.n) return;
        const long long key = a.keys[i];
        long long k = key;
        const int r = (int)(k % a.corr_uselen);
        k /= a.corr_uselen;
        const int zi = (int)(k % a.numz);
        k /= a.numz;
        const int wi = (int)(k % a.numw);
        k /= a.numw;
        const int stage = (int)(k % a.nstages);
        const long long rstep = k / a.nstages;
        const int numharm = 1 << stage;
        const double frlo = floor(pzl_calc_required_r(1.0, a.startr[rstep]));
        PzlHostHit h;
        h.key = key;
        h.power = a.pows[i];
        h.rr = (frlo + r * (double)PZL_DR) / (double)numharm;
        h.sigma = (float)pzl_candidate_sigma((double)h.power, numharm, a.numindep[stage]);
        h.zi = (short)zi;
        h.wi = (short)wi;
        h.stage = stage;
        a.out[i] = h;
    }
}

// ---------------------------------------------------------------------------
// Plane utilities behind the per-plane accel.h calls (add_ffdotpows,
// search_ffdotpows): fund[w][z][r] += sub[wind[w]][zind[z]][rind[r]]
// (accel_utils.c:1356-1377) and threshold + ordered compaction
// (accel_utils.c:1507-1555 scan order w, z, r).
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_add_plane(float *fund, int numws, int numzs, int numrs, const float *sub,
                                                   int sub_numzs, int sub_numrs, const int *winds,
                                                   const uint16_t *zinds, const uint16_t *rinds)
{
    const long long n = (long long)numws * numzs * numrs;
    for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        const int r = (int)(i % numrs);
        const int z = (int)((i / numrs) % numzs);
        const int w = (int)(i / ((long long)numrs * numzs));
        fund[i] += sub[((long long)winds[w] * sub_numzs + zinds[z]) * sub_numrs + rinds[r]];
    }
}

__global__ __launch_bounds__(256) void k_threshold(const float *plane, long long n, float powcut, long long *idx,
                                                   unsigned long long *count, long long cap)
{
    const unsigned long long lanemask_lt = (1ull << (__lane_id())) - 1ull;
    for (long long base = blockIdx.x * 256ll; base < n; base += (long long)gridDim.x * 256) {
        const long long i = base + threadIdx.x;
        const bool hit = i < n && plane[i] > powcut;
        const unsigned long long mask = __ballot(hit);
        if (mask) {
            const int leader = __ffsll((long long)mask) - 1;
            unsigned long long slot0 = 0;
            if ((int)__lane_id() == leader) slot0 = atomicAdd(count, (unsigned long long)__popcll(mask));
            slot0 = __shfl(slot0, leader);
            const unsigned long long slot = slot0 + __popcll(mask & lanemask_lt);
            if (hit && (long long)slot < cap) idx[slot] = i;
        }
    }
}

extern "C" hipError_t pzl_launch_add_plane(float *fund, int numws, int numzs, int numrs, const float *sub,
                                           int sub_numzs, int sub_numrs, const int *winds, const uint16_t *zinds,
                                           const uint16_t *rinds, hipStream_t st)
{
    const long long n = (long long)numws * numzs * numrs;
    const long long blocks = std::min<long long>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_add_plane, dim3((unsigned)std::max<long long>(blocks, 1)), dim3(256), 0, st, fund, numws,
                       numzs, numrs, sub, sub_numzs, sub_numrs, winds, zinds, rinds);
    return hipGetLastError();
}

extern "C" hipError_t pzl_launch_threshold(const float *plane, long long n, float powcut, long long *idx,
                                           unsigned long long *count, long long cap, hipStream_t st)
{
    const long long blocks = std::min<long long>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_threshold, dim3((unsigned)std::max<long long>(blocks, 1)), dim3(256), 0, st, plane, n,
                       powcut, idx, count, cap);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Launchers (C linkage for the host engine)
// ---------------------------------------------------------------------------

extern "C" hipError_t pzl_launch_hitprep(const long long *keys, const float *pows, long long n,
                                         const double *startr, int corr_uselen, int numz, int numw, int nstages,
                                         const double *numindep, PzlHostHit *out, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    HitPrepArgs a;
    a.keys = keys;
    a.pows = pows;
    a.n = n;
    a.startr = startr;
    a.corr_uselen = corr_uselen;
    a.numz = numz;
    a.numw = numw;
    a.nstages = nstages;
    for (int s = 0; s < 6; s++) a.numindep[s] = numindep[s];
    a.out = out;
    const long long blocks = (n + 255) / 256;
    if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_hitprep, dim3((unsigned)blocks), dim3(256), 0, st, a);
    return hipGetLastError();
}

extern "C" hipError_t pzl_launch_bank_build(const PzlHarm *d_harms, int nh, int total_rows, int max_fftlen,
                                            float2 *d_bank, hipStream_t st)
{
    const size_t smem = (size_t)max_fftlen * sizeof(double2);
    hipLaunchKernelGGL(k_bank_build, dim3(total_rows), dim3(BANK_NT), smem, st, d_harms, nh, d_bank);
    return hipGetLastError();
}

extern "C" hipError_t pzl_launch_segments(const PzlBatchArgs *a, int max_fftlen, hipStream_t st)
{
    const int numdata = max_fftlen / 2;
    if (numdata > SEG_MAXDATA) return hipErrorInvalidValue;  // the transform's registers hold no more
    int n2 = 1;
    while (n2 < numdata) n2 <<= 1;
    const size_t smem = (size_t)numdata * sizeof(double2) + (size_t)n2 * sizeof(float);
    return pzl_launch((const void *)k_segments, dim3(a->nb * a->nh), dim3(SEG_NT), smem, st, *a);
}

// One correlation variant: kernel-row length N, NT threads, G rows per
// workgroup, ping-pong LDS or not, split rows or not, radices.  One workgroup
// per work unit (the grid is not persistent, so that kernels of concurrent
// searches interleave); its twiddles come from a table built once per device.
template <int N, int NT, int G, bool PING, int SH, int PADSH, int... Rs>
static hipError_t launch_corr_t(const PzlCorrArgs *a, const int *class_nrows, hipStream_t st)
{
    using V = CorrV<N, NT, G, PING, SH, PADSH, Rs...>;
    PzlCorrArgs a2;
    if (!corr_units_of(*a, class_nrows, NT, G, a2)) return hipErrorInvalidValue;
    if (a2.total_items < 1) return hipSuccess;
    hipError_t e = V::table(st, &a2.twtab);
    if (e != hipSuccess) return e;
    return pzl_launch((const void *)k_corr<N, NT, G, PING, SH, PADSH, Rs...>, dim3((unsigned)a2.total_items), dim3(NT),
                      V::smem(), st, a2);
}

// ---------------------------------------------------------------------------
// Grouped correlation: the classes of several FFT lengths in one launch
// ---------------------------------------------------------------------------
//
// A group is a list of variants with one thread count, one per FFT length,
// longest first.  The launch holds the work units of every class the caller
// passes, class after class in the group's order: each workgroup finds its
// class from its index (uniform, so the whole workgroup takes one variant's
// code) and runs that unit.  The short lengths, which alone fill a small part
// of the chip, then share one launch and its tail with the longer ones.

#define PZL_CORR_GROUP_MAX 10

struct PzlCorrGroupArgs {
    PzlCorrArgs c[PZL_CORR_GROUP_MAX];    // per member: its class (total_items 0: absent)
    long long begin[PZL_CORR_GROUP_MAX + 1];  // first workgroup of each member
};

template <int K, typename V, typename... Vs>
__device__ __forceinline__ void group_unit(const PzlCorrGroupArgs &g, long long wg, float2 *sm)
{
    if (wg < g.begin[K + 1]) {
        V::unit(g.c[K], wg - g.begin[K], sm);
        return;
    }
    if constexpr (sizeof...(Vs) > 0) group_unit<K + 1, Vs...>(g, wg, sm);
}

// WMIN: waves per SIMD the registers must allow (1: the compiler's choice)
template <int NT, int WMIN, typename... Vs>
__global__ __launch_bounds__(NT, WMIN) void k_corr_group(PzlCorrGroupArgs g)
{
    extern __shared__ float2 smem[];
    group_unit<0, Vs...>(g, blockIdx.x, smem);
}

template <int NT, int WMIN, typename... Vs> struct CorrGroup {
    static constexpr int M = sizeof...(Vs);
    static_assert(M <= PZL_CORR_GROUP_MAX, "too many members");
    static_assert(((Vs::NT == NT) && ...), "members share the thread count");
    static constexpr int lens[M] = {Vs::N...};
    static int member(int n)
    {
        for (int k = 0; k < M; k++)
            if (lens[k] == n) return k;
        return -1;
    }
    // members K..: their classes' units (cls[k] < 0: absent), in order
    template <int K, typename V, typename... Rest>
    static hipError_t fill(PzlCorrGroupArgs &g, long long &acc, const int *cls, const PzlCorrArgs *const *args,
                           const int *const *class_nrows, hipStream_t st)
    {
        g.begin[K] = acc;
        g.c[K] = PzlCorrArgs{};
        if (cls[K] >= 0) {
            if (!corr_units_of(*args[cls[K]], class_nrows[cls[K]], NT, V::G, g.c[K])) return hipErrorInvalidValue;
            if (g.c[K].total_items > 0) {
                hipError_t e = V::table(st, &g.c[K].twtab);
                if (e != hipSuccess) return e;
            }
            acc += g.c[K].total_items;
        }
        if constexpr (sizeof...(Rest) > 0) return fill<K + 1, Rest...>(g, acc, cls, args, class_nrows, st);
        return hipSuccess;
    }
    static hipError_t launch(int ncls, const int *fftlens, const PzlCorrArgs *const *args,
                             const int *const *class_nrows, hipStream_t st)
    {
        int cls[M];
        for (int k = 0; k < M; k++) cls[k] = -1;
        for (int c = 0; c < ncls; c++) {
            const int k = member(fftlens[c]);
            if (k < 0 || cls[k] >= 0) return hipErrorInvalidValue;  // not a member, or a length twice
            cls[k] = c;
        }
        PzlCorrGroupArgs g{};
        long long acc = 0;
        hipError_t e = fill<0, Vs...>(g, acc, cls, args, class_nrows, st);
        if (e != hipSuccess) return e;
        for (int k = M; k <= PZL_CORR_GROUP_MAX; k++) g.begin[k] = acc;
        if (acc < 1) return hipSuccess;
        if (acc > 0x7fffffffLL) return hipErrorInvalidValue;
        const size_t sm = std::max({Vs::smem()...});
        return pzl_launch((const void *)k_corr_group<NT, WMIN, Vs...>, dim3((unsigned)acc), dim3(NT), sm, st, g);
    }
    static int has(int n) { return member(n) >= 0; }
};

// The groups, each with a member for every FFT length from 128 to 2048 (the
// lengths shorter than the longest of a search).
using CorrGroup256 =
    CorrGroup<256, 1, CorrV<2048, 256, 1, true, 0, 3, 8, 8, 8, 4>, CorrV<1280, 256, 2, true, 0, 4, 16, 16, 5>,
              CorrV<1024, 256, 2, true, 0, 5, 8, 8, 4, 4>, CorrV<768, 256, 4, true, 0, 4, 12, 4, 4, 4>,
              CorrV<512, 256, 4, true, 0, 4, 8, 8, 8>, CorrV<384, 256, 8, true, 0, 4, 6, 4, 4, 4>,
              CorrV<256, 256, 8, true, 0, 4, 8, 8, 4>, CorrV<192, 256, 16, true, 0, 4, 12, 4, 4>,
              CorrV<128, 256, 16, true, 0, 4, 8, 4, 4>>;

// The same with fewer rows per workgroup for the lengths whose variants above
// hold many values per thread: every member stays within 128 VGPRs, so the
// group keeps 4 waves per SIMD.
using CorrGroup256Lean =
    CorrGroup<256, 4, CorrV<2048, 256, 1, true, 0, 5, 8, 8, 4, 8>, CorrV<1280, 256, 1, true, 0, 4, 4, 8, 8, 5>,
              CorrV<1024, 256, 2, true, 0, 5, 8, 8, 4, 4>, CorrV<768, 256, 1, true, 0, 4, 3, 4, 4, 4, 4>,
              CorrV<512, 256, 4, true, 0, 4, 8, 8, 8>, CorrV<384, 256, 4, true, 0, 4, 8, 8, 6>,
              CorrV<256, 256, 8, true, 0, 4, 8, 8, 4>, CorrV<192, 256, 8, true, 0, 4, 8, 4, 6>,
              CorrV<128, 256, 16, true, 0, 4, 8, 4, 4>>;

struct CorrGroupEntry {
    const char *name;
    hipError_t (*launch)(int, const int *, const PzlCorrArgs *const *, const int *const *, hipStream_t);
    int (*has)(int);
};

static const CorrGroupEntry g_corr_groups[] = {
    {"256t: 2048:1r 1280:2r 1024:2r 768:4r 512:4r 384:8r 256:8r 192:16r 128:16r", CorrGroup256::launch,
     CorrGroup256::has},
    {"256t lean: 2048:1r 1280:1r 1024:2r 768:1r 512:4r 384:4r 256:8r 192:8r 128:16r", CorrGroup256Lean::launch,
     CorrGroup256Lean::has},
};
static const int g_ncorr_groups = (int)(sizeof(g_corr_groups) / sizeof(g_corr_groups[0]));

extern "C" int pzl_corr_num_groups(void) { return g_ncorr_groups; }

extern "C" const char *pzl_corr_group_name(int group)
{
    return group >= 0 && group < g_ncorr_groups ? g_corr_groups[group].name : nullptr;
}

extern "C" int pzl_corr_group_has(int group, int n)
{
    return group >= 0 && group < g_ncorr_groups && g_corr_groups[group].has(n);
}

extern "C" hipError_t pzl_launch_corr_group(int group, int ncls, const int *fftlens, const PzlCorrArgs *const *args,
                                            const int *const *class_nrows, hipStream_t st)
{
    if (group < 0 || group >= g_ncorr_groups || ncls < 0) return hipErrorInvalidValue;
    return g_corr_groups[group].launch(ncls, fftlens, args, class_nrows, st);
}

typedef hipError_t (*CorrLaunch)(const PzlCorrArgs *, const int *, hipStream_t);
struct CorrVariant {
    int n;
    CorrLaunch fn;
    const char *name;
};

#define PZL_CV(N, NT, G, PING, ...) \
    {N, launch_corr_t<N, NT, G, PING, 0, 4, __VA_ARGS__>, #N ":" #NT "t:" #G "r:" #PING ":" #__VA_ARGS__}
// the same with each row split into its even and odd outputs (two N/2-point FFTs)
#define PZL_CVS(N, NT, G, PING, ...) \
    {N, launch_corr_t<N, NT, G, PING, 1, 4, __VA_ARGS__>, #N ":" #NT "t:" #G "r:" #PING ":split:" #__VA_ARGS__}
// split rows with LDS padding of one element per 2^PADSH
#define PZL_CVSP(N, NT, G, PING, PADSH, ...)                                                                  \
    {N, launch_corr_t<N, NT, G, PING, 1, PADSH, __VA_ARGS__>,                                                 \
     #N ":" #NT "t:" #G "r:" #PING ":split:pad" #PADSH ":" #__VA_ARGS__}
// the default with LDS padding of one element per 2^PADSH
#define PZL_CVP(N, NT, G, PING, PADSH, ...)                                                                   \
    {N, launch_corr_t<N, NT, G, PING, 0, PADSH, __VA_ARGS__>,                                                 \
     #N ":" #NT "t:" #G "r:" #PING ":pad" #PADSH ":" #__VA_ARGS__}
// Candidates per FFT length; the first of each length is the default, the
// autotuner may pick another on the device.
static const CorrVariant g_corr_variants[] = {
    PZL_CV(128, 128, 16, true, 16, 8),
    PZL_CV(128, 256, 16, true, 8, 4, 4),
    PZL_CV(192, 256, 16, true, 12, 4, 4),
    PZL_CV(192, 64, 4, true, 12, 4, 4),
    PZL_CV(192, 256, 8, true, 8, 4, 6),
    PZL_CV(256, 128, 8, true, 16, 16),
    PZL_CV(256, 256, 8, true, 8, 8, 4),
    PZL_CV(384, 256, 8, true, 6, 4, 4, 4),
    PZL_CV(384, 64, 2, true, 12, 4, 4, 2),
    PZL_CV(384, 256, 4, true, 8, 8, 6),
    PZL_CV(512, 128, 4, true, 8, 8, 8),
    PZL_CV(512, 256, 4, true, 8, 8, 8),
    PZL_CV(512, 128, 4, true, 16, 16, 2),
    PZL_CV(768, 256, 4, true, 12, 4, 4, 4),
    PZL_CV(768, 64, 1, true, 12, 4, 4, 4),
    PZL_CV(768, 256, 1, true, 3, 4, 4, 4, 4),
    PZL_CV(1024, 128, 2, true, 16, 16, 4),
    PZL_CV(1024, 256, 2, true, 8, 8, 4, 4),
    PZL_CV(1024, 64, 1, true, 16, 16, 4),
    PZL_CVS(1024, 128, 1, true, 8, 8, 8),
    PZL_CVP(1024, 256, 2, true, 5, 8, 8, 4, 4),
    PZL_CVP(1024, 128, 1, true, 5, 8, 8, 4, 4),
    PZL_CVP(1024, 256, 2, true, 5, 8, 4, 4, 8),
    PZL_CV(1280, 128, 1, true, 16, 16, 5),
    PZL_CV(1280, 256, 2, true, 16, 16, 5),
    PZL_CV(1280, 256, 1, true, 4, 8, 8, 5),
    PZL_CVP(1280, 128, 1, true, 3, 16, 16, 5),
    PZL_CVP(1280, 128, 1, true, 5, 16, 16, 5),
    PZL_CVP(1280, 320, 2, true, 5, 8, 8, 4, 5),
    PZL_CV(2048, 128, 1, true, 16, 16, 8),
    PZL_CV(2048, 256, 1, true, 8, 8, 8, 4),
    PZL_CV(2048, 256, 2, false, 16, 16, 8),
    PZL_CVP(2048, 256, 1, true, 3, 8, 8, 8, 4),
    PZL_CVP(2048, 256, 1, true, 5, 8, 8, 8, 4),
    PZL_CVP(2048, 256, 1, true, 6, 8, 8, 8, 4),
    PZL_CVP(2048, 256, 1, true, 5, 8, 8, 4, 8),
    PZL_CVP(2048, 256, 1, true, 5, 4, 8, 8, 8),
    PZL_CVP(2048, 128, 1, true, 3, 16, 16, 8),
    PZL_CVS(2048, 128, 1, true, 16, 16, 4),
    PZL_CVS(2048, 256, 2, false, 16, 16, 4),
    PZL_CVSP(2048, 256, 1, true, 5, 8, 8, 4, 4),
    PZL_CVSP(2048, 256, 1, true, 5, 8, 4, 4, 8),
    PZL_CVSP(2048, 256, 1, true, 5, 4, 8, 8, 4),
    PZL_CV(4096, 256, 1, false, 16, 16, 16),
    PZL_CV(4096, 512, 1, false, 8, 8, 8, 8),
    PZL_CVS(4096, 256, 1, false, 16, 16, 8),
    PZL_CVSP(4096, 512, 1, true, 5, 8, 8, 8, 4),
    PZL_CVP(4096, 512, 1, true, 5, 8, 8, 8, 8),
};
static const int g_ncorr_variants = (int)(sizeof(g_corr_variants) / sizeof(g_corr_variants[0]));

static const CorrVariant *corr_variant(int n, int v)
{
    int k = 0;
    for (int i = 0; i < g_ncorr_variants; i++)
        if (g_corr_variants[i].n == n) {
            if (k == v) return &g_corr_variants[i];
            k++;
        }
    return nullptr;
}

extern "C" int pzl_corr_supported(int n) { return corr_variant(n, 0) != nullptr; }

extern "C" int pzl_corr_num_variants(int n)
{
    int k = 0;
    for (int i = 0; i < g_ncorr_variants; i++) k += g_corr_variants[i].n == n;
    return k;
}

extern "C" const char *pzl_corr_variant_name(int n, int v)
{
    const CorrVariant *cv = corr_variant(n, v);
    return cv ? cv->name : "";
}

extern "C" hipError_t pzl_launch_corr(int N, int variant, const PzlCorrArgs *a, const int *class_nrows,
                                      hipStream_t st)
{
    const CorrVariant *cv = corr_variant(N, variant);
    if (!cv) return hipErrorInvalidValue;
    return cv->fn(a, class_nrows, st);
}

template <int RPT> static hipError_t launch_harmsum_t(const PzlSumArgs *a, hipStream_t st)
{
    using SS = SumShape<RPT>;
    const int chunks = (a->corr_uselen + SS::RCH - 1) / SS::RCH;
    const long long nblk = (long long)a->nb * a->numw * ((a->numz + SUM_ZB - 1) / SUM_ZB) * chunks;
    const dim3 g((unsigned)(8 * ((nblk + 7) / 8))), t(SS::NT);
    switch (a->nstages) {
    case 1: hipLaunchKernelGGL((k_harmsum<1, RPT>), g, t, 0, st, *a); break;
    case 2: hipLaunchKernelGGL((k_harmsum<2, RPT>), g, t, 0, st, *a); break;
    case 3: hipLaunchKernelGGL((k_harmsum<3, RPT>), g, t, 0, st, *a); break;
    case 4: hipLaunchKernelGGL((k_harmsum<4, RPT>), g, t, 0, st, *a); break;
    case 5: hipLaunchKernelGGL((k_harmsum<5, RPT>), g, t, 0, st, *a); break;
    case 6: hipLaunchKernelGGL((k_harmsum<6, RPT>), g, t, 0, st, *a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

extern "C" hipError_t pzl_launch_harmsum(const PzlSumArgs *a, hipStream_t st)
{
    // buffer offsets are 32-bit; the planes of a set carry PZL_PLANE_SLACK floats behind them
    if ((a->plane_stride + PZL_PLANE_SLACK) * 4 >= (1LL << 31)) return hipErrorInvalidValue;
    switch (a->rpt) {
    case 1: return launch_harmsum_t<1>(a, st);
    case 2: return launch_harmsum_t<2>(a, st);
    case 4: return launch_harmsum_t<4>(a, st);
    default: return hipErrorInvalidValue;
    }
}
