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

#include <hipcub/hipcub.hpp>

#include "pzl_fft.h"
#include "pzl_kernels.h"
#include "pzl_math.h"
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

    if (fabs(w) >= 1e-4) {
        // gen_w_response (responses.c:333-418): float samples of a cubic-phase
        // chirp, zero-padded by numbetween; only numkern bins of its real FFT
        // are kept, so they are computed directly as a DFT in double.
        const int npts = pzl_next2_to_n(6 * khw + 200 + numkern / PZL_NUMBETWEEN);
        const double dt = 1.0 / (double)npts, amp = 2.0 * dt;
        const int fbar = npts / 4;
        const double f = (double)fbar - 0.5 * z + w / 12.0, fd = (z - 0.5 * w) / 2.0, fdd = w / 6.0;
        float *samp = (float *)sbuf;
        for (int t = tid; t < npts; t += BANK_NT) {
            const double tt = t * dt;
            samp[t] = (float)(amp * cos(6.283185307179586477 * (tt * (tt * (tt * fdd + fd) + f))));
        }
        __syncthreads();
        const int nfft = npts * PZL_NUMBETWEEN;  // real FFT length
        const int k0 = fbar * PZL_NUMBETWEEN - half;
#pragma unroll
        for (int i = 0; i < BANK_MAXV; i++) {
            const int m = tid + i * BANK_NT;
            double ar = 0.0, ai = 0.0;
            if (m < numkern) {
                const long long k = k0 + m;
                // X[k] = sum_t samp[t] exp(-2 pi i t k / nfft); the rotation is
                // re-anchored exactly every 64 samples.
                double sr, cr;
                sincospi(-2.0 * (double)k / (double)nfft, &sr, &cr);
                for (int t0 = 0; t0 < npts; t0 += 64) {
                    double s0, c0;
                    const long long ph = ((long long)t0 * k) % nfft;
                    sincospi(-2.0 * (double)ph / (double)nfft, &s0, &c0);
                    const int tend = t0 + 64 < npts ? t0 + 64 : npts;
                    for (int t = t0; t < tend; t++) {
                        const double d = (double)samp[t];
                        ar += d * c0;
                        ai += d * s0;
                        const double nc = c0 * cr - s0 * sr;
                        s0 = c0 * sr + s0 * cr;
                        c0 = nc;
                    }
                }
            }
            resp[i] = double2{ar, ai};
        }
        __syncthreads();
    } else if (fabs(z) >= 1e-4) {
        // gen_z_response (responses.c:228-316), roffset = 0
        const double PI = 3.14159265358979323846;
        const double absz = fabs(z);
        const int signz = (z < 0.0) ? -1 : 1;
        const double zd = signz * 1.41421356237309504880 / sqrt(absz);
        const double cons = zd / 2.0, pibyz = PI / z;
        const double startr = -0.5 * z + numkern / (double)(2 * PZL_NUMBETWEEN);
        const double delta = -1.0 / PZL_NUMBETWEEN;
#pragma unroll
        for (int i = 0; i < BANK_MAXV; i++) {
            const int m = tid + i * BANK_NT;
            double re = 0.0, im = 0.0;
            if (m < numkern) {
                const double r = startr + m * delta;
                const double yy = r * zd, zz = yy + z * zd, xx = pibyz * r * r;
                double fsy, fcy, fsz, fcz;
                pzl_fresnel(yy, &fsy, &fcy);
                pzl_fresnel(zz, &fsz, &fcz);
                const double c = cos(xx), s = sin(xx);
                const double tr = signz * (fcz - fcy), ti = fsy - fsz;
                re = (tr * c - ti * s) * cons;
                im = -(tr * s + ti * c) * cons;
            }
            resp[i] = double2{re, im};
        }
    } else {
        // gen_r_response (responses.c:159-225), roffset = 0: sinc * e^{i r}
#pragma unroll
        for (int i = 0; i < BANK_MAXV; i++) {
            const int m = tid + i * BANK_NT;
            double re = 0.0, im = 0.0;
            if (m < numkern) {
                const double x = numkern / (double)(2 * PZL_NUMBETWEEN) - m * 0.5;  // r / pi
                if (m == half) {
                    re = 1.0;
                    im = 0.0;
                } else {
                    double s, c;
                    sincospi(x, &s, &c);
                    const double sinc = s / (3.14159265358979323846 * x);
                    re = c * sinc;
                    im = s * sinc;
                }
            }
            resp[i] = double2{re, im};
        }
    }
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
    lds_fft_f64<-1>(sbuf, fftlen, tid, BANK_NT);
    float2 *out = bank + H.bank_off + (long long)row * fftlen;
    for (int k = tid; k < fftlen; k += BANK_NT) out[k] = float2{(float)sbuf[k].x, (float)sbuf[k].y};
}

// ---------------------------------------------------------------------------
// Segments: gather, normalise, spread + forward FFT, r-index tables
// ---------------------------------------------------------------------------

#define SEG_NT 256

__global__ __launch_bounds__(SEG_NT) void k_segments(PzlBatchArgs a)
{
    extern __shared__ double2 sbuf[];  // fftlen/2 complex doubles, then keys
    const int tid = threadIdx.x;
    const int b = blockIdx.x / a.nh, h = blockIdx.x % a.nh;
    const PzlHarm H = a.harms[h];
    const long long rstep = a.rstep0 + b;
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
        // bitonic sort of the powers; the median is sorted element (n-1)/2
        for (int k = 2; k <= n2; k <<= 1) {
            for (int j = k >> 1; j > 0; j >>= 1) {
                for (int i = tid; i < n2; i += SEG_NT) {
                    const int ixj = i ^ j;
                    if (ixj > i) {
                        const float ki = keys[i], kj = keys[ixj];
                        const bool up = ((i & k) == 0);
                        if ((ki > kj) == up) {
                            keys[i] = kj;
                            keys[ixj] = ki;
                        }
                    }
                }
                __syncthreads();
            }
        }
        norm = 1.0 / sqrt((double)keys[(numdata - 1) / 2] / log(2.0));
    }
    __syncthreads();
    for (int i = tid; i < numdata; i += SEG_NT) {
        const double2 x = sbuf[i];
        sbuf[i] = double2{(double)(float)(x.x * norm), (double)(float)(x.y * norm)};
    }
    __syncthreads();
    // FFT of the fftlen-point spread array == numdata-point FFT, repeated.
    lds_fft_f64<-1>(sbuf, numdata, tid, SEG_NT);
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

__host__ __device__ __forceinline__ int lds_pad(int p) { return p + (p >> 4); }

struct SlotDesc {
    const float2 *seg;
    const float2 *kern;
    float *plane;
    int offset, numrs, valid;
};

// Stockham pass with radix R on G FFTs of length N held in LDS; Ns is the
// product of the radices already applied.  The last pass turns the
// natural-order outputs into powers and stores the wanted r range.
template <int N, int G, int NT, int Ns, int R, int... Rest>
__device__ __forceinline__ void corr_pass(float2 *data, const float2 *tw, const SlotDesc *sd, int tid, float norm)
{
    constexpr int NB = N / R;
    constexpr int PER = G * NB / NT;
    static_assert(PER * NT == G * NB, "pass does not tile the workgroup");
    constexpr bool last = (sizeof...(Rest) == 0);
    constexpr int TWS = N / (Ns * R);
    float2 v[PER][R];
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const int bf = tid + i * NT;
        const int g = bf / NB, j = bf % NB;
        const int step = (j % Ns) * TWS;
#pragma unroll
        for (int r = 0; r < R; r++) {
            float2 x = data[lds_pad(g * N + j + r * NB)];
            if (r > 0) x = cmul(x, tw[step * r]);
            v[i][r] = x;
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const int bf = tid + i * NT;
        const int g = bf / NB, j = bf % NB;
        dftR<R, 1>(v[i]);
        if constexpr (!last) {
            const int idxD = (j / Ns) * Ns * R + (j % Ns);
#pragma unroll
            for (int r = 0; r < R; r++) data[lds_pad(g * N + idxD + r * Ns)] = v[i][r];
        } else {
            const SlotDesc d = sd[g];
            if (d.valid) {
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const int ro = j + r * NB - d.offset;  // natural-order index minus offset
                    if (ro >= 0 && ro < d.numrs) {
                        const float2 y = v[i][r];
                        d.plane[ro] = (y.x * y.x + y.y * y.y) * norm;
                    }
                }
            }
        }
    }
    if constexpr (!last) {
        __syncthreads();
        corr_pass<N, G, NT, Ns * R, Rest...>(data, tw, sd, tid, norm);
    }
}

template <int N, int G, int NT, int R0, int... Rest>
__global__ __launch_bounds__(NT) void k_corr(PzlCorrArgs a)
{
    extern __shared__ float2 smem[];
    float2 *data = smem;                     // G*N (+pad)
    float2 *tw = smem + lds_pad(G * N) + 1;  // N twiddles exp(+2 pi i m / N)
    __shared__ SlotDesc sd[G];
    const int tid = threadIdx.x;
    const float norm = (float)(1.0 / ((double)N * (double)N));

    for (int m = tid; m < N; m += NT) {
        double s, c;
        sincospi(2.0 * (double)m / (double)N, &s, &c);
        tw[m] = float2{(float)c, (float)s};
    }

    for (long long grp = blockIdx.x; grp * G < a.total_items; grp += gridDim.x) {
        __syncthreads();
        if (tid < G) {
            const long long item = grp * G + tid;
            SlotDesc d;
            d.valid = 0;
            d.seg = nullptr;
            d.kern = nullptr;
            d.plane = nullptr;
            d.offset = 0;
            d.numrs = 0;
            if (item < a.total_items) {
                // items: for each harmonic of this size class, rows x batch
                long long acc = 0;
                for (int c = 0; c < a.nclass; c++) {
                    const int h = a.class_harms[c];
                    const PzlHarm H = a.harms[h];
                    const long long cnt = (long long)H.nrows * a.nb;
                    if (item < acc + cnt) {
                        const long long loc = item - acc;
                        const int row = (int)(loc / a.nb), b = (int)(loc % a.nb);
                        d.valid = 1;
                        d.seg = a.segs + (long long)b * a.seg_stride + H.seg_off;
                        d.kern = a.bank + H.bank_off + (long long)row * N;
                        d.plane = a.planes + (long long)b * a.plane_stride + H.plane_off +
                                  (long long)row * H.rs_stride;
                        d.offset = 2 * H.halfwidth;
                        d.numrs = a.hr_tab[b * a.nh + h].numrs;
                        break;
                    }
                    acc += cnt;
                }
            }
            sd[tid] = d;
        }
        __syncthreads();

        // pass 0 (Ns = 1): P[k] = seg[k mod N/2] * conj(kern[k]) straight from HBM/L2
        {
            constexpr int R = R0;
            constexpr int NB = N / R;
            constexpr int PER = G * NB / NT;
            static_assert(PER * NT == G * NB, "pass 0 does not tile the workgroup");
#pragma unroll
            for (int i = 0; i < PER; i++) {
                const int bf = tid + i * NT;
                const int g = bf / NB, j = bf % NB;
                const SlotDesc d = sd[g];
                float2 v[R];
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const int k = j + r * NB;
                    float2 x = float2{0.f, 0.f};
                    if (d.valid) {
                        const float2 pd = d.seg[k < N / 2 ? k : k - N / 2];
                        const float2 kk = d.kern[k];
                        // re = dr*kr + di*ki, im = di*kr - dr*ki (accel_utils.c:1262-1265)
                        x = float2{pd.x * kk.x + pd.y * kk.y, pd.y * kk.x - pd.x * kk.y};
                    }
                    v[r] = x;
                }
                dftR<R, 1>(v);
#pragma unroll
                for (int r = 0; r < R; r++) data[lds_pad(g * N + j * R + r)] = v[r];
            }
        }
        __syncthreads();
        corr_pass<N, G, NT, R0, Rest...>(data, tw, sd, tid, norm);
    }
}

// ---------------------------------------------------------------------------
// Harmonic sum, threshold, compaction
// ---------------------------------------------------------------------------

#define SUM_NT 256

// rint((m * j) / 2^s) with round-half-even, in integers: the subharmonic r
// index of fundamental point m = 2*fullrlo + ii (accel_utils.c:1115-1122 with
// integer fullrlo; rint rounds half to even).
__device__ __forceinline__ long long rint_dyadic(long long m, int j, int s)
{
    const long long prod = m * j;
    long long q = prod >> s;
    if (s > 0) {
        const long long rem = prod & ((1LL << s) - 1), half = 1LL << (s - 1);
        if (rem > half || (rem == half && (q & 1))) q++;
    }
    return q;
}

// One workgroup: SUM_NT consecutive r of one (rstep, w, z) row of the
// fundamental.  All subharmonic powers of a cell are loaded first and then
// added in the reference's order (stage 1..S, odd harmonics ascending).
template <int NSTAGES>
__global__ __launch_bounds__(SUM_NT) void k_harmsum(PzlSumArgs a)
{
    constexpr int NSUB = (1 << (NSTAGES - 1)) - 1;
    const int tid = threadIdx.x;
    const int chunks = (a.corr_uselen + SUM_NT - 1) / SUM_NT;
    const long long blk = blockIdx.x;
    const int chunk = (int)(blk % chunks);
    const long long rowid = blk / chunks;  // (b, w, z)
    const int z = (int)(rowid % a.numz);
    const int w = (int)((rowid / a.numz) % a.numw);
    const int b = (int)(rowid / ((long long)a.numz * a.numw));
    const int r = chunk * SUM_NT + tid;
    const bool active = r < a.corr_uselen;
    const int rr = active ? r : a.corr_uselen - 1;
    const long long rstep = a.rstep0 + b;
    const float *base = a.planes + (long long)b * a.plane_stride;
    const PzlHarm F = a.harms[0];
    float acc = base[F.plane_off + ((long long)w * a.numz + z) * F.rs_stride + rr];

    float v[NSUB > 0 ? NSUB : 1];
    if constexpr (NSUB > 0) {
        const long long m = a.int_rlo ? (long long)(2.0 * a.startr[rstep]) + rr : 0;
#pragma unroll
        for (int q = 0; q < NSUB; q++) {
            const int h = q + 1;
            const PzlHarm H = a.harms[h];
            const float *row = base + H.plane_off +
                               ((long long)a.winds[h * a.numw + w] * H.numz + a.zinds[h * a.numz + z]) * H.rs_stride;
            int rind;
            if (a.int_rlo)
                rind = (int)(rint_dyadic(m, H.harmnum, H.stage) - 2 * a.hr_tab[b * a.nh + h].rlo);
            else
                rind = a.rinds[((long long)b * a.nh + h) * a.corr_uselen + rr];
            v[q] = row[rind];
        }
    }
    const unsigned long long lanemask_lt = (1ull << (__lane_id())) - 1ull;
    int q = 0;
#pragma unroll
    for (int s = 0; s < NSTAGES; s++) {
        if (s > 0) {
#pragma unroll
            for (int k = 0; k < (1 << (s - 1)); k++, q++) acc += v[q];
        }
        const bool hit = active && (acc > a.powcut[s]);
        const unsigned long long mask = __ballot(hit);
        if (mask) {
            const int cnt = __popcll(mask);
            const int leader = __ffsll((long long)mask) - 1;
            unsigned long long slot0 = 0;
            if ((int)__lane_id() == leader) slot0 = atomicAdd(a.nhits, (unsigned long long)cnt);
            slot0 = __shfl(slot0, leader);
            if (hit) {
                const unsigned long long slot = slot0 + __popcll(mask & lanemask_lt);
                if (slot < a.maxhits) {
                    a.hit_keys[slot] =
                        ((((rstep * NSTAGES + s) * a.numw + w) * (long long)a.numz + z) * a.corr_uselen + r);
                    a.hit_pows[slot] = acc;
                }
            }
        }
    }
    if (a.sums_out && active)
        a.sums_out[(((long long)b * a.numw + w) * a.numz + z) * a.corr_uselen + r] = acc;
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
    for (long long i = blockIdx.x * 256ll + threadIdx.x; i < a.n; i += (long long)gridDim.x * 256) {
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

extern "C" hipError_t pzl_sort_hits(void *temp, size_t *temp_bytes, const long long *keys_in, long long *keys_out,
                                    const float *pows_in, float *pows_out, long long n, int end_bit,
                                    hipStream_t st)
{
    return hipcub::DeviceRadixSort::SortPairs(temp, *temp_bytes, keys_in, keys_out, pows_in, pows_out, (int)n, 0,
                                              end_bit, st);
}

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
    long long blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
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
    int n2 = 1;
    while (n2 < numdata) n2 <<= 1;
    const size_t smem = (size_t)numdata * sizeof(double2) + (size_t)n2 * sizeof(float);
    hipLaunchKernelGGL(k_segments, dim3(a->nb * a->nh), dim3(SEG_NT), smem, st, *a);
    return hipGetLastError();
}

template <int N, int G, int NT, int... Rs>
static hipError_t launch_corr_t(const PzlCorrArgs *a, int nblocks, hipStream_t st)
{
    const size_t smem = (size_t)(lds_pad(G * N) + 1 + N) * sizeof(float2);
    long long groups = (a->total_items + G - 1) / G;
    int grid = (int)(groups < nblocks ? groups : nblocks);
    if (grid < 1) return hipSuccess;
    hipLaunchKernelGGL((k_corr<N, G, NT, Rs...>), dim3(grid), dim3(NT), smem, st, *a);
    return hipGetLastError();
}

extern "C" int pzl_corr_supported(int n)
{
    switch (n) {
    case 128: case 192: case 256: case 384: case 512: case 768: case 1024: case 1280: case 2048: case 4096:
        return 1;
    default:
        return 0;
    }
}

extern "C" hipError_t pzl_launch_corr(int N, const PzlCorrArgs *a, int nblocks, hipStream_t st)
{
    switch (N) {
    case 128: return launch_corr_t<128, 16, 256, 8, 4, 4>(a, nblocks, st);
    case 192: return launch_corr_t<192, 16, 256, 3, 4, 4, 4>(a, nblocks, st);
    case 256: return launch_corr_t<256, 8, 256, 8, 8, 4>(a, nblocks, st);
    case 384: return launch_corr_t<384, 8, 256, 6, 4, 4, 4>(a, nblocks, st);
    case 512: return launch_corr_t<512, 4, 256, 8, 8, 8>(a, nblocks, st);
    case 768: return launch_corr_t<768, 4, 256, 3, 4, 4, 4, 4>(a, nblocks, st);
    case 1024: return launch_corr_t<1024, 2, 256, 8, 8, 4, 4>(a, nblocks, st);
    case 1280: return launch_corr_t<1280, 4, 256, 5, 4, 4, 4, 4>(a, nblocks, st);
    case 2048: return launch_corr_t<2048, 1, 256, 8, 8, 8, 4>(a, nblocks, st);
    case 4096: return launch_corr_t<4096, 1, 512, 8, 8, 8, 8>(a, nblocks, st);
    default: return hipErrorInvalidValue;
    }
}

extern "C" hipError_t pzl_launch_harmsum(const PzlSumArgs *a, hipStream_t st)
{
    const int chunks = (a->corr_uselen + SUM_NT - 1) / SUM_NT;
    const long long nblk = (long long)a->nb * a->numw * a->numz * chunks;
    const dim3 g((unsigned)nblk), t(SUM_NT);
    switch (a->nstages) {
    case 1: hipLaunchKernelGGL(k_harmsum<1>, g, t, 0, st, *a); break;
    case 2: hipLaunchKernelGGL(k_harmsum<2>, g, t, 0, st, *a); break;
    case 3: hipLaunchKernelGGL(k_harmsum<3>, g, t, 0, st, *a); break;
    case 4: hipLaunchKernelGGL(k_harmsum<4>, g, t, 0, st, *a); break;
    case 5: hipLaunchKernelGGL(k_harmsum<5>, g, t, 0, st, *a); break;
    case 6: hipLaunchKernelGGL(k_harmsum<6>, g, t, 0, st, *a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
