// pzl_optimize.hip -- candidate optimisation of accelsearch_CPU2 on the GPU:
// the harmonic-summed, oversampled F/Fdot/Fdotdot volume around each
// candidate and the Fourier interpolations of its local power and peak
// derivatives (optimize_accelcand, accel_utils.c:573-657; max_rzw_arr_harmonics,
// maximize_rzw.c:85-191; corr_rzw_vol / rzw_interp, rzwinterp.c:3-229;
// get_localpower3d / get_derivs3d, characteristics.c:75-191).
//
//   k_opt_interp    one point interpolation per workgroup: the response at
//                   the point's fractional offset, then its correlation with
//                   2 khw bins of the spectrum
//   k_opt_datafft   per (candidate, harmonic): the spectrum segment of the
//                   volume, spread by 8 and FFT'd
//   k_opt_volrow    per (candidate, harmonic, w, z): the 8x-oversampled
//                   response, placed and FFT'd, times the conjugate of the
//                   segment's transform, inverse FFT, and the powers of the 24
//                   fine r steps the harmonic sum reads
//   k_opt_sum       per candidate: the powers of its harmonics divided by
//                   their local powers, summed in the reference's order, and
//                   the first maximum of the volume
#include <hip/hip_runtime.h>

#include "pzl_fft.h"
#include "pzl_optimize.h"
#include "pzl_response.h"

#define OPT_NT 256
#define OPT_INTERP_MAXV 4   // interpolation responses: numkern <= 1024
#define OPT_VOL_MAXV 16     // volume responses: numkern <= 4096
#define OPT_SAMP_MAX 8192   // chirp samples of a w response

template <typename T> __device__ T block_sum(T v, T *red)
{
    const int tid = threadIdx.x;
    red[tid] = v;
    __syncthreads();
    for (int s = OPT_NT / 2; s > 0; s >>= 1) {
        if (tid < s) red[tid] += red[tid + s];
        __syncthreads();
    }
    const T r = red[0];
    __syncthreads();
    return r;
}

// rzw_interp (rzwinterp.c:145-229) and rz_interp (rzinterp.c:144-230): the
// response's values sum against data[int(r) - khw + m], m < 2 khw, clipped to
// the data.
__global__ __launch_bounds__(OPT_NT) void k_opt_interp(const float2 *__restrict__ data, int numdata,
                                                       const PzlInterpJob *__restrict__ jobs,
                                                       float2 *__restrict__ ans)
{
    __shared__ float samp[OPT_SAMP_MAX];
    __shared__ double red[OPT_NT];
    const int tid = threadIdx.x;
    const PzlInterpJob J = jobs[blockIdx.x];
    const double r = J.r, z = J.z, w = J.w;
    const int khw = J.khw;
    if (r > numdata - 1.0 || r < 0.0) {
        if (tid == 0) ans[blockIdx.x] = float2{0.f, 0.f};
        return;
    }
    double dint;
    const double frac = modf(r, &dint);
    const int intfreq = (int)dint;
    if (fabs(w) < 1E-4 && fabs(z) < 1E-4) {
        // rz_interp's shortcut at (nearly) integer frequencies
        if (frac < 1E-5) {
            if (tid == 0) ans[blockIdx.x] = data[intfreq];
            return;
        }
        if ((1.0 - frac) < 1E-5) {
            if (tid == 0) ans[blockIdx.x] = data[intfreq + 1];
            return;
        }
    }
    const int numkern = 2 * khw;
    double2 resp[OPT_INTERP_MAXV];
    // |w| < 1e-4 gives gen_z_response (and that gen_r_response for |z| < 1e-4)
    pzl_gen_response<OPT_NT, OPT_INTERP_MAXV>(frac, 1, z, w, numkern, tid, samp, resp);
    int lodata = intfreq - khw, loresp = 0;
    if (lodata < 0) {
        loresp = -lodata;
        lodata = 0;
    }
    const int hidata = intfreq + khw - 1;
    const int hiresp = (hidata > numdata - 1) ? numkern - hidata + numdata - 1 : numkern;
    double ar = 0.0, ai = 0.0;
#pragma unroll
    for (int i = 0; i < OPT_INTERP_MAXV; i++) {
        const int m = tid + i * OPT_NT;
        if (m >= loresp && m < hiresp) {
            const float2 d = data[intfreq - khw + m];
            ar += (double)d.x * resp[i].x + (double)d.y * resp[i].y;
            ai += (double)d.y * resp[i].x - (double)d.x * resp[i].y;
        }
    }
    ar = block_sum(ar, red);
    ai = block_sum(ai, red);
    if (tid == 0) ans[blockIdx.x] = float2{(float)ar, (float)ai};
}

// corr_complex with RAW data (corr_routines.c:140-190): numbins = fftlen / 8
// bins from lobin - khw (zeros outside the data), spread by 8, forward FFT.
__global__ __launch_bounds__(OPT_NT) void k_opt_datafft(const float2 *__restrict__ data, int numdata,
                                                        const PzlVolJob *__restrict__ jobs,
                                                        double2 *__restrict__ dfft, int stride)
{
    extern __shared__ double2 sbuf[];
    const int tid = threadIdx.x;
    const PzlVolJob J = jobs[blockIdx.x];
    const int fftlen = J.fftlen, numbins = fftlen / PZL_OPT_NB, beginbin = J.lobin - J.khw;
    for (int k = tid; k < fftlen; k += OPT_NT) sbuf[k] = double2{0.0, 0.0};
    __syncthreads();
    for (int b = tid; b < numbins; b += OPT_NT) {
        const int idx = beginbin + b;
        if (idx >= 0 && idx < numdata) {
            const float2 x = data[idx];
            sbuf[b * PZL_OPT_NB] = double2{(double)x.x, (double)x.y};
        }
    }
    __syncthreads();
    lds_fft_f64<-1>(sbuf, fftlen, tid, OPT_NT);
    double2 *out = dfft + (long long)blockIdx.x * stride;
    for (int k = tid; k < fftlen; k += OPT_NT) out[k] = sbuf[k];
}

// One (candidate harmonic, w, z) row of corr_rzw_vol (rzwinterp.c:86-139):
// gen_w_response(0, 8, z, w, 16 khw), place_complex_kernel, forward FFT,
// complex_corr_conv CORR (correlations.c:56-64: data x conj(kernel) / fftlen),
// inverse FFT, chop 8 khw points; then the float powers (maximize_rzw.c:144-155)
// of the rows' 24 fine r steps.
__global__ __launch_bounds__(OPT_NT) void k_opt_volrow(const PzlVolJob *__restrict__ jobs,
                                                       const double2 *__restrict__ dfft, int stride,
                                                       float *__restrict__ powv, float *__restrict__ tail)
{
    extern __shared__ double2 sbuf[];
    const int tid = threadIdx.x;
    constexpr int NZW = PZL_OPT_NUMZ * PZL_OPT_NUMW;
    const int job = blockIdx.x / NZW, ij = blockIdx.x % NZW;
    const int wi = ij / PZL_OPT_NUMZ, zj = ij % PZL_OPT_NUMZ;
    const PzlVolJob &J = jobs[job];
    const int fftlen = J.fftlen, khw = J.khw;
    const double z = J.zlo + zj * J.dz, w = J.wlo + wi * J.dw;
    const int numkern = 2 * PZL_OPT_NB * khw, half = numkern / 2;
    double2 resp[OPT_VOL_MAXV];
    pzl_gen_response<OPT_NT, OPT_VOL_MAXV>(0.0, PZL_OPT_NB, z, w, numkern, tid, (float *)sbuf, resp);
    for (int k = tid; k < fftlen; k += OPT_NT) sbuf[k] = double2{0.0, 0.0};
    __syncthreads();
#pragma unroll
    for (int i = 0; i < OPT_VOL_MAXV; i++) {
        const int m = tid + i * OPT_NT;
        if (m < numkern) sbuf[(m >= half) ? (m - half) : (fftlen - half + m)] = resp[i];
    }
    __syncthreads();
    lds_fft_f64<-1>(sbuf, fftlen, tid, OPT_NT);
    const double2 *D = dfft + (long long)job * stride;
    const double norm = 1.0 / fftlen;
    for (int k = tid; k < fftlen; k += OPT_NT) {
        const double2 d = D[k], kk = sbuf[k];
        sbuf[k] = double2{(d.x * kk.x + d.y * kk.y) * norm, (d.y * kk.x - kk.y * d.x) * norm};
    }
    __syncthreads();
    lds_fft_f64<1>(sbuf, fftlen, tid, OPT_NT);
    // the r steps of the sum, and the row's last PZL_OPT_TAIL powers: the sum
    // reads a negative step index in the preceding row of the contiguous volume
    int e = -1;
    if (tid < PZL_OPT_NUMR) e = J.rind[tid];
    else if (tid < PZL_OPT_NUMR + PZL_OPT_TAIL) e = J.numgoodbins - PZL_OPT_TAIL + (tid - PZL_OPT_NUMR);
    if (e >= 0) {
        const double2 v = sbuf[PZL_OPT_NB * khw + e];
        const float ar = (float)v.x, ai = (float)v.y;
        const float a2 = ar * ar, b2 = ai * ai;
        if (tid < PZL_OPT_NUMR) powv[((long long)job * NZW + ij) * PZL_OPT_NUMR + tid] = a2 + b2;
        else tail[((long long)job * NZW + ij) * PZL_OPT_TAIL + (tid - PZL_OPT_NUMR)] = a2 + b2;
    }
}

// Per candidate: powsum = sum over its harmonics, highest first, of
// pow / locpow (float, maximize_rzw.c:150-153), and the first maximum in
// (w, z, r) order with maxpow starting at 0 (:161-175).  The volume is one
// contiguous [w][z][r] block in the reference, so a negative r index reads
// the end of the preceding row; before the first row (memory outside the
// block there) the power is taken as 0.
__global__ __launch_bounds__(OPT_NT) void k_opt_sum(const float *__restrict__ powv,
                                                    const float *__restrict__ tail,
                                                    const PzlVolJob *__restrict__ jobs,
                                                    const float *__restrict__ locpow,
                                                    const int *__restrict__ job0, const int *__restrict__ nharm,
                                                    int *__restrict__ best)
{
    constexpr int NZW = PZL_OPT_NUMZ * PZL_OPT_NUMW;
    constexpr int NCELL = PZL_OPT_NUMZ * PZL_OPT_NUMW * PZL_OPT_NUMR;
    __shared__ float rv[OPT_NT];
    __shared__ int ri[OPT_NT];
    const int tid = threadIdx.x, c = blockIdx.x;
    const int j0 = job0[c], nh = nharm[c];
    float bv = 0.0f;
    int bi = 0;
    for (int cell = tid; cell < NCELL; cell += OPT_NT) {
        const int row = cell / PZL_OPT_NUMR, kk = cell % PZL_OPT_NUMR;
        float acc = 0.0f;
        for (int h = 0; h < nh; h++) {
            const int job = j0 + h;
            const int ri = jobs[job].rind[kk];
            float p = 0.0f;
            if (ri >= 0) p = powv[(long long)job * NCELL + cell];
            else if (row > 0) p = tail[((long long)job * NZW + row - 1) * PZL_OPT_TAIL + PZL_OPT_TAIL + ri];
            acc += p / locpow[job];
        }
        if (acc > bv) {  // cells of a thread ascend: strict > keeps the first
            bv = acc;
            bi = cell;
        }
    }
    rv[tid] = bv;
    ri[tid] = bi;
    __syncthreads();
    for (int s = OPT_NT / 2; s > 0; s >>= 1) {
        if (tid < s) {
            const float ov = rv[tid + s];
            const int oi = ri[tid + s];
            if (ov > rv[tid] || (ov == rv[tid] && oi < ri[tid] && ov > 0.0f)) {
                rv[tid] = ov;
                ri[tid] = oi;
            }
        }
        __syncthreads();
    }
    if (tid == 0) best[c] = rv[0] > 0.0f ? ri[0] : 0;
}

extern "C" hipError_t pzl_launch_opt_interp(const float2 *data, int numdata, const PzlInterpJob *jobs, int njobs,
                                            float2 *ans, hipStream_t st)
{
    if (njobs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_opt_interp, dim3(njobs), dim3(OPT_NT), 0, st, data, numdata, jobs, ans);
    return hipGetLastError();
}

extern "C" hipError_t pzl_launch_opt_volume(const float2 *data, int numdata, const PzlVolJob *jobs, int njobs,
                                            int max_fftlen, double2 *dfft, float *powv, float *tail,
                                            hipStream_t st)
{
    if (njobs <= 0) return hipSuccess;
    const size_t smem = sizeof(double2) * (size_t)max_fftlen;
    hipLaunchKernelGGL(k_opt_datafft, dim3(njobs), dim3(OPT_NT), smem, st, data, numdata, jobs, dfft, max_fftlen);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_opt_volrow, dim3(njobs * PZL_OPT_NUMZ * PZL_OPT_NUMW), dim3(OPT_NT), smem, st, jobs,
                       dfft, max_fftlen, powv, tail);
    return hipGetLastError();
}

extern "C" hipError_t pzl_launch_opt_sum(const float *powv, const float *tail, const PzlVolJob *jobs,
                                         const float *locpow, const int *job0, const int *nharm, int ncand,
                                         int *best, hipStream_t st)
{
    if (ncand <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_opt_sum, dim3(ncand), dim3(OPT_NT), 0, st, powv, tail, jobs, locpow, job0, nharm, best);
    return hipGetLastError();
}
