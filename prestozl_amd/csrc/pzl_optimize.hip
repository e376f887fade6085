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
//   k_opt_volrow    per (candidate, harmonic, w, z): the 8x-oversampled
//                   response (five FFTs of the chirp for a w response) and its
//                   correlation with the spectrum at the 24 fine r steps the
//                   harmonic sum reads, as their powers
//   k_opt_sum       per candidate: the powers of its harmonics divided by
//                   their local powers, summed in the reference's order, and
//                   the first maximum of the volume
#include <hip/hip_runtime.h>

#include "pzl_fft.h"
#include "pzl_optimize.h"
#include "pzl_response.h"

#define OPT_NT 256
#define OPT_INTERP_MAXV 4   // interpolation responses: numkern <= 1024
#define OPT_SAMP_MAX 8192   // chirp samples of a w response
#define OPT_RESP_SAMP (PZL_OPT_RESP_FFT_MAX / OPT_NT)  // chirp samples per thread, residue FFTs

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

// Unnormalised FFT (SIGN -1 forward, +1 inverse) of n = 2^k <= PZL_OPT_RESP_FFT_MAX
// points in LDS: Stockham passes of radix 4 (one of radix 2 last when k is
// odd), each staging its butterflies in registers between two barriers, with
// twiddles exp(SIGN 2 pi i j / (Ns R)) read from tw[j] = exp(-2 pi i j / PZL_OPT_TWN).
// Every thread of the workgroup must call; ends with a barrier.
template <int SIGN>
__device__ void opt_fft_pow2(double2 *buf, int n, const double2 *__restrict__ tw, int tid)
{
    constexpr int B4 = PZL_OPT_RESP_FFT_MAX / 4 / OPT_NT, B2 = PZL_OPT_RESP_FFT_MAX / 2 / OPT_NT;
    int Ns = 1;
    while (Ns < n) {
        if ((n / Ns) % 4 == 0) {
            const int nbf = n >> 2, step = PZL_OPT_TWN / (Ns * 4);
            double2 v[B4][4];
#pragma unroll
            for (int b = 0; b < B4; b++) {
                const int j = tid + b * OPT_NT;
                if (j < nbf) {
                    const int jm = j & (Ns - 1);
                    v[b][0] = buf[j];
#pragma unroll
                    for (int r = 1; r < 4; r++) {
                        double2 t = tw[jm * r * step];
                        if (SIGN > 0) t.y = -t.y;
                        v[b][r] = cmul(buf[j + r * nbf], t);
                    }
                }
            }
            __syncthreads();
#pragma unroll
            for (int b = 0; b < B4; b++) {
                const int j = tid + b * OPT_NT;
                if (j < nbf) {
                    dft4<SIGN>(v[b]);
                    const int jm = j & (Ns - 1), o = ((j - jm) << 2) + jm;
#pragma unroll
                    for (int r = 0; r < 4; r++) buf[o + r * Ns] = v[b][r];
                }
            }
            __syncthreads();
            Ns <<= 2;
        } else {
            const int nbf = n >> 1, step = PZL_OPT_TWN / (Ns * 2);
            double2 v[B2][2];
#pragma unroll
            for (int b = 0; b < B2; b++) {
                const int j = tid + b * OPT_NT;
                if (j < nbf) {
                    const int jm = j & (Ns - 1);
                    double2 t = tw[jm * step];
                    if (SIGN > 0) t.y = -t.y;
                    v[b][0] = buf[j];
                    v[b][1] = cmul(buf[j + nbf], t);
                }
            }
            __syncthreads();
#pragma unroll
            for (int b = 0; b < B2; b++) {
                const int j = tid + b * OPT_NT;
                if (j < nbf) {
                    dft2<SIGN>(v[b][0], v[b][1]);
                    const int jm = j & (Ns - 1), o = ((j - jm) << 1) + jm;
                    buf[o] = v[b][0];
                    buf[o + Ns] = v[b][1];
                }
            }
            __syncthreads();
            Ns <<= 1;
        }
    }
}

// tw[j] = exp(-2 pi i j / PZL_OPT_TWN)
__global__ __launch_bounds__(OPT_NT) void k_opt_twiddles(double2 *__restrict__ tw)
{
    const int j = blockIdx.x * OPT_NT + threadIdx.x;
    if (j < PZL_OPT_TWN) {
        double s, c;
        sincospi(-2.0 * (double)j / (double)PZL_OPT_TWN, &s, &c);
        tw[j] = double2{c, s};
    }
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

// One (candidate harmonic, w, z) row of corr_rzw_vol (rzwinterp.c:86-139)
// at the r steps the harmonic sum reads.  The reference correlates the data
// segment (fftlen / 8 bins from lobin - khw, spread by 8) with the numkern =
// 16 khw point response gen_w_response(0, 8, z, w, numkern) through FFTs of
// fftlen points (complex_corr_conv CORR, correlations.c:56-64) and keeps the
// outputs 8 khw + e, 0 <= e < numgoodbins, where the circular wrap does not
// reach; output e is the direct sum of 2 khw terms
//   out(e) = sum_j data[lobin - khw + b0 + j] conj(resp[8 (b0 + j) - e]),
//   b0 = ceil(e / 8), 0 <= j < 2 khw,
// so the row evaluates only the outputs read: the 24 fine r steps of the sum
// and the row's last PZL_OPT_TAIL (a negative step index reads the end of the
// preceding row of the reference's contiguous volume).  Their float powers
// follow maximize_rzw.c:144-155.
//
// The w response (|w| >= 1e-4) is the numkern bins from k0 = 8 fbar - 8 khw
// (fbar = npts / 4) of the 8x zero-padded real FFT of npts float chirp
// samples s[t] (responses.c:333-418).  With k = 8 q + p,
//   X[k] = sum_t s[t] exp(-2 pi i t k / (8 npts)) = Y_p[q],
//   Y_p = FFT_npts(s[t] exp(-2 pi i t p / (8 npts))),
// and s being real, X[8 q + 8 - p] = conj(Y_p[npts - 1 - q]): five npts-point
// FFTs give every residue.  Output e reads residue p = -e mod 8 at q = fbar -
// khw + j (mod npts: X is periodic in k; where k0 < 0 the reference copies from
// before its array).  The z and r responses are evaluated per value into LDS.
// The LDS holds ldsn double2; a row needing more sets *err and writes nothing.
__global__ __launch_bounds__(OPT_NT, 2) void k_opt_volrow(const float2 *__restrict__ data, int numdata,
                                                       const PzlVolJob *__restrict__ jobs,
                                                       const double2 *__restrict__ tw, int ldsn,
                                                       float *__restrict__ powv, float *__restrict__ tail,
                                                       int *__restrict__ err, int nz, int nzw)
{
    extern __shared__ double2 sbuf[];
    constexpr int NOUT = PZL_OPT_NUMR + PZL_OPT_TAIL;  // 32 outputs per row
    constexpr int GL = 32, NG = OPT_NT / GL;           // lanes per output, outputs in flight
    __shared__ int outs[NOUT];      // e of each output slot, -1 if not needed
    __shared__ int lst[5][NOUT];    // output slots of each residue pass
    __shared__ int lcnt[5];
    const int tid = threadIdx.x, lane = tid % GL, grp = tid / GL;
    const int job = blockIdx.x / nzw, ij = blockIdx.x % nzw;
    const int wi = ij / nz, zj = ij % nz;
    const PzlVolJob &J = jobs[job];
    const int khw = J.khw, beginbin = J.lobin - khw, nterm = 2 * khw;
    const double z = J.zlo + zj * J.dz, w = J.wlo + wi * J.dw;
    const int numkern = 2 * PZL_OPT_NB * khw;
    const bool wresp = fabs(w) >= 1e-4;
    if (tid < NOUT) {
        int e;
        if (tid < PZL_OPT_NUMR) e = J.rind[tid] >= 0 ? J.rind[tid] : -1;
        else e = J.numgoodbins - PZL_OPT_TAIL + (tid - PZL_OPT_NUMR);
        outs[tid] = e;
    }
    __syncthreads();
    if (tid == 0) {
        for (int p = 0; p < 5; p++) lcnt[p] = 0;
        for (int o = 0; o < NOUT; o++) {
            const int e = outs[o];
            if (e < 0) continue;
            const int c = (-e) & (PZL_OPT_NB - 1);
            const int p = wresp ? (c <= PZL_OPT_NB / 2 ? c : PZL_OPT_NB - c) : 0;
            lst[p][lcnt[p]++] = o;
        }
    }
    // out(e) from the response values in sbuf: pass p of the w response holds
    // Y_p (mirror classes read npts - 1 - q), the z and r responses hold resp[m]
    auto emit = [&](int p, int npts) {
        const int cnt = lcnt[p];
        for (int base = 0; base < cnt; base += NG) {
            const int i = base + grp;
            double ar = 0.0, ai = 0.0;
            int o = -1;
            if (i < cnt) {
                o = lst[p][i];
                const int e = outs[o], b0 = (e + PZL_OPT_NB - 1) / PZL_OPT_NB;
                const bool mirror = wresp && ((-e) & (PZL_OPT_NB - 1)) != p;
                const int q0 = npts / 4 - khw;
                for (int j = lane; j < nterm; j += GL) {
                    const int b = beginbin + b0 + j;
                    const float2 d = (b >= 0 && b < numdata) ? data[b] : float2{0.f, 0.f};
                    double yr, yi;  // conj of the response value
                    if (wresp) {
                        const int q = (q0 + j) & (npts - 1);
                        const double2 y = sbuf[mirror ? npts - 1 - q : q];
                        yr = y.x;
                        yi = mirror ? y.y : -y.y;
                    } else {
                        const double2 y = sbuf[PZL_OPT_NB * (b0 + j) - e];
                        yr = y.x;
                        yi = -y.y;
                    }
                    ar += (double)d.x * yr - (double)d.y * yi;
                    ai += (double)d.x * yi + (double)d.y * yr;
                }
            }
#pragma unroll
            for (int off = GL / 2; off > 0; off >>= 1) {
                ar += __shfl_xor(ar, off, GL);
                ai += __shfl_xor(ai, off, GL);
            }
            if (o >= 0 && lane == 0) {
                const float fr = (float)ar, fi = (float)ai;
                const float pw = fr * fr + fi * fi;
                if (o < PZL_OPT_NUMR) powv[((long long)job * nzw + ij) * PZL_OPT_NUMR + o] = pw;
                else tail[((long long)job * nzw + ij) * PZL_OPT_TAIL + (o - PZL_OPT_NUMR)] = pw;
            }
        }
    };
    if (wresp) {
        const int npts = pzl_next2_to_n(6 * pzl_w_resp_halfwidth(z, w) + 200 + numkern / PZL_OPT_NB);
        if (npts > ldsn) {  // the launch sized the LDS for shorter chirps
            if (tid == 0) atomicOr(err, 1);
            return;
        }
        const double dt = 1.0 / (double)npts, amp = 2.0 * dt;
        const int fbar = npts / 4;
        const double f = (double)fbar + 0.0 - 0.5 * z + w / 12.0, fd = (z - 0.5 * w) / 2.0, fdd = w / 6.0;
        float smp[OPT_RESP_SAMP];
#pragma unroll
        for (int i = 0; i < OPT_RESP_SAMP; i++) {
            const int t = tid + i * OPT_NT;
            const double tt = t * dt;
            smp[i] = (t < npts) ? (float)(amp * cos(6.283185307179586477 * (tt * (tt * (tt * fdd + fd) + f)))) : 0.0f;
        }
        const int tstep = PZL_OPT_TWN / (PZL_OPT_NB * npts);
        for (int p = 0; p <= PZL_OPT_NB / 2; p++) {
            __syncthreads();  // the previous pass has read sbuf
#pragma unroll
            for (int i = 0; i < OPT_RESP_SAMP; i++) {
                const int t = tid + i * OPT_NT;
                if (t < npts) {
                    const double2 e = tw[t * p * tstep];
                    sbuf[t] = double2{(double)smp[i] * e.x, (double)smp[i] * e.y};
                }
            }
            __syncthreads();
            opt_fft_pow2<-1>(sbuf, npts, tw, tid);
            emit(p, npts);
        }
    } else {
        if (numkern > ldsn) {
            if (tid == 0) atomicOr(err, 1);
            return;
        }
        for (int m = tid; m < numkern; m += OPT_NT) sbuf[m] = pzl_zr_response_at(0.0, PZL_OPT_NB, z, numkern, m);
        __syncthreads();
        emit(0, 0);
    }
}

// Per candidate: powsum = sum over its harmonics, highest first, of
// pow / locpow (float, maximize_rzw.c:150-153), and the first maximum in
// (w, z, r) order with maxpow starting at 0 (:161-175).  The volume is one
// contiguous [w][z][r] block in the reference, so a negative r index reads
// the end of the preceding row; before the first row (memory outside the
// block there) the power is taken as 0.  With a local power of 1 and one
// harmonic this is max_rzw_arr's raw-power maximum (maximize_rzw.c:31-45).
// The grid has nzw (w, z) rows; bestv (if not null) gets the maximum.
__global__ __launch_bounds__(OPT_NT) void k_opt_sum(const float *__restrict__ powv,
                                                    const float *__restrict__ tail,
                                                    const PzlVolJob *__restrict__ jobs,
                                                    const float *__restrict__ locpow,
                                                    const int *__restrict__ job0, const int *__restrict__ nharm,
                                                    int *__restrict__ best, float *__restrict__ bestv, int nzw)
{
    const int NZW = nzw, NCELL = nzw * PZL_OPT_NUMR;
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
    if (tid == 0) {
        best[c] = rv[0] > 0.0f ? ri[0] : 0;
        if (bestv) bestv[c] = rv[0];
    }
}

extern "C" hipError_t pzl_launch_opt_interp(const float2 *data, int numdata, const PzlInterpJob *jobs, int njobs,
                                            float2 *ans, hipStream_t st)
{
    if (njobs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_opt_interp, dim3(njobs), dim3(OPT_NT), 0, st, data, numdata, jobs, ans);
    return hipGetLastError();
}

extern "C" hipError_t pzl_launch_opt_twiddles(double2 *tw, hipStream_t st)
{
    hipLaunchKernelGGL(k_opt_twiddles, dim3(PZL_OPT_TWN / OPT_NT), dim3(OPT_NT), 0, st, tw);
    return hipGetLastError();
}

extern "C" hipError_t pzl_launch_opt_volume(const float2 *data, int numdata, const PzlVolJob *jobs, int njobs,
                                            int ldsn, int nz, int nw, const double2 *tw, float *powv, float *tail,
                                            int *err, hipStream_t st)
{
    if (njobs <= 0) return hipSuccess;
    if (ldsn <= 0 || ldsn > PZL_OPT_RESP_FFT_MAX || nz < 1 || nw < 1 || nz > PZL_OPT_NUMZ || nw > PZL_OPT_NUMW)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_opt_volrow, dim3(njobs * nz * nw), dim3(OPT_NT), sizeof(double2) * (size_t)ldsn, st, data,
                       numdata, jobs, tw, ldsn, powv, tail, err, nz, nz * nw);
    return hipGetLastError();
}

extern "C" hipError_t pzl_launch_opt_sum(const float *powv, const float *tail, const PzlVolJob *jobs,
                                         const float *locpow, const int *job0, const int *nharm, int ncand,
                                         int nzw, int *best, float *bestv, hipStream_t st)
{
    if (ncand <= 0) return hipSuccess;
    if (nzw < 1 || nzw > PZL_OPT_NUMZ * PZL_OPT_NUMW) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_opt_sum, dim3(ncand), dim3(OPT_NT), 0, st, powv, tail, jobs, locpow, job0, nharm, best,
                       bestv, nzw);
    return hipGetLastError();
}
