// pzl_dedisp.hip -- prepsubband's dedispersion on the GPU (the producer of
// the DM trials the search reads): DownsampDataArray and getDedisp
// (prepsubband_kernels.cu:23-90), driven by prepsubband_cu.c:1143-1164.
//
//   k_dd_downsamp   ds[t][s] = (sum_{j < bs} in[t bs + j][s]) / bs, one thread
//                   per output, subbands fastest (coalesced)
//   k_dd_transpose  the previous and the current block, [t][nsub] each, into
//                   one subband-major window S[s][0 .. 2 worklen)
//   k_dd_sum        out[dm][x] = sum_s S[s][x + offsets[dm][s]], subbands in
//                   ascending order (the reference's float summation order),
//                   8 DMs per workgroup sharing each subband's window in
//                   LDS, output tiles mapped XCD by XCD (k_dd_sum_direct:
//                   one DM per workgroup, for blocks of many subbands)
//
// Both reference kernels read time-major data; the sums are formed in the same
// order from the same values, so results are bit-identical.
#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "pzl_accel.h"
#include "pzl_plan.h"

#define DD_NT 256
#define DD_TILE 32
#define DD_XPT 4      // outputs per thread in k_dd_sum, DD_NT apart
#define DD_DG 8       // DMs per k_dd_sum workgroup
#define DD_SPREAD 256  // widest offset spread of a group's subband served from LDS
#define DD_NSUB 256    // most subbands whose offsets k_dd_sum keeps in LDS
#define DD_XPD 4       // consecutive outputs per thread in k_dd_sum_direct

namespace {

__global__ __launch_bounds__(DD_NT) void k_dd_downsamp(float *__restrict__ out, const float *__restrict__ in,
                                                       int worklen, int bs, int nsub)
{
    const long long g = (long long)blockIdx.x * DD_NT + threadIdx.x;
    if (g >= (long long)worklen * nsub) return;
    const int t = (int)(g / nsub), s = (int)(g - (long long)t * nsub);
    const float *p = in + (long long)t * bs * nsub + s;
    float v = 0.0f;
    for (int j = 0; j < bs; j++) v += p[(long long)j * nsub];
    out[g] = v / (float)bs;
}

// S[s][t] = last[t][s] for t < worklen, current[t - worklen][s] after, by
// 32 x 32 tiles through LDS.
__global__ __launch_bounds__(DD_TILE * 8) void k_dd_transpose(float *__restrict__ S, const float *__restrict__ last,
                                                              const float *__restrict__ cur, int worklen, int nsub)
{
    __shared__ float tile[DD_TILE][DD_TILE + 1];
    const int t0 = blockIdx.x * DD_TILE, s0 = blockIdx.y * DD_TILE;
    const int tx = threadIdx.x % DD_TILE, ty = threadIdx.x / DD_TILE;  // 32 x 8
    for (int r = ty; r < DD_TILE; r += 8) {
        const int t = t0 + r, s = s0 + tx;
        float v = 0.0f;
        if (t < 2 * worklen && s < nsub)
            v = t < worklen ? last[(long long)t * nsub + s] : cur[(long long)(t - worklen) * nsub + s];
        tile[r][tx] = v;
    }
    __syncthreads();
    for (int r = ty; r < DD_TILE; r += 8) {
        const int s = s0 + r, t = t0 + tx;
        if (s < nsub && t < 2 * worklen) S[(long long)s * 2 * worklen + t] = tile[tx][r];
    }
}

// out[dm][x] for DD_NT * DD_XPD consecutive x of one DM per workgroup, for
// blocks of more than DD_NSUB subbands: every DM reads its subbands' windows
// directly (one 16-byte load per thread and subband), the 1-D grid mapped XCD
// by XCD as below.  Subbands in ascending order; a bad offset sets *err and
// that DM's outputs are left unwritten.
__global__ __launch_bounds__(DD_NT) void k_dd_sum_direct(float *__restrict__ out, long long out_stride,
                                                  const float *__restrict__ S, const int *__restrict__ offsets,
                                                  int worklen, int nsub, int ntiles, int numdms,
                                                  int *__restrict__ err)
{
    const int w = blockIdx.x, nb = gridDim.x;
    int tile, dm;
    if (ntiles % 8 == 0 && nb % 8 == 0) {
        const int per = ntiles / 8, j = w / 8;  // the j-th workgroup of XCD w % 8
        tile = (w % 8) * per + j % per;
        dm = j / per;
    } else {
        tile = w % ntiles;
        dm = w / ntiles;
    }
    if (dm >= numdms) return;
    // DD_XPD consecutive outputs per thread, read with 16-byte loads per
    // subband (4-byte aligned: the offsets are arbitrary)
    const int x0 = tile * DD_NT * DD_XPD + threadIdx.x * DD_XPD;
    const bool full = x0 + DD_XPD <= worklen;
    const int *off = offsets + (long long)dm * nsub;
    const long long w2 = 2LL * worklen;
    float v[DD_XPD];
#pragma unroll
    for (int k = 0; k < DD_XPD; k++) v[k] = 0.0f;
    bool bad = false;
    for (int s = 0; s < nsub; s++) {
        const int o = off[s];
        bad |= (o < 0 || o > worklen);
        const float *row = S + s * w2 + (bad ? 0 : o) + x0;
        if (full) {
#pragma unroll
            for (int k = 0; k < DD_XPD; k += 4) {
                float4 q;
                __builtin_memcpy(&q, row + k, sizeof(q));
                v[k] += q.x;
                v[k + 1] += q.y;
                v[k + 2] += q.z;
                v[k + 3] += q.w;
            }
        } else {
#pragma unroll
            for (int k = 0; k < DD_XPD; k++)
                if (x0 + k < worklen) v[k] += row[k];
        }
    }
    if (bad) {
        if (threadIdx.x == 0) atomicOr(err, 1);
        return;
    }
#pragma unroll
    for (int k = 0; k < DD_XPD; k++)
        if (x0 + k < worklen) out[(long long)dm * out_stride + x0 + k] = v[k];
}

// out[dm][x] for DD_DG consecutive DMs and DD_NT * DD_XPT outputs per
// workgroup (thread t takes x0 + k DD_NT + t, k < DD_XPT: consecutive lanes,
// consecutive words).  Neighbouring DMs' offsets differ little, so per subband
// the union of the group's windows, S[s][x0 + lo .. x0 + X + hi) with lo / hi
// the group's smallest / largest offset, is read from global memory once into
// LDS (two buffers alternate: one barrier per subband) and every DM of the
// group reads its own shifted part of it; the next subband's window is read
// into registers while one is summed.  A subband whose offsets spread wider
// than DD_SPREAD is read from global memory directly.  The block pair's
// window is larger than one XCD's L2, and workgroups are dispatched to the 8
// XCDs in turn: the 1-D grid is mapped so that each XCD takes a contiguous
// range of output tiles for every DM group, and its part of the window stays
// in its L2.  Subbands are summed in ascending order (the reference's order
// and values: the same bits).  An offset outside [0, worklen] sets *err and
// that DM's outputs are left unwritten.
__global__ __launch_bounds__(DD_NT) void k_dd_sum(float *__restrict__ out, long long out_stride,
                                                  const float *__restrict__ S, const int *__restrict__ offsets,
                                                  int worklen, int nsub, int ntiles, int numdms,
                                                  int *__restrict__ err)
{
    constexpr int X = DD_NT * DD_XPT, WIN = X + DD_SPREAD;
    __shared__ float win[2][WIN];
    const int w = blockIdx.x, nb = gridDim.x, tid = threadIdx.x;
    int tile, grp;
    if (ntiles % 8 == 0 && nb % 8 == 0) {
        const int per = ntiles / 8, j = w / 8;  // the j-th workgroup of XCD w % 8
        tile = (w % 8) * per + j % per;
        grp = j / per;
    } else {
        tile = w % ntiles;
        grp = w / ntiles;
    }
    const int dm0 = grp * DD_DG;
    if (dm0 >= numdms || nsub > DD_NSUB) return;  // the launch sends larger blocks to k_dd_sum_direct
    const int ndm = min(DD_DG, numdms - dm0);
    const int x0 = tile * X;
    const long long w2 = 2LL * worklen;
    float v[DD_DG][DD_XPT];
#pragma unroll
    for (int d = 0; d < DD_DG; d++)
#pragma unroll
        for (int k = 0; k < DD_XPT; k++) v[d][k] = 0.0f;
    // the group's offsets (a bad one as 0) and each subband's smallest and
    // largest, in LDS
    __shared__ int soff[DD_DG][DD_NSUB], slo[DD_NSUB], shi[DD_NSUB];
    __shared__ unsigned sbad;
    if (tid == 0) sbad = 0;
    __syncthreads();
    for (int i = tid; i < DD_DG * nsub; i += DD_NT) {
        const int d = i / nsub, s = i - d * nsub;
        int od = d < ndm ? offsets[(long long)(dm0 + d) * nsub + s] : 0;
        if (od < 0 || od > worklen) {
            atomicOr(&sbad, 1u << d);
            od = 0;
        }
        soff[d][s] = od;
    }
    __syncthreads();
    for (int s = tid; s < nsub; s += DD_NT) {
        int lo = worklen, hi = 0;
        for (int d = 0; d < ndm; d++) {
            lo = min(lo, soff[d][s]);
            hi = max(hi, soff[d][s]);
        }
        slo[s] = lo;
        shi[s] = hi;
    }
    __syncthreads();
    const unsigned bad = sbad;  // bit d: DM dm0 + d has an offset outside [0, worklen]
    // subband s + 1's window (LDS path) is fetched into registers while
    // subband s is summed: the global loads' latency overlaps the sums
    constexpr int NPRE = WIN / DD_NT;
    float pre[NPRE];
    int lon = 0, hin = 0;
    auto fetch = [&](int s) {
        lon = slo[s];
        hin = shi[s];
        if (hin - lon <= DD_SPREAD) {
            // positions x0 + lo + j, j < X + hi - lo (0 beyond the row's end,
            // which no written output reads)
            const int span = X + hin - lon;
            const long long base = (long long)x0 + lon;
            const float *row = S + s * w2;
#pragma unroll
            for (int i = 0; i < NPRE; i++) {
                const int j = tid + i * DD_NT;
                pre[i] = (j < span && base + j < w2) ? row[base + j] : 0.0f;
            }
        }
    };
    fetch(0);
    int buf = 0;
    for (int s = 0; s < nsub; s++) {
        const int lo = lon;
        const bool shared = hin - lon <= DD_SPREAD;
        if (shared) {
#pragma unroll
            for (int i = 0; i < NPRE; i++) win[buf][tid + i * DD_NT] = pre[i];
            __syncthreads();
        }
        if (s + 1 < nsub) fetch(s + 1);
        if (shared) {
#pragma unroll
            for (int d = 0; d < DD_DG; d++)
                if (d < ndm) {
                    const float *wd = &win[buf][tid + soff[d][s] - lo];
#pragma unroll
                    for (int k = 0; k < DD_XPT; k++) v[d][k] += wd[k * DD_NT];
                }
            buf ^= 1;
        } else {
            const float *row = S + s * w2;
#pragma unroll
            for (int d = 0; d < DD_DG; d++)
                if (d < ndm) {
                    const int od = soff[d][s];
#pragma unroll
                    for (int k = 0; k < DD_XPT; k++) {
                        const int x = x0 + k * DD_NT + tid;
                        if (x < worklen) v[d][k] += row[od + x];
                    }
                }
        }
    }
    if (bad && tid == 0) atomicOr(err, 1);
#pragma unroll
    for (int d = 0; d < DD_DG; d++) {
        if (d >= ndm || (bad >> d) & 1u) continue;
#pragma unroll
        for (int k = 0; k < DD_XPT; k++) {
            const int x = x0 + k * DD_NT + tid;
            if (x < worklen) out[(long long)(dm0 + d) * out_stride + x] = v[d][k];
        }
    }
}

struct DdWork {
    std::mutex mu;
    float *S = nullptr;
    long long cap = 0;
    int *err = nullptr;
    float *zero = nullptr;  // a block of zeros after the last one of a series
    long long zcap = 0;
};
std::mutex g_dd_mu;
std::vector<DdWork *> g_dd_work;

DdWork *dd_work(int device)
{
    std::lock_guard<std::mutex> lk(g_dd_mu);
    if ((int)g_dd_work.size() <= device) g_dd_work.resize(device + 1, nullptr);
    if (!g_dd_work[device]) g_dd_work[device] = new DdWork();
    return g_dd_work[device];
}

}  // namespace

using namespace pzl;

extern "C" {

int pzl_downsamp_data_array(float *currentdsdata_gpu, const float *currentdata_gpu, int worklen, int bs, int nsub)
{
    if (!currentdsdata_gpu || !currentdata_gpu) return fail(PZL_ERR_ARG, "downsamp: null data");
    if (worklen < 1 || bs < 1 || nsub < 1) return fail(PZL_ERR_ARG, "downsamp: worklen, bs, nsub must be >= 1");
    const long long total = (long long)worklen * nsub;
    if (total * bs >= (1LL << 40)) return fail(PZL_ERR_ARG, "downsamp: block too large");
    hipLaunchKernelGGL(k_dd_downsamp, dim3((unsigned)((total + DD_NT - 1) / DD_NT)), dim3(DD_NT), 0, nullptr,
                       currentdsdata_gpu, currentdata_gpu, worklen, bs, nsub);
    PZL_HIPCHK(hipGetLastError());
    PZL_HIPCHK(hipStreamSynchronize(nullptr));
    return 0;
}

// Enqueue the transpose and the sums of one block pair on the null stream
// (the workspace and the error flag are the caller's, under w->mu).
static int dedisp_enqueue(DdWork *w, const float *cur, const float *last, float *out, int worklen, int nsub,
                          int numdms, const int *offsets, long long out_stride)
{
    hipLaunchKernelGGL(k_dd_transpose, dim3((2 * worklen + DD_TILE - 1) / DD_TILE, (nsub + DD_TILE - 1) / DD_TILE),
                       dim3(DD_TILE * 8), 0, nullptr, w->S, last, cur, worklen, nsub);
    PZL_HIPCHK(hipGetLastError());
    const int ntiles = (worklen + DD_NT * DD_XPT - 1) / (DD_NT * DD_XPT);
    const int ngroups = (numdms + DD_DG - 1) / DD_DG;
    if (nsub <= DD_NSUB) {
        hipLaunchKernelGGL(k_dd_sum, dim3((unsigned)ntiles * ngroups), dim3(DD_NT), 0, nullptr, out, out_stride,
                           w->S, offsets, worklen, nsub, ntiles, numdms, w->err);
    } else {
        const int nt = (worklen + DD_NT * DD_XPD - 1) / (DD_NT * DD_XPD);
        hipLaunchKernelGGL(k_dd_sum_direct, dim3((unsigned)nt * numdms), dim3(DD_NT), 0, nullptr, out, out_stride,
                           w->S, offsets, worklen, nsub, nt, numdms, w->err);
    }
    PZL_HIPCHK(hipGetLastError());
    return 0;
}

// Argument checks and the workspace for blocks of (worklen, nsub).
static int dedisp_prepare(DdWork *w, int worklen, int nsub, int numdms, long long out_stride)
{
    if (worklen < 1 || nsub < 1 || numdms < 1 || (long long)numdms * ((worklen + 1023) / 1024) >= (1LL << 31) ||
        out_stride < worklen)
        return fail(PZL_ERR_ARG, "dedisp: worklen %d, nsub %d, numdms %d, row stride %lld (>= worklen)", worklen, nsub,
                    numdms, out_stride);
    if ((long long)nsub * 2 * worklen >= (1LL << 31)) return fail(PZL_ERR_ARG, "dedisp: block too large");
    const long long need = (long long)nsub * 2 * worklen;
    if (need > w->cap) {
        (void)hipFree(w->S);
        w->S = nullptr;
        w->cap = 0;
        PZL_HIPCHK(hipMalloc(&w->S, sizeof(float) * need));
        w->cap = need;
    }
    if (!w->err) PZL_HIPCHK(hipMalloc(&w->err, sizeof(int)));
    PZL_HIPCHK(hipMemsetAsync(w->err, 0, sizeof(int), nullptr));
    return 0;
}

static int dedisp_finish(DdWork *w, int worklen)
{
    int err = 0;
    PZL_HIPCHK(hipMemcpy(&err, w->err, sizeof(int), hipMemcpyDeviceToHost));
    if (err) return fail(PZL_ERR_ARG, "dedisp: an offset lies outside [0, worklen = %d]", worklen);
    return 0;
}

int pzl_get_dedisp_strided(const float *currentdsdata_gpu, const float *lastdsdata_gpu, float *outdata_gpu,
                           int worklen, int nsub, int numdms, const int *offsets_gpu, long long out_stride)
{
    if (!currentdsdata_gpu || !lastdsdata_gpu || !outdata_gpu || !offsets_gpu)
        return fail(PZL_ERR_ARG, "dedisp: null data");
    int device = 0;
    PZL_HIPCHK(hipGetDevice(&device));
    DdWork *w = dd_work(device);
    std::lock_guard<std::mutex> lk(w->mu);
    int rc = dedisp_prepare(w, worklen, nsub, numdms, out_stride);
    if (!rc)
        rc = dedisp_enqueue(w, currentdsdata_gpu, lastdsdata_gpu, outdata_gpu, worklen, nsub, numdms, offsets_gpu,
                            out_stride);
    return rc ? rc : dedisp_finish(w, worklen);
}

int pzl_get_dedisp(const float *currentdsdata_gpu, const float *lastdsdata_gpu, float *outdata_gpu, int worklen,
                   int nsub, int numdms, const int *offsets_gpu)
{
    return pzl_get_dedisp_strided(currentdsdata_gpu, lastdsdata_gpu, outdata_gpu, worklen, nsub, numdms, offsets_gpu,
                                  worklen);
}

int pzl_dedisperse(const float *subbands_gpu, long long nsamp, int nsub, int worklen, int numdms,
                   const int *offsets_gpu, float *out_gpu)
{
    if (!subbands_gpu || !out_gpu || !offsets_gpu) return fail(PZL_ERR_ARG, "dedisperse: null data");
    if (worklen < 1 || nsamp < worklen || nsamp % worklen)
        return fail(PZL_ERR_ARG, "dedisperse: %lld samples are not whole blocks of %d", nsamp, worklen);
    int device = 0;
    PZL_HIPCHK(hipGetDevice(&device));
    DdWork *w = dd_work(device);
    std::lock_guard<std::mutex> lk(w->mu);
    int rc = dedisp_prepare(w, worklen, nsub, numdms, nsamp);
    if (rc) return rc;
    const long long nblk = nsamp / worklen, blk = (long long)worklen * nsub;
    if (blk > w->zcap) {
        (void)hipFree(w->zero);
        w->zero = nullptr;
        w->zcap = 0;
        PZL_HIPCHK(hipMalloc(&w->zero, sizeof(float) * blk));
        PZL_HIPCHK(hipMemset(w->zero, 0, sizeof(float) * blk));
        w->zcap = blk;
    }
    // every block pair enqueued in order (the window workspace is reused in
    // stream order), one wait and one check at the end
    for (long long k = 0; k < nblk && rc == 0; k++) {
        const float *last = subbands_gpu + k * blk;
        const float *cur = k + 1 < nblk ? subbands_gpu + (k + 1) * blk : w->zero;
        rc = dedisp_enqueue(w, cur, last, out_gpu + k * worklen, worklen, nsub, numdms, offsets_gpu, nsamp);
    }
    return rc ? rc : dedisp_finish(w, worklen);
}

}  // extern "C"
