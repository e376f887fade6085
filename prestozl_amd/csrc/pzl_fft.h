// pzl_fft.h -- small-radix DFT butterflies and a block-cooperative LDS FFT
// for gfx950.  Conventions: X[k] = sum_j x[j] exp(SIGN * 2 pi i j k / n),
// unnormalised (SIGN = -1 forward, +1 inverse), as FFTW's fftwf_plan_dft_1d
// which the reference calls (fftcalls.c:71-135, accel_utils.c:1222-1269).
#pragma once

#include <hip/hip_runtime.h>

#define PZL_FFT_HD __host__ __device__ __forceinline__

template <typename T> struct cpx_t;
template <> struct cpx_t<float> { typedef float2 type; };
template <> struct cpx_t<double> { typedef double2 type; };

template <typename C> PZL_FFT_HD C cadd(C a, C b) { return C{a.x + b.x, a.y + b.y}; }
template <typename C> PZL_FFT_HD C csub(C a, C b) { return C{a.x - b.x, a.y - b.y}; }
template <typename C> PZL_FFT_HD C cmul(C a, C b)
{
    return C{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// multiply by SIGN * i
template <int SIGN, typename C> PZL_FFT_HD C cmul_si(C a)
{
    return SIGN > 0 ? C{-a.y, a.x} : C{a.y, -a.x};
}

// In-register DFTs of size R on v[0..R-1], outputs in natural order.
template <int SIGN, typename C> PZL_FFT_HD void dft2(C &a, C &b)
{
    C t = a;
    a = cadd(t, b);
    b = csub(t, b);
}

template <int SIGN, typename C> PZL_FFT_HD void dft4(C *v)
{
    C s0 = cadd(v[0], v[2]), d0 = csub(v[0], v[2]);
    C s1 = cadd(v[1], v[3]), d1 = cmul_si<SIGN>(csub(v[1], v[3]));
    v[0] = cadd(s0, s1);
    v[2] = csub(s0, s1);
    v[1] = cadd(d0, d1);
    v[3] = csub(d0, d1);
}

template <int SIGN, typename C> PZL_FFT_HD void dft8(C *v)
{
    typedef decltype(v[0].x) R;
    const R h = (R)0.70710678118654752440;
    C e[4] = {v[0], v[2], v[4], v[6]};
    C o[4] = {v[1], v[3], v[5], v[7]};
    dft4<SIGN>(e);
    dft4<SIGN>(o);
    // o[k] *= exp(SIGN 2 pi i k / 8)
    // k=1: (1 + SIGN i) h ; k=2: SIGN i ; k=3: (-1 + SIGN i) h
    C o1 = o[1], o3 = o[3];
    o[1] = C{(o1.x - (R)SIGN * o1.y) * h, ((R)SIGN * o1.x + o1.y) * h};
    o[2] = cmul_si<SIGN>(o[2]);
    o[3] = C{(-o3.x - (R)SIGN * o3.y) * h, ((R)SIGN * o3.x - o3.y) * h};
#pragma unroll
    for (int k = 0; k < 4; k++) {
        v[k] = cadd(e[k], o[k]);
        v[k + 4] = csub(e[k], o[k]);
    }
}

template <int SIGN, typename C> PZL_FFT_HD void dft3(C *v)
{
    typedef decltype(v[0].x) R;
    const R c1 = (R)-0.5, s1 = (R)SIGN * (R)0.86602540378443864676;
    C t1 = cadd(v[1], v[2]), t2 = csub(v[1], v[2]);
    C m = C{v[0].x + c1 * t1.x, v[0].y + c1 * t1.y};
    C j = C{-s1 * t2.y, s1 * t2.x};  // i * s1 * t2
    v[0] = cadd(v[0], t1);
    v[1] = cadd(m, j);
    v[2] = csub(m, j);
}

template <int SIGN, typename C> PZL_FFT_HD void dft5(C *v)
{
    typedef decltype(v[0].x) R;
    const R c1 = (R)0.30901699437494742410, c2 = (R)-0.80901699437494742410;
    const R s1 = (R)SIGN * (R)0.95105651629515357212, s2 = (R)SIGN * (R)0.58778525229247312917;
    C a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
    C a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
    C x0 = v[0];
    C m1 = C{x0.x + c1 * a1.x + c2 * a2.x, x0.y + c1 * a1.y + c2 * a2.y};
    C m2 = C{x0.x + c2 * a1.x + c1 * a2.x, x0.y + c2 * a1.y + c1 * a2.y};
    // n1 = i (s1 b1 + s2 b2), n2 = i (s2 b1 - s1 b2)
    C q1 = C{s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y};
    C q2 = C{s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y};
    C n1 = C{-q1.y, q1.x}, n2 = C{-q2.y, q2.x};
    v[0] = C{x0.x + a1.x + a2.x, x0.y + a1.y + a2.y};
    v[1] = cadd(m1, n1);
    v[4] = csub(m1, n1);
    v[2] = cadd(m2, n2);
    v[3] = csub(m2, n2);
}

template <int SIGN, typename C> PZL_FFT_HD void dft6(C *v)
{
    // 6 = 2 x 3 (Good-Thomas free: 3 then 2 with twiddles)
    C a[3] = {v[0], v[2], v[4]}, b[3] = {v[1], v[3], v[5]};
    dft3<SIGN>(a);
    dft3<SIGN>(b);
    typedef decltype(v[0].x) R;
    const R h = (R)0.5, s = (R)SIGN * (R)0.86602540378443864676;
    // b[k] *= exp(SIGN 2 pi i k / 6)
    C b1 = b[1], b2 = b[2];
    b[1] = C{b1.x * h - b1.y * s, b1.x * s + b1.y * h};
    b[2] = C{-b2.x * h - b2.y * s, b2.x * s - b2.y * h};
#pragma unroll
    for (int k = 0; k < 3; k++) {
        v[k] = cadd(a[k], b[k]);
        v[k + 3] = csub(a[k], b[k]);
    }
}

// exp(SIGN 2 pi i m / R) at compile time: cos and sin by Taylor series on
// [0, pi/4] after octant reduction, in double, rounded once to the working type.
constexpr double ct_pi = 3.14159265358979323846264338327950288;
constexpr double ct_taylor_sin(double x)
{
    double term = x, sum = x;
    for (int k = 1; k < 30; k++) {
        term *= -x * x / ((2.0 * k) * (2.0 * k + 1.0));
        sum += term;
    }
    return sum;
}
constexpr double ct_taylor_cos(double x)
{
    double term = 1.0, sum = 1.0;
    for (int k = 1; k < 30; k++) {
        term *= -x * x / ((2.0 * k - 1.0) * (2.0 * k));
        sum += term;
    }
    return sum;
}
// cos / sin of 2 pi m / R for integer m, R > 0 (exact octant symmetries)
constexpr double ct_cos2pi(long m, long R)
{
    m = ((m % R) + R) % R;
    // angle = 2 pi m / R; reduce with the 8 octants in integer arithmetic
    const long o = (8 * m) / R;            // octant 0..7
    const double f = (double)(8 * m - o * R) / (double)R;  // fraction within the octant
    const double a = f * ct_pi / 4.0;
    switch (o) {
    case 0: return ct_taylor_cos(a);
    case 1: return ct_taylor_sin(ct_pi / 4.0 - a);
    case 2: return -ct_taylor_sin(a);
    case 3: return -ct_taylor_cos(ct_pi / 4.0 - a);
    case 4: return -ct_taylor_cos(a);
    case 5: return -ct_taylor_sin(ct_pi / 4.0 - a);
    case 6: return ct_taylor_sin(a);
    default: return ct_taylor_cos(ct_pi / 4.0 - a);
    }
}
constexpr double ct_sin2pi(long m, long R) { return ct_cos2pi(4 * m - R, 4 * R); }

// cos and sin of 2 pi m / R, m < R, as compile-time tables
template <int R> struct ct_roots {
    double c[R], s[R];
    constexpr ct_roots() : c(), s()
    {
        for (int m = 0; m < R; m++) {
            c[m] = ct_cos2pi(m, R);
            s[m] = ct_sin2pi(m, R);
        }
    }
};

// DFT of odd size R by the symmetric pairs j, R - j:
//   X[k] = v0 + sum_j (v_j + v_{R-j}) cos(2 pi j k / R)
//             + SIGN i sum_j (v_j - v_{R-j}) sin(2 pi j k / R),  j = 1 .. (R-1)/2,
// and X[R - k] the same with the sine sum's sign reversed.
template <int R, int SIGN, typename C> PZL_FFT_HD void dft_odd(C *v)
{
    typedef decltype(v[0].x) T;
    constexpr int H = (R - 1) / 2;
    constexpr ct_roots<R> rt{};
    C sp[H], dm[H];
#pragma unroll
    for (int j = 1; j <= H; j++) {
        sp[j - 1] = cadd(v[j], v[R - j]);
        dm[j - 1] = csub(v[j], v[R - j]);
    }
    const C x0 = v[0];
    C y0 = x0;
#pragma unroll
    for (int j = 0; j < H; j++) y0 = cadd(y0, sp[j]);
#pragma unroll
    for (int k = 1; k <= H; k++) {
        T ar = x0.x, ai = x0.y, br = 0, bi = 0;
#pragma unroll
        for (int j = 1; j <= H; j++) {
            const T c = (T)rt.c[(j * k) % R], sn = (T)rt.s[(j * k) % R];
            ar += sp[j - 1].x * c;
            ai += sp[j - 1].y * c;
            br += dm[j - 1].x * sn;
            bi += dm[j - 1].y * sn;
        }
        v[k] = C{ar - (T)SIGN * bi, ai + (T)SIGN * br};
        v[R - k] = C{ar + (T)SIGN * bi, ai - (T)SIGN * br};
    }
    v[0] = y0;
}

// x * exp(SIGN 2 pi i M / R) with the trivial factors done without multiplies
template <int M, int R, int SIGN, typename C> PZL_FFT_HD C twc(C x)
{
    typedef decltype(x.x) T;
    constexpr int m = ((M % R) + R) % R;
    if constexpr (m == 0) {
        return x;
    } else if constexpr (4 * m == R) {
        return cmul_si<SIGN>(x);
    } else if constexpr (2 * m == R) {
        return C{-x.x, -x.y};
    } else if constexpr (4 * m == 3 * R) {
        return cmul_si<-SIGN>(x);
    } else {
        constexpr T c = (T)ct_cos2pi(m, R);
        constexpr T s = (T)(SIGN * ct_sin2pi(m, R));
        return C{x.x * c - x.y * s, x.x * s + x.y * c};
    }
}

template <int R, int SIGN, typename C> PZL_FFT_HD void dftR(C *v);

// DFT of size R1*R2 as R2 DFTs of size R1, twiddles, R1 DFTs of size R2:
// X[k1 + R1 k2] = sum_n2 W_R2^(n2 k2) W_R^(n2 k1) sum_n1 x[R2 n1 + n2] W_R1^(n1 k1)
template <int R1, int R2, int SIGN, int N2 = 0, typename C> PZL_FFT_HD void comp_stage1(C (*t)[R1], C *v)
{
    if constexpr (N2 < R2) {
#pragma unroll
        for (int n1 = 0; n1 < R1; n1++) t[N2][n1] = v[R2 * n1 + N2];
        dftR<R1, SIGN>(t[N2]);
        comp_stage1<R1, R2, SIGN, N2 + 1>(t, v);
    }
}
template <int R1, int R2, int SIGN, int N2, int K1, typename C> PZL_FFT_HD void comp_tw(C (*t)[R1])
{
    if constexpr (N2 < R2) {
        if constexpr (K1 < R1) {
            t[N2][K1] = twc<N2 * K1, R1 * R2, SIGN>(t[N2][K1]);
            comp_tw<R1, R2, SIGN, N2, K1 + 1>(t);
        } else {
            comp_tw<R1, R2, SIGN, N2 + 1, 0>(t);
        }
    }
}
template <int R1, int R2, int SIGN, typename C> PZL_FFT_HD void dft_comp(C *v)
{
    C t[R2][R1];
    comp_stage1<R1, R2, SIGN>(t, v);
    comp_tw<R1, R2, SIGN, 1, 1>(t);
#pragma unroll
    for (int k1 = 0; k1 < R1; k1++) {
        C u[R2];
#pragma unroll
        for (int n2 = 0; n2 < R2; n2++) u[n2] = t[n2][k1];
        dftR<R2, SIGN>(u);
#pragma unroll
        for (int k2 = 0; k2 < R2; k2++) v[k1 + R1 * k2] = u[k2];
    }
}

template <int R, int SIGN, typename C> PZL_FFT_HD void dftR(C *v)
{
    if constexpr (R == 1) return;
    else if constexpr (R == 2) dft2<SIGN>(v[0], v[1]);
    else if constexpr (R == 3) dft3<SIGN>(v);
    else if constexpr (R == 4) dft4<SIGN>(v);
    else if constexpr (R == 5) dft5<SIGN>(v);
    else if constexpr (R == 6) dft6<SIGN>(v);
    else if constexpr (R == 8) dft8<SIGN>(v);
    else if constexpr (R == 10) dft_comp<2, 5, SIGN>(v);
    else if constexpr (R == 12) dft_comp<4, 3, SIGN>(v);
    else if constexpr (R == 16) dft_comp<4, 4, SIGN>(v);
    else if constexpr (R == 20) dft_comp<4, 5, SIGN>(v);
    else static_assert(R == 2, "radix not provided");
}

// One Stockham pass of radix R over n points in LDS: each of NT threads reads
// its butterflies' inputs (twiddled by exp(SIGN 2 pi i jm r / (Ns R)), from
// sincospi in double) into registers, the block synchronises, then writes them
// in place.  R and the butterflies per thread are compile-time, so the
// register arrays need no scratch; they are sized for n <= NMAX.
template <int SIGN, int R, int NT, int NMAX>
__device__ __forceinline__ void lds_fft_pass_f64(double2 *buf, int n, int Ns, int tid)
{
    constexpr int MAXB = (NMAX / R + NT - 1) / NT;  // butterflies per thread
    const int nbf = n / R;
    double2 v[MAXB][R];
#pragma unroll
    for (int b = 0; b < MAXB; b++) {
        const int j = tid + b * NT;
        if (j < nbf) {
            const int jm = j % Ns;
#pragma unroll
            for (int r = 0; r < R; r++) {
                double2 x = buf[j + r * nbf];
                if (Ns > 1 && r > 0) {
                    double sn, cs;
                    sincospi((double)SIGN * 2.0 * (double)(jm * r) / (double)(Ns * R), &sn, &cs);
                    x = cmul(x, double2{cs, sn});
                }
                v[b][r] = x;
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < MAXB; b++) {
        const int j = tid + b * NT;
        if (j < nbf) {
            if constexpr (R == 2) dft2<SIGN>(v[b][0], v[b][1]);
            else if constexpr (R == 3) dft3<SIGN>(v[b]);
            else if constexpr (R == 4) dft4<SIGN>(v[b]);
            else dft5<SIGN>(v[b]);
            const int jm = j % Ns;
            const int idxD = (j / Ns) * Ns * R + jm;
#pragma unroll
            for (int r = 0; r < R; r++) buf[idxD + r * Ns] = v[b][r];
        }
    }
    __syncthreads();
}

// Block-cooperative double-precision Stockham FFT of n points held in LDS
// (buf, n <= NMAX <= 4096), any n = 2^a 3^b 5^c, by NT threads: passes of
// radix 4 while 4 divides what is left, then 2, 3, 5.  Used by the
// kernel-bank build and the segment transforms; a smaller NMAX holds fewer
// butterflies per thread in registers.
template <int SIGN, int NT, int NMAX = 4096>
__device__ __forceinline__ void lds_fft_f64(double2 *buf, int n, int tid)
{
    static_assert(NMAX <= 4096, "LDS transforms hold at most 4096 points");
    int Ns = 1;
    while (Ns < n) {
        const int rem = n / Ns;
        if (rem % 4 == 0) {
            lds_fft_pass_f64<SIGN, 4, NT, NMAX>(buf, n, Ns, tid);
            Ns *= 4;
        } else if (rem % 2 == 0) {
            lds_fft_pass_f64<SIGN, 2, NT, NMAX>(buf, n, Ns, tid);
            Ns *= 2;
        } else if (rem % 3 == 0) {
            lds_fft_pass_f64<SIGN, 3, NT, NMAX>(buf, n, Ns, tid);
            Ns *= 3;
        } else {
            lds_fft_pass_f64<SIGN, 5, NT, NMAX>(buf, n, Ns, tid);
            Ns *= 5;
        }
    }
}
