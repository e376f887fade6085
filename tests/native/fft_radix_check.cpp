// Host check of the in-register DFTs of pzl_fft.h: every radix, both signs,
// against a direct double-precision DFT.  Prints one line per (radix, sign)
// with the largest error relative to the largest output; exits 1 on failure.
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "pzl_fft.h"

template <int R, int SIGN> static double check()
{
    double worst = 0;
    unsigned s = 12345u + 97u * R + (SIGN > 0);
    for (int trial = 0; trial < 50; trial++) {
        double2 v[R], x[R];
        for (int i = 0; i < R; i++) {
            s = s * 1664525u + 1013904223u;
            const double a = (s >> 8) / 16777216.0 - 0.5;
            s = s * 1664525u + 1013904223u;
            const double b = (s >> 8) / 16777216.0 - 0.5;
            x[i] = double2{a, b};
            v[i] = x[i];
        }
        dftR<R, SIGN>(v);
        double maxo = 0, maxe = 0;
        for (int k = 0; k < R; k++) {
            double re = 0, im = 0;
            for (int j = 0; j < R; j++) {
                const double ang = SIGN * 2.0 * M_PI * (double)((long)j * k % R) / R;
                re += x[j].x * cos(ang) - x[j].y * sin(ang);
                im += x[j].x * sin(ang) + x[j].y * cos(ang);
            }
            maxo = fmax(maxo, hypot(re, im));
            maxe = fmax(maxe, hypot(v[k].x - re, v[k].y - im));
        }
        worst = fmax(worst, maxe / maxo);
    }
    printf("radix %2d sign %+d max rel err %.3e\n", R, SIGN, worst);
    return worst;
}

int main()
{
    double w = 0;
    w = fmax(w, check<2, -1>()); w = fmax(w, check<2, 1>());
    w = fmax(w, check<3, -1>()); w = fmax(w, check<3, 1>());
    w = fmax(w, check<4, -1>()); w = fmax(w, check<4, 1>());
    w = fmax(w, check<5, -1>()); w = fmax(w, check<5, 1>());
    w = fmax(w, check<6, -1>()); w = fmax(w, check<6, 1>());
    w = fmax(w, check<8, -1>()); w = fmax(w, check<8, 1>());
    w = fmax(w, check<10, -1>()); w = fmax(w, check<10, 1>());
    w = fmax(w, check<12, -1>()); w = fmax(w, check<12, 1>());
    w = fmax(w, check<16, -1>()); w = fmax(w, check<16, 1>());
    w = fmax(w, check<20, -1>()); w = fmax(w, check<20, 1>());
    // the compile-time twiddle table: exp(2 pi i m / R) for all m < R of the radices used
    double tw = 0;
    for (int R : {8, 10, 12, 16, 20, 24, 256}) {
        for (int m = 0; m < R; m++) {
            tw = fmax(tw, fabs(ct_cos2pi(m, R) - cos(2 * M_PI * m / R)));
            tw = fmax(tw, fabs(ct_sin2pi(m, R) - sin(2 * M_PI * m / R)));
        }
    }
    printf("compile-time twiddles max abs err %.3e\n", tw);
    return (w < 1e-14 && tw < 1e-15) ? 0 : 1;
}
