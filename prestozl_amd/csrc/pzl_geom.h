// pzl_geom.h -- search geometry shared by host code and HIP kernels.
//
// The index arithmetic is the reference's, kept bit-for-bit because every
// subharmonic gather depends on it:
//   calc_required_r/z/w, index_from_r/z/w   accel_utils.c:54-97
//   NEAREST_INT                             accel_utils.c:36
//   per-(rstep, harmonic) rlo / numrs       accel_utils.c:1106-1140
// Device code evaluates these in double, which is IEEE-exact on gfx950, so
// the tables built on the GPU equal the ones the host would build.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define PZL_HD __host__ __device__ __forceinline__
#else
#define PZL_HD inline
#endif

#define PZL_NUMBETWEEN 2
#define PZL_DR 0.5
#define PZL_RDR 2
#define PZL_DZ 2
#define PZL_RDZ 0.5
#define PZL_DW 20
#define PZL_RDW 0.05
#define PZL_DBLCORRECT 1e-14
#define PZL_MAXHARMS 32  // numharm up to 32 (6 stages), as the reference

PZL_HD int pzl_nearest_int(double x) { return (int)(x < 0 ? x - 0.5 : x + 0.5); }
PZL_HD double pzl_calc_required_r(double hf, double rfull) { return rint(PZL_RDR * rfull * hf) * PZL_DR; }
PZL_HD int pzl_calc_required_z(double hf, double zfull) { return pzl_nearest_int(PZL_RDZ * zfull * hf) * PZL_DZ; }
PZL_HD int pzl_calc_required_w(double hf, double wfull) { return pzl_nearest_int(PZL_RDW * wfull * hf) * PZL_DW; }
PZL_HD int pzl_index_from_r(double r, double lor) { return (int)((r - lor) * PZL_RDR + PZL_DBLCORRECT); }
PZL_HD int pzl_index_from_z(double z, double loz) { return (int)((z - loz) * PZL_RDZ + PZL_DBLCORRECT); }
PZL_HD int pzl_index_from_w(double w, double low) { return (int)((w - low) * PZL_RDW + PZL_DBLCORRECT); }

// Per-(rstep, harmonic) r geometry (accel_utils.c:1104-1140).
struct PzlHarmR {
    long long rlo;   // lowest subharmonic r bin present (floor(drlo))
    int numrs;       // interbinned r points in the subharmonic plane
};

PZL_HD PzlHarmR pzl_harm_r(double hf, int fundamental, double fullrlo, double fullrhi, int corr_uselen)
{
    PzlHarmR o;
    double drlo = pzl_calc_required_r(hf, fullrlo);
    double drhi = pzl_calc_required_r(hf, fullrhi);
    o.rlo = (long long)floor(drlo);
    o.numrs = (int)((ceil(drhi) - floor(drlo)) * PZL_RDR + PZL_DBLCORRECT) + 1;
    if (fundamental) o.numrs = corr_uselen;
    else if (o.numrs % PZL_RDR) o.numrs = (o.numrs / PZL_RDR + 1) * PZL_RDR;
    return o;
}

// rinds[ii]: subharmonic r index of fundamental r point ii (accel_utils.c:1115-1122)
PZL_HD int pzl_rind(double hf, double fullrlo, long long sub_rlo, int ii)
{
    double rr = fullrlo + ii * PZL_DR;
    return pzl_index_from_r(pzl_calc_required_r(hf, rr), (double)sub_rlo);
}

// Description of one harmonic of a search, as the kernels see it.
struct PzlHarm {
    int numharm, harmnum, stage;  // harmonic = harmnum / numharm, summed at 'stage'
    double hf;                    // harmnum / numharm
    int zmax, wmax;               // largest |z|, |w| of this harmonic's kernels
    int numz, numw, nrows;        // kernel grid, nrows = numz * numw
    int zlo, wlo;                 // calc_required_z/w of the fundamental's zlo/wlo
    int fftlen;                   // correlation length
    int halfwidth;                // kern[0][0].kern_half_width = segment offset in bins
    int max_numrs;                // largest numrs over all rsteps of the search
    int rs_stride;                // floats between consecutive (w,z) rows of a plane
    long long bank_off;           // complex offset of this harmonic's kernels in the bank
    long long plane_off;          // float offset of this plane inside one rstep's planes
    long long seg_off;            // complex offset of this harmonic's segment inside one rstep
};
