// accel_output.cpp -- accelsearch_CPU2's candidate stage for the accel.h
// drop-in:
//   optimize_accelcand (+ the batch the driver uses)  accel_utils.c:573-657
//   calc_props, calc_rzwerrs                          characteristics.c:193-265, :353-398
//   nice_output_1/2 (value(error) strings)            output.c:6-210
//   output_fundamentals, output_harmonics             accel_utils.c:659-1047
// The optimisation itself runs on the GPU (pzl_optimize_spectrum); this file
// holds the host logic and the text/binary output formats, which keep the
// reference's columns and number formats so that downstream readers
// (sifting, prepfold -accelcand) parse them unchanged.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "accel.h"
#include "pzl_accel.h"
#include "pzl_stats.h"

extern "C" int pzl_optimize_accelcands(GSList *list, accelobs *obs);

namespace {

constexpr double kPi = 3.14159265358979323846, kTwoPi = 2.0 * kPi;
constexpr double kSOL = 299792458.0;  // presto.h:39
constexpr double kDblCorrect = 1e-14;  // presto.h:21
constexpr int kNoteLen = 20;

struct RzwErrs {  // rzwerrs, presto.h:234-248
    double p, pd, pdd, f, fd, fdd;
    float perr, pderr, pdderr, ferr, fderr, fdderr;
};

double copysign_like(double a, double b) { return b < 0.0 ? -fabs(a) : fabs(a); }

int twon_to_index(int n)  // accelsearch_lib.c:293-303
{
    int x = 0;
    while (n > 1) {
        n >>= 1;
        x++;
    }
    return x;
}

void center_string(char *out, const char *in, int width)  // accel_utils.c:659-679
{
    const int len = (int)strlen(in);
    memset(out, ' ', width);
    out[width] = '\0';
    if (len >= width) strncpy(out, in, width);
    else strncpy(out + (width - len) / 2, in, len);
}

void write_val_with_err(FILE *f, double val, double err, int numerr, int width)  // accel_utils.c:681-696
{
    char tmp[80], ctr[80];
    tmp[0] = '\0';
    if (numerr == 1) nice_output_1(tmp, val, err, 0);
    else if (numerr == 2) nice_output_2(tmp, val, err, 0);
    else printf("\numerr = %d is out-of-range (1-2) in write_val_with_err()\n", numerr);
    center_string(ctr, tmp, width);
    fprintf(f, "%s  ", ctr);
}

void calc_rzwerrs(const fourierprops *props, double T, RzwErrs *r)  // characteristics.c:353-398
{
    if (T <= 0.0) {
        printf("\n T = %f (out-of-bounds) in calc_rzwerrs().\n\n", T);
        exit(-1);
    }
    const double T2 = T * T, T3 = T2 * T, r2 = props->r * props->r, sr2 = (double)props->rerr * props->rerr;
    const double r4 = r2 * r2, z2 = props->z * props->z, sz2 = (double)props->zerr * props->zerr;
    const double tmp = props->r * props->w - 3 * z2;
    r->f = props->r / T;
    r->ferr = (float)(props->rerr / T);
    r->fd = props->z / T2;
    r->fderr = (float)(props->zerr / T2);
    r->fdd = props->w / T3;
    r->fdderr = (float)(props->werr / T3);
    r->p = T / props->r;
    r->perr = (float)(T * props->rerr / r2);
    r->pd = -props->z / r2;
    r->pderr = (float)sqrt(4 * z2 * sr2 / (r4 * r2) + sz2 / r4);
    r->pdd = props->w == 0.0 ? 0.0 : (2 * z2 - props->r * props->w) / (T * r2 * props->r);
    r->pdderr = (float)sqrt(((double)props->werr * props->werr * r4 + 16 * sz2 * r2 + 4 * tmp * tmp * sr2) /
                            (r4 * r4 * T2));
}

// accel_utils.c:99-136: "SL? of Cand i" (a stronger candidate within 15 bins
// at another z) and "H k of Cand i" notes
void compare_rzw_cands(const fourierprops *list, int n, char *notes)
{
    char tmp[40];
    for (int ii = 0; ii < n; ii++) {
        for (int jj = 0; jj < n; jj++) {
            if (ii == jj) continue;
            if (fabs(list[ii].r - list[jj].r) < 15.0 && fabs(list[ii].z - list[jj].z) > 1.0 &&
                list[ii].pow > list[jj].pow) {
                if (strncmp(notes + jj * kNoteLen, "                      ", kNoteLen) == 0) {
                    snprintf(tmp, sizeof tmp, "SL? of Cand %d", ii + 1);
                    strncpy(notes + jj * kNoteLen, tmp, kNoteLen);
                }
                continue;
            }
            for (int kk = 1; kk < 61; kk++) {
                if ((fabs(list[ii].r - list[jj].r / kk) < list[jj].rerr * 3) &&
                    (fabs(list[ii].z - list[jj].z / kk) < list[jj].zerr * 2)) {
                    if (strncmp(notes + jj * kNoteLen, "                      ", kNoteLen) == 0) {
                        snprintf(tmp, sizeof tmp, "H %d of Cand %d", kk, ii + 1);
                        strncpy(notes + jj * kNoteLen, tmp, kNoteLen);
                        break;
                    }
                }
            }
        }
    }
}

void alloc_cand_arrays(accelcand *c)  // accel_utils.c:578-582
{
    if (c->pows) return;
    c->pows = (double *)calloc(c->numharm, sizeof(double));
    c->hirs = (double *)calloc(c->numharm, sizeof(double));
    c->hizs = (double *)calloc(c->numharm, sizeof(double));
    c->hiws = (double *)calloc(c->numharm, sizeof(double));
    c->derivs = (rderivs *)calloc(c->numharm, sizeof(rderivs));
}

}  // namespace

extern "C" {

// "value(error)" with a one-digit error (output.c:6-112)
int nice_output_1(char *output, double val, double err, int len)
{
    char temp[64];
    const int errexp = fabs(err) == 0.0 ? 0 : (int)floor(log10(fabs(err)));
    int e = errexp;
    int errval = (int)floor(fabs(err) * pow(10.0, (double)(-e)) + kDblCorrect + 0.5);
    if (errval == 10) {
        errval = 1;
        e++;
    }
    const double rndval = pow(10.0, (double)e) * floor(val * pow(10.0, (double)(-e)) + 0.5);
    int nint;
    if (fabs(val) == 0.0) {
        nint = 1;
    } else {
        nint = (int)ceil(log10(fabs(val)));
        if (nint == 0) nint++;
    }
    const int nfrac = -e;
    int totprec = nint + nfrac;
    const int outexp = fabs(rndval) == 0.0 ? 0 : (int)floor(log10(fabs(rndval)));
    const double outmant = rndval * pow(10.0, (double)(-outexp));
    if (fabs(1.0 - outmant) < kDblCorrect || fabs(-1.0 - outmant) < kDblCorrect) totprec++;
    if ((outexp >= 0 && e > 0) && outexp > e)
        snprintf(temp, sizeof temp, "% .*f(%d)x10^%d", totprec - 1, copysign_like(outmant, rndval), errval, outexp);
    else if ((outexp >= 0 && e > 0) && outexp == e)
        snprintf(temp, sizeof temp, "% d(%d)x10^%d", (int)copysign_like(outmant, rndval), errval, outexp);
    else if (outexp < -4 && outexp >= e)
        snprintf(temp, sizeof temp, "% .*f(%d)x10^%d", totprec - 1, copysign_like(outmant, rndval), errval, outexp);
    else if (outexp < e && e != 0)
        snprintf(temp, sizeof temp, "% d(%d)x10^%d", (int)(copysign_like(outmant, rndval) + kDblCorrect), errval, e);
    else if (nfrac == 0 && fabs(rndval) < 1.0e-15)
        snprintf(temp, sizeof temp, "% d(%d)", (int)fabs(rndval), errval);
    else if (fabs(rndval) <= kDblCorrect && e < -5)
        snprintf(temp, sizeof temp, "0.0(%d)x10^%d", errval, e + 1);
    else
        snprintf(temp, sizeof temp, "% .*f(%d)", nfrac, rndval, errval);
    if (len == 0) snprintf(output, 64, "%s", temp);
    else snprintf(output, 64, "%*s", len, temp);
    return (int)strlen(output);
}

// "value(error)" with a two-digit error (output.c:114-210)
int nice_output_2(char *output, double val, double err, int len)
{
    char temp[64];
    int errexp = fabs(err) == 0.0 ? 0 : (int)floor(log10(fabs(err)));
    int errval = (int)floor(fabs(err) * pow(10.0, (double)(-errexp + 1)) + kDblCorrect + 0.5);
    if (errval == 100) {
        errval = 10;
        errexp++;
    }
    double rndval = pow(10.0, (double)(errexp - 1)) * floor(val * pow(10.0, (double)(-errexp + 1)) + 0.5);
    const int outexp = rndval == 0.0 ? 0 : (int)floor(log10(fabs(rndval)));
    const int valexp = val == 0.0 ? 0 : (int)floor(log10(fabs(val)));
    double outmant = rndval * pow(10.0, (double)(-outexp));
    if (outexp < -4) {
        const int tmp = outexp - errexp + 1;
        if (tmp < 2) {
            if (tmp == 0)
                snprintf(temp, sizeof temp, "% .1f(%d.%d)x10^%d", outmant / 10, errval / 10, errval % 10, outexp + 1);
            else
                snprintf(temp, sizeof temp, "% .1f(%d.%d)x10^%d", outmant, errval / 10, errval % 10, outexp);
        } else {
            snprintf(temp, sizeof temp, "% .*f(%.2d)x10^%d", tmp, outmant, errval, outexp);
        }
    } else if (errexp - valexp > 1 && (errexp != 0 && errexp != 1)) {
        snprintf(temp, sizeof temp, "0.0(%d.%d)x10^%d", errval / 10, errval % 10, errexp);
    } else if (errexp == 0) {
        if (fabs(rndval) < 0.1) rndval = fabs(rndval);
        snprintf(temp, sizeof temp, "% .1f(%d.%d)", rndval, errval / 10, errval % 10);
    } else if (errexp < 0) {
        if (fabs(rndval) * pow(10.0, (double)(-errexp + 1)) < 0.1) rndval = fabs(rndval);
        if (fabs(rndval) <= kDblCorrect && errexp < -4)
            snprintf(temp, sizeof temp, "0.0(%d.%d)x10^%d", errval / 10, errval % 10, errexp);
        else
            snprintf(temp, sizeof temp, "% .*f(%.2d)", -errexp + 1, rndval, errval);
    } else if (errexp - outexp == 0) {
        if (errexp == 1)
            snprintf(temp, sizeof temp, "% d(%.2d)", (int)rndval, errval);
        else
            snprintf(temp, sizeof temp, "% .1f(%d.%d)x10^%d", outmant, errval / 10, errval % 10, errexp);
    } else if (errexp - outexp == 1) {
        if (fabs(outmant) < 0.99999999999999) outmant = fabs(outmant);
        snprintf(temp, sizeof temp, "% .1f(%d.%d)x10^%d", outmant / 10.0, errval / 10, errval % 10, errexp);
    } else if ((errexp > 1 && outexp > 1) && errexp - outexp <= -1) {
        snprintf(temp, sizeof temp, "% .*f(%.2d)x10^%d", outexp - errexp + 1, outmant, errval, outexp);
    } else if (errexp == 1 && outexp > 0) {
        snprintf(temp, sizeof temp, "% d(%.2d)", (int)rndval, errval);
    } else {
        printf("This is an undefined condition in nice_output2().\n");
        exit(1);
    }
    if (len == 0) snprintf(output, 64, "%s", temp);
    else snprintf(output, 64, "%*s", len, temp);
    return (int)strlen(output);
}

void calc_props(rderivs data, double r, double z, double w, fourierprops *result)
{
    if (fabs(data.pow) < kDblCorrect) {
        printf("\n data.pow = %f (out-of-bounds) in calc_props().\n\n", data.pow);
        exit(-1);
    }
    if (data.locpow < kDblCorrect) {
        printf("\n data.locpow = %f (out-of-bounds) in calc_props().\n\n", data.locpow);
        exit(-1);
    }
    const double tmppow = data.pow / data.locpow, tmpd2pow = data.d2pow / data.locpow;
    result->pur = (float)(sqrt(1.5 * fabs(tmpd2pow) / tmppow) / kPi);
    if (fabs(tmppow) < kDblCorrect) result->pur = 1.0f;
    result->purerr = (float)(1.0 / (result->pur * sqrt(10.0 * tmppow)));
    result->r = r;
    result->rerr = (float)(3.0 / (kPi * result->pur * sqrt(6.0 * tmppow)));
    result->z = z;
    result->zerr = (float)(3.0 * sqrt(10.0) / (kPi * result->pur * result->pur * sqrt(tmppow)));
    result->w = w;
    result->werr = (float)(6.0 * sqrt(105.0) / (kPi * result->pur * result->pur * result->pur * sqrt(tmppow)));
    result->pow = (float)tmppow;
    result->powerr = (float)sqrt(2.0 * tmppow);
    result->sig = (float)sqrt(2.0 * tmppow - log(kPi * tmppow));
    result->rawpow = (float)data.pow;
    result->phs = (float)data.phs;
    result->phserr = (float)(1.0 / result->powerr);
    result->cen = (float)(-data.dphs / kTwoPi);
    result->cenerr = (float)(1.0 / sqrt(24.0 * tmppow));
    result->locpow = (float)data.locpow;
}

// All candidates of a list at once (the loop of accelsearch_lib.c:921-930 as
// one GPU batch).  With harmonic polishing: the rzw volume of the summed
// harmonics for w > 0 searches, the (r, z) simplex for w = 0 ones
// (accel_utils.c:584-608).  With -noharmpolish: every harmonic of every
// candidate on its own, max_rzw_arr or max_rz_arr at (r, z, w) * (ii + 1)
// (accel_utils.c:610-652).  Returns 0, or -1 when the spectrum is not in
// memory (the reference then reads the file, max_rzw_file).  create_accelobs
// holds the whole spectrum in obs->fft when it maps the file, reads a .dat,
// and also when it is asked not to map (it then reads the file whole).
int pzl_optimize_accelcands(GSList *list, accelobs *obs)
{
    if (!(obs->mmap_file || obs->dat_input || obs->fft)) return -1;
    const bool polish = obs->use_harmonic_polishing != 0;
    std::vector<PzlOptCand> oc;
    std::vector<accelcand *> cs;
    for (GSList *p = list; p; p = p->next) {
        accelcand *c = (accelcand *)p->data;
        cs.push_back(c);
        for (int ii = 0; ii < (polish ? 1 : c->numharm); ii++) {
            PzlOptCand o;
            memset(&o, 0, sizeof o);
            if (polish) {
                o.r = c->r - obs->lobin;
                o.z = c->z;
                o.w = c->w;
                o.numharm = c->numharm;
            } else {
                o.r = c->r * (ii + 1) - obs->lobin;
                o.z = c->z * (ii + 1);
                o.w = c->w * (ii + 1);
                o.numharm = 1;
            }
            o.fdot_only = obs->numw == 0;
            oc.push_back(o);
        }
    }
    if (oc.empty()) return 0;
    const int rc = polish ? pzl_optimize_spectrum(0, (const float *)obs->fft, obs->numbins, oc.data(),
                                                  (long long)oc.size())
                          : pzl_maximize_spectrum(0, (const float *)obs->fft, obs->numbins, oc.data(),
                                                  (long long)oc.size());
    if (rc != 0) {
        printf("\nGPU candidate optimisation failed: %s\n\n", pzl_last_error());
        exit(1);
    }
    size_t k = 0;
    for (accelcand *c : cs) {
        alloc_cand_arrays(c);
        for (int ii = 0; ii < c->numharm; ii++) {
            const PzlOptCand &o = polish ? oc[k] : oc[k + ii];
            if (polish) {
                c->hirs[ii] = (o.ropt + obs->lobin) * (ii + 1);
                c->hizs[ii] = o.zopt * (ii + 1);
                c->hiws[ii] = obs->numw ? o.wopt * (ii + 1) : 0.0;
            } else {
                c->hirs[ii] = o.ropt + obs->lobin;
                c->hizs[ii] = o.zopt;
                // w = 0: max_rz_arr sets no w (the reference leaves the
                // uncleared gen_dvect value); 0 here
                c->hiws[ii] = obs->numw ? o.wopt : 0.0;
            }
            const int h = polish ? ii : 0;
            const PzlRDerivs &d = o.derivs[h];
            c->derivs[ii] = rderivs{d.pow, d.phs, d.dpow, d.dphs, d.d2pow, d.d2phs, d.locpow};
            c->pows[ii] = o.pows[h];
        }
        k += polish ? 1 : (size_t)c->numharm;
        c->sigma = (float)pzl_candidate_sigma(c->power, c->numharm,
                                              (double)obs->numindep[twon_to_index(c->numharm)]);
    }
    return 0;
}

void optimize_accelcand(accelcand *cand, accelobs *obs)
{
    GSList one;
    one.data = cand;
    one.next = nullptr;
    if (pzl_optimize_accelcands(&one, obs) != 0) {
        printf("\noptimize_accelcand: the spectrum is not in memory\n\n");
        exit(1);
    }
}

void output_fundamentals(fourierprops *props, GSList *list, accelobs *obs, infodata *idata)
{
    (void)idata;  // the pulsar-catalogue comparison of comp_psr_to_cand is not available
    const int numcols = 13;
    const int widths[13] = {4, 5, 6, 8, 4, 16, 15, 15, 15, 11, 11, 15, 20};
    const int errors[13] = {0, 0, 0, 0, 0, 1, 1, 2, 1, 2, 2, 2, 0};
    static const char *titles1[] = {"",         "",          "Summed",     "Coherent", "Num",
                                    "Period",   "Frequency", "FFT 'r'",    "Freq Deriv", "FFT 'z'",
                                    "FFT 'w'",  "Accel",     ""};
    static const char *titles2[] = {"Cand",   "Sigma", "Power",  "Power",  "Harm",    "(ms)", "(Hz)",
                                    "(bin)",  "(Hz/s)", "(bins)", "(bins)", "(m/s^2)", "Notes"};
    char tmp[80], ctr[80];
    const int numcands = (int)g_slist_length(list);
    if (obs->workfile && !obs->dat_input) fclose(obs->workfile);
    obs->workfile = fopen(obs->accelnm, "w");
    if (!obs->workfile) {
        printf("\nCannot open '%s' for writing.\n\n", obs->accelnm);
        exit(1);
    }
    FILE *f = obs->workfile;
    std::vector<char> notes((size_t)numcands * kNoteLen + 1, ' ');
    compare_rzw_cands(props, numcands, notes.data());
    for (int row = 0; row < 2; row++) {
        const char **titles = row == 0 ? titles1 : titles2;
        for (int ii = 0; ii < numcols - 1; ii++) {
            if (obs->numw == 0 && ii == 10) continue;  // no jerk column
            center_string(ctr, titles[ii], widths[ii]);
            fprintf(f, "%s  ", ctr);
        }
        center_string(ctr, titles[numcols - 1], widths[numcols - 1]);
        fprintf(f, "%s\n", ctr);
    }
    for (int ii = 0; ii < numcols - 1; ii++) {
        if (obs->numw == 0 && ii == 10) continue;
        memset(tmp, '-', widths[ii]);
        tmp[widths[ii]] = '\0';
        fprintf(f, "%s--", tmp);
    }
    memset(tmp, '-', widths[numcols - 1]);
    tmp[widths[numcols - 1]] = '\0';
    fprintf(f, "%s\n", tmp);

    int ii = 0;
    for (GSList *p = list; p; p = p->next, ii++) {
        const accelcand *cand = (const accelcand *)p->data;
        RzwErrs errs;
        calc_rzwerrs(props + ii, obs->T, &errs);
        // coherently summed power (accel_utils.c:806-826)
        double cr = 0.0, ci = 0.0;
        const double phs0 = cand->derivs[0].phs;
        for (int jj = 0; jj < cand->numharm; jj++) {
            const rderivs &h = cand->derivs[jj];
            const double amp = obs->nph > 0.0 ? sqrt(h.pow / obs->nph) : sqrt(h.pow / h.locpow);
            const double phscorr = phs0 - fmod((jj + 1.0) * phs0, kTwoPi);
            cr += amp * cos(h.phs + phscorr);
            ci += amp * sin(h.phs + phscorr);
        }
        const double coherent_pow = cr * cr + ci * ci;
        snprintf(tmp, sizeof tmp, "%-4d", ii + 1);
        center_string(ctr, tmp, widths[0]);
        fprintf(f, "%s  ", ctr);
        snprintf(tmp, sizeof tmp, "%.2f", cand->sigma);
        center_string(ctr, tmp, widths[1]);
        fprintf(f, "%s  ", ctr);
        snprintf(tmp, sizeof tmp, "%.2f", cand->power);
        center_string(ctr, tmp, widths[2]);
        fprintf(f, "%s  ", ctr);
        snprintf(tmp, sizeof tmp, "%.2f", coherent_pow);
        center_string(ctr, tmp, widths[3]);
        fprintf(f, "%s  ", ctr);
        snprintf(tmp, sizeof tmp, "%d", cand->numharm);
        center_string(ctr, tmp, widths[4]);
        fprintf(f, "%s  ", ctr);
        write_val_with_err(f, errs.p * 1000.0, errs.perr * 1000.0, errors[5], widths[5]);
        write_val_with_err(f, errs.f, errs.ferr, errors[6], widths[6]);
        write_val_with_err(f, props[ii].r, props[ii].rerr, errors[7], widths[7]);
        write_val_with_err(f, errs.fd, errs.fderr, errors[8], widths[8]);
        write_val_with_err(f, props[ii].z, props[ii].zerr, errors[9], widths[9]);
        if (obs->numw) write_val_with_err(f, props[ii].w, props[ii].werr, errors[10], widths[10]);
        const double accel = props[ii].z * kSOL / (obs->T * obs->T * errs.f);
        const double accelerr = props[ii].zerr * kSOL / (obs->T * obs->T * errs.f);
        write_val_with_err(f, accel, accelerr, errors[11], widths[11]);
        fprintf(f, "  %.20s\n", notes.data() + ii * kNoteLen);
        fflush(f);
    }
    fprintf(f, "\n\n");
}

void output_harmonics(GSList *list, accelobs *obs, infodata *idata)
{
    (void)idata;
    const int numcols = 15;
    const int widths[15] = {5, 4, 5, 15, 11, 18, 13, 12, 9, 12, 9, 12, 10, 10, 20};
    const int errors[15] = {0, 0, 0, 2, 0, 2, 0, 2, 0, 2, 0, 2, 2, 2, 0};
    static const char *titles1[] = {"",        "",         "",        "Power /", "Raw",
                                    "FFT 'r'", "Pred 'r'", "FFT 'z'", "Pred 'z'", "FFT 'w'",
                                    "Pred 'w'", "Phase",   "Centroid", "Purity",  ""};
    static const char *titles2[] = {"Cand",   "Harm",   "Sigma",  "Loc Pow", "Power",  "(bin)",   "(bin)", "(bins)",
                                    "(bins)", "(bins)", "(bins)", "(rad)",   "(0-1)", "<p> = 1", "Notes"};
    char tmp[80], ctr[80], notes[kNoteLen + 1];
    FILE *f = obs->workfile;
    if (!f) return;
    for (int row = 0; row < 2; row++) {
        for (int ii = 0; ii < numcols - 1; ii++) {
            if (obs->numw == 0 && (ii == 9 || ii == 10)) continue;
            const char *t = row == 0 ? titles1[ii] : (obs->nph > 0.0 && ii == 3 ? "NumPhot" : titles2[ii]);
            center_string(ctr, t, widths[ii]);
            fprintf(f, "%s  ", ctr);
        }
        center_string(ctr, row == 0 ? titles1[numcols - 1] : titles2[numcols - 1], widths[numcols - 1]);
        fprintf(f, "%s\n", ctr);
    }
    for (int ii = 0; ii < numcols - 1; ii++) {
        if (obs->numw == 0 && (ii == 9 || ii == 10)) continue;
        memset(tmp, '-', widths[ii]);
        tmp[widths[ii]] = '\0';
        fprintf(f, "%s--", tmp);
    }
    memset(tmp, '-', widths[numcols - 1]);
    tmp[widths[numcols - 1]] = '\0';
    fprintf(f, "%s\n", tmp);
    memset(notes, ' ', kNoteLen);
    notes[kNoteLen] = '\0';
    int ii = 0;
    for (GSList *p = list; p; p = p->next, ii++) {
        accelcand *cand = (accelcand *)p->data;
        for (int jj = 0; jj < cand->numharm; jj++) {
            fourierprops props;
            rderivs d = cand->derivs[jj];
            if (obs->nph > 0.0) d.locpow = obs->nph;
            calc_props(d, cand->hirs[jj], cand->hizs[jj], cand->hiws[jj], &props);
            if (jj == 0) snprintf(tmp, sizeof tmp, " %-4d", ii + 1);
            else snprintf(tmp, sizeof tmp, "     ");
            center_string(ctr, tmp, widths[0]);
            fprintf(f, "%s  ", ctr);
            snprintf(tmp, sizeof tmp, "%-4d", jj + 1);
            center_string(ctr, tmp, widths[1]);
            fprintf(f, "%s  ", ctr);
            snprintf(tmp, sizeof tmp, "%.2f", pzl_candidate_sigma(props.pow, 1, 1));
            center_string(ctr, tmp, widths[2]);
            fprintf(f, "%s  ", ctr);
            write_val_with_err(f, props.pow, props.powerr, errors[3], widths[3]);
            snprintf(tmp, sizeof tmp, "%.3g", props.rawpow);
            center_string(ctr, tmp, widths[4]);
            fprintf(f, "%s  ", ctr);
            write_val_with_err(f, props.r, props.rerr, errors[5], widths[5]);
            snprintf(tmp, sizeof tmp, "%.2f", cand->r * (jj + 1));
            center_string(ctr, tmp, widths[6]);
            fprintf(f, "%s  ", ctr);
            write_val_with_err(f, props.z, props.zerr, errors[7], widths[7]);
            snprintf(tmp, sizeof tmp, "%.2f", cand->z * (jj + 1));
            center_string(ctr, tmp, widths[8]);
            fprintf(f, "%s  ", ctr);
            if (obs->numw) {
                write_val_with_err(f, props.w, props.werr, errors[9], widths[9]);
                snprintf(tmp, sizeof tmp, "%.2f", cand->w * (jj + 1));
                center_string(ctr, tmp, widths[10]);
                fprintf(f, "%s  ", ctr);
            }
            write_val_with_err(f, props.phs, props.phserr, errors[11], widths[11]);
            write_val_with_err(f, props.cen, props.cenerr, errors[12], widths[12]);
            write_val_with_err(f, props.pur, props.purerr, errors[13], widths[13]);
            fprintf(f, "  %.20s\n", notes);
            fflush(f);
        }
    }
    fprintf(f, "\n\n");
    fclose(f);
    obs->workfile = nullptr;
    // the .inf file follows the tables (accel_utils.c:1041-1045)
    const std::string inf = std::string(obs->rootfilenm) + ".inf";
    FILE *in = fopen(inf.c_str(), "r");
    FILE *out = fopen(obs->accelnm, "a");
    if (in && out) {
        char buf[4096];
        size_t n;
        while ((n = fread(buf, 1, sizeof buf, in)) > 0) fwrite(buf, 1, n, out);
    }
    if (in) fclose(in);
    if (out) fclose(out);
}

}  // extern "C"
