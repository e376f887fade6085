// accel_api.cpp -- the accel.h drop-in (libaccelsearch_lib.so exports).
//
//   parseCmdline                accelsearch_cmd.c:1259 (flags, defaults :20-96)
//   create_accelobs             accel_utils.c:1636-2037 (.fft mmap, .inf, sizing)
//   create/free_subharminfos    accel_utils.c:236-325 (shapes; kernels on the GPU)
//   subharm_fderivs_vol         accel_utils.c:1094-1293 (plane computed on the GPU)
//   add_ffdotpows[_ptrs]        accel_utils.c:1356-1406 (GPU add of host planes)
//   search_ffdotpows            accel_utils.c:1507-1555 (GPU threshold, host list)
//   sort_accelcands, eliminate_harmonics, free/print_accelcand
//                               accel_utils.c:327-380, :482-571, :1049-1056
//   accelsearch_CPU1/GPU/CPU2   accelsearch_lib.c:360-980 (CPU2's candidate
//                               stage and outputs: accel_output.cpp)
// Error behaviour follows the reference: bad inputs print a message and exit;
// accelsearch_GPU returns 1 when the hit buffer overflows.
#include <fcntl.h>
#include <hip/hip_runtime.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "accel.h"
#include "pzl_accel.h"
#include "pzl_kernels.h"
#include "pzl_math.h"
#include "pzl_stats.h"

extern "C" int pzl_optimize_accelcands(GSList *list, accelobs *obs);  // accel_output.cpp

namespace {

[[noreturn]] void die(int code, const char *fmt, const char *arg = "")
{
    printf(fmt, arg);
    fflush(stdout);
    exit(code);
}

// ---------------------------------------------------------------- .inf
// readinf (ioinf.c:121-250): the fields in file order; N and dt follow the
// telescope block (7 lines unless the data set is artificial).
void read_inf_values(const char *root, infodata *d)
{
    std::string fn = std::string(root) + ".inf";
    FILE *f = fopen(fn.c_str(), "r");
    if (!f) die(1, "\nCannot open '%s'\n\n", fn.c_str());
    std::vector<std::string> vals;
    char line[512];
    while (fgets(line, sizeof line, f)) {
        char *eq = (strlen(line) > 40 && line[40] == '=') ? line + 40 : strrchr(line, '=');
        if (!eq) continue;
        std::string v(eq + 1);
        size_t a = v.find_first_not_of(" \t\r\n"), b = v.find_last_not_of(" \t\r\n");
        vals.push_back(a == std::string::npos ? std::string() : v.substr(a, b - a + 1));
    }
    fclose(f);
    memset(d, 0, sizeof *d);
    if (vals.size() < 4) die(1, "\nMalformed .inf file '%s'\n\n", fn.c_str());
    snprintf(d->name, sizeof d->name, "%s", vals[0].c_str());
    snprintf(d->telescope, sizeof d->telescope, "%s", vals[1].c_str());
    size_t k = 2;
    if (strcmp(d->telescope, "None (Artificial Data Set)") != 0) {
        if (vals.size() < 11) die(1, "\nMalformed .inf file '%s'\n\n", fn.c_str());
        snprintf(d->instrument, sizeof d->instrument, "%s", vals[2].c_str());
        snprintf(d->object, sizeof d->object, "%s", vals[3].c_str());
        snprintf(d->observer, sizeof d->observer, "%s", vals[6].c_str());
        d->bary = atoi(vals[8].c_str());
        k = 9;
    } else {
        d->mjd_i = -1;
        snprintf(d->object, sizeof d->object, "fake pulsar");
    }
    d->N = strtod(vals[k].c_str(), nullptr);
    d->dt = strtod(vals[k + 1].c_str(), nullptr);
    if (!(d->N > 0) || !(d->dt > 0)) die(1, "\nBad N or dt in '%s'\n\n", fn.c_str());
    d->numonoff = 1;
    d->onoff[0] = 0;
    d->onoff[1] = d->N - 1;
}

PzlConfig config_from(const accelobs *obs, const Cmdline *cmd)
{
    PzlConfig c{};
    c.N = (double)obs->N;
    c.dt = obs->dt;
    c.numbins = obs->numbins;
    c.lobin = obs->lobin;
    c.zmax = cmd->zmax;
    c.wmax = cmd->wmaxP ? cmd->wmax : 0;
    c.wmaxP = cmd->wmaxP;
    c.numharm = 1 << (obs->numharmstages - 1);
    c.sigma = obs->sigma;
    c.floP = cmd->floP;
    c.fhiP = cmd->fhiP;
    c.rloP = cmd->rloP;
    c.rhiP = cmd->rhiP;
    c.flo = cmd->flo;
    c.fhi = cmd->fhi;
    c.rlo = cmd->rlo;
    c.rhi = cmd->rhi;
    c.nph = obs->nph;
    c.batchsize = cmd->batchsizeP ? cmd->batchsize : 8;
    return c;
}

// ---------------------------------------------------------------- plane API
// subharm_fderivs_vol and friends keep one small plan per search shape and
// spectrum so that consecutive calls reuse its buffers and kernel bank.
struct ApiPlan {
    void *plan = nullptr;
    const void *fft = nullptr;
};
std::mutex g_api_mu;
std::map<std::string, ApiPlan> g_api_plans;

std::string shape_key(const accelobs *obs)
{
    char buf[512];
    snprintf(buf, sizeof buf, "%p|%lld|%lld|%lld|%.17g|%.17g|%.17g|%.17g|%.17g|%.17g|%d|%d|%.9g|%.9g",
             (const void *)obs->fft, obs->N, obs->numbins, obs->lobin, obs->dt, obs->rlo, obs->rhi, obs->zhi,
             obs->whi, (double)obs->highestbin, obs->numharmstages, obs->numw, (double)obs->sigma,
             (double)obs->nph);
    return buf;
}

void *api_plan(accelobs *obs)
{
    std::lock_guard<std::mutex> lk(g_api_mu);
    const std::string key = shape_key(obs);
    auto it = g_api_plans.find(key);
    if (it != g_api_plans.end()) return it->second.plan;
    PzlConfig c{};
    c.N = (double)obs->N;
    c.dt = obs->dt;
    c.numbins = obs->numbins;
    c.lobin = obs->lobin;
    c.zmax = (int)obs->zhi;
    c.wmax = (int)obs->whi;
    c.wmaxP = obs->numw > 0;
    c.numharm = 1 << (obs->numharmstages - 1);
    c.sigma = obs->sigma;
    c.rloP = 1;
    c.rlo = obs->rlo;
    c.rhiP = 1;
    c.rhi = (double)obs->highestbin;
    c.nph = obs->nph;
    c.batchsize = 1;
    c.maxhits = 1;
    void *plan = nullptr;
    if (pzl_plan_create(&c, 0, &plan) != 0) die(1, "\nGPU plan creation failed: %s\n\n", pzl_last_error());
    if (pzl_set_spectrum(plan, (const float *)obs->fft, obs->numbins) != 0)
        die(1, "\nSpectrum upload failed: %s\n\n", pzl_last_error());
    g_api_plans[key] = ApiPlan{plan, obs->fft};
    return plan;
}

float ***gen_f3Darr(long nhgts, long nrows, long ncols)
{
    float ***c = (float ***)malloc(sizeof(float **) * nhgts);
    c[0] = (float **)malloc(sizeof(float *) * nhgts * nrows);
    c[0][0] = (float *)malloc(sizeof(float) * nhgts * nrows * ncols);
    for (long i = 0; i < nhgts; i++) {
        c[i] = c[0] + i * nrows;
        for (long j = 0; j < nrows; j++) c[i][j] = c[0][0] + (i * nrows + j) * ncols;
    }
    return c;
}

// GPU scratch for the per-plane calls
struct Scratch {
    float *a = nullptr, *b = nullptr;
    long long na = 0, nb = 0;
    void *maps = nullptr;
    long long nmaps = 0;
    long long *idx = nullptr;
    long long nidx = 0;
    unsigned long long *cnt = nullptr;
    hipStream_t st = nullptr;
} g_scr;
std::mutex g_scr_mu;

template <typename T> void ensure(T **p, long long *have, long long need)
{
    if (*have >= need) return;
    if (*p) (void)hipFree(*p);
    if (hipMalloc((void **)p, sizeof(T) * need) != hipSuccess) die(1, "\nhipMalloc of %s failed\n\n", "scratch");
    *have = need;
}

void chk(hipError_t e, const char *what)
{
    if (e != hipSuccess) die(1, "\nHIP error in %s\n\n", what);
}

// insert_new_accelcand (accel_utils.c:382-480) on a GSList
accelcand *new_cand(float power, float sigma, int numharm, double r, double z, double w)
{
    accelcand *c = (accelcand *)calloc(1, sizeof(accelcand));
    c->power = power;
    c->sigma = sigma;
    c->numharm = numharm;
    c->r = r;
    c->z = z;
    c->w = w;
    return c;
}

GSList *insert_cand(GSList *list, float power, float sigma, int numharm, double rr, double zz, double ww, int *added)
{
    const double close = ACCEL_CLOSEST_R;
    *added = 0;
    if (!list) {
        GSList *n = g_slist_alloc();
        n->data = new_cand(power, sigma, numharm, rr, zz, ww);
        *added = 1;
        return n;
    }
    GSList *tmp = list, *prev = nullptr;
    while (tmp->next && ((accelcand *)tmp->data)->r < rr) {
        prev = tmp;
        tmp = tmp->next;
    }
    const double next_diff = fabs(rr - ((accelcand *)tmp->data)->r);
    const double prev_diff = prev ? fabs(rr - ((accelcand *)prev->data)->r) : close + 1.0;
    if (prev_diff < close) {
        if (((accelcand *)prev->data)->sigma < sigma) {
            free_accelcand(prev->data, nullptr);
            prev->data = new_cand(power, sigma, numharm, rr, zz, ww);
            *added = 1;
        }
        if (next_diff < close && ((accelcand *)tmp->data)->sigma < sigma) {
            free_accelcand(tmp->data, nullptr);
            if (*added) {
                list = g_slist_remove_link(list, tmp);
                g_slist_free_1(tmp);
            } else {
                tmp->data = new_cand(power, sigma, numharm, rr, zz, ww);
                *added = 1;
            }
        }
    } else if (next_diff < close) {
        if (((accelcand *)tmp->data)->sigma < sigma) {
            free_accelcand(tmp->data, nullptr);
            tmp->data = new_cand(power, sigma, numharm, rr, zz, ww);
            *added = 1;
        }
    } else {
        GSList *n = g_slist_alloc();
        n->data = new_cand(power, sigma, numharm, rr, zz, ww);
        *added = 1;
        if (!tmp->next && ((accelcand *)tmp->data)->r < rr - close) {
            tmp->next = n;
            return list;
        }
        if (prev) {
            prev->next = n;
            n->next = tmp;
        } else {
            n->next = list;
            return n;
        }
    }
    return list;
}

int twon_to_index(int n)
{
    int x = 0;
    while (n > 1) {
        n >>= 1;
        x++;
    }
    return x;
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------- statistics
double candidate_sigma(double power, int numsum, double numtrials)
{
    return pzl_candidate_sigma(power, numsum, numtrials);
}

double power_for_sigma(double sigma, int numsum, double numtrials)
{
    return pzl_power_for_sigma(sigma, numsum, numtrials);
}

// ---------------------------------------------------------------- command line
Cmdline *parseCmdline(int argc, char **argv)
{
    Cmdline *c = (Cmdline *)calloc(1, sizeof(Cmdline));
    // defaults (accelsearch_cmd.c:20-96)
    c->batchsizeP = 1;
    c->batchsize = 8;
    c->batchsizeC = 1;
    c->ncpusP = 1;
    c->ncpus = 1;
    c->ncpusC = 1;
    c->lobinP = 1;
    c->lobin = 0;
    c->lobinC = 1;
    c->numharmP = 1;
    c->numharm = 8;
    c->numharmC = 1;
    c->zmaxP = 1;
    c->zmax = 200;
    c->zmaxC = 1;
    c->sigmaP = 1;
    c->sigma = 2.0f;
    c->sigmaC = 1;
    c->floP = 1;
    c->flo = 1.0;
    c->floC = 1;
    c->fhiP = 1;
    c->fhi = 10000.0;
    c->fhiC = 1;
    c->baryvP = 1;
    c->baryv = 0.0;
    c->baryvC = 1;
    std::string full;
    for (int i = 0; i < argc; i++) {
        if (i) full += " ";
        full += argv[i];
    }
    c->full_cmd_line = strdup(full.c_str());
    std::vector<char *> rest;
    auto need = [&](int &i) -> const char * {
        if (i + 1 >= argc) die(1, "\nmissing argument after option '%s'\n\n", argv[i]);
        return argv[++i];
    };
    for (int i = 1; i < argc; i++) {
        const char *a = argv[i];
        if (a[0] != '-' || a[1] == 0) {
            rest.push_back(argv[i]);
            continue;
        }
#define INTOPT(name, P, V, C, LO, HI)                                                 \
    if (!strcmp(a, "-" name)) {                                                        \
        c->P = 1;                                                                      \
        c->V = (int)strtol(need(i), nullptr, 10);                                      \
        c->C = 1;                                                                      \
        if (c->V < (LO) || c->V > (HI)) die(2, "\nvalue of option '%s' out of range\n\n", a); \
        continue;                                                                      \
    }
#define DBLOPT(name, P, V, C, LO, HI)                                                 \
    if (!strcmp(a, "-" name)) {                                                        \
        c->P = 1;                                                                      \
        c->V = strtod(need(i), nullptr);                                               \
        c->C = 1;                                                                      \
        if (c->V < (LO) || c->V > (HI)) die(2, "\nvalue of option '%s' out of range\n\n", a); \
        continue;                                                                      \
    }
#define FLAG(name, P)             \
    if (!strcmp(a, "-" name)) {   \
        c->P = 1;                 \
        continue;                 \
    }
        // option ranges of accelsearch_cmd.c:1280-1428
        INTOPT("batchsize", batchsizeP, batchsize, batchsizeC, 1, INT_MAX)
        INTOPT("ncpus", ncpusP, ncpus, ncpusC, 1, INT_MAX)
        INTOPT("lobin", lobinP, lobin, lobinC, 0, INT_MAX)
        INTOPT("numharm", numharmP, numharm, numharmC, 1, 32)
        INTOPT("zmax", zmaxP, zmax, zmaxC, 0, 1200)
        INTOPT("wmax", wmaxP, wmax, wmaxC, 0, 4000)
        DBLOPT("rlo", rloP, rlo, rloC, 0.0, HUGE_VAL)
        DBLOPT("rhi", rhiP, rhi, rhiC, 0.0, HUGE_VAL)
        DBLOPT("flo", floP, flo, floC, 0.0, HUGE_VAL)
        DBLOPT("fhi", fhiP, fhi, fhiC, 0.0, HUGE_VAL)
        DBLOPT("baryv", baryvP, baryv, baryvC, -0.1, 0.1)
        if (!strcmp(a, "-sigma")) {
            c->sigmaP = 1;
            c->sigma = (float)strtod(need(i), nullptr);
            c->sigmaC = 1;
            if (c->sigma < 1.0f || c->sigma > 30.0f) die(2, "\nvalue of option '%s' out of range\n\n", a);
            continue;
        }
        if (!strcmp(a, "-zaplist")) {
            c->zaplistP = 1;
            c->zaplist = strdup(need(i));
            c->zaplistC = 1;
            continue;
        }
        FLAG("inmem", inmemP)
        FLAG("photon", photonP)
        FLAG("median", medianP)
        FLAG("locpow", locpowP)
        FLAG("otheropt", otheroptP)
        FLAG("noharmpolish", noharmpolishP)
        FLAG("noharmremove", noharmremoveP)
        FLAG("cuda", cudaP)
#undef INTOPT
#undef DBLOPT
#undef FLAG
        die(1, "\nunknown option '%s'\n\n", a);
    }
    // As in the reference, -flo/-fhi are on by default and take precedence over
    // -rlo/-rhi in create_accelobs (accel_utils.c:1908-1960).
    if (rest.empty()) die(1, "%s", "\nneed an input .fft file\n\n");
    c->argc = (int)rest.size();
    c->argv = (char **)malloc(sizeof(char *) * (rest.size() + 1));
    for (size_t i = 0; i < rest.size(); i++) c->argv[i] = rest[i];
    c->argv[rest.size()] = nullptr;
    return c;
}

// ---------------------------------------------------------------- accelobs
void create_accelobs(accelobs *obs, infodata *idata, Cmdline *cmd, int usemmap)
{
    memset(obs, 0, sizeof *obs);
    const char *fn = cmd->argv[0];
    const char *dot = strrchr(fn, '.');
    if (!dot) die(0, "\nInput file ('%s') must be an '.fft' or '.[s]dat' file!\n\n", fn);
    if (strcmp(dot + 1, "fft") != 0) die(0, "\nInput file ('%s') must be an '.fft' file here (no .dat input)\n\n", fn);
    obs->rootfilenm = strndup(fn, dot - fn);
    obs->use_harmonic_polishing = cmd->noharmpolishP ? 0 : 1;
    read_inf_values(obs->rootfilenm, idata);
    obs->dat_input = 0;
    obs->fftfile = fopen(fn, "rb");
    if (!obs->fftfile) die(1, "\nCannot open '%s'\n\n", fn);
    struct stat sb;
    fstat(fileno(obs->fftfile), &sb);
    obs->numbins = (long long)(sb.st_size / (long long)sizeof(fcomplex));
    if (usemmap) {
        fclose(obs->fftfile);
        obs->fftfile = nullptr;
        obs->mmap_file = open(fn, O_RDONLY);
        if (obs->mmap_file == -1) die(-1, "\nError in open() of '%s'\n\n", fn);
        obs->fft = (fcomplex *)mmap(0, sizeof(fcomplex) * obs->numbins, PROT_READ, MAP_SHARED, obs->mmap_file, 0);
        if (obs->fft == MAP_FAILED) die(-1, "\nError in mmap() of '%s'\n\n", fn);
    } else {
        // the GPU path needs the spectrum in memory: read it whole
        obs->fft = (fcomplex *)malloc(sizeof(fcomplex) * obs->numbins);
        if (fread(obs->fft, sizeof(fcomplex), obs->numbins, obs->fftfile) != (size_t)obs->numbins)
            die(1, "\nShort read of '%s'\n\n", fn);
    }
    if (cmd->zmax % ACCEL_DZ) cmd->zmax = (cmd->zmax / ACCEL_DZ + 1) * ACCEL_DZ;
    if (cmd->wmaxP && cmd->wmax % ACCEL_DW) cmd->wmax = (cmd->wmax / ACCEL_DW + 1) * ACCEL_DW;
    obs->N = (long long)idata->N;
    obs->dt = idata->dt;
    obs->T = idata->dt * idata->N;
    if (cmd->photonP) {
        obs->nph = obs->fft[0].r;
        printf("Normalizing powers using %.0f photons.\n\n", obs->nph);
    } else {
        obs->nph = 0.0;
        if (cmd->locpowP) die(1, "%s", "\n-locpow normalisation is not provided by this engine\n\n");
        obs->norm_type = 0;
    }
    obs->lobin = cmd->lobin;
    if (obs->lobin > 0) {
        obs->nph = 0.0;
        if (cmd->lobin > obs->numbins - 1)
            die(1, "%s", "\n'lobin' is greater than the total number of\n   frequencies in the data set.  Exiting.\n\n");
    }
    if (cmd->numharm != 1 && cmd->numharm != 2 && cmd->numharm != 4 && cmd->numharm != 8 && cmd->numharm != 16 &&
        cmd->numharm != 32)
        die(1, "\n'numharm' must be a power-of-two!  Exiting%s\n\n");
    // geometry and thresholds from the engine (one implementation)
    PzlConfig c{};
    c.N = (double)obs->N;
    c.dt = obs->dt;
    c.numbins = obs->numbins;
    c.lobin = obs->lobin;
    c.zmax = cmd->zmax;
    c.wmax = cmd->wmaxP ? cmd->wmax : 0;
    c.wmaxP = cmd->wmaxP;
    c.numharm = cmd->numharm;
    c.sigma = cmd->sigma;
    c.floP = cmd->floP;
    c.fhiP = cmd->fhiP;
    c.rloP = cmd->rloP;
    c.rhiP = cmd->rhiP;
    c.flo = cmd->flo;
    c.fhi = cmd->fhi;
    c.rlo = cmd->rlo;
    c.rhi = cmd->rhi;
    PzlGeometry g;
    if (pzl_geometry_from_config(&c, &g) != 0) die(1, "\n%s\n\n", pzl_last_error());
    obs->numharmstages = g.numharmstages;
    obs->dz = ACCEL_DZ;
    obs->numz = g.numz;
    obs->whi = g.whi;
    obs->wlo = g.wlo;
    obs->dw = cmd->wmaxP ? ACCEL_DW : 0.0;
    obs->numw = g.numw;
    const size_t rootlen = strlen(obs->rootfilenm) + 45;
    obs->candnm = (char *)calloc(rootlen, 1);
    obs->accelnm = (char *)calloc(rootlen, 1);
    obs->workfilenm = (char *)calloc(rootlen, 1);
    if (obs->numw) {
        snprintf(obs->candnm, rootlen, "%s_ACCEL_%d_JERK_%d.cand", obs->rootfilenm, cmd->zmax, cmd->wmax);
        snprintf(obs->accelnm, rootlen, "%s_ACCEL_%d_JERK_%d", obs->rootfilenm, cmd->zmax, cmd->wmax);
        snprintf(obs->workfilenm, rootlen, "%s_ACCEL_%d_JERK_%d.txtcand", obs->rootfilenm, cmd->zmax, cmd->wmax);
    } else {
        snprintf(obs->candnm, rootlen, "%s_ACCEL_%d.cand", obs->rootfilenm, cmd->zmax);
        snprintf(obs->accelnm, rootlen, "%s_ACCEL_%d", obs->rootfilenm, cmd->zmax);
        snprintf(obs->workfilenm, rootlen, "%s_ACCEL_%d.txtcand", obs->rootfilenm, cmd->zmax);
    }
    obs->workfile = fopen(obs->workfilenm, "w");
    obs->numbetween = ACCEL_NUMBETWEEN;
    obs->rlo = g.rlo;
    obs->rhi = g.rhi;
    obs->highestbin = g.highestbin;
    obs->dr = ACCEL_DR;
    obs->zhi = g.zhi;
    obs->zlo = g.zlo;
    obs->sigma = cmd->sigma;
    obs->powcut = (float *)malloc(obs->numharmstages * sizeof(float));
    obs->numindep = (long long *)malloc(obs->numharmstages * sizeof(long long));
    for (int ii = 0; ii < obs->numharmstages; ii++) {
        obs->powcut[ii] = g.powcut[ii];
        obs->numindep[ii] = g.numindep[ii];
    }
    obs->numzap = 0;
    obs->maxkernlen = g.maxkernlen;
    obs->fftlen = g.fftlen;
    obs->corr_uselen = g.corr_uselen;
    // the GPU engine handles every shape; the in-memory variant is not used
    obs->inmem = 0;
    obs->ffdotplane = nullptr;
}

void free_accelobs(accelobs *obs)
{
    // The plane API's plans are keyed by the spectrum's address: drop this
    // spectrum's plans before its memory is released, so that another file
    // mapped or allocated at the same address later cannot find them.
    {
        std::lock_guard<std::mutex> lk(g_api_mu);
        for (auto it = g_api_plans.begin(); it != g_api_plans.end();) {
            if (it->second.fft == (const void *)obs->fft) {
                pzl_plan_destroy(it->second.plan);
                it = g_api_plans.erase(it);
            } else {
                ++it;
            }
        }
    }
    if (obs->mmap_file) {
        munmap(obs->fft, sizeof(fcomplex) * obs->numbins);
        close(obs->mmap_file);
    } else {
        free(obs->fft);
        if (obs->fftfile) fclose(obs->fftfile);
    }
    if (obs->workfile) fclose(obs->workfile);
    free(obs->powcut);
    free(obs->numindep);
    free(obs->rootfilenm);
    free(obs->candnm);
    free(obs->accelnm);
    free(obs->workfilenm);
    obs->fft = nullptr;
    obs->workfile = nullptr;
}

// ---------------------------------------------------------------- subharminfos
static void init_subharminfo(int numharm, int harmnum, int zmax, int wmax, subharminfo *shi, accelobs *obs)
{
    const double hf = (double)harmnum / (double)numharm;
    shi->numharm = numharm;
    shi->harmnum = harmnum;
    shi->zmax = pzl_calc_required_z(hf, zmax);
    shi->wmax = pzl_calc_required_w(hf, wmax);
    shi->rinds = nullptr;
    shi->zinds = nullptr;
    if (numharm > 1) {
        shi->rinds = (unsigned short *)malloc(obs->corr_uselen * sizeof(unsigned short));
        shi->zinds = (unsigned short *)malloc(obs->corr_uselen * sizeof(unsigned short));
    }
    int fftlen;
    if (numharm == 1 && harmnum == 1) {
        fftlen = obs->fftlen;
    } else {
        const int bins_needed = (int)ceil(obs->corr_uselen * hf) + 2;
        const int end_effects =
            2 * ACCEL_NUMBETWEEN *
            pzl_w_resp_halfwidth(pzl_calc_required_z(hf, zmax), pzl_calc_required_w(hf, wmax));
        fftlen = pzl_next_good_fftlen(bins_needed + end_effects);
    }
    shi->numkern_zdim = (shi->zmax / ACCEL_DZ) * 2 + 1;
    shi->numkern_wdim = (shi->wmax / ACCEL_DW) * 2 + 1;
    shi->numkern = shi->numkern_zdim * shi->numkern_wdim;
    shi->kern = (kernel **)malloc(sizeof(kernel *) * shi->numkern_wdim);
    shi->kern[0] = (kernel *)malloc(sizeof(kernel) * shi->numkern);
    for (int ii = 1; ii < shi->numkern_wdim; ii++) shi->kern[ii] = shi->kern[ii - 1] + shi->numkern_zdim;
    for (int ii = 0; ii < shi->numkern_wdim; ii++)
        for (int jj = 0; jj < shi->numkern_zdim; jj++) {
            kernel *k = &shi->kern[ii][jj];
            k->z = -shi->zmax + jj * ACCEL_DZ;
            k->w = -shi->wmax + ii * ACCEL_DW;
            k->fftlen = fftlen;
            k->numbetween = ACCEL_NUMBETWEEN;
            k->kern_half_width = pzl_w_resp_halfwidth((double)k->z, (double)k->w);
            k->numgoodbins = fftlen - 2 * ACCEL_NUMBETWEEN * k->kern_half_width;
            k->data = nullptr;  // FFT'd kernels live in GPU memory
        }
}

subharminfo **create_subharminfos(accelobs *obs)
{
    subharminfo **shis = (subharminfo **)malloc(obs->numharmstages * sizeof(subharminfo *));
    shis[0] = (subharminfo *)malloc(2 * sizeof(subharminfo));
    init_subharminfo(1, 1, (int)obs->zhi, (int)obs->whi, &shis[0][0], obs);
    for (int ii = 1; ii < obs->numharmstages; ii++) {
        const int harmtosum = 1 << ii;
        shis[ii] = (subharminfo *)malloc(harmtosum * sizeof(subharminfo));
        for (int jj = 1; jj < harmtosum; jj += 2)
            init_subharminfo(harmtosum, jj, (int)obs->zhi, (int)obs->whi, &shis[ii][jj - 1], obs);
    }
    return shis;
}

static void free_subharminfo(subharminfo *shi)
{
    free(shi->kern[0]);
    free(shi->kern);
    if (shi->numharm > 1) {
        free(shi->rinds);
        free(shi->zinds);
    }
}

void free_subharminfos(accelobs *obs, subharminfo **shis)
{
    for (int ii = 1; ii < obs->numharmstages; ii++) {
        const int harmtosum = 1 << ii;
        for (int jj = 1; jj < harmtosum; jj += 2) free_subharminfo(&shis[ii][jj - 1]);
        free(shis[ii]);
    }
    free_subharminfo(&shis[0][0]);
    free(shis[0]);
    free(shis);
}

// ---------------------------------------------------------------- amplitudes
fcomplex *get_fourier_amplitudes(long long lobin, int numbins, accelobs *obs)
{
    fcomplex *out = (fcomplex *)calloc(numbins, sizeof(fcomplex));
    long long offset = 0;
    if (lobin - obs->lobin < 0) offset = llabs(lobin - obs->lobin);
    const long long firstbin = (lobin - obs->lobin) + offset;
    long long newnumbins = numbins - offset;
    if (firstbin + newnumbins > obs->numbins) newnumbins -= firstbin + newnumbins - obs->numbins;
    if (newnumbins > 0) memcpy(out + offset, obs->fft + firstbin, sizeof(fcomplex) * newnumbins);
    return out;
}

// ---------------------------------------------------------------- planes
ffdotpows *subharm_fderivs_vol(int numharm, int harmnum, double fullrlo, double fullrhi, subharminfo *shi,
                               accelobs *obs)
{
    (void)fullrhi;  // the plane spans corr_uselen fundamental points from fullrlo
    void *plan = api_plan(obs);
    int idx = 0;
    if (numharm > 1) idx = (numharm >> 1) + (harmnum - 1) / 2;
    PzlHarmInfo hi;
    if (pzl_plan_harm(plan, idx, &hi) != 0) die(1, "\n%s\n\n", pzl_last_error());
    int numrs = 0;
    long long rlo = 0;
    if (pzl_compute_plane_at(plan, fullrlo, idx, nullptr, 0, &numrs, &rlo) != 0)
        die(1, "\nGPU plane failed: %s\n\n", pzl_last_error());
    ffdotpows *ffd = (ffdotpows *)malloc(sizeof(ffdotpows));
    const double hf = (double)harmnum / (double)numharm;
    ffd->rlo = rlo;
    ffd->zlo = pzl_calc_required_z(hf, obs->zlo);
    ffd->wlo = pzl_calc_required_w(hf, obs->wlo);
    ffd->numrs = numrs;
    ffd->numzs = hi.numz;
    ffd->numws = hi.numw;
    ffd->powers = gen_f3Darr(hi.numw, hi.numz, numrs);
    if (pzl_compute_plane_at(plan, fullrlo, idx, ffd->powers[0][0], (long long)hi.numw * hi.numz * numrs, &numrs,
                             &rlo) != 0)
        die(1, "\nGPU plane failed: %s\n\n", pzl_last_error());
    if (numharm > 1) {
        for (int ii = 0; ii < obs->corr_uselen; ii++) shi->rinds[ii] = (unsigned short)pzl_rind(hf, fullrlo, rlo, ii);
        for (int ii = 0; ii < obs->numz; ii++)
            shi->zinds[ii] =
                (unsigned short)pzl_index_from_z(pzl_calc_required_z(hf, obs->zlo + ii * ACCEL_DZ), ffd->zlo);
    }
    ffd->rinds = shi->rinds;
    ffd->zinds = shi->zinds;
    return ffd;
}

void free_ffdotpows(ffdotpows *ffd)
{
    free(ffd->powers[0][0]);
    free(ffd->powers[0]);
    free(ffd->powers);
    free(ffd);
}

void add_ffdotpows(ffdotpows *fundamental, ffdotpows *subharmonic, int numharm, int harmnum)
{
    std::lock_guard<std::mutex> lk(g_scr_mu);
    const double hf = (double)harmnum / (double)numharm;
    const long long nf = (long long)fundamental->numws * fundamental->numzs * fundamental->numrs;
    const long long ns = (long long)subharmonic->numws * subharmonic->numzs * subharmonic->numrs;
    std::vector<int> winds(fundamental->numws);
    for (int ii = 0; ii < fundamental->numws; ii++)
        winds[ii] = pzl_index_from_w(pzl_calc_required_w(hf, fundamental->wlo + ii * ACCEL_DW), subharmonic->wlo);
    if (!g_scr.st) chk(hipStreamCreateWithFlags(&g_scr.st, hipStreamNonBlocking), "stream");
    ensure(&g_scr.a, &g_scr.na, nf);
    ensure(&g_scr.b, &g_scr.nb, ns);
    const long long mapbytes = sizeof(int) * fundamental->numws + 2 * (fundamental->numzs + fundamental->numrs);
    long long have = g_scr.nmaps;
    char *maps = (char *)g_scr.maps;
    ensure(&maps, &have, mapbytes);
    g_scr.maps = maps;
    g_scr.nmaps = have;
    int *d_w = (int *)maps;
    uint16_t *d_z = (uint16_t *)(maps + sizeof(int) * fundamental->numws);
    uint16_t *d_r = d_z + fundamental->numzs;
    chk(hipMemcpyAsync(g_scr.a, fundamental->powers[0][0], sizeof(float) * nf, hipMemcpyHostToDevice, g_scr.st), "copy");
    chk(hipMemcpyAsync(g_scr.b, subharmonic->powers[0][0], sizeof(float) * ns, hipMemcpyHostToDevice, g_scr.st), "copy");
    chk(hipMemcpyAsync(d_w, winds.data(), sizeof(int) * fundamental->numws, hipMemcpyHostToDevice, g_scr.st), "copy");
    chk(hipMemcpyAsync(d_z, subharmonic->zinds, 2 * fundamental->numzs, hipMemcpyHostToDevice, g_scr.st), "copy");
    chk(hipMemcpyAsync(d_r, subharmonic->rinds, 2 * fundamental->numrs, hipMemcpyHostToDevice, g_scr.st), "copy");
    chk(pzl_launch_add_plane(g_scr.a, fundamental->numws, fundamental->numzs, fundamental->numrs, g_scr.b,
                             subharmonic->numzs, subharmonic->numrs, d_w, d_z, d_r, g_scr.st),
        "add_ffdotpows");
    chk(hipMemcpyAsync(fundamental->powers[0][0], g_scr.a, sizeof(float) * nf, hipMemcpyDeviceToHost, g_scr.st),
        "copy");
    chk(hipStreamSynchronize(g_scr.st), "sync");
}

void add_ffdotpows_ptrs(ffdotpows *fundamental, ffdotpows *subharmonic, int numharm, int harmnum)
{
    add_ffdotpows(fundamental, subharmonic, numharm, harmnum);
}

GSList *search_ffdotpows(ffdotpows *ffdot, int numharm, accelobs *obs, GSList *cands)
{
    std::vector<long long> idx;
    const float powcut = obs->powcut[twon_to_index(numharm)];
    const long long numindep = obs->numindep[twon_to_index(numharm)];
    const long long n = (long long)ffdot->numws * ffdot->numzs * ffdot->numrs;
    {
        std::lock_guard<std::mutex> lk(g_scr_mu);
        if (!g_scr.st) chk(hipStreamCreateWithFlags(&g_scr.st, hipStreamNonBlocking), "stream");
        ensure(&g_scr.a, &g_scr.na, n);
        ensure(&g_scr.idx, &g_scr.nidx, n);
        long long one = g_scr.cnt ? 1 : 0;
        ensure(&g_scr.cnt, &one, 1);
        chk(hipMemcpyAsync(g_scr.a, ffdot->powers[0][0], sizeof(float) * n, hipMemcpyHostToDevice, g_scr.st), "copy");
        chk(hipMemsetAsync(g_scr.cnt, 0, sizeof(unsigned long long), g_scr.st), "memset");
        chk(pzl_launch_threshold(g_scr.a, n, powcut, g_scr.idx, g_scr.cnt, n, g_scr.st), "search_ffdotpows");
        unsigned long long cnt = 0;
        chk(hipMemcpyAsync(&cnt, g_scr.cnt, sizeof cnt, hipMemcpyDeviceToHost, g_scr.st), "copy");
        chk(hipStreamSynchronize(g_scr.st), "sync");
        idx.resize(cnt);
        if (cnt) chk(hipMemcpy(idx.data(), g_scr.idx, sizeof(long long) * cnt, hipMemcpyDeviceToHost), "copy");
    }
    std::sort(idx.begin(), idx.end());  // the reference's w, z, r scan order
    const float *pw = ffdot->powers[0][0];
    for (long long i : idx) {
        const int kk = (int)(i % ffdot->numrs), jj = (int)((i / ffdot->numrs) % ffdot->numzs);
        const int ii = (int)(i / ((long long)ffdot->numrs * ffdot->numzs));
        const float pow = pw[i];
        const float sig = (float)pzl_candidate_sigma(pow, numharm, (double)numindep);
        const double rr = (ffdot->rlo + kk * (double)ACCEL_DR) / (double)numharm;
        const double zz = (ffdot->zlo + jj * (double)ACCEL_DZ) / (double)numharm;
        const double ww = (ffdot->wlo + ii * (double)ACCEL_DW) / (double)numharm;
        int added = 0;
        cands = insert_cand(cands, pow, sig, numharm, rr, zz, ww, &added);
        if (added && !obs->dat_input && obs->workfile)
            fprintf(obs->workfile, "%-7.2f  %-7.4f  %-2d  %-14.4f  %-14.9f  %-10.4f %-10.4f\n", pow, sig, numharm, rr,
                    rr / obs->T, zz, ww);
    }
    return cands;
}

// ---------------------------------------------------------------- candidate lists
void free_accelcand(gpointer data, gpointer user_data)
{
    (void)user_data;
    accelcand *c = (accelcand *)data;
    if (c->pows) {
        free(c->pows);
        free(c->hirs);
        free(c->hizs);
        free(c->hiws);
        free(c->derivs);
    }
    free(c);
}

void print_accelcand(gpointer data, gpointer user_data)
{
    (void)user_data;
    accelcand *c = (accelcand *)data;
    printf("sigma: %-7.4f  pow: %-7.2f  harm: %-2d  r: %-14.4f  z: %-10.4f  w: %-10.2f\n", c->sigma, c->power,
           c->numharm, c->r, c->z, c->w);
}

static gint compare_accelcand_sigma(gconstpointer ca, gconstpointer cb)
{
    const accelcand *a = (const accelcand *)ca, *b = (const accelcand *)cb;
    const int r = (a->sigma < b->sigma) - (a->sigma > b->sigma);
    if (r) return r;
    return (a->power < b->power) - (a->power > b->power);
}

GSList *sort_accelcands(GSList *list) { return g_slist_sort(list, compare_accelcand_sigma); }

GSList *eliminate_harmonics(GSList *cands, int *numcands)
{
    static const double ratios[][2] = {{3, 2},  {5, 2},  {2, 3},  {4, 3}, {5, 3}, {3, 4}, {5, 4},  {2, 5},
                                       {3, 5},  {4, 5},  {5, 6},  {2, 7}, {3, 7}, {4, 7}, {3, 8},  {5, 8},
                                       {2, 9},  {3, 10}, {2, 11}, {3, 11}, {2, 13}, {3, 13}, {2, 15}};
    const int maxharm = 16;
    const double tooclose = 1.5;
    GSList *cur = cands;
    while (cur && cur->next) {
        const accelcand *c = (const accelcand *)cur->data;
        GSList *other = cur->next;
        do {
            const accelcand *o = (const accelcand *)other->data;
            bool remove = false;
            for (int ii = 1; ii <= maxharm && !remove; ii++)
                if (fabs(c->r / ii - o->r) < tooclose) remove = true;
            for (int ii = 1; ii <= maxharm && !remove; ii++)
                if (fabs(c->r * ii - o->r) < tooclose) remove = true;
            for (size_t k = 0; k < sizeof ratios / sizeof ratios[0] && !remove; k++)
                if (fabs(c->r * ratios[k][0] / ratios[k][1] - o->r) < tooclose) remove = true;
            if (remove) {
                GSList *dead = other;
                other = other->next;
                free_accelcand(dead->data, nullptr);
                cands = g_slist_remove_link(cands, dead);
                g_slist_free_1(dead);
                *numcands = *numcands - 1;
            } else {
                other = other->next;
            }
        } while (other);
        if (cur->next) cur = cur->next;
    }
    return cands;
}

// ---------------------------------------------------------------- the pipeline stages
void accelsearch_CPU1(int argc, char *argv[], subharminfo ***subharminfs_ptr, accelobs *obs_ptr,
                      infodata *idata_ptr, Cmdline **cmd_ptr)
{
    if (argc == 1) die(1, "%s", "\nusage: accelsearch [options] file.fft\n\n");
    Cmdline *cmd = parseCmdline(argc, argv);
    *cmd_ptr = cmd;
    create_accelobs(obs_ptr, idata_ptr, cmd, 1);
    if (cmd->zaplistP) die(1, "%s", "\n-zaplist applies to .dat input only\n\n");
    printf("\nGenerating correlation kernels:\n");
    *subharminfs_ptr = create_subharminfos(obs_ptr);
    if (cmd->otheroptP) die(-1, "%s", "otheroptP is not supported\n");
    if (!obs_ptr->dat_input) printf("  Working candidates in a test format are in '%s'.\n\n", obs_ptr->workfilenm);
}

int accelsearch_GPU(accelobs obs, subharminfo **subharminfs, GSList **cands_ptr, Cmdline *cmd)
{
    PzlConfig c = config_from(&obs, cmd);
    void *plan = nullptr;
    if (pzl_plan_create(&c, 0, &plan) != 0) die(1, "\nGPU plan creation failed: %s\n\n", pzl_last_error());
    if (pzl_set_spectrum(plan, (const float *)obs.fft, obs.numbins) != 0)
        die(1, "\nSpectrum upload failed: %s\n\n", pzl_last_error());
    long long ncand = 0;
    int rc = pzl_search(plan, 0, -1, nullptr, 0, &ncand, nullptr);
    if (rc == PZL_ERR_OVERFLOW) {
        printf("\nThe hit buffer overflowed; no candidates are returned.\n\n");
        pzl_plan_destroy(plan);
        free_subharminfos(&obs, subharminfs);
        *cands_ptr = nullptr;
        return 1;
    }
    if (rc != 0) die(1, "\nGPU search failed: %s\n\n", pzl_last_error());
    std::vector<PzlCand> v(std::max<long long>(ncand, 1));
    if (pzl_get_cands(plan, v.data(), ncand, &ncand) != 0) die(1, "\nGPU search failed: %s\n\n", pzl_last_error());
    GSList *list = nullptr, *tail = nullptr;
    for (long long i = 0; i < ncand; i++) {
        GSList *n = g_slist_alloc();
        n->data = new_cand(v[i].power, v[i].sigma, v[i].numharm, v[i].r, v[i].z, v[i].w);
        if (tail) tail->next = n;
        else list = n;
        tail = n;
    }
    pzl_plan_destroy(plan);
    free_subharminfos(&obs, subharminfs);
    *cands_ptr = list;
    return 0;
}

void accelsearch_CPU2(GSList **cands, accelobs *obs, infodata *idata, Cmdline *cmd)
{
    int numcands = (int)g_slist_length(*cands);
    printf("numcands after while:%d\n", numcands);
    if (numcands) {
        *cands = sort_accelcands(*cands);
        if (cmd->numharm > 1 && !cmd->noharmremoveP) *cands = eliminate_harmonics(*cands, &numcands);
        if (pzl_optimize_accelcands(*cands, obs) != 0) die(1, "\nThe spectrum is not in memory\n\n");
        // the fundamentals' properties, with the search-grid errors
        // (accelsearch_lib.c:932-948), then the text and binary outputs
        std::vector<fourierprops> props((size_t)numcands);
        int ii = 0;
        for (GSList *p = *cands; p; p = p->next, ii++) {
            const accelcand *c = (const accelcand *)p->data;
            calc_props(c->derivs[0], c->r, c->z, c->w, &props[ii]);
            props[ii].rerr = (float)(PZL_DR) / c->numharm;
            props[ii].zerr = (float)(PZL_DZ) / c->numharm;
            props[ii].werr = (float)(PZL_DW) / c->numharm;
        }
        output_fundamentals(props.data(), *cands, obs, idata);
        output_harmonics(*cands, obs, idata);
        FILE *f = fopen(obs->candnm, "wb");
        if (!f || fwrite(props.data(), sizeof(fourierprops), (size_t)numcands, f) != (size_t)numcands)
            die(1, "\nCannot write '%s'\n\n", obs->candnm);
        fclose(f);
        printf("\n\n");
    } else {
        printf("No candidates above sigma = %.2f were found.\n\n", obs->sigma);
    }
    printf("Final candidates in binary format are in '%s'.\n", obs->candnm);
    printf("Final Candidates in a text format are in '%s'.\n\n", obs->accelnm);
    free_accelobs(obs);
    g_slist_foreach(*cands, free_accelcand, nullptr);
    g_slist_free(*cands);
    *cands = nullptr;
}

}  // extern "C"
