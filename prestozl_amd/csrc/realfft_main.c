/* realfft_main.c -- the realfft command (realfft.c:31-275), in-core path, with
 * the transform on the GPU (pzl_realfft).  For each file: a name ending in
 * .fft is inverse-transformed to .dat, anything else (.dat) forward to .fft,
 * in the input's directory or -outdir; -fwd / -inv force the direction (the
 * suffixes stay as they are, as in the reference), -delete removes the input.
 * The whole series is held in device memory, so the reference's out-of-core
 * two-pass path (-disk, files above MAXREALFFT) has no counterpart here. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "pzl_accel.h"

static void usage(void)
{
    fprintf(stderr, "usage: realfft [-fwd | -inv] [-delete] [-outdir dir] file.dat|file.fft ...\n");
}

int main(int argc, char *argv[])
{
    int force = 0, del = 0, nfiles = 0, rc = 0;
    const char *outdir = NULL;
    for (int i = 1; i < argc; i++) {
        if (!strcmp(argv[i], "-fwd")) force = -1;
        else if (!strcmp(argv[i], "-inv")) force = 1;
        else if (!strcmp(argv[i], "-delete")) del = 1;
        else if (!strcmp(argv[i], "-outdir") && i + 1 < argc) outdir = argv[++i];
        else if (argv[i][0] == '-') {
            usage();
            return 1;
        } else nfiles++;
    }
    if (!nfiles) {
        usage();
        return 1;
    }
    for (int i = 1; i < argc; i++) {
        if (argv[i][0] == '-') {
            if (!strcmp(argv[i], "-outdir")) i++;
            continue;
        }
        const char *path = argv[i];
        const char *base = strrchr(path, '/');
        base = base ? base + 1 : path;
        const char *dot = strrchr(base, '.');
        const size_t rootlen = dot ? (size_t)(dot - base) : strlen(base);
        const int is_fft = dot && !strcmp(dot, ".fft");
        int isign = is_fft ? 1 : -1;
        if (force) isign = force;
        const char *dir = outdir;
        char *indir = NULL;
        if (!dir) {
            const size_t dl = (size_t)(base - path);
            indir = (char *)calloc(dl + 2, 1);
            if (dl) memcpy(indir, path, dl - 1);
            else strcpy(indir, ".");
            dir = indir;
        }
        char *out = (char *)calloc(strlen(dir) + rootlen + 8, 1);
        sprintf(out, "%s/%.*s.%s", dir, (int)rootlen, base, is_fft ? "dat" : "fft");
        free(indir);

        FILE *f = fopen(path, "rb");
        if (!f) {
            fprintf(stderr, "realfft: cannot open '%s'\n", path);
            free(out);
            rc = 1;
            continue;
        }
        fseek(f, 0, SEEK_END);
        const long long bytes = ftell(f);
        fseek(f, 0, SEEK_SET);
        if (bytes % (isign == -1 ? (long long)sizeof(float) : 2LL * (long long)sizeof(float))) {
            fprintf(stderr, "realfft: '%s' does not hold whole %s values; skipped\n", path,
                    isign == -1 ? "float" : "complex");
            fclose(f);
            free(out);
            rc = 1;
            continue;
        }
        const long long n = bytes / (long long)sizeof(float);
        float *data = (float *)malloc(sizeof(float) * (size_t)(n ? n : 1));
        if (!data || (long long)fread(data, sizeof(float), (size_t)n, f) != n) {
            fprintf(stderr, "realfft: cannot read '%s'\n", path);
            fclose(f);
            free(data);
            free(out);
            rc = 1;
            continue;
        }
        fclose(f);
        printf("%s: %lld floats, %s transform -> '%s'\n", path, n, isign == -1 ? "forward" : "inverse", out);
        if (pzl_realfft(data, n, isign) != 0) {
            fprintf(stderr, "realfft: %s\n", pzl_last_error());
            free(data);
            free(out);
            rc = 1;
            continue;
        }
        FILE *o = fopen(out, "wb");
        if (!o || (long long)fwrite(data, sizeof(float), (size_t)n, o) != n) {
            fprintf(stderr, "realfft: cannot write '%s'\n", out);
            rc = 1;
        }
        if (o) fclose(o);
        if (del && rc == 0) remove(path);
        free(data);
        free(out);
    }
    return rc;
}
