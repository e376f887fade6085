/* accelsearch_main.c -- the accelsearch_cu command-line driver
 * (accelsearch_cu.c:22-34): the three pipeline stages of the accel.h API,
 * linked against libaccelsearch_lib.so.
 *
 *   accelsearch [options] file.fft
 *
 * Exit status: 0, or 1 when the GPU stage reported a hit-buffer overflow
 * (the reference ignores that return value; scripts here can see it).
 */
#include "accel.h"

int main(int argc, char *argv[])
{
    subharminfo **subharminfs;
    accelobs obs;
    infodata idata;
    GSList *cands = NULL;
    Cmdline *cmd;

    accelsearch_CPU1(argc, argv, &subharminfs, &obs, &idata, &cmd);
    int too_large = accelsearch_GPU(obs, subharminfs, &cands, cmd);
    accelsearch_CPU2(&cands, &obs, &idata, cmd);
    return too_large ? 1 : 0;
}
