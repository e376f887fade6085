/* accelsearch_pipeline_main.c -- the three-stage pipeline of
 * bin/accelsearch_pipeline_cu.py (process_tasks / task_t1 / task_t2 /
 * task_t3, :296-407) in C, over the accel.h drop-in: one thread runs
 * accelsearch_CPU1 on each file in turn, a second accelsearch_GPU, a third
 * accelsearch_CPU2, handing each file's state on through one-slot queues as
 * the reference does (queue.Queue(maxsize=1)).  Different files are in
 * different stages at once, so this is also the check that the drop-in is
 * safe under that overlap.
 *
 *   accelsearch_pipeline [options] -- file1.fft file2.fft ...
 *
 * The options are given to every file (argv = options + file, as the
 * reference's task arguments).  Exit status: the number of files whose GPU
 * stage reported a hit-buffer overflow (their CPU2 is skipped, as task_t2
 * does for a non-zero error code).
 */
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "accel.h"

typedef struct {
    int file;  /* index, -1 = end of stream */
    subharminfo **subharminfs;
    accelobs obs;
    infodata idata;
    Cmdline *cmd;
    GSList *cands;
} Job;

/* A one-slot queue: put blocks while the slot is full. */
typedef struct {
    pthread_mutex_t mu;
    pthread_cond_t cv;
    int full;
    Job *job;
} Slot;

static void slot_init(Slot *s)
{
    pthread_mutex_init(&s->mu, NULL);
    pthread_cond_init(&s->cv, NULL);
    s->full = 0;
    s->job = NULL;
}

static void slot_put(Slot *s, Job *j)
{
    pthread_mutex_lock(&s->mu);
    while (s->full) pthread_cond_wait(&s->cv, &s->mu);
    s->job = j;
    s->full = 1;
    pthread_cond_broadcast(&s->cv);
    pthread_mutex_unlock(&s->mu);
}

static Job *slot_get(Slot *s)
{
    pthread_mutex_lock(&s->mu);
    while (!s->full) pthread_cond_wait(&s->cv, &s->mu);
    Job *j = s->job;
    s->full = 0;
    pthread_cond_broadcast(&s->cv);
    pthread_mutex_unlock(&s->mu);
    return j;
}

static int g_nopt, g_nfiles, g_failed;
static char **g_opts, **g_files;
static Slot g_q12, g_q23;

static void *stage1(void *arg)
{
    (void)arg;
    for (int f = 0; f < g_nfiles; f++) {
        Job *j = (Job *)calloc(1, sizeof(Job));
        int argc = g_nopt + 2;
        char **argv = (char **)calloc((size_t)argc + 1, sizeof(char *));
        argv[0] = (char *)"accelsearch";
        for (int k = 0; k < g_nopt; k++) argv[k + 1] = g_opts[k];
        argv[argc - 1] = g_files[f];
        j->file = f;
        accelsearch_CPU1(argc, argv, &j->subharminfs, &j->obs, &j->idata, &j->cmd);
        slot_put(&g_q12, j);
    }
    Job *end = (Job *)calloc(1, sizeof(Job));
    end->file = -1;
    slot_put(&g_q12, end);
    return NULL;
}

static void *stage2(void *arg)
{
    (void)arg;
    for (;;) {
        Job *j = slot_get(&g_q12);
        if (j->file < 0) {
            slot_put(&g_q23, j);
            return NULL;
        }
        if (accelsearch_GPU(j->obs, j->subharminfs, &j->cands, j->cmd) != 0) {
            fprintf(stderr, "accelsearch_pipeline: '%s': hit-buffer overflow, no candidates\n", g_files[j->file]);
            g_failed++;
            free(j);
            continue;
        }
        slot_put(&g_q23, j);
    }
}

static void *stage3(void *arg)
{
    (void)arg;
    for (;;) {
        Job *j = slot_get(&g_q23);
        if (j->file < 0) {
            free(j);
            return NULL;
        }
        accelsearch_CPU2(&j->cands, &j->obs, &j->idata, j->cmd);
        free(j);
    }
}

int main(int argc, char *argv[])
{
    int sep = -1;
    for (int k = 1; k < argc; k++)
        if (!strcmp(argv[k], "--")) {
            sep = k;
            break;
        }
    if (sep < 0 || sep == argc - 1) {
        fprintf(stderr, "usage: accelsearch_pipeline [options] -- file1.fft [file2.fft ...]\n");
        return 2;
    }
    g_opts = argv + 1;
    g_nopt = sep - 1;
    g_files = argv + sep + 1;
    g_nfiles = argc - sep - 1;
    slot_init(&g_q12);
    slot_init(&g_q23);
    pthread_t t1, t2, t3;
    pthread_create(&t1, NULL, stage1, NULL);
    pthread_create(&t2, NULL, stage2, NULL);
    pthread_create(&t3, NULL, stage3, NULL);
    pthread_join(t1, NULL);
    pthread_join(t2, NULL);
    pthread_join(t3, NULL);
    return g_failed;
}
