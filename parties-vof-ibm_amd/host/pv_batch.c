/*
 * pv_batch.c -- many independent cases of uneven size advanced together on
 * one GPU (no counterpart in the reference).
 *
 * Each case is an ordinary handle with its own HIP stream; pv_batch_rk3
 * drives every case from its own host thread, so the kernels of different
 * cases overlap on the device while each case runs exactly the code path it
 * runs alone (its results are bit for bit those of a solo run).  With
 * pv_batch_set_threads(bt, T), 0 < T < cases, T host threads share the
 * cases instead: thread t steps cases t, t + T, ... one step each in turn.
 */
#include "pv_host.h"
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

struct pv_batch {
    int n, nthreads;     /* nthreads 0: one host thread per case */
    pv_bag **bags;
};

typedef struct {
    pv_bag *bag;         /* one case, all steps (one thread per case) */
    pv_batch *bt;        /* or cases first, first + stride, ... step by step */
    int first, stride, nsteps, rc, bad;
    char err[512];
} batch_job;

int pv_batch_create(const pv_params *prms, int ncases, int device, pv_batch **out)
{
    pv_batch *bt;
    int k, rc;
    if (!prms || ncases < 1 || !out) return pv_set_error(PV_EINVAL, "bad batch arguments");
    *out = NULL;
    bt = (pv_batch *)calloc(1, sizeof *bt);
    if (!bt) return pv_set_error(PV_ENOMEM, "host allocation failed");
    bt->bags = (pv_bag **)calloc(ncases, sizeof(pv_bag *));
    if (!bt->bags) { free(bt); return pv_set_error(PV_ENOMEM, "host allocation failed"); }
    bt->n = ncases;
    for (k = 0; k < ncases; k++) {
        rc = pv_create(&prms[k], device, &bt->bags[k]);
        if (rc) {
            pv_batch_destroy(bt);
            return rc;
        }
    }
    *out = bt;
    return PV_OK;
}

int pv_batch_destroy(pv_batch *bt)
{
    int k;
    if (!bt) return PV_OK;
    for (k = 0; k < bt->n; k++)
        if (bt->bags[k]) pv_destroy(bt->bags[k]);
    free(bt->bags);
    free(bt);
    return PV_OK;
}

int pv_batch_size(const pv_batch *bt) { return bt ? bt->n : 0; }

pv_bag *pv_batch_case(pv_batch *bt, int k)
{
    if (!bt || k < 0 || k >= bt->n) return NULL;
    return bt->bags[k];
}

static void *batch_worker(void *arg)
{
    batch_job *j = (batch_job *)arg;
    int s, k;
    if (j->bag) {
        j->rc = pv_Temporal_int_rk3(j->bag, j->nsteps);
        if (j->rc) snprintf(j->err, sizeof j->err, "%s", pv_last_error());
        return NULL;
    }
    for (s = 0; s < j->nsteps && !j->rc; s++)
        for (k = j->first; k < j->bt->n && !j->rc; k += j->stride)
            if ((j->rc = pv_Temporal_int_rk3(j->bt->bags[k], 1))) {
                j->bad = k;
                snprintf(j->err, sizeof j->err, "%s", pv_last_error());
            }
    return NULL;
}

int pv_batch_set_threads(pv_batch *bt, int nthreads)
{
    if (!bt || nthreads < 0) return pv_set_error(PV_EINVAL, "bad batch arguments");
    bt->nthreads = nthreads;
    return PV_OK;
}

int pv_batch_rk3(pv_batch *bt, int nsteps)
{
    batch_job *jobs;
    pthread_t *th;
    int k, rc = PV_OK, started = 0, nt;
    if (!bt || nsteps < 0) return pv_set_error(PV_EINVAL, "bad batch arguments");
    nt = (bt->nthreads > 0 && bt->nthreads < bt->n) ? bt->nthreads : bt->n;
    jobs = (batch_job *)calloc(nt, sizeof *jobs);
    th = (pthread_t *)calloc(nt, sizeof *th);
    if (!jobs || !th) { free(jobs); free(th); return pv_set_error(PV_ENOMEM, "host allocation failed"); }
    for (k = 0; k < nt; k++) {
        jobs[k].bag = (nt == bt->n) ? bt->bags[k] : NULL;
        jobs[k].bt = bt;
        jobs[k].first = k;
        jobs[k].stride = nt;
        jobs[k].bad = k;
        jobs[k].nsteps = nsteps;
        if (pthread_create(&th[k], NULL, batch_worker, &jobs[k]) != 0) {
            jobs[k].rc = PV_EINVAL;
            snprintf(jobs[k].err, sizeof jobs[k].err, "thread start failed");
            break;
        }
        started++;
    }
    for (k = 0; k < started; k++) pthread_join(th[k], NULL);
    for (k = 0; k < nt; k++)
        if (jobs[k].rc && rc == PV_OK) rc = pv_set_error(jobs[k].rc, "case %d: %s", jobs[k].bad, jobs[k].err);
    free(jobs);
    free(th);
    return rc;
}

/* pv_autotune_mg on every MG-PCG case, one case after another (each timed
 * alone); wlevels[k], sweeps[k] = its choice, -1 for a case without MG-PCG */
int pv_batch_autotune_mg(pv_batch *bt, int *wlevels, int *sweeps)
{
    int k, rc;
    if (!bt) return pv_set_error(PV_EINVAL, "null batch");
    for (k = 0; k < bt->n; k++) {
        int w = -1, nu = -1, ps = bt->bags[k]->prm.pressure_solver;
        if (ps == PV_PSOLVE_MG_PCG || ps == PV_PSOLVE_LEARNED_MG)
            if ((rc = pv_autotune_mg(bt->bags[k], &w, &nu))) return rc;
        if (wlevels) wlevels[k] = w;
        if (sweeps) sweeps[k] = nu;
    }
    return PV_OK;
}

/* pv_autotune on every FFT-PCG case; choices[k] = chosen graph chunk (0 = eager) */
int pv_batch_autotune(pv_batch *bt, int *choices)
{
    int k, rc;
    if (!bt) return pv_set_error(PV_EINVAL, "null batch");
    for (k = 0; k < bt->n; k++) {
        int c = 0;
        if (bt->bags[k]->prm.pressure_solver != PV_PSOLVE_FFT_PCG) {
            if (choices) choices[k] = -1;
            continue;
        }
        if ((rc = pv_autotune(bt->bags[k], &c))) return rc;
        if (choices) choices[k] = c;
    }
    return PV_OK;
}
