/*
 * pv_learn_host.c -- training data, training loop and warm start of the learned
 * pressure solve (no counterpart in the reference).
 *
 * Data: while collecting, every plain FFT-PCG pressure solve stores at the
 * chosen iterations k the iterate phi_k, the residual r_k, the
 * FFT-preconditioned z_k and the network features of (r_k, z_k, rho).
 * After convergence (phi*), e_k = phi* - phi_k solves A e_k = r_k; the target
 * is the source correction (L e_k - r_k) / s_r (L: the constant-coefficient
 * Laplacian, M_fft = L^-1), cut with the features into random crops of side
 * cs (kept in HBM).
 * Training: Adam on random batches of crops, loss on the crop interior
 * (6 pixels from the edges: the network's receptive-field radius).
 */
#include "pv_host.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define HIPCHK(call)                                                                 \
    do {                                                                             \
        hipError_t e_ = (call);                                                      \
        if (e_ != hipSuccess)                                                        \
            return pv_set_error(PV_EHIP, "%s failed: %s (%s:%d)", #call,             \
                                hipGetErrorString(e_), __FILE__, __LINE__);          \
    } while (0)
#define RC(x) do { int rc_ = (x); if (rc_) return rc_; } while (0)

#define CNN_BORDER 6

static unsigned long long sm64(unsigned long long *s)
{
    unsigned long long z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

void pv_learn_free(pv_bag *b)
{
    pv_learn *L = b->learn;
    if (!L) return;
    hipFree(L->snapmem); hipFree(L->feat); hipFree(L->tgt); hipFree(L->cout); hipFree(L->part);
    hipFree(L->ds_in); hipFree(L->ds_tgt); hipFree(L->d_org); hipFree(L->d_idx);
    free(L);
    b->learn = NULL;
}

static int learn_alloc(pv_bag *b)
{
    pv_learn *L;
    size_t npix = (size_t)b->NXM * b->NYM;
    if (b->learn) return PV_OK;
    L = (pv_learn *)calloc(1, sizeof *L);
    if (!L) return pv_set_error(PV_ENOMEM, "host allocation failed");
    b->learn = L;
    HIPCHK(hipMalloc((void **)&L->tgt, npix * sizeof(float)));
    HIPCHK(hipMalloc((void **)&L->cout, npix * sizeof(float)));
    HIPCHK(hipMalloc((void **)&L->part, 3 * (size_t)b->NYM * sizeof(double)));
    HIPCHK(hipMalloc((void **)&L->feat, npix * 3 * sizeof(float)));   /* warm-start features; grown by collect */
    return PV_OK;
}

int pv_set_pressure_solver(pv_bag *b, int solver)
{
    if (!b) return pv_set_error(PV_EINVAL, "null handle");
    if (solver < PV_PSOLVE_JACOBI_PCG || solver > PV_PSOLVE_LEARNED_MG)
        return pv_set_error(PV_EINVAL, "unknown pressure_solver %d", solver);
    b->prm.pressure_solver = solver;
    return PV_OK;
}

int pv_learn_collect_begin(pv_bag *b, const int *snap_iters, int nsnap, int crops_per_snap, int crop_size,
                           int max_samples, unsigned long long seed)
{
    pv_learn *L;
    size_t npix = (size_t)b->NXM * b->NYM, fe;
    int k;
    if (!b || !snap_iters || nsnap < 1 || nsnap > PV_LEARN_MAXSNAP || crops_per_snap < 1 || max_samples < 1)
        return pv_set_error(PV_EINVAL, "bad collector arguments");
    if (crop_size < 32 || crop_size % 16 || crop_size > b->NXM || crop_size > b->NYM)
        return pv_set_error(PV_EINVAL, "crop size %d: multiple of 16, >= 32, within the grid", crop_size);
    for (k = 0; k < nsnap; k++)
        if (snap_iters[k] < 0 || (k && snap_iters[k] <= snap_iters[k - 1]))
            return pv_set_error(PV_EINVAL, "snapshot iterations must be increasing and >= 0");
    HIPCHK(hipSetDevice(b->device));
    RC(learn_alloc(b));
    L = b->learn;
    hipFree(L->snapmem); hipFree(L->feat); hipFree(L->ds_in); hipFree(L->ds_tgt); hipFree(L->d_org);
    L->snapmem = NULL; L->feat = NULL; L->ds_in = NULL; L->ds_tgt = NULL; L->d_org = NULL;
    fe = pv_field_elems(b);
    HIPCHK(hipMalloc((void **)&L->snapmem, 3 * (size_t)nsnap * fe * sizeof(double)));
    for (k = 0; k < nsnap; k++) {
        L->sphi[k] = pv_field_origin(b, L->snapmem + (3 * (size_t)k) * fe);
        L->sz[k] = pv_field_origin(b, L->snapmem + (3 * (size_t)k + 1) * fe);
        L->sr[k] = pv_field_origin(b, L->snapmem + (3 * (size_t)k + 2) * fe);
        L->snap_iter[k] = snap_iters[k];
    }
    HIPCHK(hipMalloc((void **)&L->feat, (size_t)nsnap * npix * 3 * sizeof(float)));
    HIPCHK(hipMalloc((void **)&L->ds_in, (size_t)max_samples * crop_size * crop_size * 3 * sizeof(float)));
    HIPCHK(hipMalloc((void **)&L->ds_tgt, (size_t)max_samples * crop_size * crop_size * sizeof(float)));
    HIPCHK(hipMalloc((void **)&L->d_org, 2 * (size_t)crops_per_snap * sizeof(int)));
    L->nsnap = nsnap;
    L->cs = crop_size;
    L->crops_per_snap = crops_per_snap;
    L->ds_cap = max_samples;
    L->ds_n = 0;
    L->rng = seed;
    L->collecting = 1;
    return PV_OK;
}

int pv_learn_collect_end(pv_bag *b, int *nsamples)
{
    if (!b || !b->learn) return pv_set_error(PV_EINVAL, "no collector (pv_learn_collect_begin)");
    HIPCHK(hipStreamSynchronize(b->stream));
    b->learn->collecting = 0;
    if (nsamples) *nsamples = b->learn->ds_n;
    return PV_OK;
}

/* snapshot si of the running FFT-PCG: phi_k, z_k, r_k, s_r and the features
 * of (r_k, z_k) */
int pv_learn_snapshot(pv_bag *b, int si)
{
    pv_learn *L = b->learn;
    size_t npix = (size_t)b->NXM * b->NYM;
    pvk_learn_copy(&b->g, b->f[PV_F_PHI], L->sphi[si], b->stream);
    pvk_learn_copy(&b->g, b->pcg.z, L->sz[si], b->stream);
    pvk_learn_copy(&b->g, b->pcg.r, L->sr[si], b->stream);
    pvk_learn_stats(&b->g, b->d_sc, b->pcg.r, b->pcg.z, NULL, NULL, L->part, b->stream);
    pvk_learn_feat(&b->g, b->d_sc, b->pcg.r, b->pcg.z, b->f[PV_F_RHO], L->feat + (size_t)si * npix * 3, b->stream);
    HIPCHK(hipGetLastError());
    if (pv_read_scalars(b)) return PV_EHIP;
    L->sig_r[si] = b->h_sc->red[5];
    return PV_OK;
}

/* after convergence: targets of the snapshots taken and their crops */
int pv_learn_harvest(pv_bag *b, int nsnap)
{
    pv_learn *L = b->learn;
    size_t npix = (size_t)b->NXM * b->NYM;
    int si, k, *org;
    org = (int *)malloc(2 * (size_t)L->crops_per_snap * sizeof(int));
    if (!org) return pv_set_error(PV_ENOMEM, "host allocation failed");
    for (si = 0; si < nsnap && L->ds_n < L->ds_cap; si++) {
        int n = L->crops_per_snap;
        if (n > L->ds_cap - L->ds_n) n = L->ds_cap - L->ds_n;
        pvk_learn_target(&b->g, b->d_sc, b->f[PV_F_PHI], L->sphi[si], L->sr[si], L->sig_r[si], L->tgt, b->stream);
        for (k = 0; k < n; k++) {
            org[2 * k] = (int)(sm64(&L->rng) % (unsigned long long)(b->NXM - L->cs + 1));
            org[2 * k + 1] = (int)(sm64(&L->rng) % (unsigned long long)(b->NYM - L->cs + 1));
        }
        if (hipMemcpyAsync(L->d_org, org, 2 * (size_t)n * sizeof(int), hipMemcpyHostToDevice, b->stream) !=
                hipSuccess ||
            hipStreamSynchronize(b->stream) != hipSuccess) {
            free(org);
            return pv_set_error(PV_EHIP, "crop origin upload failed");
        }
        pvk_learn_crop(L->feat + (size_t)si * npix * 3, L->tgt, b->NXM, L->cs, L->d_org, n, L->ds_n, L->ds_in,
                       L->ds_tgt, b->stream);
        L->ds_n += n;
    }
    free(org);
    HIPCHK(hipGetLastError());
    return PV_OK;
}

int pv_learn_train(pv_bag *b, int steps, int batch, float lr, unsigned long long seed, double *loss_first,
                   double *loss_last)
{
    pv_learn *L = b ? b->learn : NULL;
    size_t cs2, npix;
    int t, k, *idx, nblk;
    float *bin, *bout, *btgt, *lp;
    unsigned long long s = seed;
    if (!L || L->ds_n < 1) return pv_set_error(PV_EINVAL, "no training samples (pv_learn_collect_*)");
    if (!b->cnn_ready) return pv_set_error(PV_EINVAL, "network not initialised (pv_cnn_init)");
    if (steps < 1 || batch < 1) return pv_set_error(PV_EINVAL, "steps and batch must be >= 1");
    HIPCHK(hipSetDevice(b->device));
    cs2 = (size_t)L->cs * L->cs;
    npix = (size_t)batch * cs2;
    /* every step's sample indices are drawn up front and uploaded once, in
     * stream order (the stream does not synchronise with the null stream) */
    HIPCHK(hipStreamSynchronize(b->stream));
    hipFree(L->d_idx);
    L->d_idx = NULL;
    HIPCHK(hipMalloc((void **)&L->d_idx, (size_t)steps * batch * sizeof(int)));
    if (b->cnn_io_cap < npix * 5) {
        hipFree(b->cnn_io);
        b->cnn_io = NULL;
        b->cnn_io_cap = 0;
        HIPCHK(hipMalloc((void **)&b->cnn_io, npix * 5 * sizeof(float)));
        b->cnn_io_cap = npix * 5;
    }
    if (b->cnn.cap_pix < npix) {
        /* reserve through a throw-away evaluation of the right shape */
        float *tmp = (float *)calloc(npix * 3, sizeof(float));
        if (!tmp) return pv_set_error(PV_ENOMEM, "host allocation failed");
        k = pv_cnn_eval(b, tmp, NULL, NULL, NULL, NULL, batch, L->cs, L->cs, 0, 0);
        free(tmp);
        RC(k);
    }
    bin = b->cnn_io; bout = bin + 3 * npix; btgt = bout + npix;
    idx = (int *)malloc((size_t)steps * batch * sizeof(int));
    nblk = (int)(cs2 / 256) * batch;
    lp = (float *)malloc(nblk * sizeof(float));
    if (!idx || !lp) { free(idx); free(lp); return pv_set_error(PV_ENOMEM, "host allocation failed"); }
    for (k = 0; k < steps * batch; k++) idx[k] = (int)(sm64(&s) % (unsigned long long)L->ds_n);
    if (hipMemcpy(L->d_idx, idx, (size_t)steps * batch * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
        free(idx); free(lp);
        return pv_set_error(PV_EHIP, "batch index upload failed");
    }
    for (t = 1; t <= steps; t++) {
        pvk_learn_gather(L->ds_in, L->ds_tgt, L->cs, L->d_idx + (size_t)(t - 1) * batch, batch, bin, btgt,
                         b->stream);
        pvk_cnn_forward(&b->cnn, bin, bout, batch, L->cs, L->cs, 0, b->stream);
        pvk_cnn_backward(&b->cnn, bin, bout, btgt, batch, L->cs, L->cs, 0, CNN_BORDER, b->stream);
        if (t == 1 || t == steps) {
            double sum = 0.0;
            if (hipMemcpyAsync(lp, b->cnn.lpart, nblk * sizeof(float), hipMemcpyDeviceToHost, b->stream) !=
                    hipSuccess ||
                hipStreamSynchronize(b->stream) != hipSuccess) {
                free(idx); free(lp);
                return pv_set_error(PV_EHIP, "loss readback failed");
            }
            for (k = 0; k < nblk; k++) sum += lp[k];
            if (t == 1 && loss_first) *loss_first = sum;
            if (t == steps && loss_last) *loss_last = sum;
        }
        pvk_cnn_adam(&b->cnn, lr, t, b->stream);
    }
    free(idx);
    free(lp);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(b->stream));
    return PV_OK;
}

/* out = r + s_r c(r, z, rho) for the current (r, z = M_fft r) of the solve:
 * the right-hand side whose M_fft solve is the learned approximation of
 * A^-1 r */
int pv_learn_source(pv_bag *b, double *out)
{
    pv_learn *L;
    size_t npix = (size_t)b->NXM * b->NYM;
    if (b->NXM % 16 || b->NYM % 16) return pv_set_error(PV_EINVAL, "learned solve needs NXM, NYM multiples of 16");
    RC(learn_alloc(b));
    L = b->learn;
    if (b->cnn.cap_pix < npix) {
        float *big = (float *)calloc(npix * 3, sizeof(float));
        int rc;
        if (!big) return pv_set_error(PV_ENOMEM, "host allocation failed");
        /* grows the activation tensors to the grid (a forward on zeros) */
        rc = pv_cnn_eval(b, big, NULL, NULL, NULL, NULL, 1, b->NYM, b->NXM, 1, 0);
        free(big);
        RC(rc);
    }
    pvk_learn_stats(&b->g, b->d_sc, b->pcg.r, b->pcg.z, NULL, NULL, L->part, b->stream);
    pvk_learn_feat(&b->g, b->d_sc, b->pcg.r, b->pcg.z, b->f[PV_F_RHO], L->feat, b->stream);
    pvk_cnn_forward(&b->cnn, L->feat, L->cout, 1, b->NYM, b->NXM, 1, b->stream);
    pvk_learn_addr(&b->g, b->d_sc, b->pcg.r, L->cout, out, b->stream);
    HIPCHK(hipGetLastError());
    return PV_OK;
}

int pv_learn_start_residual(pv_bag *b, double *res)
{
    if (!b || !res) return pv_set_error(PV_EINVAL, "null argument");
    *res = b->learn_start_res;
    return PV_OK;
}
