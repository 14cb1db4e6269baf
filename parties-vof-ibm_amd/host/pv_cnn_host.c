/*
 * pv_cnn_host.c -- host side of the learned pressure preconditioner network
 * (kernels in csrc/pv_cnn.hip): parameter initialisation, buffer sizing,
 * parameter I/O and batch evaluation / training steps.  No counterpart in
 * the reference; it exists for the learned FFT-FCG pressure solve.
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

static unsigned long long sm64(unsigned long long *s)
{
    unsigned long long z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static double gauss(unsigned long long *s)
{
    double u1 = ((double)(sm64(s) >> 11) + 0.5) / 9007199254740992.0;
    double u2 = ((double)(sm64(s) >> 11) + 0.5) / 9007199254740992.0;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

void pv_cnn_free(pv_bag *b)
{
    pvk_cnn *c = &b->cnn;
    int k;
    hipFree(c->w); hipFree(c->g); hipFree(c->m); hipFree(c->v);
    hipFree(c->wb); hipFree(c->wt);
    for (k = 0; k < 5; k++) hipFree(c->act[k]);
    hipFree(c->gz[0]); hipFree(c->gz[1]);
    hipFree(c->gout); hipFree(c->part); hipFree(c->lpart); hipFree(c->rpart);
    hipFree(b->cnn_io);
    memset(c, 0, sizeof *c);
    b->cnn_io = NULL;
    b->cnn_io_cap = 0;
    b->cnn_ready = 0;
}

/* activation / gradient tensors for npix = B * H * W pixels */
static int cnn_reserve(pv_bag *b, size_t npix)
{
    pvk_cnn *c = &b->cnn;
    size_t tiles = npix / 256 + 1, act = npix * 32 * sizeof(unsigned short);
    int k;
    if (npix <= c->cap_pix) return PV_OK;
    for (k = 0; k < 5; k++) { hipFree(c->act[k]); c->act[k] = NULL; }
    hipFree(c->gz[0]); hipFree(c->gz[1]); hipFree(c->gout); hipFree(c->part); hipFree(c->lpart);
    c->gout = c->part = c->lpart = NULL;
    c->gz[0] = c->gz[1] = NULL;
    c->cap_pix = 0;
    if (!c->rpart) HIPCHK(hipMalloc((void **)&c->rpart, (size_t)PVK_CNN_RSPLIT * 9248 * sizeof(float)));
    for (k = 0; k < 5; k++) HIPCHK(hipMalloc((void **)&c->act[k], act));
    HIPCHK(hipMalloc((void **)&c->gz[0], act));
    HIPCHK(hipMalloc((void **)&c->gz[1], act));
    HIPCHK(hipMalloc((void **)&c->gout, npix * sizeof(float)));
    HIPCHK(hipMalloc((void **)&c->part, tiles * 9248 * sizeof(float)));
    HIPCHK(hipMalloc((void **)&c->lpart, tiles * sizeof(float)));
    c->cap_pix = npix;
    return PV_OK;
}

static int io_reserve(pv_bag *b, size_t nfloat)
{
    if (nfloat <= b->cnn_io_cap) return PV_OK;
    hipFree(b->cnn_io);
    b->cnn_io = NULL;
    HIPCHK(hipMalloc((void **)&b->cnn_io, nfloat * sizeof(float)));
    b->cnn_io_cap = nfloat;
    return PV_OK;
}

/* He-normal weights (std sqrt(2 / fan_in)), zero biases; the output layer is
 * scaled by 0.1 so the untrained network adds a small correction. */
int pv_cnn_init(pv_bag *b, unsigned long long seed)
{
    pvk_cnn *c;
    float *h;
    unsigned long long s = seed;
    int k, l;
    if (!b) return pv_set_error(PV_EINVAL, "null handle");
    HIPCHK(hipSetDevice(b->device));
    if (!b->cnn_ready) {
        c = &b->cnn;
        memset(c, 0, sizeof *c);
        HIPCHK(hipMalloc((void **)&c->w, PVK_CNN_NPARAM * sizeof(float)));
        HIPCHK(hipMalloc((void **)&c->g, PVK_CNN_NPARAM * sizeof(float)));
        HIPCHK(hipMalloc((void **)&c->m, PVK_CNN_NPARAM * sizeof(float)));
        HIPCHK(hipMalloc((void **)&c->v, PVK_CNN_NPARAM * sizeof(float)));
        HIPCHK(hipMalloc((void **)&c->wb, 4 * 9216 * sizeof(unsigned short)));
        HIPCHK(hipMalloc((void **)&c->wt, 4 * 9216 * sizeof(unsigned short)));
        b->cnn_ready = 1;
    }
    c = &b->cnn;
    h = (float *)calloc(PVK_CNN_NPARAM, sizeof(float));
    if (!h) return pv_set_error(PV_ENOMEM, "host allocation failed");
    for (k = 0; k < 864; k++) h[k] = (float)(gauss(&s) * sqrt(2.0 / 27.0));
    for (l = 0; l < 4; l++)
        for (k = 0; k < 9216; k++) h[896 + l * 9248 + k] = (float)(gauss(&s) * sqrt(2.0 / 288.0));
    for (k = 0; k < 288; k++) h[37888 + k] = (float)(0.1 * gauss(&s) * sqrt(2.0 / 288.0));
    k = pv_cnn_set_params(b, h);
    free(h);
    return k;
}

int pv_cnn_set_params(pv_bag *b, const float *host)
{
    if (!b || !b->cnn_ready) return pv_set_error(PV_EINVAL, "network not initialised (pv_cnn_init)");
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipStreamSynchronize(b->stream));   /* work in flight may still read the weights */
    HIPCHK(hipMemcpy(b->cnn.w, host, PVK_CNN_NPARAM * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMemset(b->cnn.m, 0, PVK_CNN_NPARAM * sizeof(float)));
    HIPCHK(hipMemset(b->cnn.v, 0, PVK_CNN_NPARAM * sizeof(float)));
    pvk_cnn_pack(&b->cnn, b->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(b->stream));
    return PV_OK;
}

int pv_cnn_get_params(pv_bag *b, float *host)
{
    if (!b || !b->cnn_ready) return pv_set_error(PV_EINVAL, "network not initialised (pv_cnn_init)");
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipStreamSynchronize(b->stream));
    HIPCHK(hipMemcpy(host, b->cnn.w, PVK_CNN_NPARAM * sizeof(float), hipMemcpyDeviceToHost));
    return PV_OK;
}

static int check_shape(int B, int H, int W)
{
    if (B < 1 || H < 16 || W < 16 || H % 16 || W % 16)
        return pv_set_error(PV_EINVAL, "batch shape %d x %d x %d: B >= 1, H and W multiples of 16", B, H, W);
    return PV_OK;
}

/* Forward of a host batch; with tgt: loss and the full parameter gradient
 * (grad may be NULL to keep it on the device for pv_cnn_adam_step). */
int pv_cnn_eval(pv_bag *b, const float *in3, const float *tgt, float *out, float *grad, double *loss, int B, int H,
                int W, int periodic_x, int border)
{
    size_t npix = (size_t)B * H * W;
    float *din, *dout, *dtgt;
    int rc;
    if (!b || !b->cnn_ready) return pv_set_error(PV_EINVAL, "network not initialised (pv_cnn_init)");
    if ((rc = check_shape(B, H, W))) return rc;
    HIPCHK(hipSetDevice(b->device));
    if ((rc = cnn_reserve(b, npix))) return rc;
    if ((rc = io_reserve(b, npix * 5))) return rc;
    din = b->cnn_io; dout = din + 3 * npix; dtgt = dout + npix;
    HIPCHK(hipMemcpyAsync(din, in3, 3 * npix * sizeof(float), hipMemcpyHostToDevice, b->stream));
    pvk_cnn_forward(&b->cnn, din, dout, B, H, W, periodic_x, b->stream);
    HIPCHK(hipGetLastError());
    if (tgt) {
        HIPCHK(hipMemcpyAsync(dtgt, tgt, npix * sizeof(float), hipMemcpyHostToDevice, b->stream));
        pvk_cnn_backward(&b->cnn, din, dout, dtgt, B, H, W, periodic_x, border, b->stream);
        HIPCHK(hipGetLastError());
    }
    if (out) HIPCHK(hipMemcpyAsync(out, dout, npix * sizeof(float), hipMemcpyDeviceToHost, b->stream));
    if (tgt && grad)
        HIPCHK(hipMemcpyAsync(grad, b->cnn.g, PVK_CNN_NPARAM * sizeof(float), hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    if (tgt && loss) {
        int nblk = (int)(((size_t)H * W + 255) / 256) * B, k;
        float *lp = (float *)malloc(nblk * sizeof(float));
        double sum = 0.0;
        if (!lp) return pv_set_error(PV_ENOMEM, "host allocation failed");
        HIPCHK(hipMemcpy(lp, b->cnn.lpart, nblk * sizeof(float), hipMemcpyDeviceToHost));
        for (k = 0; k < nblk; k++) sum += lp[k];
        free(lp);
        *loss = sum;
    }
    return PV_OK;
}

int pv_cnn_adam_step(pv_bag *b, float lr, int t)
{
    if (!b || !b->cnn_ready) return pv_set_error(PV_EINVAL, "network not initialised (pv_cnn_init)");
    if (t < 1) return pv_set_error(PV_EINVAL, "Adam step count starts at 1");
    HIPCHK(hipSetDevice(b->device));
    pvk_cnn_adam(&b->cnn, lr, t, b->stream);
    HIPCHK(hipGetLastError());
    return PV_OK;
}
