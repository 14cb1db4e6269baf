/*
 * pv_ic.c -- initial volume-fraction fields on the canonical layout.
 *
 * The reference initialisers (IO/VoF_Init.c:877-953 Rayleigh-Taylor,
 * :1032-1100 planar rising bubble) average 0.5 (1 + tanh(s / (2 sqrt 2 Cn)))
 * over 6 x 6 sub-samples of each cell of the G range; s is the signed
 * distance to the interface, positive in the liquid (F = 1).  The seeded
 * generators below reuse that profile for fields the reference has no
 * initialiser for (multi-mode RT, random bubbles and droplets).
 */
#include "pv_host.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define PI_REF 3.14159265358979323846

typedef double (*sdist_fn)(double x, double y, const void *ctx);

static int fill_profile(const pv_params *prm_in, double *F, sdist_fn sd, const void *ctx)
{
    pv_params p = *prm_in;
    pv_metrics m;
    double slab, denom, dx, dy;
    const int samples = 6;
    const double inv_samples = 1.0 / (double)samples;
    const double sample_weight = inv_samples * inv_samples;
    int NX, NY, i, j, P;
    if (!F || p.NXM < 1 || p.NYM < 1) return pv_set_error(PV_EINVAL, "bad arguments");
    pv_derive_params(&p, &slab);
    pv_host_metrics(&p, &m);
    NX = p.NXM + 1; NY = p.NYM + 1; P = p.NXM + 7;
    denom = 2.0 * sqrt(2.0) * p.Cn;
    dx = m.dx_c[3];
    dy = m.dy_c[3];
    memset(F, 0, sizeof(double) * (size_t)P * (size_t)(p.NYM + 7));
    for (j = 0; j < NY; ++j) {
        const double y_min = m.yc[j + 3] - 0.5 * dy;
        for (i = 0; i < NX; ++i) {
            const double x_min = m.xc[i + 3] - 0.5 * dx;
            double f_cell = 0.0;
            int sj, si;
            for (sj = 0; sj < samples; ++sj) {
                const double y = y_min + (sj + 0.5) * dy * inv_samples;
                for (si = 0; si < samples; ++si) {
                    const double x = x_min + (si + 0.5) * dx * inv_samples;
                    const double s = sd(x, y, ctx);
                    f_cell += sample_weight * 0.5 * (1.0 + tanh(s / (denom + 1e-30)));
                }
            }
            F[(size_t)(j + 3) * P + (i + 3)] = f_cell;
        }
    }
    pv_host_metrics_free(&m);
    return PV_OK;
}

/* ---- planar rising bubble (VoF_Init.c:1032-1100) ---- */
typedef struct { double R, xc, yc; } bubble1_ctx;
static double sd_bubble1(double x, double y, const void *c)
{
    const bubble1_ctx *b = (const bubble1_ctx *)c;
    double r = sqrt((x - b->xc) * (x - b->xc) + (y - b->yc) * (y - b->yc));
    return r - b->R;
}
int pv_ic_rising_bubble(const pv_params *prm, double *F)
{
    bubble1_ctx c;
    double Lx = prm->xmax - prm->xmin;
    c.R = 0.25 * Lx;
    c.xc = 0.5 * (prm->xmin + prm->xmax);
    c.yc = prm->ymin + 2.0 * c.R;
    return fill_profile(prm, F, sd_bubble1, &c);
}

/* ---- single-mode Rayleigh-Taylor (VoF_Init.c:877-953) ---- */
typedef struct { double xmin, Lx, y_mid, amp; } rt_ctx;
static double sd_rt(double x, double y, const void *c)
{
    const rt_ctx *r = (const rt_ctx *)c;
    double eta = r->y_mid + r->amp * cos(2.0 * PI_REF * (x - r->xmin) / r->Lx);
    return y - eta;
}
int pv_ic_rayleigh_taylor(const pv_params *prm, double *F)
{
    rt_ctx c;
    c.xmin = prm->xmin;
    c.Lx = prm->xmax - prm->xmin;
    c.y_mid = 0.5 * (prm->ymin + prm->ymax);
    c.amp = 0.1 * c.Lx;
    return fill_profile(prm, F, sd_rt, &c);
}

/* ---- splitmix64 stream ---- */
static unsigned long long sm64(unsigned long long *s)
{
    unsigned long long z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static double u01(unsigned long long *s) { return (double)(sm64(s) >> 11) * (1.0 / 9007199254740992.0); }

/* ---- multi-mode Rayleigh-Taylor: interface y0 + sum_k a_k cos(2 pi k x / Lx + ph_k),
 *      a_k ~ U[0, amp_max], ph_k ~ U[0, 2 pi), k = 1..modes; heavy fluid on top ---- */
#define MAXMODES 64
typedef struct { double xmin, Lx, y0; int n; double a[MAXMODES], ph[MAXMODES]; } mrt_ctx;
static double sd_mrt(double x, double y, const void *c)
{
    const mrt_ctx *r = (const mrt_ctx *)c;
    double eta = r->y0;
    int k;
    for (k = 0; k < r->n; k++) eta += r->a[k] * cos(2.0 * PI_REF * (k + 1) * (x - r->xmin) / r->Lx + r->ph[k]);
    return y - eta;
}
int pv_ic_multimode_rt(const pv_params *prm, unsigned long long seed, int modes, double amp_max, double y0,
                       double *F)
{
    mrt_ctx c;
    unsigned long long s = seed;
    int k;
    if (modes < 1 || modes > MAXMODES) return pv_set_error(PV_EINVAL, "modes must be 1..%d", MAXMODES);
    c.xmin = prm->xmin; c.Lx = prm->xmax - prm->xmin; c.y0 = y0; c.n = modes;
    for (k = 0; k < modes; k++) { c.a[k] = amp_max * u01(&s); c.ph[k] = 2.0 * PI_REF * u01(&s); }
    return fill_profile(prm, F, sd_mrt, &c);
}

/* ---- seeded bubbles / droplets, x-periodic distance, gap >= 4 Cn between
 *      surfaces and >= 4 Cn wall clearance; draw order: count, then per
 *      candidate (radius, x, y) ---- */
#define MAXB 4096
typedef struct { int n, gas_inside; double Lx; double x[MAXB], y[MAXB], r[MAXB]; } bub_ctx;
static double sd_bub(double x, double y, const void *c)
{
    const bub_ctx *b = (const bub_ctx *)c;
    double best = 1e300;
    int k;
    for (k = 0; k < b->n; k++) {
        double dx = fabs(x - b->x[k]);
        double d;
        if (dx > 0.5 * b->Lx) dx = b->Lx - dx;
        d = sqrt(dx * dx + (y - b->y[k]) * (y - b->y[k])) - b->r[k];
        if (d < best) best = d;
    }
    return b->gas_inside ? best : -best;
}
int pv_ic_bubbles(const pv_params *prm, unsigned long long seed, int nmin, int nmax, double rmin, double rmax,
                  int gas_inside, double *F, int *nplaced)
{
    bub_ctx *c;                       /* ~96 KiB: on the heap, so calls may run concurrently */
    pv_params p = *prm;
    unsigned long long s = seed;
    double slab, gap;
    int target, tries = 0, k, rc;
    if (nmin < 1 || nmax < nmin || nmax > MAXB) return pv_set_error(PV_EINVAL, "bubble count out of range");
    c = (bub_ctx *)calloc(1, sizeof *c);
    if (!c) return pv_set_error(PV_ENOMEM, "pv_ic_bubbles: out of host memory");
    pv_derive_params(&p, &slab);
    gap = 4.0 * p.Cn;
    c->Lx = p.xmax - p.xmin;
    c->gas_inside = gas_inside;
    target = nmin + (int)(u01(&s) * (double)(nmax - nmin + 1));
    if (target > nmax) target = nmax;
    while (c->n < target && tries < 200000) {
        double r = rmin + (rmax - rmin) * u01(&s);
        double x = p.xmin + c->Lx * u01(&s);
        double y = p.ymin + (r + gap) + ((p.ymax - p.ymin) - 2.0 * (r + gap)) * u01(&s);
        int ok = 1;
        tries++;
        for (k = 0; k < c->n && ok; k++) {
            double dx = fabs(x - c->x[k]);
            if (dx > 0.5 * c->Lx) dx = c->Lx - dx;
            if (sqrt(dx * dx + (y - c->y[k]) * (y - c->y[k])) < r + c->r[k] + gap) ok = 0;
        }
        if (!ok) continue;
        c->x[c->n] = x; c->y[c->n] = y; c->r[c->n] = r; c->n++;
    }
    if (nplaced) *nplaced = c->n;
    rc = fill_profile(prm, F, sd_bub, c);
    free(c);
    return rc;
}
