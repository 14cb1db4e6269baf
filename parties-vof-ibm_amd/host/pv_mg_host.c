/*
 * pv_mg_host.c -- level hierarchy and work set of the multigrid-preconditioned
 * pressure CG (no counterpart in the reference; its pressure CG is
 * preconditioned by the FFT solve of psolve_fft.c or by Jacobi).
 *
 * Level 0 is the pressure grid itself (z, r and the face coefficients of the
 * CG, device pitch P).  Levels halve both sizes while the level has more than
 * PVK_MG_LDS_CELLS cells (the first level at or below that size, c0, and all
 * levels after it run in one workgroup from LDS), then on while both sizes
 * are even and the level has more than PVK_MG_BOTTOM_MIN cells or more than
 * PVK_MG_BOTTOM_NX columns; the last level must have at most PVK_MG_BOTTOM
 * cells and PVK_MG_BOTTOM_NX columns (the band of its factorisation).  A
 * grid that does not halve that way has no MG-PCG (mg_ok = 0).
 *
 * Two hierarchies of the same sizes are kept: mg (FP64 V-cycle) and mgf
 * (FP32 V-cycle inside the FP64 CG; its level 0 keeps FP32 copies of the
 * pre-smoothed x and the faces and reads r / writes z in FP64).  Handles of
 * mixed precision use mgf unless pv_set_mg_precision says otherwise.  Every
 * level has a pre-smoothed x (xp) apart from its result x.  Level 0's pitch is
 * the field pitch P, the others' their width.
 */
#include "pv_host.h"
#include <stdlib.h>
#include <string.h>

int pv_mg_create(pv_bag *b)
{
    pvk_mg *m = &b->mg, *mf = &b->mgf;
    int lx[PVK_MG_MAXLEV], ly[PVK_MG_MAXLEV], nl = 1, c0, l, nrz, nx = b->NXM, ny = b->NYM;
    size_t cells = 0, total, totalf, nc0 = (size_t)nx * ny;
    double *q;
    float *qf;
    memset(m, 0, sizeof *m);
    memset(mf, 0, sizeof *mf);
    memset(&b->mgw, 0, sizeof b->mgw);
    b->mgmem = NULL;
    b->mgfmem = NULL;
    b->mg_ok = 0;
    b->mg_chunk = 4;
    b->mg_bottom_age[0] = b->mg_bottom_age[1] = -1;
    b->mgp = m;
    lx[0] = nx;
    ly[0] = ny;
    while (lx[nl - 1] * ly[nl - 1] > PVK_MG_LDS_CELLS) {
        if ((lx[nl - 1] & 1) || (ly[nl - 1] & 1) || nl == PVK_MG_MAXLEV) return PV_OK;
        lx[nl] = lx[nl - 1] / 2;
        ly[nl] = ly[nl - 1] / 2;
        nl++;
    }
    c0 = nl - 1;
    while ((lx[nl - 1] * ly[nl - 1] > PVK_MG_BOTTOM_MIN || lx[nl - 1] > PVK_MG_BOTTOM_NX) && !(lx[nl - 1] & 1) &&
           !(ly[nl - 1] & 1) && nl < PVK_MG_MAXLEV) {
        lx[nl] = lx[nl - 1] / 2;
        ly[nl] = ly[nl - 1] / 2;
        nl++;
    }
    if (lx[nl - 1] * ly[nl - 1] > PVK_MG_BOTTOM || lx[nl - 1] > PVK_MG_BOTTOM_NX || nl - 1 - c0 > PVK_MG_LOW)
        return PV_OK;
    /* the smoothing levels (all but the last) pair their cells along x */
    for (l = 0; l + 1 < nl; l++)
        if (lx[l] & 1) return PV_OK;
    m->nlev = mf->nlev = nl;
    m->c0 = mf->c0 = c0;
    m->nu = mf->nu = 2;
    mf->fp32 = 1;
    for (l = 0; l < nl; l++) {
        m->L[l].nx = mf->L[l].nx = lx[l];
        m->L[l].ny = mf->L[l].ny = ly[l];
        m->L[l].pitch = mf->L[l].pitch = lx[l];
    }
    for (l = 1; l < nl; l++) cells += (size_t)lx[l] * ly[l];
    /* per-tile partials sized for the smallest level-0 tiles (the most) */
    m->sh0 = mf->sh0 = 2;
    nrz = pvk_mg_nrz(m);
    m->sh0 = mf->sh0 = -1;
    total = 5 * cells + 2 * (size_t)b->g.P * ny + 3 * (size_t)PVK_MG_BOTTOM * PVK_MG_BOTTOM +
            2 * (size_t)PVK_MG_BOTTOM * (PVK_MG_BOTTOM_NX + 3) + 5 * (size_t)ny + 3 * (size_t)nrz;
    totalf = 5 * cells + 3 * nc0 + (size_t)PVK_MG_BOTTOM * PVK_MG_BOTTOM;
    if (hipMalloc((void **)&b->mgmem, total * sizeof(double)) != hipSuccess) {
        b->mgmem = NULL;
        return pv_set_error(PV_ENOMEM, "MG-PCG work allocation failed");
    }
    if (hipMalloc((void **)&b->mgfmem, totalf * sizeof(float)) != hipSuccess) {
        b->mgfmem = NULL;
        return pv_set_error(PV_ENOMEM, "MG-PCG work allocation failed");
    }
    /* FP64 V-cycle: level 0 is the CG's z, r and faces */
    q = b->mgmem;
    m->L[0].pitch = b->g.P;
    m->L[0].x = b->f[PV_FI_Z];
    m->L[0].f = b->f[PV_F_P_RES];
    m->L[0].bx = b->f[PV_FI_PBX];
    m->L[0].by = b->f[PV_FI_PBY];
    m->L[0].xp = q; q += (size_t)b->g.P * ny;
    for (l = 1; l < nl; l++) {
        size_t n = (size_t)lx[l] * ly[l];
        m->L[l].x = q; q += n;
        m->L[l].f = q; q += n;
        m->L[l].bx = q; q += n;
        m->L[l].by = q; q += n;
        m->L[l].xp = q; q += n;
    }
    m->G = q; q += (size_t)PVK_MG_BOTTOM * PVK_MG_BOTTOM;
    m->cm = q; q += PVK_MG_BOTTOM;
    m->Lb = q; q += (size_t)PVK_MG_BOTTOM * (PVK_MG_BOTTOM_NX + 2);    /* the band, then 1 / diag */
    m->H = q; q += (size_t)PVK_MG_BOTTOM * PVK_MG_BOTTOM;
    m->p0 = b->g.P;
    /* FP32 V-cycle: level 0 has its own x and face copies; r in, z out */
    qf = b->mgfmem;
    mf->L[0].x = NULL;                 /* the result goes to z0 */
    mf->L[0].xp = qf; qf += nc0;
    mf->L[0].f = NULL;
    mf->L[0].bx = qf; qf += nc0;
    mf->L[0].by = qf; qf += nc0;
    for (l = 1; l < nl; l++) {
        size_t n = (size_t)lx[l] * ly[l];
        mf->L[l].x = qf; qf += n;
        mf->L[l].f = qf; qf += n;
        mf->L[l].bx = qf; qf += n;
        mf->L[l].by = qf; qf += n;
        mf->L[l].xp = qf; qf += n;
    }
    mf->G = q; q += (size_t)PVK_MG_BOTTOM * PVK_MG_BOTTOM;
    mf->cm = q; q += PVK_MG_BOTTOM;
    mf->Lb = q; q += (size_t)PVK_MG_BOTTOM * (PVK_MG_BOTTOM_NX + 2);
    mf->H = qf; qf += (size_t)PVK_MG_BOTTOM * PVK_MG_BOTTOM;
    mf->r0 = b->f[PV_F_P_RES];
    mf->z0 = b->f[PV_FI_Z];
    mf->p0 = b->g.P;
    {
        pvk_pcg *w = &b->mgw;
        w->phi = b->f[PV_F_PHI];
        w->r = b->f[PV_F_P_RES];
        w->z = b->f[PV_FI_Z];
        w->d[0] = b->f[PV_F_P_D];
        w->d[1] = b->f[PV_FI_D2];
        w->Ad = b->f[PV_F_P_AD];
        w->bx = b->f[PV_FI_PBX];
        w->by = b->f[PV_FI_PBY];
        w->rhs = b->f[PV_F_P_RHS];
        w->rho = b->f[PV_F_RHO];
        w->dad[0] = q; q += ny;
        w->dad[1] = q; q += ny;
        w->rr = q; q += ny;
        w->mean = q; q += ny;
        w->sad = q; q += ny;
        w->rz[0] = q; q += nrz;
        w->rz[1] = q; q += nrz;
        w->r2 = q; q += (size_t)b->g.P * ny;     /* r_1, r_3, ...: pitch P like r */
        w->rrt = q; q += nrz;
        nrz = pvk_mg_nrz(m);                     /* the tiles in use */
        w->nrrt = (c0 > 0) ? nrz : 0;
        w->ndad = ny;
        w->nrr = ny;
        w->nrz = nrz;
        w->nmean = ny;
        w->band = 1;
        w->rrow = 1;
        w->rdct = 1;
        w->flex = 0;
    }
    /* the faces of level c0 go to LDS too when the whole set fits in 144 KiB */
    m->topf = mf->topf = 1;
    if (pvk_mg_lds_bytes(m) > 144 * 1024) m->topf = 0;
    if (pvk_mg_lds_bytes(mf) > 144 * 1024) mf->topf = 0;
    b->mgp = (b->prm.precision == PV_PREC_MIXED) ? mf : m;
    b->mg_ok = 1;
    return PV_OK;
}

void pv_mg_destroy(pv_bag *b)
{
    if (b->mgmem) hipFree(b->mgmem);
    if (b->mgfmem) hipFree(b->mgfmem);
    b->mgmem = NULL;
    b->mgfmem = NULL;
    b->mg_ok = 0;
}

int pv_set_mg_precision(pv_bag *b, int fp32)
{
    if (!b) return pv_set_error(PV_EINVAL, "null handle");
    if (fp32 < 0 || fp32 > 1) return pv_set_error(PV_EINVAL, "MG precision must be 0 (FP64) or 1 (FP32)");
    b->mgp = fp32 ? &b->mgf : &b->mg;
    return PV_OK;
}

int pv_set_pressure_warm_start(pv_bag *b, int mode)
{
    if (!b) return pv_set_error(PV_EINVAL, "null handle");
    if (mode < 0 || mode > 1) return pv_set_error(PV_EINVAL, "warm start mode must be 0 or 1 (got %d)", mode);
    b->p_warm = mode;
    b->have_prev[0] = b->have_prev[1] = b->have_prev[2] = 0;
    memset(b->warm_wins, 0, sizeof b->warm_wins);
    memset(b->warm_solves, 0, sizeof b->warm_solves);
    return PV_OK;
}

int pv_set_mg_tile0(pv_bag *b, int shape)
{
    if (!b) return pv_set_error(PV_EINVAL, "null handle");
    if (shape < -1 || shape > 2) return pv_set_error(PV_EINVAL, "MG level-0 tile shape must be -1, 0, 1 or 2 (got %d)", shape);
    if (!b->mg_ok) return pv_set_error(PV_EINVAL, "MG-PCG is not available for this grid size");
    b->mg.sh0 = b->mgf.sh0 = shape;
    {
        const int nrz = pvk_mg_nrz(&b->mg);
        b->mgw.nrz = nrz;
        b->mgw.nrrt = (b->mg.c0 > 0) ? nrz : 0;
    }
    return PV_OK;
}

int pv_set_momentum_rows(pv_bag *b, int rows)
{
    if (!b) return pv_set_error(PV_EINVAL, "null handle");
    if (rows != 1 && rows != 2 && rows != 4) return pv_set_error(PV_EINVAL, "momentum rows per block must be 1, 2 or 4 (got %d)", rows);
    b->mcg.band = rows;
    b->mcg.nb[0] = (b->NYM + rows - 1) / rows;
    b->mcg.nb[1] = (b->NYM - 1 + rows - 1) / rows;
    return PV_OK;
}

int pv_set_momentum_warm_start(pv_bag *b, int mode)
{
    if (!b) return pv_set_error(PV_EINVAL, "null handle");
    if (mode < 0 || mode > 1) return pv_set_error(PV_EINVAL, "momentum warm start must be 0 or 1 (got %d)", mode);
    b->m_warm = mode;
    b->have_mprev[0] = b->have_mprev[1] = b->have_mprev[2] = 0;
    return PV_OK;
}

int pv_set_momentum_preconditioner(pv_bag *b, int kind)
{
    if (!b) return pv_set_error(PV_EINVAL, "null handle");
    if (kind < 0 || kind > 1) return pv_set_error(PV_EINVAL, "momentum preconditioner must be 0 or 1 (got %d)", kind);
    if (kind == 1 && !b->ph.precision)
        return pv_set_error(PV_EINVAL, "the polynomial momentum preconditioner needs the mixed-precision form "
                                       "(it reads the FP32 operator coefficients)");
    b->m_pc = kind;
    return PV_OK;
}

int pv_set_momentum_cg_steps(pv_bag *b, int exact)
{
    if (!b) return pv_set_error(PV_EINVAL, "null handle");
    if (exact < 0 || exact > 1) return pv_set_error(PV_EINVAL, "momentum CG steps must be 0 or 1 (got %d)", exact);
    b->m_exact = exact;
    return PV_OK;
}

int pv_set_ch_solver(pv_bag *b, int spectral)
{
    if (!b) return pv_set_error(PV_EINVAL, "null handle");
    if (spectral < 0 || spectral > 1) return pv_set_error(PV_EINVAL, "CH solver must be 0 or 1 (got %d)", spectral);
    if (spectral && !(b->fft.ok && b->fft.dct_ok))
        return pv_set_error(PV_EINVAL, "spectral CH solve needs the FFT tables and a power-of-two NYM");
    b->ch_spectral = spectral;
    return PV_OK;
}

int pv_set_mg_cycle(pv_bag *b, int wlevels)
{
    int l, mask = 0;
    if (!b) return pv_set_error(PV_EINVAL, "null handle");
    if (wlevels < 0 || wlevels > PVK_MG_LOW + 1)
        return pv_set_error(PV_EINVAL, "MG W levels must be 0..%d (got %d)", PVK_MG_LOW + 1, wlevels);
    if (!b->mg_ok) return pv_set_error(PV_EINVAL, "MG-PCG is not available for this grid size");
    for (l = b->mg.c0; l < b->mg.c0 + wlevels && l < b->mg.nlev - 1; l++) mask |= 1 << l;
    b->mg.wmask = b->mgf.wmask = mask;
    return PV_OK;
}

int pv_set_mg_sweeps(pv_bag *b, int nu)
{
    if (!b) return pv_set_error(PV_EINVAL, "null handle");
    if (nu < 1 || nu > 2) return pv_set_error(PV_EINVAL, "MG sweeps must be 1 or 2 (got %d)", nu);
    b->mg.nu = b->mgf.nu = nu;
    return PV_OK;
}
