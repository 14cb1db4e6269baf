/*
 * pv_io.c -- the reference's Data / Resume HDF5 files for the 2-D
 * VOF_DIFFUSE path (include/parties_vof_io.h), written from and read into a
 * device subdomain through pv_get_field / pv_set_field.
 *
 * One rank holds the whole domain, so every field dataset is written whole:
 * (NZ, NY, NX) = (2, NYM+1, NXM+1) doubles, the G range of the reference's
 * Output_h5_flow_variable (IO/Output.c:1141-1197), both k slabs equal in 2-D.
 */
#include "parties_vof_io.h"

#include <hdf5.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static __thread char g_err[512];

const char *pv_io_last_error(void) { return g_err; }

static int io_error(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

typedef struct {
    pv_bag *bag;
    pv_params prm;
    int NX, NY, NZ;          /* G-range extents: NXM+1, NYM+1, 2 */
    size_t W, H;             /* canonical frame: NXM+7 by NYM+7 */
    double *canon;           /* one canonical field */
    double *g;               /* one dataset: NZ * NY * NX */
} io_ctx;

static int ctx_open(io_ctx *c, pv_bag *bag)
{
    memset(c, 0, sizeof *c);
    if (!bag) return io_error(PV_EINVAL, "null handle");
    if (pv_get_params(bag, &c->prm) != PV_OK) return io_error(PV_EINVAL, "%s", pv_last_error());
    c->bag = bag;
    c->NX = c->prm.NXM + 1; c->NY = c->prm.NYM + 1; c->NZ = 2;
    c->W = (size_t)c->prm.NXM + 7; c->H = (size_t)c->prm.NYM + 7;
    c->canon = (double *)malloc(c->W * c->H * sizeof(double));
    c->g = (double *)malloc((size_t)c->NZ * c->NY * c->NX * sizeof(double));
    if (!c->canon || !c->g) {
        free(c->canon); free(c->g);
        return io_error(PV_ENOMEM, "host allocation failed");
    }
    return PV_OK;
}

static void ctx_close(io_ctx *c)
{
    free(c->canon);
    free(c->g);
}

/* ---- writing ----------------------------------------------------------- */
static int write_dataset(hid_t f, const char *name, hid_t type, int ndim, const hsize_t *dim, const void *data)
{
    hid_t sp = H5Screate_simple(ndim, dim, NULL), ds;
    herr_t st;
    if (sp < 0) return io_error(PV_EHIP, "H5Screate_simple for %s failed", name);
    ds = H5Dcreate2(f, name, type, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    H5Sclose(sp);
    if (ds < 0) return io_error(PV_EHIP, "H5Dcreate2 %s failed", name);
    st = H5Dwrite(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
    H5Dclose(ds);
    return st < 0 ? io_error(PV_EHIP, "H5Dwrite %s failed", name) : PV_OK;
}

static int write_scalar_d(hid_t f, const char *name, double v)
{
    hsize_t d = 1;
    return write_dataset(f, name, H5T_NATIVE_DOUBLE, 1, &d, &v);
}

static int write_scalar_i(hid_t f, const char *name, int v)
{
    hsize_t d = 1;
    return write_dataset(f, name, H5T_NATIVE_INT, 1, &d, &v);
}

/* a device field's G range, both k slabs (field < 0: zeros) */
static int write_field(io_ctx *c, hid_t f, const char *name, int field)
{
    const hsize_t dim[3] = {(hsize_t)c->NZ, (hsize_t)c->NY, (hsize_t)c->NX};
    int i, j, k;
    if (field >= 0) {
        if (pv_get_field(c->bag, field, c->canon) != PV_OK) return io_error(PV_EHIP, "%s", pv_last_error());
        for (k = 0; k < c->NZ; k++)
            for (j = 0; j < c->NY; j++)
                for (i = 0; i < c->NX; i++)
                    c->g[((size_t)k * c->NY + j) * c->NX + i] = c->canon[(size_t)(j + 3) * c->W + (size_t)(i + 3)];
    } else {
        memset(c->g, 0, (size_t)c->NZ * c->NY * c->NX * sizeof(double));
    }
    return write_dataset(f, name, H5T_NATIVE_DOUBLE, 3, dim, c->g);
}

static hid_t create_file(const char *filename)
{
    return H5Fcreate(filename, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
}

static int make_group(hid_t f, const char *name)
{
    hid_t g = H5Gcreate2(f, name, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (g < 0) return io_error(PV_EHIP, "H5Gcreate2 %s failed", name);
    H5Gclose(g);
    return PV_OK;
}

#define TRY(x)                         \
    do {                               \
        if ((rc = (x)) != PV_OK) goto done; \
    } while (0)

/* Output.c:277-774 (VOF, VOF_DIFFUSE, SURFACE_TENSION; no CONC / LES / RANS) */
int pv_Output_h5_data(pv_bag *bag, const char *filename)
{
    io_ctx c;
    hid_t f = -1;
    int rc, i, k;
    pv_stats st;
    double *coord = NULL, slab;
    if ((rc = ctx_open(&c, bag)) != PV_OK) return rc;
    if (pv_get_stats(bag, &st) != PV_OK) { rc = io_error(PV_EHIP, "%s", pv_last_error()); goto done; }
    f = create_file(filename);
    if (f < 0) { rc = io_error(PV_EINVAL, "cannot create %s", filename); goto done; }
    TRY(write_scalar_d(f, "/time", st.time));
    TRY(make_group(f, "/grid"));
    TRY(write_scalar_i(f, "/grid/NX", c.NX));
    TRY(write_scalar_i(f, "/grid/NY", c.NY));
    TRY(write_scalar_i(f, "/grid/NZ", c.NZ));
    {   /* Grid.c:845-861 coordinates, index 0 .. N-1 */
        const int nmax = (c.NX > c.NY) ? c.NX : c.NY;
        const double dx = (c.prm.xmax - c.prm.xmin) / (double)(c.NX - 1);
        const double dy = (c.prm.ymax - c.prm.ymin) / (double)(c.NY - 1);
        hsize_t d;
        coord = (double *)malloc(sizeof(double) * (size_t)nmax);
        if (!coord) { rc = io_error(PV_ENOMEM, "host allocation failed"); goto done; }
        d = (hsize_t)c.NX;
        for (i = 0; i < c.NX; i++) coord[i] = c.prm.xmin + i * dx;
        TRY(write_dataset(f, "/grid/xu", H5T_NATIVE_DOUBLE, 1, &d, coord));
        for (i = 0; i < c.NX; i++) coord[i] = c.prm.xmin + (0.5 + i) * dx;
        TRY(write_dataset(f, "/grid/xc", H5T_NATIVE_DOUBLE, 1, &d, coord));
        d = (hsize_t)c.NY;
        for (i = 0; i < c.NY; i++) coord[i] = c.prm.ymin + i * dy;
        TRY(write_dataset(f, "/grid/yv", H5T_NATIVE_DOUBLE, 1, &d, coord));
        for (i = 0; i < c.NY; i++) coord[i] = c.prm.ymin + (0.5 + i) * dy;
        TRY(write_dataset(f, "/grid/yc", H5T_NATIVE_DOUBLE, 1, &d, coord));
        /* z: zmin = 0, dz = Lz / (NZ - 1) = the slab thickness (IO/Input.c:40-70) */
        slab = c.prm.twod_slab_thickness;
        d = (hsize_t)c.NZ;
        for (k = 0; k < c.NZ; k++) coord[k] = 0.0 + k * slab;
        TRY(write_dataset(f, "/grid/zw", H5T_NATIVE_DOUBLE, 1, &d, coord));
        for (k = 0; k < c.NZ; k++) coord[k] = 0.0 + (0.5 + k) * slab;
        TRY(write_dataset(f, "/grid/zc", H5T_NATIVE_DOUBLE, 1, &d, coord));
    }
    TRY(write_field(&c, f, "/u", PV_F_U));
    TRY(write_field(&c, f, "/v", PV_F_V));
    TRY(write_field(&c, f, "/w", -1));
    TRY(write_field(&c, f, "/p", PV_F_P));
    TRY(make_group(f, "/VOF"));
    TRY(write_field(&c, f, "/VOF/F", PV_F_C_L));           /* the phase cache's mirror of C_L */
    TRY(write_field(&c, f, "/VOF/C_L", PV_F_C_L));
    TRY(write_field(&c, f, "/VOF/C_S", -1));
    TRY(write_field(&c, f, "/VOF/C_G", PV_F_C_G));
    TRY(write_field(&c, f, "/VOF/psi", PV_F_PSI));
    TRY(write_field(&c, f, "/VOF/psi_LG", PV_F_PSI_LG));
    TRY(write_field(&c, f, "/VOF/lap_C", PV_F_LAP_C));
    TRY(write_field(&c, f, "/VOF/bulk_S", PV_F_BULK_S));
    TRY(write_field(&c, f, "/VOF/ch_rhs_n", PV_F_CH_RHS_N));
    TRY(write_field(&c, f, "/VOF/ch_rhs_nm1", PV_F_CH_RHS_NM1));
    TRY(write_field(&c, f, "/VOF/normal_x", -1));
    TRY(write_field(&c, f, "/VOF/normal_y", -1));
    TRY(write_field(&c, f, "/VOF/normal_z", -1));
    TRY(write_field(&c, f, "/VOF/rho", PV_F_RHO));
    TRY(write_field(&c, f, "/VOF/mu", PV_F_MU));
    TRY(write_field(&c, f, "/VOF/f_sigma_old_x", PV_F_FSX_OLD));
    TRY(write_field(&c, f, "/VOF/f_sigma_old_y", PV_F_FSY_OLD));
    TRY(write_field(&c, f, "/VOF/f_sigma_old_z", -1));
done:
    if (f >= 0) H5Fclose(f);
    free(coord);
    ctx_close(&c);
    return rc;
}

/* Output.c:44-276 (VOF, VOF_DIFFUSE, SURFACE_TENSION) */
int pv_Output_h5_resume(pv_bag *bag, const char *filename, const pv_io_counters *counters)
{
    io_ctx c;
    hid_t f = -1;
    int rc;
    pv_stats st;
    pv_io_counters zero;
    hsize_t dt = PV_IO_NTIMER;
    if ((rc = ctx_open(&c, bag)) != PV_OK) return rc;
    if (!counters) { memset(&zero, 0, sizeof zero); counters = &zero; }
    if (pv_get_stats(bag, &st) != PV_OK) { rc = io_error(PV_EHIP, "%s", pv_last_error()); goto done; }
    f = create_file(filename);
    if (f < 0) { rc = io_error(PV_EINVAL, "cannot create %s", filename); goto done; }
    TRY(make_group(f, "/Resume"));
    TRY(write_scalar_d(f, "/time", st.time));
    TRY(write_scalar_d(f, "/dt", st.dt));
    TRY(write_scalar_d(f, "/dt_old", st.dt_old));
    TRY(write_scalar_i(f, "/ntime", st.ntime));
    TRY(write_scalar_i(f, "/noutput", counters->noutput));
    TRY(write_scalar_i(f, "/noutput_2d", counters->noutput_2d));
    TRY(write_scalar_d(f, "/output_time", counters->output_time));
    TRY(write_scalar_d(f, "/output_time_2d", counters->output_time_2d));
    TRY(write_dataset(f, "/timer", H5T_NATIVE_DOUBLE, 1, &dt, counters->timer));
    TRY(write_field(&c, f, "/Resume/F", PV_F_C_L));
    TRY(write_field(&c, f, "/Resume/C_L", PV_F_C_L));
    TRY(write_field(&c, f, "/Resume/C_S", -1));
    TRY(write_field(&c, f, "/Resume/ch_rhs_n", PV_F_CH_RHS_N));
    TRY(write_field(&c, f, "/Resume/ch_rhs_nm1", PV_F_CH_RHS_NM1));
    TRY(write_field(&c, f, "/Resume/mu", PV_F_MU));
    TRY(write_field(&c, f, "/Resume/rho", PV_F_RHO));
    TRY(write_field(&c, f, "/Resume/flux_x", PV_F_FLUX_X));
    TRY(write_field(&c, f, "/Resume/flux_y", PV_F_FLUX_Y));
    TRY(write_field(&c, f, "/Resume/flux_z", -1));
    TRY(write_field(&c, f, "/Resume/normal_x", -1));
    TRY(write_field(&c, f, "/Resume/normal_y", -1));
    TRY(write_field(&c, f, "/Resume/normal_z", -1));
    TRY(write_field(&c, f, "/Resume/psi_LG", PV_F_PSI_LG));
    TRY(write_field(&c, f, "/Resume/f_sigma_old_x", PV_F_FSX_OLD));
    TRY(write_field(&c, f, "/Resume/f_sigma_old_y", PV_F_FSY_OLD));
    TRY(write_field(&c, f, "/Resume/f_sigma_old_z", -1));
done:
    if (f >= 0) H5Fclose(f);
    ctx_close(&c);
    return rc;
}

/* ---- reading ----------------------------------------------------------- */
static int exists(hid_t f, const char *name)
{
    /* every component of the path must exist for H5Lexists to be asked */
    char buf[256];
    const char *p = name;
    if (strlen(name) >= sizeof buf) return 0;
    while ((p = strchr(p + 1, '/')) != NULL) {
        memcpy(buf, name, (size_t)(p - name));
        buf[p - name] = 0;
        if (H5Lexists(f, buf, H5P_DEFAULT) <= 0) return 0;
    }
    return H5Lexists(f, name, H5P_DEFAULT) > 0;
}

/* Resume.c:692-733: the whole dataset, dimensions checked */
static int read_dataset(hid_t f, const char *name, hid_t type, int ndim, const hsize_t *dim, void *data)
{
    hid_t ds, sp;
    hsize_t df[3];
    int nd, i;
    herr_t st;
    if (!exists(f, name)) return io_error(PV_EINVAL, "dataset %s missing", name);
    ds = H5Dopen2(f, name, H5P_DEFAULT);
    if (ds < 0) return io_error(PV_EINVAL, "open dataset %s failed", name);
    sp = H5Dget_space(ds);
    nd = H5Sget_simple_extent_ndims(sp);
    if (nd != ndim || nd > 3) {
        H5Sclose(sp); H5Dclose(ds);
        return io_error(PV_EINVAL, "dataset %s has %d dimensions, expected %d", name, nd, ndim);
    }
    H5Sget_simple_extent_dims(sp, df, NULL);
    H5Sclose(sp);
    for (i = 0; i < ndim; i++)
        if (df[i] != dim[i]) {
            H5Dclose(ds);
            return io_error(PV_EINVAL, "dataset %s: dimension %d is %d, expected %d", name, i, (int)df[i],
                            (int)dim[i]);
        }
    st = H5Dread(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
    H5Dclose(ds);
    return st < 0 ? io_error(PV_EHIP, "read %s failed", name) : PV_OK;
}

/* Resume_h5_flow_variable (Resume.c:813-900): the G range of slab k = 0
 * into a device field; the ghosts keep the device field's values until the
 * caller's exchange / boundary routine fills them */
static int read_field(io_ctx *c, hid_t f, const char *name, int field)
{
    const hsize_t dim[3] = {(hsize_t)c->NZ, (hsize_t)c->NY, (hsize_t)c->NX};
    int i, j, rc;
    if ((rc = read_dataset(f, name, H5T_NATIVE_DOUBLE, 3, dim, c->g)) != PV_OK) return rc;
    if (pv_get_field(c->bag, field, c->canon) != PV_OK) return io_error(PV_EHIP, "%s", pv_last_error());
    for (j = 0; j < c->NY; j++)
        for (i = 0; i < c->NX; i++) c->canon[(size_t)(j + 3) * c->W + (size_t)(i + 3)] = c->g[(size_t)j * c->NX + i];
    if (pv_set_field(c->bag, field, c->canon) != PV_OK) return io_error(PV_EHIP, "%s", pv_last_error());
    return PV_OK;
}

static int zero_field(io_ctx *c, int field)
{
    return pv_fill_field(c->bag, field, 0.0) == PV_OK ? PV_OK : io_error(PV_EHIP, "%s", pv_last_error());
}

static int comm(io_ctx *c, int field, int component)
{
    return pv_Communication_update_ghost_nodes_flow_variable(c->bag, field, component, 3) == PV_OK
               ? PV_OK : io_error(PV_EHIP, "%s", pv_last_error());
}

int pv_Resume_h5(pv_bag *bag, const char *resume_file, const char *data_file, pv_io_counters *counters)
{
    io_ctx c;
    hid_t f = -1;
    int rc, ntime = 0, have_rho, have_mu, i, j;
    double t = 0.0, dt = 0.0, dt_old = 0.0;
    hsize_t one = 1, ntim = PV_IO_NTIMER;
    pv_io_counters tmp;
    const int O = 'o', C_ = 'c';                    /* VOLUME_FRACTION, cell type */
    if ((rc = ctx_open(&c, bag)) != PV_OK) return rc;
    if (!counters) counters = &tmp;
    memset(counters, 0, sizeof *counters);
    /* Resume.c:31-338: scalars and the VOF fields */
    f = H5Fopen(resume_file, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) { rc = io_error(PV_EINVAL, "cannot open %s", resume_file); goto done; }
    TRY(read_dataset(f, "/time", H5T_NATIVE_DOUBLE, 1, &one, &t));
    TRY(read_dataset(f, "/dt", H5T_NATIVE_DOUBLE, 1, &one, &dt));
    TRY(read_dataset(f, "/dt_old", H5T_NATIVE_DOUBLE, 1, &one, &dt_old));
    TRY(read_dataset(f, "/ntime", H5T_NATIVE_INT, 1, &one, &ntime));
    TRY(read_dataset(f, "/noutput", H5T_NATIVE_INT, 1, &one, &counters->noutput));
    TRY(read_dataset(f, "/noutput_2d", H5T_NATIVE_INT, 1, &one, &counters->noutput_2d));
    TRY(read_dataset(f, "/output_time", H5T_NATIVE_DOUBLE, 1, &one, &counters->output_time));
    TRY(read_dataset(f, "/output_time_2d", H5T_NATIVE_DOUBLE, 1, &one, &counters->output_time_2d));
    TRY(read_dataset(f, "/timer", H5T_NATIVE_DOUBLE, 1, &ntim, counters->timer));
    TRY(read_field(&c, f, "/Resume/F", PV_F_F));
    TRY(comm(&c, PV_F_F, O));
    if (exists(f, "/Resume/C_L")) TRY(read_field(&c, f, "/Resume/C_L", PV_F_C_L));
    else {
        if (pv_get_field(bag, PV_F_F, c.canon) != PV_OK || pv_set_field(bag, PV_F_C_L, c.canon) != PV_OK) {
            rc = io_error(PV_EHIP, "%s", pv_last_error());
            goto done;
        }
    }
    TRY(comm(&c, PV_F_C_L, O));
    if (exists(f, "/Resume/ch_rhs_n")) TRY(read_field(&c, f, "/Resume/ch_rhs_n", PV_F_CH_RHS_N));
    else TRY(zero_field(&c, PV_F_CH_RHS_N));
    TRY(comm(&c, PV_F_CH_RHS_N, O));
    if (exists(f, "/Resume/ch_rhs_nm1")) TRY(read_field(&c, f, "/Resume/ch_rhs_nm1", PV_F_CH_RHS_NM1));
    else TRY(zero_field(&c, PV_F_CH_RHS_NM1));
    TRY(comm(&c, PV_F_CH_RHS_NM1, O));
    have_mu = exists(f, "/Resume/mu");
    have_rho = exists(f, "/Resume/rho");
    if (have_mu) { TRY(read_field(&c, f, "/Resume/mu", PV_F_MU)); TRY(comm(&c, PV_F_MU, O)); }
    if (have_rho) { TRY(read_field(&c, f, "/Resume/rho", PV_F_RHO)); TRY(comm(&c, PV_F_RHO, O)); }
    if (exists(f, "/Resume/flux_x")) TRY(read_field(&c, f, "/Resume/flux_x", PV_F_FLUX_X));
    else TRY(zero_field(&c, PV_F_FLUX_X));
    TRY(comm(&c, PV_F_FLUX_X, 'p'));
    if (exists(f, "/Resume/flux_y")) TRY(read_field(&c, f, "/Resume/flux_y", PV_F_FLUX_Y));
    else TRY(zero_field(&c, PV_F_FLUX_Y));
    TRY(comm(&c, PV_F_FLUX_Y, 'q'));
    if (exists(f, "/Resume/f_sigma_old_x")) TRY(read_field(&c, f, "/Resume/f_sigma_old_x", PV_F_FSX_OLD));
    else TRY(zero_field(&c, PV_F_FSX_OLD));
    if (exists(f, "/Resume/f_sigma_old_y")) TRY(read_field(&c, f, "/Resume/f_sigma_old_y", PV_F_FSY_OLD));
    else TRY(zero_field(&c, PV_F_FSY_OLD));
    if (!have_rho || !have_mu) {
        if (pv_VOF_update_density_viscosity(bag) != PV_OK) { rc = io_error(PV_EHIP, "%s", pv_last_error()); goto done; }
    }
    /* C_G = clamp01(1 - C_L - C_S) over the G range, then its exchange */
    if (pv_get_field(bag, PV_F_C_L, c.canon) != PV_OK) { rc = io_error(PV_EHIP, "%s", pv_last_error()); goto done; }
    {
        double *cg = (double *)malloc(c.W * c.H * sizeof(double));
        if (!cg) { rc = io_error(PV_ENOMEM, "host allocation failed"); goto done; }
        if (pv_get_field(bag, PV_F_C_G, cg) != PV_OK) { free(cg); rc = io_error(PV_EHIP, "%s", pv_last_error()); goto done; }
        for (j = 0; j < c.NY; j++)
            for (i = 0; i < c.NX; i++) {
                const size_t o = (size_t)(j + 3) * c.W + (size_t)(i + 3);
                double g = 1.0 - c.canon[o] - 0.0;
                cg[o] = (g < 0.0) ? 0.0 : ((g > 1.0) ? 1.0 : g);
            }
        rc = pv_set_field(bag, PV_F_C_G, cg);
        free(cg);
        if (rc != PV_OK) { rc = io_error(PV_EHIP, "%s", pv_last_error()); goto done; }
    }
    TRY(comm(&c, PV_F_C_G, O));
    H5Fclose(f);
    f = -1;
    /* Resume.c:341-401: u, v, (w), p from the Data file */
    f = H5Fopen(data_file, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) { rc = io_error(PV_EINVAL, "cannot open %s", data_file); goto done; }
    TRY(read_field(&c, f, "/u", PV_F_U));
    TRY(read_field(&c, f, "/v", PV_F_V));
    TRY(read_field(&c, f, "/p", PV_F_P));
    if (pv_Velocity_update_boundaries(bag, PV_F_U, PV_VEL_TYPE_NORMAL) != PV_OK ||
        pv_Velocity_update_boundaries(bag, PV_F_V, PV_VEL_TYPE_NORMAL) != PV_OK) {
        rc = io_error(PV_EHIP, "%s", pv_last_error());
        goto done;
    }
    TRY(comm(&c, PV_F_P, C_));
    /* the time state of Parameters; the next step starts at RK stage 0 */
    if (pv_set_time_state(bag, t, dt, dt_old, 0) != PV_OK || pv_set_step_count(bag, ntime) != PV_OK) {
        rc = io_error(PV_EHIP, "%s", pv_last_error());
        goto done;
    }
    rc = PV_OK;
done:
    if (f >= 0) H5Fclose(f);
    ctx_close(&c);
    return rc;
}
