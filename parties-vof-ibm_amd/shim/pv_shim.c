/*
 * pv_shim.c -- the reference routines' own signatures over the C ABI of
 * parties_vof.h (declared in include/parties_vof_shim.h).
 *
 * A registry maps the reference's pointers (Cart3d_bag*, its MAC_grid*,
 * Parameters*, Pressure*, Velocity* and double*** arrays) onto a pv_bag and
 * its field ids.  Each routine looks up its arguments and calls the pv_
 * entry point that replaces it; nothing here touches array values except
 * pvs_push / pvs_pull, which move a bound array between the reference's
 * host layout and the device.
 */
#include "parties_vof_shim.h"

#include <pthread.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MAXBAGS 64
#define MAXARRS 1024

typedef struct {
    Cart3d_bag *db;
    pv_bag *h;
    MAC_grid *grid;
    Parameters *params;
    Pressure *p;
    Velocity *u, *v;
    double time_max;
    int NXM, NYM;
} bag_entry;

typedef struct {
    double ***arr;
    Cart3d_bag *db;
    int field, range;
} arr_entry;

static pthread_mutex_t g_mu = PTHREAD_MUTEX_INITIALIZER;
static __thread char g_err[512];
static bag_entry g_bags[MAXBAGS];
static arr_entry g_arrs[MAXARRS];
static pvs_fatal_fn g_fatal = NULL;
static void *g_fatal_ctx = NULL;

const char *pvs_last_error(void) { return g_err; }

static int shim_error(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

void pvs_set_fatal(pvs_fatal_fn fn, void *ctx)
{
    pthread_mutex_lock(&g_mu);
    g_fatal = fn;
    g_fatal_ctx = ctx;
    pthread_mutex_unlock(&g_mu);
}

/* Display_throw_error's role (IO/Display.c:910-937) */
static void fatal(const char *where, const char *what)
{
    char msg[768];
    pvs_fatal_fn fn;
    void *ctx;
    snprintf(msg, sizeof msg, "%s: %s", where, what);
    pthread_mutex_lock(&g_mu);
    fn = g_fatal;
    ctx = g_fatal_ctx;
    pthread_mutex_unlock(&g_mu);
    if (fn) {
        fn(msg, ctx);
        return;
    }
    fprintf(stderr, "PARTIES device path error: %s\n", msg);
    exit(EXIT_FAILURE);
}

/* a pv_ status: 0, or the fatal path with the library's message */
static int check(const char *where, int rc)
{
    if (rc != PV_OK) fatal(where, pv_last_error());
    return rc;
}

static bag_entry *find_bag_locked(const Cart3d_bag *db)
{
    int k;
    if (!db) return NULL;
    for (k = 0; k < MAXBAGS; k++)
        if (g_bags[k].db == db) return &g_bags[k];
    return NULL;
}

static int lookup_bag(const char *where, const Cart3d_bag *db, bag_entry *out)
{
    bag_entry *e;
    pthread_mutex_lock(&g_mu);
    e = find_bag_locked(db);
    if (e) *out = *e;
    pthread_mutex_unlock(&g_mu);
    if (!e) {
        fatal(where, "Cart3d_bag is not bound to a device subdomain (pvs_bind)");
        return -1;
    }
    return 0;
}

/* the bag whose grid or params pointer this is (routines without a bag) */
static int lookup_by_grid(const char *where, const MAC_grid *grid, const Parameters *params, bag_entry *out)
{
    int k, found = 0;
    pthread_mutex_lock(&g_mu);
    for (k = 0; k < MAXBAGS && !found; k++)
        if (g_bags[k].db && ((grid && g_bags[k].grid == grid) || (params && g_bags[k].params == params))) {
            *out = g_bags[k];
            found = 1;
        }
    pthread_mutex_unlock(&g_mu);
    if (!found) {
        fatal(where, "MAC_grid / Parameters do not belong to a bound Cart3d_bag");
        return -1;
    }
    return 0;
}

static int lookup_arr(const char *where, double ***arr, const Cart3d_bag *db, arr_entry *out)
{
    int k, found = 0;
    pthread_mutex_lock(&g_mu);
    for (k = 0; k < MAXARRS && !found; k++)
        if (g_arrs[k].arr == arr && arr && (!db || g_arrs[k].db == db)) {
            *out = g_arrs[k];
            found = 1;
        }
    pthread_mutex_unlock(&g_mu);
    if (!found) {
        fatal(where, "array is not bound to a device field of this Cart3d_bag (pvs_bind_field)");
        return -1;
    }
    return 0;
}

int pvs_bind(Cart3d_bag *db, pv_bag *h, MAC_grid *grid, Parameters *params, Pressure *p, Velocity *u,
             Velocity *v)
{
    pv_params prm;
    bag_entry *e;
    int k;
    if (!db || !h) return shim_error(PV_EINVAL, "pvs_bind: null bag or handle");
    if (pv_get_params(h, &prm) != PV_OK) return shim_error(PV_EINVAL, "pvs_bind: %s", pv_last_error());
    pthread_mutex_lock(&g_mu);
    e = find_bag_locked(db);
    for (k = 0; k < MAXBAGS && !e; k++)
        if (!g_bags[k].db) e = &g_bags[k];
    if (e) {
        e->db = db; e->h = h; e->grid = grid; e->params = params; e->p = p; e->u = u; e->v = v;
        e->time_max = 0.0;
        e->NXM = prm.NXM; e->NYM = prm.NYM;
    }
    pthread_mutex_unlock(&g_mu);
    return e ? PV_OK : shim_error(PV_ENOMEM, "pvs_bind: more than %d bags", MAXBAGS);
}

int pvs_bind_field(Cart3d_bag *db, double ***arr, int field, int range)
{
    arr_entry *slot = NULL;
    int k, rc = PV_OK;
    if (!db || !arr) return shim_error(PV_EINVAL, "pvs_bind_field: null bag or array");
    if (field < 0 || field >= PV_NFIELDS) return shim_error(PV_EINVAL, "pvs_bind_field: bad field %d", field);
    if (range != PVS_L_RANGE && range != PVS_G_RANGE)
        return shim_error(PV_EINVAL, "pvs_bind_field: bad range %d", range);
    pthread_mutex_lock(&g_mu);
    if (!find_bag_locked(db)) {
        rc = shim_error(PV_EINVAL, "pvs_bind_field: bag not bound");
    } else {
        for (k = 0; k < MAXARRS; k++)
            if (g_arrs[k].arr == arr) { slot = &g_arrs[k]; break; }
        for (k = 0; k < MAXARRS && !slot; k++)
            if (!g_arrs[k].arr) slot = &g_arrs[k];
        if (!slot) rc = shim_error(PV_ENOMEM, "pvs_bind_field: more than %d arrays", MAXARRS);
        else { slot->arr = arr; slot->db = db; slot->field = field; slot->range = range; }
    }
    pthread_mutex_unlock(&g_mu);
    return rc;
}

int pvs_unbind(Cart3d_bag *db)
{
    bag_entry *e;
    int k;
    pthread_mutex_lock(&g_mu);
    e = find_bag_locked(db);
    if (e) memset(e, 0, sizeof *e);
    for (k = 0; k < MAXARRS; k++)
        if (g_arrs[k].db == db) memset(&g_arrs[k], 0, sizeof g_arrs[k]);
    pthread_mutex_unlock(&g_mu);
    return e ? PV_OK : shim_error(PV_EINVAL, "pvs_unbind: bag not bound");
}

pv_bag *pvs_handle(Cart3d_bag *db)
{
    bag_entry *e;
    pv_bag *h = NULL;
    pthread_mutex_lock(&g_mu);
    e = find_bag_locked(db);
    if (e) h = e->h;
    pthread_mutex_unlock(&g_mu);
    return h;
}

int pvs_set_time_max(Cart3d_bag *db, double time_max)
{
    bag_entry *e;
    pthread_mutex_lock(&g_mu);
    e = find_bag_locked(db);
    if (e) e->time_max = time_max;
    pthread_mutex_unlock(&g_mu);
    return e ? PV_OK : shim_error(PV_EINVAL, "pvs_set_time_max: bag not bound");
}

/* index bounds of a range in the reference's arrays (Grid.c:62-64,
 * :187-202 with NZM = 1: NZ = 2) */
static void range_bounds(int range, int NXM, int NYM, int *k0, int *k1, int *j0, int *j1, int *i0, int *i1)
{
    if (range == PVS_L_RANGE) {
        *k0 = -3; *k1 = 2 + 3; *j0 = -3; *j1 = NYM + 1 + 3; *i0 = -3; *i1 = NXM + 1 + 3;
    } else {
        *k0 = 0; *k1 = 2; *j0 = 0; *j1 = NYM + 1; *i0 = 0; *i1 = NXM + 1;
    }
}

static int transfer(const char *where, Cart3d_bag *db, double ***arr, int to_device)
{
    bag_entry b;
    arr_entry a;
    double *canon;
    int k0, k1, j0, j1, i0, i1, i, j, k, rc;
    size_t W, n;
    {   /* a binding error is returned, not sent down the fatal path */
        bag_entry *e;
        int found = 0;
        pthread_mutex_lock(&g_mu);
        e = find_bag_locked(db);
        if (e) {
            b = *e;
            for (k = 0; k < MAXARRS && !found; k++)
                if (arr && g_arrs[k].arr == arr && g_arrs[k].db == db) { a = g_arrs[k]; found = 1; }
        }
        pthread_mutex_unlock(&g_mu);
        if (!e) return shim_error(PV_EINVAL, "%s: Cart3d_bag is not bound", where);
        if (!found) return shim_error(PV_EINVAL, "%s: array is not bound to this Cart3d_bag", where);
    }
    W = (size_t)b.NXM + 7;
    n = W * ((size_t)b.NYM + 7);
    canon = (double *)malloc(n * sizeof(double));
    if (!canon) return shim_error(PV_ENOMEM, "%s: out of host memory", where);
    range_bounds(a.range, b.NXM, b.NYM, &k0, &k1, &j0, &j1, &i0, &i1);
    /* read-modify-write: a G-range array leaves the device field's ghosts as they are */
    rc = pv_get_field(b.h, a.field, canon);
    if (rc != PV_OK) rc = shim_error(rc, "%s: %s", where, pv_last_error());
    if (rc == PV_OK && to_device) {
        for (j = j0; j < j1; j++)
            for (i = i0; i < i1; i++) canon[(size_t)(j + 3) * W + (size_t)(i + 3)] = arr[0][j][i];
        rc = pv_set_field(b.h, a.field, canon);
        if (rc != PV_OK) rc = shim_error(rc, "%s: %s", where, pv_last_error());
    } else if (rc == PV_OK) {
        for (k = k0; k < k1; k++)
            for (j = j0; j < j1; j++)
                for (i = i0; i < i1; i++) arr[k][j][i] = canon[(size_t)(j + 3) * W + (size_t)(i + 3)];
    }
    free(canon);
    return rc;
}

int pvs_push(Cart3d_bag *db, double ***arr) { return transfer("pvs_push", db, arr, 1); }
int pvs_pull(Cart3d_bag *db, double ***arr) { return transfer("pvs_pull", db, arr, 0); }

/* ------------------------------------------------------------------ */
/* the reference's routines                                            */
/* ------------------------------------------------------------------ */
#define BAG(where, db, b) \
    bag_entry b;          \
    if (lookup_bag(where, db, &b)) return
#define BAGV(where, db, b, bad) \
    bag_entry b;                \
    if (lookup_bag(where, db, &b)) return bad
#define FIELD(where, arr, db, a) \
    arr_entry a;                 \
    if (lookup_arr(where, arr, db, &a)) return

/* Temporal_int.c:270-448: while time - time_max < dt, one RK3 step (the CFL
 * step at stage 0 once time != 0, three substeps, time += dt * BETA2[rk]) */
int Temporal_int_rk3(Cart3d_bag *data_bag, Debug_trace *dtrace)
{
    pv_stats st;
    BAGV("Temporal_int_rk3", data_bag, b, -1);
    (void)dtrace;
    if (check("Temporal_int_rk3", pv_get_stats(b.h, &st))) return -1;
    while (st.time - b.time_max < st.dt) {
        if (check("Temporal_int_rk3", pv_Temporal_int_rk3(b.h, 1))) return -1;
        if (check("Temporal_int_rk3", pv_get_stats(b.h, &st))) return -1;
    }
    return 0;
}

void Temporal_int_all_the_equations(Cart3d_bag *data_bag, Debug_trace *dtrace)
{
    BAG("Temporal_int_all_the_equations", data_bag, b);
    (void)dtrace;
    check("Temporal_int_all_the_equations", pv_Temporal_int_all_the_equations(b.h));
}

double Dtime_cfl(Cart3d_bag *data_bag)
{
    double dt = -1.0;
    BAGV("Dtime_cfl", data_bag, b, -1.0);
    if (check("Dtime_cfl", pv_Dtime_cfl(b.h, &dt))) return -1.0;
    return dt;
}

void Communication_update_ghost_nodes_flow_variable(double ***data, char component, int pnodes,
                                                    Cart3d_bag *data_bag)
{
    static const char *W = "Communication_update_ghost_nodes_flow_variable";
    BAG(W, data_bag, b);
    FIELD(W, data, data_bag, a);
    check(W, pv_Communication_update_ghost_nodes_flow_variable(b.h, a.field, (unsigned char)component, pnodes));
}

/* no-argument routines of one bag */
#define SIMPLE(name, call)                     \
    void name(Cart3d_bag *db)                  \
    {                                          \
        BAG(#name, db, b);                     \
        check(#name, call(b.h));               \
    }

SIMPLE(VOF_DIFFUSE_init, pv_VOF_DIFFUSE_init)
SIMPLE(VOF_DIFFUSE_update_phase_cache, pv_VOF_DIFFUSE_update_phase_cache)
SIMPLE(VOF_DIFFUSE_compute_bulk_S, pv_VOF_DIFFUSE_compute_bulk_S)
SIMPLE(VOF_DIFFUSE_compute_psi_LG, pv_VOF_DIFFUSE_compute_psi_LG)
SIMPLE(VOF_DIFFUSE_step, pv_VOF_DIFFUSE_step)
SIMPLE(VOF_DIFFUSE_compute_f_sigma, pv_VOF_DIFFUSE_compute_f_sigma)
SIMPLE(VOF_DIFFUSE_apply_f_sigma_old, pv_VOF_apply_f_sigma_old)
SIMPLE(VOF_DIFFUSE_update_density_viscosity, pv_VOF_update_density_viscosity)
SIMPLE(VOF_update_density_viscosity, pv_VOF_update_density_viscosity)
SIMPLE(VOF_apply_f_sigma_old, pv_VOF_apply_f_sigma_old)
SIMPLE(VOF_swap_f_sigma, pv_VOF_swap_f_sigma)
SIMPLE(Velocity_add_gravity_2_RHS, pv_Velocity_add_gravity_2_RHS)
SIMPLE(Velocity_cell_center, pv_Velocity_cell_center)
SIMPLE(Velocity_u_set_implicit_explicit, pv_Velocity_u_set_implicit_explicit)
SIMPLE(Velocity_v_set_implicit_explicit, pv_Velocity_v_set_implicit_explicit)
SIMPLE(Velocity_u_set_RHS, pv_Velocity_u_set_RHS)
SIMPLE(Velocity_v_set_RHS, pv_Velocity_v_set_RHS)
SIMPLE(Pressure_set_RHS, pv_Pressure_set_RHS)
SIMPLE(Pressure_compute_preconditioner, pv_Pressure_compute_preconditioner)
SIMPLE(Pressure_project_velocity, pv_Pressure_project_velocity)

void Velocity_w_set_implicit_explicit(Cart3d_bag *data_bag)
{
    BAG("Velocity_w_set_implicit_explicit", data_bag, b);
    (void)b;
}

void Velocity_w_set_RHS(Cart3d_bag *data_bag)
{
    BAG("Velocity_w_set_RHS", data_bag, b);
    (void)b;
}

void VOF_DIFFUSE_set_boundary_values(double ***f, Cart3d_bag *db)
{
    BAG("VOF_DIFFUSE_set_boundary_values", db, b);
    FIELD("VOF_DIFFUSE_set_boundary_values", f, db, a);
    check("VOF_DIFFUSE_set_boundary_values", pv_VOF_DIFFUSE_set_boundary_values(b.h, a.field));
}

void VOF_DIFFUSE_compute_laplacian(double ***in, double ***lap, Cart3d_bag *db)
{
    BAG("VOF_DIFFUSE_compute_laplacian", db, b);
    FIELD("VOF_DIFFUSE_compute_laplacian", in, db, ai);
    FIELD("VOF_DIFFUSE_compute_laplacian", lap, db, ao);
    check("VOF_DIFFUSE_compute_laplacian", pv_VOF_DIFFUSE_compute_laplacian(b.h, ai.field, ao.field));
}

void VOF_DIFFUSE_advect_WENO5(Cart3d_bag *db, double ***rhs_out)
{
    BAG("VOF_DIFFUSE_advect_WENO5", db, b);
    FIELD("VOF_DIFFUSE_advect_WENO5", rhs_out, db, a);
    check("VOF_DIFFUSE_advect_WENO5", pv_VOF_DIFFUSE_advect_WENO5(b.h, a.field));
}

void VOF_DIFFUSE_explicit_diffusion(Cart3d_bag *db, double ***rhs_out)
{
    BAG("VOF_DIFFUSE_explicit_diffusion", db, b);
    FIELD("VOF_DIFFUSE_explicit_diffusion", rhs_out, db, a);
    check("VOF_DIFFUSE_explicit_diffusion", pv_VOF_DIFFUSE_explicit_diffusion(b.h, a.field));
}

void VOF_DIFFUSE_solve_implicit_biharmonic(Cart3d_bag *db, double ***rhs_explicit)
{
    int its = 0;
    double res = 0.0;
    BAG("VOF_DIFFUSE_solve_implicit_biharmonic", db, b);
    FIELD("VOF_DIFFUSE_solve_implicit_biharmonic", rhs_explicit, db, a);
    check("VOF_DIFFUSE_solve_implicit_biharmonic",
          pv_VOF_DIFFUSE_solve_implicit_biharmonic(b.h, a.field, &its, &res));
}

void Velocity_update_boundaries(double ***data, char component, int type, Cart3d_bag *data_bag)
{
    BAG("Velocity_update_boundaries", data_bag, b);
    FIELD("Velocity_update_boundaries", data, data_bag, a);
    (void)component;     /* the field decides u or v, as the reference's own caller pairs them */
    check("Velocity_update_boundaries", pv_Velocity_update_boundaries(b.h, a.field, type));
}

int Velocity_solve_cg(Velocity *vel, Cart3d_bag *data_bag)
{
    int its = 0, comp;
    BAGV("Velocity_solve_cg", data_bag, b, -1);
    if (vel && vel == b.u) comp = 'u';
    else if (vel && vel == b.v) comp = 'v';
    else {
        fatal("Velocity_solve_cg", "Velocity is neither the bag's u nor its v (w is inactive in 2-D)");
        return -1;
    }
    if (check("Velocity_solve_cg", pv_Velocity_solve_cg(b.h, comp, &its))) return -1;
    return its;
}

int Pressure_solve_cg(Cart3d_bag *data_bag)
{
    int its = 0;
    BAGV("Pressure_solve_cg", data_bag, b, -1);
    if (check("Pressure_solve_cg", pv_Pressure_solve_cg(b.h, &its))) return -1;
    return its;
}

void Pressure_operator_variableCoeff(double ***Aphi, double ***phi, double ***rho, MAC_grid *grid,
                                     Parameters *params, Cart3d_bag *data_bag)
{
    static const char *W = "Pressure_operator_variableCoeff";
    BAG(W, data_bag, b);
    FIELD(W, Aphi, data_bag, ao);
    FIELD(W, phi, data_bag, ai);
    FIELD(W, rho, data_bag, ar);
    (void)grid;
    (void)params;
    if (ar.field != PV_F_RHO) {
        fatal(W, "rho must be the bag's density array (the device operator's coefficients)");
        return;
    }
    check(W, pv_Pressure_operator_variableCoeff(b.h, ao.field, ai.field));
}

void Pressure_apply_BCs(double ***phi, MAC_grid *grid, Parameters *params)
{
    bag_entry b;
    arr_entry a;
    if (lookup_by_grid("Pressure_apply_BCs", grid, params, &b)) return;
    if (lookup_arr("Pressure_apply_BCs", phi, b.db, &a)) return;
    check("Pressure_apply_BCs", pv_Pressure_apply_BCs(b.h, a.field));
}

double Pressure_compute_velocity_divergence(Cart3d_bag *data_bag)
{
    double d = -1.0;
    BAGV("Pressure_compute_velocity_divergence", data_bag, b, -1.0);
    if (check("Pressure_compute_velocity_divergence", pv_Pressure_compute_velocity_divergence(b.h, &d)))
        return -1.0;
    return d;
}

int Pressure_solve_x_period(Pressure *p, MAC_grid *grid, Parameters *params)
{
    bag_entry b;
    if (lookup_by_grid("Pressure_solve_x_period", grid, params, &b)) return -1;
    if (p && b.p && p != b.p) {
        fatal("Pressure_solve_x_period", "Pressure does not belong to the bag of this grid");
        return -1;
    }
    if (check("Pressure_solve_x_period", pv_Pressure_solve_fft(b.h, PV_F_P_RHS, PV_F_PHI))) return -1;
    return 0;
}
