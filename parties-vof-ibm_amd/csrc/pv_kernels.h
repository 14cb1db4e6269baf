/*
 * pv_kernels.h -- interface between the C host layer (host directory) and the
 * gfx950 kernels (csrc/).  Plain C: device pointers, sizes, a stream.
 *
 * Device field layout: one 2-D slab per field, row pitch P doubles (multiple
 * of 16, i.e. 128 B), ghost depth 3.  Every field pointer handed to a
 * launcher points at cell (0,0); cell (i,j), i,j in [-3, N+3], is at
 * f[j*P + i].  Row i = 0 starts on a 128-byte boundary.
 */
#ifndef PV_KERNELS_H
#define PV_KERNELS_H
#include <stddef.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Device-resident scalars of one handle (read and written by kernels, so a
 * whole substep runs without host round trips). */
typedef struct pvk_scal {
    double dt, dt_old, time;
    /* Cahn-Hilliard biharmonic CG (VOF_DIFFUSE.c:339-429) */
    double ch_rzb[2];        /* <r,r> of iterations n (slot n & 1) and n - 1 */
    double ch_bnorm, ch_rel;
    int    ch_iter, ch_done;
    /* momentum Jacobi-PCG per component 0 = u, 1 = v (msolve_cg.c:525-722) */
    double m_rzb[2][2];      /* [component][n & 1]: <r, z> of iterations n and n - 1 */
    double m_rms[2];
    int    m_iter[2], m_done[2];
    /* pressure PCG (psolve_cg.c:130-293 or FFT-PCG) */
    double p_gamma, p_alpha, p_beta, p_res, p_bnorm, p_mean;
    int    p_iter, p_done;
    double clamp_delta, div_max, max_idt, dt_new;
    double clamp_delta_int;  /* the clamp's correction over interior cells only */
    double red[8];           /* raw reduction results, for tests */
} pvk_scal;

/* Geometry and metrics (device pointers offset so that m[i] is valid for
 * i in [-3, N+2]); Grid.c:385-420 arrays. */
typedef struct pvk_grid {
    int NXM, NYM, P;
    int nbx, nby;            /* interior launch grid (64 x 4 cell tiles) */
    double slab;             /* dummy_z_slab_thickness (TwodOps.c:330-338) */
    const double *idx_u, *idx_c, *dx_u, *dx_c;
    const double *idy_v, *idy_c, *dy_v, *dy_c;
} pvk_grid;

/* Physical and solver constants (IO/default.inp names). */
typedef struct pvk_phys {
    double Re, iRe, Pe_CH, Cn, We, rho_g, mu_g, bih;
    double gx, gy, Ri, dp_dx, rho1, rho2, cfl, max_dt, default_dt;
    double ch_tol, CG_ETOL, P_CG_ETOL, p_rel_tol;
    int    weno_order, constant_dt, CG_MAXIT, P_CG_MAXIT;
    int    ch_iter_max, p_max_iter, precision, pad;
} pvk_phys;

/* Reduction scratch: per-block partial sums (deterministic two-stage sums). */
typedef struct pvk_red {
    double *part;            /* >= 4 * nblocks doubles */
    int nblocks;
} pvk_red;

/* FFT-x / DCT-y constant-coefficient solve tables (csrc/pv_fft.hip). */
typedef struct pvk_fft {
    int ok;                   /* 0: sizes not supported (NXM must be 2^a 3^b, 16 | NXM) */
    int mx;                   /* NXM / 2: length of the complex row transforms */
    int pow2x;                /* 1: mx is a power of two (shift-indexed passes) */
    int dct_ok;               /* 1: NYM is a power of two (the y transforms of the CH start) */
    int logmx, logmy;         /* log2(NXM/2), log2(NYM/2) where they are powers of two */
    const double *rtx, *rty;  /* (cos, sin)(2 pi k / N), k <= N/2 */
    const double *dcty;       /* (cos, sin)(pi k / (2 NYM)), k < NYM */
    const double *lamx, *lamy;/* 1-D eigenvalue magnitudes */
    const double *wx, *wy;    /* (cos, -sin)(2 pi t / M), t < M: exp(-2 pi i t / M) */
    double *spec;             /* NXM * NYM spectral scratch, row-major */
    /* tridiagonal solves along y on the columns of spec (csrc/pv_pcg.hip,
     * k_tri_cols): row coefficients as_j, an_j (psolve_fft.c:467-591), the
     * modified wavenumber of each halfcomplex column, and the launch shape
     * (tri_kw columns x tri_s segments of tri_l rows per block) */
    const double *tri_as, *tri_an, *tri_mw;
    int tri_kw, tri_lkw, tri_s, tri_l;
} pvk_fft;

/* Work set of the fused FFT-preconditioned PCG (csrc/pv_pcg.hip). */
typedef struct pvk_pcg {
    double *phi, *r, *z, *d[2], *Ad;  /* fields, cell (0,0); d double-buffered */
    double *bx, *by;                  /* operator face coefficients (per solve) */
    const double *rhs, *rho;
    double *specT;                    /* NXM * NYM transposed spectrum */
    double *dad[2], *rr, *rz[2], *mean;  /* per-block partial sums; dad by iteration parity */
    double *zad;                      /* <A d_n, z_{n+1}> partials of the flexible (learned) form */
    double *sad;                      /* non-NULL: partials of sum(A d), one per row, so the residual
                                         update can keep r at zero mean (MG-PCG) */
    double *r2;                       /* MG-PCG with tiled levels: the residual alternates between r
                                         (r_0, r_2, ...) and r2 (r_1, r_3, ...): the level-0 smoothing
                                         forms r_{n+1} from r_n and reads halos of it */
    double *rrt;                      /* <r_{n+1}, r_{n+1}> partials of those kernels, one per tile */
    int nrrt;                         /* their count (0: r is updated in place, partials in rr) */
    int ndad, nrr, nrz, nmean;        /* partial counts */
    int flex;                         /* 1: flexible CG, beta = -<z_{n+1}, A d_n> / <d_n, A d_n> */
    int band, rrow, rdct;             /* rows per block: bands, FFT rows, DCT rows */
    int keep_pin;                     /* 1: the y solve keeps the reference's pin of mode (0,0) at j = 0
                                         (the standalone solve); 0: it also removes the mean of that
                                         column (the PCG preconditioner's pseudo-inverse) */
    double chb;                       /* > 0: the FFT stages solve (a_dt + chb L^2) z = r (Cahn-Hilliard,
                                         a_dt of RK stage chstage) instead of L z = r, and run even
                                         after the pressure solve's stop flag */
    int chstage;
} pvk_pcg;

/* Work sets of the fused CG loops (csrc/pv_cg.hip).  Partial-sum arrays
 * hold one double per band. */
typedef struct pvk_ccg {             /* Cahn-Hilliard biharmonic CG */
    double *p[2];                    /* search direction, double-buffered */
    double *bb, *rr[2], *pap;        /* partials: ||b||^2, ||r||^2 (by parity), <p, Ap> */
    int nb, band;                    /* bands (= blocks) and rows per band */
    int npap;                        /* <p, Ap> partials: one per 128 x 8 tile */
    int nt0;                         /* start partials (bb, rr[0] at n = 0): one per 64 x 16 tile */
} pvk_ccg;

#define PVK_MOM_BAND 4                /* most rows per block of the momentum CG kernels (1, 2 or 4;
                                         pvk_mcg.band, default 2) */
typedef struct pvk_mcg {             /* momentum Jacobi-PCG, [0] = u, [1] = v */
    double *x[2], *r[2], *dg[2], *Ad[2];
    double *d[2][2];                 /* [component][n & 1]; d[c][0] is the public field */
    const double *rhs[2];
    const double *rho, *mu;
    double *rz[2][2], *dad[2], *rr[2];
    int nb[2], band;                 /* nb[c] = 0 leaves component c out */
    int maxit;
    float *cf[2][5];                 /* FP32 form: operator coefficients (centre, E, W, N, S) of each
                                        unknown, pitch P, formed by the start kernel */
    int exact;                       /* 0: EPS = 1e-14 added to the step denominators (the
                                        reference); 1: the denominators as they are, a zero
                                        one giving a zero step */
    int pc;                          /* 0: Jacobi (the reference); 1: first-order Neumann
                                        polynomial of the Jacobi splitting (FP32 form) */
    double *r2[2], *z[2];            /* pc = 1: r of odd iterations, preconditioned residual */
    double tol, in_tot[2];           /* CG_ETOL; 1 / number of unknowns */
} pvk_mcg;

/* Learned preconditioner network (csrc/pv_cnn.hip): parameters, Adam state
 * and the activation / gradient tensors of the largest batch seen so far. */
#define PVK_CNN_NPARAM 38177
#define PVK_CNN_RSPLIT 32                /* row groups of the gradient reduction */
typedef struct pvk_cnn {
    float *w, *g, *m, *v;            /* FP32 parameters, gradient, Adam moments */
    unsigned short *wb, *wt;         /* BF16 32 -> 32 weights: forward, flipped transpose */
    unsigned short *act[5];          /* BF16 NHWC activations of layers 1..5 */
    unsigned short *gz[2];           /* BF16 NHWC gradients w.r.t. pre-activations */
    float *gout;                     /* dLoss / d out */
    float *part, *lpart;             /* per-block partials: gradients, loss */
    float *rpart;                    /* PVK_CNN_RSPLIT x 9248 row-group sums of part */
    size_t cap_pix;                  /* B * H * W the tensors hold */
} pvk_cnn;

/* ---- ghost fills (kind codes in pv_kernels.hip k_ghost) ---- */
void pvk_bc_ch(const pvk_grid *g, double *f, hipStream_t s);          /* VOF_DIFFUSE_set_boundary_values */
void pvk_halo_cell(const pvk_grid *g, double *f, hipStream_t s);
void pvk_halo_u(const pvk_grid *g, double *f, hipStream_t s);
/* x exchange of pnodes (1..3) ghost layers over rows -pnodes .. NYM+pnodes;
 * ulike: u-type wrap (pnodes+1 layers below, image column 0 included) */
void pvk_xwrap(const pvk_grid *g, double *f, int ulike, int pnodes, hipStream_t s);
void pvk_vel_bc(const pvk_grid *g, double *f, int comp, int type, hipStream_t s);
void pvk_p_bc(const pvk_grid *g, double *f, hipStream_t s);          /* Pressure_apply_BCs + cell halo */
void pvk_ghost_multi(const pvk_grid *g, double *f0, double *f1, double *f2, double *f3, int nf, int kind,
                     hipStream_t s);
void pvk_fill(double *f, size_t n, double val, hipStream_t s);
void pvk_copy(const pvk_grid *g, const double *src, double *dst, hipStream_t s);

/* ---- Cahn-Hilliard ---- */
void pvk_laplacian(const pvk_grid *g, const double *in, double *out, hipStream_t s);
void pvk_bulk_S(const pvk_grid *g, const pvk_phys *ph, const double *C, double *S, hipStream_t s);
void pvk_weno_advect(const pvk_grid *g, const pvk_phys *ph, const double *C, const double *u,
                     const double *v, double *rhs, hipStream_t s);
void pvk_explicit_diffusion(const pvk_grid *g, const pvk_phys *ph, const double *C, double *rhs,
                            hipStream_t s);
/* b -= (Cn^2/Pe) L^2 C (standalone biharmonic solve) */
void pvk_ch_sub_bih(const pvk_grid *g, const pvk_phys *ph, const double *C, double *b, hipStream_t s);
/* fused biharmonic CG (csrc/pv_cg.hip): start from the explicit CH
 * right-hand side (b == NULL: WENO5 + diffusion + RK3 assembly) or from a
 * given b; then iterations n = 0, 1, ...: half 0 = apply, half 1 = update */
/* b == NULL: the explicit right-hand side and the CG start from C (bout, if
 * not NULL, receives b); else r = p = b - A C for a given b */
void pvk_ccg_init(const pvk_grid *g, const pvk_phys *ph, int stage, pvk_scal *sc, const pvk_ccg *w,
                  const double *C, const double *u, const double *v, double *rn, const double *b,
                  double *r, double *p, double *bout, hipStream_t s);
void pvk_ccg_iter(const pvk_grid *g, const pvk_phys *ph, int stage, pvk_scal *sc, const pvk_ccg *w,
                  int half, int n, double *x, double *r, double *Ap, hipStream_t s);
void pvk_clamp_props(const pvk_grid *g, const pvk_phys *ph, double *C, double *G, double *rho,
                     double *mu, const pvk_red *rd, pvk_scal *sc, hipStream_t s);
void pvk_phase_props(const pvk_grid *g, const pvk_phys *ph, const double *C, double *G, double *rho,
                     double *mu, hipStream_t s);
void pvk_psi_fsigma(const pvk_grid *g, const pvk_phys *ph, const double *C, const double *G,
                    double *lap, double *psi, double *fx, double *fy, hipStream_t s);
void pvk_psi_only(const pvk_grid *g, const pvk_phys *ph, const double *C, double *psi, hipStream_t s);
void pvk_clamp_init(const pvk_grid *g, const double *F, double *C, hipStream_t s);
/* temporal start of the momentum solves: mode 0 s = x, x += scale * d; mode 1 d = x - s */
void pvk_mom_warm(const pvk_grid *g, int mode, double scale, double *u, double *v, double *su, double *sv,
                  double *du, double *dv, hipStream_t s);
/* C = clamp01(F) over the whole L range, ghosts included */
void pvk_clamp_slab(const pvk_grid *g, const double *F, double *C, hipStream_t s);

/* ---- momentum ---- */
void pvk_cell_center(const pvk_grid *g, const double *u, const double *v, double *ub, double *vb,
                     hipStream_t s);
void pvk_mom_terms(const pvk_grid *g, const pvk_phys *ph, int stage, const pvk_scal *sc, int comp,
                   const double *u, const double *v, const double *mu, const double *rho,
                   const double *p, double *eold, double *vold, const double *fold,
                   double *rhs, double *ex_out, double *im_out, double *ve_out, int write_terms,
                   int add_force, hipStream_t s);
void pvk_mom_force(const pvk_grid *g, const pvk_phys *ph, int comp, const double *rho, const double *fold,
                   double *rhs, int which, hipStream_t s);
void pvk_mom_matvec(const pvk_grid *g, const pvk_phys *ph, int stage, const pvk_scal *sc, int comp,
                    const double *x, const double *rho, const double *mu, double *Ax, hipStream_t s);
/* fused momentum Jacobi-PCG for the components with nb[c] > 0 */
void pvk_mcg_init(const pvk_grid *g, const pvk_phys *ph, int stage, pvk_scal *sc, const pvk_mcg *w,
                  hipStream_t s);
/* pc = 1: z_0 = M r_0 into d[c][0] and <r_0, z_0> after pvk_mcg_init; after the
 * loop, the final residual back into r[c] when it ended in r2[c] */
void pvk_mcg_pc_start(const pvk_grid *g, const pvk_mcg *w, hipStream_t s);
void pvk_mcg_pc_finish(const pvk_grid *g, const pvk_scal *sc, const pvk_mcg *w, hipStream_t s);
void pvk_mcg_iter(const pvk_grid *g, const pvk_phys *ph, int stage, pvk_scal *sc, const pvk_mcg *w,
                  int half, int n, hipStream_t s);

/* ---- pressure ---- */
void pvk_p_rhs2(const pvk_grid *g, int stage, const pvk_scal *sc, const double *u, const double *v,
                const double *rho, double *rhs, double *M, hipStream_t s);
void pvk_p_operator(const pvk_grid *g, const double *rho, const double *x, double *Ax, hipStream_t s);
void pvk_p_cg_init2(const pvk_grid *g, const double *rhs, const double *M, double *phi, double *r, double *d,
                    int use_M, int shift_mean, const pvk_scal *sc, const pvk_red *rd, hipStream_t s);
void pvk_p_cg_apply(const pvk_grid *g, pvk_scal *sc, const double *rho, const double *d, double *Ad,
                    const pvk_red *rd, hipStream_t s);
void pvk_p_cg_update(const pvk_grid *g, pvk_scal *sc, double *phi, const double *d, double *r,
                     const double *Ad, const double *M, int use_M, const pvk_red *rd,
                     hipStream_t s);
void pvk_p_cg_dupdate(const pvk_grid *g, pvk_scal *sc, const double *z, const double *M, int use_M,
                      double *d, hipStream_t s);
void pvk_p_scalar2(const pvk_grid *g, pvk_scal *sc, int which, int mode, int maxit, double tol,
                   const pvk_red *rd, int nb, hipStream_t s);
void pvk_dot(const pvk_grid *g, const double *a, const double *b, const pvk_red *rd, hipStream_t s);
void pvk_dot_sc(const pvk_grid *g, const pvk_scal *sc, const double *a, const double *b,
                const pvk_red *rd, hipStream_t s);
void pvk_p_mean(const pvk_grid *g, const double *x, const pvk_red *rd, hipStream_t s);
void pvk_p_mean_finalize(const pvk_grid *g, pvk_scal *sc, const pvk_red *rd, hipStream_t s);
void pvk_p_sub_mean(const pvk_grid *g, const pvk_scal *sc, double *x, int ghosts, hipStream_t s);
void pvk_project(const pvk_grid *g, int stage, const pvk_scal *sc, const double *rho,
                 const double *phi, double *p, double *u, double *v, hipStream_t s);
void pvk_divergence(const pvk_grid *g, const double *u, const double *v, const pvk_red *rd,
                    hipStream_t s);
void pvk_div_finalize(const pvk_grid *g, pvk_scal *sc, const pvk_red *rd, hipStream_t s);

/* fused FFT-PCG (csrc/pv_pcg.hip): start (mean removal, r0, forward rows)
 * and the six per-iteration stages; n = -1 preconditions the start residual */
void pvk_pcg_start(const pvk_grid *g, const pvk_fft *F, pvk_scal *sc, const pvk_pcg *w, double tol,
                   hipStream_t s);
void pvk_pcg_stage(const pvk_grid *g, const pvk_fft *F, pvk_scal *sc, const pvk_pcg *w, int stage, int n,
                   double tol, int maxit, hipStream_t s);
/* clear the solve's stop flag (the stages skip their work once it is set) */
void pvk_pcg_clear_done(pvk_scal *sc, hipStream_t s);

/* sc->red[slot] = sum of the <r, z> partials rz[rzs] */
void pvk_pcg_gamma(pvk_scal *sc, const pvk_pcg *w, int rzs, int slot, hipStream_t s);
void pvk_pcg_prep(const pvk_grid *g, pvk_scal *sc, const pvk_pcg *w, hipStream_t s);
void pvk_pcg_apply(const pvk_grid *g, pvk_scal *sc, const pvk_pcg *w, int n, hipStream_t s);

/* ---- multigrid-preconditioned CG (csrc/pv_mg.hip) ----
 * Level 0 is the pressure grid; levels 1.. are compact (pitch nx).  The
 * V-cycle's arrays are of its type (FP64, or FP32 with fp32 = 1).  FP64:
 * level 0 x = z, f = r, faces = the PCG faces (pitch P).  FP32: level 0 x and
 * faces are FP32 copies (pitch nx), r is read from r0 and z written to z0
 * (pitch p0).  Levels c0.. (<= PVK_MG_LDS_CELLS cells) run inside one
 * workgroup; the last level (halving stops at <= PVK_MG_BOTTOM_MIN cells and
 * <= PVK_MG_BOTTOM_NX columns, or at odd sizes; <= PVK_MG_BOTTOM cells) is
 * solved with the grounded inverse G. */
#define PVK_MG_MAXLEV 16
#define PVK_MG_LDS_CELLS 4096
#define PVK_MG_BOTTOM 256
#define PVK_MG_BOTTOM_MIN 256
#define PVK_MG_BOTTOM_NX 64            /* the bottom's width: the band of its factorisation */
#define PVK_MG_LOW 3                  /* most levels below the first in-LDS level */
typedef struct pvk_mglev {
    int nx, ny, pitch;
    void *x, *f, *bx, *by;            /* result x, right-hand side, faces */
    void *xp;                         /* tiled levels: x after the pre-smoothing, kept apart from x so
                                         that no block of the post kernel reads, as its halo, cells
                                         another block of the same launch is overwriting */
} pvk_mglev;
typedef struct pvk_mg {
    int nlev, c0, nu, fp32;           /* levels, first in-LDS level, sweeps per side, FP32 V-cycle */
    int topf;                         /* faces of level c0 copied to LDS too (when they fit) */
    int wmask;                        /* in-LDS levels (bit l) with two coarse corrections (a W there) */
    int sh0;                          /* level-0 tile shape: 0 = 64 x 16, 1 = 32 x 16, 2 = 32 x 8,
                                         -1 = by width as the other levels */
    pvk_mglev L[PVK_MG_MAXLEV];
    const double *r0;                 /* FP32: level-0 right-hand side (r), pitch p0 */
    double *z0;                       /* FP32: level-0 result (z), pitch p0 */
    int p0;
    double *G;                        /* the grounded inverse G, nb x nb (set-up) */
    double *cm;                       /* nb column means of G (set-up) */
    double *Lb;                       /* the band Cholesky factor of the grounded bottom operator and
                                         1 / its diagonal (set-up) */
    void *H;                          /* the bottom solve's P G P, nb x nb, in the cycle's type */
} pvk_mg;
/* faces of every level; with bottom != 0 also the bottom matrix P G P */
void pvk_mg_setup(const pvk_mg *m, const double *bx0, const double *by0, int pf, int bottom, hipStream_t s);
int  pvk_mg_nrz(const pvk_mg *m);
size_t pvk_mg_lds_bytes(const pvk_mg *m);
void pvk_mg_pre(const pvk_mg *m, int l, pvk_scal *sc, const pvk_pcg *w, int n, double tol, int maxit,
                hipStream_t s);
void pvk_mg_coarse(const pvk_mg *m, const pvk_grid *g, pvk_scal *sc, const pvk_pcg *w, int n, double tol,
                   int maxit, hipStream_t s);
void pvk_mg_post(const pvk_mg *m, int l, pvk_scal *sc, const pvk_pcg *w, int n, hipStream_t s);
void pvk_mg_check(pvk_scal *sc, const pvk_pcg *w, int n, double tol, int maxit, hipStream_t s);
void pvk_mg_take(const pvk_grid *g, const pvk_scal *sc, const double *src, double *phi, hipStream_t s);
void pvk_mg_start(const pvk_grid *g, pvk_scal *sc, const pvk_pcg *w, int warm, hipStream_t s);
void pvk_mg_update(const pvk_grid *g, pvk_scal *sc, const pvk_pcg *w, int n, hipStream_t s);
/* out = A x with the pressure operator of the fused PCG (faces from its start) */
void pvk_pcg_opx(const pvk_grid *g, const pvk_pcg *w, const double *x, double *out, hipStream_t s);

/* learned-solve glue (csrc/pv_learn.hip); sc->red[5..7] = s_r, s_z, mean(e) */
void pvk_learn_stats(const pvk_grid *g, pvk_scal *sc, const double *r, const double *z, const double *a,
                     const double *b, double *part, hipStream_t s);
void pvk_learn_feat(const pvk_grid *g, const pvk_scal *sc, const double *r, const double *z, const double *rho,
                    float *in3, hipStream_t s);


void pvk_learn_target(const pvk_grid *g, const pvk_scal *sc, const double *phistar, const double *phik,
                      const double *r, double sr, float *tgt, hipStream_t s);
void pvk_learn_addr(const pvk_grid *g, const pvk_scal *sc, const double *r, const float *c, double *out,
                    hipStream_t s);
void pvk_learn_dots(const pvk_grid *g, const double *z, const double *r, const double *Ad, double *rzp, double *zadp,
                    hipStream_t s);
void pvk_learn_copy(const pvk_grid *g, const double *src, double *dst, hipStream_t s);
void pvk_learn_axpy(const pvk_grid *g, const double *src, double *dst, hipStream_t s);   /* dst += src */
void pvk_learn_crop(const float *in3, const float *tgt, int NXM, int cs, const int *org, int n, int base,
                    float *ds_in, float *ds_tgt, hipStream_t s);
void pvk_learn_gather(const float *ds_in, const float *ds_tgt, int cs, const int *idx, int B, float *bin,
                      float *btgt, hipStream_t s);

/* learned preconditioner network (csrc/pv_cnn.hip) */
void pvk_cnn_pack(const pvk_cnn *c, hipStream_t s);
void pvk_cnn_forward(const pvk_cnn *c, const float *in3, float *out, int B, int H, int W, int px, hipStream_t s);
void pvk_cnn_backward(const pvk_cnn *c, const float *in3, const float *out, const float *tgt, int B, int H, int W,
                      int px, int border, hipStream_t s);
void pvk_cnn_adam(const pvk_cnn *c, float lr, int t, hipStream_t s);

/* ---- time step ---- */
void pvk_cfl(const pvk_grid *g, const double *u, const double *v, const pvk_red *rd, hipStream_t s);
void pvk_cfl_finalize2(const pvk_grid *g, const pvk_phys *ph, pvk_scal *sc, double dx_min,
                       const pvk_red *rd, hipStream_t s);
void pvk_take_dt(pvk_scal *sc, hipStream_t s);
void pvk_advance_time(pvk_scal *sc, int stage, hipStream_t s);

#ifdef __cplusplus
}
#endif
#endif
