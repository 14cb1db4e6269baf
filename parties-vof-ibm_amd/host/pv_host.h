/*
 * pv_host.h -- private state of one pv_bag (host layer only).
 */
#ifndef PV_HOST_H
#define PV_HOST_H
#include "../../include/parties_vof.h"
#include "../csrc/pv_kernels.h"

/* device-only work arrays after the public fields */
enum pv_internal_field {
    PV_FI_DG_U = PV_NFIELDS,   /* Jacobi diagonal + EPS, u (msolve_cg.c:103-249) */
    PV_FI_DG_V,                /* same, v */
    PV_FI_TMP,                 /* scratch */
    PV_FI_Z,                   /* preconditioned residual of the pressure PCG */
    PV_FI_D2,                  /* second direction buffer of the fused pressure PCG */
    PV_FI_CHP2,                /* second direction buffer of the biharmonic CG */
    PV_FI_DU2,                 /* second direction buffers of the momentum PCG */
    PV_FI_DV2,
    PV_FI_PBX,                 /* pressure operator face coefficients, x and y */
    PV_FI_PBY,
    PV_FI_LTMP,                /* right-hand side of the learned preconditioner's FFT solve */
    PV_FI_PHI0,                /* pressure correction of the last solve of RK3 stage 0, 1, 2 */
    PV_FI_PHI1,
    PV_FI_PHI2,
    PV_FI_MU0,                 /* momentum solve of RK3 stage 0, 1, 2: the last increment of u ... */
    PV_FI_MU1,
    PV_FI_MU2,
    PV_FI_MV0,                 /* ... and of v */
    PV_FI_MV1,
    PV_FI_MV2,
    PV_FI_MUS,                 /* u and v before the momentum solve */
    PV_FI_MVS,
    PV_FI_MCF0,                /* FP32 momentum operator coefficients, two arrays per field: */
    PV_FI_MCF1,                /* u in MCF0..2, v in MCF3..5 (five used of six each) */
    PV_FI_MCF2,
    PV_FI_MCF3,
    PV_FI_MCF4,
    PV_FI_MCF5,
    PV_FI_MZU,                 /* momentum polynomial preconditioner: z of u and v ... */
    PV_FI_MZV,
    PV_FI_MRU2,                /* ... and the residuals of odd iterations */
    PV_FI_MRV2,
    PV_FI_END
};
#define PV_NINTERNAL (PV_FI_END - PV_NFIELDS)
#define PV_TPOOL 256                /* event pairs per timer between syncs */
#define PV_TSAMP 8192               /* launch durations kept per timer */

typedef struct pv_metrics {
    int nx, ny;
    double *xu, *xc, *yv, *yc;
    double *dx_u, *dx_c, *idx_u, *idx_c;
    double *dy_v, *dy_c, *idy_v, *idy_c;
} pv_metrics;

/* learned-solve training collector (host/pv_learn_host.c) */
#define PV_LEARN_MAXSNAP 8
typedef struct pv_learn {
    int collecting, nsnap, snap_iter[PV_LEARN_MAXSNAP];
    int cs, crops_per_snap, ds_cap, ds_n;
    unsigned long long rng;
    double *snapmem, *sphi[PV_LEARN_MAXSNAP], *sz[PV_LEARN_MAXSNAP], *sr[PV_LEARN_MAXSNAP];
    double sig_r[PV_LEARN_MAXSNAP];  /* RMS of r_k at each snapshot */
    float *feat;                   /* [nsnap][NYM][NXM][3] */
    float *tgt, *cout;             /* [NYM][NXM] target / network output */
    double *part;                  /* 3 NYM row partials */
    float *ds_in, *ds_tgt;         /* dataset crops [n][cs][cs][3], [n][cs][cs] */
    int *d_org, *d_idx;
} pv_learn;

struct pv_bag {
    pv_params prm;
    double slab;
    int device;
    int NXM, NYM, NXT, NYT, P;
    size_t alloc_elems;
    double *slabmem;               /* all fields, one allocation */
    double *f[PV_FI_END];          /* cell (0,0) of each field */
    double *metricmem;
    pv_metrics hm;                 /* host copy of the metrics */
    double dx_min, h_idx_u0, h_idy_v0;
    pvk_grid g;
    pvk_phys ph;
    pvk_red rd;
    pvk_scal *d_sc;                /* device scalars */
    pvk_scal *h_sc;                /* pinned host mirror */
    hipStream_t stream;
    /* host view of the time state */
    double time, dt, dt_old;
    int stage, ntime, time_nonzero;
    pv_stats st;
    int ch_chunk, mom_chunk[2], p_chunk;
    /* kernel timers */
    int timers_on;             /* 0 off, else sample period: time every k-th launch per timer */
    long long tseen[PV_NTIMERS];
    int tactive[PV_NTIMERS];
    hipEvent_t tev[PV_NTIMERS][2][PV_TPOOL];
    int tpending[PV_NTIMERS];
    double tms[PV_NTIMERS];
    long long tcount[PV_NTIMERS];
    float *tsamp;                   /* PV_NTIMERS x PV_TSAMP sampled durations (ms) */
    int tnsamp[PV_NTIMERS];
    /* FFT-x / DCT-y preconditioner */
    pvk_fft fft;
    double *fftmem;
    pvk_pcg pcg;
    double *pcgmem;
    pvk_mg mg, mgf;                /* multigrid preconditioner, FP64 and FP32 V-cycle (mg_ok) */
    pvk_mg *mgp;                   /* the one in use */
    pvk_pcg mgw;                   /* CG work set of the MG-PCG: fields + its own partials */
    double *mgmem;
    float *mgfmem;
    int mg_ok, mg_chunk;
    int mg_bottom_age[2];          /* solves since the bottom matrix of mg / mgf was built; -1: not built */
    int p_warm;                    /* 1: start each MG solve from the stage's last correction when better */
    int ch_spectral;               /* 1: biharmonic CG started from the exact spectral solve */
    int have_prev[3];              /* PV_FI_PHI0 + stage holds a correction */
    int m_warm;                    /* 1: start each momentum solve from u + the stage's last increment */
    int m_exact;                   /* 1: momentum CG steps without the EPS = 1e-14 denominators */
    int m_pc;                      /* momentum preconditioner: 0 Jacobi, 1 Neumann polynomial */
    int have_mprev[3];             /* PV_FI_MU0 / MV0 + stage hold an increment ... */
    double mprev_dt[3];            /* ... of a step of this dt */
    int warm_wins[3], warm_solves[3];  /* per stage: the temporal start's consecutive wins, solves since
                                          it became trusted */
    int net_losses, net_skip;      /* consecutive losses of the network's start; solves left to skip it */
    int net_backoff;               /* length of the next skip (16, doubling to 256 while it keeps losing) */
    pvk_ccg ccg;                   /* fused biharmonic CG work set */
    pvk_cnn cnn;                   /* learned preconditioner network (pv_cnn_init) */
    int cnn_ready;
    float *cnn_io;                 /* device staging of batch inputs / targets / outputs */
    size_t cnn_io_cap;
    pvk_mcg mcg;                   /* fused momentum PCG work set */
    double *cgmem;
    pv_learn *learn;               /* collector / learned-solve state, or NULL */
    int use_graphs, graph_chunk;   /* pressure FFT-PCG chunks replayed from hipGraphs */
    hipGraphExec_t *pgraph;        /* one executable graph per chunk index */
    int npgraph;
    hipGraphExec_t *mggraph;       /* the same for MG-PCG iterations */
    int nmggraph;
    const pvk_mg *mggraph_key;     /* hierarchy and sweeps the MG graphs were captured with */
    int mggraph_nu, mggraph_w;     /* sweeps and W levels the MG graphs were captured with */
    int mggraph_sh;                /* and the level-0 tile shape */
    double learn_start_res;        /* relative residual after the last warm start */
};

int  pv_set_error(int code, const char *fmt, ...);
void pv_derive_params(pv_params *p, double *slab);
void pv_host_metrics(const pv_params *p, pv_metrics *m);
void pv_host_metrics_free(pv_metrics *m);
int  pv_sync_scalars_to_device(pv_bag *b);
int  pv_read_scalars(pv_bag *b);
double *pv_dev_field(pv_bag *b, int f);
int  pv_fft_create(pv_bag *b);
int  pv_cg_create(pv_bag *b);
int  pv_mg_create(pv_bag *b);
int  pv_set_mg_precision(pv_bag *b, int fp32);
int  pv_set_ch_solver(pv_bag *b, int spectral);
int  pv_set_mg_tile0(pv_bag *b, int shape);
int  pv_set_momentum_rows(pv_bag *b, int rows);
int  pv_set_momentum_preconditioner(pv_bag *b, int kind);
int  pv_set_fft_tri_shape(pv_bag *b, int rows, int cols);
void pv_mg_destroy(pv_bag *b);
void pv_cnn_free(pv_bag *b);
int  pv_learn_snapshot(pv_bag *b, int si);
int  pv_learn_harvest(pv_bag *b, int nsnap);
int  pv_learn_source(pv_bag *b, double *out);
void pv_learn_free(pv_bag *b);
void pv_graphs_free(pv_bag *b);
size_t pv_field_elems(const pv_bag *b);
size_t pv_field_front(const pv_bag *b);
double *pv_field_origin(const pv_bag *b, double *base);
void pv_tbegin(pv_bag *b, int t);
void pv_tend(pv_bag *b, int t);
int  pv_timers_flush(pv_bag *b);
void pv_fft_destroy(pv_bag *b);

#endif
