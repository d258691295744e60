/*
 * rtmlib_modC.h -- C ABI of the MI355X (gfx950) elastic RTM library
 * (elastic_amd/librtm_hip.so).
 *
 * Conventions shared by every entry point (the reference's own layout,
 * W/gpu.h:8-14):
 *   - a field (nz,nx) is column-major FP32: index iz + nz*ix, 0-based;
 *   - a history (nz,nx,nt) is iz + nz*ix + nz*nx*it;
 *   - receiver data per component is (nt,nrx): it + nt*ix
 *     (W/rtmlib-mod.f90:76, W/rtm.f90:82-86);
 *   - all pointers are HOST pointers unless the name ends in _dev;
 *   - functions returning int return 0 on success or a negative RTM_E_* code;
 *     rtm_hip_last_error() gives a message for the last failure on the
 *     calling thread.
 */
#ifndef RTMLIB_MODC_H
#define RTMLIB_MODC_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  RTM_OK = 0,
  RTM_E_ARG = -1,       /* bad dimensions, null pointer, index out of range */
  RTM_E_NOMEM = -2,     /* device or pinned-host allocation failed */
  RTM_E_HIP = -3,       /* a HIP runtime call failed */
  RTM_E_NONFINITE = -4, /* non-finite input or a wavefield blew up */
  RTM_E_CFL = -5        /* dt above the stability limit of the 10th-order scheme */
};

/* ------------------------------------------------------------------------
 * Drop-in adjoint. Replaces the reference's
 *   void rtm_op3_c_(int nt,int nz,int nx,int zrec, float*Vx0,...,float*aux_m2m3_c)
 * defined at W/rtmlib-modC.c:29-31, bound from Fortran at
 * W/rtmlib-modF.f90:29-34 (interface rtm_op3_c ... bind(c,name='rtm_op3_c_'),
 * ints by value, arrays dimension(*)), called at W/rtmlib-mod.f90:158.
 * Semantics: for it = nt-3 .. 0 (0-based),
 *   X[it] = X[it] + X[it+2] + Op^T(Y[it+1])       (gather form, no sponge)
 * in place on the five (nz,nx,nt) histories. The five (nz,nx) coefficient
 * arrays are already scaled by dt/dx (W/rtm-mod-adj.f90:111-113). zrec is
 * accepted and unused, as in the reference. Errors are printed to stderr and
 * the call returns, as the reference's void entry point does.
 * ---------------------------------------------------------------------- */
void rtm_op3_c_(int nt, int nz, int nx, int zrec, float *Vx0, float *Vz0, float *sigmaxx0,
                float *sigmazz0, float *sigmaxz0, float *m1_x, float *m1_z, float *aux_m2_c,
                float *aux_m3_c, float *aux_m2m3_c);

/* Same contract as rtm_op3_c_, returning RTM_OK or RTM_E_*. No counterpart
 * in the reference, which returns void (W/rtmlib-modC.c:29); this is the
 * status-returning form callers should bind when they can check errors. */
int rtm_op3_hip(int nt, int nz, int nx, int zrec, float *Vx0, float *Vz0, float *sigmaxx0,
                float *sigmazz0, float *sigmaxz0, const float *m1_x, const float *m1_z,
                const float *aux_m2_c, const float *aux_m3_c, const float *aux_m2m3_c);

/* Allocation-free form of rtm_op3_hip: same arguments, semantics and status
 * codes, bit-for-bit the same results. No counterpart in the reference; it
 * is for callers that cannot allocate device memory (serving loops,
 * harnesses that forbid hipMalloc). It calls no hipMalloc / hipFree: its
 * device memory is one per-GPU library array of 128 MiB, so nz*nx is limited
 * to about 1.67e6 cells (RTM_E_ARG beyond). The histories stay in host
 * memory and pass through a 3-level device ring, one level at a time. Calls
 * that overlap on one GPU share that array and race. */
int rtm_op3_hip_noalloc(int nt, int nz, int nx, int zrec, float *Vx0, float *Vz0, float *sigmaxx0,
                        float *sigmazz0, float *sigmaxz0, const float *m1_x, const float *m1_z,
                        const float *aux_m2_c, const float *aux_m3_c, const float *aux_m2m3_c);

/* ------------------------------------------------------------------------
 * Host mirror of the Fortran operator module rtm_mod_adj
 * (W/rtm-mod-adj.f90). rtm_init_c_ mirrors rtm_init (:43-117): Fortran
 * 1-based xsource, zsource, zrec on the (nz,nx) computational grid, Lame
 * input (m1=rho, m2=lambda, m3=mu; lame must be 1), nsource=250 Ricker of
 * peak frequency fp (W/rtm-mod-adj.f90:100), sponge width abc (K covers
 * abc+5 cells per edge, W/abc-mod.f90:35). dt is taken as given (the
 * reference driver derives it, W/rtm.f90:66). free_surface is accepted and
 * unused as in the reference. rtm_op_c_ mirrors rtm_op (:120-129):
 *   adj=1: RTM. Vx..sigmaxz are the (nt,nx) receiver data on row zrec; the
 *          forward source wavefield, the adjoint of the data and the Lame
 *          imaging condition (W/image-mod-rtm.f90:146-162) run on the GPU;
 *          dm1=drho, dm2=dlambda, dm3=dmu (nz,nx) are written.
 *   adj=0: Born modelling (:273-375): dm1..dm3 are read and Vx..sigmaxz are
 *          written.
 * The handle owns its device memory until rtm_final_c_.
 * ---------------------------------------------------------------------- */
typedef struct rtm_ctx rtm_ctx;

int rtm_init_c_(rtm_ctx **ctx, int nz, float dz, int nx, float dx, float dt, int xsource,
                int zsource, int zrec, float fp, int nt, int abc, int free_surface,
                const float *m1, const float *m2, const float *m3, int lame);
int rtm_op_c_(rtm_ctx *ctx, int adj, float *dm1, float *dm2, float *dm3, float *Vx, float *Vz,
              float *sigmaxx, float *sigmazz, float *sigmaxz);
int rtm_final_c_(rtm_ctx *ctx);

/* Options of an rtm_ctx (no counterpart in the reference, whose choices are
 * compile-time): */
enum {
  RTM_OPT_ADJ_ABC = 1,    /* 0: adjoint without sponge (W/ drop-in, default);
                             1: exact adjoint of the sponged forward;
                             2: the R/Src adjoint, K applied in place to the
                                stored levels it+1 and it+2 before every step
                                (R/Src/rtm-mod-adj.f90:418-430) */
  RTM_OPT_SNAP_EVERY = 2, /* image every j-th level, weight j (default 1) */
  RTM_OPT_SNAP_PREC = 3,  /* source snapshots: 0 fp32 (default), 1 bf16, 2 fp16 */
  RTM_OPT_COMP_IMAGE = 4, /* 1: compensated (Neumaier) FP32 image sums */
  RTM_OPT_REC_X0 = 5,     /* first receiver column (0-based, default 0) */
  RTM_OPT_REC_NX = 6      /* receiver count (default nx) */
};
int rtm_set_option_c_(rtm_ctx *ctx, int key, int value);

/* ------------------------------------------------------------------------
 * Lower-level single-shot entry points on an explicit computational grid
 * (0-based cells). They take the dt/dx-scaled coefficient arrays of
 * rtm_op3_c_ and let callers (the Python harness, batch drivers) place the
 * grid inside a padded domain.
 * ---------------------------------------------------------------------- */
typedef struct {
  int nz, nx, nt;     /* computational grid and number of time levels */
  float dx, dt;
  int xs, zs;         /* explosive source cell */
  int zr, rx0, nrx;   /* receiver row, first column, count */
  int abc;            /* sponge width (<=0: none) */
  int adj_abc;        /* 0: adjoint without sponge, 1: exact adjoint sponge, 2: R/Src sponge */
  int snap_every;     /* imaging decimation j >= 1 */
  int snap_prec;      /* 0 fp32, 1 bf16, 2 fp16 (per-snapshot max-abs scale) */
  int comp_image;     /* compensated image accumulation */
} rtm_shot_t;

/* Forward source propagation with sponge and explosive Ricker source
 * (W/rtm-mod-adj.f90:178-260) recording all five components on row zr
 * (W/rtm-mod-adj.f90:361-375). wavelet[it-1] = source(it). rec: 5*(nt,nrx),
 * component order Vx, Vz, sxx, szz, sxz. last: optional 10*(nz,nx) = the
 * final two levels (level nt-1 then nt), fields in the same order. */
int rtm_forward_hip(const rtm_shot_t *s, const float *coef5, const float *wavelet, float *rec,
                    float *last);

/* Wavefield movie of one forward run (the reference's snap_wfld,
 * W/snap-mod-rtm.f90:154-283): frame k = 1..nsnap is level k*dsnap with
 * dsnap = nt / nsnap, the five fields of the window [z0, z0+wnz) x
 * [x0, x0+wnx) (0-based cells). movie is [nsnap][5][wnx][wnz] (each field
 * block in the reference layout); *dsnap receives the frame spacing. */
int rtm_forward_movie_hip(const rtm_shot_t *s, const float *coef5, const float *wavelet, int nsnap,
                          int z0, int x0, int wnz, int wnx, float *movie, int *dsnap);

/* Full RTM of one shot (forward + adjoint + Lame imaging). coef5 holds the
 * five (nz,nx) arrays m1_x, m1_z, aux_m2, aux_m3, aux_m2m3 back to back.
 * rec: 5*(nt,nrx) observed data. img: 3*(nz,nx) = drho, dlambda, dmu. */
int rtm_migrate_hip(const rtm_shot_t *s, const float *coef5, const float *wavelet,
                    const float *rec, float *img);

/* Born modelling of one shot (W/rtm-mod-adj.f90:273-375): dm 3*(nz,nx)
 * (drho, dlambda, dmu) -> rec 5*(nt,nrx). */
int rtm_born_hip(const rtm_shot_t *s, const float *coef5, const float *wavelet,
                 const float *dm, float *rec);

/* Staggered-grid coefficients from rho, lambda, mu (W/model-mod.f90:21-89,
 * W/rtmlib-mod.f90:51-58, W/rtm-mod-adj.f90:111-113): writes coef5. */
int rtm_model_op(int nz, int nx, float dx, float dt, const float *rho, const float *lam,
                 const float *mu, float *coef5);

/* Ricker wavelet of W/source-mod.f90:31-44 (nsource samples, peak at
 * nsource/2): out[it-1] = source(it), it=1..nt. */
int rtm_ricker(int nt, float dt, float fp, int nsource, float *out);

/* Sponge K of W/abc-mod.f90:59-78 applied in place to one (nz,nx) field on
 * the GPU (test and tooling entry point). */
int abc_k_hip(int nz, int nx, int abc, float *field);

/* Stability check in FP64: returns RTM_OK when vmax*dt/dx is below the
 * 10th-order staggered limit 1/(sqrt(2)*sum|c_k|), else RTM_E_CFL. */
int rtm_cfl_check(float vmax, float dt, float dx, double *courant, double *limit);

/* ------------------------------------------------------------------------
 * Batched multi-shot RTM (no counterpart in the reference, which runs one
 * shot per process). Every shot has its own computational grid (uneven
 * aperture windows allowed); all inputs are made resident in HBM by
 * rtm_batch_create, shots are processed in groups sized to the free device
 * memory, and each time level of a group is one kernel launch over all of its
 * shots (optionally replayed from HIP graphs). Images of each shot are kept
 * and the region [img_z0, img_z0+img_nz) x [img_x0, img_x0+img_nx) of each is
 * added (in FP64) into a stacked image at (stack_z0, stack_x0).
 * ---------------------------------------------------------------------- */
typedef struct rtm_batch rtm_batch;

typedef struct {
  int nz, nx;                 /* computational grid of this shot */
  int xs, zs, zr, rx0, nrx;   /* source cell, receiver row and columns */
  int img_z0, img_x0, img_nz, img_nx;  /* region stacked into the image */
  int stack_z0, stack_x0;
} rtm_batch_shot_t;

typedef struct {
  int nt;
  float dx, dt;
  int abc, adj_abc;   /* adj_abc: 0 none, 1 exact adjoint sponge, 2 R/Src sponge */
  int snap_every;   /* imaging decimation j */
  int snap_prec;    /* 0 fp32, 1 bf16, 2 fp16 */
  int comp_image;   /* compensated image sums */
  int txr;          /* tile width 1/2/4/8 (0 = autotuned, or RTM_TXR when set) */
  int use_graph;    /* capture time loops in HIP graphs */
  int max_group;    /* cap on shots per group (0 = memory bound) */
  int stack_nz, stack_nx;  /* stacked image size (0 = none) */
  double mem_fraction;     /* share of free device memory for working sets (0 = 0.85) */
} rtm_batch_cfg_t;

int rtm_batch_create(rtm_batch **b, const rtm_batch_cfg_t *cfg, int nshots,
                     const rtm_batch_shot_t *shots, const float *const *coef5,
                     const float *wavelet);
/* Second coefficient set per shot, used by rtm_batch_synthesize. */
int rtm_batch_set_true_model(rtm_batch *b, const float *const *coef5_true);
/* Observed data of every shot on the GPU: forward(true) - forward(migration). */
int rtm_batch_synthesize(rtm_batch *b);
int rtm_batch_set_data(rtm_batch *b, int shot, const float *rec);  /* 5*(nt,nrx) */
int rtm_batch_get_data(rtm_batch *b, int shot, float *rec);
/* RTM of every shot (device-resident inputs; the timed work). */
int rtm_batch_migrate(rtm_batch *b);
int rtm_batch_get_image(rtm_batch *b, int shot, float *img3);  /* 3*(nz,nx) */
int rtm_batch_get_stack(rtm_batch *b, float *img3);            /* 3*(stack_nz,stack_nx) */
/* [0] fwd ms [1] fwd launches [2] adj ms [3] adj launches [4] total ms
 * [5] fwd cell-updates [6] adj cell-updates [7] shots per group */
int rtm_batch_info(rtm_batch *b, double *out8);
int rtm_batch_destroy(rtm_batch *b);
/* Tile width and kernel form the batch runs with (form 0: one level per
 * launch on tiles, 1: two levels per launch, 2: one level per launch,
 * strips marching along x, 3: the marching forward with the tile adjoint). */
int rtm_batch_txr(rtm_batch *b);
int rtm_batch_tblock(rtm_batch *b);
/* Strip width (columns) of the first group's marching launches (the
 * adjoint's for form 2, the forward's for form 3). A group's
 * width, per kernel, is the narrowest even width from 16 to 64 whose strips
 * fit one round of that kernel's resident workgroups, else the autotuner's
 * choice among 16, 24, 28 and 32; RTM_MARCH_W fixes one width for all
 * groups, RTM_MARCH_FIT=0 turns the fitting off. */
int rtm_batch_march_w(rtm_batch *b);
/* Tile autotuner: runs a short copy of the workload (the shots of the full
 * run's first group, at most 24 or RTM_TUNE_MAXSHOTS; max(48, 8*snap_every)
 * levels; same options) with tile widths 1, 2, 4, 8 at one level per
 * launch, 1, 2, 4 at two levels per launch (temporal blocking) and the
 * marching forms (strips of 16, 24, 28 and 32 columns), and returns the fastest in
 * *txr, *tblock;
 * ms8[tblock*4 + log2(txr)] receives the times of the first seven and ms8[7]
 * that of the faster marching form, 2 or 3 (ms, -1 for candidates not run).
 * The result
 * is cached per process and, when RTM_TUNE_CACHE names a file, in it;
 * rtm_batch_create with cfg->txr == 0 uses the cached or tuned choice unless
 * RTM_TXR / RTM_TBLOCK are set. */
int rtm_batch_autotune(const rtm_batch_cfg_t *cfg, int nshots, const rtm_batch_shot_t *shots,
                       const float *const *coef5, const float *wavelet, int *txr, int *tblock,
                       double *ms8);

/* Streaming kernel over nbytes, timed over reps launches with HIP events:
 * achieved GB/s (bytes read + written) and ms per launch. mode 0: copy,
 * 1: copy with non-temporal loads/stores, 2: read only, 3: write only. Used to
 * measure the attainable HBM bandwidth and to calibrate the FETCH_SIZE /
 * WRITE_SIZE counters against a kernel of known traffic. */
int rtm_bandwidth_probe(size_t nbytes, int reps, int mode, double *gbps, double *ms_per_copy);

/* Library lifecycle: optional explicit device selection and release of the
 * cached device state of the calling process. */
int rtm_hip_init(int device);
int rtm_hip_finalize(void);
const char *rtm_hip_last_error(void);

/* Per-phase GPU timings (ms, measured with HIP events on the launch stream)
 * and launch counts of the last rtm_migrate_hip / batch run on this thread:
 * [0] forward-step kernel ms, [1] forward launches, [2] adjoint-step kernel
 * ms, [3] adjoint launches, [4] whole-run ms. */
int rtm_last_timings(double *out5);

#ifdef __cplusplus
}
#endif
#endif /* RTMLIB_MODC_H */
