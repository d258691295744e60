/*
 * learned_abc.h -- learned absorbing layer (part of librtm_hip.so).
 *
 * No counterpart in the reference: its only boundary treatment is the
 * multiplicative sponge K (W/abc-mod.f90:59-78, 40 points). The learned layer
 * replaces it by an 8-cell band inside the 5-cell rim of the computational
 * grid. For every band cell a small network maps local wavefield values to
 * an additive correction of the five fields of the new level:
 *   X(L)[cell] = plain stencil update (no sponge) + s * NN(features / s)
 *
 * Features of a band cell (LABC_NFEAT = 245, padded to LABC_KPAD = 248):
 *   0..224   the 5 fields of level L-1 on a 9 x 5 patch: 9 cells from the
 *            cell inwards along the edge normal (slot kn) times 5 tangential
 *            offsets m = -2..2, feature (field*5 + m+2)*9 + kn (the inward
 *            steps of one field and offset are consecutive features);
 *            stresses are divided by LABC_Z0 (an impedance scale,
 *            Pa / (m/s));
 *   225..229 the cell's own 5 fields of level L-2 (same scaling);
 *   230..233 one-hot edge (top, bottom, left, right);
 *   234..241 one-hot depth in the band;
 *   242..244 the cell's material from the dt/dx-scaled coefficients:
 *            sqrt(m1_x*aux_m2m3) = vp dt/dx, sqrt(m1_x*aux_m3) = vs dt/dx,
 *            sqrt(aux_m2m3/m1_x)/LABC_Z0 = rho vp / LABC_Z0.
 * Wavefield features (0..229) are divided by the cell's scale s = max-abs
 * over them, so the correction is equivariant to the wavefield amplitude.
 * Network: 245 -> 32 -> 32 -> 5, exact GELU, BF16 weights and BF16-rounded
 * features with FP32 accumulation (9093 parameters).
 * Band cells: top rows [R, R+W) and bottom rows [nz-R-W, nz-R) over columns
 * [R, nx-R) (corners belong to them), left columns [R, R+W) and right
 * columns [nx-R-W, nx-R) over rows [R+W, nz-R-W); R = 5, W = 8.
 */
#ifndef LEARNED_ABC_H
#define LEARNED_ABC_H

#include "rtmlib_modC.h"

#ifdef __cplusplus
extern "C" {
#endif

#define LABC_RIM 5
#define LABC_WIDTH 8
#define LABC_NFEAT 245
#define LABC_KPAD 248
#define LABC_HID 32
#define LABC_NOUT 5
#define LABC_NPARAM (LABC_NFEAT * LABC_HID + LABC_HID + LABC_HID * LABC_HID + LABC_HID + \
                     LABC_HID * LABC_NOUT + LABC_NOUT)
#define LABC_Z0 6.0e6f

typedef struct labc_net labc_net;

/* Parameters in order W1 (245x32, row-major [k][j]), b1 (32), W2 (32x32),
 * b2 (32), W3 (32x5), b3 (5). */
int labc_net_create(labc_net **net, unsigned seed);
int labc_net_destroy(labc_net *net);
int labc_net_get_params(labc_net *net, float *params);
int labc_net_set_params(labc_net *net, const float *params);
/* Number of band cells of an nz x nx grid. */
int labc_cell_count(int nz, int nx);
/* Band cell c -> (iz, ix, edge 0..3, depth 0..7). */
int labc_cell_coords(int nz, int nx, int c, int *iz, int *ix, int *edge, int *depth);
/* Evaluation of the network on a feature matrix feat [n][LABC_KPAD] (the
 * normalised features): out [n][5]. mode must be 0 (the training forward,
 * 32 samples per block); propagation uses its own inference kernel. */
int labc_net_forward(labc_net *net, const float *feat, int n, float *out, int mode);
/* Loss 0.5*mean((out - target)^2) over n*5 values and its gradient [LABC_NPARAM]. */
int labc_net_grad(labc_net *net, const float *feat, const float *target, int n, float *grad,
                  float *loss);
/* Features of the band cells of one grid from two levels and the grid's
 * coefficients (host arrays, each 5 fields (nz,nx) back to back):
 * feat [ncell][LABC_KPAD], scale [ncell]. */
int labc_features(int nz, int nx, const float *lev_m1, const float *lev_m2, const float *coef5,
                  float *feat, float *scale);
/* The same for an adjoint state: levels L+1 and L+2 in place of L-1 and
 * L-2, the velocities (not the stresses) divided by the impedance scale. */
int labc_features_adj(int nz, int nx, const float *lev_p1, const float *lev_p2, const float *coef5,
                      float *feat, float *scale);

/* Training set on the device. */
typedef struct labc_data labc_data;
int labc_data_create(labc_data **d, int capacity);
int labc_data_destroy(labc_data *d);
int labc_data_size(labc_data *d);
/* Copy rows [row0, row0+n) of the set to the host: feat [n][LABC_KPAD],
 * tgt [n][5] (either may be NULL). */
int labc_data_get(labc_data *d, int row0, int n, float *feat, float *tgt);
/* Teacher-forced pairs from one wide-sponge run: the wide grid (pad_wide
 * cells of padding, sponge abc_wide) and the band grid (pad = R + W) of the
 * same physical model share the physical region. At each listed level L the
 * band grid's levels L-1, L-2 are the windowed wide levels (rim zeroed); the
 * target is wide(L) - plain band-grid update(L) on the band cells, divided by
 * the cell scale. coef5_wide / coef5_small: dt/dx-scaled coefficients of the
 * two grids; wavelet[it-1] = source(it). noise > 0 multiplies the band cells
 * of the two windowed levels by 1 + noise*u (u uniform in [-1,1), hashed
 * from seed, level, cell and field) before the features and the plain update
 * are formed, so the network learns to pull perturbed states back.
 * abc_band > 0 (<= LABC_WIDTH) trains the layer to work with a sponge of that
 * width in the band: the band grid's update applies it, and the wide grid
 * applies the same factors at the band's position (continued outwards), so
 * the target is the exterior contribution of a damped wavefield. The levels
 * must lie in [2, nt] (level 2 gives no pair), each listed once
 * (RTM_E_ARG otherwise). */
int labc_collect(labc_data *d, int nz_phys, int nx_phys, int pad_wide, int abc_wide, int nt,
                 float dt, int zs_phys, int xs_phys, const float *coef5_wide,
                 const float *coef5_small, const float *wavelet, const int *levels, int nlev,
                 float noise, unsigned seed, int abc_band);
/* Teacher-forced pairs of the adjoint (receiver-side) propagation, for a
 * layer in the adjoint runs. The wide grid records receivers at depth
 * zr_phys over every physical column in a forward run (sponges as in
 * labc_collect), then propagates them back with the exact adjoint of its
 * sponge. At each listed level L the band grid's levels L+1, L+2 are the
 * windowed wide levels (rim zeroed, noise as in labc_collect); the features
 * are those of labc_features_adj; the target is wide(L) - the band grid's
 * adjoint update(L) (the same injected data, the exact adjoint of the band
 * sponge when abc_band > 0) on the band cells, divided by the cell scale
 * (velocities also by the impedance scale). The levels must lie in [1, nt],
 * each listed once (RTM_E_ARG otherwise). coef5_wide_obs (NULL: none)
 * makes the injected data those of a migration: the wide run of that
 * (observed) model minus the wide run of coef5_wide, the model of the
 * adjoint runs. */
int labc_collect_adj(labc_data *d, int nz_phys, int nx_phys, int pad_wide, int abc_wide, int nt,
                     float dt, int zs_phys, int xs_phys, int zr_phys, const float *coef5_wide,
                     const float *coef5_small, const float *wavelet, const int *levels, int nlev,
                     float noise, unsigned seed, int abc_band, const float *coef5_wide_obs);
/* Adam on random minibatches: loss per step in losses[steps]. */
int labc_train(labc_net *net, labc_data *d, int steps, int batch, float lr, unsigned seed,
               float *losses);
/* Mean loss over the whole data set. */
int labc_eval(labc_net *net, labc_data *d, float *loss);

/* One forward propagation of a band grid with the learned layer (net) and
 * the sponge s->abc (0 = none; both may act together), or, with
 * net = NULL, with the sponge alone: records receivers like rtm_forward_hip
 * and returns the last two levels. With a network the grid must be at least
 * 2 * (LABC_RIM + LABC_WIDTH) + 1 cells on each side (RTM_E_ARG otherwise). */
int labc_forward(labc_net *net, const rtm_shot_t *s, const float *coef5, const float *wavelet,
                 float *rec, float *last);
/* Device time of the last labc_forward time loop (ms). */
double labc_last_forward_ms(void);

/* The learned layer in a batch (no counterpart in the reference): with a
 * network, every forward run of the batch (data synthesis, the source
 * wavefield of the imaging) corrects the band cells of each level L >= 3
 * after its plain update and clears the 5-cell rim, in one launch per
 * level over the group's shots, the network's first layer on BF16 matrix
 * cores. The shots' grids should be padded by LABC_RIM + LABC_WIDTH cells
 * and the batch's sponge width (cfg abc) be the band sponge the network was
 * trained with; the adjoint keeps the batch's adjoint sponge option (and
 * takes a network of its own through rtm_batch_set_learned_adjoint). NULL
 * drops the layer. The network must outlive the batch's runs. Every grid
 * must be at least 2 * (LABC_RIM + LABC_WIDTH) + 1 cells on each side
 * (RTM_E_ARG otherwise; after an error the batch has no learned layer in
 * either direction). */
int rtm_batch_set_learned(rtm_batch *b, labc_net *net);
/* The learned layer in the batch's adjoint runs (no counterpart in the
 * reference): a network trained on adjoint states (labc_collect_adj)
 * corrects the band cells of every adjoint level after its step and clears
 * the rim, one level per launch. The imaging of a level happens in its step:
 * only the images of band cells (outside the physical region) miss the
 * correction. The batch's adjoint sponge should be the exact adjoint
 * (adj_abc 1) of the band sponge the network was trained with; the R/Src
 * adjoint sponge (adj_abc 2) is refused with RTM_E_ARG. NULL drops the
 * layer; grid sizes as for rtm_batch_set_learned. */
int rtm_batch_set_learned_adjoint(rtm_batch *b, labc_net *net);

#ifdef __cplusplus
}
#endif
#endif /* LEARNED_ABC_H */
