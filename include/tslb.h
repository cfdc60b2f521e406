/*
 * tslb.h -- C ABI of the MI355X TSLB library (libtslb.so).
 *
 * The thread-safe recursive-regularized D3Q27 lattice Boltzmann step (TSLB) for
 * one- and two-component flow on one GPU.  The state lives on the device as
 * structure-of-arrays moment fields (rho, u, v, w, Pi^neq (6), phi); one step is
 * a fused moment-space collide + pull-stream.  Host arrays are Fortran-ordered
 * (1:nx,1:ny,1:nz), i fastest, and stay owned by the caller; nothing is
 * retained after a call returns.
 *
 * Conventions: every call returns 0 on success and a negative TSLB_E* code on
 * failure; tslb_last_error() gives the message of the last failure on the
 * calling thread.  A context is not re-entrant: one host thread drives it.
 * Work is queued on the context's own HIP stream (default priority);
 * tslb_sync() waits for it.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the reference repository root).  In the reference, these are no-argument
 * Fortran subroutines on module globals, inside an `!$acc data` region that
 * owns device residency (recursive_2c_TSLB3D.f90:316-317,841).
 */
#ifndef TSLB_H
#define TSLB_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TSLB_OK 0
#define TSLB_E_ARG (-1)      /* invalid argument */
#define TSLB_E_HIP (-2)      /* HIP runtime error */
#define TSLB_E_NOMEM (-3)    /* device allocation failed */
#define TSLB_E_STATE (-4)    /* call not valid in the current state */

#define TSLB_FORCE_ADDITIVE 0 /* f_l += w_l c_l.F / cs^2  (recursive_TSLB3D.f90:571-707) */
#define TSLB_FORCE_SHIFT 1    /* u_eq = u + F            (lb_kernels_module.f90:585-587)  */

/* ncomp value of the colour-gradient model (turbulent_benchmark/couette/
 * recursive_colorTSLB3D.f90:719-959): field TSLB_PHI then holds the colour
 * psi = (rhoA - rhoB)/rho, sigma is the surface tension. */
#define TSLB_COLOUR 3

#define TSLB_STORAGE_FP32 0
/* BF16 and FP16 (rho-1 x64 and Pi^neq x4096 to stay in FP16's normal
 * range): the step stores rho-1, u, v, w and phi with stochastic rounding
 * (unbiased, random bits from a hash of node and step), the rest and all
 * uploads to nearest even. */
#define TSLB_STORAGE_BF16 1
#define TSLB_STORAGE_FP16 2
/* 16-bit signed fixed point per field: k = round(value * 2^m), |k| <= 32767,
 * with a power-of-two step 2^-m per field group (rho-1; u, v, w; the six
 * Pi^neq; phi): defaults 2^-17, 2^-17, 2^-18 and 2^-14, ranges +-1/4, +-1/4,
 * +-1/8 and +-2 (see tslb_set_storage_range); values beyond are clipped and
 * counted (tslb_storage_clipped).  The step rounds rho-1, u, v, w and phi
 * stochastically (k = floor(value * 2^m + r), r uniform in [0, 1) in steps
 * of 2^-16 from a hash of node and step: unbiased), Pi^neq and all uploads
 * to nearest. */
#define TSLB_STORAGE_FX16 3

/* Moment field indices (the order of tslb_upload_moments' packed buffer). */
#define TSLB_RHO 0
#define TSLB_U 1
#define TSLB_V 2
#define TSLB_W 3
#define TSLB_PXX 4
#define TSLB_PYY 5
#define TSLB_PZZ 6
#define TSLB_PXY 7
#define TSLB_PXZ 8
#define TSLB_PYZ 9
#define TSLB_PHI 10
#define TSLB_NFIELDS 11

/* Run parameters.  Replaces the hard-coded parameter block of the drivers
 * (recursive_2c_TSLB3D.f90:20-52, recursive_TSLB3D.f90:20-48). */
typedef struct tslb_params {
  int32_t nx, ny, nz;    /* lattice extent (physical nodes, no ghosts)            */
  int32_t ncomp;         /* 1: single component; 2: two components, phase field with
                            sharpening and CSF (the reference's TSLB 2C path);
                            TSLB_COLOUR: two components, colour gradient with
                            perturbation and recolouring                           */
  int32_t periodic[3];   /* 1: periodic axis; 0: halfway bounce-back wall just
                            outside both faces of that axis                         */
  int32_t force_mode;    /* TSLB_FORCE_*; two-component runs always shift          */
  int32_t storage;       /* TSLB_STORAGE_*: device storage of the moment fields     */
  int32_t device;        /* HIP device ordinal                                      */
  float tau;             /* relaxation time, omega = 1/tau                          */
  float sigma;           /* surface tension (CSF), recursive_2c_TSLB3D.f90:27        */
  float sharp_c;         /* phase-field sharpening coefficient, :28                  */
  float force[3];        /* constant body force fx, fy, fz                           */
  float lin_force;       /* linear forcing coefficient A: u_eq += A u (shift mode)   */
  float tau_b;           /* colour model: relaxation time of component B (<= 0.5:
                            tau); per node omega from the harmonic-mean viscosity   */
  float beta;            /* colour model: recolouring parameter (0..1)               */
} tslb_params;

typedef struct tslb_ctx tslb_ctx;

/* Fill *p with defaults (periodic 16^3, one component, tau 0.8, FP32). */
void tslb_default_params(tslb_params *p);

/* Allocate the device state (two moment buffers, normals, scratch).
 * Replaces the allocate block and the `!$acc data copy(...)` region:
 * recursive_2c_TSLB3D.f90:54-59,316-317. */
int tslb_create(tslb_ctx **out, const tslb_params *p);

/* Free the device state.  Replaces `!$acc end data` (recursive_2c_TSLB3D.f90:841). */
int tslb_destroy(tslb_ctx *ctx);

/* Geometry: isfluid (nx,ny,nz) int32, 1 fluid / 0 solid.  Solid nodes keep
 * their moments; fluid neighbours see halfway bounce-back.  Replaces the
 * geometry block recursive_2c_TSLB3D.f90:79-85 and bcs_all_bback_2c
 * (boundary_cds_3D_module.f90:10-75).  NULL clears the mask (all fluid). */
int tslb_set_isfluid(tslb_ctx *ctx, const int32_t *isfluid);

/* Host -> device moments.  `m` holds TSLB_NFIELDS arrays of nx*ny*nz floats,
 * packed in TSLB_* order.  Replaces `!$acc update device` of the moment arrays
 * (recursive_2c_TSLB3D.f90:316-317). */
int tslb_upload_moments(tslb_ctx *ctx, const float *m);
/* One field (TSLB_*), nx*ny*nz floats. */
int tslb_upload_field(tslb_ctx *ctx, int field, const float *data);

/* Device -> host moments (same layout); used at output steps.  Replaces
 * `!$acc update host(rho,phi,u,v,w)` (recursive_2c_TSLB3D.f90:586,597). */
int tslb_download_moments(tslb_ctx *ctx, float *m);
int tslb_download_field(tslb_ctx *ctx, int field, float *data);

/* Moment-space checkpoint, in place of the population dump and reload
 * dump_distros_1c_3d / read_distros_1c_3d (print_module.f90:700-723).  The
 * file holds a 64-byte header (magic "TSLBCKPT", version 1, nx, ny, nz,
 * ncomp, storage, bytes per value, field count, the caller's step; version 4
 * for the 16-bit formats, whose 4 float group scales (FX16: 2^m) and the
 * library's step count (uint32, the seed of the stochastic rounding) follow
 * the header and are restored on reading; version 1 files of FP16 and BF16
 * storage are read too and continue with the context's step count) and the 11 stored fields verbatim (storage type,
 * field 0 as rho-1), so a run that
 * reads it continues bit for bit as one that never stopped.  The isfluid map
 * and plane BCs are not part of it: the caller sets them again.  Reading
 * checks lattice, component count and storage against the context. */
int tslb_checkpoint_write(tslb_ctx *ctx, const char *path, int64_t step);
int tslb_checkpoint_read(tslb_ctx *ctx, const char *path, int64_t *step);

/* Colour-gradient normals n = grad(phi)/(|grad(phi)|+1e-9) of the current
 * state, 3*nx*ny*nz floats (nx, ny, nz components); for ncomp =
 * TSLB_COLOUR the unnormalised colour gradient grad(psi).  Computed over the
 * whole lattice from the current state when called (the same bits the step
 * uses).  Replaces `!$acc update host(normx,...)`. */
int tslb_download_normals(tslb_ctx *ctx, float *n3);

/* Initial state from the uploaded rho, u, v, w, phi: Pi^neq = 0 (f = feq at
 * every fluid node) and, when prepush_phi != 0, phi advected once by the
 * unsharpened D3Q7 equilibrium, as the reference writes g already pushed.
 * Replaces recursive_2c_TSLB3D.f90:106-287 and initial_conditions
 * (TSLB_2C-mpi/lb_kernels_module.f90:938-1142). */
int tslb_init_equilibrium(tslb_ctx *ctx, int prepush_phi);

/* Advance nsteps full steps: (two-component) colour-gradient normals, then the
 * fused moment-space collide + pull-stream with periodic wrap, halfway
 * bounce-back and forcing, then the moment-space plane BCs if set.  Replaces
 * one pass of the time loop recursive_2c_TSLB3D.f90:344-840 (K1 moments,
 * K2 normals, K3 fused collide+push, bcs_2c_turbulent_jet / pbcs). */
int tslb_step(tslb_ctx *ctx, int nsteps);

/* Split entry points, for bisection against the reference's kernel order. */
/* K2 colour gradient (recursive_2c_TSLB3D.f90:543-577): for the phase field
 * the curvature pass of the step (the fused pass forms the normals itself);
 * for ncomp = TSLB_COLOUR nothing runs (the fused pass forms the colour
 * gradient), so the split sequence is the same for every model. */
int tslb_normals(tslb_ctx *ctx);
/* K3 fused collide + stream followed by K1 moments of the streamed
 * populations, as one moment-space pass (recursive_2c_TSLB3D.f90:607-836 then
 * :346-542 of the next step); uses the normals of the last tslb_normals().
 * The moment-space plane BCs of tslb_set_zplane_bc follow it, as in
 * tslb_step, so tslb_normals + tslb_collide_stream equals one tslb_step. */
int tslb_collide_stream(tslb_ctx *ctx);

/* Moment-space z-plane boundary conditions applied after every step:
 * plane z=1 gets rho, Pi from z=2, u=v=0, w=w_in(x,y); plane z=nz is a
 * zero-gradient copy of z=nz-1.  Replaces bcs_TSLB_only_z_turbojet
 * (TSLB_newBCS/boundary_cds_3D_module.f90:1233-1292).  w_in is nx*ny floats
 * (NULL disables the BC); for a two-component run phi_in sets phi on z=1
 * (NULL keeps the copy from z=2). */
int tslb_set_zplane_bc(tslb_ctx *ctx, const float *w_in, const float *phi_in);

/* Turbulent-jet boundary, the root driver's bcs_2c_turbulent_jet
 * (boundary_cds_3D_module.f90:862-1049): planes k=1 and k=nz are walls
 * (every node solid in tslb_set_isfluid), z not periodic.  Each step, after
 * the collide-stream, the 9 populations crossing from plane 1 into plane 2
 * become the regularized inlet values
 *   feq_l(rho(y), u(x)) + (1-omega) w_l a^neq_l(Pi(y), u(y)),  x = y - c_l,
 * with u(x) the velocity stored on the inlet plane (set it with
 * tslb_upload_plane); the outlet (plane nz) first copies w from plane nz-1,
 * then does the same for the populations crossing into plane nz-1.  Phase
 * field: walls neither reflect nor send g (g_5 into plane 2 is
 * phi(x)(1+3w(x))/8, g_6 into plane nz-1 is phi(y)(1-0.15)/8).  Enable it
 * before tslb_init_equilibrium so that the initial phi push follows the same
 * wall rule.  Exclusive with tslb_set_zplane_bc.  enable = 0 disables it. */
int tslb_set_jet_bc(tslb_ctx *ctx, int enable);

/* Host -> device one z-plane k (0-based, nx*ny floats, i fastest) of a
 * moment field, e.g. the inflow velocity w and phi on the inlet plane.
 * Replaces `!$acc update device(w(...,1))` of the inflow refresh
 * (boundary_cds_3D_module.f90:866-877). */
int tslb_upload_plane(tslb_ctx *ctx, int field, int k, const float *data);

/* FX16 storage: a half range R for the group of `field` (rho-1; u, v, w;
 * the six Pi^neq; phi).  The group takes the finest power-of-two step 2^-m
 * whose range 32767 2^-m covers R; values beyond saturate.  The stored
 * fields of the group are re-encoded.  No counterpart in
 * the reference (which stores FP32, vars_single_3D_module.f90:4): the
 * reduced-precision storage of the moment fields. */
int tslb_set_storage_range(tslb_ctx *ctx, int field, float half_range);

/* FX16 storage: how many stored values were clipped at their range since the
 * context was created (the fused pass counts nodes, uploads count values);
 * 0 for other storage.  Synchronous.  No counterpart in the reference. */
int tslb_storage_clipped(tslb_ctx *ctx, int64_t *count);

/* Seeded noise: field += amp * (2 r - 1) with r = tslb_rand_noseeded(i, j, k, l)
 * at 1-based node coordinates, computed on the device.  Replaces the seeded
 * initial perturbation of initial_conditions
 * (TSLB_2C-mpi/lb_kernels_module.f90:970-977).  Field TSLB_RHO perturbs rho. */
int tslb_seed_noise(tslb_ctx *ctx, int field, float amp, int l);

/* FP64 diagnostics over fluid nodes: out[0] sum rho, [1] sum phi,
 * [2] sum rho|u|^2/2, [3..5] sum rho u, [6] enstrophy sum |curl u|^2/2
 * (central differences).  Synchronous. */
int tslb_reduce(tslb_ctx *ctx, double out[7]);

/* tslb_reduce without allocation, host copy or synchronisation: the same
 * seven FP64 sums of the current state, written to out_dev (7 doubles of
 * DEVICE memory) by two kernels queued on the context's stream; read them
 * after tslb_sync (or an event on that stream).  No counterpart in the
 * reference: for callers that cannot allocate or block, such as code under
 * HIP graph capture and tight serving loops.  The per-block partials live in
 * a fixed device array of the library (1024 blocks x 7 doubles per GPU, held
 * until the process ends); the kernels stride over any lattice, so no
 * lattice exceeds it.  Calls that overlap on one GPU share that array and
 * race: serialise them (one stream, or events). */
int tslb_reduce_noalloc(tslb_ctx *ctx, double *out_dev);

/* Wait for all queued work of the context. */
int tslb_sync(tslb_ctx *ctx);

/* Kernel timing with HIP events on the context stream: when enabled every
 * fused collide-stream launch is bracketed by events.  tslb_timing_read()
 * synchronises and returns the total milliseconds and the launch count of the
 * fused kernel (k=0) and the normals kernel (k=1) since the last reset. */
int tslb_timing_enable(tslb_ctx *ctx, int on);
int tslb_timing_read(tslb_ctx *ctx, int k, double *total_ms, int64_t *launches);

/* Device memory of the current device, bytes (hipMemGetInfo).  Replaces
 * get_memory_gpu (print_module.f90:546-592). */
int tslb_mem_info(size_t *free_bytes, size_t *total_bytes);
/* Host memory: peak resident set (MB) and total RAM (MB).  Replaces get_mem /
 * get_totalram (TSLB_newBCS/get_mem.c:23-39, get_ram.c:49-66). */
int tslb_host_mem_mb(double *peak_rss_mb, double *total_ram_mb);

/* Counter-based RNG, REAL(4) result in [0,1): TSLB_newBCS/vars_single_3D_module.f90:102-165. */
float tslb_rand_noseeded(int i, int j, int k, int l);
/* Box-Muller Gaussian from two draws: vars_single_3D_module.f90:78-100. */
float tslb_gauss_noseeded(int i, int j, int k, int l);

/* Bandwidth calibration: a plain float4 copy of the context's moment buffer
 * into its second buffer (11 fields, read + write), best of 3 timed launches;
 * *gbs = bytes moved / time in GB/s.  No counterpart in the reference; it
 * measures the HBM ceiling the roofline fraction is quoted against.  The
 * state buffer is unchanged. */
int tslb_stream_copy(tslb_ctx *ctx, double *gbs);

/* Bytes of device memory the context holds. */
int tslb_device_bytes(tslb_ctx *ctx, size_t *bytes);

/* Message of the last failure on this thread (always NUL-terminated). */
int tslb_last_error(char *buf, int len);

/* Library version string, e.g. "tslb 0.1 gfx950". */
const char *tslb_version(void);

/* ---------------------------------------------------------------------------
 * Learned subgrid closure.  No counterpart in the reference (whose LES options
 * stop at the fixed Smagorinsky-free TSLB step): a per-node MLP, 10 inputs
 * (strain rate Sxx, Syy, Szz, Sxy, Sxz, Syz from the isotropic D3Q27 gradient
 * of the velocity -- the stencil of the colour gradient,
 * recursive_2c_TSLB3D.f90:555-566 --, grad(phi), phi) -> 64 -> 64 -> 64 -> 7
 * (tanh-form GELU), outputs nu_t = s0 softplus(o0) and six stress corrections
 * dsigma_k = s_{k+1} o_{k+1}.  The collision then relaxes with
 * omega = 1/(tau + 3 nu_t) and adds dsigma to the post-collision
 * non-equilibrium stress.  BF16 matrix-core products, FP32 accumulation.
 *
 * Parameters (TSLB_CLOSURE_NPARAM floats, natural layout, each layer's weight
 * in_features x out_features row-major): W1 (10x64), b1, W2 (64x64), b2, W3,
 * b3, W4 (64x7), b4, in_scale (10; inputs are multiplied by it), out_scale (7).
 *
 * Samples for training: tslb_closure_add_samples box-filters a lattice by
 * factor^3 blocks (rho_bar, Favre velocity, phi_bar) and stores, per coarse
 * node, the 10 inputs of the filtered field and rho_bar with the subgrid
 * stress tau_ab = <rho u_a u_b> - rho_bar u~_a u~_b.  The training loss is
 * mean over samples of sum_k c_k (M_k - tau_k)^2 / t_scale^2 with
 * M = -2 rho_bar nu_t S + dsigma, c = (1,1,1,2,2,2); FP64 loss sum.
 * --------------------------------------------------------------------------- */
#define TSLB_CLOSURE_NPARAM 9496
typedef struct tslb_closure tslb_closure;
/* A model with room for max_samples training samples, on the current device. */
int tslb_closure_create(tslb_closure **out, int64_t max_samples);
int tslb_closure_destroy(tslb_closure *m);
int tslb_closure_set_params(tslb_closure *m, const float *theta, int n);
int tslb_closure_get_params(tslb_closure *m, float *theta, int n);
/* Append the samples of the lattice's current state filtered by factor^3
 * blocks (factor divides nx, ny, nz; ncomp 1 or 2). */
int tslb_closure_add_samples(tslb_closure *m, tslb_ctx *ctx, int factor);
int tslb_closure_num_samples(tslb_closure *m, int64_t *ns);
int tslb_closure_clear_samples(tslb_closure *m);
/* Stored samples, SoA: feat[10][ns], aux[7][ns] (rho_bar, tau[6]). */
int tslb_closure_get_samples(tslb_closure *m, float *feat, float *aux);
int tslb_closure_set_samples(tslb_closure *m, const float *feat, const float *aux, int64_t ns);
/* Outputs (nu_t, dsigma[6]) of every stored sample, SoA out[7][ns]. */
int tslb_closure_forward(tslb_closure *m, float *out);
/* Loss and gradient (TSLB_CLOSURE_NPARAM floats; the scale entries are 0) over
 * the samples idx[0..nb) (idx NULL: all samples); grad may be NULL. */
int tslb_closure_loss_grad(tslb_closure *m, const int32_t *idx, int64_t nb, float t_scale, double *loss,
                           float *grad);
/* steps Adam steps (beta 0.9, 0.999, eps 1e-8) on random batches of `batch`
 * (1 .. 2^20) samples drawn from (seed, step); *last_loss = loss of the last
 * batch.  The Adam state carries over between calls until set_params. */
int tslb_closure_train(tslb_closure *m, int steps, int batch, float lr, float t_scale, uint32_t seed,
                       double *last_loss);
/* Attach a model to a lattice (ncomp 1 or 2, no jet boundary): before every
 * collision the closure fields are recomputed from the current state.  NULL
 * detaches.  The model must outlive the attachment. */
int tslb_set_closure(tslb_ctx *ctx, tslb_closure *m);
/* Fixed closure fields clo[7][n] (nu_t, dsigma) used by every collision; NULL
 * detaches. */
int tslb_set_closure_fields(tslb_ctx *ctx, const float *clo);
/* Closure fields clo[7][n] the next collision uses (recomputed for a model). */
int tslb_download_closure(tslb_ctx *ctx, float *clo);

#ifdef __cplusplus
}
#endif
#endif /* TSLB_H */
