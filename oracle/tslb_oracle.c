/*
 * tslb_oracle.c -- CPU oracle for the thread-safe recursive-regularized D3Q27
 * lattice Boltzmann step (TSLB), one and two components.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is compiled into oracle/liboracle_tslb.so
 * and is loaded by the test-suite and by the CPU-baseline leg of bench.py as the
 * checker.  The product (the HIP library) never links, loads or calls it.
 *
 * It restates the reference's population-push algorithm (reference paths are
 * relative to the reference repository root):
 *   - lattice tables           TSLB_newBCS/vars_single_3D_module.f90:21-43
 *   - equilibrium feq_l        recursive_2c_TSLB3D.f90:378-536, :639-806
 *                              (27 closed-form polynomials; restated below in
 *                              Hermite form, identical as polynomials)
 *   - K1 moments (+phi)        recursive_2c_TSLB3D.f90:346-542
 *   - K2 colour gradient       recursive_2c_TSLB3D.f90:555-571
 *   - K3 collide + push, CSF   recursive_2c_TSLB3D.f90:607-836
 *     1C additive forcing      recursive_TSLB3D.f90:571-707
 *     1C/2C velocity-shift     TSLB_2C-mpi/lb_kernels_module.f90:585-587,769-771
 *   - bounce-back on solids    boundary_cds_3D_module.f90:10-75
 *   - jet inlet / outlet       boundary_cds_3D_module.f90:862-1049
 *   - colour-gradient model    turbulent_benchmark/couette/recursive_colorTSLB3D.f90:
 *                              719-959 (restated with its defects corrected)
 *   - periodic ghost wrap      TSLB_newBCS/boundary_cds_3D_module.f90:1294-1877
 *   - moment-space z-plane BC  TSLB_newBCS/boundary_cds_3D_module.f90:1233-1292
 *                              (bcs_TSLB_only_z_turbojet; driver order
 *                              TSLB_newBCS/recursive_TSbcs_TSLB3.f90:362-771)
 *   - counter-based RNG        TSLB_newBCS/vars_single_3D_module.f90:78-165
 *   - initial state (f=feq, g pre-pushed)  recursive_2c_TSLB3D.f90:106-287
 * and adds a moment-space pull restatement of the same step (orc_pull_step_*)
 * used to pin the equivalence the GPU design relies on.
 *
 * The file is compiled twice: -DREAL=float -DSFX=f and -DREAL=double -DSFX=d.
 *
 * Storage conventions (shared with the product's C-ABI): every field is a
 * Fortran-ordered (1:nx,1:ny,1:nz) array, i fastest.  Populations live in a
 * ghosted array (0:nx+1,0:ny+1,0:nz+1) per direction, as in the reference.
 *
 * Boundary handling.  For each interior fluid destination y and direction l,
 * the source is s = y - c_l.  Periodic axes wrap s.  If s leaves the domain
 * along a non-periodic axis, or s is a solid node (isfluid == 0), the
 * population is the halfway bounce-back of the one y itself sent towards s:
 * f_l(y) = f*_opp(l)(y).  This is the combined effect of bcs_all_bback_2c and
 * pbcs in the reference.  Colour-gradient stencils read phi at the nearest
 * interior node across a non-periodic face, and a zero normal on solid nodes
 * (the reference never updates normals there).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef REAL
#define REAL double
#define SFX d
#endif

#define CAT2(a, b) a##_##b
#define CAT(a, b) CAT2(a, b)
#define FN(n) CAT(n, SFX)

typedef REAL real;

/* D3Q27 ordering, TSLB_newBCS/vars_single_3D_module.f90:23-30 */
static const int CX[27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 0,
                           0, 1, -1, -1, 1, 1, -1, 1, -1, -1, 1, 1, -1};
static const int CY[27] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 1, -1, 1,
                           -1, 0, 0, 0, 0, 1, -1, -1, 1, -1, 1, -1, 1};
static const int CZ[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1,
                           1, 1, -1, 1, -1, 1, -1, 1, -1, 1, -1, -1, 1};
static const int OPP[27] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14,
                            13, 16, 15, 18, 17, 20, 19, 22, 21, 24, 23, 26, 25};

static real wgt(int l) {
  /* p0..p3, vars_single_3D_module.f90:31-36 */
  if (l == 0) return (real)(8.0 / 27.0);
  if (l <= 6) return (real)(2.0 / 27.0);
  if (l <= 18) return (real)(1.0 / 54.0);
  return (real)(1.0 / 216.0);
}

enum { F_RHO = 0, F_U, F_V, F_W, F_PXX, F_PYY, F_PZZ, F_PXY, F_PXZ, F_PYZ, F_PHI, NFIELD };

typedef struct {
  int nx, ny, nz, ncomp, per[3], force_mode;
  int jet;            /* turbulent-jet inlet/outlet planes (orc_set_jet) */
  real tau_b, beta;   /* colour-gradient model (ncomp 3): tau of B, recolouring */
  real *fa;           /* colour model: component-A populations [27][ng] */
  real *apops;        /* colour model pull scratch: [27][n] */
  real omega, sigma, sharp_c, fx, fy, fz, lin_a;
  size_t n, ng;
  real *f, *g;        /* [27][ng], [7][ng] */
  real *m[NFIELD];    /* moments, [n] each */
  real *nrm[3];       /* colour-gradient normals */
  real *fst[3];       /* CSF velocity shift sigma*kappa*n (pull form scratch) */
  int32_t *isf;       /* 1 fluid, 0 solid */
  real *pops;         /* pull scratch: [27][n] */
  real *gpops;        /* pull scratch: [7][n] */
  real *clo;          /* learned closure fields nu_t, dsigma[6]: [7][n] (NULL: none) */
  real *zw_in;        /* moment-space z-plane BC: inflow w of plane k = 0 (NULL: off) */
  real *zphi_in;      /* and its phi (NULL: phi copied from k = 1) */
} ctx_t;

/* ---------------------------------------------------------------------- */
/* Equilibrium and non-equilibrium Hermite terms                           */
/* ---------------------------------------------------------------------- */

/* feq_l(rho,u): 3rd-order Hermite equilibrium on D3Q27.  Identical, as a
 * polynomial, to recursive_2c_TSLB3D.f90:378-536 (e.g. l=1 at :379):
 * w_l rho [1 + 3 cu + 9/2 cu^2 - 3/2 u^2 + 9/2 cu (cu^2 - u^2)]. */
static real feq_l(int l, real rho, real u, real v, real w) {
  real cu = (real)CX[l] * u + (real)CY[l] * v + (real)CZ[l] * w;
  real uu = u * u + v * v + w * w;
  real poly = (real)1 + (real)3 * cu + (real)4.5 * cu * cu - (real)1.5 * uu +
              (real)4.5 * cu * (cu * cu - uu);
  return wgt(l) * rho * poly;
}

/* a^neq_l(Pi,u) without the weight: 2nd-order Hermite projection of Pi^neq
 * plus the recursive 3rd-order term a3 = sym(u (x) Pi).  Identical, as a
 * polynomial, to fneq1 at recursive_2c_TSLB3D.f90:640-806 (l=1 at :647). */
static real aneq_l(int l, const real *P, real u, real v, real w) {
  /* P = {pxx, pyy, pzz, pxy, pxz, pyz} */
  real cx = (real)CX[l], cy = (real)CY[l], cz = (real)CZ[l];
  real cPc = cx * cx * P[0] + cy * cy * P[1] + cz * cz * P[2] +
             (real)2 * (cx * cy * P[3] + cx * cz * P[4] + cy * cz * P[5]);
  real tr = P[0] + P[1] + P[2];
  real cu = cx * u + cy * v + cz * w;
  real uPc = u * (P[0] * cx + P[3] * cy + P[4] * cz) +
             v * (P[3] * cx + P[1] * cy + P[5] * cz) +
             w * (P[4] * cx + P[5] * cy + P[2] * cz);
  return (real)4.5 * cPc - (real)1.5 * tr +
         (real)13.5 * (cu * cPc - cu * tr / (real)3 - (real)2 * uPc / (real)3);
}

void FN(orc_feq)(real rho, real u, real v, real w, real *out) {
  for (int l = 0; l < 27; l++) out[l] = feq_l(l, rho, u, v, w);
}

void FN(orc_aneq)(const real *P6, real u, real v, real w, real *out) {
  for (int l = 0; l < 27; l++) out[l] = aneq_l(l, P6, u, v, w);
}

/* ---------------------------------------------------------------------- */
/* Indexing                                                                */
/* ---------------------------------------------------------------------- */

static inline size_t idx(const ctx_t *c, int i, int j, int k) {
  return (size_t)i + (size_t)c->nx * ((size_t)j + (size_t)c->ny * (size_t)k);
}
/* ghosted index, i in [-1, nx] */
static inline size_t gidx(const ctx_t *c, int i, int j, int k) {
  return (size_t)(i + 1) +
         (size_t)(c->nx + 2) * ((size_t)(j + 1) + (size_t)(c->ny + 2) * (size_t)(k + 1));
}
static inline int wrap1(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }

/* wrap s along periodic axes; return 0 if s leaves a non-periodic axis */
static inline int wrap_src(const ctx_t *c, int *i, int *j, int *k) {
  int d[3] = {c->nx, c->ny, c->nz};
  int *p[3] = {i, j, k};
  for (int a = 0; a < 3; a++) {
    if (*p[a] < 0 || *p[a] >= d[a]) {
      if (!c->per[a]) return 0;
      *p[a] = wrap1(*p[a], d[a]);
    }
  }
  return 1;
}

/* phi read for stencils: wrap periodic axes, clamp non-periodic ones */
static inline real phi_at(const ctx_t *c, int i, int j, int k) {
  int d[3] = {c->nx, c->ny, c->nz};
  int p[3] = {i, j, k};
  for (int a = 0; a < 3; a++) {
    if (p[a] < 0 || p[a] >= d[a]) {
      if (c->per[a]) p[a] = wrap1(p[a], d[a]);
      else p[a] = p[a] < 0 ? 0 : d[a] - 1;
    }
  }
  return c->m[F_PHI][idx(c, p[0], p[1], p[2])];
}

/* normal read for stencils: wrap periodic axes, zero outside / on solids */
static inline real nrm_at(const ctx_t *c, int a, int i, int j, int k) {
  if (!wrap_src(c, &i, &j, &k)) return (real)0;
  return c->nrm[a][idx(c, i, j, k)];
}

/* ---------------------------------------------------------------------- */
/* Counter-based RNG: TSLB_newBCS/vars_single_3D_module.f90:102-165.        */
/* Integer-exact; the float steps are exact multiples of 2^-24.             */
/* ---------------------------------------------------------------------- */
#if defined(ORC_DEFINE_RNG)
float orc_rand_noseeded(int i, int j, int k, int l) {
  int isub = i % 178 + 1, jsub = j % 178 + 1, ksub = k % 178 + 1;
  int lsub = l % 169, msub;
  float s, t, u33, u97, csub, uni;
  const float cc = 362436.0f / 16777216.0f;
  const float cd = 7654321.0f / 16777216.0f;
  const float cm = 16777213.0f / 16777216.0f;
  s = 0.0f;
  t = 0.5f;
  for (int jj = 0; jj < 24; jj++) {
    msub = ((isub * jsub) % 179 * ksub) % 179;
    isub = jsub;
    jsub = ksub;
    ksub = msub;
    lsub = (53 * lsub + 1) % 169;
    if ((lsub * msub) % 64 >= 32) s = s + t;
    t = 0.5f * t;
  }
  u97 = s;
  s = 0.0f;
  t = 0.5f;
  for (int jj = 0; jj < 24; jj++) {
    msub = ((isub * jsub) % 179 * ksub) % 179;
    isub = jsub;
    jsub = ksub;
    ksub = msub;
    lsub = (53 * lsub + 1) % 169;
    if ((lsub * msub) % 64 >= 32) s = s + t;
    t = 0.5f * t;
  }
  u33 = s;
  uni = u97 - u33;
  if (uni < 0.0f) uni = uni + 1.0f;
  csub = cc - cd;
  if (csub < 0.0f) csub = csub + cm;
  uni = uni - csub;
  if (uni < 0.0f) uni = uni + 1.0f;
  return uni;
}

/* Box-Muller on two draws, vars_single_3D_module.f90:78-100 (REAL(4)). */
float orc_gauss_noseeded(int i, int j, int k, int l) {
  float d1 = orc_rand_noseeded(i, j, k, l);
  int kk = (int)lroundf(d1 * 50.0f); /* Fortran nint: half away from zero */
  float d2 = orc_rand_noseeded(i, j, k, l + kk);
  d1 = d1 * (1.0f - 1e-30f) + 1e-30f;
  return sqrtf(-2.0f * logf(d1)) * cosf(2.0f * 3.14159265358979323846f * d2);
}
#endif

/* ---------------------------------------------------------------------- */
/* Context                                                                 */
/* ---------------------------------------------------------------------- */

void *FN(orc_create)(int nx, int ny, int nz, int ncomp, const int *periodic, double tau,
                     double sigma, double sharp_c, const double *force, int force_mode,
                     double lin_a) {
  ctx_t *c = (ctx_t *)calloc(1, sizeof(ctx_t));
  if (!c) return NULL;
  c->nx = nx;
  c->ny = ny;
  c->nz = nz;
  c->ncomp = ncomp;
  for (int a = 0; a < 3; a++) c->per[a] = periodic[a];
  c->omega = (real)(1.0 / tau);
  c->sigma = (real)sigma;
  c->sharp_c = (real)sharp_c;
  c->fx = (real)force[0];
  c->fy = (real)force[1];
  c->fz = (real)force[2];
  c->force_mode = force_mode;
  c->lin_a = (real)lin_a;
  c->n = (size_t)nx * ny * nz;
  c->ng = (size_t)(nx + 2) * (ny + 2) * (nz + 2);
  c->f = (real *)calloc(27 * c->ng, sizeof(real));
  c->g = (real *)calloc(7 * c->ng, sizeof(real));
  c->tau_b = (real)tau;
  c->beta = (real)0.7;
  c->fa = ncomp == 3 ? (real *)calloc(27 * c->ng, sizeof(real)) : NULL;
  for (int q = 0; q < NFIELD; q++) c->m[q] = (real *)calloc(c->n, sizeof(real));
  for (int a = 0; a < 3; a++) {
    c->nrm[a] = (real *)calloc(c->n, sizeof(real));
    c->fst[a] = (real *)calloc(c->n, sizeof(real));
  }
  c->isf = (int32_t *)malloc(c->n * sizeof(int32_t));
  for (size_t p = 0; p < c->n; p++) c->isf[p] = 1;
  c->pops = NULL;
  c->gpops = NULL;
  return c;
}

void FN(orc_destroy)(void *h) {
  ctx_t *c = (ctx_t *)h;
  if (!c) return;
  free(c->f);
  free(c->g);
  for (int q = 0; q < NFIELD; q++) free(c->m[q]);
  for (int a = 0; a < 3; a++) {
    free(c->nrm[a]);
    free(c->fst[a]);
  }
  free(c->isf);
  free(c->pops);
  free(c->gpops);
  free(c->fa);
  free(c->apops);
  free(c->clo);
  free(c->zw_in);
  free(c->zphi_in);
  free(c);
}

/* Learned closure (no reference counterpart): per-node nu_t and stress
 * correction dsigma; the collision of node p uses omega = 1/(tau + 3 nu_t)
 * and the post-collision non-equilibrium stress (1 - omega) Pi + dsigma,
 * which a^neq, linear in Pi, turns into populations.  NULL clears. */
void FN(orc_set_closure)(void *h, const real *clo) {
  ctx_t *c = (ctx_t *)h;
  if (!clo) {
    free(c->clo);
    c->clo = NULL;
    return;
  }
  if (!c->clo) c->clo = (real *)malloc(7 * c->n * sizeof(real));
  memcpy(c->clo, clo, 7 * c->n * sizeof(real));
}

void FN(orc_set_isfluid)(void *h, const int32_t *isf) {
  ctx_t *c = (ctx_t *)h;
  memcpy(c->isf, isf, c->n * sizeof(int32_t));
}

/* ---------------------------------------------------------------------- */
/* Moment-space z-plane BC, bcs_TSLB_only_z_turbojet                         */
/* (TSLB_newBCS/boundary_cds_3D_module.f90:1233-1292).  The reference calls  */
/* it after the moments kernel and before the fused collide + push, which    */
/* runs over every node, boundary planes included                            */
/* (TSLB_newBCS/recursive_TSbcs_TSLB3.f90:563,617-769).  Its moments kernel  */
/* skips the two non-fluid planes, so their moments are exactly what this   */
/* BC writes:                                                                 */
/*   k = 1  (0 here):     rho, Pi from k = 2; u = v = 0; w = the inflow w    */
/*                        (:1262-1273; w(i,j,1) is never recomputed, so it    */
/*                        keeps the last inflow draw, :1253-1255)             */
/*   k = nz (nz-1 here):  rho, u, v, w, Pi from k = nz-1 (:1275-1285)          */
/* Here every node is fluid and z is not periodic, so K3/K1 also produce     */
/* moments on both planes (from bounced-back populations beyond them); the   */
/* BC overwrites all of them, so the result is the reference's.  Two-         */
/* component runs (no reference counterpart: the newBCS driver is one-        */
/* component) take phi of plane 0 from phi_in when given, else from k = 1,   */
/* and copy phi of plane nz-1 from nz-2.  The inflow redraw every 10 steps   */
/* (:1237-1259) is the caller's: it passes the new plane here.               */
/* ---------------------------------------------------------------------- */
static void zplane_bc(ctx_t *c) {
  if (!c->zw_in) return;
  const size_t plane = (size_t)c->nx * c->ny;
  const size_t lo1 = plane, hi = plane * (size_t)(c->nz - 1), hi1 = plane * (size_t)(c->nz - 2);
  for (size_t p = 0; p < plane; p++) {
    for (int q = 0; q < NFIELD; q++) c->m[q][hi + p] = c->m[q][hi1 + p];
    for (int q = 0; q < NFIELD; q++) {
      if (q == F_U || q == F_V) c->m[q][p] = 0;
      else if (q == F_W) c->m[q][p] = c->zw_in[p];
      else if (q == F_PHI && c->zphi_in) c->m[q][p] = c->zphi_in[p];
      else c->m[q][p] = c->m[q][lo1 + p];
    }
  }
}

/* Enable the z-plane BC with inflow w_in (nx*ny) and, for two components,
 * phi_in (or NULL); NULL w_in disables it.  Applied at once to the current
 * moments, then after every step. */
void FN(orc_set_zplane_bc)(void *h, const real *w_in, const real *phi_in) {
  ctx_t *c = (ctx_t *)h;
  const size_t plane = (size_t)c->nx * c->ny;
  if (!w_in) {
    free(c->zw_in);
    free(c->zphi_in);
    c->zw_in = c->zphi_in = NULL;
    return;
  }
  if (!c->zw_in) c->zw_in = (real *)malloc(plane * sizeof(real));
  memcpy(c->zw_in, w_in, plane * sizeof(real));
  if (phi_in && c->ncomp == 2) {
    if (!c->zphi_in) c->zphi_in = (real *)malloc(plane * sizeof(real));
    memcpy(c->zphi_in, phi_in, plane * sizeof(real));
  } else {
    free(c->zphi_in);
    c->zphi_in = NULL;
  }
  zplane_bc(c);
}

/* fields: 11 arrays in the order rho,u,v,w,pxx,pyy,pzz,pxy,pxz,pyz,phi */
void FN(orc_get_moments)(void *h, real *out) {
  ctx_t *c = (ctx_t *)h;
  for (int q = 0; q < NFIELD; q++) memcpy(out + (size_t)q * c->n, c->m[q], c->n * sizeof(real));
}
void FN(orc_set_moments)(void *h, const real *in) {
  ctx_t *c = (ctx_t *)h;
  for (int q = 0; q < NFIELD; q++) memcpy(c->m[q], in + (size_t)q * c->n, c->n * sizeof(real));
}
void FN(orc_get_normals)(void *h, real *out) {
  ctx_t *c = (ctx_t *)h;
  for (int a = 0; a < 3; a++) memcpy(out + (size_t)a * c->n, c->nrm[a], c->n * sizeof(real));
}

/* ---------------------------------------------------------------------- */
/* K1: moments (+phi), recursive_2c_TSLB3D.f90:346-542                     */
/* ---------------------------------------------------------------------- */
static void moments_from_pops(const ctx_t *c, const real *fl, real *rho_o, real *u_o, real *v_o,
                              real *w_o, real *P) {
  real rho = 0, jx = 0, jy = 0, jz = 0;
  for (int l = 0; l < 27; l++) rho += fl[l];
  real px = 0, nx_ = 0, py = 0, ny_ = 0, pz = 0, nz_ = 0;
  for (int l = 0; l < 27; l++) {
    if (CX[l] > 0) px += fl[l];
    if (CX[l] < 0) nx_ += fl[l];
    if (CY[l] > 0) py += fl[l];
    if (CY[l] < 0) ny_ += fl[l];
    if (CZ[l] > 0) pz += fl[l];
    if (CZ[l] < 0) nz_ += fl[l];
  }
  jx = (px - nx_) / rho;
  jy = (py - ny_) / rho;
  jz = (pz - nz_) / rho;
  for (int q = 0; q < 6; q++) P[q] = 0;
  for (int l = 1; l < 27; l++) {
    real fneq = fl[l] - feq_l(l, rho, jx, jy, jz);
    real cx = (real)CX[l], cy = (real)CY[l], cz = (real)CZ[l];
    if (CX[l]) P[0] += fneq;
    if (CY[l]) P[1] += fneq;
    if (CZ[l]) P[2] += fneq;
    if (CX[l] && CY[l]) P[3] += cx * cy * fneq;
    if (CX[l] && CZ[l]) P[4] += cx * cz * fneq;
    if (CY[l] && CZ[l]) P[5] += cy * cz * fneq;
  }
  (void)c;
  *rho_o = rho;
  *u_o = jx;
  *v_o = jy;
  *w_o = jz;
}

static void k1_moments(ctx_t *c) {
  const int nx = c->nx, ny = c->ny, nz = c->nz;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        size_t p = idx(c, i, j, k);
        if (!c->isf[p]) continue;
        size_t gp = gidx(c, i, j, k);
        real fl[27], P[6], rho, u, v, w;
        for (int l = 0; l < 27; l++) fl[l] = c->f[(size_t)l * c->ng + gp];
        moments_from_pops(c, fl, &rho, &u, &v, &w, P);
        c->m[F_RHO][p] = rho;
        c->m[F_U][p] = u;
        c->m[F_V][p] = v;
        c->m[F_W][p] = w;
        for (int q = 0; q < 6; q++) c->m[F_PXX + q][p] = P[q];
        if (c->ncomp == 2) {
          real phi = 0;
          for (int l = 0; l < 7; l++) phi += c->g[(size_t)l * c->ng + gp];
          c->m[F_PHI][p] = phi;
        }
        if (c->ncomp == 3) {
          real ra = 0;
          for (int l = 0; l < 27; l++) ra += c->fa[(size_t)l * c->ng + gp];
          c->m[F_PHI][p] = (real)2 * ra / rho - (real)1;
        }
      }
}

/* ---------------------------------------------------------------------- */
/* K2: colour gradient and normals, recursive_2c_TSLB3D.f90:555-571         */
/* grad = (1/cs^2) sum_l w_l c_l phi(x+c_l), grouped as the reference:      */
/* 3*(p1*face + p2*edges + p3*corners).                                      */
/* ---------------------------------------------------------------------- */
static void grad_stencil(const ctx_t *c, int i, int j, int k, int which, real *g3) {
  const real p1 = wgt(1), p2 = wgt(7), p3 = wgt(19);
  for (int a = 0; a < 3; a++) {
    real face = 0, edge_p = 0, edge_m = 0, corn_p = 0, corn_m = 0;
    real fp = 0, fm = 0;
    for (int l = 1; l < 27; l++) {
      int ca = a == 0 ? CX[l] : (a == 1 ? CY[l] : CZ[l]);
      if (ca == 0) continue;
      int nnz = (CX[l] != 0) + (CY[l] != 0) + (CZ[l] != 0);
      real val = which < 0 ? phi_at(c, i + CX[l], j + CY[l], k + CZ[l])
                           : nrm_at(c, which, i + CX[l], j + CY[l], k + CZ[l]);
      if (nnz == 1) {
        if (ca > 0) fp = val; else fm = val;
      } else if (nnz == 2) {
        if (ca > 0) edge_p += val; else edge_m += val;
      } else {
        if (ca > 0) corn_p += val; else corn_m += val;
      }
    }
    face = fp - fm;
    g3[a] = (real)3 * (p1 * face + p2 * (edge_p - edge_m) + p3 * (corn_p - corn_m));
  }
}

static void k2_normals(ctx_t *c) {
  const int nx = c->nx, ny = c->ny, nz = c->nz;
  const real eps = (real)1e-9f;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        size_t p = idx(c, i, j, k);
        if (!c->isf[p]) {
          c->nrm[0][p] = c->nrm[1][p] = c->nrm[2][p] = 0;
          continue;
        }
        real g3[3];
        grad_stencil(c, i, j, k, -1, g3);
        real ss = g3[0] * g3[0] + g3[1] * g3[1] + g3[2] * g3[2];
        real mod = sizeof(real) == 4 ? (real)sqrtf((float)ss) : (real)sqrt((double)ss);
        c->nrm[0][p] = g3[0] / (mod + eps);
        c->nrm[1][p] = g3[1] / (mod + eps);
        c->nrm[2][p] = g3[2] / (mod + eps);
      }
}

/* Gated curvature at fluid node p, recursive_2c_TSLB3D.f90:614-631:
 * kappa = -div(n) where phi(1-phi)|n|^2 > 0.2, else 0 */
static real node_curv(const ctx_t *c, int i, int j, int k) {
  size_t p = idx(c, i, j, k);
  real phi = c->m[F_PHI][p];
  real n0 = c->nrm[0][p], n1 = c->nrm[1][p], n2 = c->nrm[2][p];
  real curv = 0;
  real indicator = phi * ((real)1 - phi) * (n0 * n0 + n1 * n1 + n2 * n2);
  if (indicator > (real)0.2f) {
    real dx[3], dy[3], dz[3];
    grad_stencil(c, i, j, k, 0, dx);
    grad_stencil(c, i, j, k, 1, dy);
    grad_stencil(c, i, j, k, 2, dz);
    curv = -(dx[0] + dy[1] + dz[2]);
  }
  return curv;
}

/* CSF shift sigma*kappa*n at fluid node p, recursive_2c_TSLB3D.f90:614-636 */
static void csf_shift(const ctx_t *c, int i, int j, int k, real *s3) {
  size_t p = idx(c, i, j, k);
  real n0 = c->nrm[0][p], n1 = c->nrm[1][p], n2 = c->nrm[2][p];
  real curv = node_curv(c, i, j, k);
  s3[0] = c->sigma * curv * n0;
  s3[1] = c->sigma * curv * n1;
  s3[2] = c->sigma * curv * n2;
}

/* K2 on the current phi, then the gated curvature of every fluid node (0
 * on solids and where the gate is closed) into out (n values). */
void FN(orc_curvature)(void *h, real *out) {
  ctx_t *c = (ctx_t *)h;
  k2_normals(c);
#pragma omp parallel for schedule(static)
  for (int k = 0; k < c->nz; k++)
    for (int j = 0; j < c->ny; j++)
      for (int i = 0; i < c->nx; i++) {
        size_t p = idx(c, i, j, k);
        out[p] = c->isf[p] ? node_curv(c, i, j, k) : (real)0;
      }
}

/* ---------------------------------------------------------------------- */
/* Post-collision populations of one fluid node: recursive-regularized      */
/* collision, recursive_2c_TSLB3D.f90:637-806 (2C), recursive_TSLB3D.f90:   */
/* 556-707 (1C additive forcing), lb_kernels_module.f90:585-587 (shift).     */
/* ---------------------------------------------------------------------- */
static void collide_node(const ctx_t *c, size_t p, const real *shift, real *fo, real *go) {
  real rho = c->m[F_RHO][p], u = c->m[F_U][p], v = c->m[F_V][p], w = c->m[F_W][p];
  real P[6];
  for (int q = 0; q < 6; q++) P[q] = c->m[F_PXX + q][p];
  real us = u, vs = v, ws = w;
  if (c->ncomp == 2) {
    us = u + shift[0];
    vs = v + shift[1];
    ws = w + shift[2];
  }
  if (c->ncomp == 2 || c->force_mode == 1) {
    us = us + c->fx + c->lin_a * u;
    vs = vs + c->fy + c->lin_a * v;
    ws = ws + c->fz + c->lin_a * w;
  }
  real om1 = (real)1 - c->omega;
  if (c->clo) {
    const real om = (real)1 / ((real)1 / c->omega + (real)3 * c->clo[p]);
    for (int q = 0; q < 6; q++) P[q] = ((real)1 - om) * P[q] + c->clo[(size_t)(1 + q) * c->n + p];
    om1 = (real)1;
  }
  for (int l = 0; l < 27; l++) {
    real val = feq_l(l, rho, us, vs, ws) + om1 * aneq_l(l, P, u, v, w) * wgt(l);
    if (c->ncomp == 1 && c->force_mode == 0 && l > 0)
      val += (c->fx * (real)CX[l] + c->fy * (real)CY[l] + c->fz * (real)CZ[l]) * (wgt(l) * (real)3);
    fo[l] = val;
  }
  if (c->ncomp == 2 && go) {
    /* D3Q7 phase field with sharpening, recursive_2c_TSLB3D.f90:808-831 */
    real phi = c->m[F_PHI][p];
    const real p0g = (real)0.25, p1g = (real)0.125;
    real n3[3] = {c->nrm[0][p], c->nrm[1][p], c->nrm[2][p]};
    real u3[3] = {u, v, w};
    go[0] = p0g * phi;
    for (int a = 0; a < 3; a++) {
      real sh = p1g * c->sharp_c * phi * ((real)1 - phi) * n3[a];
      go[1 + 2 * a] = p1g * phi * ((real)1 + (real)3 * u3[a]) + sh;
      go[2 + 2 * a] = p1g * phi * ((real)1 - (real)3 * u3[a]) - sh;
    }
  }
}

/* ---------------------------------------------------------------------- */
/* Turbulent-jet boundary, bcs_2c_turbulent_jet (boundary_cds_3D_module.f90: */
/* 862-1049).  Planes k = 0 (inlet) and k = nz-1 (outlet) are walls.  After  */
/* K3 the solids bounce f back (:878-930) but not g: a g population pushed    */
/* into a wall is lost, and a wall sends none (its slots keep their zero      */
/* allocation value).  Then the 9 populations crossing from the inlet plane   */
/* into plane 1 are replaced by regularized inlet values                      */
/*   f_l(y) = feq_l(rho(y), u(x)) + (1-omega) w_l a^neq_l(Pi(y), u(y)),       */
/* x = y - c_l on the inlet plane (:932-990), and g_5(y) = phi(x)(1+3w(x))/8  */
/* (:992-993).  On the outlet plane w(x) = w(x - e_z) (:996), then the 9      */
/* populations crossing into plane nz-2 take the same form (:997-1044) and    */
/* g_6(y) = phi(y)(1 - 3*0.05)/8 (:1048-1049).  Destinations across a         */
/* non-periodic x/y face keep their bounce-back value.                        */
/* ---------------------------------------------------------------------- */
static real jet_value(int l, real omega, real rho_y, const real *P_y, real uy, real vy, real wy,
                      real ux, real vx, real wx) {
  return feq_l(l, rho_y, ux, vx, wx) + ((real)1 - omega) * aneq_l(l, P_y, uy, vy, wy) * wgt(l);
}
static real jet_pop(const ctx_t *c, int l, size_t py, real ux, real vx, real wx) {
  real P[6];
  for (int q = 0; q < 6; q++) P[q] = c->m[F_PXX + q][py];
  return jet_value(l, c->omega, c->m[F_RHO][py], P, c->m[F_U][py], c->m[F_V][py], c->m[F_W][py], ux,
                   vx, wx);
}
/* the 27 jet values for destination moments (rho_y, P_y, u_y) and a wall
 * velocity u_x (exposed for the pins against the reference's polynomials) */
void FN(orc_jet_values)(real omega, real rho_y, const real *P_y, const real *u_y, const real *u_x,
                        real *out) {
  for (int l = 0; l < 27; l++)
    out[l] = jet_value(l, omega, rho_y, P_y, u_y[0], u_y[1], u_y[2], u_x[0], u_x[1], u_x[2]);
}

/* in-plane wrap of a destination column; 0 across a non-periodic face */
static int wrap_xy(const ctx_t *c, int *i, int *j) {
  if (*i < 0 || *i >= c->nx) {
    if (!c->per[0]) return 0;
    *i = wrap1(*i, c->nx);
  }
  if (*j < 0 || *j >= c->ny) {
    if (!c->per[1]) return 0;
    *j = wrap1(*j, c->ny);
  }
  return 1;
}

void FN(orc_set_jet)(void *h, int enable) { ((ctx_t *)h)->jet = enable ? 1 : 0; }

/* jet values of one inlet (k = 0) or outlet (k = nz-1) node x, push form */
static void jet_push_node(ctx_t *c, int i, int j, int outlet) {
  const int kx = outlet ? c->nz - 1 : 0, ky = outlet ? c->nz - 2 : 1, sz = outlet ? -1 : 1;
  const size_t px = idx(c, i, j, kx);
  const real ux = c->m[F_U][px], vx = c->m[F_V][px], wx = c->m[F_W][px];
  for (int l = 1; l < 27; l++) {
    if (CZ[l] != sz) continue;
    int yi = i + CX[l], yj = j + CY[l];
    if (!wrap_xy(c, &yi, &yj)) continue;
    c->f[(size_t)l * c->ng + gidx(c, yi, yj, ky)] = jet_pop(c, l, idx(c, yi, yj, ky), ux, vx, wx);
  }
  if (c->ncomp == 2) {
    const real p1g = (real)0.125;
    if (!outlet)
      c->g[(size_t)5 * c->ng + gidx(c, i, j, 1)] = p1g * c->m[F_PHI][px] * ((real)1 + (real)3 * wx);
    else
      c->g[(size_t)6 * c->ng + gidx(c, i, j, ky)] =
          p1g * c->m[F_PHI][idx(c, i, j, ky)] * ((real)1 - (real)3 * (real)0.05);
  }
}

static void jet_bc_push(ctx_t *c) {
  const int nx = c->nx, ny = c->ny, nz = c->nz;
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) c->m[F_W][idx(c, i, j, nz - 1)] = c->m[F_W][idx(c, i, j, nz - 2)];
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      jet_push_node(c, i, j, 0);
      jet_push_node(c, i, j, 1);
    }
}

/* ---------------------------------------------------------------------- */
/* Colour-gradient model (ncomp 3): two components A and B carried as the   */
/* total populations f = fA + fB and the A populations fA, with the colour  */
/* field psi = (rhoA - rhoB)/rho stored in place of phi.  Formula source:   */
/* turbulent_benchmark/couette/recursive_colorTSLB3D.f90:719-959, which     */
/* does not compile; restated with its defects corrected:                   */
/*  - psi = (rhoA - rhoB)/rho (the file inverts it, :511);                  */
/*  - u from the summed momenta over rho = rhoA + rhoB (:494-508);          */
/*  - the perturbation is quadratic in c.grad(psi) (the file's linear form  */
/*    breaks momentum conservation, :772 ff):                               */
/*      Omega_0 = (5/27) A |g|,  Omega_l = A ((c_l.g)^2 - |g|^2)/(k_l |g|),  */
/*      k_l = 27, 108, 432 for |c_l|^2 = 1, 2, 3 (the file's coefficients,  */
/*    :764-956), which conserves mass and momentum and gives the interface  */
/*    stress (A/(9|g|))(g g - |g|^2 I); its surface tension is 2A/(9 omega), */
/*    so A = (9/2) sigma omega makes sigma the surface tension (the file's   */
/*    st_coeff = (9/4) sigma omega, :760, would give sigma/2);               */
/*  - Latva-Kokko recolouring with cos = c.g/(|c||g|) (the file divides by  */
/*    |g|^2, :773):  fA_l = (rhoA/rho) f_l + beta (rhoA rhoB/rho^2) w_l rho  */
/*    (c_l.g)/(|c_l||g|), both only where |g| > 0.001 (:763);              */
/*  - per-node omega from the harmonic mean viscosity (:756-758) of         */
/*    nu_A = (tau - 1/2)/3 and nu_B = (tau_b - 1/2)/3;                       */
/*  - additive body force w_l (F.c_l)/cs^2 (the file's fx*p1dcssq).          */
/* The gradient g = (1/cs^2) sum_l w_l c_l psi(x + c_l) is the stencil of   */
/* grad_stencil.                                                             */
/* ---------------------------------------------------------------------- */
void FN(orc_set_colour)(void *h, double tau_b, double beta) {
  ctx_t *c = (ctx_t *)h;
  c->tau_b = (real)tau_b;
  c->beta = (real)beta;
}

static void k2_colour_grad(ctx_t *c) {
  const int nx = c->nx, ny = c->ny, nz = c->nz;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        size_t p = idx(c, i, j, k);
        if (!c->isf[p]) {
          c->nrm[0][p] = c->nrm[1][p] = c->nrm[2][p] = 0;
          continue;
        }
        real g3[3];
        grad_stencil(c, i, j, k, -1, g3);
        for (int a = 0; a < 3; a++) c->nrm[a][p] = g3[a];
      }
}

/* total and A post-collision populations of one fluid node from its moments
 * (rho, u, Pi, psi) and colour gradient g */
static void colour_values(real rho, real u, real v, real w, const real *P, real psi, const real *g3,
                          real omega_a, real omega_b, real sigma, real beta, real fx, real fy, real fz,
                          real *fo, real *fao) {
  real ca = (real)0.5 * ((real)1 + psi);
  if (ca < 0) ca = 0;
  if (ca > 1) ca = 1;
  const real cb = (real)1 - ca;
  const real nua = ((real)1 / omega_a - (real)0.5) / (real)3;
  const real nub = ((real)1 / omega_b - (real)0.5) / (real)3;
  const real nu = (real)1 / (ca / nua + cb / nub);
  const real omega = (real)1 / ((real)3 * nu + (real)0.5);
  const real mg2 = g3[0] * g3[0] + g3[1] * g3[1] + g3[2] * g3[2];
  const real mg = sizeof(real) == 4 ? (real)sqrtf((float)mg2) : (real)sqrt((double)mg2);
  const int pert = mg > (real)0.001;
  const real A = (real)4.5 * sigma * omega;
  for (int l = 0; l < 27; l++) {
    real val = feq_l(l, rho, u, v, w) + ((real)1 - omega) * aneq_l(l, P, u, v, w) * wgt(l);
    if (l > 0) val += (fx * (real)CX[l] + fy * (real)CY[l] + fz * (real)CZ[l]) * (wgt(l) * (real)3);
    real om2 = 0;
    if (pert) {
      if (l == 0) {
        val += (real)5 * A * mg / (real)27;
      } else {
        const int nnz = (CX[l] != 0) + (CY[l] != 0) + (CZ[l] != 0);
        const real k = nnz == 1 ? (real)27 : (nnz == 2 ? (real)108 : (real)432);
        const real cg = (real)CX[l] * g3[0] + (real)CY[l] * g3[1] + (real)CZ[l] * g3[2];
        val += A * (cg * cg - mg2) / (k * mg);
        const real cl = nnz == 1 ? (real)1 : (nnz == 2 ? (real)1.4142135623730951 : (real)1.7320508075688772);
        om2 = beta * ca * cb * wgt(l) * rho * cg / (cl * mg);
      }
    }
    fo[l] = val;
    fao[l] = ca * val + om2;
  }
}

void FN(orc_colour_values)(double omega_a, double omega_b, double sigma, double beta, real rho,
                           const real *u3, const real *P6, real psi, const real *g3, real *fo,
                           real *fao) {
  colour_values(rho, u3[0], u3[1], u3[2], P6, psi, g3, (real)omega_a, (real)omega_b, (real)sigma,
                (real)beta, 0, 0, 0, fo, fao);
}

static void colour_node(const ctx_t *c, size_t p, real *fo, real *fao) {
  real P[6], g3[3] = {c->nrm[0][p], c->nrm[1][p], c->nrm[2][p]};
  for (int q = 0; q < 6; q++) P[q] = c->m[F_PXX + q][p];
  colour_values(c->m[F_RHO][p], c->m[F_U][p], c->m[F_V][p], c->m[F_W][p], P, c->m[F_PHI][p], g3,
                c->omega, (real)1 / c->tau_b, c->sigma, c->beta, c->fx, c->fy, c->fz, fo, fao);
}

/* K3 push on fluid nodes, then boundary resolve */
static void k3_push(ctx_t *c) {
  const int nx = c->nx, ny = c->ny, nz = c->nz;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        size_t p = idx(c, i, j, k);
        if (!c->isf[p]) continue;
        real s3[3] = {0, 0, 0}, fo[27], go[7];
        if (c->ncomp == 3) {
          real fao[27];
          colour_node(c, p, fo, fao);
          for (int l = 0; l < 27; l++) {
            const size_t d = (size_t)l * c->ng + gidx(c, i + CX[l], j + CY[l], k + CZ[l]);
            c->f[d] = fo[l];
            c->fa[d] = fao[l];
          }
          continue;
        }
        if (c->ncomp == 2) csf_shift(c, i, j, k, s3);
        collide_node(c, p, s3, fo, go);
        for (int l = 0; l < 27; l++)
          c->f[(size_t)l * c->ng + gidx(c, i + CX[l], j + CY[l], k + CZ[l])] = fo[l];
        if (c->ncomp == 2)
          for (int l = 0; l < 7; l++)
            c->g[(size_t)l * c->ng + gidx(c, i + CX[l], j + CY[l], k + CZ[l])] = go[l];
      }
}

/* Boundary resolve for array a with nq directions (27 for f, 7 for g). */
static void resolve(ctx_t *c, real *a, int nq) {
  const int nx = c->nx, ny = c->ny, nz = c->nz;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        size_t p = idx(c, i, j, k);
        if (!c->isf[p]) continue;
        size_t gp = gidx(c, i, j, k);
        for (int l = 1; l < nq; l++) {
          int si = i - CX[l], sj = j - CY[l], sk = k - CZ[l];
          int ri = si, rj = sj, rk = sk;
          int inside = wrap_src(c, &si, &sj, &sk);
          if (!inside || !c->isf[idx(c, si, sj, sk)]) {
            /* halfway bounce-back: the population y sent towards s returns;
             * the jet boundary reflects f only */
            if (c->jet && nq == 7) a[(size_t)l * c->ng + gp] = 0;
            else a[(size_t)l * c->ng + gp] = a[(size_t)OPP[l] * c->ng + gidx(c, ri, rj, rk)];
          } else if (si != ri || sj != rj || sk != rk) {
            /* periodic wrap: the source pushed into its ghost neighbour */
            a[(size_t)l * c->ng + gp] =
                a[(size_t)l * c->ng + gidx(c, si + CX[l], sj + CY[l], sk + CZ[l])];
          }
        }
      }
}

/* One reference step: K2, K3, boundary, K1. */
void FN(orc_step)(void *h, int nsteps) {
  ctx_t *c = (ctx_t *)h;
  for (int s = 0; s < nsteps; s++) {
    if (c->ncomp == 2) k2_normals(c);
    if (c->ncomp == 3) k2_colour_grad(c);
    k3_push(c);
    resolve(c, c->f, 27);
    if (c->ncomp == 2) resolve(c, c->g, 7);
    if (c->ncomp == 3) resolve(c, c->fa, 27);
    if (c->jet) jet_bc_push(c);
    k1_moments(c);
    zplane_bc(c);
  }
}

/* Initial state, recursive_2c_TSLB3D.f90:87-287: moments set on every node,
 * f = feq at fluid nodes, g pre-pushed (when prepush), then K1. */
void FN(orc_init)(void *h, const real *rho, const real *u, const real *v, const real *w,
                  const real *phi, int prepush) {
  ctx_t *c = (ctx_t *)h;
  const int nx = c->nx, ny = c->ny, nz = c->nz;
  for (size_t p = 0; p < c->n; p++) {
    c->m[F_RHO][p] = rho[p];
    c->m[F_U][p] = u[p];
    c->m[F_V][p] = v[p];
    c->m[F_W][p] = w[p];
    for (int q = 0; q < 6; q++) c->m[F_PXX + q][p] = 0;
    c->m[F_PHI][p] = phi ? phi[p] : (real)0;
    c->nrm[0][p] = c->nrm[1][p] = c->nrm[2][p] = 0;
  }
  memset(c->f, 0, 27 * c->ng * sizeof(real));
  memset(c->g, 0, 7 * c->ng * sizeof(real));
  if (c->fa) memset(c->fa, 0, 27 * c->ng * sizeof(real));
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        size_t p = idx(c, i, j, k);
        if (!c->isf[p]) continue;
        size_t gp = gidx(c, i, j, k);
        for (int l = 0; l < 27; l++)
          c->f[(size_t)l * c->ng + gp] = feq_l(l, c->m[F_RHO][p], c->m[F_U][p], c->m[F_V][p], c->m[F_W][p]);
        if (c->ncomp == 3) {
          const real ca = (real)0.5 * ((real)1 + c->m[F_PHI][p]);
          for (int l = 0; l < 27; l++)
            c->fa[(size_t)l * c->ng + gp] = ca * c->f[(size_t)l * c->ng + gp];
        }
        if (c->ncomp == 2) {
          real ph = c->m[F_PHI][p];
          real u3[3] = {c->m[F_U][p], c->m[F_V][p], c->m[F_W][p]};
          c->g[gp] = (real)0.25 * ph;
          for (int a = 0; a < 3; a++) {
            real gp_ = (real)0.125 * ph * ((real)1 + (real)3 * u3[a]);
            real gm_ = (real)0.125 * ph * ((real)1 - (real)3 * u3[a]);
            int lp = 1 + 2 * a, lm = 2 + 2 * a;
            if (prepush) {
              c->g[(size_t)lp * c->ng + gidx(c, i + CX[lp], j + CY[lp], k + CZ[lp])] = gp_;
              c->g[(size_t)lm * c->ng + gidx(c, i + CX[lm], j + CY[lm], k + CZ[lm])] = gm_;
            } else {
              c->g[(size_t)lp * c->ng + gp] = gp_;
              c->g[(size_t)lm * c->ng + gp] = gm_;
            }
          }
        }
      }
  if (c->ncomp == 2 && prepush) resolve(c, c->g, 7);
  k1_moments(c);
}

/* ---------------------------------------------------------------------- */
/* Moment-space pull restatement (the form the GPU kernels implement):      */
/*   m_{t+1}(y) = M[ f_l(y) ],  f_l(y) = R_l(m_t(y - c_l))  or, when the     */
/*   source is solid / outside a wall, R_opp(l)(m_t(y)).                      */
/* It computes exactly the same post-collision values as orc_step and sums  */
/* them in the same order, so both forms agree bit for bit.                 */
/* ---------------------------------------------------------------------- */
void FN(orc_pull_step)(void *h, int nsteps) {
  ctx_t *c = (ctx_t *)h;
  const int nx = c->nx, ny = c->ny, nz = c->nz;
  if (!c->pops) c->pops = (real *)malloc(27 * c->n * sizeof(real));
  if (!c->gpops) c->gpops = (real *)malloc(7 * c->n * sizeof(real));
  if (c->ncomp == 3 && !c->apops) c->apops = (real *)malloc(27 * c->n * sizeof(real));
  real *nm[NFIELD];
  for (int q = 0; q < NFIELD; q++) nm[q] = (real *)malloc(c->n * sizeof(real));
  for (int s = 0; s < nsteps; s++) {
    if (c->ncomp == 2) k2_normals(c);
    if (c->ncomp == 3) k2_colour_grad(c);
    /* post-collision populations at every fluid source, node-local */
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; k++)
      for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
          size_t p = idx(c, i, j, k);
          if (!c->isf[p]) continue;
          real s3[3] = {0, 0, 0}, fo[27], go[7];
          if (c->ncomp == 3) {
            real fao[27];
            colour_node(c, p, fo, fao);
            for (int l = 0; l < 27; l++) {
              c->pops[(size_t)l * c->n + p] = fo[l];
              c->apops[(size_t)l * c->n + p] = fao[l];
            }
            continue;
          }
          if (c->ncomp == 2) csf_shift(c, i, j, k, s3);
          collide_node(c, p, s3, fo, go);
          for (int l = 0; l < 27; l++) c->pops[(size_t)l * c->n + p] = fo[l];
          if (c->ncomp == 2)
            for (int l = 0; l < 7; l++) c->gpops[(size_t)l * c->n + p] = go[l];
        }
    /* pull + moments */
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; k++)
      for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
          size_t p = idx(c, i, j, k);
          if (!c->isf[p]) {
            for (int q = 0; q < NFIELD; q++) nm[q][p] = c->m[q][p];
            continue;
          }
          real fl[27], gl[7], al[27], P[6], rho, u, v, w;
          for (int l = 0; l < 27; l++) {
            int si = i - CX[l], sj = j - CY[l], sk = k - CZ[l];
            int inside = wrap_src(c, &si, &sj, &sk);
            if (l > 0 && (!inside || !c->isf[idx(c, si, sj, sk)])) {
              fl[l] = c->pops[(size_t)OPP[l] * c->n + p];
              if (c->ncomp == 3) al[l] = c->apops[(size_t)OPP[l] * c->n + p];
              if (l < 7 && c->ncomp == 2)
                gl[l] = c->jet ? (real)0 : c->gpops[(size_t)OPP[l] * c->n + p];
              /* jet inlet / outlet: the source column x on the wall plane */
              const int in_l = c->jet && k == 1 && CZ[l] == 1;
              const int out_l = c->jet && k == nz - 2 && CZ[l] == -1;
              int xi = i - CX[l], xj = j - CY[l];
              if ((in_l || out_l) && wrap_xy(c, &xi, &xj)) {
                const size_t px = idx(c, xi, xj, in_l ? 0 : nz - 1);
                const real wx = in_l ? c->m[F_W][px] : c->m[F_W][idx(c, xi, xj, nz - 2)];
                fl[l] = jet_pop(c, l, p, c->m[F_U][px], c->m[F_V][px], wx);
                if (c->ncomp == 2 && l == 5)
                  gl[l] = (real)0.125 * c->m[F_PHI][px] * ((real)1 + (real)3 * wx);
                if (c->ncomp == 2 && l == 6)
                  gl[l] = (real)0.125 * c->m[F_PHI][p] * ((real)1 - (real)3 * (real)0.05);
              }
            } else {
              size_t sp = idx(c, si, sj, sk);
              fl[l] = c->pops[(size_t)l * c->n + sp];
              if (c->ncomp == 3) al[l] = c->apops[(size_t)l * c->n + sp];
              if (l < 7 && c->ncomp == 2) gl[l] = c->gpops[(size_t)l * c->n + sp];
            }
          }
          moments_from_pops(c, fl, &rho, &u, &v, &w, P);
          nm[F_RHO][p] = rho;
          nm[F_U][p] = u;
          nm[F_V][p] = v;
          nm[F_W][p] = w;
          for (int q = 0; q < 6; q++) nm[F_PXX + q][p] = P[q];
          if (c->ncomp == 2) {
            real phi = 0;
            for (int l = 0; l < 7; l++) phi += gl[l];
            nm[F_PHI][p] = phi;
          } else if (c->ncomp == 3) {
            real ra = 0;
            for (int l = 0; l < 27; l++) ra += al[l];
            nm[F_PHI][p] = (real)2 * ra / rho - (real)1;
          } else {
            nm[F_PHI][p] = c->m[F_PHI][p];
          }
        }
    if (c->jet)
      for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) nm[F_W][idx(c, i, j, nz - 1)] = c->m[F_W][idx(c, i, j, nz - 2)];
    for (int q = 0; q < NFIELD; q++) memcpy(c->m[q], nm[q], c->n * sizeof(real));
    zplane_bc(c);
  }
  for (int q = 0; q < NFIELD; q++) free(nm[q]);
}

/* ---------------------------------------------------------------------- */
/* FP64 reductions over fluid nodes:                                        */
/*  out[0]=sum rho, [1]=sum phi, [2]=sum rho|u|^2/2, [3..5]=sum rho u,       */
/*  [6]=enstrophy sum |curl u|^2/2 (central differences, periodic wrap).     */
/* ---------------------------------------------------------------------- */
static double fld(const ctx_t *c, int q, int i, int j, int k) {
  int d[3] = {c->nx, c->ny, c->nz}, p[3] = {i, j, k};
  for (int a = 0; a < 3; a++) {
    if (p[a] < 0 || p[a] >= d[a]) {
      if (c->per[a]) p[a] = wrap1(p[a], d[a]);
      else p[a] = p[a] < 0 ? 0 : d[a] - 1;
    }
  }
  return (double)c->m[q][idx(c, p[0], p[1], p[2])];
}

void FN(orc_reduce)(void *h, double *out) {
  ctx_t *c = (ctx_t *)h;
  double s[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int k = 0; k < c->nz; k++)
    for (int j = 0; j < c->ny; j++)
      for (int i = 0; i < c->nx; i++) {
        size_t p = idx(c, i, j, k);
        if (!c->isf[p]) continue;
        double r = c->m[F_RHO][p], u = c->m[F_U][p], v = c->m[F_V][p], w = c->m[F_W][p];
        s[0] += r;
        s[1] += c->ncomp == 3 ? 0.5 * r * (1.0 + (double)c->m[F_PHI][p]) : (double)c->m[F_PHI][p];
        s[2] += 0.5 * r * (u * u + v * v + w * w);
        s[3] += r * u;
        s[4] += r * v;
        s[5] += r * w;
        double wx = 0.5 * (fld(c, F_W, i, j + 1, k) - fld(c, F_W, i, j - 1, k)) -
                    0.5 * (fld(c, F_V, i, j, k + 1) - fld(c, F_V, i, j, k - 1));
        double wy = 0.5 * (fld(c, F_U, i, j, k + 1) - fld(c, F_U, i, j, k - 1)) -
                    0.5 * (fld(c, F_W, i + 1, j, k) - fld(c, F_W, i - 1, j, k));
        double wz = 0.5 * (fld(c, F_V, i + 1, j, k) - fld(c, F_V, i - 1, j, k)) -
                    0.5 * (fld(c, F_U, i, j + 1, k) - fld(c, F_U, i, j - 1, k));
        s[6] += 0.5 * (wx * wx + wy * wy + wz * wz);
      }
  for (int q = 0; q < 7; q++) out[q] = s[q];
}

#if defined(ORC_DEFINE_RNG)
void orc_set_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n);
#else
  (void)n;
#endif
}
int orc_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
#endif
