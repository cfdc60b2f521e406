// tslb.hip -- MI355X (gfx950) TSLB library: HIP kernels and the C ABI of
// include/tslb.h.
//
// State: SoA moment fields on the device, Fortran (i fastest) order, two
// buffers (source / destination of a step).  Field 0 holds rho-1 (shifted
// density), so that the FP32 sums of the fused kernel keep their precision;
// the host interface converts to rho.
//
// One step (reference: recursive_2c_TSLB3D.f90:344-840) is, in moment space,
//   m_{t+1}(y) = M[ f_l(y) ],   f_l(y) = R_l(m_t(y - c_l)),
// where R_l is the recursive-regularized post-collision population
// (feq_l(rho,u_eq) + (1-omega) w_l a^neq_l(Pi,u) [+ forcing]) and M takes the
// moments (rho, u, Pi^neq; phi from the D3Q7 phase populations).  Bounce-back
// replaces R_l(m(y-c_l)) by R_opp(l)(m(y)) when y-c_l is solid or beyond a wall.
// Because the reference's collide kernel reads only node-local moments and
// writes each destination once, this restatement is exact.
//
// Kernels
//   k_fused     LDS-tiled 2.5D collide + pull-stream.  A 64x8 block marches a
//               z-chunk; per source plane it evaluates the 27 (+7) populations
//               of every node of the (64+2)x(8+2) footprint once, scatters the
//               24 in-plane-moving ones (+2 phase fluxes) through LDS, and each
//               destination column gathers its populations and accumulates
//               moment sums for three destination planes in registers.  Phase
//               field and colour model: the normals or the colour gradient
//               come from three phi (psi) planes kept in an LDS ring.
//   k_normals   dense normals and curvature or colour gradient (downloads,
//               boundary passes); sparse curvature (phase-field step).
//   k_prepush   initial one-step advection of phi (reference writes g pushed).
//   k_links     per-node bounce-back link masks from an isfluid map.
//   k_zplane_bc moment-space inlet/outlet planes.
//   k_reduce*   FP64 diagnostics.
#include <hip/hip_runtime.h>

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/resource.h>
#include <sys/sysinfo.h>

#include <mutex>
#include <type_traits>
#include <string>
#include <utility>
#include <vector>

#include "tslb.h"

namespace {

// ---------------------------------------------------------------------------
// Lattice tables (TSLB_newBCS/vars_single_3D_module.f90:21-43)
// ---------------------------------------------------------------------------
constexpr int kCX[27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 0,
                         0, 1, -1, -1, 1, 1, -1, 1, -1, -1, 1, 1, -1};
constexpr int kCY[27] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 1, -1, 1,
                         -1, 0, 0, 0, 0, 1, -1, -1, 1, -1, 1, -1, 1};
constexpr int kCZ[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1,
                         1, 1, -1, 1, -1, 1, -1, 1, -1, 1, -1, -1, 1};
constexpr int kOPP[27] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14,
                          13, 16, 15, 18, 17, 20, 19, 22, 21, 24, 23, 26, 25};

__host__ __device__ constexpr float wgt(int l) {
  return l == 0 ? (float)(8.0 / 27.0)
                : (l <= 6 ? (float)(2.0 / 27.0) : (l <= 18 ? (float)(1.0 / 54.0) : (float)(1.0 / 216.0)));
}

// direction with velocity (cx,cy,cz)
__host__ __device__ constexpr int dir_of(int cx, int cy, int cz) {
  int r = -1;
  for (int l = 0; l < 27; l++)
    if (kCX[l] == cx && kCY[l] == cy && kCZ[l] == cz) r = l;
  return r;
}

// LDS slot of a direction that moves in the x-y plane (all but 0, 5, 6):
// 8 (c_z + 1) + d(c_x, c_y), d the in-plane direction index (row-major over
// c_y, c_x with (0,0) left out), so the three populations of one in-plane
// direction sit 8 slots apart and a ring source writes its entering ones
// through three per-lane addresses with constant offsets
__host__ __device__ constexpr int dslot(int cx, int cy) {
  return (cy + 1) * 3 + (cx + 1) - ((cy + 1) * 3 + (cx + 1) > 4 ? 1 : 0);
}
__host__ __device__ constexpr int qslot(int l) {
  return 8 * (kCZ[l] + 1) + dslot(kCX[l], kCY[l]);
}

// 27-bit masks of the directions whose source y - c_l lies beyond a face
__host__ __device__ constexpr uint32_t dirmask(int axis, int sign) {
  uint32_t m = 0;
  for (int l = 0; l < 27; l++) {
    int c = axis == 0 ? kCX[l] : (axis == 1 ? kCY[l] : kCZ[l]);
    if (c == sign) m |= 1u << l;
  }
  return m;
}
constexpr uint32_t kAllDirs = (1u << 27) - 2u;  // bits 1..26
constexpr uint32_t kSolid = 1u << 31;

// tile geometry of k_fused
// (build-time override: -DTSLB_TILE_X=.. -DTSLB_TILE_Y=..)
#ifndef TSLB_TILE_X
#define TSLB_TILE_X 64
#endif
#ifndef TSLB_TILE_Y
#define TSLB_TILE_Y 8
#endif
constexpr int TX = TSLB_TILE_X, TY = TSLB_TILE_Y;
static_assert(TX % 64 == 0 || (TX * TY) % 64 == 0, "blocks are whole wavefronts");
constexpr int SW = TX + 2, SH = TY + 2, NSLOT = SW * SH;
constexpr int NRING = 2 * SW + 2 * TY;
constexpr int NTHR = TX * TY;
static_assert(NRING <= NTHR && NSLOT <= 2 * NTHR, "one ring slot and at most two footprint slots per thread");
// phi footprint (two-cell halo) of the gradient stencils
constexpr int PW = TX + 4, PH = TY + 4, NPSLOT = PW * PH;
static_assert(NPSLOT <= 2 * NTHR, "at most two phi footprint slots per thread");
// The interface run list and the curvature flags are kept per half-tile (the
// left and right 32 columns of a tile-plane: a wave is one 64-node row, so a
// ballot splits at lane 32), and the curvature pass runs on half-tiles
constexpr int HX = TX / 2, HNT = HX * TY;
constexpr int HSW = HX + 2, HNSLOT = HSW * SH, HPW = HX + 4, HNPSLOT = HPW * PH;
static_assert(TX == 64, "half-tiles split a 64-lane row ballot");
static_assert(HNPSLOT <= 2 * HNT && HNSLOT <= 2 * HNT && 2 * 64 <= NTHR, "half-tile slots per thread");

struct Params {
  int nx, ny, nz;
  int px, py, pz;  // periodic flags
  int ncomp;
  int shift;       // velocity-shift forcing
  float om1;       // 1 - omega
  float sigma, sharp_c;
  float fx, fy, fz, lin_a;
  int z_chunk, tiles_x, tiles_y, nzc;
  long long fs;             // element stride between the fields of a state buffer
  int gwall;                // 1: walls absorb phase populations (jet boundary)
  float inua, inub;         // colour model: 1/nu of components A and B
  float beta;               // colour model: recolouring parameter
  float tau;                // relaxation time (learned closure: omega = 1/(tau + 3 nu_t))
  float sc[4];   // storage scale per field group, sgrp(q) (1 for FP32)
  float isc[4];  // decode scale: 1 / sc
  uint32_t step;            // fused launches so far (16-bit storage: the seed of the unbiased rounding)
  unsigned int *sat;        // FX16: count of stored values clipped at their range (device)
};


// Moment storage: FP32, FP16, BF16 or FX16 (the fused pass rounds rho - 1,
// u, v, w and phi of the 16-bit formats stochastically, everything else
// rounds to nearest even).  Reduced storage keeps every field pre-multiplied
// by a per-group scale (Params::sc) so that the shifted density and Pi^neq
// sit in FP16's normal range; compute is always FP32.  FX16 stores each
// field as a 16-bit signed fixed-point number k = round(v 2^m), |k| <=
// 32767: uniform absolute resolution 2^-m (m per field group) instead of 11
// (FP16) or 8 (BF16) significant bits.  The scales sc = 2^m and isc = 2^-m are
// powers of two, so decode and encode are exact inverses: a value that the
// step leaves unchanged is stored again as the same k, also under the
// stochastic rounding below.  (With a scale such as 32767 / R the product
// isc sc 32767 differs from 1 by ~1e-7, and unbiased rounding carried that
// bias into every step: the phase mass of the 512^3 emulsion drifted by
// 1e-3 over 20 000 steps.)
// Storage scales are shared by field groups: rho - 1; u, v, w; the six
// Pi^neq; phi (psi).  Four scales instead of eleven keep them in scalar
// registers through the fused pass (the per-field form spilled them into
// vector lanes and reloaded them every plane).
__host__ __device__ constexpr int sgrp(int q) { return q == 0 ? 0 : (q < 4 ? 1 : (q < 10 ? 2 : 3)); }

struct bf16_t {
  uint16_t b;
};
struct fx16_t {
  int16_t v;
};
__device__ __forceinline__ float to_f(float x) { return x; }
__device__ __forceinline__ float to_f(_Float16 x) { return (float)x; }
__device__ __forceinline__ float to_f(bf16_t x) { return __uint_as_float((uint32_t)x.b << 16); }
// FX16 decode: one convert that sign-extends the low 16 bits of the
// zero-extended load (SDWA).  Written out because the compiler folds a sign
// extension into the load (buffer_load_sshort) at its point of use, and the
// raw loads of one source were then issued one at a time, each waited out
// before the next.  Code 0 is 0, so a load past the buffer's end reads 0.
__device__ __forceinline__ float to_f(fx16_t x) {
  float f;
  asm("v_cvt_f32_i32_sdwa %0, sext(%1) dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:WORD_0"
      : "=v"(f)
      : "v"((uint32_t)(uint16_t)x.v));
  return f;
}
template <typename S>
__device__ __forceinline__ S from_f(float v);
template <>
__device__ __forceinline__ float from_f<float>(float v) { return v; }
template <>
__device__ __forceinline__ _Float16 from_f<_Float16>(float v) { return (_Float16)v; }
template <>
__device__ __forceinline__ bf16_t from_f<bf16_t>(float v) {
  uint32_t u = __float_as_uint(v);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return bf16_t{(uint16_t)(u >> 16)};
}
// FP16 and BF16 with stochastic rounding (v_cvt_sr_f16_f32 and
// v_cvt_sr_bf16_f32): the random bits r decide between the two neighbours
// in proportion to the distances, so the stored value is unbiased.
typedef _Float16 f16x2_t __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
template <typename S>
__device__ S from_f_sr(float v, uint32_t r);
template <>
__device__ __forceinline__ _Float16 from_f_sr<_Float16>(float v, uint32_t r) {
  return __builtin_amdgcn_cvt_sr_f16_f32(f16x2_t{}, v, r, false)[0];
}
template <>
__device__ __forceinline__ bf16_t from_f_sr<bf16_t>(float v, uint32_t r) {
  return bf16_t{__builtin_bit_cast(uint16_t, __builtin_amdgcn_cvt_sr_bf16_f32(bf16x2_t{}, v, r, false)[0])};
}
// FX16 encode by the magic number C = 1.5 2^23: for |y| < 2^22 the float
// y + C lies in [2^23, 2^24), where the spacing is 1, so the one rounding of
// the sum is round-to-nearest-even of y, and its low 16 bits are k in two's
// complement.  No clamp: |y| < 32767.5 is the caller's to ensure.
constexpr float kFx16Magic = 12582912.f;
__device__ __forceinline__ fx16_t fx16_code(float t) {
  return fx16_t{(int16_t)(uint16_t)__float_as_uint(t + kFx16Magic)};
}
// round to nearest, clamped to +-32767
template <>
__device__ __forceinline__ fx16_t from_f<fx16_t>(float v) {
  return fx16_code(__builtin_amdgcn_fmed3f(v, -32767.f, 32767.f));
}
// FX16 with unbiased (stochastic) rounding, in integers: Y = y 2^16 rounded
// to nearest (exact once |y| >= 128), k = floor((Y + w) / 2^16) for a word
// w uniform in [0, 2^16): k exceeds floor(Y / 2^16) with probability
// (Y mod 2^16) / 2^16, so the mean of k is Y / 2^16.  Round-to-nearest
// alone is biased wherever a field is nearly uniform (every node of a
// resting phase rounds the same way), which drifts the conserved sums, and
// it drops every per-step change below half a step (a slowly forced flow
// stagnates); the fused pass stores rho - 1, u, v, w and phi this way.  (The
// same in floats, round(y - 1/2 + (w + 1/2) 2^-16), loses the dither's low
// bits in the sum once |y| is large: a bias of ~2.5e-4 steps at phi ~ 1,
// which drifted the emulsion's phase mass by 1e-4 over 20 000 steps.)
// |y| <= 32766 keeps Y + w within int32 and k within +-32767.
__device__ __forceinline__ fx16_t fx16_dither(float y, uint32_t w16) {
  const int Y = (int)rintf(y * 65536.f);
  return fx16_t{(int16_t)((Y + (int)w16) >> 16)};
}
// The random words of the unbiased rounding: a hash of node and step.
// Two properties matter.  (1) The words must average to uniform over the
// lattice step by step: with the seed XORed into the node index, the set of
// hashed values over a power-of-two lattice is the same every step, and the
// small deviation of its mean from uniform (~1e-5) repeated every step (a
// phase-mass drift of 1e-3 over 20 000 steps of the 512^3 emulsion); adding
// the seed moves the window of hashed values instead.  (2) A node's words
// must be unrelated from one step to the next: with one multiplication
// (node + step K) times a constant, the high half moved by a constant each
// step, and the rounding of u, fed back through the flow, drifted a uniform
// flow by 0.03 steps per step.  mix32 is the two-round multiply-xorshift
// finaliser "lowbias32" (C. Wellons, hash-prospector), applied to the step
// (a scalar) and to node + seed.
__host__ __device__ constexpr uint32_t mix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x7FEB352Du;
  h ^= h >> 15;
  h *= 0x846CA68Bu;
  h ^= h >> 16;
  return h;
}
__device__ __forceinline__ uint32_t node_hash(uint32_t node, uint32_t step) { return mix32(node + mix32(step)); }
// 32 more bits from a mixed word
__device__ __forceinline__ uint32_t node_rehash(uint32_t h) {
  h = (h ^ (h >> 15)) * 0x2C1B3C6Du;
  return h ^ (h >> 12);
}
// the fields that the fused pass stores with unbiased rounding (every
// 16-bit format) and the hash word each takes
__host__ __device__ constexpr bool dithered(int q) { return q < 4 || q == TSLB_PHI; }
__host__ __device__ constexpr int dither_word(int q) { return q < 4 ? q : 4; }
__device__ __forceinline__ uint32_t rot16(uint32_t h) { return (h << 16) | (h >> 16); }

// a stored value of field group g, decoded
template <typename S>
__device__ __forceinline__ float dec(S x, const Params &p, int g) {
  if constexpr (sizeof(S) == 4) return to_f(x);
  else return to_f(x) * p.isc[g];
}

template <typename S>
struct FieldsT {
  S *f[TSLB_NFIELDS];
};
template <typename S>
struct CFieldsT {
  const S *f[TSLB_NFIELDS];
};

// scaled field access
template <typename S>
__device__ __forceinline__ float ldf(const S *__restrict__ a, int id, const Params &p, int q) {
  return dec(a[id], p, sgrp(q));
}
template <typename S>
__device__ __forceinline__ void stf(S *__restrict__ a, int id, float v, const Params &p, int q) {
  if constexpr (sizeof(S) == 4) {
    a[id] = from_f<S>(v);
  } else {
    if constexpr (std::is_same<S, fx16_t>::value) {
      if (!(fabsf(v * p.sc[sgrp(q)]) <= 32767.5f) && p.sat) atomicAdd(p.sat, 1u);
    }
    a[id] = from_f<S>(v * p.sc[sgrp(q)]);
  }
}

__device__ __forceinline__ int wrapc(int g, int n) {
  int r = g % n;
  return r < 0 ? r + n : r;
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2.  Map dispatch order to tiles so that consecutive tiles (which share
// halo rows and columns) run on the same XCD at about the same time: XCD x
// gets the contiguous tile range [x*nblk/8, (x+1)*nblk/8).
__device__ __forceinline__ int xcd_tile(int b, int nblk) {
  constexpr int NXCD = 8;
  const int per = nblk / NXCD;
  if (b >= per * NXCD) return b;  // remainder blocks keep their order
  return (b % NXCD) * per + b / NXCD;
}

// coordinate of a source along one axis; false beyond a non-periodic face
__device__ __forceinline__ bool src_coord(int g, int n, int per, int &out) {
  if (g >= 0 && g < n) {
    out = g;
    return true;
  }
  if (!per) return false;
  out = wrapc(g, n);
  return true;
}

template <int C>
__device__ __forceinline__ float cm(float a) {
  if constexpr (C == 0) return 0.f;
  else if constexpr (C > 0) return a;
  else return -a;
}

// ---------------------------------------------------------------------------
// Post-collision populations of one node, shifted by the weights:
//   h_l = f*_l - w_l,  f*_l = feq_l(rho,u_eq) + (1-omega) w_l a^neq_l(Pi,u) [+F]
// (recursive_2c_TSLB3D.f90:637-806, recursive_TSLB3D.f90:556-707).
// ---------------------------------------------------------------------------
// Destination of the populations of one source: directions that move in the
// x-y plane go to the LDS slot of the source, 0, 5 and 6 stay in registers.
struct PopSink {
  float (*lds)[NSLOT];
  int slot;
  float h0, h5, h6;
  template <int L>
  __device__ __forceinline__ void put(float v) {
    if constexpr (L == 0) h0 = v;
    else if constexpr (L == 5) h5 = v;
    else if constexpr (L == 6) h6 = v;
    else lds[qslot(L)][slot] = v;
  }
  // phase populations 1..4 (in-plane movers) go to LDS slots 24..27
  template <int G>
  __device__ __forceinline__ void putg(float v) {
    lds[23 + G][slot] = v;
  }
};
// All 27 (+7) populations of one node into registers (boundary kernels).
struct ArraySink {
  float h[27];
  float g[7];
  template <int L>
  __device__ __forceinline__ void put(float v) {
    h[L] = v;
  }
  template <int G>
  __device__ __forceinline__ void putg(float v) {
    g[G] = v;
  }
};

// Shifted post-collision populations by the D3Q27 tensor-product transform.
// Raw moments M_abc = sum_l cx^a cy^b cz^c h_l (a,b,c in {0,1,2}) of
// h = f* - w follow from the Hermite coefficients of f*: a0 = rho,
// a1 = rho V (+F), a2 = rho VV + (1-omega) Pi, a3 = rho VVV + (1-omega) sym(u Pi)
// on the seven D3Q27 third-order components, zero above; with cs^2 = 1/3:
//   M_2.. = a2_xx + cs^2 a0 (and products of such factors per axis).
// Each M_abc is stored pre-multiplied by 2^-(number of nonzero exponents) so
// that the 1D inverse (m0, m1, m2) -> f(-1) = m2 - m1, f(0) = m0 - 2 m2,
// f(+1) = m2 + m1 costs three operations per triple.
// The raw moments M[a][b][c] (scaled as above) of one source from its
// moments (rho - 1, u, Pi), the equilibrium velocity V and the additive force.
template <bool ADD>
__device__ __forceinline__ void raw_moments(float drho, float ux, float uy, float uz, float Pxx, float Pyy,
                                            float Pzz, float Pxy, float Pxz, float Pyz, float Vx, float Vy,
                                            float Vz, float fx, float fy, float fz, float om1,
                                            float (&M)[3][3][3]) {
  const float rho = 1.f + drho;
  const float Bxx = om1 * Pxx, Byy = om1 * Pyy, Bzz = om1 * Pzz;
  const float Bxy = om1 * Pxy, Bxz = om1 * Pxz, Byz = om1 * Pyz;
  const float rVx = rho * Vx, rVy = rho * Vy, rVz = rho * Vz;
  float a1x = rVx, a1y = rVy, a1z = rVz;
  if constexpr (ADD) {
    a1x += fx;
    a1y += fy;
    a1z += fz;
  }
  const float rVxy = rVx * Vy, rVxz = rVx * Vz, rVyz = rVy * Vz;
  const float a2xx = fmaf(rVx, Vx, Bxx), a2yy = fmaf(rVy, Vy, Byy), a2zz = fmaf(rVz, Vz, Bzz);
  const float a2xy = rVxy + Bxy, a2xz = rVxz + Bxz, a2yz = rVyz + Byz;
  // rho Va Va Vb taken as (rho Va Vb) Va
  const float a3xxy = fmaf(rVxy, Vx, fmaf(uy, Bxx, 2.f * ux * Bxy));
  const float a3xxz = fmaf(rVxz, Vx, fmaf(uz, Bxx, 2.f * ux * Bxz));
  const float a3xyy = fmaf(rVxy, Vy, fmaf(ux, Byy, 2.f * uy * Bxy));
  const float a3yyz = fmaf(rVyz, Vy, fmaf(uz, Byy, 2.f * uy * Byz));
  const float a3xzz = fmaf(rVxz, Vz, fmaf(ux, Bzz, 2.f * uz * Bxz));
  const float a3yzz = fmaf(rVyz, Vz, fmaf(uy, Bzz, 2.f * uz * Byz));
  const float a3xyz = fmaf(rVxy, Vz, fmaf(ux, Byz, fmaf(uy, Bxz, uz * Bxy)));
  constexpr float c3 = 1.f / 3.f, c9 = 1.f / 9.f;
  // M[a][b][c], scaled by 2^-(#nonzero)
  M[0][0][0] = drho;
  M[1][0][0] = 0.5f * a1x;
  M[0][1][0] = 0.5f * a1y;
  M[0][0][1] = 0.5f * a1z;
  M[2][0][0] = 0.5f * fmaf(c3, drho, a2xx);
  M[0][2][0] = 0.5f * fmaf(c3, drho, a2yy);
  M[0][0][2] = 0.5f * fmaf(c3, drho, a2zz);
  M[1][1][0] = 0.25f * a2xy;
  M[1][0][1] = 0.25f * a2xz;
  M[0][1][1] = 0.25f * a2yz;
  M[2][1][0] = 0.25f * fmaf(c3, a1y, a3xxy);
  M[2][0][1] = 0.25f * fmaf(c3, a1z, a3xxz);
  M[1][2][0] = 0.25f * fmaf(c3, a1x, a3xyy);
  M[0][2][1] = 0.25f * fmaf(c3, a1z, a3yyz);
  M[1][0][2] = 0.25f * fmaf(c3, a1x, a3xzz);
  M[0][1][2] = 0.25f * fmaf(c3, a1y, a3yzz);
  M[1][1][1] = 0.125f * a3xyz;
  // products of per-axis factors from the scaled entries above:
  //   M220 = 1/4 (c9 drho + c3 (a2xx + a2yy)) = c3/2 (M200 + M020) - c9/4 drho,
  //   M222 = 1/8 (c27 drho + c9 (a2xx + a2yy + a2zz)) = c9/4 (M200 + M020 + M002) - drho/108,
  //   M221 = 1/8 (c9 a1z + c3 (a3xxz + a3yyz)) = c3/2 (M201 + M021) - c9/4 M001
  const float e9 = (0.25f * c9) * drho;
  const float s200_020 = M[2][0][0] + M[0][2][0];
  M[2][2][0] = fmaf(0.5f * c3, s200_020, -e9);
  M[2][0][2] = fmaf(0.5f * c3, M[2][0][0] + M[0][0][2], -e9);
  M[0][2][2] = fmaf(0.5f * c3, M[0][2][0] + M[0][0][2], -e9);
  M[2][1][1] = (0.125f * c3) * a2yz;
  M[1][2][1] = (0.125f * c3) * a2xz;
  M[1][1][2] = (0.125f * c3) * a2xy;
  M[2][2][1] = fmaf(0.5f * c3, M[2][0][1] + M[0][2][1], -(0.25f * c9) * M[0][0][1]);
  M[2][1][2] = fmaf(0.5f * c3, M[2][1][0] + M[0][1][2], -(0.25f * c9) * M[0][1][0]);
  M[1][2][2] = fmaf(0.5f * c3, M[1][2][0] + M[1][0][2], -(0.25f * c9) * M[1][0][0]);
  M[2][2][2] = fmaf(0.25f * c9, s200_020 + M[0][0][2], drho * (-1.f / 108.f));
}

template <bool ADD, typename Sink>
__device__ __forceinline__ void all_pops_mt(const float *m, float Vx, float Vy, float Vz,
                                            const Params &p, Sink &h, float om1) {
  float M[3][3][3];
  raw_moments<ADD>(m[0], m[1], m[2], m[3], m[4], m[5], m[6], m[7], m[8], m[9], Vx, Vy, Vz, p.fx, p.fy, p.fz,
                   om1, M);
  // pass z: G[a][b][cz]
  float G[3][3][3];
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) {
      const float m0 = M[a][b][0], m1 = M[a][b][1], m2 = M[a][b][2];
      G[a][b][0] = m2 - m1;
      G[a][b][1] = fmaf(-2.f, m2, m0);
      G[a][b][2] = m2 + m1;
    }
  // pass y: K[a][cy][cz]
  float K[3][3][3];
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int cz = 0; cz < 3; cz++) {
      const float m0 = G[a][0][cz], m1 = G[a][1][cz], m2 = G[a][2][cz];
      K[a][0][cz] = m2 - m1;
      K[a][1][cz] = fmaf(-2.f, m2, m0);
      K[a][2][cz] = m2 + m1;
    }
  // pass x and store
#define MT_ROW(CY, CZ)                                              \
  {                                                                 \
    const float m0 = K[0][CY + 1][CZ + 1], m1 = K[1][CY + 1][CZ + 1]; \
    const float m2 = K[2][CY + 1][CZ + 1];                           \
    h.template put<dir_of(-1, CY, CZ)>(m2 - m1);                     \
    h.template put<dir_of(0, CY, CZ)>(fmaf(-2.f, m2, m0));          \
    h.template put<dir_of(1, CY, CZ)>(m2 + m1);                      \
  }
  MT_ROW(-1, -1) MT_ROW(-1, 0) MT_ROW(-1, 1)
  MT_ROW(0, -1) MT_ROW(0, 0) MT_ROW(0, 1)
  MT_ROW(1, -1) MT_ROW(1, 0) MT_ROW(1, 1)
#undef MT_ROW
}


// A ring source of k_fused (one node outside the tile's 64x8 columns) sends
// only the 9 populations with c_P = sgn into the tile, P the axis normal to
// its side of the ring (y for the rows above and below the tile, x for the
// columns beside it).  Those 9 take the P pass at c_P = sgn first (one
// operation per (a, c) pair), then the z pass and the in-plane pass along the
// other axis Q on three triples: 27 operations instead of the 81 of all 27.
// Columns run the same arithmetic with x and y exchanged in the inputs, so
// the M[a][b][c] below carry Q, P, z exponents.  rq[q + 1] addresses the LDS
// slot of the c_z = -1 population moving q along Q; c_z = 0 and +1 sit 8 and
// 16 slots further (qslot).
// Adj: adj(q, cz, v) returns the stored population of the direction with
// c_Q = q, c_z = cz (the colour model adds its perturbation there).
struct NoAdj {
  __device__ __forceinline__ float operator()(int, int, float v) const { return v; }
};
template <bool ADD, typename Adj = NoAdj>
__device__ __forceinline__ void ring_pops(const float *m, float Vx, float Vy, float Vz, const Params &p,
                                          float om1, bool col, float sgn, float *const (&rq)[3],
                                          const Adj &adj = Adj{}) {
  const float uq = col ? m[2] : m[1], up = col ? m[1] : m[2];
  const float Pqq = col ? m[5] : m[4], Ppp = col ? m[4] : m[5];
  const float Pqz = col ? m[9] : m[8], Ppz = col ? m[8] : m[9];
  const float Vq = col ? Vy : Vx, Vp = col ? Vx : Vy;
  float fq = 0.f, fp = 0.f;
  if constexpr (ADD) {
    fq = col ? p.fy : p.fx;
    fp = col ? p.fx : p.fy;
  }
  float M[3][3][3];
  raw_moments<ADD>(m[0], uq, up, m[3], Pqq, Ppp, m[6], m[7], Pqz, Ppz, Vq, Vp, Vz, fq, fp, p.fz, om1, M);
  float K[3][3];
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int c = 0; c < 3; c++) K[a][c] = fmaf(sgn, M[a][1][c], M[a][2][c]);
  float G[3][3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const float m0 = K[a][0], m1 = K[a][1], m2 = K[a][2];
    G[a][0] = m2 - m1;
    G[a][1] = fmaf(-2.f, m2, m0);
    G[a][2] = m2 + m1;
  }
#pragma unroll
  for (int cz = 0; cz < 3; cz++) {
    const float m0 = G[0][cz], m1 = G[1][cz], m2 = G[2][cz];
    rq[0][cz * 8 * NSLOT] = adj(-1, cz - 1, m2 - m1);
    rq[1][cz * 8 * NSLOT] = adj(0, cz - 1, fmaf(-2.f, m2, m0));
    rq[2][cz * 8 * NSLOT] = adj(1, cz - 1, m2 + m1);
  }
}

// Colour model, ring sources: the entering direction with c_Q = q, c_P =
// sgn, c_z = cz, its weight, 1/|c| and c.g (the ring side's P axis is y for
// rows, x for columns).
struct RingDir {
  bool col;
  float sgn, gx, gy, gz;
  __device__ __forceinline__ float cg(int q, int cz) const {
    const float cx = col ? sgn : (float)q, cy = col ? (float)q : sgn;
    return fmaf(cx, gx, fmaf(cy, gy, (float)cz * gz));
  }
  // c_P != 0: one, two or three nonzero components
  __device__ __forceinline__ static float w(int q, int cz) {
    const int nnz = 1 + (q != 0) + (cz != 0);
    return nnz == 1 ? 2.f / 27.f : (nnz == 2 ? 1.f / 54.f : 1.f / 216.f);
  }
  __device__ __forceinline__ static float inv(int q, int cz) {
    const int nnz = 1 + (q != 0) + (cz != 0);
    return nnz == 1 ? 1.f : (nnz == 2 ? 0.70710678118654752f : 0.57735026918962576f);
  }
};
// the colour-gradient perturbation of ColourSink on one entering population
struct ColourRingAdj {
  RingDir d;
  float K, g2;
  __device__ __forceinline__ float operator()(int q, int cz, float v) const {
    const float cg = d.cg(q, cz);
    return fmaf(K, RingDir::w(q, cz) * fmaf(cg, cg, -g2), v);
  }
};

// ---------------------------------------------------------------------------
// Colour-gradient stencil, grouped as the reference (recursive_2c_TSLB3D.f90:
// 555-566): 3*(p1*(face+ - face-) + p2*(edges+ - edges-) + p3*(corners+ -
// corners-)) along axis A, directions visited in l order.
// ---------------------------------------------------------------------------
template <int A, typename Get>
__device__ __forceinline__ float grad_axis(const Get &get) {
  float fp = 0.f, fm = 0.f, ep = 0.f, em = 0.f, cp = 0.f, cmn = 0.f;
#pragma unroll
  for (int l = 1; l < 27; l++) {
    const int ca = A == 0 ? kCX[l] : (A == 1 ? kCY[l] : kCZ[l]);
    if (ca == 0) continue;
    const int nnz = (kCX[l] != 0) + (kCY[l] != 0) + (kCZ[l] != 0);
    const float val = get(kCX[l], kCY[l], kCZ[l]);
    if (nnz == 1) {
      if (ca > 0) fp = val; else fm = val;
    } else if (nnz == 2) {
      if (ca > 0) ep += val; else em += val;
    } else {
      if (ca > 0) cp += val; else cmn += val;
    }
  }
  return 3.f * (wgt(1) * (fp - fm) + wgt(7) * (ep - em) + wgt(19) * (cp - cmn));
}

// Unit normal n = grad(phi)/(|grad(phi)| + 1e-9) (recursive_2c_TSLB3D.f90:
// 555-571) from the same stencil, separated.  Its weights 2/27, 1/54 and
// 1/216 (faces, edges, corners) are W(dx) W(dy) W(dz) / 216 with W(0) = 4,
// W(+-1) = 1, so with the in-plane sums of plane z (centre value pc(z))
//   Px(z) = sum_dy W(dy) (phi(x+1, y+dy, z) - phi(x-1, y+dy, z)),  Py alike,
//   Ps(z) = sum_dx,dy W(dx) W(dy) (phi(x+dx, y+dy, z) - pc(z))
// the gradient is (3/216) (Px(z-1) + 4 Px(z) + Px(z+1), Py(..),
// Ps(z+1) - Ps(z-1) + 36 (pc(z+1) - pc(z-1))).  Every term is a difference
// of nearby phi before any weighted sum, so the sums carry no cancellation:
// in a bulk phase (|grad phi| ~ 1e-4, phi ~ 1) the FP32 normal stays within
// ~1e-6 of the FP64 one, where summing each side of the stencil first (as
// the reference does) loses ~1e-4.  A march along z forms Px, Py, Ps once
// per plane from the 3x3 in-plane neighbourhood (centre c of phi planes of
// row pitch PITCH) and carries the z-combination in registers.
struct PlaneGrad {
  float x, y, s, c;
};
template <int PITCH = PW>
__device__ __forceinline__ PlaneGrad plane_grad(const float *pl, int c) {
  const float a = pl[c - PITCH - 1], b = pl[c - PITCH], d = pl[c - PITCH + 1];
  const float e = pl[c - 1], f = pl[c], g = pl[c + 1];
  const float h = pl[c + PITCH - 1], i = pl[c + PITCH], j = pl[c + PITCH + 1];
  PlaneGrad r;
  r.x = (d - a) + fmaf(4.f, g - e, j - h);
  r.y = (h - a) + fmaf(4.f, i - b, j - d);
  const float cm = fmaf(4.f, e - f, (a - f) + (h - f)), c0 = (b - f) + (i - f);
  const float cp = fmaf(4.f, g - f, (d - f) + (j - f));
  r.s = fmaf(4.f, c0, cm + cp);
  r.c = f;
  return r;
}
// running z-combination of one column: rx = Px(z-1) + 4 Px(z), lx = Px(z)
// (and y alike), sm = Ps(z-1), sz = Ps(z)
struct GradRun {
  float rx, lx, ry, ly, sm, sz;
};
// a, b: the sums of planes z-1 and z
__device__ __forceinline__ void grad_run_init(GradRun &g, const PlaneGrad &a, const PlaneGrad &b) {
  g.rx = fmaf(4.f, b.x, a.x);
  g.lx = b.x;
  g.ry = fmaf(4.f, b.y, a.y);
  g.ly = b.y;
  g.sm = a.s;
  g.sz = b.s;
}
// the unit normal (NORM) or the gradient of plane z from the sums of plane
// z+1 and the centre value of plane z-1 (pcm), and the advance to z+1; the
// sums give 72 grad(phi), so the 1e-9 floor of the normal becomes 7.2e-8
template <bool NORM = true>
__device__ __forceinline__ void grad_run_step(GradRun &g, const PlaneGrad &n, float pcm, float *nn) {
  const float gx = g.rx + n.x, gy = g.ry + n.y, gz = fmaf(36.f, n.c - pcm, n.s - g.sm);
  if constexpr (NORM) {
    const float inv = __builtin_amdgcn_rcpf(__builtin_amdgcn_sqrtf(fmaf(gx, gx, fmaf(gy, gy, gz * gz))) + 7.2e-8f);
    nn[0] = gx * inv;
    nn[1] = gy * inv;
    nn[2] = gz * inv;
  } else {
    constexpr float k = 1.f / 72.f;
    nn[0] = gx * k;
    nn[1] = gy * k;
    nn[2] = gz * k;
  }
  g.rx = fmaf(4.f, n.x, g.lx);
  g.lx = n.x;
  g.ry = fmaf(4.f, n.y, g.ly);
  g.ly = n.y;
  g.sm = g.sz;
  g.sz = n.s;
}
// the unit normal (NORM) or the gradient at centre c from phi planes z-1, z,
// z+1: the operations of the march in the same order, so every pass gets
// the same bits
template <int PITCH = PW, bool NORM = true>
__device__ __forceinline__ void sep_normal(const float *P0, const float *P1, const float *P2, int c,
                                           float &n0, float &n1, float &n2) {
  GradRun g;
  grad_run_init(g, plane_grad<PITCH>(P0, c), plane_grad<PITCH>(P1, c));
  float nn[3];
  grad_run_step<NORM>(g, plane_grad<PITCH>(P2, c), P0[c], nn);
  n0 = nn[0];
  n1 = nn[1];
  n2 = nn[2];
}

// phi stencil coordinate along one axis: wrapped across a periodic face,
// the nearest interior node across a wall
__device__ __forceinline__ int clampw(int g, int n, int per) {
  if (g >= 0 && g < n) return g;
  return per ? wrapc(g, n) : (g < 0 ? 0 : n - 1);
}

// ---------------------------------------------------------------------------
// Source evaluation for k_fused: loads the moments of one node and returns its
// shifted post-collision populations (and phase populations).
// ---------------------------------------------------------------------------
template <int NC, bool CLO = false>
struct Src {
  float m[NC >= 2 ? 11 : 10];  // drho,u,v,w,pxx,pyy,pzz,pxy,pxz,pyz(,phi or psi)
  float nn[3];                 // normals (phase field) or colour gradient (colour model)
  float kap;                   // gated curvature (phase field)
  float clo[CLO ? 7 : 1];      // learned closure: nu_t, dsigma (xx,yy,zz,xy,xz,yz)
};

// Raw buffer access.  A descriptor (gfx9 layout: 32-bit data format, no
// stride) covers one plane of one field, so a node is a loop-invariant 32-bit
// byte offset from a scalar base, with no per-access 64-bit address
// arithmetic; an access past the plane reads 0 and a store there is dropped.
constexpr int kRsrcWord3 = 0x00020000;
typedef __amdgpu_buffer_rsrc_t rsrc_t;
__device__ __forceinline__ rsrc_t rsrc(const void *base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes, kRsrcWord3);
}
template <typename S>
__device__ __forceinline__ S bld(rsrc_t r, uint32_t off);
template <>
__device__ __forceinline__ float bld<float>(rsrc_t r, uint32_t off) {
  return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
}
template <>
__device__ __forceinline__ _Float16 bld<_Float16>(rsrc_t r, uint32_t off) {
  return __builtin_bit_cast(_Float16, __builtin_amdgcn_raw_buffer_load_b16(r, (int)off, 0, 0));
}
template <>
__device__ __forceinline__ bf16_t bld<bf16_t>(rsrc_t r, uint32_t off) {
  return bf16_t{__builtin_amdgcn_raw_buffer_load_b16(r, (int)off, 0, 0)};
}
template <>
__device__ __forceinline__ fx16_t bld<fx16_t>(rsrc_t r, uint32_t off) {
  return fx16_t{(int16_t)__builtin_amdgcn_raw_buffer_load_b16(r, (int)off, 0, 0)};
}
// CP: cache-policy bits of the store; 2 (nt) marks it non-temporal: the
// fused pass writes the next state once and reads none of it back, and its
// stores then leave the L2 to the halo rows the neighbouring tiles re-read
template <int CP = 0>
__device__ __forceinline__ void bst(rsrc_t r, uint32_t off, float v) {
  __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, (int)off, 0, CP);
}
template <int CP = 0>
__device__ __forceinline__ void bst(rsrc_t r, uint32_t off, _Float16 v) {
  __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, v), r, (int)off, 0, CP);
}
template <int CP = 0>
__device__ __forceinline__ void bst(rsrc_t r, uint32_t off, bf16_t v) {
  __builtin_amdgcn_raw_buffer_store_b16(v.b, r, (int)off, 0, CP);
}
template <int CP = 0>
__device__ __forceinline__ void bst(rsrc_t r, uint32_t off, fx16_t v) {
  __builtin_amdgcn_raw_buffer_store_b16((unsigned short)v.v, r, (int)off, 0, CP);
}
__device__ __forceinline__ uint32_t bld_u32(rsrc_t r, uint32_t off) {
  return __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0);
}
__device__ __forceinline__ uint8_t bld_u8(rsrc_t r, uint32_t off) {
  return __builtin_amdgcn_raw_buffer_load_b8(r, (int)off, 0, 0);
}

template <typename T>
__device__ __forceinline__ T ld_off(const T *__restrict__ base, uint32_t byte_off) {
  return *(const T *)((const char *)base + byte_off);
}
template <typename T>
__device__ __forceinline__ void st_off(T *__restrict__ base, uint32_t byte_off, T v) {
  *(T *)((char *)base + byte_off) = v;
}

// One source plane gz: byte bases of the plane in the state buffer (field 0)
// and in the normals buffer (n0); field q sits q field strides further.
struct PlaneSrc {
  const char *f;      // state plane, field 0
  const char *n;      // normals plane, n0
  const char *c;      // closure plane, nu_t
  uint32_t pbS, pb4;  // plane bytes (storage type, float)
  long long fsS, fs4; // field stride in bytes (storage type, float)
};

// Loads of one source node.  offS / off4 are in-plane byte offsets; off4 is
// kOOB where the tile-plane has no interface (the buffer then returns 0
// without a memory access).
constexpr uint32_t kOOB = 0x80000000u;
// The ten hydrodynamic moments of a source are loaded; phi (psi) of the two
// component models comes from the fused pass's LDS ring instead.
// Reduced storage: the raw values of one source, loaded before any is
// converted (left to itself the scheduler reuses one register per load and
// waits out each round trip).
template <int NC, typename S>
struct RawSrc {
  S v[10];
};
template <int NC, typename S>
__device__ __forceinline__ void load_raw(const PlaneSrc &ps, uint32_t offS, RawSrc<NC, S> &r) {
#pragma unroll
  for (int q = 0; q < 10; q++) r.v[q] = bld<S>(rsrc(ps.f + q * ps.fsS, ps.pbS), offS);
}
template <int NC, typename S, bool CLO>
__device__ __forceinline__ void convert_raw(const RawSrc<NC, S> &r, const Params &p, Src<NC, CLO> &s) {
#pragma unroll
  for (int q = 0; q < 10; q++) s.m[q] = dec(r.v[q], p, sgrp(q));
}

// FP32 storage: the moments of one source straight into place.
template <int NC, bool CLO>
__device__ __forceinline__ void load_m32(const PlaneSrc &ps, uint32_t offS, Src<NC, CLO> &s) {
#pragma unroll
  for (int q = 0; q < 10; q++) s.m[q] = bld<float>(rsrc(ps.f + q * ps.fsS, ps.pbS), offS);
}

// The other loads of one source: gated curvature, closure.
template <int NC, bool CLO>
__device__ __forceinline__ void load_extra(const PlaneSrc &ps, uint32_t off4, uint32_t offC, Src<NC, CLO> &s) {
  // phase field: the gated curvature (the normals are computed by the caller)
  if constexpr (NC == 2) s.kap = bld<float>(rsrc(ps.n + 3 * ps.fs4, ps.pb4), off4);
  if constexpr (CLO) {
#pragma unroll
    for (int q = 0; q < 7; q++) s.clo[q] = bld<float>(rsrc(ps.c + q * ps.fs4, ps.pb4), offC);
  }
}

// Learned closure in the collision: omega = 1/(tau + 3 nu_t) and the stress
// correction dsigma added to the post-collision non-equilibrium stress, so
// that the populations are built from Pi* = (1 - omega) Pi + dsigma (and its
// third-order terms).  Returns the moments to build from and their 1 - omega.
template <int NC, bool CLO>
__device__ __forceinline__ const float *closure_moments(const Src<NC, CLO> &s, const Params &p,
                                                        float (&buf)[10], float &om1) {
  if constexpr (!CLO) {
    om1 = p.om1;
    return s.m;
  } else {
    const float o1 = 1.f - __builtin_amdgcn_rcpf(fmaf(3.f, s.clo[0], p.tau));  // one ulp
#pragma unroll
    for (int q = 0; q < 4; q++) buf[q] = s.m[q];
#pragma unroll
    for (int q = 4; q < 10; q++) buf[q] = fmaf(o1, s.m[q], s.clo[q - 3]);
    om1 = 1.f;
    return buf;
  }
}

// Colour-gradient perturbation added to each population on its way to the
// sink (turbulent_benchmark/couette/recursive_colorTSLB3D.f90:764-956, in its
// momentum-conserving quadratic form):
//   Omega_l = K [w_l ((c_l.g)^2 - |g|^2) + (2/3) |g|^2 delta_l0],
// K = A/(2|g|), A = (9/2) sigma omega, i.e. (5/27) A |g| at rest and
// A ((c.g)^2 - |g|^2)/(k_l |g|) with k_l = 27, 108, 432; K = 0 below the
// |g| > 0.001 threshold.
template <typename Inner>
struct ColourSink {
  Inner &in;
  // K w_l ((c_l.g)^2 - |g|^2) needs only 13 distinct (c.g)^2 - |g|^2, all
  // from the squares and pair products of g: faces -(g_b^2 + g_c^2), edges
  // -g_c^2 +- 2 g_a g_b, corners 2 (sx sy gx gy + sx sz gx gz + sy sz gy gz)
  // (4 sign patterns up to a common sign), the rest (10/27) |g|^2
  float kf, ke, kc, k0;  // K w by class, and K (10/27) |g|^2
  float fa[3];           // faces x, y, z
  float ep[3], em[3];    // edges xy, xz, yz with sa sb = +1, -1
  float cn[4];           // corners (+,+,+), (+,+,-), (+,-,+), (-,+,+)
  __device__ __forceinline__ ColourSink(Inner &i, float K, float gx, float gy, float gz, float g2) : in(i) {
    const float x2 = gx * gx, y2 = gy * gy, z2 = gz * gz;
    const float pxy = 2.f * gx * gy, pxz = 2.f * gx * gz, pyz = 2.f * gy * gz;
    kf = K * (2.f / 27.f);
    ke = K * (1.f / 54.f);
    kc = K * (1.f / 216.f);
    k0 = K * (10.f / 27.f) * g2;
    fa[0] = -(y2 + z2);
    fa[1] = -(x2 + z2);
    fa[2] = -(x2 + y2);
    ep[0] = pxy - z2;
    em[0] = -(pxy + z2);
    ep[1] = pxz - y2;
    em[1] = -(pxz + y2);
    ep[2] = pyz - x2;
    em[2] = -(pyz + x2);
    cn[0] = pxy + pxz + pyz;
    cn[1] = pxy - (pxz + pyz);
    cn[2] = pxz - (pxy + pyz);
    cn[3] = pyz - (pxy + pxz);
  }
  template <int L>
  __device__ __forceinline__ void put(float v) {
    constexpr int cx = kCX[L], cy = kCY[L], cz = kCZ[L];
    constexpr int nnz = (cx != 0) + (cy != 0) + (cz != 0);
    float r;
    if constexpr (nnz == 0) {
      r = v + k0;
    } else if constexpr (nnz == 1) {
      r = fmaf(kf, fa[cx != 0 ? 0 : (cy != 0 ? 1 : 2)], v);
    } else if constexpr (nnz == 2) {
      constexpr int pr = cz == 0 ? 0 : (cy == 0 ? 1 : 2);  // xy, xz, yz
      constexpr int sab = (cz == 0 ? cx * cy : (cy == 0 ? cx * cz : cy * cz));
      r = fmaf(ke, sab > 0 ? ep[pr] : em[pr], v);
    } else {
      // sign pattern up to a common sign: (sx sy, sx sz) = (+,+) -> 0,
      // (+,-) -> 1, (-,+) -> 2, (-,-) -> 3
      constexpr int sxy = cx * cy, sxz = cx * cz;
      constexpr int k = sxy > 0 ? (sxz > 0 ? 0 : 1) : (sxz > 0 ? 2 : 3);
      r = fmaf(kc, cn[k], v);
    }
    in.template put<L>(r);
  }
  template <int G>
  __device__ __forceinline__ void putg(float) {}
};


// Colour model, per node (recursive_colorTSLB3D.f90:719-959, corrected):
// omega from the harmonic-mean viscosity of the two components (:756-758),
// 1/nu = cA/nuA + cB/nuB = x and omega = 1/(3 nu + 1/2) = x/(3 + x/2); the
// perturbation strength K = (9/4) sigma omega/|g| and the recolouring factor
// bc = beta cA cB rho/|g| where |g| > 0.001 (0 elsewhere).  Hardware
// reciprocal and reciprocal square root (one ulp) in place of four
// correctly rounded divisions and a square root.
struct ColourNode {
  float ca, omega, K, bc, g2;
};
__device__ __forceinline__ ColourNode colour_node(const float *m, float gx, float gy, float gz, const Params &p) {
  ColourNode c;
  c.ca = fminf(fmaxf(0.5f * (1.f + m[10]), 0.f), 1.f);
  const float cb = 1.f - c.ca;
  const float x = fminf(fmaf(c.ca, p.inua, cb * p.inub), 3e30f);  // nu = 0 (tau = 1/2): omega = 2
  c.omega = x * __builtin_amdgcn_rcpf(fmaf(0.5f, x, 3.f));
  c.g2 = fmaf(gx, gx, fmaf(gy, gy, gz * gz));
  const bool pert = c.g2 > 1e-6f;
  const float ig = __builtin_amdgcn_rsqf(c.g2);
  c.K = pert ? 2.25f * p.sigma * c.omega * ig : 0.f;
  c.bc = pert ? p.beta * c.ca * cb * (1.f + m[0]) * ig : 0.f;
  return c;
}

template <int NC, bool ADD, bool CLO, typename Sink>
__device__ __forceinline__ void eval_src(const Src<NC, CLO> &s, const Params &p, Sink &h, float *gq) {
  const float *m = s.m;
  float cbuf[10], om1;
  const float *mc = closure_moments(s, p, cbuf, om1);
  if constexpr (NC == 2) {
    const float phi = m[10];
    const float n0 = s.nn[0], n1 = s.nn[1], n2 = s.nn[2];
    // CSF velocity shift sigma*kappa*n, recursive_2c_TSLB3D.f90:614-636; the
    // gated curvature comes from the normals pass
    const float curv = s.kap;
    const float sx = p.sigma * curv * n0, sy = p.sigma * curv * n1, sz = p.sigma * curv * n2;
    const float esx = sx + p.fx + p.lin_a * m[1];
    const float esy = sy + p.fy + p.lin_a * m[2];
    const float esz = sz + p.fz + p.lin_a * m[3];
    all_pops_mt<false>(mc, m[1] + esx, m[2] + esy, m[3] + esz, p, h, om1);
    // D3Q7 phase field with sharpening, recursive_2c_TSLB3D.f90:808-831
    // g_{+-a} = phi/8 +- A_a with A_a = (3/8) phi u_a + sh n_a: the sink gets
    // A_x and A_y (the destination adds phi/8 of the source from the phi
    // footprint), gq returns g0, A_x, A_y, phi/8, g5, g6
    const float p1g = 0.125f;
    const float sh = p1g * p.sharp_c * phi * (1.f - phi);
    const float base = p1g * phi;
    const float ax = fmaf(3.f * base, m[1], sh * n0);
    const float ay = fmaf(3.f * base, m[2], sh * n1);
    const float az = fmaf(3.f * base, m[3], sh * n2);
    h.template putg<1>(ax);
    h.template putg<2>(ay);
    gq[0] = 0.25f * phi;
    gq[1] = ax;
    gq[2] = ay;
    gq[3] = base;
    gq[5] = base + az;
    gq[6] = base - az;
  } else if constexpr (NC == 3) {
    // colour model: per-node omega, perturbation and recolouring factor
    const float gx = s.nn[0], gy = s.nn[1], gz = s.nn[2];
    const ColourNode c = colour_node(m, gx, gy, gz, p);
    ColourSink<Sink> cs(h, c.K, gx, gy, gz, c.g2);
    all_pops_mt<ADD>(m, m[1], m[2], m[3], p, cs, 1.f - c.omega);
    gq[0] = c.ca;
    gq[1] = c.bc;
    gq[2] = gx;
    gq[3] = gy;
    gq[4] = gz;
  } else {
    if (p.shift) {
      all_pops_mt<false>(mc, m[1] + (p.fx + p.lin_a * m[1]), m[2] + (p.fy + p.lin_a * m[2]),
                         m[3] + (p.fz + p.lin_a * m[3]), p, h, om1);
    } else {
      all_pops_mt<ADD>(mc, m[1], m[2], m[3], p, h, om1);
    }
  }
}

// A ring source (one- and two-component models): eval_src restricted to the
// 9 populations entering the tile (ring_pops) and, for the phase field, the
// flux A_P along the side's normal axis (rf: LDS slot 24 for A_x, 25 for A_y).
template <int NC, bool ADD, bool CLO>
__device__ __forceinline__ void eval_ring(const Src<NC, CLO> &s, const Params &p, bool col, float sgn,
                                          float *const (&rq)[3], float *rf) {
  const float *m = s.m;
  float cbuf[10], om1;
  const float *mc = closure_moments(s, p, cbuf, om1);
  if constexpr (NC == 3) {
    // the colour model (eval_src's, on the 9 entering populations); rf
    // returns cA, bc, gx, gy, gz for the recolouring pass
    const float gx = s.nn[0], gy = s.nn[1], gz = s.nn[2];
    const ColourNode c = colour_node(m, gx, gy, gz, p);
    const ColourRingAdj adj{RingDir{col, sgn, gx, gy, gz}, c.K, c.g2};
    ring_pops<ADD>(m, m[1], m[2], m[3], p, 1.f - c.omega, col, sgn, rq, adj);
    rf[0] = c.ca;
    rf[1] = c.bc;
    rf[2] = gx;
    rf[3] = gy;
    rf[4] = gz;
  } else if constexpr (NC == 2) {
    const float phi = m[10];
    const float n0 = s.nn[0], n1 = s.nn[1], n2 = s.nn[2];
    const float curv = s.kap;
    const float sx = p.sigma * curv * n0, sy = p.sigma * curv * n1, sz = p.sigma * curv * n2;
    const float esx = sx + p.fx + p.lin_a * m[1];
    const float esy = sy + p.fy + p.lin_a * m[2];
    const float esz = sz + p.fz + p.lin_a * m[3];
    ring_pops<false>(mc, m[1] + esx, m[2] + esy, m[3] + esz, p, om1, col, sgn, rq);
    const float sh = 0.125f * p.sharp_c * phi * (1.f - phi);
    const float base = 0.125f * phi;
    *rf = fmaf(3.f * base, col ? m[1] : m[2], sh * (col ? n0 : n1));
  } else {
    if (p.shift) {
      ring_pops<false>(mc, m[1] + (p.fx + p.lin_a * m[1]), m[2] + (p.fy + p.lin_a * m[2]),
                       m[3] + (p.fz + p.lin_a * m[3]), p, om1, col, sgn, rq);
    } else {
      ring_pops<ADD>(mc, m[1], m[2], m[3], p, om1, col, sgn, rq);
    }
  }
}

// moment sums of shifted populations for one destination plane
struct Acc {
  float s0, sx, sy, sz, sxx, syy, szz, sxy, sxz, syz, phi;
};
// The sums start at -0: -0 + x is x for every x, so the compiler drops the
// first add of each sum, where +0 + x (which turns -0 into +0) costs one
// instruction per field and plane.
__device__ __forceinline__ void acc_zero(Acc &a) {
  a.s0 = a.sx = a.sy = a.sz = a.sxx = a.syy = a.szz = a.sxy = a.sxz = a.syz = a.phi = -0.f;
}

// accumulate the 9 populations of one c_z layer, p[cy+1][cx+1]
template <int CZ>
__device__ __forceinline__ void acc_layer(Acc &a, const float (&p)[3][3]) {
  float rs[3], rd[3], rx[3];
#pragma unroll
  for (int r = 0; r < 3; r++) {
    rx[r] = p[r][2] + p[r][0];
    rs[r] = rx[r] + p[r][1];
    rd[r] = p[r][2] - p[r][0];
  }
  const float U0 = rs[0] + rs[1] + rs[2];
  const float U1 = rs[2] - rs[0];
  const float U2 = rs[2] + rs[0];
  const float X1 = rd[0] + rd[1] + rd[2];
  const float X1Y = rd[2] - rd[0];
  const float X2 = rx[0] + rx[1] + rx[2];
  a.s0 += U0;
  a.sy += U1;
  a.syy += U2;
  a.sx += X1;
  a.sxy += X1Y;
  a.sxx += X2;
  if constexpr (CZ != 0) {
    a.sz += cm<CZ>(U0);
    a.szz += U0;
    a.sxz += cm<CZ>(X1);
    a.syz += cm<CZ>(U1);
  }
}

// gather the c_z = CZ layer for destination column (tx,ty) from LDS
template <int CZ, bool MASKED>
__device__ __forceinline__ void gather_layer(Acc &a, const float (*lds)[NSLOT], int tx, int ty,
                                             float own, uint32_t bm) {
  float p[3][3];
#pragma unroll
  for (int cy = -1; cy <= 1; cy++)
#pragma unroll
    for (int cx = -1; cx <= 1; cx++) {
      const int l = dir_of(cx, cy, CZ);
      float v;
      if (cx == 0 && cy == 0) v = own;
      else v = lds[qslot(l)][(ty + 1 - cy) * SW + (tx + 1 - cx)];
      if (MASKED && ((bm >> l) & 1u)) v = 0.f;
      p[cy + 1][cx + 1] = v;
    }
  acc_layer<CZ>(a, p);
}

// bounce-back replacements for plane of the own source: the population a
// destination sent towards an invalid source comes back, f_l(y) = h_opp(l)(y).
template <int CZ>
__device__ __forceinline__ void bb_layer(Acc &a, const float (*lds)[NSLOT], int tx, int ty,
                                         float h0, float h5, float h6, uint32_t bm) {
  float p[3][3];
  const int own = (ty + 1) * SW + (tx + 1);
#pragma unroll
  for (int cy = -1; cy <= 1; cy++)
#pragma unroll
    for (int cx = -1; cx <= 1; cx++) {
      const int l = dir_of(cx, cy, CZ);
      const int o = kOPP[l];
      float v = 0.f;
      if ((bm >> l) & 1u) {
        if (o == 0) v = h0;
        else if (o == 5) v = h5;
        else if (o == 6) v = h6;
        else v = lds[qslot(o)][own];
      }
      p[cy + 1][cx + 1] = v;
    }
  acc_layer<CZ>(a, p);
}

// Colour model, second LDS pass: the A population of each direction from
// the total one (shifted h, as stored) with Latva-Kokko recolouring,
//   fA_l = cA (h_l + w_l) + bc w_l (c_l.g)/|c_l|,  bc = beta cA cB rho/|g|,
// for all 27 directions of one source from shared terms:
// fA_l = cA h_l + (cA w_l + bc w_l (c_l.g)/|c_l|), where the second part
// takes 4 values of cA w_l and, up to sign, 13 of the c.g term (faces g_a,
// edges g_a +- g_b, corners gx +- gy +- gz)
struct Recolour {
  float ca, cw0, cwf, cwe, cwc;
  float tf[3], tep[3], tem[3], tcn[4];
  __device__ __forceinline__ explicit Recolour(const float *c) {
    ca = c[0];
    const float bc = c[1], gx = c[2], gy = c[3], gz = c[4];
    cw0 = ca * (8.f / 27.f);
    cwf = ca * (2.f / 27.f);
    cwe = ca * (1.f / 54.f);
    cwc = ca * (1.f / 216.f);
    const float bf = bc * (2.f / 27.f), be = bc * (0.70710678118654752f / 54.f),
                bn = bc * (0.57735026918962576f / 216.f);
    tf[0] = bf * gx;
    tf[1] = bf * gy;
    tf[2] = bf * gz;
    tep[0] = be * (gx + gy);
    tem[0] = be * (gx - gy);
    tep[1] = be * (gx + gz);
    tem[1] = be * (gx - gz);
    tep[2] = be * (gy + gz);
    tem[2] = be * (gy - gz);
    const float xpy = gx + gy, xmy = gx - gy;
    tcn[0] = bn * (xpy + gz);
    tcn[1] = bn * (xpy - gz);
    tcn[2] = bn * (xmy + gz);
    tcn[3] = bn * (xmy - gz);
  }
  template <int L>
  __device__ __forceinline__ float a(float h) const {
    constexpr int cx = kCX[L], cy = kCY[L], cz = kCZ[L];
    constexpr int nnz = (cx != 0) + (cy != 0) + (cz != 0);
    if constexpr (nnz == 0) {
      return fmaf(ca, h, cw0);
    } else if constexpr (nnz == 1) {
      constexpr int ax = cx != 0 ? 0 : (cy != 0 ? 1 : 2);
      constexpr int sg = cx + cy + cz;
      return fmaf(ca, h, sg > 0 ? cwf + tf[ax] : cwf - tf[ax]);
    } else if constexpr (nnz == 2) {
      // c.g = sa (g_a + sa sb g_b) for the pair (a, b)
      constexpr int pr = cz == 0 ? 0 : (cy == 0 ? 1 : 2);
      constexpr int sa = pr == 2 ? cy : cx, sb = pr == 0 ? cy : cz;
      const float t = sa * sb > 0 ? tep[pr] : tem[pr];
      return fmaf(ca, h, sa > 0 ? cwe + t : cwe - t);
    } else {
      // c.g = sx (gx + sx sy gy + sx sz gz)
      constexpr int sxy = cx * cy, sxz = cx * cz;
      constexpr int k = sxy > 0 ? (sxz > 0 ? 0 : 1) : (sxz > 0 ? 2 : 3);
      return fmaf(ca, h, cx > 0 ? cwc + tcn[k] : cwc - tcn[k]);
    }
  }
};

// colour model: rewrite the 24 in-plane total populations of one LDS slot as
// A populations (c = {cA, bc, gx, gy, gz})
template <int L>
__device__ __forceinline__ void colour_rewrite(float (*lds)[NSLOT], int slot, const Recolour &rc) {
  if constexpr (L < 27) {
    if constexpr (L != 5 && L != 6) lds[qslot(L)][slot] = rc.template a<L>(lds[qslot(L)][slot]);
    colour_rewrite<L + 1>(lds, slot, rc);
  }
}

// colour model, a ring source's second pass: its 9 entering total
// populations become recoloured A populations (Recolour's, by direction)
__device__ __forceinline__ void colour_rewrite_ring(float *const (&rq)[3], const float *c, bool col, float sgn) {
  const RingDir d{col, sgn, c[2], c[3], c[4]};
#pragma unroll
  for (int cz = -1; cz <= 1; cz++)
#pragma unroll
    for (int q = -1; q <= 1; q++) {
      float &v = rq[q + 1][(cz + 1) * 8 * NSLOT];
      const float w = RingDir::w(q, cz);
      v = fmaf(c[0], v + w, c[1] * (w * RingDir::inv(q, cz)) * d.cg(q, cz));
    }
}

// colour model: A populations of the c_z = CZ layer arriving at (tx, ty)
template <int CZ, bool MASKED>
__device__ __forceinline__ float sum_layer(const float (*lds)[NSLOT], int tx, int ty, float own,
                                           uint32_t bm) {
  float s = -0.f;
#pragma unroll
  for (int cy = -1; cy <= 1; cy++)
#pragma unroll
    for (int cx = -1; cx <= 1; cx++) {
      const int l = dir_of(cx, cy, CZ);
      float v = (cx == 0 && cy == 0) ? own : lds[qslot(l)][(ty + 1 - cy) * SW + (tx + 1 - cx)];
      if (MASKED && ((bm >> l) & 1u)) v = 0.f;
      s += v;
    }
  return s;
}

// colour model: bounce-back of the A populations of the own node
template <int CZ>
__device__ __forceinline__ float bb_sum_layer(const float (*lds)[NSLOT], int tx, int ty, float a0,
                                              float a5, float a6, uint32_t bm) {
  float s = 0.f;
  const int own = (ty + 1) * SW + (tx + 1);
#pragma unroll
  for (int cy = -1; cy <= 1; cy++)
#pragma unroll
    for (int cx = -1; cx <= 1; cx++) {
      const int l = dir_of(cx, cy, CZ);
      const int o = kOPP[l];
      if ((bm >> l) & 1u) s += o == 0 ? a0 : (o == 5 ? a5 : (o == 6 ? a6 : lds[qslot(o)][own]));
    }
  return s;
}

// Runs of interface tile-planes are cut at every kRunMax-th plane of a
// z-chunk, so that no block of the curvature pass marches more than kRunMax
// planes: the long runs of a chunk otherwise set the length of the launch.
// (Cutting the runs of chunks with interface in half their planes at every
// 32nd plane instead, to save the runs' start-up, measured slower at 1.2 and
// 15 % volume fraction: 0.142 -> 0.168 and 0.97 -> 1.02 ms per step.)
constexpr int kRunMax = 8;
constexpr uint64_t kRunCut = 0x0101010101010101ull;  // every 8th plane of a chunk
static_assert(kRunMax == 8, "kRunCut marks every 8th plane");
// starts of the runs in a 64-plane hit mask, and the length of the run at k
__device__ __forceinline__ uint64_t run_starts(uint64_t all) { return (all & ~(all << 1)) | (all & kRunCut); }
__device__ __forceinline__ int run_len(uint64_t all, int k) {
  const uint64_t above = ~(all >> k);  // first clear bit at or above k ends the run
  const int len = above ? __ffsll((long long)above) - 1 : 64 - k;
  return min(len, kRunMax - (k & (kRunMax - 1)));
}

// Appends the interface runs of one tile's z-chunk [z0, z0 + nz) to the
// list, for both half-tiles: wm[2 w + h] holds wave w's plane mask of half h.
// Entry = 64 (2 (z tiles_plane + tile) + h) + length - 1.  Every thread of
// the block calls it (one barrier; the caller has stored wm and synchronised).
__device__ __forceinline__ void append_runs(const uint64_t *wm, int *sbase, int tid, int z0, int nz,
                                            int tiles_plane, int tile, int *cnt, int *list) {
  uint64_t all0 = 0, all1 = 0;
#pragma unroll
  for (int w = 0; w < NTHR / 64; w++) {
    all0 |= wm[2 * w];
    all1 |= wm[2 * w + 1];
  }
  const uint64_t st0 = run_starts(all0), st1 = run_starts(all1);
  const int n0 = __popcll(st0);
  if (tid == 0) *sbase = (st0 | st1) ? atomicAdd(cnt, n0 + __popcll(st1)) : 0;
  __syncthreads();
  // thread k appends the run of half 0 starting at plane z0 + k, thread
  // 64 + k that of half 1, each at its rank among the starts
  const int h = tid >> 6, k = tid & 63;
  if (h < 2 && k < nz) {
    const uint64_t st = h ? st1 : st0;
    if ((st >> k) & 1ull) {
      const int len = run_len(h ? all1 : all0, k);
      list[*sbase + (h ? n0 : 0) + __popcll(st & ((1ull << k) - 1ull))] =
          (2 * ((z0 + k) * tiles_plane + tile) + h) * 64 + len - 1;
    }
  }
}

template <int NC, bool MASK, bool ADD, int WPE, typename S, bool CLO = false, bool R9 = true>
__global__ __launch_bounds__(NTHR) __attribute__((amdgpu_waves_per_eu(WPE, 8))) void k_fused(
    const S *__restrict__ src, S *__restrict__ dst, const float *__restrict__ nrm,
    const uint8_t *__restrict__ flags, const uint32_t *__restrict__ links, const float *__restrict__ clo,
    int *__restrict__ ilist, uint8_t *__restrict__ flags_next, Params p) {
  constexpr int NQ = NC == 2 ? 26 : 24;
  constexpr int NF = NC >= 2 ? 11 : 10;
  __shared__ float lds[NQ][NSLOT];
  // phase field: phi planes zs-1, zs, zs+1 of the (TX+4)x(TY+4) footprint
  constexpr int NPR = NC >= 2 ? 3 : 1;
  __shared__ float phr[NPR][NC >= 2 ? NPSLOT : 1];
  __shared__ uint64_t ilist_wm[2 * (NTHR / 64)];
  __shared__ int ilist_base;

  const int tid = threadIdx.x;
  const int tx = tid % TX, ty = tid / TX;
  int b = xcd_tile(blockIdx.x, gridDim.x);
  const int tile_x = b % p.tiles_x;
  b /= p.tiles_x;
  const int tile_y = b % p.tiles_y;
  const int zc = b / p.tiles_y;
  const int x0 = tile_x * TX, y0 = tile_y * TY;
  const int z0 = zc * p.z_chunk;
  const int z1 = min(z0 + p.z_chunk, p.nz);
  const int x = x0 + tx, y = y0 + ty;
  const bool active = x < p.nx && y < p.ny;
  const size_t plane = (size_t)p.nx * p.ny;
  const size_t tiles_plane = (size_t)p.tiles_x * p.tiles_y;
  const uint32_t pbS = (uint32_t)(plane * sizeof(S)), pb4 = (uint32_t)(plane * 4);
  const long long fsS = p.fs * (long long)sizeof(S), fs4 = p.fs * 4;

  // own source column
  int ox = 0, oy = 0;
  const bool oxy = src_coord(x, p.nx, p.px, ox) && src_coord(y, p.ny, p.py, oy);
  // ring source slot of this thread
  int rsx = 0, rsy = 0;
  if (tid < NRING) {
    if (tid < SW) { rsx = tid; rsy = 0; }
    else if (tid < 2 * SW) { rsx = tid - SW; rsy = SH - 1; }
    else if (tid < 2 * SW + TY) { rsx = 0; rsy = 1 + (tid - 2 * SW); }
    else { rsx = SW - 1; rsy = 1 + (tid - 2 * SW - TY); }
  }
  // a ring source's side (columns: c_P along x; rows: along y), the sign of
  // the c_P that enters the tile, and the LDS addresses of its 9 entering
  // populations (c_Q = -1, 0, 1 at c_z = -1; c_z = 0 and 1 sit 8 and 16
  // slots further)
  auto ring_geom = [&](bool &col, float &sgn, float *(&rq)[3]) {
    col = tid >= 2 * SW;
    sgn = (tid < SW || (tid >= 2 * SW && tid < 2 * SW + TY)) ? 1.f : -1.f;
    const int sg = sgn > 0.f ? 1 : -1, rs = rsy * SW + rsx;
    rq[0] = &lds[col ? dslot(sg, -1) : dslot(-1, sg)][rs];
    rq[1] = &lds[col ? dslot(sg, 0) : dslot(0, sg)][rs];
    rq[2] = &lds[col ? dslot(sg, 1) : dslot(1, sg)][rs];
  };
  int rx = 0, ry = 0;
  const bool rxy = tid < NRING && src_coord(x0 + rsx - 1, p.nx, p.px, rx) &&
                   src_coord(y0 + rsy - 1, p.ny, p.py, ry);
  // in-plane element offsets (loop-invariant) and flag tiles of both columns
  const uint32_t o_xy = (uint32_t)(ox + p.nx * oy), r_xy = (uint32_t)(rx + p.nx * ry);
  const uint32_t d_xy = active ? (uint32_t)(x + p.nx * y) : 0u;
  // curvature flag bytes of both source columns: one per half-tile
  const int o_tile = 2 * ((oy / TY) * p.tiles_x + ox / TX) + (ox % TX) / HX;
  const int r_tile = 2 * ((ry / TY) * p.tiles_x + rx / TX) + (rx % TX) / HX;

  // phase field: phi footprint slots tid and tid + NTHR of this thread, and
  // their in-plane byte offsets (nearest interior node across a wall)
  const bool ptwo = NC >= 2 && tid + NTHR < NPSLOT;
  auto phi_off = [&](int sl) -> uint32_t {
    const int gx = clampw(x0 + sl % PW - 2, p.nx, p.px);
    const int gy = clampw(y0 + sl / PW - 2, p.ny, p.py);
    return (uint32_t)(gx + p.nx * gy) * (uint32_t)sizeof(S);
  };
  const uint32_t pofs0 = NC >= 2 ? phi_off(tid) : 0u, pofs1 = ptwo ? phi_off(tid + NTHR) : 0u;
  S pr0{}, pr1{};  // raw storage values, converted at commit
  auto phi_issue = [&](int z) {
    const rsrc_t r = rsrc((const char *)src + TSLB_PHI * fsS + (size_t)clampw(z, p.nz, p.pz) * pbS, pbS);
    pr0 = bld<S>(r, pofs0);
    if (ptwo) pr1 = bld<S>(r, pofs1);
  };
  auto phi_commit = [&](float *d) {
    d[tid] = dec(pr0, p, sgrp(TSLB_PHI));
    if (ptwo) d[tid + NTHR] = dec(pr1, p, sgrp(TSLB_PHI));
  };
  // Phase field: the normals of both sources come from the separated
  // gradient stencil, marched along z with the source planes (footprint
  // centres co_c and cr_c of the phi planes).
  const int co_c = (ty + 2) * PW + tx + 2, cr_c = (rsy + 1) * PW + rsx + 1;
  GradRun gro, grr;

  // ghost-wall bounce-back bits of this column
  uint32_t wall_xy = 0;
  if (!p.px) {
    if (x == 0) wall_xy |= dirmask(0, 1);
    if (x == p.nx - 1) wall_xy |= dirmask(0, -1);
  }
  if (!p.py) {
    if (y == 0) wall_xy |= dirmask(1, 1);
    if (y == p.ny - 1) wall_xy |= dirmask(1, -1);
  }
  auto plane_mask = [&](int z) -> uint32_t {
    uint32_t m = wall_xy;
    if (!p.pz) {
      if (z == 0) m |= dirmask(2, 1);
      if (z == p.nz - 1) m |= dirmask(2, -1);
    }
    if constexpr (MASK) {
      if (active && z >= 0 && z < p.nz) m |= bld_u32(rsrc(links + plane * z, pb4), d_xy * 4u);
    }
    return m;
  };

  // source fetch of one plane: own column and (tid < NRING) the ring slot
  auto fetch = [&](int zs, Src<NC, CLO> &o, bool &ook, Src<NC, CLO> &r, bool &rok) {
    int gz = 0;
    const bool zok = src_coord(zs, p.nz, p.pz, gz);
    PlaneSrc ps;
    ps.f = (const char *)src + (size_t)gz * pbS;
    ps.n = (const char *)nrm + (size_t)gz * pb4;
    ps.c = (const char *)clo + (size_t)gz * pb4;
    ps.pbS = pbS;
    ps.pb4 = pb4;
    ps.fsS = fsS;
    ps.fs4 = fs4;
    const rsrc_t fl = rsrc(flags + 2 * tiles_plane * (size_t)gz, (uint32_t)(2 * tiles_plane));
    const rsrc_t lk = rsrc(links + plane * (size_t)gz, pb4);
    // Per source: link mask, curvature flag, moments, the flag-dependent
    // loads.  (Issuing both sources' masks and flags ahead of all moments
    // measured slower: 3.10 against 3.00 ms for 2C FP32.)
    RawSrc<NC, S> rwo, rwr;  // reduced storage: converted after all loads
    ook = zok && oxy;
    if (ook) {
      if constexpr (MASK) ook = (bld_u32(lk, o_xy * 4u) & kSolid) == 0u;
      bool un = NC != 2;  // flags gate the phase-field curvature only
      if constexpr (NC == 2) un = bld_u8(fl, (uint32_t)o_tile) != 0;
      if (ook) {
        if constexpr (sizeof(S) == 4) load_m32(ps, o_xy * 4u, o);
        else load_raw<NC, S>(ps, o_xy * (uint32_t)sizeof(S), rwo);
        load_extra(ps, un ? o_xy * 4u : kOOB, o_xy * 4u, o);
      }
    }
    rok = zok && rxy;
    if (rok) {
      if constexpr (MASK) rok = (bld_u32(lk, r_xy * 4u) & kSolid) == 0u;
      bool un = NC != 2;
      if constexpr (NC == 2) un = bld_u8(fl, (uint32_t)r_tile) != 0;
      if (rok) {
        if constexpr (sizeof(S) == 4) load_m32(ps, r_xy * 4u, r);
        else load_raw<NC, S>(ps, r_xy * (uint32_t)sizeof(S), rwr);
        load_extra(ps, un ? r_xy * 4u : kOOB, r_xy * 4u, r);
      }
    }
    if constexpr (sizeof(S) != 4) {
      __builtin_amdgcn_sched_barrier(0);
      if (ook) convert_raw(rwo, p, o);
      if (rok) convert_raw(rwr, p, r);
    }
  };

  // Phase field: the next step's list of interface runs, as k_iface_list
  // builds it, from the planes this block finishes (when ilist is given),
  // one plane mask per half-tile
  uint64_t imask0 = 0, imask1 = 0;

  // One source plane zs.  aA, aB and aC accumulate the destination planes
  // zs-1 (completed and written here), zs and zs+1; bmA..bmC are their
  // bounce-back masks.  The loop below renames the three accumulators instead
  // of copying them.  Phase field: pm, p0 and pp hold phi of planes zs-1, zs
  // and zs+1; plane zs+2 is loaded meanwhile and replaces zs-1 in pm.
  auto body = [&](int zs, Acc &aA, Acc &aB, Acc &aC, uint32_t &bmA, uint32_t &bmB, uint32_t &bmC,
                  float *pm, const float *p0, const float *pp) {
    if constexpr (NC >= 2) {
      if (zs + 2 <= z1 + 1) phi_issue(zs + 2);  // planes up to z1 + 1 are read
    }
    Src<NC, CLO> co, cr;
    bool co_ok, cr_ok;
    // two-component models in FP32 storage: the plane's loads issue at raised
    // wave priority, ahead of the other block's arithmetic, so that their
    // latency starts counting early (2C emulsion 2.66 -> 2.63 ms, colour
    // sphere -0.8 %; one component 0.8 % slower, left out).  With 16-bit
    // storage, whose loads are issued in one batch before any conversion,
    // neither this hint nor the ring-wave one below pays for the phase-field
    // model or one component (2C emulsion without them: FP16 2.67 -> 2.60 ms,
    // BF16 2.75 -> 2.62 ms; channel BF16 0.58 -> 0.54 ms); the colour model
    // keeps both (FP16 sphere 0.459 -> 0.448 ms with them)
    if constexpr (NC >= 2 && (sizeof(S) == 4 || NC == 3)) __builtin_amdgcn_s_setprio(2);
    fetch(zs, co, co_ok, cr, cr_ok);
    // the waves that also evaluate ring sources keep priority 1 until the
    // barrier (they set when it releases), the others drop to 0
    const bool ring_wave = (tid >> 6) < (NRING + 63) / 64;
    if (ring_wave && (sizeof(S) == 4 || NC == 3)) __builtin_amdgcn_s_setprio(1);
    else __builtin_amdgcn_s_setprio(0);
    float gq[7];
    float h0 = 0.f, h5 = 0.f, h6 = 0.f, g0 = 0.f, g5 = 0.f, g6 = 0.f;
    float gax = 0.f, gay = 0.f, gb = 0.f;  // phase field: own A_x, A_y, phi/8
    float oc[5] = {0.f, 0.f, 0.f, 0.f, 0.f}, rc[5] = {0.f, 0.f, 0.f, 0.f, 0.f};  // colour: cA, bc, g
    // phase field: unit normals; colour model: the colour gradient
    float nno[3], nnr[3];
    if constexpr (NC >= 2) {
      grad_run_step<NC == 2>(gro, plane_grad(pp, co_c), pm[co_c], nno);
      if (tid < NRING) grad_run_step<NC == 2>(grr, plane_grad(pp, cr_c), pm[cr_c], nnr);
    }
    // ---- own source slot
    if (co_ok) {
      if constexpr (NC >= 2) {
        co.nn[0] = nno[0];
        co.nn[1] = nno[1];
        co.nn[2] = nno[2];
        co.m[NF - 1] = p0[co_c];
      }
      PopSink sk{lds, (ty + 1) * SW + (tx + 1), 0.f, 0.f, 0.f};
      eval_src<NC, ADD>(co, p, sk, gq);
      h0 = sk.h0;
      h5 = sk.h5;
      h6 = sk.h6;
      if constexpr (NC == 2) {
        g0 = gq[0];
        gax = gq[1];
        gay = gq[2];
        gb = gq[3];
        g5 = gq[5];
        g6 = gq[6];
      }
      if constexpr (NC == 3)
        for (int k = 0; k < 5; k++) oc[k] = gq[k];
    }
    // ---- halo ring slot
    if (cr_ok) {
      if constexpr (NC >= 2) {
        cr.nn[0] = nnr[0];
        cr.nn[1] = nnr[1];
        cr.nn[2] = nnr[2];
        cr.m[NF - 1] = p0[cr_c];
      }
      if constexpr (R9) {
        // the 9 populations (and the phase flux, or the colour model's
        // recolouring inputs) that enter the tile
        bool col;
        float sgn;
        float *rq[3];
        ring_geom(col, sgn, rq);
        if constexpr (NC == 3) eval_ring<NC, ADD>(cr, p, col, sgn, rq, rc);
        else eval_ring<NC, ADD>(cr, p, col, sgn, rq, &lds[col ? 24 : 25][rsy * SW + rsx]);
      } else {
        PopSink sk{lds, rsy * SW + rsx, 0.f, 0.f, 0.f};
        eval_src<NC, ADD>(cr, p, sk, gq);
        if constexpr (NC == 3)
          for (int k = 0; k < 5; k++) rc[k] = gq[k];
      }
    }
    __builtin_amdgcn_s_setprio(0);
    __syncthreads();
    if constexpr (NC >= 2) phi_commit(pm);  // read again after the closing barrier

    bmC = plane_mask(zs + 1);
    acc_zero(aC);
    if (active) {
      if ((bmA | bmB | bmC) == 0u) {
        gather_layer<-1, false>(aA, lds, tx, ty, h6, 0u);
        gather_layer<0, false>(aB, lds, tx, ty, h0, 0u);
        gather_layer<1, false>(aC, lds, tx, ty, h5, 0u);
        if constexpr (NC == 2) {
          // g_{+-a} of the in-plane neighbours: phi/8 from the phi plane, +-A_a from LDS
          aA.phi += g6;
          aB.phi += g0 + fmaf(0.125f, p0[(ty + 2) * PW + tx + 1], lds[24][(ty + 1) * SW + tx]) +
                    fmaf(0.125f, p0[(ty + 2) * PW + tx + 3], -lds[24][(ty + 1) * SW + tx + 2]) +
                    fmaf(0.125f, p0[(ty + 1) * PW + tx + 2], lds[25][ty * SW + tx + 1]) +
                    fmaf(0.125f, p0[(ty + 3) * PW + tx + 2], -lds[25][(ty + 2) * SW + tx + 1]);
          aC.phi += g5;
        }
      } else {
        gather_layer<-1, true>(aA, lds, tx, ty, h6, bmA);
        gather_layer<0, true>(aB, lds, tx, ty, h0, bmB);
        gather_layer<1, true>(aC, lds, tx, ty, h5, bmC);
        bb_layer<-1>(aB, lds, tx, ty, h0, h5, h6, bmB);
        bb_layer<0>(aB, lds, tx, ty, h0, h5, h6, bmB);
        bb_layer<1>(aB, lds, tx, ty, h0, h5, h6, bmB);
        if constexpr (NC == 2) {
          const float gown[7] = {g0, gb + gax, gb - gax, gb + gay, gb - gay, g5, g6};
          // phase populations: l=6 into plane zs-1, l=0..4 into zs, l=5 into zs+1
          if (!((bmA >> 6) & 1u)) aA.phi += g6;
          float s = g0;
          if (!((bmB >> 1) & 1u)) s += fmaf(0.125f, p0[(ty + 2) * PW + tx + 1], lds[24][(ty + 1) * SW + tx]);
          if (!((bmB >> 2) & 1u)) s += fmaf(0.125f, p0[(ty + 2) * PW + tx + 3], -lds[24][(ty + 1) * SW + tx + 2]);
          if (!((bmB >> 3) & 1u)) s += fmaf(0.125f, p0[(ty + 1) * PW + tx + 2], lds[25][ty * SW + tx + 1]);
          if (!((bmB >> 4) & 1u)) s += fmaf(0.125f, p0[(ty + 3) * PW + tx + 2], -lds[25][(ty + 2) * SW + tx + 1]);
          if (!p.gwall) {
#pragma unroll
            for (int l = 1; l <= 6; l++)
              if ((bmB >> l) & 1u) s += gown[kOPP[l]];
          }
          aB.phi += s;
          if (!((bmC >> 5) & 1u)) aC.phi += g5;
        }
      }
    }
    if constexpr (NC == 3) {
      // second LDS pass: every source slot turns its total populations into
      // the recoloured A populations, and destinations sum those
      __syncthreads();
      if (ring_wave) __builtin_amdgcn_s_setprio(1);  // the ring rewrites set this barrier too
      float a0 = 0.f, a5 = 0.f, a6 = 0.f;
      if (co_ok) {
        const Recolour ro(oc);
        colour_rewrite<1>(lds, (ty + 1) * SW + (tx + 1), ro);
        a0 = ro.a<0>(h0);
        a5 = ro.a<5>(h5);
        a6 = ro.a<6>(h6);
      }
      if (cr_ok) {
        if constexpr (R9) {
          bool col;
          float sgn;
          float *rq[3];
          ring_geom(col, sgn, rq);
          colour_rewrite_ring(rq, rc, col, sgn);
        } else {
          colour_rewrite<1>(lds, rsy * SW + rsx, Recolour(rc));
        }
      }
      __builtin_amdgcn_s_setprio(0);
      __syncthreads();
      if (active) {
        if ((bmA | bmB | bmC) == 0u) {
          aA.phi += sum_layer<-1, false>(lds, tx, ty, a6, 0u);
          aB.phi += sum_layer<0, false>(lds, tx, ty, a0, 0u);
          aC.phi += sum_layer<1, false>(lds, tx, ty, a5, 0u);
        } else {
          aA.phi += sum_layer<-1, true>(lds, tx, ty, a6, bmA);
          aB.phi += sum_layer<0, true>(lds, tx, ty, a0, bmB) +
                    bb_sum_layer<-1>(lds, tx, ty, a0, a5, a6, bmB) +
                    bb_sum_layer<0>(lds, tx, ty, a0, a5, a6, bmB) +
                    bb_sum_layer<1>(lds, tx, ty, a0, a5, a6, bmB);
          aC.phi += sum_layer<1, true>(lds, tx, ty, a5, bmC);
        }
      }
    }

    // ---- finalize destination plane zs-1
    const int zd = zs - 1;
    bool hit = false;  // phase field: may the new plane carry curvature here?
    if (active && zd >= z0 && zd < z1) {
      const size_t zb = (size_t)zd * pbS;
      const uint32_t offS = d_xy * (uint32_t)sizeof(S);
      if (MASK && (bmA & kSolid)) {
#pragma unroll
        for (int q = 0; q < NF; q++)
          bst<2>(rsrc((char *)dst + zb + q * fsS, pbS), offS,
              bld<S>(rsrc((const char *)src + zb + q * fsS, pbS), offS));
        if constexpr (NC == 2) {
          const float f = dec(bld<S>(rsrc((const char *)src + zb + TSLB_PHI * fsS, pbS), offS), p, sgrp(TSLB_PHI));
          hit = f * (1.f - f) > 0.19f;
        }
      } else {
        const float rho = 1.f + aA.s0;
        // 1/rho: v_rcp_f32 refined by one Newton step in residual form.  The
        // residual 1 - rho inv is exact in the fma, so the step keeps all of
        // it and gives the correctly rounded reciprocal in nearly every case.
        // (The product form inv (2 - rho inv) rounds 2 - rho inv next to 1,
        // which biases 1/rho low: u then lost ~1.7e-8 of itself per step, a
        // kinetic-energy drift of 3.4e-5 over 1000 Taylor-Green steps.)
        float inv = __builtin_amdgcn_rcpf(rho);
        inv = fmaf(inv, fmaf(-rho, inv, 1.f), inv);
        const float u = aA.sx * inv, v = aA.sy * inv, w = aA.sz * inv;
        const float t = aA.s0 * (1.f / 3.f);
        const float out[11] = {aA.s0,
                               u,
                               v,
                               w,
                               aA.sxx - t - rho * u * u,
                               aA.syy - t - rho * v * v,
                               aA.szz - t - rho * w * w,
                               aA.sxy - rho * u * v,
                               aA.sxz - rho * u * w,
                               aA.syz - rho * v * w,
                               NC == 3 ? fmaf(2.f * aA.phi, inv, -1.f) : aA.phi};
        if constexpr (std::is_same<S, fx16_t>::value) {
          // the value in steps, y = v 2^m
          float x[NF];
#pragma unroll
          for (int q = 0; q < NF; q++) x[q] = out[q] * p.sc[sgrp(q)];
          // Past |y| = 32766 the values are clamped so that neither rounding
          // leaves +-32767 (dithered ones to 32766, see fx16_dither); a node
          // with a value beyond its field's range, |y| > 32767.5, is counted.
          float mx = 0.f;
#pragma unroll
          for (int q = 0; q < NF; q++) mx = fmaxf(mx, fabsf(x[q]));
          if (mx > 32766.f) {
            bool clip = false;
#pragma unroll
            for (int q = 0; q < NF; q++) clip |= fabsf(x[q]) > 32767.5f;
            if (clip && p.sat) atomicAdd(p.sat, 1u);
#pragma unroll
            for (int q = 0; q < NF; q++)
              x[q] = __builtin_amdgcn_fmed3f(x[q], -32767.f, dithered(q) ? 32766.f : 32767.f);
          }
          const uint32_t h1 = node_hash(d_xy + (uint32_t)plane * (uint32_t)zd, p.step);
          const uint32_t h2 = node_rehash(h1);
          const uint32_t w[5] = {h1 & 0xFFFFu, h1 >> 16, h2 & 0xFFFFu, h2 >> 16, (h1 ^ h2) & 0xFFFFu};
#pragma unroll
          for (int q = 0; q < NF; q++)
            bst<2>(rsrc((char *)dst + zb + q * fsS, pbS), offS,
                   dithered(q) ? fx16_dither(x[q], w[dither_word(q)]) : fx16_code(x[q]));
        } else if constexpr (sizeof(S) == 4) {
#pragma unroll
          for (int q = 0; q < NF; q++) bst<2>(rsrc((char *)dst + zb + q * fsS, pbS), offS, from_f<S>(out[q]));
        } else {
          // FP16, BF16: rho - 1, u, v, w and phi with stochastic rounding,
          // 32 random bits each
          const uint32_t h1 = node_hash(d_xy + (uint32_t)plane * (uint32_t)zd, p.step);
          const uint32_t h2 = node_rehash(h1);
          const uint32_t rw[5] = {h1, h2, rot16(h1), rot16(h2), h1 ^ h2};
#pragma unroll
          for (int q = 0; q < NF; q++) {
            const float v = out[q] * p.sc[sgrp(q)];
            bst<2>(rsrc((char *)dst + zb + q * fsS, pbS), offS,
                   dithered(q) ? from_f_sr<S>(v, rw[dither_word(q)]) : from_f<S>(v));
          }
        }
        if constexpr (NC == 2) hit = out[10] * (1.f - out[10]) > 0.19f;
      }
    }
    if constexpr (NC == 2) {
      // the test of k_iface_list on the new state (storage rounding of phi
      // stays far inside the 0.19 margin), one bit per plane and wave
      const uint64_t bal = __ballot(hit);
      if (zd >= z0 && zd < z1) {
        if ((uint32_t)bal) imask0 |= 1ull << (zd - z0);
        if (bal >> 32) imask1 |= 1ull << (zd - z0);
      }
    }
    __syncthreads();
  };

  Acc a0, a1, a2;
  acc_zero(a0);
  acc_zero(a1);
  uint32_t m0 = 0, m1 = plane_mask(z0 - 1), m2 = 0;
  float *R0 = phr[0], *R1 = phr[NPR > 1 ? 1 : 0], *R2 = phr[NPR > 2 ? 2 : 0];
  if constexpr (NC >= 2) {
    for (int k = 0; k < 3; k++) {
      phi_issue(z0 - 2 + k);
      phi_commit(phr[k < NPR ? k : 0]);
    }
    __syncthreads();
    grad_run_init(gro, plane_grad(R0, co_c), plane_grad(R1, co_c));
    if (tid < NRING) grad_run_init(grr, plane_grad(R0, cr_c), plane_grad(R1, cr_c));
  }
  for (int zs = z0 - 1; zs <= z1; zs += 3) {
    body(zs, a0, a1, a2, m0, m1, m2, R0, R1, R2);
    if (zs + 1 > z1) break;
    body(zs + 1, a1, a2, a0, m1, m2, m0, R1, R2, R0);
    if (zs + 2 > z1) break;
    body(zs + 2, a2, a0, a1, m2, m0, m1, R2, R0, R1);
  }
  if constexpr (NC == 2) {
    if (ilist) {
      const int tile = tile_y * p.tiles_x + tile_x;
      if ((tid & 63) == 0) {
        ilist_wm[2 * (tid >> 6)] = imask0;
        ilist_wm[2 * (tid >> 6) + 1] = imask1;
      }
      if (tid < 2 * (z1 - z0)) flags_next[2 * ((z0 + (tid >> 1)) * tiles_plane + tile) + (tid & 1)] = 0;
      __syncthreads();
      append_runs(ilist_wm, &ilist_base, tid, z0, z1 - z0, (int)tiles_plane, tile, ilist, ilist + 1);
    }
  }
}

// K2 + curvature, LDS-tiled.  A 64x8 block marches a z-chunk.  It keeps four
// phi planes of its (64+4)x(8+4) footprint and four normal planes of its
// (64+2)x(8+2) footprint in LDS rings, writes n = grad(phi)/(|grad(phi)|+1e-9)
// (recursive_2c_TSLB3D.f90:555-571) and kappa = -div(n) at nodes past the
// interface gate phi(1-phi)|n|^2 > 0.2 (0 elsewhere, :614-631).  Across a
// non-periodic face phi is read at the nearest interior node and n is 0; n is
// 0 on solid nodes.  The phi plane three ahead is loaded into registers while
// the current plane is computed.  RAW (colour model) writes the unnormalised
// gradient of the colour field psi and no curvature.
//
// The library launches it dense: `dense` writes n and kappa everywhere (and
// flags every tile-plane), RAW always is.  A plane whose phi window z-1..z+1
// is uniform skips the stencil; uniformity of each loaded plane is voted per
// wave from registers.  Its sparse mode (kappa only, each (tile, plane) with
// a flag byte, 1 when any of its kappa is nonzero, gated windows skipped)
// ran the phase-field step before the run-list pair k_iface_list +
// k_curv_list replaced it.
constexpr int NWAVE = NTHR / 64;
template <typename S, bool RAW>
__global__ __launch_bounds__(NTHR) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_normals_tiled(const S *__restrict__ phi,
                                                         float *__restrict__ n0,
                                                         float *__restrict__ n1,
                                                         float *__restrict__ n2,
                                                         float *__restrict__ kap,
                                                         uint8_t *__restrict__ flags,
                                                         const uint32_t *__restrict__ links,
                                                         int has_mask, int dense, Params p) {
  __shared__ float ph[4][NPSLOT];
  __shared__ float nl[4][3][NSLOT];
  __shared__ float wv[NWAVE][3];  // per wave: plane uniform (1/0), its value, interface (1/0)
  const int tid = threadIdx.x;
  const int tx = tid % TX, ty = tid / TX;
  int b = xcd_tile(blockIdx.x, gridDim.x);
  const int tile_x = b % p.tiles_x;
  b /= p.tiles_x;
  const int tile_y = b % p.tiles_y;
  const int zc = b / p.tiles_y;
  const int x0 = tile_x * TX, y0 = tile_y * TY;
  const int z0 = zc * p.z_chunk;
  const int z1 = min(z0 + p.z_chunk, p.nz);
  const int x = x0 + tx, y = y0 + ty;
  const bool active = x < p.nx && y < p.ny;
  const size_t plane = (size_t)p.nx * p.ny;
  const size_t tiles_plane = (size_t)p.tiles_x * p.tiles_y;
  // ring slot of plane z >= z0 - 2
  auto r4 = [z0](int z) { return (z - z0 + 4) & 3; };

  // phi slots of this thread (tid, tid + NTHR) as in-plane byte offsets
  const bool two = tid + NTHR < NPSLOT;
  auto phi_off = [&](int sl) -> uint32_t {
    const int gx = clampw(x0 + sl % PW - 2, p.nx, p.px);
    const int gy = clampw(y0 + sl / PW - 2, p.ny, p.py);
    return (uint32_t)(gx + p.nx * gy) * (uint32_t)sizeof(S);
  };
  const uint32_t pofs0 = phi_off(tid), pofs1 = two ? phi_off(tid + NTHR) : 0u;
  // is a phi slot one of this tile's own nodes?
  auto phi_own = [&](int sl) -> bool {
    const int sx = sl % PW - 2, sy = sl / PW - 2;
    return sx >= 0 && sx < TX && sy >= 0 && sy < TY && x0 + sx < p.nx && y0 + sy < p.ny;
  };
  const bool pown0 = phi_own(tid), pown1 = two && phi_own(tid + NTHR);
  const uint32_t pbS = (uint32_t)(plane * sizeof(S)), pb4 = (uint32_t)(plane * 4);
  // phi planes are loaded PQ planes ahead of their commit into a register
  // queue of raw storage values (converted at commit, so that no instruction
  // waits on a load before its plane is due)
  struct PhiQ {
    S a, b;
  };
  auto issue = [&](int z, PhiQ &q) {
    if (z > z1 + 1) return;  // planes up to z1 + 1 are read
    const rsrc_t base = rsrc(phi + plane * (size_t)clampw(z, p.nz, p.pz), pbS);
    q.a = bld<S>(base, pofs0);
    if (two) q.b = bld<S>(base, pofs1);
  };
  float pr0 = 0.f, pr1 = 0.f;
  auto take = [&](const PhiQ &q) {
    pr0 = dec(q.a, p, sgrp(TSLB_PHI));
    pr1 = two ? dec(q.b, p, sgrp(TSLB_PHI)) : 0.f;
  };
  // store the loaded plane z and post this wave's votes: are all its values
  // equal, and may one of the tile's own nodes carry curvature?  The gate
  // phi(1-phi)|n|^2 > 0.2 with |n| < 1 needs phi(1-phi) > 0.2; 0.19 leaves
  // room for rounding.
  auto commit = [&](int z) {
    float *d = ph[r4(z)];
    d[tid] = pr0;
    if (two) d[tid + NTHR] = pr1;
    const float v = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(pr0)));
    const bool all = __all(pr0 == v && (!two || pr1 == v));
    const bool hit = (pown0 && pr0 * (1.f - pr0) > 0.19f) || (pown1 && pr1 * (1.f - pr1) > 0.19f);
    const bool ihit = __any(hit);
    if ((tid & 63) == 0) {
      wv[tid >> 6][0] = all ? 1.f : 0.f;
      wv[tid >> 6][1] = v;
      wv[tid >> 6][2] = ihit ? 1.f : 0.f;
    }
  };
  // after a barrier: is the plane of the last commit one value (and which),
  // and does it hold interface nodes of this tile (iface)?
  auto read_vote = [&](float &val, bool &iface) -> bool {
    const float v0 = wv[0][1];
    bool u = true, i = false;
#pragma unroll
    for (int w = 0; w < NWAVE; w++) {
      u = u && wv[w][0] != 0.f && wv[w][1] == v0;
      i = i || wv[w][2] != 0.f;
    }
    val = v0;
    iface = i;
    return u;
  };

  // normal slots of this thread (tid, tid + NTHR < NSLOT)
  const bool ntwo = tid + NTHR < NSLOT;
  auto nslot = [&](int sl, int &c, uint32_t &xy, bool &ok) {
    const int sx = sl % SW, sy = sl / SW;
    int gx = 0, gy = 0;
    ok = src_coord(x0 + sx - 1, p.nx, p.px, gx) && src_coord(y0 + sy - 1, p.ny, p.py, gy);
    c = (sy + 1) * PW + (sx + 1);
    xy = ok ? (uint32_t)(gx + p.nx * gy) * 4u : 0u;
  };
  int nc0 = 0, nc1 = 0;
  uint32_t nxy0 = 0, nxy1 = 0;
  bool nok0 = false, nok1 = false;
  nslot(tid, nc0, nxy0, nok0);
  if (ntwo) nslot(tid + NTHR, nc1, nxy1, nok1);

  // normals of plane z on the (TX+2)x(TY+2) footprint, from phi planes z-1..z+1
  auto normals_plane = [&](int z, bool zero) {
    int gz = 0;
    const bool zin = src_coord(z, p.nz, p.pz, gz);
    const float *P0 = ph[r4(z - 1)], *P1 = ph[r4(z)], *P2 = ph[r4(z + 1)];
    float(*dn)[NSLOT] = nl[r4(z)];
    const rsrc_t lk = rsrc(links + plane * (size_t)gz, pb4);
    auto one = [&](int sl, int c, uint32_t xy, bool okxy) {
      float a = 0.f, bb = 0.f, cc = 0.f;
      if (!zero && zin && okxy) {
        bool ok = true;
        if (has_mask) ok = (bld_u32(lk, xy) & kSolid) == 0u;
        if (ok) sep_normal<PW, !RAW>(P0, P1, P2, c, a, bb, cc);
      }
      dn[0][sl] = a;
      dn[1][sl] = bb;
      dn[2][sl] = cc;
    };
    one(tid, nc0, nxy0, nok0);
    if (ntwo) one(tid + NTHR, nc1, nxy1, nok1);
  };

  // uniformity (u*), value (v*) and interface votes (i*) of phi planes
  // z-2 .. z+2 for the current z.  Sparse phase-field output (gate) needs the
  // curvature of plane z only where i(z), and so n of plane z only where one
  // of i(z-1), i(z), i(z+1) holds; elsewhere both stay zero.
  const bool gate = !RAW && !dense;
  bool um2, um1, u0, up1, up2;
  float vm2, vm1, v0, vp1, vp2;
  bool i0, ip1, ip2;
  bool us[4], is[4];
  float vs[4];
  PhiQ q0, q1, q2, q3;  // plane z lives in q[(z - z0 + 2) & 3]
  issue(z0 - 2, q0);
  issue(z0 - 1, q1);
  issue(z0, q2);
  issue(z0 + 1, q3);
#pragma unroll
  for (int k = 0; k < 4; k++) {
    PhiQ &q = k == 0 ? q0 : (k == 1 ? q1 : (k == 2 ? q2 : q3));
    take(q);
    commit(z0 - 2 + k);
    issue(z0 + 2 + k, q);
    __syncthreads();
    us[k] = read_vote(vs[k], is[k]);
    __syncthreads();
  }
  um2 = us[0];
  vm2 = vs[0];
  um1 = us[1];
  vm1 = vs[1];
  u0 = us[2];
  v0 = vs[2];
  up1 = us[3];
  vp1 = vs[3];
  i0 = is[2];
  ip1 = is[3];
  normals_plane(z0 - 1, (um2 && um1 && u0 && vm2 == vm1 && vm1 == v0) || (gate && !i0));
  normals_plane(z0, (um1 && u0 && up1 && vm1 == v0 && v0 == vp1) || (gate && !(i0 || ip1)));
  __syncthreads();
  // one destination plane z; q holds plane z + 2 and is refilled with z + 6
  auto body = [&](int z, PhiQ &q) {
    take(q);
    commit(z + 2);
    issue(z + 6, q);
    __syncthreads();
    up2 = read_vote(vp2, ip2);
    normals_plane(z + 1, (u0 && up1 && up2 && v0 == vp1 && vp1 == vp2) || (gate && !(i0 || ip1 || ip2)));
    __syncthreads();
    const bool zero_z = (um1 && u0 && up1 && vm1 == v0 && v0 == vp1) || (gate && !i0);
    int any = 0;
    float a = 0.f, bb = 0.f, c = 0.f, curv = 0.f;
    if (!zero_z && active) {
      const int own = (ty + 1) * SW + (tx + 1);
      const int zm = r4(z - 1), z0s = r4(z), zp = r4(z + 1);
      a = nl[z0s][0][own];
      bb = nl[z0s][1][own];
      c = nl[z0s][2][own];
      const float f = ph[z0s][(ty + 2) * PW + (tx + 2)];
      const float ind = f * (1.f - f) * (a * a + bb * bb + c * c);
      if (!RAW && ind > 0.2f) {
        auto nat = [&](int comp) {
          return [&, comp](int dx, int dy, int dz) -> float {
            const int zz = dz < 0 ? zm : (dz > 0 ? zp : z0s);
            return nl[zz][comp][(ty + 1 + dy) * SW + (tx + 1 + dx)];
          };
        };
        const float dxx = grad_axis<0>(nat(0));
        const float dyy = grad_axis<1>(nat(1));
        const float dzz = grad_axis<2>(nat(2));
        curv = -(dxx + dyy + dzz);
      }
      if constexpr (RAW) any = (a != 0.f) | (bb != 0.f) | (c != 0.f);
      else any = curv != 0.f;
    }
    // block-uniform: skip the vote for a uniform window
    const int flag = zero_z ? 0 : __syncthreads_or(any);
    if ((flag || dense) && active) {
      const size_t zb = plane * (size_t)z;
      const uint32_t off = (uint32_t)(x + p.nx * y) * 4u;
      if (RAW || dense) {
        bst(rsrc(n0 + zb, pb4), off, a);
        bst(rsrc(n1 + zb, pb4), off, bb);
        bst(rsrc(n2 + zb, pb4), off, c);
      }
      if constexpr (!RAW) bst(rsrc(kap + zb, pb4), off, curv);
    }
    if (tid < 2) flags[2 * (tiles_plane * z + tile_y * p.tiles_x + tile_x) + tid] = (uint8_t)(flag || dense);
    um2 = um1;
    vm2 = vm1;
    um1 = u0;
    vm1 = v0;
    u0 = up1;
    v0 = vp1;
    up1 = up2;
    vp1 = vp2;
    i0 = ip1;
    ip1 = ip2;
  };
  for (int z = z0; z < z1; z += 4) {
    body(z, q0);
    if (z + 1 >= z1) break;
    body(z + 1, q1);
    if (z + 2 >= z1) break;
    body(z + 2, q2);
    if (z + 3 >= z1) break;
    body(z + 3, q3);
  }
}

// Sparse curvature of the phase field in two passes.  The gate
// phi(1-phi)|n|^2 > 0.2 with |n| < 1 needs phi(1-phi) > 0.2 at the node
// itself (0.19 leaves room for rounding), and in the emulsion only a thin
// shell of tile-planes holds such nodes.
//
// k_iface_list: per (tile, plane), a vote over the tile's own nodes; runs of
// consecutive tile-planes that may carry curvature are appended to a list,
// every flag byte of the chunk is cleared.  A 64x8 block marches a z-chunk
// (at most 64 planes) with the phi plane four ahead in flight; each wave
// keeps a 64-bit mask of its planes with a hit (a ballot, no barrier), the
// block ORs the eight masks once at the end and appends its runs with one
// atomic on the counter (a vector atomic); runs are cut at every 8th plane
// of the chunk.  Entry: 64 (z_first * tiles per plane + tile) + run length - 1.
template <typename S>
__global__ __launch_bounds__(NTHR) void k_iface_list(const S *__restrict__ phi, uint8_t *__restrict__ flags,
                                                     int *__restrict__ cnt, int *__restrict__ list, Params p) {
  __shared__ uint64_t wmask[2 * (NTHR / 64)];
  __shared__ int base;
  const int tid = threadIdx.x;
  const int tx = tid % TX, ty = tid / TX;
  int b = xcd_tile(blockIdx.x, gridDim.x);
  const int tile_x = b % p.tiles_x;
  b /= p.tiles_x;
  const int tile_y = b % p.tiles_y;
  const int zc = b / p.tiles_y;
  const int z0 = zc * p.z_chunk;
  const int z1 = min(z0 + p.z_chunk, p.nz);
  const int x = tile_x * TX + tx, y = tile_y * TY + ty;
  const bool active = x < p.nx && y < p.ny;
  const size_t plane = (size_t)p.nx * p.ny;
  const int tiles_plane = p.tiles_x * p.tiles_y;
  const int tile = tile_y * p.tiles_x + tile_x;
  const uint32_t pbS = (uint32_t)(plane * sizeof(S));
  const uint32_t off = active ? (uint32_t)(x + p.nx * y) * (uint32_t)sizeof(S) : kOOB;
  auto ld = [&](int z) -> S {
    if (z >= z1) return S{};
    return bld<S>(rsrc(phi + plane * (size_t)z, pbS), off);
  };
  uint64_t mask0 = 0, mask1 = 0;  // wave-uniform, per half-tile: bit k for plane z0 + k
  S q0 = ld(z0), q1 = ld(z0 + 1), q2 = ld(z0 + 2), q3 = ld(z0 + 3);
  auto body = [&](int z, S &q) {
    const float f = dec(q, p, sgrp(TSLB_PHI));
    q = ld(z + 4);
    const bool hit = active && f * (1.f - f) > 0.19f;
    const uint64_t bal = __ballot(hit);
    if ((uint32_t)bal) mask0 |= 1ull << (z - z0);
    if (bal >> 32) mask1 |= 1ull << (z - z0);
  };
  for (int z = z0; z < z1; z += 4) {
    body(z, q0);
    if (z + 1 >= z1) break;
    body(z + 1, q1);
    if (z + 2 >= z1) break;
    body(z + 2, q2);
    if (z + 3 >= z1) break;
    body(z + 3, q3);
  }
  if ((tid & 63) == 0) {
    wmask[2 * (tid >> 6)] = mask0;
    wmask[2 * (tid >> 6) + 1] = mask1;
  }
  if (tid < 2 * (z1 - z0)) flags[2 * ((z0 + (tid >> 1)) * tiles_plane + tile) + (tid & 1)] = 0;
  __syncthreads();
  append_runs(wmask, &base, tid, z0, z1 - z0, tiles_plane, tile, cnt, list);
}

// k_curv_list: curvature of the listed runs of tile-planes.  Per run a block
// marches its planes as k_normals_tiled does: phi of planes z-3..z+3 on the
// (TX+4)x(TY+4) footprint in a six-plane LDS ring (the plane three ahead
// loads into registers meanwhile), n of planes z-1..z+2 on the (TX+2)x(TY+2)
// footprint in a four-plane ring, kappa = -div(n) at the own nodes past the
// gate -- the arithmetic of k_normals_tiled, so the values are the same bits.
// Every listed tile-plane gets its kappa (zeros included) and flag byte 1,
// so one barrier per plane suffices.  Blocks stride over the list (its
// length read from the counter), so every wave ends its loop.  The plane
// loop is unrolled over the longest run (ring slots at constant offsets) and
// held to 80 VGPRs, which keeps three blocks per CU (51 KiB of LDS each).
template <typename S>
__global__ __launch_bounds__(HNT) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_curv_list(const S *__restrict__ phi, float *__restrict__ kap,
                                                    uint8_t *__restrict__ flags, const int *__restrict__ cnt,
                                                    const int *__restrict__ list, const uint32_t *__restrict__ links,
                                                    int has_mask, Params p) {
  // rings deep enough that a slot is rewritten only two barriers after its
  // last read, so one barrier per plane orders every LDS access
  __shared__ float ph[6][HNPSLOT];
  __shared__ float nl[4][3][HNSLOT];
  const int tid = threadIdx.x;
  const int tx = tid % HX, ty = tid / HX;
  const size_t plane = (size_t)p.nx * p.ny;
  const int tiles_plane = p.tiles_x * p.tiles_y;
  const uint32_t pbS = (uint32_t)(plane * sizeof(S)), pb4 = (uint32_t)(plane * 4);
  const bool two = tid + HNT < HNPSLOT;
  const bool ntwo = tid + HNT < HNSLOT;
  const int n_entries = *cnt;
  for (int e = blockIdx.x; e < n_entries; e += gridDim.x) {
    const int ent = list[e];
    const int len = (ent & 63) + 1, half = (ent >> 6) & 1, tp0 = ent >> 7;
    const int zs = tp0 / tiles_plane, tile = tp0 - zs * tiles_plane;
    const int x0 = (tile % p.tiles_x) * TX + half * HX, y0 = (tile / p.tiles_x) * TY;
    const int x = x0 + tx, y = y0 + ty;
    const bool active = x < p.nx && y < p.ny;
    auto phi_off = [&](int sl) -> uint32_t {
      const int gx = clampw(x0 + sl % HPW - 2, p.nx, p.px);
      const int gy = clampw(y0 + sl / HPW - 2, p.ny, p.py);
      return (uint32_t)(gx + p.nx * gy) * (uint32_t)sizeof(S);
    };
    const uint32_t o0 = phi_off(tid), o1 = two ? phi_off(tid + HNT) : 0u;
    auto rph = [zs](int z) { return (z - zs + 12) % 6; };  // phi planes z >= zs - 2
    auto rnl = [zs](int z) { return (z - zs + 4) & 3; };    // normal planes z >= zs - 1
    S qa{}, qb{};
    auto issue_to = [&](int z, S &ra, S &rb) {
      const rsrc_t r = rsrc(phi + plane * (size_t)clampw(z, p.nz, p.pz), pbS);
      ra = bld<S>(r, o0);
      if (two) rb = bld<S>(r, o1);
    };
    auto issue = [&](int z) { issue_to(z, qa, qb); };
    auto commit_from = [&](int z, S ra, S rb) {
      const float a = dec(ra, p, sgrp(TSLB_PHI)), c = two ? dec(rb, p, sgrp(TSLB_PHI)) : 0.f;
      ph[rph(z)][tid] = a;
      if (two) ph[rph(z)][tid + HNT] = c;
    };
    auto commit = [&](int z) { commit_from(z, qa, qb); };
    // n of plane zz from the in-plane sums of phi plane zz+1, the centre of
    // plane zz-1 and the z-march
    // of each normal slot (g0: slot tid, g1: slot tid + HNT)
    const int c0 = (tid / HSW + 1) * HPW + tid % HSW + 1;
    const int c1 = ntwo ? ((tid + HNT) / HSW + 1) * HPW + (tid + HNT) % HSW + 1 : c0;
    GradRun g0, g1;
    // per normal slot, fixed over the run: in-plane validity and the byte
    // offset of its link word
    auto slot_xy = [&](int sl, uint32_t &lofs) -> bool {
      int gx = 0, gy = 0;
      const bool ok = src_coord(x0 + sl % HSW - 1, p.nx, p.px, gx) && src_coord(y0 + sl / HSW - 1, p.ny, p.py, gy);
      lofs = (uint32_t)(gx + p.nx * gy) * 4u;
      return ok;
    };
    uint32_t lofs0 = 0, lofs1 = 0;
    const bool okxy0 = slot_xy(tid, lofs0);
    const bool okxy1 = ntwo && slot_xy(tid + HNT, lofs1);
    auto normals = [&](int zz) {
      int gz = 0;
      const bool zin = src_coord(zz, p.nz, p.pz, gz);
      const rsrc_t lk = rsrc(links + plane * (size_t)gz, pb4);
      const float *P0 = ph[rph(zz - 1)], *P2 = ph[rph(zz + 1)];
      float(*dn)[HNSLOT] = nl[rnl(zz)];
      auto one = [&](int sl, int c, GradRun &g, bool okxy, uint32_t lofs) {
        float nn[3];
        grad_run_step(g, plane_grad<HPW>(P2, c), P0[c], nn);
        bool ok = zin && okxy;
        if (ok && has_mask) ok = (bld_u32(lk, lofs) & kSolid) == 0u;
        dn[0][sl] = ok ? nn[0] : 0.f;
        dn[1][sl] = ok ? nn[1] : 0.f;
        dn[2][sl] = ok ? nn[2] : 0.f;
      };
      one(tid, c0, g0, okxy0, lofs0);
      if (ntwo) one(tid + HNT, c1, g1, okxy1, lofs1);
    };
    // prologue: phi planes zs-2..zs+2 (all five loads in flight at once), n
    // of planes zs-1 and zs
    {
      S pa[5]{}, pb[5]{};
#pragma unroll
      for (int k = 0; k < 5; k++) issue_to(zs - 2 + k, pa[k], pb[k]);
      if (len > 1) issue(zs + 3);
#pragma unroll
      for (int k = 0; k < 5; k++) commit_from(zs - 2 + k, pa[k], pb[k]);
    }
    __syncthreads();
    grad_run_init(g0, plane_grad<HPW>(ph[rph(zs - 2)], c0), plane_grad<HPW>(ph[rph(zs - 1)], c0));
    if (ntwo) grad_run_init(g1, plane_grad<HPW>(ph[rph(zs - 2)], c1), plane_grad<HPW>(ph[rph(zs - 1)], c1));
    normals(zs - 1);
    normals(zs);
    // unrolled over the longest run, so that every ring slot index is a
    // constant offset from the run's start
#pragma unroll
    for (int k = 0; k < kRunMax; k++) {
      if (k >= len) break;
      const int z = zs + k;
      normals(z + 1);
      // phi of plane z+3 into the slot of plane z-3, whose last reads (n of
      // plane z-2, the gate of plane z-3) came before the previous barrier
      if (k + 1 < len) {
        commit(z + 3);
        if (k + 2 < len) issue(z + 4);
      }
      __syncthreads();
      float curv = 0.f;
      if (active) {
        const int own = (ty + 1) * HSW + (tx + 1);
        const float(*n0)[HNSLOT] = nl[rnl(z)];
        const float a = n0[0][own], bb = n0[1][own], c = n0[2][own];
        const float f = ph[rph(z)][(ty + 2) * HPW + (tx + 2)];
        const float ind = f * (1.f - f) * (a * a + bb * bb + c * c);
        if (ind > 0.2f) {
          const int zm = rnl(z - 1), z0s = rnl(z), zp = rnl(z + 1);
          auto nat = [&](int comp) {
            return [&, comp](int dx, int dy, int dz) -> float {
              const int zz = dz < 0 ? zm : (dz > 0 ? zp : z0s);
              return nl[zz][comp][(ty + 1 + dy) * HSW + (tx + 1 + dx)];
            };
          };
          const float dxx = grad_axis<0>(nat(0));
          const float dyy = grad_axis<1>(nat(1));
          const float dzz = grad_axis<2>(nat(2));
          curv = -(dxx + dyy + dzz);
        }
      }
      // every listed tile-plane is flagged and gets its kappa (zeros
      // included): no block-wide vote, no second barrier
      if (active) bst(rsrc(kap + plane * (size_t)z, pb4), (uint32_t)(x + p.nx * y) * 4u, curv);
      if (tid == 0) flags[2 * ((size_t)z * tiles_plane + tile) + half] = 1;
    }
    __syncthreads();  // LDS reuse by the next entry
  }
}

// plain copy for bandwidth calibration: one float4 per thread (16 B per lane,
// fully coalesced), launched with a grid that covers the buffer in one pass.
// On a 5.9 GB buffer this moves 6.6 TB/s, where a grid-stride loop over 8192
// blocks of 4 float4 per thread moves 5.4 TB/s: short-lived workgroups keep
// more loads in flight across the 8 XCDs than persistent ones.
typedef float v4f __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_copy(const v4f *__restrict__ a, v4f *__restrict__ b,
                                              long long n4) {
  const long long stride = (long long)gridDim.x * 256;
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n4; i += stride)
    __builtin_nontemporal_store(__builtin_nontemporal_load(a + i), b + i);
}

// ---------------------------------------------------------------------------
// Initial phi: one advection by the unsharpened D3Q7 equilibrium
// (recursive_2c_TSLB3D.f90:256-283 writes g(i+1,..,1) etc.), bounce-back at
// invalid sources.
// ---------------------------------------------------------------------------
template <typename S>
__global__ __launch_bounds__(256) void k_prepush(const S *__restrict__ phi,
                                                 const S *__restrict__ u,
                                                 const S *__restrict__ v,
                                                 const S *__restrict__ w,
                                                 S *__restrict__ out,
                                                 const uint32_t *__restrict__ links, int has_mask,
                                                 Params p) {
  const int n = p.nx * p.ny * p.nz;
  const int id = blockIdx.x * 256 + threadIdx.x;
  if (id >= n) return;
  if (has_mask && (links[id] & kSolid)) {
    out[id] = phi[id];
    return;
  }
  const int x = id % p.nx;
  const int y = (id / p.nx) % p.ny;
  const int z = id / (p.nx * p.ny);
  const S *vel[3] = {u, v, w};
  float own[7];
  const float ph0 = ldf(phi, id, p, TSLB_PHI);
  own[0] = 0.25f * ph0;
  for (int a = 0; a < 3; a++) {
    const float ua = ldf(vel[a], id, p, TSLB_U + a);
    own[1 + 2 * a] = 0.125f * ph0 * (1.f + 3.f * ua);
    own[2 + 2 * a] = 0.125f * ph0 * (1.f - 3.f * ua);
  }
  float s = own[0];
  for (int l = 1; l <= 6; l++) {
    int sx, sy, sz;
    bool ok = src_coord(x - kCX[l], p.nx, p.px, sx) && src_coord(y - kCY[l], p.ny, p.py, sy) &&
              src_coord(z - kCZ[l], p.nz, p.pz, sz);
    int sid = ok ? sx + p.nx * sy + p.nx * p.ny * sz : 0;
    if (ok && has_mask && (links[sid] & kSolid)) ok = false;
    if (!ok) {
      if (!p.gwall) s += own[kOPP[l]];
    } else {
      const int a = (l - 1) / 2;
      const float sgn = (l & 1) ? 1.f : -1.f;
      s += 0.125f * ldf(phi, sid, p, TSLB_PHI) * (1.f + sgn * 3.f * ldf(vel[a], sid, p, TSLB_U + a));
    }
  }
  stf(out, id, s, p, TSLB_PHI);
}

// ---------------------------------------------------------------------------
// Link masks from isfluid: bit l set when the source y - c_l of fluid node y
// is a solid node; solid nodes carry kSolid and all direction bits.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_links(const int32_t *__restrict__ isf,
                                               uint32_t *__restrict__ links, Params p) {
  const int n = p.nx * p.ny * p.nz;
  const int id = blockIdx.x * 256 + threadIdx.x;
  if (id >= n) return;
  if (isf[id] == 0) {
    links[id] = kSolid | kAllDirs;
    return;
  }
  const int x = id % p.nx;
  const int y = (id / p.nx) % p.ny;
  const int z = id / (p.nx * p.ny);
  uint32_t m = 0;
  for (int l = 1; l < 27; l++) {
    int sx, sy, sz;
    if (src_coord(x - kCX[l], p.nx, p.px, sx) && src_coord(y - kCY[l], p.ny, p.py, sy) &&
        src_coord(z - kCZ[l], p.nz, p.pz, sz)) {
      if (isf[sx + p.nx * sy + p.nx * p.ny * sz] == 0) m |= 1u << l;
    }
  }
  links[id] = m;
}

// ---------------------------------------------------------------------------
// Moment-space z-plane BC (TSLB_newBCS/boundary_cds_3D_module.f90:1233-1292):
// plane 0 <- plane 1 with u = v = 0 and w = w_in; plane nz-1 <- plane nz-2.
// ---------------------------------------------------------------------------
template <typename S>
__global__ __launch_bounds__(256) void k_zplane_bc(FieldsT<S> f, const float *__restrict__ w_in,
                                                   const float *__restrict__ phi_in, int ncomp,
                                                   Params p) {
  const int plane = p.nx * p.ny;
  const int id = blockIdx.x * 256 + threadIdx.x;
  if (id >= plane) return;
  const int lo = id, lo1 = id + plane;
  const int hi = id + plane * (p.nz - 1), hi1 = id + plane * (p.nz - 2);
  const int nf = ncomp >= 2 ? TSLB_NFIELDS : 10;
  for (int q = 0; q < nf; q++) f.f[q][hi] = f.f[q][hi1];
  for (int q = 0; q < nf; q++) {
    if (q == TSLB_U || q == TSLB_V) stf(f.f[q], lo, 0.f, p, q);
    else if (q == TSLB_W) stf(f.f[q], lo, w_in[id], p, q);
    else if (q == TSLB_PHI && phi_in) stf(f.f[q], lo, phi_in[id], p, q);
    else f.f[q][lo] = f.f[q][lo1];
  }
}

// ---------------------------------------------------------------------------
// Turbulent-jet boundary (bcs_2c_turbulent_jet, boundary_cds_3D_module.f90:
// 862-1049).  Planes 0 (inlet) and nz-1 (outlet) are walls; k_fused has
// bounced back the populations that cross from them into planes 1 and nz-2
// (and, with Params::gwall, let no phase population cross).  This pass
// replaces, in moment space, each such population of a fluid node y by the
// regularized boundary value
//   h_l(y) = feq_l(rho(y), u(x)) + (1-omega) w_l a^neq_l(Pi(y), u(y)) - w_l,
// x = y - c_l on the wall plane (:932-983, :991-1041), by subtracting its
// bounce-back value h_opp(l)(y) from the shifted sums of y and adding the new
// one, then refinishing the moments.  At the outlet u(x) takes w from the
// plane below x (:992), and that w is stored on the outlet plane.  Phase
// field: g_5(y) = phi(x)(1 + 3 w(x))/8 at the inlet (:985-988), g_6(y) =
// phi(y)(1 - 3*0.05)/8 at the outlet (:1046-1048).  Populations whose x lies
// across a non-periodic x/y face keep their bounce-back value.  One thread per
// column and side; time-t moments from src, time-t+1 moments in dst.
// ---------------------------------------------------------------------------
template <int NC, bool ADD, typename S>
__global__ __launch_bounds__(256) void k_jet_bc(const S *__restrict__ src, S *__restrict__ dst,
                                                const float *__restrict__ nrm,
                                                const uint8_t *__restrict__ flags,
                                                const uint32_t *__restrict__ links, Params p) {
  const int plane = p.nx * p.ny;
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= 2 * plane) return;
  const int side = t / plane, xy = t - side * plane;
  const int i = xy % p.nx, j = xy / p.nx;
  const int zy = side ? p.nz - 2 : 1, zx = side ? p.nz - 1 : 0;
  const int sz = side ? -1 : 1;  // c_z of the populations crossing into plane zy
  auto at = [&](int q, int id) -> float { return ldf(src + (size_t)q * p.fs, id, p, q); };
  if (side) {
    // outlet wall plane: w(x) = w(x - e_z), boundary_cds_3D_module.f90:992
    stf(dst + (size_t)TSLB_W * p.fs, xy + plane * zx, at(TSLB_W, xy + plane * zy), p, TSLB_W);
  }
  const int y = xy + plane * zy;
  const uint32_t ly = links[y];
  if (ly & kSolid) return;

  // time-t state of y and its shifted post-collision populations, as k_fused
  // evaluated them for the bounce-back
  Src<NC> sy;
#pragma unroll
  for (int q = 0; q < (NC == 2 ? 11 : 10); q++) sy.m[q] = at(q, y);
  if constexpr (NC == 2) {
    // n from phi as the fused pass computes it; planes zy +- 1 are in range
    float ph3[3][9];
#pragma unroll
    for (int dz = -1; dz <= 1; dz++)
#pragma unroll
      for (int dy = -1; dy <= 1; dy++)
#pragma unroll
        for (int dx = -1; dx <= 1; dx++) {
          const int gi = clampw(i + dx, p.nx, p.px), gj = clampw(j + dy, p.ny, p.py);
          ph3[dz + 1][(dy + 1) * 3 + dx + 1] = at(TSLB_PHI, gi + p.nx * gj + plane * (zy + dz));
        }
    sep_normal<3>(ph3[0], ph3[1], ph3[2], 4, sy.nn[0], sy.nn[1], sy.nn[2]);
    const int tile = (j / TY) * p.tiles_x + i / TX;
    const bool un = flags[2 * ((size_t)zy * p.tiles_x * p.tiles_y + tile) + (i % TX) / HX] != 0;
    sy.kap = un ? nrm[3 * (size_t)p.fs + y] : 0.f;
  }
  ArraySink own;
  float gq[7];
  eval_src<NC, ADD>(sy, p, own, gq);

  // shifted sums of y from its finished moments
  auto dat = [&](int q) -> float { return ldf(dst + (size_t)q * p.fs, y, p, q); };
  const float s0 = dat(TSLB_RHO), rho = 1.f + s0;
  const float u1 = dat(TSLB_U), v1 = dat(TSLB_V), w1 = dat(TSLB_W);
  const float t3 = s0 * (1.f / 3.f);
  float sm[10] = {s0,
                 rho * u1,
                 rho * v1,
                 rho * w1,
                 dat(TSLB_PXX) + t3 + rho * u1 * u1,
                 dat(TSLB_PYY) + t3 + rho * v1 * v1,
                 dat(TSLB_PZZ) + t3 + rho * w1 * w1,
                 dat(TSLB_PXY) + rho * u1 * v1,
                 dat(TSLB_PXZ) + rho * u1 * w1,
                 dat(TSLB_PYZ) + rho * v1 * w1};
  bool changed = false;
#pragma unroll
  for (int l = 1; l < 27; l++) {
    if (kCZ[l] != sz || !((ly >> l) & 1u)) continue;
    int xi = 0, xj = 0;
    if (!src_coord(i - kCX[l], p.nx, p.px, xi) || !src_coord(j - kCY[l], p.ny, p.py, xj)) continue;
    const int x = xi + p.nx * xj + plane * zx;
    const float ux = at(TSLB_U, x), vx = at(TSLB_V, x);
    const float wx = side ? at(TSLB_W, xi + p.nx * xj + plane * zy) : at(TSLB_W, x);
    ArraySink bnd;
    all_pops_mt<false>(sy.m, ux, vx, wx, p, bnd, p.om1);
    const float d = bnd.h[l] - own.h[kOPP[l]];
    const float cx = (float)kCX[l], cy = (float)kCY[l], cz = (float)kCZ[l];
    sm[0] += d;
    sm[1] += cx * d;
    sm[2] += cy * d;
    sm[3] += cz * d;
    sm[4] += cx * cx * d;
    sm[5] += cy * cy * d;
    sm[6] += cz * cz * d;
    sm[7] += cx * cy * d;
    sm[8] += cx * cz * d;
    sm[9] += cy * cz * d;
    changed = true;
  }
  if (changed) {
    const float r = 1.f + sm[0];
    const float inv = 1.f / r;
    const float u = sm[1] * inv, v = sm[2] * inv, w = sm[3] * inv;
    const float tt = sm[0] * (1.f / 3.f);
    const float out[10] = {sm[0], u, v, w, sm[4] - tt - r * u * u, sm[5] - tt - r * v * v,
                           sm[6] - tt - r * w * w, sm[7] - r * u * v, sm[8] - r * u * w, sm[9] - r * v * w};
#pragma unroll
    for (int q = 0; q < 10; q++) stf(dst + (size_t)q * p.fs, y, out[q], p, q);
  }
  if constexpr (NC == 2) {
    float add;
    if (!side) {
      const int x = xy + plane * zx;
      add = 0.125f * at(TSLB_PHI, x) * (1.f + 3.f * at(TSLB_W, x));
    } else {
      add = 0.125f * sy.m[TSLB_PHI] * (1.f - 3.f * 0.05f);
    }
    stf(dst + (size_t)TSLB_PHI * p.fs, y, dat(TSLB_PHI) + add, p, TSLB_PHI);
  }
}

// ---------------------------------------------------------------------------
// Counter-based RNG on the device (TSLB_newBCS/vars_single_3D_module.f90:
// 102-165); integer-exact, identical to the host tslb_rand_noseeded.
// ---------------------------------------------------------------------------
__host__ __device__ inline float rand_noseeded_impl(int i, int j, int k, int l) {
  int isub = i % 178 + 1, jsub = j % 178 + 1, ksub = k % 178 + 1;
  int lsub = l % 169, msub;
  const float cc = 362436.0f / 16777216.0f;
  const float cd = 7654321.0f / 16777216.0f;
  const float cmod = 16777213.0f / 16777216.0f;
  float u[2];
  for (int pass = 0; pass < 2; pass++) {
    float s = 0.0f, t = 0.5f;
    for (int jj = 0; jj < 24; jj++) {
      msub = ((isub * jsub) % 179 * ksub) % 179;
      isub = jsub;
      jsub = ksub;
      ksub = msub;
      lsub = (53 * lsub + 1) % 169;
      if ((lsub * msub) % 64 >= 32) s = s + t;
      t = 0.5f * t;
    }
    u[pass] = s;
  }
  float uni = u[0] - u[1];
  if (uni < 0.0f) uni = uni + 1.0f;
  float csub = cc - cd;
  if (csub < 0.0f) csub = csub + cmod;
  uni = uni - csub;
  if (uni < 0.0f) uni = uni + 1.0f;
  return uni;
}

// field += amp * (2 r - 1), r = rand_noseeded(i, j, k, l) with 1-based i, j, k
// (seeding as TSLB_2C-mpi/lb_kernels_module.f90:970-977)
template <typename S>
__global__ __launch_bounds__(256) void k_seed_noise(S *__restrict__ a, int q, float amp, int l, Params p) {
  const int n = p.nx * p.ny * p.nz;
  const int id = blockIdx.x * 256 + threadIdx.x;
  if (id >= n) return;
  const int x = id % p.nx;
  const int y = (id / p.nx) % p.ny;
  const int z = id / (p.nx * p.ny);
  const float r = rand_noseeded_impl(x + 1, y + 1, z + 1, l);
  stf(a, id, ldf(a, id, p, q) + amp * (2.0f * r - 1.0f), p, q);
}

// host-format float <-> storage, with the field's scale; field 0 carries rho-1
template <typename S>
__global__ __launch_bounds__(256) void k_to_storage(const float *__restrict__ in, S *__restrict__ out,
                                                    int n, int q, Params p) {
  const int id = blockIdx.x * 256 + threadIdx.x;
  if (id >= n) return;
  const float v = q == TSLB_RHO ? in[id] - 1.0f : in[id];
  stf(out, id, v, p, q);
}
template <typename S>
__global__ __launch_bounds__(256) void k_from_storage(const S *__restrict__ in, float *__restrict__ out,
                                                      int n, int q, Params p) {
  const int id = blockIdx.x * 256 + threadIdx.x;
  if (id >= n) return;
  const float v = ldf(in, id, p, q);
  out[id] = q == TSLB_RHO ? 1.0f + v : v;
}

__global__ __launch_bounds__(256) void k_fill(float *__restrict__ a, float v, int n) {
  const int id = blockIdx.x * 256 + threadIdx.x;
  if (id < n) a[id] = v;
}

#include "closure_kernels.h"

// ---------------------------------------------------------------------------
// FP64 diagnostics: per-block partial sums, then one block sums the partials.
// ---------------------------------------------------------------------------
constexpr int NRED = 7;
template <typename S>
__global__ __launch_bounds__(256) void k_reduce_part(CFieldsT<S> f, const uint32_t *__restrict__ links,
                                                     int has_mask, Params p,
                                                     double *__restrict__ part) {
  __shared__ double sh[NRED][256];
  const int n = p.nx * p.ny * p.nz;
  double acc[NRED] = {0, 0, 0, 0, 0, 0, 0};
  for (int id = blockIdx.x * 256 + threadIdx.x; id < n; id += gridDim.x * 256) {
    if (has_mask && (links[id] & kSolid)) continue;
    const int x = id % p.nx;
    const int y = (id / p.nx) % p.ny;
    const int z = id / (p.nx * p.ny);
    const double rho = 1.0 + (double)ldf(f.f[0], id, p, 0);
    const double u = ldf(f.f[1], id, p, 1), v = ldf(f.f[2], id, p, 2), w = ldf(f.f[3], id, p, 3);
    acc[0] += rho;
    if (p.ncomp == 2) acc[1] += (double)ldf(f.f[TSLB_PHI], id, p, TSLB_PHI);
    if (p.ncomp == TSLB_COLOUR) acc[1] += 0.5 * rho * (1.0 + (double)ldf(f.f[TSLB_PHI], id, p, TSLB_PHI));
    acc[2] += 0.5 * rho * (u * u + v * v + w * w);
    acc[3] += rho * u;
    acc[4] += rho * v;
    acc[5] += rho * w;
    auto at = [&](int q, int dx, int dy, int dz) -> double {
      int xx = x + dx, yy = y + dy, zz = z + dz;
      if (xx < 0 || xx >= p.nx) xx = p.px ? wrapc(xx, p.nx) : (xx < 0 ? 0 : p.nx - 1);
      if (yy < 0 || yy >= p.ny) yy = p.py ? wrapc(yy, p.ny) : (yy < 0 ? 0 : p.ny - 1);
      if (zz < 0 || zz >= p.nz) zz = p.pz ? wrapc(zz, p.nz) : (zz < 0 ? 0 : p.nz - 1);
      return (double)ldf(f.f[q], xx + p.nx * yy + p.nx * p.ny * zz, p, q);
    };
    const double wx = 0.5 * (at(3, 0, 1, 0) - at(3, 0, -1, 0)) - 0.5 * (at(2, 0, 0, 1) - at(2, 0, 0, -1));
    const double wy = 0.5 * (at(1, 0, 0, 1) - at(1, 0, 0, -1)) - 0.5 * (at(3, 1, 0, 0) - at(3, -1, 0, 0));
    const double wz = 0.5 * (at(2, 1, 0, 0) - at(2, -1, 0, 0)) - 0.5 * (at(1, 0, 1, 0) - at(1, 0, -1, 0));
    acc[6] += 0.5 * (wx * wx + wy * wy + wz * wz);
  }
  for (int q = 0; q < NRED; q++) sh[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int q = 0; q < NRED; q++) sh[q][threadIdx.x] += sh[q][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int q = 0; q < NRED; q++) part[(size_t)blockIdx.x * NRED + q] = sh[q][0];
}

__global__ __launch_bounds__(256) void k_reduce_final(const double *__restrict__ part, int nblk,
                                                      double *__restrict__ out) {
  __shared__ double sh[NRED][256];
  double acc[NRED] = {0, 0, 0, 0, 0, 0, 0};
  for (int b = threadIdx.x; b < nblk; b += 256)
    for (int q = 0; q < NRED; q++) acc[q] += part[(size_t)b * NRED + q];
  for (int q = 0; q < NRED; q++) sh[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int q = 0; q < NRED; q++) sh[q][threadIdx.x] += sh[q][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int q = 0; q < NRED; q++) out[q] = sh[q][0];
}

// Scratch of tslb_reduce_noalloc: the per-block partial sums of
// k_reduce_part, in a file-scope device array so that the entry point
// allocates nothing.  Its grid is capped at kScratchBlocks and strides over
// any lattice.  Every GPU the library runs on holds its own copy (57 KiB) at
// its own address until the process ends; the host never uses the array's
// name as an address, only the pointer hipGetSymbolAddress returns for the
// current device (scratch_setup, called by tslb_create).
constexpr int kScratchBlocks = 1024;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double scratch[kScratchBlocks * NRED];

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

// device address of `scratch` on each device ordinal, looked up once with
// that device current
double *g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mu;

int scratch_setup(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return fail(TSLB_E_ARG, "device ordinal beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch[dev]) return TSLB_OK;
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return fail(TSLB_E_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
  g_scratch[dev] = (double *)ptr;
  return TSLB_OK;
}

#define HIPCK(expr)                                                                 \
  do {                                                                              \
    hipError_t e_ = (expr);                                                         \
    if (e_ != hipSuccess)                                                           \
      return fail(TSLB_E_HIP, std::string(#expr " failed: ") + hipGetErrorString(e_)); \
  } while (0)

constexpr int kRedBlocks = 1024;

}  // namespace

struct tslb_ctx {
  tslb_params prm;
  Params kp;
  hipStream_t stream = nullptr;
  size_t n = 0, fs = 0;  // nodes, field stride (elements)
  size_t es = 4;         // bytes per stored moment value
  void *buf[2] = {nullptr, nullptr};
  int cur = 0;
  int phi_buf = 0;  // one component: the buffer that holds phi
  float *nrm = nullptr;  // n0, n1, n2, kappa: 4 * fs floats
  uint8_t *flags = nullptr;  // per (half-tile, plane): kappa (phase field) or gradient nonzero there
  int *ilist = nullptr;      // phase field: counter, then the runs of half-tile-planes that may carry curvature
  uint8_t *flags_alt = nullptr;  // phase field: flags of the next step (cleared by the fused pass)
  int list_ok = 0;           // phase field: ilist describes the current state
  int ncu = 256;             // compute units of the device
  int curv_grid = 1536;      // k_curv_list blocks: as many as are resident at once
  uint32_t *links = nullptr;
  int has_mask = 0;
  int normals_fresh = 0;
  double *part = nullptr, *red = nullptr;
  float *w_in = nullptr, *phi_in = nullptr;
  unsigned int *sat = nullptr;  // FX16: clipped-value counter (kp.sat)
  int zbc = 0, zbc_phi = 0;
  int jet = 0;     // turbulent-jet inlet/outlet planes (k_jet_bc after k_fused)
  int zwalls = 0;  // bit 0: plane 0 all solid; bit 1: plane nz-1 all solid
  int variant = 2;  // bit 1: cap VGPRs at 128 (two blocks per CU); bit 2: for comparisons, the
                    // phase-field FP32 kernel with the full 27-population ring evaluation
  // learned closure: 0 none, 1 fixed per-node fields, 2 recomputed before
  // every collision by `model` from the current state
  int clo_mode = 0;
  float *clo = nullptr;    // nu_t, dsigma[6]: 7 * fs floats
  float *cfeat = nullptr;  // closure inputs of every node: 10 * n floats
  struct tslb_closure *model = nullptr;
  size_t bytes = 0;
  int timing = 0;
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending[2];
  double ms[2] = {0, 0};
  int64_t launches[2] = {0, 0};
};

namespace {

// Default FX16 steps 2^-m per group, range 32767 2^-m: rho - 1 2^-17 (within
// +-1/4: the start-up pressure waves of a sharp interface reach ~0.1),
// velocities 2^-17 (+-1/4, at the lattice's stable limit), Pi^neq 2^-18
// (+-1/8), phi (or the colour psi) 2^-14 (+-2: a sharp droplet's start-up
// overshoots phi to ~1.15).  Clipped values are counted
// (tslb_storage_clipped).
// the encode scale sc of field group g and its decode isc = 1 / sc (both
// powers of two for FX16)
void set_scale(Params &k, int g, float sc) {
  k.sc[g] = sc;
  k.isc[g] = 1.f / sc;
}

int fx16_shift(int g) {
  if (g == sgrp(TSLB_RHO)) return 17;
  if (g == sgrp(TSLB_U)) return 17;
  if (g == sgrp(TSLB_PXX)) return 18;
  return 14;
}
// the finest power-of-two step 2^-m whose range 32767 2^-m covers R
int fx16_shift_for(float half_range) {
  int m = (int)floorf(log2f(32767.f / half_range));
  return m < -30 ? -30 : (m > 40 ? 40 : m);
}

// call fn with a value of the storage type of the context
template <typename Fn>
auto with_storage(const tslb_ctx *c, Fn &&fn) {
  switch (c->prm.storage) {
    case TSLB_STORAGE_FP16: return fn(_Float16{});
    case TSLB_STORAGE_BF16: return fn(bf16_t{});
    case TSLB_STORAGE_FX16: return fn(fx16_t{});
    default: return fn(float{});
  }
}

template <typename S>
S *field_ptr(void *base, size_t fs, int q) {
  return (S *)base + (size_t)q * fs;
}
template <typename S>
FieldsT<S> fields_of(void *base, size_t fs) {
  FieldsT<S> f;
  for (int q = 0; q < TSLB_NFIELDS; q++) f.f[q] = field_ptr<S>(base, fs, q);
  return f;
}
template <typename S>
CFieldsT<S> cfields_of(const void *base, size_t fs) {
  CFieldsT<S> f;
  for (int q = 0; q < TSLB_NFIELDS; q++) f.f[q] = (const S *)base + (size_t)q * fs;
  return f;
}

int dev_alloc(tslb_ctx *c, void **p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(TSLB_E_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed: " +
                                  hipGetErrorString(e));
  }
  c->bytes += bytes;
  return TSLB_OK;
}

hipEvent_t take_event(tslb_ctx *c) {
  hipEvent_t e = nullptr;
  if (!c->ev_pool.empty()) {
    e = c->ev_pool.back();
    c->ev_pool.pop_back();
  } else {
    (void)hipEventCreate(&e);
  }
  return e;
}

int drain_timing(tslb_ctx *c) {
  for (int k = 0; k < 2; k++) {
    for (auto &pr : c->pending[k]) {
      HIPCK(hipEventSynchronize(pr.second));
      float ms = 0.f;
      HIPCK(hipEventElapsedTime(&ms, pr.first, pr.second));
      c->ms[k] += ms;
      c->launches[k] += 1;
      c->ev_pool.push_back(pr.first);
      c->ev_pool.push_back(pr.second);
    }
    c->pending[k].clear();
  }
  return TSLB_OK;
}

// zero a field range of the state buffer (any storage type: all-zero bits)
void zero_fields(tslb_ctx *c, void *base, int q0, int q1) {
  const size_t words = (size_t)(q1 - q0) * c->fs * c->es / 4;
  float *start = (float *)((char *)base + (size_t)q0 * c->fs * c->es);
  hipLaunchKernelGGL(k_fill, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, c->stream, start, 0.f,
                     (int)words);
}

int launch_closure(tslb_ctx *c);  // closure_api.h

int launch_normals(tslb_ctx *c, int dense = 0) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (c->timing) {
    e0 = take_event(c);
    e1 = take_event(c);
    HIPCK(hipEventRecord(e0, c->stream));
  }
  const int nblk = c->kp.tiles_x * c->kp.tiles_y * c->kp.nzc;
  with_storage(c, [&](auto tag) {
    using S = decltype(tag);
    if (c->prm.ncomp == TSLB_COLOUR)
      hipLaunchKernelGGL((k_normals_tiled<S, true>), dim3(nblk), dim3(NTHR), 0, c->stream,
                         field_ptr<S>(c->buf[c->cur], c->fs, TSLB_PHI), c->nrm, c->nrm + c->fs,
                         c->nrm + 2 * c->fs, c->nrm + 3 * c->fs, c->flags, c->links, c->has_mask, 1,
                         c->kp);
    else if (dense)
      hipLaunchKernelGGL((k_normals_tiled<S, false>), dim3(nblk), dim3(NTHR), 0, c->stream,
                         field_ptr<S>(c->buf[c->cur], c->fs, TSLB_PHI), c->nrm, c->nrm + c->fs,
                         c->nrm + 2 * c->fs, c->nrm + 3 * c->fs, c->flags, c->links, c->has_mask,
                         1, c->kp);
    else {
      // sparse curvature: list the tile-planes that may carry it, then
      // compute those only
      const S *phi = field_ptr<S>(c->buf[c->cur], c->fs, TSLB_PHI);
      if (!c->list_ok) {  // else the last fused pass listed the runs and cleared the flags
        hipLaunchKernelGGL(k_fill, dim3(1), dim3(256), 0, c->stream, (float *)c->ilist, 0.f, 1);
        hipLaunchKernelGGL((k_iface_list<S>), dim3(nblk), dim3(NTHR), 0, c->stream, phi, c->flags, c->ilist,
                           c->ilist + 1, c->kp);
      }
      hipLaunchKernelGGL((k_curv_list<S>), dim3(c->curv_grid), dim3(HNT), 0, c->stream, phi, c->nrm + 3 * c->fs,
                         c->flags, c->ilist, c->ilist + 1, c->links, c->has_mask, c->kp);
    }
    return 0;
  });
  HIPCK(hipGetLastError());
  if (c->timing) {
    HIPCK(hipEventRecord(e1, c->stream));
    c->pending[1].push_back({e0, e1});
  }
  c->normals_fresh = dense ? 2 : 1;
  return TSLB_OK;
}

int launch_fused(tslb_ctx *c) {
  const Params &p = c->kp;
  const int nblk = p.tiles_x * p.tiles_y * p.nzc;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (c->timing) {
    e0 = take_event(c);
    e1 = take_event(c);
    HIPCK(hipEventRecord(e0, c->stream));
  }
  const bool add = p.ncomp != 2 && !p.shift && (p.fx != 0.f || p.fy != 0.f || p.fz != 0.f);
  const bool w4 = (c->variant & 2) != 0;
  // phase field: the fused pass lists the next step's interface runs unless
  // a boundary pass still changes phi after it
  const bool build = p.ncomp == 2 && !c->jet && !c->zbc;
  int *ilist = build ? c->ilist : nullptr;
  if (build) hipLaunchKernelGGL(k_fill, dim3(1), dim3(256), 0, c->stream, (float *)c->ilist, 0.f, 1);
  with_storage(c, [&](auto tag) {
    using S = decltype(tag);
    const S *s = (const S *)c->buf[c->cur];
    S *d = (S *)c->buf[1 - c->cur];
#define TSLB_LAUNCH(NC, MASK, ADD, WPE)                                                          \
  do {                                                                                             \
    if (c->clo_mode) {                                                                             \
      if constexpr (NC != TSLB_COLOUR)                                                             \
        hipLaunchKernelGGL((k_fused<NC, MASK, ADD, 4, S, true>), dim3(nblk), dim3(NTHR), 0,        \
                           c->stream, s, d, c->nrm, c->flags, c->links, c->clo, ilist, c->flags_alt, p); \
    } else {                                                                                       \
      bool done = false;                                                                           \
      if constexpr (NC == 2 && !MASK && WPE == 4 && sizeof(S) == 4) {                               \
        if (c->variant & 4) {                                                                      \
          hipLaunchKernelGGL((k_fused<NC, MASK, ADD, WPE, S, false, false>), dim3(nblk), dim3(NTHR), \
                             0, c->stream, s, d, c->nrm, c->flags, c->links, (const float *)nullptr, \
                             ilist, c->flags_alt, p);                                              \
          done = true;                                                                             \
        }                                                                                          \
      }                                                                                            \
      if (!done)                                                                                   \
        hipLaunchKernelGGL((k_fused<NC, MASK, ADD, WPE, S>), dim3(nblk), dim3(NTHR), 0, c->stream,  \
                           s, d, c->nrm, c->flags, c->links, (const float *)nullptr, ilist,        \
                           c->flags_alt, p);                                                       \
    }                                                                                              \
  } while (0)
    // FP32 storage keeps both register budgets selectable for sweeps (capped
    // at 128 VGPRs for two blocks per CU, or uncapped); reduced storage is capped
#define TSLB_LAUNCH_V(NC, MASK, ADD)                                      \
  do {                                                                    \
    if constexpr (sizeof(S) == 4) {                                       \
      if (w4) TSLB_LAUNCH(NC, MASK, ADD, 4);                              \
      else TSLB_LAUNCH(NC, MASK, ADD, 1);                                 \
    } else {                                                              \
      TSLB_LAUNCH(NC, MASK, ADD, 4);                                      \
    }                                                                     \
  } while (0)
    if (p.ncomp == 2) {
      if (c->has_mask) TSLB_LAUNCH_V(2, true, false);
      else TSLB_LAUNCH_V(2, false, false);
    } else if (p.ncomp == TSLB_COLOUR) {
      if (c->has_mask) {
        if (add) TSLB_LAUNCH_V(3, true, true);
        else TSLB_LAUNCH_V(3, true, false);
      } else {
        if (add) TSLB_LAUNCH_V(3, false, true);
        else TSLB_LAUNCH_V(3, false, false);
      }
    } else if (add) {
      if (c->has_mask) TSLB_LAUNCH_V(1, true, true);
      else TSLB_LAUNCH_V(1, false, true);
    } else {
      if (c->has_mask) TSLB_LAUNCH_V(1, true, false);
      else TSLB_LAUNCH_V(1, false, false);
    }
#undef TSLB_LAUNCH_V
#undef TSLB_LAUNCH
    return 0;
  });
  HIPCK(hipGetLastError());
  if (c->timing) {
    HIPCK(hipEventRecord(e1, c->stream));
    c->pending[0].push_back({e0, e1});
    if (c->pending[0].size() >= 512) {
      int rc = drain_timing(c);
      if (rc) return rc;
    }
  }
  if (c->jet) {
    const int plane = p.nx * p.ny;
    const bool add1 = p.ncomp == 1 && !p.shift && (p.fx != 0.f || p.fy != 0.f || p.fz != 0.f);
    with_storage(c, [&](auto tag) {
      using S = decltype(tag);
      const S *s = (const S *)c->buf[c->cur];
      S *d = (S *)c->buf[1 - c->cur];
      const dim3 g((unsigned)((2 * plane + 255) / 256));
      if (p.ncomp == 2)
        hipLaunchKernelGGL((k_jet_bc<2, false, S>), g, dim3(256), 0, c->stream, s, d, c->nrm, c->flags, c->links, p);
      else if (add1)
        hipLaunchKernelGGL((k_jet_bc<1, true, S>), g, dim3(256), 0, c->stream, s, d, c->nrm, c->flags, c->links, p);
      else
        hipLaunchKernelGGL((k_jet_bc<1, false, S>), g, dim3(256), 0, c->stream, s, d, c->nrm, c->flags, c->links, p);
      return 0;
    });
    HIPCK(hipGetLastError());
  }
  // one component: phi stays in buf[phi_buf] (check_ctx moves it when needed)
  c->cur = 1 - c->cur;
  c->kp.step += 1;
  if (build) {
    // the next curvature pass writes the flags the fused pass just cleared
    std::swap(c->flags, c->flags_alt);
    c->list_ok = 1;
  } else {
    c->list_ok = 0;
  }
  c->normals_fresh = 0;
  return TSLB_OK;
}

int launch_zbc(tslb_ctx *c) {
  if (!c->zbc) return TSLB_OK;
  const int plane = c->prm.nx * c->prm.ny;
  with_storage(c, [&](auto tag) {
    using S = decltype(tag);
    hipLaunchKernelGGL(k_zplane_bc<S>, dim3((plane + 255) / 256), dim3(256), 0, c->stream,
                       fields_of<S>(c->buf[c->cur], c->fs), c->w_in, c->zbc_phi ? c->phi_in : nullptr,
                       c->prm.ncomp, c->kp);
    return 0;
  });
  HIPCK(hipGetLastError());
  return TSLB_OK;
}

// One-component steps leave phi where it was (the fused pass neither reads
// nor writes it); every other entry point first brings it into the current
// buffer, since they read it there or use the other buffer as scratch.
int check_ctx_step(tslb_ctx *c) {
  if (!c) return fail(TSLB_E_ARG, "null context");
  HIPCK(hipSetDevice(c->prm.device));
  return TSLB_OK;
}
int check_ctx(tslb_ctx *c) {
  int rc = check_ctx_step(c);
  if (rc) return rc;
  c->list_ok = 0;  // the state may change: the next curvature pass lists afresh
  if (c->prm.ncomp == 1 && c->phi_buf != c->cur) {
    HIPCK(hipMemcpyAsync((char *)c->buf[c->cur] + TSLB_PHI * c->fs * c->es,
                         (char *)c->buf[c->phi_buf] + TSLB_PHI * c->fs * c->es, c->n * c->es,
                         hipMemcpyDeviceToDevice, c->stream));
    c->phi_buf = c->cur;
  }
  return TSLB_OK;
}

}  // namespace

extern "C" {

void tslb_default_params(tslb_params *p) {
  if (!p) return;
  memset(p, 0, sizeof(*p));
  p->nx = p->ny = p->nz = 16;
  p->ncomp = 1;
  p->periodic[0] = p->periodic[1] = p->periodic[2] = 1;
  p->force_mode = TSLB_FORCE_ADDITIVE;
  p->storage = TSLB_STORAGE_FP32;
  p->device = 0;
  p->tau = 0.8f;
  p->beta = 0.7f;
}

int tslb_create(tslb_ctx **out, const tslb_params *p) {
  if (!out || !p) return fail(TSLB_E_ARG, "null argument");
  *out = nullptr;
  if (p->nx < 1 || p->ny < 1 || p->nz < 1)
    return fail(TSLB_E_ARG, "lattice extents must be >= 1");
  const long long n = (long long)p->nx * p->ny * p->nz;
  if (n > 0x7fffffffLL) return fail(TSLB_E_ARG, "nx*ny*nz must fit in int32");
  // kernels address one z-plane of a field through a buffer descriptor with
  // 32-bit byte offsets
  if ((long long)p->nx * p->ny * 4 > 0x7fffffffLL) return fail(TSLB_E_ARG, "nx*ny*4 bytes must fit in int32");
  if (p->ncomp == 2 && (long long)((p->nx + TX - 1) / TX) * ((p->ny + TY - 1) / TY) * p->nz * 64 > 0x7fffffffLL)
    return fail(TSLB_E_ARG, "too many tile-planes for the curvature run list");
  if (p->ncomp != 1 && p->ncomp != 2 && p->ncomp != TSLB_COLOUR)
    return fail(TSLB_E_ARG, "ncomp must be 1, 2 or TSLB_COLOUR");
  if (p->ncomp == TSLB_COLOUR && p->tau_b > 0.f && !(p->tau_b > 0.5f))
    return fail(TSLB_E_ARG, "tau_b must be > 0.5 (or <= 0 for tau)");
  if (!(p->tau > 0.5f)) return fail(TSLB_E_ARG, "tau must be > 0.5");
  if (p->storage != TSLB_STORAGE_FP32 && p->storage != TSLB_STORAGE_FP16 &&
      p->storage != TSLB_STORAGE_BF16 && p->storage != TSLB_STORAGE_FX16)
    return fail(TSLB_E_ARG, "storage must be TSLB_STORAGE_FP32, _FP16, _BF16 or _FX16");
  if (p->force_mode != TSLB_FORCE_ADDITIVE && p->force_mode != TSLB_FORCE_SHIFT)
    return fail(TSLB_E_ARG, "force_mode must be TSLB_FORCE_ADDITIVE or TSLB_FORCE_SHIFT");
  int ndev = 0;
  HIPCK(hipGetDeviceCount(&ndev));
  if (p->device < 0 || p->device >= ndev) return fail(TSLB_E_ARG, "device ordinal out of range");
  HIPCK(hipSetDevice(p->device));
  {
    const int rc0 = scratch_setup(p->device);
    if (rc0) return rc0;
  }

  tslb_ctx *c = new tslb_ctx();
  c->prm = *p;
  c->n = (size_t)n;
  c->fs = (c->n + 63) & ~(size_t)63;
  Params &k = c->kp;
  k.fs = (long long)c->fs;
  k.nx = p->nx;
  k.ny = p->ny;
  k.nz = p->nz;
  k.px = p->periodic[0] ? 1 : 0;
  k.py = p->periodic[1] ? 1 : 0;
  k.pz = p->periodic[2] ? 1 : 0;
  k.ncomp = p->ncomp;
  // the colour model takes its body force additively (recursive_colorTSLB3D.f90's fx*p1dcssq)
  k.shift = (p->ncomp == 2 || (p->ncomp == 1 && p->force_mode == TSLB_FORCE_SHIFT)) ? 1 : 0;
  k.inua = 3.f / (p->tau - 0.5f);
  k.inub = 3.f / ((p->tau_b > 0.5f ? p->tau_b : p->tau) - 0.5f);
  k.beta = p->beta;
  k.om1 = 1.f - 1.f / p->tau;
  k.tau = p->tau;
  k.sigma = p->sigma;
  k.sharp_c = p->sharp_c;
  k.fx = p->force[0];
  k.fy = p->force[1];
  k.fz = p->force[2];
  k.lin_a = p->lin_force;
  k.tiles_x = (p->nx + TX - 1) / TX;
  k.tiles_y = (p->ny + TY - 1) / TY;
  {
    // z-chunks of 64 planes (measured best at 512^3 among 32, 64 and the
    // whole depth), shorter when that would leave fewer than two blocks per CU
    int ncu = 256;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, p->device) == hipSuccess && prop.multiProcessorCount > 0)
      ncu = prop.multiProcessorCount;
    c->ncu = ncu;
    // the curvature blocks stride over the run list with equal shares, so
    // the grid is the number that is resident at once: one block more per
    // CU would start only when another ends and double the launch's length
    int bpc = 0;
    with_storage(c, [&](auto tag) {
      using S = decltype(tag);
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, k_curv_list<S>, HNT, 0) != hipSuccess) bpc = 0;
      return 0;
    });
    if (const char *v = getenv("TSLB_CURV_BPC")) bpc = atoi(v);
    c->curv_grid = (bpc > 0 ? bpc : 6) * ncu;
    const long long tiles = (long long)k.tiles_x * k.tiles_y;
    int zc = 64;
    while (zc > 8 && tiles * ((p->nz + zc - 1) / zc) < 2LL * ncu) zc /= 2;
    if (const char *v = getenv("TSLB_ZCHUNK")) zc = atoi(v);
    if (zc < 1) zc = 1;
    if (zc > 64) zc = 64;  // k_iface_list keeps one bit per plane of a chunk
    if (zc > p->nz) zc = p->nz;
    k.z_chunk = zc;
  }
  k.nzc = (p->nz + k.z_chunk - 1) / k.z_chunk;
  // storage scales per field group: FP16 keeps rho-1 (x64) and Pi^neq
  // (x4096) in its normal range; BF16 has FP32's exponent range and FP32
  // needs none; FX16 steps are powers of two per group (defaults in
  // fx16_shift, changed by tslb_set_storage_range)
  c->es = p->storage == TSLB_STORAGE_FP32 ? 4 : 2;
  for (int g = 0; g < 4; g++) {
    float sc = 1.f;
    if (p->storage == TSLB_STORAGE_FP16) {
      if (g == sgrp(TSLB_RHO)) sc = 64.f;
      if (g == sgrp(TSLB_PXX)) sc = 4096.f;
    }
    if (p->storage == TSLB_STORAGE_FX16) sc = ldexpf(1.f, fx16_shift(g));
    set_scale(k, g, sc);
  }

  if (const char *v = getenv("TSLB_FUSED_VARIANT")) c->variant = atoi(v) & 7;
  int rc = TSLB_OK;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return fail(TSLB_E_HIP, std::string("hipStreamCreate failed: ") + hipGetErrorString(e));
  }
  const size_t fb = c->fs * TSLB_NFIELDS * c->es;
  if ((rc = dev_alloc(c, (void **)&c->buf[0], fb)) || (rc = dev_alloc(c, (void **)&c->buf[1], fb)) ||
      (p->ncomp >= 2 && (rc = dev_alloc(c, (void **)&c->nrm, 4 * c->fs * sizeof(float)))) ||
      (p->ncomp >= 2 &&
       (rc = dev_alloc(c, (void **)&c->flags, 2 * (size_t)k.tiles_x * k.tiles_y * p->nz))) ||
      (p->ncomp == 2 &&
       (rc = dev_alloc(c, (void **)&c->ilist, (2 * (size_t)k.tiles_x * k.tiles_y * p->nz + 1) * sizeof(int)))) ||
      (p->ncomp == 2 &&
       (rc = dev_alloc(c, (void **)&c->flags_alt, 2 * (size_t)k.tiles_x * k.tiles_y * p->nz))) ||
      (p->storage == TSLB_STORAGE_FX16 && (rc = dev_alloc(c, (void **)&c->sat, 256))) ||
      (rc = dev_alloc(c, (void **)&c->part, (size_t)kRedBlocks * NRED * sizeof(double))) ||
      (rc = dev_alloc(c, (void **)&c->red, NRED * sizeof(double)))) {
    std::string msg = g_err;
    tslb_destroy(c);
    return fail(rc, msg);
  }
  // zero state: rho-1 = 0, all moments 0
  for (int b = 0; b < 2; b++) zero_fields(c, c->buf[b], 0, TSLB_NFIELDS);
  if (c->sat) hipLaunchKernelGGL(k_fill, dim3(1), dim3(256), 0, c->stream, (float *)c->sat, 0.f, 1);
  k.sat = c->sat;
  if (c->nrm) {
    const int tot = (int)(4 * c->fs);
    hipLaunchKernelGGL(k_fill, dim3((tot + 255) / 256), dim3(256), 0, c->stream, c->nrm, 0.f, tot);
  }
  e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    tslb_destroy(c);
    return fail(TSLB_E_HIP, std::string("initial fill failed: ") + hipGetErrorString(e));
  }
  *out = c;
  return TSLB_OK;
}

int tslb_destroy(tslb_ctx *c) {
  if (!c) return TSLB_OK;
  (void)hipSetDevice(c->prm.device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (int k = 0; k < 2; k++) {
    for (auto &pr : c->pending[k]) {
      (void)hipEventDestroy(pr.first);
      (void)hipEventDestroy(pr.second);
    }
  }
  for (auto e : c->ev_pool) (void)hipEventDestroy(e);
  (void)hipFree(c->buf[0]);
  (void)hipFree(c->buf[1]);
  (void)hipFree(c->nrm);
  (void)hipFree(c->flags);
  (void)hipFree(c->flags_alt);
  (void)hipFree(c->ilist);
  (void)hipFree(c->links);
  (void)hipFree(c->part);
  (void)hipFree(c->red);
  (void)hipFree(c->w_in);
  (void)hipFree(c->sat);
  (void)hipFree(c->phi_in);
  (void)hipFree(c->clo);
  (void)hipFree(c->cfeat);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return TSLB_OK;
}

int tslb_set_isfluid(tslb_ctx *c, const int32_t *isf) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!isf) {
    c->has_mask = 0;
    c->zwalls = 0;
    c->jet = 0;
    c->kp.gwall = 0;
    return TSLB_OK;
  }
  {
    const size_t plane = (size_t)c->prm.nx * c->prm.ny;
    int w0 = 1, w1 = 1;
    for (size_t q = 0; q < plane; q++) {
      if (isf[q]) w0 = 0;
      if (isf[plane * (c->prm.nz - 1) + q]) w1 = 0;
    }
    c->zwalls = w0 | (w1 << 1);
    if (c->jet && c->zwalls != 3) {
      c->jet = 0;
      c->kp.gwall = 0;
    }
  }
  if (!c->links) {
    if ((rc = dev_alloc(c, (void **)&c->links, c->n * sizeof(uint32_t)))) return rc;
  }
  int32_t *tmp = nullptr;
  HIPCK(hipMalloc((void **)&tmp, c->n * sizeof(int32_t)));
  hipError_t e = hipMemcpyAsync(tmp, isf, c->n * sizeof(int32_t), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_links, dim3((unsigned)((c->n + 255) / 256)), dim3(256), 0, c->stream, tmp,
                       c->links, c->kp);
    e = hipStreamSynchronize(c->stream);
  }
  (void)hipFree(tmp);
  if (e != hipSuccess) return fail(TSLB_E_HIP, std::string("set_isfluid: ") + hipGetErrorString(e));
  c->has_mask = 1;
  return TSLB_OK;
}

int tslb_upload_field(tslb_ctx *c, int field, const float *data) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (field < 0 || field >= TSLB_NFIELDS || !data) return fail(TSLB_E_ARG, "bad field or pointer");
  // stage the host floats in the second buffer (scratch between steps), then
  // convert into the state buffer (rho -> rho-1, storage type and scale)
  float *tmp = (float *)c->buf[1 - c->cur];
  HIPCK(hipMemcpyAsync(tmp, data, c->n * sizeof(float), hipMemcpyHostToDevice, c->stream));
  with_storage(c, [&](auto tag) {
    using S = decltype(tag);
    hipLaunchKernelGGL(k_to_storage<S>, dim3((unsigned)((c->n + 255) / 256)), dim3(256), 0, c->stream, tmp,
                       field_ptr<S>(c->buf[c->cur], c->fs, field), (int)c->n, field, c->kp);
    return 0;
  });
  HIPCK(hipGetLastError());
  HIPCK(hipStreamSynchronize(c->stream));
  c->normals_fresh = 0;
  return TSLB_OK;
}

int tslb_upload_moments(tslb_ctx *c, const float *m) {
  if (!m) return fail(TSLB_E_ARG, "null pointer");
  for (int q = 0; q < TSLB_NFIELDS; q++) {
    int rc = tslb_upload_field(c, q, m + (size_t)q * (c ? c->n : 0));
    if (rc) return rc;
  }
  return TSLB_OK;
}

int tslb_download_field(tslb_ctx *c, int field, float *data) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (field < 0 || field >= TSLB_NFIELDS || !data) return fail(TSLB_E_ARG, "bad field or pointer");
  float *tmp = (float *)c->buf[1 - c->cur];
  with_storage(c, [&](auto tag) {
    using S = decltype(tag);
    hipLaunchKernelGGL(k_from_storage<S>, dim3((unsigned)((c->n + 255) / 256)), dim3(256), 0, c->stream,
                       (const S *)field_ptr<S>(c->buf[c->cur], c->fs, field), tmp, (int)c->n, field, c->kp);
    return 0;
  });
  HIPCK(hipGetLastError());
  HIPCK(hipMemcpyAsync(data, tmp, c->n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  HIPCK(hipStreamSynchronize(c->stream));
  return TSLB_OK;
}

int tslb_download_moments(tslb_ctx *c, float *m) {
  if (!m) return fail(TSLB_E_ARG, "null pointer");
  for (int q = 0; q < TSLB_NFIELDS; q++) {
    int rc = tslb_download_field(c, q, m + (size_t)q * (c ? c->n : 0));
    if (rc) return rc;
  }
  return TSLB_OK;
}

// Moment-space checkpoint: header + the stored fields verbatim.
struct CkptHeader {
  char magic[8];  // "TSLBCKPT"
  int32_t version, nx, ny, nz, ncomp, storage, elem_bytes, nfields;
  int64_t step;
  int32_t pad[4];
};
static_assert(sizeof(CkptHeader) == 64, "checkpoint header is 64 bytes");

// Checkpoint versions: 1 for FP32 storage; 4 for the 16-bit formats, whose
// 4 group scales (FX16: power-of-two steps) and step count (the seed of the
// stochastic rounding) follow the header.
constexpr int32_t kCkpt16 = 4;
static int32_t ckpt_version(int storage) { return storage == TSLB_STORAGE_FP32 ? 1 : kCkpt16; }

int tslb_checkpoint_write(tslb_ctx *c, const char *path, int64_t step) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!path) return fail(TSLB_E_ARG, "null path");
  CkptHeader h;
  memset(&h, 0, sizeof(h));
  memcpy(h.magic, "TSLBCKPT", 8);
  h.version = ckpt_version(c->prm.storage);
  h.nx = c->prm.nx;
  h.ny = c->prm.ny;
  h.nz = c->prm.nz;
  h.ncomp = c->prm.ncomp;
  h.storage = c->prm.storage;
  h.elem_bytes = (int32_t)c->es;
  h.nfields = TSLB_NFIELDS;
  h.step = step;
  FILE *fp = fopen(path, "wb");
  if (!fp) return fail(TSLB_E_ARG, std::string("cannot open ") + path + " for writing");
  const size_t bytes = c->n * c->es;
  std::vector<char> host(bytes);
  bool ok = fwrite(&h, sizeof(h), 1, fp) == 1;
  if (ok && h.version == kCkpt16) ok = fwrite(c->kp.sc, sizeof(float), 4, fp) == 4;
  if (ok && h.version == kCkpt16) ok = fwrite(&c->kp.step, sizeof(uint32_t), 1, fp) == 1;
  for (int q = 0; ok && q < TSLB_NFIELDS; q++) {
    hipError_t e = hipMemcpyAsync(host.data(), (char *)c->buf[c->cur] + (size_t)q * c->fs * c->es, bytes,
                                  hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
      fclose(fp);
      return fail(TSLB_E_HIP, std::string("checkpoint copy: ") + hipGetErrorString(e));
    }
    ok = fwrite(host.data(), 1, bytes, fp) == bytes;
  }
  if (fclose(fp) != 0) ok = false;
  if (!ok) return fail(TSLB_E_ARG, std::string("short write to ") + path);
  return TSLB_OK;
}

int tslb_checkpoint_read(tslb_ctx *c, const char *path, int64_t *step) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!path) return fail(TSLB_E_ARG, "null path");
  FILE *fp = fopen(path, "rb");
  if (!fp) return fail(TSLB_E_ARG, std::string("cannot open ") + path);
  CkptHeader h;
  if (fread(&h, sizeof(h), 1, fp) != 1 || memcmp(h.magic, "TSLBCKPT", 8) != 0) {
    fclose(fp);
    return fail(TSLB_E_ARG, std::string(path) + " is not a TSLB checkpoint");
  }
  // version 1 files of FP16 and BF16 storage (written before the step count
  // was stored) hold the same fixed scales: they are read, and the run goes
  // on with the context's own step count
  const bool legacy16 =
      h.version == 1 && (h.storage == TSLB_STORAGE_FP16 || h.storage == TSLB_STORAGE_BF16);
  if (h.version != ckpt_version(h.storage) && !legacy16) {
    fclose(fp);
    if (h.storage == TSLB_STORAGE_FX16 && h.version == 3)
      return fail(TSLB_E_ARG, std::string(path) + " holds FX16 codes of the earlier range-scaled format (version 3), "
                                                  "which power-of-two steps cannot decode exactly");
    return fail(TSLB_E_ARG, std::string(path) + " is not a TSLB checkpoint");
  }
  if (h.nx != c->prm.nx || h.ny != c->prm.ny || h.nz != c->prm.nz || h.ncomp != c->prm.ncomp ||
      h.storage != c->prm.storage || h.elem_bytes != (int32_t)c->es || h.nfields != TSLB_NFIELDS) {
    fclose(fp);
    return fail(TSLB_E_ARG, "checkpoint lattice, components or storage differ from the context");
  }
  float sc[4];
  uint32_t kstep = 0;
  if (h.version == kCkpt16 && (fread(sc, sizeof(float), 4, fp) != 4 || fread(&kstep, sizeof(uint32_t), 1, fp) != 1)) {
    fclose(fp);
    return fail(TSLB_E_ARG, std::string("truncated checkpoint ") + path);
  }
  if (h.version == kCkpt16) {
    for (int g = 0; g < 4; g++) {
      int e;
      if (!(sc[g] > 0.f) || !(sc[g] < 1e30f) || frexpf(sc[g], &e) != 0.5f) {
        fclose(fp);
        return fail(TSLB_E_ARG, std::string(path) + " holds an invalid storage scale");
      }
    }
    for (int g = 0; g < 4; g++) set_scale(c->kp, g, sc[g]);
    c->kp.step = kstep;
  }
  const size_t bytes = c->n * c->es;
  std::vector<char> host(bytes);
  for (int q = 0; q < TSLB_NFIELDS; q++) {
    if (fread(host.data(), 1, bytes, fp) != bytes) {
      fclose(fp);
      return fail(TSLB_E_ARG, std::string("truncated checkpoint ") + path);
    }
    hipError_t e = hipMemcpyAsync((char *)c->buf[c->cur] + (size_t)q * c->fs * c->es, host.data(), bytes,
                                  hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
      fclose(fp);
      return fail(TSLB_E_HIP, std::string("checkpoint copy: ") + hipGetErrorString(e));
    }
  }
  fclose(fp);
  if (step) *step = h.step;
  c->normals_fresh = 0;
  return TSLB_OK;
}

int tslb_download_normals(tslb_ctx *c, float *n3) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!n3) return fail(TSLB_E_ARG, "null pointer");
  if (c->prm.ncomp < 2)
    return fail(TSLB_E_STATE, "normals (colour model: colour gradients) exist only for two-component runs");
  // the stepping pass writes n only where it is nonzero; recompute densely
  if (c->normals_fresh != 2 && (rc = launch_normals(c, 1))) return rc;
  for (int a = 0; a < 3; a++)
    HIPCK(hipMemcpyAsync(n3 + (size_t)a * c->n, c->nrm + (size_t)a * c->fs, c->n * sizeof(float),
                         hipMemcpyDeviceToHost, c->stream));
  HIPCK(hipStreamSynchronize(c->stream));
  return TSLB_OK;
}

int tslb_init_equilibrium(tslb_ctx *c, int prepush) {
  int rc = check_ctx(c);
  if (rc) return rc;
  void *base = c->buf[c->cur];
  const int nb = (int)((c->n + 255) / 256);
  zero_fields(c, base, TSLB_PXX, TSLB_PYZ + 1);
  if (c->prm.ncomp == 2 && prepush) {
    with_storage(c, [&](auto tag) {
      using S = decltype(tag);
      S *tmp = field_ptr<S>(c->buf[1 - c->cur], c->fs, TSLB_PHI);
      hipLaunchKernelGGL(k_prepush<S>, dim3(nb), dim3(256), 0, c->stream, field_ptr<S>(base, c->fs, TSLB_PHI),
                         field_ptr<S>(base, c->fs, TSLB_U), field_ptr<S>(base, c->fs, TSLB_V),
                         field_ptr<S>(base, c->fs, TSLB_W), tmp, c->links, c->has_mask, c->kp);
      return 0;
    });
    HIPCK(hipGetLastError());
    HIPCK(hipMemcpyAsync((char *)base + TSLB_PHI * c->fs * c->es,
                         (char *)c->buf[1 - c->cur] + TSLB_PHI * c->fs * c->es, c->n * c->es,
                         hipMemcpyDeviceToDevice, c->stream));
  }
  HIPCK(hipStreamSynchronize(c->stream));
  c->normals_fresh = 0;
  return TSLB_OK;
}

int tslb_normals(tslb_ctx *c) {
  int rc = check_ctx_step(c);
  if (rc) return rc;
  if (c->prm.ncomp < 2) return TSLB_OK;
  if (c->prm.ncomp == TSLB_COLOUR) {  // the fused pass forms the colour gradient itself
    c->normals_fresh = 1;
    return TSLB_OK;
  }
  return launch_normals(c);
}

int tslb_collide_stream(tslb_ctx *c) {
  int rc = check_ctx_step(c);
  if (rc) return rc;
  if (c->prm.ncomp >= 2 && !c->normals_fresh) return fail(TSLB_E_STATE, "tslb_normals() must precede tslb_collide_stream()");
  if (c->clo_mode == 2 && (rc = launch_closure(c))) return rc;
  // the same fused pass and boundary planes as one tslb_step
  if ((rc = launch_fused(c))) return rc;
  return launch_zbc(c);
}

int tslb_step(tslb_ctx *c, int nsteps) {
  int rc = check_ctx_step(c);
  if (rc) return rc;
  if (nsteps < 0) return fail(TSLB_E_ARG, "nsteps < 0");
  for (int s = 0; s < nsteps; s++) {
    if (c->clo_mode == 2 && (rc = launch_closure(c))) return rc;
    if (c->prm.ncomp == 2 && (rc = launch_normals(c))) return rc;
    if ((rc = launch_fused(c))) return rc;
    if ((rc = launch_zbc(c))) return rc;
  }
  return TSLB_OK;
}

int tslb_set_jet_bc(tslb_ctx *c, int enable) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!enable) {
    c->jet = 0;
    c->kp.gwall = 0;
    return TSLB_OK;
  }
  if (c->prm.nz < 3 || c->prm.periodic[2])
    return fail(TSLB_E_ARG, "jet boundary needs nz >= 3 and a non-periodic z axis");
  if (!c->has_mask || c->zwalls != 3)
    return fail(TSLB_E_STATE, "jet boundary needs every node of planes k=1 and k=nz solid (tslb_set_isfluid)");
  if (c->zbc) return fail(TSLB_E_STATE, "jet boundary and tslb_set_zplane_bc are exclusive");
  if (c->clo_mode) return fail(TSLB_E_STATE, "jet boundary and the learned closure are exclusive");
  if (c->prm.ncomp == TSLB_COLOUR) return fail(TSLB_E_STATE, "the jet boundary is defined for the phase-field model");
  c->jet = 1;
  c->kp.gwall = 1;
  return TSLB_OK;
}

int tslb_upload_plane(tslb_ctx *c, int field, int k, const float *data) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (field < 0 || field >= TSLB_NFIELDS || !data) return fail(TSLB_E_ARG, "bad field or pointer");
  if (k < 0 || k >= c->prm.nz) return fail(TSLB_E_ARG, "plane index out of range");
  const size_t plane = (size_t)c->prm.nx * c->prm.ny;
  float *tmp = (float *)c->buf[1 - c->cur];
  HIPCK(hipMemcpyAsync(tmp, data, plane * sizeof(float), hipMemcpyHostToDevice, c->stream));
  with_storage(c, [&](auto tag) {
    using S = decltype(tag);
    hipLaunchKernelGGL(k_to_storage<S>, dim3((unsigned)((plane + 255) / 256)), dim3(256), 0, c->stream, tmp,
                       field_ptr<S>(c->buf[c->cur], c->fs, field) + plane * (size_t)k, (int)plane, field,
                       c->kp);
    return 0;
  });
  HIPCK(hipGetLastError());
  HIPCK(hipStreamSynchronize(c->stream));
  c->normals_fresh = 0;
  return TSLB_OK;
}

int tslb_set_zplane_bc(tslb_ctx *c, const float *w_in, const float *phi_in) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!w_in) {
    c->zbc = 0;
    return TSLB_OK;
  }
  if (c->prm.nz < 3 || c->prm.periodic[2]) return fail(TSLB_E_ARG, "z-plane BC needs nz >= 3 and a non-periodic z axis");
  if (c->jet) return fail(TSLB_E_STATE, "tslb_set_zplane_bc and the jet boundary are exclusive");
  const size_t plane = (size_t)c->prm.nx * c->prm.ny;
  if (!c->w_in && (rc = dev_alloc(c, (void **)&c->w_in, plane * sizeof(float)))) return rc;
  if (!c->phi_in && (rc = dev_alloc(c, (void **)&c->phi_in, plane * sizeof(float)))) return rc;
  HIPCK(hipMemcpyAsync(c->w_in, w_in, plane * sizeof(float), hipMemcpyHostToDevice, c->stream));
  c->zbc_phi = 0;
  if (phi_in && c->prm.ncomp == 2) {
    HIPCK(hipMemcpyAsync(c->phi_in, phi_in, plane * sizeof(float), hipMemcpyHostToDevice, c->stream));
    c->zbc_phi = 1;
  }
  c->zbc = 1;
  if ((rc = launch_zbc(c))) return rc;
  HIPCK(hipStreamSynchronize(c->stream));
  c->normals_fresh = 0;
  return TSLB_OK;
}

int tslb_set_storage_range(tslb_ctx *c, int field, float half_range) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (c->prm.storage != TSLB_STORAGE_FX16) return fail(TSLB_E_STATE, "storage ranges apply to TSLB_STORAGE_FX16");
  if (field < 0 || field >= TSLB_NFIELDS) return fail(TSLB_E_ARG, "bad field");
  if (!(half_range > 0.f) || !(half_range < 1e30f)) return fail(TSLB_E_ARG, "half_range must be positive and finite");
  // the range belongs to the field's group; re-encode each stored field of
  // the group: decode with the old scale into the second buffer (scratch
  // between steps), encode with the new one
  const Params old = c->kp;
  set_scale(c->kp, sgrp(field), ldexpf(1.f, fx16_shift_for(half_range)));
  float *tmp = (float *)c->buf[1 - c->cur];
  const dim3 g((unsigned)((c->n + 255) / 256));
  for (int q = 0; q < TSLB_NFIELDS; q++) {
    if (sgrp(q) != sgrp(field)) continue;
    fx16_t *f = field_ptr<fx16_t>(c->buf[c->cur], c->fs, q);
    hipLaunchKernelGGL(k_from_storage<fx16_t>, g, dim3(256), 0, c->stream, (const fx16_t *)f, tmp, (int)c->n, q,
                       old);
    HIPCK(hipGetLastError());
    hipLaunchKernelGGL(k_to_storage<fx16_t>, g, dim3(256), 0, c->stream, (const float *)tmp, f, (int)c->n, q, c->kp);
    HIPCK(hipGetLastError());
  }
  HIPCK(hipStreamSynchronize(c->stream));
  c->normals_fresh = 0;
  return TSLB_OK;
}

int tslb_storage_clipped(tslb_ctx *c, int64_t *count) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!count) return fail(TSLB_E_ARG, "null pointer");
  *count = 0;
  if (!c->sat) return TSLB_OK;
  unsigned int v = 0;
  HIPCK(hipMemcpyAsync(&v, c->sat, sizeof(v), hipMemcpyDeviceToHost, c->stream));
  HIPCK(hipStreamSynchronize(c->stream));
  *count = (int64_t)v;
  return TSLB_OK;
}

int tslb_seed_noise(tslb_ctx *c, int field, float amp, int l) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (field < 0 || field >= TSLB_NFIELDS) return fail(TSLB_E_ARG, "bad field");
  with_storage(c, [&](auto tag) {
    using S = decltype(tag);
    hipLaunchKernelGGL(k_seed_noise<S>, dim3((unsigned)((c->n + 255) / 256)), dim3(256), 0, c->stream,
                       field_ptr<S>(c->buf[c->cur], c->fs, field), field, amp, l, c->kp);
    return 0;
  });
  HIPCK(hipGetLastError());
  HIPCK(hipStreamSynchronize(c->stream));
  c->normals_fresh = 0;
  return TSLB_OK;
}

int tslb_reduce(tslb_ctx *c, double out[7]) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!out) return fail(TSLB_E_ARG, "null pointer");
  const int nblk = (int)((c->n + 255) / 256) < kRedBlocks ? (int)((c->n + 255) / 256) : kRedBlocks;
  with_storage(c, [&](auto tag) {
    using S = decltype(tag);
    hipLaunchKernelGGL(k_reduce_part<S>, dim3(nblk), dim3(256), 0, c->stream,
                       cfields_of<S>(c->buf[c->cur], c->fs), c->links, c->has_mask, c->kp, c->part);
    return 0;
  });
  HIPCK(hipGetLastError());
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, c->stream, c->part, nblk, c->red);
  HIPCK(hipGetLastError());
  HIPCK(hipMemcpyAsync(out, c->red, NRED * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCK(hipStreamSynchronize(c->stream));
  return TSLB_OK;
}

int tslb_reduce_noalloc(tslb_ctx *c, double *out_dev) {
  int rc = check_ctx_step(c);
  if (rc) return rc;
  if (!out_dev) return fail(TSLB_E_ARG, "null output pointer");
  double *part = g_scratch[c->prm.device];
  if (!part) return fail(TSLB_E_STATE, "no scratch address for this device");
  const long long want = ((long long)c->n + 255) / 256;
  const int nblk = (int)(want < kScratchBlocks ? want : kScratchBlocks);
  if (nblk < 1 || nblk > kScratchBlocks) return fail(TSLB_E_ARG, "reduction grid beyond the scratch capacity");
  with_storage(c, [&](auto tag) {
    using S = decltype(tag);
    hipLaunchKernelGGL(k_reduce_part<S>, dim3(nblk), dim3(256), 0, c->stream,
                       cfields_of<S>(c->buf[c->cur], c->fs), c->links, c->has_mask, c->kp, part);
    return 0;
  });
  HIPCK(hipGetLastError());
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, c->stream, (const double *)part, nblk, out_dev);
  HIPCK(hipGetLastError());
  return TSLB_OK;
}

int tslb_sync(tslb_ctx *c) {
  int rc = check_ctx(c);
  if (rc) return rc;
  HIPCK(hipStreamSynchronize(c->stream));
  return TSLB_OK;
}

int tslb_timing_enable(tslb_ctx *c, int on) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if ((rc = drain_timing(c))) return rc;
  c->timing = on ? 1 : 0;
  c->ms[0] = c->ms[1] = 0.0;
  c->launches[0] = c->launches[1] = 0;
  return TSLB_OK;
}

int tslb_timing_read(tslb_ctx *c, int k, double *total_ms, int64_t *launches) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (k < 0 || k > 1 || !total_ms || !launches) return fail(TSLB_E_ARG, "bad timing query");
  if ((rc = drain_timing(c))) return rc;
  *total_ms = c->ms[k];
  *launches = c->launches[k];
  return TSLB_OK;
}

int tslb_mem_info(size_t *free_b, size_t *total_b) {
  if (!free_b || !total_b) return fail(TSLB_E_ARG, "null pointer");
  HIPCK(hipMemGetInfo(free_b, total_b));
  return TSLB_OK;
}

int tslb_host_mem_mb(double *peak_rss_mb, double *total_ram_mb) {
  if (!peak_rss_mb || !total_ram_mb) return fail(TSLB_E_ARG, "null pointer");
  struct rusage ru;
  if (getrusage(RUSAGE_SELF, &ru) != 0) return fail(TSLB_E_STATE, "getrusage failed");
  *peak_rss_mb = (double)ru.ru_maxrss / 1024.0;
  struct sysinfo si;
  if (sysinfo(&si) != 0) return fail(TSLB_E_STATE, "sysinfo failed");
  *total_ram_mb = (double)si.totalram * (double)si.mem_unit / (1024.0 * 1024.0);
  return TSLB_OK;
}

float tslb_rand_noseeded(int i, int j, int k, int l) { return rand_noseeded_impl(i, j, k, l); }

float tslb_gauss_noseeded(int i, int j, int k, int l) {
  float d1 = tslb_rand_noseeded(i, j, k, l);
  const int kk = (int)lroundf(d1 * 50.0f);
  const float d2 = tslb_rand_noseeded(i, j, k, l + kk);
  d1 = d1 * (1.0f - 1e-30f) + 1e-30f;
  return sqrtf(-2.0f * logf(d1)) * cosf(2.0f * 3.14159265358979323846f * d2);
}

int tslb_stream_copy(tslb_ctx *c, double *gbs) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!gbs) return fail(TSLB_E_ARG, "null pointer");
  const long long n4 = (long long)(c->fs * TSLB_NFIELDS * c->es / 16);
  const unsigned copy_grid = (unsigned)std::min<long long>((n4 + 255) / 256, 1LL << 30);
  if (n4 <= 0) return fail(TSLB_E_ARG, "empty state");
  hipEvent_t e0 = take_event(c), e1 = take_event(c);
  float best = 1e30f;
  for (int rep = 0; rep < 4; rep++) {
    HIPCK(hipEventRecord(e0, c->stream));
    hipLaunchKernelGGL(k_copy, dim3(copy_grid), dim3(256), 0, c->stream,
                       (const v4f *)c->buf[c->cur], (v4f *)c->buf[1 - c->cur], n4);
    HIPCK(hipGetLastError());
    HIPCK(hipEventRecord(e1, c->stream));
    HIPCK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCK(hipEventElapsedTime(&ms, e0, e1));
    if (rep > 0 && ms < best) best = ms;
  }
  c->ev_pool.push_back(e0);
  c->ev_pool.push_back(e1);
  *gbs = 2.0 * 16.0 * (double)n4 / (best * 1e-3) / 1e9;
  return TSLB_OK;
}

int tslb_device_bytes(tslb_ctx *c, size_t *bytes) {
  if (!c || !bytes) return fail(TSLB_E_ARG, "null argument");
  *bytes = c->bytes;
  return TSLB_OK;
}

int tslb_last_error(char *buf, int len) {
  if (!buf || len <= 0) return TSLB_E_ARG;
  snprintf(buf, (size_t)len, "%s", g_err.c_str());
  return TSLB_OK;
}

const char *tslb_version(void) { return "tslb 0.1 gfx950"; }

}  // extern "C"

#include "closure_api.h"
