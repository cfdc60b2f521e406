// closure_kernels.h -- device code of the learned subgrid closure, included by
// tslb.hip inside its anonymous namespace.
//
// A per-node MLP maps 10 local inputs (strain rate S from the isotropic
// gradient of the velocity, grad(phi), phi) to an eddy viscosity nu_t and six
// interfacial-stress corrections dsigma; the collision then relaxes with
// omega = 1/(tau + 3 nu_t) and adds dsigma to the post-collision stress.  The
// reference has no such closure; the network shape and the input/output sets
// follow the target description of the path (10 -> 64 -> 64 -> 64 -> 7, GELU).
//
// Matrix cores.  Every layer is one chain of v_mfma_f32_16x16x32_bf16 in the
// transposed form Z^T = W^T X^T: the A operand is the weight block (16 output
// features x 32 inputs), the B operand 32 inputs x 16 nodes, the accumulator
// 16 features x 16 nodes.  Lane l = 16 g + c holds, per MFMA,
//   A[c][8g + i], B[8g + i][c] (i < 8)  and  D[4g + i][c] (i < 4).
// The D registers of a layer become the B operand of the next one without a
// shuffle: the 32-wide K step ks of the next layer takes output blocks 2ks and
// 2ks+1, so its k = 8g + i stands for feature
//   perm(ks, g, i) = 16 (2 ks + i/4) + 4 g + i%4,
// and the weights are packed on the device in that order (k_mlp_pack).  The
// bias of layer 1 rides on a constant-1 input (k = 10); the other biases
// initialise the accumulators.
//
// Training.  k_mlp_fwd_bwd runs forward and backward per 16-node tile with
// the same chained layout (the backward products delta_{l-1} = W_l delta_l use
// a second, transposed pack) and stores the BF16 activations and deltas
// feature-major; k_mlp_wgrad contracts them over the batch on the matrix
// cores (dW = delta H^T, K = nodes) and adds its partial sums to an FP32
// gradient; k_mlp_adam updates the FP32 master weights.  The loss is summed in
// FP64.
#pragma once

constexpr int kMlpNF = 10, kMlpNH = 64, kMlpNO = 7;
// offsets in the natural parameter vector (in x out row-major per layer)
constexpr int kOffW1 = 0, kOffB1 = 640, kOffW2 = 704, kOffB2 = 4800, kOffW3 = 4864, kOffB3 = 8960,
              kOffW4 = 9024, kOffB4 = 9472, kOffIn = 9479, kOffOut = 9489, kMlpNParam = 9496;
constexpr int kFwdBlocks = 22, kBwdBlocks = 20;
constexpr float kGeluC = 0.7978845608028654f, kGeluA = 0.044715f;
// per-sample auxiliary record: rho_bar, tau_xx, tau_yy, tau_zz, tau_xy, tau_xz, tau_yz
constexpr int kAux = 7;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f32x4 mfma16(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__host__ __device__ constexpr int mlp_perm(int ks, int g, int i) {
  return 16 * (2 * ks + (i >> 2)) + 4 * g + (i & 3);
}
// z sigmoid(2u) = z / (1 + 2^(-2 u log2 e)), u = c (z + a z^3): one v_exp_f32
// and one v_rcp_f32
__device__ __forceinline__ float gelu_f(float z) {
  constexpr float c0 = -2.f * kGeluC * 1.4426950408889634f, c1 = c0 * kGeluA;
  const float e = __builtin_amdgcn_exp2f(z * fmaf(c1, z * z, c0));
  return z * __builtin_amdgcn_rcpf(1.f + e);
}
__device__ __forceinline__ float gelu_grad_f(float z) {
  const float u = kGeluC * fmaf(kGeluA * z * z, z, z);
  const float s = 1.f / (1.f + __expf(-2.f * u));
  return fmaf(z * s * (1.f - s) * 2.f * kGeluC, fmaf(3.f * kGeluA, z * z, 1.f), s);
}
__device__ __forceinline__ float softplus_f(float x) {
  return x > 20.f ? x : log1pf(__expf(x));
}

// natural parameters -> packed BF16 operand blocks (forward: 22 blocks, then
// backward: 20 blocks), 64 lanes x 8 values each
__global__ __launch_bounds__(256) void k_mlp_pack(const float *__restrict__ th, bf16x8 *__restrict__ fwd,
                                                  bf16x8 *__restrict__ bwd) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= (kFwdBlocks + kBwdBlocks) * 64) return;
  const int b = t / 64, l = t % 64, g = l >> 4, c = l & 15;
  bf16x8 v;
  for (int i = 0; i < 8; i++) {
    float x = 0.f;
    if (b < kFwdBlocks) {
      if (b < 4) {  // layer 1: A[16 mb + c][k], k = 8 g + i; k = 10 is the bias
        const int m = 16 * b + c, k = 8 * g + i;
        x = k < kMlpNF ? th[kOffW1 + k * kMlpNH + m] : (k == kMlpNF ? th[kOffB1 + m] : 0.f);
      } else if (b < 20) {  // layers 2, 3: A[16 mb + c][perm(ks, g, i)]
        const int r = (b - 4) & 7, mb = r >> 1, ks = r & 1;
        const int off = b < 12 ? kOffW2 : kOffW3;
        x = th[off + mlp_perm(ks, g, i) * kMlpNH + 16 * mb + c];
      } else {  // layer 4: 7 outputs of 16 rows
        const int ks = b - 20;
        x = c < kMlpNO ? th[kOffW4 + mlp_perm(ks, g, i) * kMlpNO + c] : 0.f;
      }
    } else {
      const int bb = b - kFwdBlocks;
      if (bb < 4) {  // W4 (64 x 7): A[m = input 16 mb + c][k = output 8 g + i]
        const int m = 16 * bb + c, k = 8 * g + i;
        x = k < kMlpNO ? th[kOffW4 + m * kMlpNO + k] : 0.f;
      } else {  // W3 then W2: A[m = input 16 mb + c][k = output perm(ks, g, i)]
        const int r = (bb - 4) & 7, mb = r >> 1, ks = r & 1;
        const int off = bb < 12 ? kOffW3 : kOffW2;
        x = th[off + (16 * mb + c) * kMlpNH + mlp_perm(ks, g, i)];
      }
    }
    v[i] = (__bf16)x;
  }
  if (b < kFwdBlocks) fwd[b * 64 + l] = v;
  else bwd[(b - kFwdBlocks) * 64 + l] = v;
}

// B operand of layer 1 for node `node` (lane group g): scaled inputs 8g..8g+7,
// the constant 1 at k = 10
__device__ __forceinline__ bf16x8 mlp_input(const float *__restrict__ feat, long long fstride,
                                            long long node, bool ok, int g,
                                            const float *__restrict__ th) {
  bf16x8 B;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const int k = 8 * g + i;
    float v = 0.f;
    if (k < kMlpNF) v = ok ? feat[k * fstride + node] * th[kOffIn + k] : 0.f;
    else if (k == kMlpNF) v = 1.f;
    B[i] = (__bf16)v;
  }
  return B;
}
// the same in two halves: raw inputs of a tile loaded ahead of their use
struct MlpRaw {
  float v[8];
};
__device__ __forceinline__ void mlp_load(MlpRaw &r, const float *__restrict__ feat, long long fstride,
                                         long long node, bool ok, int g) {
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const int k = 8 * g + i;
    r.v[i] = (k < kMlpNF && ok) ? feat[k * fstride + node] : 0.f;
  }
}
__device__ __forceinline__ bf16x8 mlp_scale(const MlpRaw &r, int g, const float *__restrict__ th) {
  bf16x8 B;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const int k = 8 * g + i;
    const float v = k < kMlpNF ? r.v[i] * th[kOffIn + k] : (k == kMlpNF ? 1.f : 0.f);
    B[i] = (__bf16)v;
  }
  return B;
}

// GELU of four accumulator blocks, packed as the two K steps of the next layer
__device__ __forceinline__ void mlp_act(const f32x4 (&z)[4], bf16x8 (&B)[2]) {
#pragma unroll
  for (int ks = 0; ks < 2; ks++)
#pragma unroll
    for (int i = 0; i < 8; i++) B[ks][i] = (__bf16)gelu_f(z[2 * ks + (i >> 2)][i & 3]);
}

// one hidden layer (64 -> 64): packed weights, block k at A[k * ST] (mb-major,
// two K steps; ST = 1 for a register array, 64 for the LDS copy)
template <int ST = 1>
__device__ __forceinline__ void mlp_hidden(const bf16x8 *A, const bf16x8 (&B)[2],
                                           const float *__restrict__ bias, int g, f32x4 (&z)[4]) {
#pragma unroll
  for (int mb = 0; mb < 4; mb++) {
    f32x4 acc = *(const f32x4 *)(bias + 16 * mb + 4 * g);
    acc = mfma16(A[(2 * mb) * ST], B[0], acc);
    acc = mfma16(A[(2 * mb + 1) * ST], B[1], acc);
    z[mb] = acc;
  }
}

// output layer (64 -> 7 of 16 rows): lane group g holds outputs 4g..4g+3
template <int ST = 1>
__device__ __forceinline__ f32x4 mlp_out(const bf16x8 *A, const bf16x8 (&B)[2],
                                         const float *__restrict__ th, int g) {
  f32x4 acc;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int q = 4 * g + i;
    acc[i] = q < kMlpNO ? th[kOffB4 + q] : 0.f;
  }
  acc = mfma16(A[0], B[0], acc);
  acc = mfma16(A[ST], B[1], acc);
  return acc;
}

// Inference: features (SoA, 10 x fstride) of n nodes -> closure fields
// (SoA, 7 x cstride): nu_t = s0 softplus(o0), dsigma_k = s_{k+1} o_{k+1}.
// One wave per 16-node tile.  The forward weight blocks (22.5 KiB) sit in LDS,
// one copy per block, read as 16-byte operands per MFMA; registers stay free
// for more waves per SIMD, which hide the MFMA -> GELU -> MFMA chain.
__global__ __launch_bounds__(256) void k_mlp_infer(const float *__restrict__ feat, long long fstride,
                                                   long long n, const bf16x8 *__restrict__ wf,
                                                   const float *__restrict__ th, float *__restrict__ clo,
                                                   long long cstride) {
  __shared__ bf16x8 wl[kFwdBlocks * 64];
  for (int i = threadIdx.x; i < kFwdBlocks * 64; i += 256) wl[i] = wf[i];
  __syncthreads();
  const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
  const bf16x8 *A = wl + lane;  // block b at A[b * 64]
  const long long ntiles = (n + 15) / 16;
  const long long nw = (long long)gridDim.x * 4;
  long long t = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  MlpRaw nxt;
  mlp_load(nxt, feat, fstride, t * 16 + c, t < ntiles && t * 16 + c < n, g);
  for (; t < ntiles; t += nw) {
    const long long node = t * 16 + c;
    const bool ok = node < n;
    // keep the LDS weight reads inside the loop (hoisted, they would occupy
    // the registers the LDS copy is meant to free)
    asm volatile("" ::: "memory");
    const MlpRaw cur = nxt;  // the next tile's inputs load while this one computes
    const long long tn = t + nw, nn = tn * 16 + c;
    mlp_load(nxt, feat, fstride, nn, tn < ntiles && nn < n, g);
    const bf16x8 B1 = mlp_scale(cur, g, th);
    f32x4 z[4];
    const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int mb = 0; mb < 4; mb++) z[mb] = mfma16(A[mb * 64], B1, zero);
    bf16x8 B[2];
    mlp_act(z, B);
    mlp_hidden<64>(A + 4 * 64, B, th + kOffB2, g, z);
    mlp_act(z, B);
    mlp_hidden<64>(A + 12 * 64, B, th + kOffB3, g, z);
    mlp_act(z, B);
    const f32x4 o = mlp_out<64>(A + 20 * 64, B, th, g);
    if (ok && g < 2) {
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int q = 4 * g + i;
        if (q < kMlpNO) {
          const float s = th[kOffOut + q];
          clo[q * cstride + node] = q == 0 ? s * softplus_f(o[i]) : s * o[i];
        }
      }
    }
  }
}

// Training, pass 1: forward and backward of the batch samples idx[0..nb).
// Stores, feature-major with node stride ldb (bf16): H0 (16 rows: scaled
// inputs, the constant 1), H1..H3 (64 rows), D1..D3 (64 rows), D4 (16 rows),
// and adds the FP64 loss sum.  inv = 1/(t_scale^2 * batch).
struct MlpTrainBufs {
  __bf16 *h0, *h1, *h2, *h3, *d1, *d2, *d3, *d4;
};
__global__ __launch_bounds__(256) void k_mlp_fwd_bwd(const float *__restrict__ feat, long long fstride,
                                                     const float *__restrict__ aux, long long astride,
                                                     const int *__restrict__ idx, int nb, int ldb,
                                                     const bf16x8 *__restrict__ wf, const bf16x8 *__restrict__ wb,
                                                     const float *__restrict__ th, float inv, MlpTrainBufs tb,
                                                     double *__restrict__ loss) {
  const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
  const int ntiles = (nb + 15) / 16;
  double lsum = 0.0;
  for (int t = blockIdx.x * 4 + (threadIdx.x >> 6); t < ntiles; t += gridDim.x * 4) {
    const int col = t * 16 + c;  // batch column
    const bool ok = col < nb;
    const long long node = ok ? idx[col] : 0;
    // ---- forward
    const bf16x8 B1 = mlp_input(feat, fstride, node, ok, g, th);
    if (g < 2) {  // rows 0..15: ten inputs, the constant 1, zeros
#pragma unroll
      for (int i = 0; i < 8; i++) tb.h0[(8 * g + i) * ldb + col] = B1[i];
    }
    const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
    f32x4 z1[4], z2[4], z3[4];
#pragma unroll
    for (int mb = 0; mb < 4; mb++) z1[mb] = mfma16(wf[mb * 64 + lane], B1, zero);
    bf16x8 B2[2], B3[2], B4[2];
    mlp_act(z1, B2);
    {
      bf16x8 A[8];
#pragma unroll
      for (int k = 0; k < 8; k++) A[k] = wf[(4 + k) * 64 + lane];
      mlp_hidden(A, B2, th + kOffB2, g, z2);
    }
    mlp_act(z2, B3);
    {
      bf16x8 A[8];
#pragma unroll
      for (int k = 0; k < 8; k++) A[k] = wf[(12 + k) * 64 + lane];
      mlp_hidden(A, B3, th + kOffB3, g, z3);
    }
    mlp_act(z3, B4);
    f32x4 o;
    {
      bf16x8 A[2] = {wf[20 * 64 + lane], wf[21 * 64 + lane]};
      o = mlp_out(A, B4, th, g);
    }
    // activations, feature-major
#pragma unroll
    for (int ks = 0; ks < 2; ks++)
#pragma unroll
      for (int i = 0; i < 8; i++) {
        const int f = mlp_perm(ks, g, i);
        tb.h1[f * ldb + col] = B2[ks][i];
        tb.h2[f * ldb + col] = B3[ks][i];
        tb.h3[f * ldb + col] = B4[ks][i];
      }
    // ---- loss and output deltas (lane group 0 holds node c's outputs 0..3,
    // group 1 outputs 4..6)
    float o4 = __shfl(o[0], c + 16), o5 = __shfl(o[1], c + 16), o6 = __shfl(o[2], c + 16);
    bf16x8 Bd;
#pragma unroll
    for (int i = 0; i < 8; i++) Bd[i] = (__bf16)0.f;
    if (g == 0) {
      float dlt[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      if (ok) {
        const float ov[7] = {o[0], o[1], o[2], o[3], o4, o5, o6};
        const float s0 = th[kOffOut];
        const float nu = s0 * softplus_f(ov[0]);
        const float rho = aux[node];
        float dnu = 0.f, l = 0.f;
#pragma unroll
        for (int k = 0; k < 6; k++) {
          const float S = feat[k * fstride + node];
          const float T = aux[(1 + k) * astride + node];
          const float sk = th[kOffOut + 1 + k];
          const float M = fmaf(-2.f * rho * nu, S, sk * ov[1 + k]);
          const float r = M - T;
          const float wk = k < 3 ? 1.f : 2.f;
          l += wk * r * r;
          const float dM = 2.f * wk * r * inv;
          dnu += dM * (-2.f * rho * S);
          dlt[1 + k] = dM * sk;
        }
        dlt[0] = dnu * s0 / (1.f + __expf(-ov[0]));
        lsum += (double)l * (double)inv;
      }
#pragma unroll
      for (int i = 0; i < 7; i++) Bd[i] = (__bf16)dlt[i];
    }
    if (g == 0) {  // rows 0..15: seven output deltas, zeros
#pragma unroll
      for (int i = 0; i < 8; i++) {
        tb.d4[i * ldb + col] = Bd[i];
        tb.d4[(8 + i) * ldb + col] = (__bf16)0.f;
      }
    }
    // ---- backward: delta_3 = (W4 delta_4) * gelu'(z3), ...
    f32x4 d[4];
#pragma unroll
    for (int mb = 0; mb < 4; mb++) {
      const f32x4 acc = mfma16(wb[mb * 64 + lane], Bd, zero);
#pragma unroll
      for (int i = 0; i < 4; i++) d[mb][i] = acc[i] * gelu_grad_f(z3[mb][i]);
    }
    auto store_delta = [&](__bf16 *dst, bf16x8 (&Bp)[2]) {
      // round to BF16 (the operand of the next product and of the weight gradient)
#pragma unroll
      for (int ks = 0; ks < 2; ks++)
#pragma unroll
        for (int i = 0; i < 8; i++) {
          Bp[ks][i] = (__bf16)d[2 * ks + (i >> 2)][i & 3];
          dst[mlp_perm(ks, g, i) * ldb + col] = Bp[ks][i];
        }
    };
    bf16x8 Bp[2];
    store_delta(tb.d3, Bp);
#pragma unroll
    for (int mb = 0; mb < 4; mb++) {
      f32x4 acc = mfma16(wb[(4 + 2 * mb) * 64 + lane], Bp[0], zero);
      acc = mfma16(wb[(4 + 2 * mb + 1) * 64 + lane], Bp[1], acc);
#pragma unroll
      for (int i = 0; i < 4; i++) d[mb][i] = acc[i] * gelu_grad_f(z2[mb][i]);
    }
    store_delta(tb.d2, Bp);
#pragma unroll
    for (int mb = 0; mb < 4; mb++) {
      f32x4 acc = mfma16(wb[(12 + 2 * mb) * 64 + lane], Bp[0], zero);
      acc = mfma16(wb[(12 + 2 * mb + 1) * 64 + lane], Bp[1], acc);
#pragma unroll
      for (int i = 0; i < 4; i++) d[mb][i] = acc[i] * gelu_grad_f(z1[mb][i]);
    }
    store_delta(tb.d1, Bp);
  }
  // FP64 loss: wave sum, one atomic per wave
  for (int off = 32; off > 0; off >>= 1) lsum += __shfl_down(lsum, off);
  if (lane == 0 && lsum != 0.0) atomicAdd(loss, lsum);
}

// Training, pass 2: weight gradients on the matrix cores.  Output blocks
// (16 x 16 of dW^T = delta H^T, or of the bias column): job j of
// kWgradJobs, K (batch columns) split into slices of ksl columns.  The FP32
// gradient is in the natural parameter layout; its bias entries get the
// products with a column of ones.
struct WgradJob {
  int layer;  // 1..4
  int mb;     // output-feature block (rows of delta)
  int nb;     // input-feature block (rows of H), -1: bias
};
constexpr int kWgradJobs = 4 + 20 + 20 + 5;
__device__ __forceinline__ WgradJob wgrad_job(int j) {
  if (j < 4) return {1, j, 0};  // 16 input rows: 10 features, the bias row, zeros
  j -= 4;
  if (j < 20) return {2, j / 5, j % 5 == 4 ? -1 : j % 5};
  j -= 20;
  if (j < 20) return {3, j / 5, j % 5 == 4 ? -1 : j % 5};
  j -= 20;
  return {4, 0, j == 4 ? -1 : j};
}
__global__ __launch_bounds__(256) void k_mlp_wgrad(MlpTrainBufs tb, int nb, int ldb, int ksl,
                                                   float *__restrict__ grad) {
  const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
  const int w = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int nslice = (nb + ksl - 1) / ksl;
  if (w >= kWgradJobs * nslice) return;
  const WgradJob jb = wgrad_job(w / nslice);
  const int k0 = (w % nslice) * ksl, k1 = min(k0 + ksl, nb);
  const __bf16 *D = jb.layer == 1 ? tb.d1 : (jb.layer == 2 ? tb.d2 : (jb.layer == 3 ? tb.d3 : tb.d4));
  const __bf16 *H = jb.layer == 1 ? tb.h0 : (jb.layer == 2 ? tb.h1 : (jb.layer == 3 ? tb.h2 : tb.h3));
  const __bf16 *drow = D + (size_t)(16 * jb.mb + c) * ldb;
  const __bf16 *hrow = H + (size_t)(16 * (jb.nb < 0 ? 0 : jb.nb) + c) * ldb;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  bf16x8 ones;
#pragma unroll
  for (int i = 0; i < 8; i++) ones[i] = (__bf16)(c == 0 ? 1.f : 0.f);
  for (int k = k0; k < k1; k += 32) {
    bf16x8 a, b;
    if (k + 32 <= k1) {
      // whole K step: 16-byte loads (rows start at multiples of 32 columns)
      a = *(const bf16x8 *)(drow + k + 8 * g);
      b = *(const bf16x8 *)(hrow + k + 8 * g);
    } else {
#pragma unroll
      for (int i = 0; i < 8; i++) {
        const int col = k + 8 * g + i;
        const bool in = col < k1;
        a[i] = in ? drow[col] : (__bf16)0.f;
        b[i] = in ? hrow[col] : (__bf16)0.f;
      }
    }
    acc = mfma16(a, jb.nb < 0 ? ones : b, acc);
  }
  // acc[i] = sum over columns of delta[16 mb + 4g + i] * H[16 nb + c]
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int jo = 16 * jb.mb + 4 * g + i;  // output feature
    const int ki = 16 * (jb.nb < 0 ? 0 : jb.nb) + c;  // input feature
    int off = -1;
    switch (jb.layer) {
      case 1:
        if (ki < kMlpNF) off = kOffW1 + ki * kMlpNH + jo;
        else if (ki == kMlpNF) off = kOffB1 + jo;
        break;
      case 2:
        if (jb.nb >= 0) off = kOffW2 + ki * kMlpNH + jo;
        else if (c == 0) off = kOffB2 + jo;
        break;
      case 3:
        if (jb.nb >= 0) off = kOffW3 + ki * kMlpNH + jo;
        else if (c == 0) off = kOffB3 + jo;
        break;
      default:
        if (jo < kMlpNO) {
          if (jb.nb >= 0) off = kOffW4 + ki * kMlpNO + jo;
          else if (c == 0) off = kOffB4 + jo;
        }
    }
    if (off >= 0 && acc[i] != 0.f) atomicAdd(grad + off, acc[i]);
  }
}

// Adam on the weights and biases (the scales stay); t is the 1-based step
__global__ __launch_bounds__(256) void k_mlp_adam(float *__restrict__ th, const float *__restrict__ grad,
                                                  float *__restrict__ m1, float *__restrict__ m2, float lr,
                                                  float b1, float b2, float eps, int t) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= kOffIn) return;
  const float gi = grad[i];
  const float a = fmaf(b1, m1[i], (1.f - b1) * gi);
  const float v = fmaf(b2, m2[i], (1.f - b2) * gi * gi);
  m1[i] = a;
  m2[i] = v;
  const float mh = a / (1.f - powf(b1, (float)t));
  const float vh = v / (1.f - powf(b2, (float)t));
  th[i] -= lr * mh / (sqrtf(vh) + eps);
}

// Field views for the filter and input kernels: the lattice state (storage
// type S, per-field scale) or a plain FP32 SoA buffer.
template <typename S>
struct StateView {
  const S *base;
  long long fs;
  __device__ __forceinline__ float get(int q, long long id, const Params &p) const {
    return dec(base[q * fs + id], p, sgrp(q));
  }
};
struct PlainView {
  const float *base;
  long long fs;
  __device__ __forceinline__ float get(int q, long long id, const Params &) const { return base[q * fs + id]; }
};

// Box filter of a fine lattice by r^3 blocks: coarse rho_bar - 1, Favre
// velocity, phi_bar (SoA, 5 x nc) and the aux record rho_bar, tau[6]
// (SoA, 7 x astride, written at column offset aoff).  tau = <rho (u-u~)(u-u~)>.
template <typename S>
__global__ __launch_bounds__(256) void k_box_filter(StateView<S> f, Params p, int r, float *__restrict__ coarse,
                                                    long long nc, float *__restrict__ aux, long long astride,
                                                    long long aoff) {
  const int ncx = p.nx / r, ncy = p.ny / r, ncz = p.nz / r;
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= nc) return;
  const int I = (int)(t % ncx), J = (int)((t / ncx) % ncy), K = (int)(t / ((long long)ncx * ncy));
  (void)ncz;
  const long long plane = (long long)p.nx * p.ny;
  double sr = 0, su = 0, sv = 0, sw = 0, sp = 0;
  for (int k = 0; k < r; k++)
    for (int j = 0; j < r; j++)
      for (int i = 0; i < r; i++) {
        const long long id = (long long)(I * r + i) + (long long)p.nx * (J * r + j) + plane * (K * r + k);
        const float rho = 1.f + f.get(TSLB_RHO, id, p);
        sr += rho;
        su += (double)rho * f.get(TSLB_U, id, p);
        sv += (double)rho * f.get(TSLB_V, id, p);
        sw += (double)rho * f.get(TSLB_W, id, p);
        if (p.ncomp == 2) sp += f.get(TSLB_PHI, id, p);
      }
  const double inv = 1.0 / ((double)r * r * r);
  const double rb = sr * inv, ut = su / sr, vt = sv / sr, wt = sw / sr;
  double txx = 0, tyy = 0, tzz = 0, txy = 0, txz = 0, tyz = 0;
  for (int k = 0; k < r; k++)
    for (int j = 0; j < r; j++)
      for (int i = 0; i < r; i++) {
        const long long id = (long long)(I * r + i) + (long long)p.nx * (J * r + j) + plane * (K * r + k);
        const double rho = 1.0 + f.get(TSLB_RHO, id, p);
        const double du = f.get(TSLB_U, id, p) - ut, dv = f.get(TSLB_V, id, p) - vt,
                     dw = f.get(TSLB_W, id, p) - wt;
        txx += rho * du * du;
        tyy += rho * dv * dv;
        tzz += rho * dw * dw;
        txy += rho * du * dv;
        txz += rho * du * dw;
        tyz += rho * dv * dw;
      }
  coarse[t] = (float)(rb - 1.0);
  coarse[nc + t] = (float)ut;
  coarse[2 * nc + t] = (float)vt;
  coarse[3 * nc + t] = (float)wt;
  coarse[4 * nc + t] = (float)(sp * inv);
  const long long a = aoff + t;
  aux[a] = (float)rb;
  aux[astride + a] = (float)(txx * inv);
  aux[2 * astride + a] = (float)(tyy * inv);
  aux[3 * astride + a] = (float)(tzz * inv);
  aux[4 * astride + a] = (float)(txy * inv);
  aux[5 * astride + a] = (float)(txz * inv);
  aux[6 * astride + a] = (float)(tyz * inv);
}

// Closure inputs of every node of a grid (nx, ny, nz, periodic flags of p):
// S from the isotropic gradient (3 sum_l w_l c_l f(x + c_l), the stencil of
// grad_axis) of (u, v, w), grad(phi), phi; the nearest
// interior node across a wall.  Velocity at fields fu..fu+2 and phi at field
// fphi of a SoA source (phi absent: fphi < 0).  Output SoA 10 x fstride at
// column offset foff.
template <typename Src>
__global__ __launch_bounds__(256) void k_features(Src f, Params p, int fu, int fphi, float *__restrict__ feat,
                                                  long long fstride, long long foff) {
  const long long n = (long long)p.nx * p.ny * p.nz;
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  const int i = (int)(t % p.nx), j = (int)((t / p.nx) % p.ny), k = (int)(t / ((long long)p.nx * p.ny));
  // 3 sum_l w_l c_l f(x + c_l) for u, v, w and phi, one neighbour at a time
  float gu[3] = {0.f, 0.f, 0.f}, gv[3] = {0.f, 0.f, 0.f}, gw[3] = {0.f, 0.f, 0.f}, gp[3] = {0.f, 0.f, 0.f};
#pragma unroll 1
  for (int l = 1; l < 27; l++) {
    const int cx = kCX[l], cy = kCY[l], cz = kCZ[l];
    const long long id = (long long)clampw(i + cx, p.nx, p.px) +
                         (long long)p.nx * (clampw(j + cy, p.ny, p.py) + (long long)p.ny * clampw(k + cz, p.nz, p.pz));
    const float w3 = 3.f * wgt(l);
    const float c[3] = {w3 * (float)cx, w3 * (float)cy, w3 * (float)cz};
    const float u = f.get(fu, id, p), v = f.get(fu + 1, id, p), w = f.get(fu + 2, id, p);
    const float ph = fphi >= 0 ? f.get(fphi, id, p) : 0.f;
#pragma unroll
    for (int a = 0; a < 3; a++) {
      gu[a] = fmaf(c[a], u, gu[a]);
      gv[a] = fmaf(c[a], v, gv[a]);
      gw[a] = fmaf(c[a], w, gw[a]);
      gp[a] = fmaf(c[a], ph, gp[a]);
    }
  }
  const float ph = fphi >= 0 ? f.get(fphi, t, p) : 0.f;
  const long long o = foff + t;
  feat[o] = gu[0];
  feat[fstride + o] = gv[1];
  feat[2 * fstride + o] = gw[2];
  feat[3 * fstride + o] = 0.5f * (gu[1] + gv[0]);
  feat[4 * fstride + o] = 0.5f * (gu[2] + gw[0]);
  feat[5 * fstride + o] = 0.5f * (gv[2] + gw[1]);
  feat[6 * fstride + o] = gp[0];
  feat[7 * fstride + o] = gp[1];
  feat[8 * fstride + o] = gp[2];
  feat[9 * fstride + o] = ph;
}
