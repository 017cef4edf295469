// Candidate prefilter network: MLP 32 -> 128 -> 128 -> 1 (ReLU), BF16 operands
// on the matrix cores (v_mfma_f32_32x32x8_bf16) with FP32 accumulation.
//
// Rounding points (the CPU check in the tests mirrors them):
//   a1 = bf16(x) . bf16(W1) + b1,  h1 = bf16(relu(a1))
//   a2 = h1 . bf16(W2) + b2,       h2 = bf16(relu(a2))
//   z  = h2 . bf16(W3) + b3
// Backward for loss w * BCE(sigmoid(z), y) / B:
//   dz = w (sigmoid(z) - y) / B
//   dW3 = h2^T dz, db3 = sum dz, dh2 = bf16((dz W3^T) * [a2 > 0])
//   dW2 = h1^T dh2, db2 = sum dh2, dh1 = bf16((dh2 W2^T) * [a1 > 0])
//   dW1 = bf16(x)^T dh1, db1 = sum dh1
// One wavefront handles a tile of 32 rows; all transposes go through LDS.
#include "pbw_internal.h"

#include <hip/hip_runtime.h>

namespace pbw {

typedef short s4v __attribute__((ext_vector_type(4)));
typedef float f16v __attribute__((ext_vector_type(16)));

__device__ __forceinline__ uint16_t f2bf(float f) {  // round to nearest even
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ f16v mfma(s4v a, s4v b, f16v c) {
  return __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, c, 0, 0, 0);
}
// Operand of lane l for k-step s: 4 consecutive bf16 at row (l % 32), columns
// 8 s + 4 (l / 32) .. +3 of a row-major [32 x ld] bf16 tile.
__device__ __forceinline__ s4v ld_op(const uint16_t* base, int ld, int s) {
  const int l = threadIdx.x & 63;
  return *(const s4v*)(base + (l & 31) * ld + 8 * s + 4 * (l >> 5));
}
// Accumulator element r of lane l holds row 8 (r / 4) + 4 (l / 32) + r % 4, column l % 32.
__device__ __forceinline__ int acc_row(int r) { return 8 * (r >> 2) + 4 * ((threadIdx.x & 63) >> 5) + (r & 3); }

struct MlpLds {
  uint16_t x[32 * 32];     // x tile, row-major [m][f]
  uint16_t xt[32 * 32];    // transposed [f][m]
  uint16_t h1[32 * 128];   // [m][k]
  uint16_t h1t[128 * 32];  // [k][m]
  uint16_t h2[32 * 128];   // [m][n]
  uint16_t d2[32 * 128];   // dh2 [m][n]
  uint16_t d2t[128 * 32];  // dh2 [n][m]
  uint16_t d1t[128 * 32];  // dh1 [k][m]
  float z[32];
  float dz[32];
};

// Forward of one 32-row tile from the x tile in LDS.  Leaves h1, h1t, h2 and z
// in LDS; mask1/mask2 get bit (t * 16 + r) set when accumulator element r of
// output tile t had a > 0 (the ReLU derivative of layers 1 and 2).
__device__ void mlp_forward_tile(const MlpDev& P, MlpLds& S, uint64_t& mask1, uint64_t& mask2) {
  const int l = threadIdx.x & 63;
  // layer 1: [32 x 32] . [32 x 128]
  for (int t = 0; t < 4; t++) {
    f16v acc = {};
#pragma unroll
    for (int s = 0; s < 4; s++) acc = mfma(ld_op(S.x, 32, s), ld_op(P.w1t + 32 * 32 * t, 32, s), acc);
    const int n = 32 * t + (l & 31);
    const float b = P.b1[n];
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const int m = acc_row(r);
      const float a = acc[r] + b;
      const uint16_t h = f2bf(a > 0.f ? a : 0.f);
      S.h1[m * 128 + n] = h;
      S.h1t[n * 32 + m] = h;
      const int bit = t * 16 + r;
      if (a > 0.f) mask1 |= 1ull << bit;
    }
  }
  __syncthreads();
  // layer 2: [32 x 128] . [128 x 128]
  for (int t = 0; t < 4; t++) {
    f16v acc = {};
#pragma unroll 2
    for (int s = 0; s < 16; s++) acc = mfma(ld_op(S.h1, 128, s), ld_op(P.w2t + 32 * 128 * t, 128, s), acc);
    const int n = 32 * t + (l & 31);
    const float b = P.b2[n];
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const int m = acc_row(r);
      const float a = acc[r] + b;
      S.h2[m * 128 + n] = f2bf(a > 0.f ? a : 0.f);
      const int bit = t * 16 + r;
      if (a > 0.f) mask2 |= 1ull << bit;
    }
  }
  __syncthreads();
  // layer 3: z[m] = h2[m] . W3 + b3; lane l sums half (l / 32) of row l % 32
  {
    const int m = l & 31, half = l >> 5;
    float zs = 0.f;
    for (int n = 64 * half; n < 64 * half + 64; n++) zs += bf2f(S.h2[m * 128 + n]) * bf2f(P.w3[n]);
    zs += __shfl_xor(zs, 32, 64);
    if (half == 0) S.z[m] = zs + P.b3[0];
  }
  __syncthreads();
}

__device__ void mlp_load_x(const uint16_t* X, const uint32_t* perm, uint64_t row0, uint64_t nrows, uint64_t N,
                           MlpLds& S) {
  const int l = threadIdx.x & 63;
  for (int e = l; e < 32 * 32; e += 64) {
    const int m = e >> 5, f = e & 31;
    uint16_t v = 0;
    if ((uint64_t)m < nrows) {
      const uint64_t idx = perm ? perm[row0 + m] : row0 + m;
      if (idx < N) v = X[idx * 32 + f];
    }
    S.x[m * 32 + f] = v;
    S.xt[f * 32 + m] = v;
  }
  __syncthreads();
}

// Inference forward.  Every layer is computed transposed, out^T = W^T . in^T:
// the accumulator of lane l then holds sample l % 32 and hidden units
// 32 t + 8 g + 4 (l / 32) + j (element 4 g + j of output tile t), which is
// exactly the B operand of k-step 4 t + g of the next layer.  So activations
// stay in registers; only W1^T and W2^T go through LDS, staged once per block
// with rows padded by 4 BF16 so the 32 rows a k-step reads sit in distinct
// banks.  Each of the block's waves takes 32-row tiles in turn.
constexpr int MLPI_WAVES = 4;
constexpr int W1T_LD = 32 + 4, W2T_LD = 128 + 4;

__device__ __forceinline__ s4v lds_op(const uint16_t* base, int ld, int s) {
  const int l = threadIdx.x & 63;
  return *(const s4v*)(base + (l & 31) * ld + 8 * s + 4 * (l >> 5));
}

__global__ __launch_bounds__(64 * MLPI_WAVES) void k_mlp_infer(MlpDev P, const uint16_t* X, uint64_t N, float* z) {
  __shared__ __align__(16) uint16_t w1t[128 * W1T_LD];
  __shared__ __align__(16) uint16_t w2t[128 * W2T_LD];
  __shared__ float b1s[128], b2s[128], w3s[128];
  for (int e = threadIdx.x; e < 128 * 32 / 4; e += 64 * MLPI_WAVES) {  // 4 BF16 per move
    const int n = e >> 3, f = (e & 7) * 4;
    *(uint64_t*)(w1t + n * W1T_LD + f) = *(const uint64_t*)(P.w1t + n * 32 + f);
  }
  for (int e = threadIdx.x; e < 128 * 128 / 4; e += 64 * MLPI_WAVES) {
    const int n = e >> 5, k = (e & 31) * 4;
    *(uint64_t*)(w2t + n * W2T_LD + k) = *(const uint64_t*)(P.w2t + n * 128 + k);
  }
  if (threadIdx.x < 128) {
    b1s[threadIdx.x] = P.b1[threadIdx.x];
    b2s[threadIdx.x] = P.b2[threadIdx.x];
    w3s[threadIdx.x] = bf2f(P.w3[threadIdx.x]);
  }
  __syncthreads();
  const int l = threadIdx.x & 63, half = l >> 5;
  const float b3 = P.b3[0];
  const uint64_t ntiles = (N + 31) / 32;
  for (uint64_t tile = (uint64_t)blockIdx.x * MLPI_WAVES + (threadIdx.x >> 6); tile < ntiles;
       tile += (uint64_t)gridDim.x * MLPI_WAVES) {
    const uint64_t row = tile * 32 + (l & 31);
    s4v xb[4];
#pragma unroll
    for (int s = 0; s < 4; s++) xb[s] = row < N ? *(const s4v*)(X + row * 32 + 8 * s + 4 * half) : s4v{0, 0, 0, 0};
    // layer 1: h1^T = bf16(relu(W1^T x^T + b1)), kept as layer 2's B operands
    s4v hb[16];
#pragma unroll
    for (int t = 0; t < 4; t++) {
      f16v acc = {};
#pragma unroll
      for (int s = 0; s < 4; s++) acc = mfma(lds_op(w1t + 32 * W1T_LD * t, W1T_LD, s), xb[s], acc);
#pragma unroll
      for (int g = 0; g < 4; g++) {
        s4v h;
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const float a = acc[4 * g + j] + b1s[32 * t + 8 * g + 4 * half + j];
          h[j] = (short)f2bf(a > 0.f ? a : 0.f);
        }
        hb[4 * t + g] = h;
      }
    }
    // layer 2 and 3: z = bf16(relu(W2^T h1^T + b2)) . W3 + b3
    float zp = 0.f;
#pragma unroll 1
    for (int t = 0; t < 4; t++) {
      f16v acc = {};
#pragma unroll
      for (int s = 0; s < 16; s++) acc = mfma(lds_op(w2t + 32 * W2T_LD * t, W2T_LD, s), hb[s], acc);
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const int n = 32 * t + 8 * (r >> 2) + 4 * half + (r & 3);
        const float a = acc[r] + b2s[n];
        zp += bf2f(f2bf(a > 0.f ? a : 0.f)) * w3s[n];
      }
    }
    zp += __shfl_xor(zp, 32, 64);
    if (half == 0 && row < N) z[row] = zp + b3;
  }
}

// Forward + backward of one 32-row tile per block; the gradient of tile w
// goes to part[w * MLP_NP ..] and its loss to loss_part[w].
__global__ __launch_bounds__(64) void k_mlp_grad(MlpDev P, const uint16_t* X, const float* Y, const uint32_t* perm,
                                                 uint64_t row0, uint64_t nrows, uint64_t N, float pos_weight,
                                                 float inv_b, float* part, float* loss_part) {
  extern __shared__ __align__(16) uint8_t dyn[];
  MlpLds& S = *(MlpLds*)dyn;
  const int l = threadIdx.x;
  const uint32_t ntiles = (uint32_t)((nrows + 31) / 32);
  for (uint32_t w = blockIdx.x; w < ntiles; w += gridDim.x) {
    const uint64_t r0 = row0 + (uint64_t)w * 32;
    const uint64_t tr = nrows - (uint64_t)w * 32 < 32 ? nrows - (uint64_t)w * 32 : 32;
    mlp_load_x(X, perm, r0, tr, N, S);
    uint64_t m1 = 0, m2 = 0;
    mlp_forward_tile(P, S, m1, m2);
    float* G = part + (uint64_t)w * MLP_NP;
    auto put = [&](uint32_t idx, float v) { G[idx] = v; };
    // dz and the loss
    if (l < 32) {
      float dz = 0.f, ls = 0.f;
      if ((uint64_t)l < tr) {
        const uint64_t idx = perm ? perm[r0 + l] : r0 + l;
        const float y = Y[idx], zz = S.z[l];
        const float wgt = y > 0.5f ? pos_weight : 1.f;
        const float sg = 1.f / (1.f + expf(-zz));
        dz = wgt * (sg - y) * inv_b;
        // BCE with logits: max(z,0) - z y + log(1 + exp(-|z|))
        ls = wgt * (fmaxf(zz, 0.f) - zz * y + log1pf(expf(-fabsf(zz))));
      }
      S.dz[l] = dz;
      float t = ls;
      for (int o = 16; o > 0; o >>= 1) t += __shfl_xor(t, o, 32);
      if (l == 0) loss_part[w] = t;
    }
    __syncthreads();
    // dW3 (columns l and l + 64), db3; dh2 = dz W3^T * [a2 > 0]
    for (int n = l; n < 128; n += 64) {
      float g = 0.f;
      for (int m = 0; m < 32; m++) g += bf2f(S.h2[m * 128 + n]) * S.dz[m];
      put(MLP_OFF_W3 + n, g);
    }
    if (l == 0) {
      float g = 0.f;
      for (int m = 0; m < 32; m++) g += S.dz[m];
      put(MLP_OFF_B3, g);
    }
    for (int t = 0; t < 4; t++) {
      const int n = 32 * t + (l & 31);
      const float w3 = bf2f(P.w3[n]);
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const int m = acc_row(r);
        const int bit = t * 16 + r;
        const bool on = (m2 >> bit) & 1ull;
        const uint16_t d = f2bf(on ? S.dz[m] * w3 : 0.f);
        S.d2[m * 128 + n] = d;
        S.d2t[n * 32 + m] = d;
      }
    }
    __syncthreads();
    // db2
    for (int n = l; n < 128; n += 64) {
      float g = 0.f;
      for (int m = 0; m < 32; m++) g += bf2f(S.d2[m * 128 + n]);
      put(MLP_OFF_B2 + n, g);
    }
    // dW2[k][n] = sum_m h1[m][k] dh2[m][n]: A = h1^T (rows k), B = dh2 (cols n)
    for (int kt = 0; kt < 4; kt++) {
      for (int nt = 0; nt < 4; nt++) {
        f16v acc = {};
#pragma unroll
        for (int s = 0; s < 4; s++) acc = mfma(ld_op(S.h1t + 32 * 32 * kt, 32, s), ld_op(S.d2t + 32 * 32 * nt, 32, s), acc);
        const int n = 32 * nt + (l & 31);
#pragma unroll
        for (int r = 0; r < 16; r++) put(MLP_OFF_W2 + (32 * kt + acc_row(r)) * 128 + n, acc[r]);
      }
    }
    // dh1 = dh2 W2^T * [a1 > 0]: A = dh2 (rows m, k over n), B[n][k] = W2[k][n] from w2 ([k][n] row-major)
    for (int t = 0; t < 4; t++) {
      f16v acc = {};
#pragma unroll 2
      for (int s = 0; s < 16; s++) acc = mfma(ld_op(S.d2, 128, s), ld_op(P.w2 + 32 * 128 * t, 128, s), acc);
      const int k = 32 * t + (l & 31);
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const int m = acc_row(r);
        const int bit = t * 16 + r;
        const bool on = (m1 >> bit) & 1ull;
        S.d1t[k * 32 + m] = f2bf(on ? acc[r] : 0.f);
      }
    }
    __syncthreads();
    // db1
    for (int k = l; k < 128; k += 64) {
      float g = 0.f;
      for (int m = 0; m < 32; m++) g += bf2f(S.d1t[k * 32 + m]);
      put(MLP_OFF_B1 + k, g);
    }
    // dW1[f][k] = sum_m x[m][f] dh1[m][k]: A = x^T (rows f), B = dh1 (cols k)
    for (int kt = 0; kt < 4; kt++) {
      f16v acc = {};
#pragma unroll
      for (int s = 0; s < 4; s++) acc = mfma(ld_op(S.xt, 32, s), ld_op(S.d1t + 32 * 32 * kt, 32, s), acc);
      const int k = 32 * kt + (l & 31);
#pragma unroll
      for (int r = 0; r < 16; r++) put(MLP_OFF_W1 + acc_row(r) * 128 + k, acc[r]);
    }
    __syncthreads();
  }
}

// Sum of the tile partials in tile order (deterministic), then either the raw
// gradient (adam == 0) or one Adam step on the FP32 master weights and the
// BF16 copies.
__global__ void k_mlp_reduce_adam(MlpDev P, const float* part, uint32_t ntiles, float* grad_out, int adam, float lr,
                                  float b1, float b2, float eps, float bc1, float bc2) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (uint32_t)MLP_NP) return;
  float g = 0.f;
  for (uint32_t w = 0; w < ntiles; w++) g += part[(uint64_t)w * MLP_NP + p];
  if (grad_out) grad_out[p] = g;
  if (!adam) return;
  float m = P.m[p] = b1 * P.m[p] + (1.f - b1) * g;
  float v = P.v[p] = b2 * P.v[p] + (1.f - b2) * g * g;
  const float mh = m / bc1, vh = v / bc2;
  const float th = P.theta[p] - lr * mh / (sqrtf(vh) + eps);
  P.theta[p] = th;
  const uint16_t hb = f2bf(th);
  if (p < (uint32_t)MLP_OFF_B1) {  // W1 [f][k]
    const uint32_t f = p / 128, k = p % 128;
    P.w1t[k * 32 + f] = hb;
  } else if (p < (uint32_t)MLP_OFF_W2) {
    P.b1[p - MLP_OFF_B1] = th;
  } else if (p < (uint32_t)MLP_OFF_B2) {  // W2 [k][n]
    const uint32_t q = p - MLP_OFF_W2, k = q / 128, n = q % 128;
    P.w2[k * 128 + n] = hb;
    P.w2t[n * 128 + k] = hb;
  } else if (p < (uint32_t)MLP_OFF_W3) {
    P.b2[p - MLP_OFF_B2] = th;
  } else if (p < (uint32_t)MLP_OFF_B3) {
    P.w3[p - MLP_OFF_W3] = hb;
  } else {
    P.b3[0] = th;
  }
}

// Writes the BF16 operand copies and FP32 biases from theta.
__global__ void k_mlp_publish(MlpDev P) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (uint32_t)MLP_NP) return;
  const float th = P.theta[p];
  const uint16_t hb = f2bf(th);
  if (p < (uint32_t)MLP_OFF_B1) {
    P.w1t[(p % 128) * 32 + p / 128] = hb;
  } else if (p < (uint32_t)MLP_OFF_W2) {
    P.b1[p - MLP_OFF_B1] = th;
  } else if (p < (uint32_t)MLP_OFF_B2) {
    const uint32_t q = p - MLP_OFF_W2, k = q / 128, n = q % 128;
    P.w2[k * 128 + n] = hb;
    P.w2t[n * 128 + k] = hb;
  } else if (p < (uint32_t)MLP_OFF_W3) {
    P.b2[p - MLP_OFF_B2] = th;
  } else if (p < (uint32_t)MLP_OFF_B3) {
    P.w3[p - MLP_OFF_W3] = hb;
  } else {
    P.b3[0] = th;
  }
}

size_t mlp_lds_bytes() { return sizeof(MlpLds); }


void launch_mlp_forward(const MlpDev& P, const uint16_t* X, uint64_t N, float* z, hipStream_t s) {
  if (N == 0) return;
  const uint64_t tiles = (N + 31) / 32;
  const uint64_t want = (tiles + MLPI_WAVES - 1) / MLPI_WAVES;
  const uint32_t blocks = (uint32_t)(want < 256 * 2 ? want : 256 * 2);  // 188 VGPRs: two 4-wave blocks per CU
  hipLaunchKernelGGL(k_mlp_infer, dim3(blocks), dim3(64 * MLPI_WAVES), 0, s, P, X, N, z);
}
void launch_mlp_grad(const MlpDev& P, const uint16_t* X, const float* Y, const uint32_t* perm, uint64_t row0,
                     uint64_t nrows, uint64_t N, float pos_weight, float inv_b, float* part, float* loss_part,
                     hipStream_t s) {
  if (nrows == 0) return;
  const uint32_t tiles = (uint32_t)((nrows + 31) / 32);
  hipLaunchKernelGGL(k_mlp_grad, dim3(tiles), dim3(64), sizeof(MlpLds), s, P, X, Y, perm, row0, nrows, N, pos_weight,
                     inv_b, part, loss_part);
}
void launch_mlp_reduce_adam(const MlpDev& P, const float* part, uint32_t ntiles, float* grad_out, int adam, float lr,
                            float b1, float b2, float eps, float bc1, float bc2, hipStream_t s) {
  hipLaunchKernelGGL(k_mlp_reduce_adam, dim3((MLP_NP + 255) / 256), dim3(256), 0, s, P, part, ntiles, grad_out, adam,
                     lr, b1, b2, eps, bc1, bc2);
}
void launch_mlp_publish(const MlpDev& P, hipStream_t s) {
  hipLaunchKernelGGL(k_mlp_publish, dim3((MLP_NP + 255) / 256), dim3(256), 0, s, P);
}

}  // namespace pbw
