// Kernels for training the Newton initial-guess MLP on the GPU (hipRTC-compiled
// once, independent of the PDE).  Matrices are row-major fp32 in memory with
// samples along columns ("feature-major"); products run on the matrix cores as
// BF16 x BF16 -> FP32 (v_mfma_f32_32x32x8bf16_1k).
//
// MFMA 32x32x8 operand layout (64 lanes, l = lane):
//   A (32 x 8):  a[i] = A[l % 32][4 (l / 32) + i]          i = 0..3
//   B (8 x 32):  b[i] = B[4 (l / 32) + i][l % 32]
//   C (32 x 32): c[j] = C[8 (j / 4) + 4 (l / 32) + j % 4][l % 32]   j = 0..15
#pragma once

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ short to_bf16(float x) {
  // round to nearest even (inputs here are finite)
  unsigned int b = __float_as_uint(x);
  b += 0x7FFFu + ((b >> 16) & 1u);
  return (short)(b >> 16);
}

__device__ __forceinline__ float bf16_to_f(short h) {
  return __uint_as_float(((unsigned int)(unsigned short)h) << 16);
}

// C = alpha op(A) op(B) + beta C;  op(A): M x K, op(B): K x N.
// A(m, k) = transA ? A[k * lda + m] : A[m * lda + k];  likewise B.
// Workgroup of 4 waves computes a 64 x 64 tile; K in chunks of 32 via LDS.
extern "C" __global__ void __launch_bounds__(256)
k_gemm_bf16(const float* __restrict__ A, int lda, int transA, const float* __restrict__ B, int ldb,
            int transB, float* __restrict__ C, int ldc, int M, int N, int K, float alpha,
            float beta) {
  __shared__ short As[64][32 + 4];
  __shared__ short Bs[32][64 + 4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  f32x16 acc;
#pragma unroll
  for (int j = 0; j < 16; j++) acc[j] = 0.0f;
  for (int k0 = 0; k0 < K; k0 += 32) {
    for (int e = tid; e < 64 * 32; e += 256) {
      int r = e / 32, k = e % 32;
      int gm = m0 + r, gk = k0 + k;
      float a = 0.0f;
      if (gm < M && gk < K) a = transA ? A[(long)gk * lda + gm] : A[(long)gm * lda + gk];
      As[r][k] = to_bf16(a);
    }
    for (int e = tid; e < 32 * 64; e += 256) {
      int k = e / 64, c = e % 64;
      int gk = k0 + k, gn = n0 + c;
      float b = 0.0f;
      if (gk < K && gn < N) b = transB ? B[(long)gn * ldb + gk] : B[(long)gk * ldb + gn];
      Bs[k][c] = to_bf16(b);
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 4; kk++) {
      bf16x4 a, b;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        a[i] = As[wr + (lane & 31)][kk * 8 + 4 * (lane >> 5) + i];
        b[i] = Bs[kk * 8 + 4 * (lane >> 5) + i][wc + (lane & 31)];
      }
      acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < 16; j++) {
    int r = m0 + wr + 8 * (j / 4) + 4 * (lane >> 5) + (j % 4);
    int c = n0 + wc + (lane & 31);
    if (r < M && c < N) {
      float v = alpha * acc[j];
      if (beta != 0.0f) v += beta * C[(long)r * ldc + c];
      C[(long)r * ldc + c] = v;
    }
  }
}

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float silu_grad(float z) {
  float s = 1.0f / (1.0f + __expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

// Z[r][c] += b[r]; H = act(Z) (act 1: SiLU, 0: identity).  H may alias Z only
// when act == 0.
extern "C" __global__ void k_bias_act(float* __restrict__ Z, const float* __restrict__ b,
                                      float* __restrict__ H, int rows, int cols, int act) {
  long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= (long)rows * cols) return;
  int r = (int)(i / cols);
  float z = Z[i] + b[r];
  Z[i] = z;
  H[i] = act ? silu(z) : z;
}

// dZ = dH * SiLU'(Z)
extern "C" __global__ void k_act_grad(float* __restrict__ dZ, const float* __restrict__ dH,
                                      const float* __restrict__ Z, long n) {
  long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  dZ[i] = dH[i] * silu_grad(Z[i]);
}

// db[r] = sum_c dZ[r][c]  (one workgroup per row)
extern "C" __global__ void __launch_bounds__(256)
k_rowsum(const float* __restrict__ dZ, float* __restrict__ db, int rows, int cols) {
  __shared__ float red[256];
  int r = blockIdx.x;
  float s = 0.0f;
  for (int c = threadIdx.x; c < cols; c += 256) s += dZ[(long)r * cols + c];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[r] = red[0];
}

// mean-squared error over rows x cols: dY = 2 (Y - T) / (rows cols), loss
// accumulated (sum of squares / (rows cols)) into *loss.
extern "C" __global__ void __launch_bounds__(256)
k_mse(const float* __restrict__ Y, const float* __restrict__ T, float* __restrict__ dY,
      float* __restrict__ loss, int rows, int cols) {
  __shared__ float red[256];
  long n = (long)rows * cols;
  long i = (long)blockIdx.x * 256 + threadIdx.x;
  float e = 0.0f;
  if (i < n) {
    float d = Y[i] - T[i];
    dY[i] = 2.0f * d / (float)n;
    e = d * d / (float)n;
  }
  red[threadIdx.x] = e;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, red[0]);
}

// Adam step (Kingma & Ba) on fp32 master weights.
extern "C" __global__ void k_adam(float* __restrict__ w, const float* __restrict__ g,
                                  float* __restrict__ m, float* __restrict__ v, long n, float lr,
                                  float b1, float b2, float eps, float c1, float c2) {
  long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  float gi = g[i];
  float mi = b1 * m[i] + (1.0f - b1) * gi;
  float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}
