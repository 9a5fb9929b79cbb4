// The potential surrogate: an MLP of 6 symmetry-adapted features -> 128 -> 128 -> 128 -> 1 with SiLU, trained and
// evaluated on the device. The two 128 x 128 hidden layers run on the matrix cores (v_mfma_f32_16x16x16_bf16:
// operands rounded to bf16, round to nearest even, products accumulated in FP32) in the forward pass and in both
// backward products (dW = A^T dZ, dA = dZ W^T); the 6 -> 128 input and 128 -> 1 output layers, the SiLU, the loss
// and Adam run in FP32 on the vector units. Rows are processed in tiles of 16 (the row count is a multiple of 16).
#pragma once

#include <hip/hip_runtime.h>

namespace spo_dev {

constexpr int kMlpIn = 6;
constexpr int kMlpWidth = 128;
// Offsets of the parameter vector: W1 (6 x 128, row k), b1, W2 (128 x 128), b2, W3, b3, W4 (128), b4
constexpr int kMlpW1 = 0;
constexpr int kMlpB1 = kMlpW1 + kMlpIn * kMlpWidth;
constexpr int kMlpW2 = kMlpB1 + kMlpWidth;
constexpr int kMlpB2 = kMlpW2 + kMlpWidth * kMlpWidth;
constexpr int kMlpW3 = kMlpB2 + kMlpWidth;
constexpr int kMlpB3 = kMlpW3 + kMlpWidth * kMlpWidth;
constexpr int kMlpW4 = kMlpB3 + kMlpWidth;
constexpr int kMlpB4 = kMlpW4 + kMlpWidth;
constexpr int kMlpParams = kMlpB4 + 1;

typedef short mlp_bf16x4 __attribute__((ext_vector_type(4)));
typedef float mlp_f32x4 __attribute__((ext_vector_type(4)));

__device__ inline short mlp_bf16(float f) {  // round to nearest even (finite inputs)
  unsigned u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return static_cast<short>(u >> 16);
}

__device__ inline float mlp_sigmoid(float z) { return 1.0f / (1.0f + __expf(-z)); }
__device__ inline float mlp_silu(float z) { return z * mlp_sigmoid(z); }
__device__ inline float mlp_silu_slope(float z) {
  const float s = mlp_sigmoid(z);
  return s * (1.0f + z * (1.0f - s));
}

// C(i, j) = sum_k A(i, k) B(k, j) over 16 x 16 tiles, one wavefront per tile; A(i, k) = a[i * a_i + k * a_k],
// B(k, j) = b[k * b_k + j * b_j]; k runs over kdim (a multiple of 16). Lane l feeds A(i0 + l % 16, k + 4 (l / 16) + c)
// and B(k + 4 (l / 16) + c, j0 + l % 16), c < 4, and receives C(i0 + 4 (l / 16) + c, j0 + l % 16).
// mode 0: c_out = C; mode 1: z = C + bias[j] to c_out, SiLU(z) to act_out.
__global__ __launch_bounds__(64) void mlp_gemm_kernel(const float* a, long a_i, long a_k, const float* b, long b_k,
                                                      long b_j, int kdim, int tiles_j, float* c_out, long c_ld,
                                                      int mode, const float* bias, float* act_out) {
  const int lane = threadIdx.x;
  const int ti = blockIdx.x / tiles_j, tj = blockIdx.x % tiles_j;
  const long i_row = 16L * ti + (lane & 15), j_col = 16L * tj + (lane & 15);
  const int kq = 4 * (lane >> 4);
  mlp_f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int k0 = 0; k0 < kdim; k0 += 16) {
    mlp_bf16x4 av, bv;
    for (int c = 0; c < 4; ++c) {
      const long k = k0 + kq + c;
      av[c] = mlp_bf16(a[i_row * a_i + k * a_k]);
      bv[c] = mlp_bf16(b[k * b_k + j_col * b_j]);
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
  }
  for (int c = 0; c < 4; ++c) {
    const long i = 16L * ti + kq + c;
    if (mode == 0) {
      c_out[i * c_ld + j_col] = acc[c];
    } else {
      const float z = acc[c] + bias[j_col];
      c_out[i * c_ld + j_col] = z;
      act_out[i * c_ld + j_col] = mlp_silu(z);
    }
  }
}

// Input layer: z1 = x W1 + b1, a1 = SiLU(z1); one thread per (row, unit)
__global__ __launch_bounds__(256) void mlp_input_kernel(const float* x, const float* p, long rows, float* z, float* act) {
  const long t = static_cast<long>(blockIdx.x) * 256 + threadIdx.x;
  if (t >= rows * kMlpWidth) return;
  const long row = t / kMlpWidth;
  const int n = static_cast<int>(t % kMlpWidth);
  float s = p[kMlpB1 + n];
  for (int k = 0; k < kMlpIn; ++k) s += x[row * kMlpIn + k] * p[kMlpW1 + k * kMlpWidth + n];
  z[t] = s;
  act[t] = mlp_silu(s);
}

// Output layer: y = a3 W4 + b4; one thread per row
__global__ __launch_bounds__(256) void mlp_output_kernel(const float* act, const float* p, long rows, float* y) {
  const long row = static_cast<long>(blockIdx.x) * 256 + threadIdx.x;
  if (row >= rows) return;
  float s = p[kMlpB4];
  for (int n = 0; n < kMlpWidth; ++n) s += act[row * kMlpWidth + n] * p[kMlpW4 + n];
  y[row] = s;
}

// Mean squared error and dL/dy = 2 (y - t) / rows; one block of 256 threads. loss may be NULL.
__global__ __launch_bounds__(256) void mlp_loss_kernel(const float* y, const float* t, long rows, float* dy,
                                                       float* loss) {
  __shared__ float part[256];
  float s = 0.0f;
  for (long r = threadIdx.x; r < rows; r += 256) {
    const float d = y[r] - t[r];
    dy[r] = 2.0f * d / static_cast<float>(rows);
    s += d * d;
  }
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0 && loss) *loss = part[0] / static_cast<float>(rows);
}

// Output layer backward: dA3 = dy W4^T, then dZ3 = dA3 SiLU'(z3) (fused); one thread per (row, unit)
__global__ __launch_bounds__(256) void mlp_output_back_kernel(const float* dy, const float* p, const float* z3,
                                                              long rows, float* dz3) {
  const long t = static_cast<long>(blockIdx.x) * 256 + threadIdx.x;
  if (t >= rows * kMlpWidth) return;
  const long row = t / kMlpWidth;
  const int n = static_cast<int>(t % kMlpWidth);
  dz3[t] = dy[row] * p[kMlpW4 + n] * mlp_silu_slope(z3[t]);
}

// dZ = dA SiLU'(z), elementwise
__global__ __launch_bounds__(256) void mlp_silu_back_kernel(const float* da, const float* z, long count, float* dz) {
  const long t = static_cast<long>(blockIdx.x) * 256 + threadIdx.x;
  if (t < count) dz[t] = da[t] * mlp_silu_slope(z[t]);
}

// Column sums over rows: out[j] = sum_r m[r * ld + j] * (w ? w[r] : 1); one block of 256 threads per column j
__global__ __launch_bounds__(256) void mlp_column_sum_kernel(const float* m, long ld, const float* w, long rows,
                                                             float* out) {
  __shared__ float part[256];
  const int j = blockIdx.x;
  float s = 0.0f;
  for (long r = threadIdx.x; r < rows; r += 256) s += m[r * ld + j] * (w ? w[r] : 1.0f);
  part[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) part[threadIdx.x] += part[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[j] = part[0];
}

// Input layer weights: g[W1 (k, n)] = sum_r x[r, k] dz1[r, n]; one block per (k, n)
__global__ __launch_bounds__(256) void mlp_input_grad_kernel(const float* x, const float* dz1, long rows, float* g) {
  __shared__ float part[256];
  const int k = blockIdx.x / kMlpWidth, n = blockIdx.x % kMlpWidth;
  float s = 0.0f;
  for (long r = threadIdx.x; r < rows; r += 256) s += x[r * kMlpIn + k] * dz1[r * kMlpWidth + n];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) g[kMlpW1 + k * kMlpWidth + n] = part[0];
}

// Adam (bias-corrected), step t >= 1
__global__ __launch_bounds__(256) void mlp_adam_kernel(float* p, const float* g, float* m, float* v, int count,
                                                       float lr, float beta1, float beta2, float eps, float c1,
                                                       float c2) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= count) return;
  const float gi = g[i];
  m[i] = beta1 * m[i] + (1.0f - beta1) * gi;
  v[i] = beta2 * v[i] + (1.0f - beta2) * gi * gi;
  p[i] -= lr * (m[i] / c1) / (sqrtf(v[i] / c2) + eps);
}

}  // namespace spo_dev
