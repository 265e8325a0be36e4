// Shift predictor: a small tanh MLP trained and evaluated on the GPU.
//
// Network: nin inputs -> `depth` hidden layers of `width` tanh units -> nout
// linear outputs.  Master weights, inputs and outputs are FP32; every matrix
// product runs on the matrix cores with BF16 operands (round to nearest even)
// and FP32 accumulation (v_mfma_f32_16x16x16bf16_1k).  Inputs are mapped
// x' = (x - in_shift) * in_scale and outputs y = y' / out_scale + out_shift,
// so training sees O(1) numbers.
//
// Flat parameter layout (FP32), dimensions padded to multiples of 16
// (nin_p = pad16(nin), nout_p = pad16(nout); padded rows/columns stay zero):
//   W_0 [width x nin_p], b_0 [width],
//   W_l [width x width], b_l [width]            l = 1 .. depth-1
//   W_o [nout_p x width], b_o [nout_p]
// W is row-major by output unit: W[o * in + i].
#include "common.h"

#define MLP_T 64   // one wave per 16 x 16 output tile

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ inline short to_bf16(float x) {
  unsigned u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);        // round to nearest even (finite inputs)
  return (short)(u >> 16);
}

enum { EPI_NONE = 0, EPI_BIAS = 1, EPI_BIAS_TANH = 2, EPI_DTANH = 3 };

// C[m][n] = sum_k A(m,k) B(k,n) with A(m,k) = A[m sam + k sak],
// B(k,n) = B[k sbk + n sbn], C row-major (ldc); M, N, K multiples of 16.
// Epilogues: + bias[n]; tanh(. + bias[n]); . * (1 - aux[m ldc + n]^2).
// grid: (N/16, M/16), block: one wave.  Lane l supplies A[m0 + l%16][k0 +
// 4(l/16) + j] and B[k0 + 4(l/16) + j][n0 + l%16], j = 0..3; accumulator
// register i holds C[m0 + 4(l/16) + i][n0 + l%16].
template <int EPI>
__global__ __launch_bounds__(MLP_T) void k_mlp_gemm(int K, const float *__restrict__ A, int sam, int sak,
                                                    const float *__restrict__ B, int sbk, int sbn,
                                                    float *__restrict__ C, int ldc, const float *__restrict__ bias,
                                                    const float *__restrict__ aux) {
  const int l = threadIdx.x, m0 = blockIdx.y * 16, n0 = blockIdx.x * 16;
  const int r = l & 15, q = 4 * (l >> 4);
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 16) {
    bf16x4 a, b;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      a[j] = to_bf16(A[(size_t)(m0 + r) * sam + (size_t)(k0 + q + j) * sak]);
      b[j] = to_bf16(B[(size_t)(k0 + q + j) * sbk + (size_t)(n0 + r) * sbn]);
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int m = m0 + q + i, n = n0 + r;
    float v = acc[i];
    if (EPI == EPI_BIAS || EPI == EPI_BIAS_TANH) v += bias[n];
    if (EPI == EPI_BIAS_TANH) v = tanhf(v);
    if (EPI == EPI_DTANH) {
      const float a = aux[(size_t)m * ldc + n];
      v *= 1.f - a * a;
    }
    C[(size_t)m * ldc + n] = v;
  }
}

// db[n] = sum_m dY[m ld + n].  One thread per n.
__global__ __launch_bounds__(256) void k_mlp_colsum(int M, int N, const float *__restrict__ dY, int ld,
                                                    float *__restrict__ db) {
  const int n = blockIdx.x * 256 + threadIdx.x;
  if (n >= N) return;
  float s = 0.f;
  for (int m = 0; m < M; ++m) s += dY[(size_t)m * ld + n];
  db[n] = s;
}

// Mean squared error over the nout real outputs and its gradient:
// dY = 2 (Y - T) / (B nout) (zero in the padded columns).  One block.
__global__ __launch_bounds__(256) void k_mlp_loss(int B, int nout, int nout_p, const float *__restrict__ Y,
                                                  const float *__restrict__ T, float *__restrict__ dY,
                                                  float *__restrict__ loss, int hist_slot) {
  __shared__ float red[256];
  float s = 0.f;
  const float sc = 2.f / (float)(B * nout);
  for (int e = threadIdx.x; e < B * nout_p; e += 256) {
    const int b = e / nout_p, o = e - b * nout_p;
    float g = 0.f;
    if (o < nout) {
      const float d = Y[e] - T[(size_t)b * nout + o];
      s += d * d;
      g = sc * d;
    }
    dY[e] = g;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const float v = red[0] / (float)(B * nout);
    loss[0] = v;
    if (hist_slot >= 0) loss[1 + hist_slot] = v;
  }
}

// Adam (Kingma & Ba 2015) on the flat parameters, step t >= 1.
__global__ __launch_bounds__(256) void k_mlp_adam(int n, float *__restrict__ p, const float *__restrict__ g,
                                                  float *__restrict__ m, float *__restrict__ v, float lr, float b1,
                                                  float b2, float eps, int t) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float mh = mi / (1.f - powf(b1, (float)t)), vh = vi / (1.f - powf(b2, (float)t));
  p[i] -= lr * mh / (sqrtf(vh) + eps);
}

__device__ inline unsigned mlp_hash(unsigned long long seed, unsigned a, unsigned b) {
  unsigned long long h = seed ^ (0x9E3779B97F4A7C15ull * (a + 1)) ^ (0xC2B2AE3D27D4EB4Full * (b + 7));
  h ^= h >> 33; h *= 0xff51afd7ed558ccdull; h ^= h >> 33; h *= 0xc4ceb9fe1a85ec53ull; h ^= h >> 33;
  return (unsigned)h;
}

// Mini-batch: rows idx(b) of the normalised data set (padded input rows of
// nin_p, target rows of nout).  idx(b) = b when B == n (the whole set, in
// order), else a hashed draw with replacement from (seed, step, b).
__global__ __launch_bounds__(256) void k_mlp_gather(int B, int n, int nin_p, int nout, const float *__restrict__ X,
                                                    const float *__restrict__ T, float *__restrict__ Xb,
                                                    float *__restrict__ Tb, unsigned long long seed, int step) {
  const int b = blockIdx.x;
  const int src = (B == n) ? b : (int)(mlp_hash(seed, (unsigned)step, (unsigned)b) % (unsigned)n);
  for (int i = threadIdx.x; i < nin_p; i += 256) Xb[(size_t)b * nin_p + i] = X[(size_t)src * nin_p + i];
  for (int o = threadIdx.x; o < nout; o += 256) Tb[(size_t)b * nout + o] = T[(size_t)src * nout + o];
}

// x' = (x - shift) * scale into rows of nin_p (zeros beyond nin); likewise
// for targets when T != nullptr.  grid: rows.
__global__ __launch_bounds__(64) void k_mlp_normalize(int nin, int nin_p, int nout, const float *__restrict__ x,
                                                      const float *__restrict__ t, const float *__restrict__ scl,
                                                      float *__restrict__ xo, float *__restrict__ to) {
  const int b = blockIdx.x, i = threadIdx.x;
  const float *in_shift = scl, *in_scale = scl + nin_p, *out_shift = scl + 2 * nin_p;
  const float *out_scale = out_shift + 16;
  if (i < nin_p) xo[(size_t)b * nin_p + i] = (i < nin) ? (x[(size_t)b * nin + i] - in_shift[i]) * in_scale[i] : 0.f;
  if (t && i < nout) to[(size_t)b * nout + i] = (t[(size_t)b * nout + i] - out_shift[i]) * out_scale[i];
}

// y = y' / out_scale + out_shift for the nout real outputs.  grid: rows.
__global__ __launch_bounds__(64) void k_mlp_denormalize(int nout, int nout_p, int nin_p, const float *__restrict__ Y,
                                                        const float *__restrict__ scl, float *__restrict__ y) {
  const int b = blockIdx.x, o = threadIdx.x;
  const float *out_shift = scl + 2 * nin_p, *out_scale = out_shift + 16;
  if (o < nout) y[(size_t)b * nout + o] = Y[(size_t)b * nout_p + o] / out_scale[o] + out_shift[o];
}

// Glorot-uniform weights from a hash of (seed, index), zero biases and zero
// padding.  One thread per parameter.
__global__ __launch_bounds__(256) void k_mlp_init(int np, float *__restrict__ p, const int *__restrict__ meta,
                                                  int nseg, unsigned long long seed) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= np) return;
  // meta: per segment {offset, rows, cols, real_rows, real_cols, is_bias}
  float v = 0.f;
  for (int s = 0; s < nseg; ++s) {
    const int *d = meta + 6 * s;
    const int off = d[0], rows = d[1], cols = d[2];
    if (i < off || i >= off + rows * cols) continue;
    const int e = i - off, rr = e / cols, cc = e - rr * cols;
    if (!d[5] && rr < d[3] && cc < d[4]) {
      const float a = sqrtf(6.f / (float)(d[3] + d[4]));
      const float u = (float)(mlp_hash(seed, (unsigned)i, 0x51u) >> 8) * (1.f / 16777216.f);
      v = (2.f * u - 1.f) * a;
    }
  }
  p[i] = v;
}
