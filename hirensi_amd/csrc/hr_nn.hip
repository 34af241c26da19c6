// Kernels of the learned gradient preconditioner: a 3D convolutional
// encoder-decoder evaluated and trained on the GPU.  Activations are FP32
// volumes [sample][channel][z][y][x]; convolution weights are BF16
// [cout][cin * 27] (3x3x3 taps, zero padding, tap t = 9 (dz+1) + 3 (dy+1) +
// (dx+1)) with FP32 accumulation on the matrix cores
// (v_mfma_f32_16x16x16_bf16); FP32 master weights for the optimiser.
//
// A convolution is the GEMM out[co][v] = sum_k W[co][k] X[k][v] over the
// columns v = (sample, voxel), with k = (ci, tap) and X[k][v] = in[ci][voxel +
// offset(tap)] gathered on the fly (no im2col buffer).  One wavefront computes
// a 16 (co) x 16 (column) tile; its lanes hold the MFMA operands: A row
// co0 + lane % 16, B column v0 + lane % 16, four K values 4 (lane / 16) .. +3
// each; D column lane % 16, rows 4 (lane / 16) .. +3.  Operands are rounded
// to BF16 (nearest even) as they are gathered.

#include <hip/hip_runtime.h>
#include <cstdint>

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (finite inputs)
__device__ __forceinline__ short to_bf16(float f) {
  unsigned int u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}

struct Vol {
  int D, H, W;   // one sample's volume
  int NB;        // samples
  __device__ __host__ __forceinline__ int n() const { return D * H * W; }
};

// a column's voxel coordinates (fixed for a lane across the K loop)
struct Col {
  int b, x, y, z, v;   // sample, coordinates, voxel index; b < 0: padding column
};

__device__ __forceinline__ Col column(const Vol& g, int c) {
  Col r;
  const int V = g.n();
  if (c >= V * g.NB) { r.b = -1; r.x = r.y = r.z = r.v = 0; return r; }
  r.b = c / V;
  r.v = c - r.b * V;
  r.x = r.v % g.W;
  r.y = (r.v / g.W) % g.H;
  r.z = r.v / (g.W * g.H);
  return r;
}

// X[k][column] for k = ci * 27 + tap of sample c.b, zero outside the volume
__device__ __forceinline__ float gather(const float* __restrict__ in, const Vol& g, int cin, int k, const Col& c) {
  if (c.b < 0 || k >= cin * 27) return 0.0f;
  const int ci = k / 27, t = k - 27 * ci;
  const int dz = t / 9 - 1, dy = (t / 3) % 3 - 1, dx = t % 3 - 1;
  const int xx = c.x + dx, yy = c.y + dy, zz = c.z + dz;
  if (xx < 0 || yy < 0 || zz < 0 || xx >= g.W || yy >= g.H || zz >= g.D) return 0.0f;
  return in[((size_t)c.b * cin + ci) * g.n() + (xx + g.W * (yy + g.H * zz))];
}

// out[b][co][v] = bias[co] + sum_k w[co][k] X[k][(b, v)]   (+ out when accumulate)
__global__ void __launch_bounds__(256) conv3d_kernel(const float* __restrict__ in, const unsigned short* __restrict__ w,
                                                     const float* __restrict__ bias, float* __restrict__ out, int cin,
                                                     int cout, Vol g, int accumulate) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int V = g.n(), K = cin * 27;
  const int v0 = (blockIdx.x * 4 + wave) * 16;
  const int co0 = blockIdx.y * 16;
  if (v0 >= V * g.NB) return;                 // whole wavefront past the last column (uniform)
  const int r = lane & 15, kq = lane >> 4;
  const Col c = column(g, v0 + r);
  const int co_a = co0 + r;
  const unsigned short* __restrict__ wrow = w + (size_t)(co_a < cout ? co_a : 0) * K;
  f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int k0 = 0; k0 < K; k0 += 16) {
    bf16x4 a, b;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int k = k0 + 4 * kq + j;
      a[j] = (short)((co_a < cout && k < K) ? wrow[k] : 0);
      b[j] = to_bf16(gather(in, g, cin, k, c));
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
  if (c.b >= 0) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int co = co0 + 4 * kq + j;
      if (co < cout) {
        const size_t o = ((size_t)c.b * cout + co) * V + c.v;
        float s = acc[j] + (bias ? bias[co] : 0.0f);
        if (accumulate) s += out[o];
        out[o] = s;
      }
    }
  }
}

// Weight gradient over a slice of the columns:
// part[s][co][k] = sum_{columns in slice s} dy[co][col] X[k][col]
__global__ void __launch_bounds__(256) conv3d_wgrad_kernel(const float* __restrict__ in, const float* __restrict__ dy,
                                                           float* __restrict__ part, int cin, int cout, Vol g,
                                                           int cslice) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int V = g.n(), K = cin * 27, NC = V * g.NB;
  const int n0 = (blockIdx.x * 4 + wave) * 16;   // k columns of the tile
  const int co0 = blockIdx.y * 16;
  const int s = blockIdx.z;
  if (n0 >= K) return;
  const int r = lane & 15, kq = lane >> 4;
  const int cb = s * cslice, ce = min(cb + cslice, NC);
  const int co_a = co0 + r;
  f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int u0 = cb; u0 < ce; u0 += 16) {
    bf16x4 a, b;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int u = u0 + 4 * kq + j;
      const Col c = column(g, u < ce ? u : NC);
      a[j] = to_bf16((co_a < cout && c.b >= 0) ? dy[((size_t)c.b * cout + co_a) * V + c.v] : 0.0f);
      b[j] = to_bf16(gather(in, g, cin, n0 + r, c));
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
  const int n = n0 + r;
  if (n < K) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int co = co0 + 4 * kq + j;
      if (co < cout) part[((size_t)s * cout + co) * K + n] = acc[j];
    }
  }
}

// dw[i] = sum_s part[s][i], in slice order
__global__ void reduce_slices_kernel(const float* __restrict__ part, int nslice, int n, float* __restrict__ dw) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float c = 0.0f;
  for (int s = 0; s < nslice; ++s) c += part[(size_t)s * n + i];
  dw[i] = c;
}

// db[co] = sum over samples and voxels of dy[b][co][v]
__global__ void __launch_bounds__(256) bias_grad_kernel(const float* __restrict__ dy, int cout, Vol g,
                                                        float* __restrict__ db) {
  __shared__ float red[256];
  const int co = blockIdx.x, V = g.n();
  float c = 0.0f;
  for (int b = 0; b < g.NB; ++b)
    for (int v = threadIdx.x; v < V; v += 256) c += dy[((size_t)b * cout + co) * V + v];
  red[threadIdx.x] = c;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[co] = red[0];
}

// wt[ci][co * 27 + (26 - t)] = w[co][ci * 27 + t]: the weights of the adjoint
// (input-gradient) convolution, whose taps are the mirrored offsets
__global__ void transpose_weights_kernel(const unsigned short* __restrict__ w, unsigned short* __restrict__ wt, int cin,
                                         int cout) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cin * cout * 27) return;
  const int t = i % 27, ci = (i / 27) % cin, co = i / (27 * cin);
  wt[((size_t)ci * cout + co) * 27 + (26 - t)] = w[i];
}

__global__ void leaky_relu_kernel(float* __restrict__ x, size_t n, float slope) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) {
    const float v = x[i];
    x[i] = v > 0.0f ? v : slope * v;
  }
}

// dy *= f'(x) for y = f(x) the activation output (y > 0 exactly when x > 0)
__global__ void leaky_relu_grad_kernel(const float* __restrict__ y, float* __restrict__ dy, size_t n, float slope) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) dy[i] = y[i] > 0.0f ? dy[i] : slope * dy[i];
}

// 2x2x2 average pooling of C volumes (D, H, W) (even sizes)
__global__ void avgpool2_kernel(const float* __restrict__ in, float* __restrict__ out, long long C, int D, int H,
                                int W) {
  const int Ho = H / 2, Wo = W / 2, Vo = (D / 2) * Ho * Wo;
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= (size_t)C * Vo) return;
  const size_t c = i / Vo;
  const int v = (int)(i - c * Vo);
  const int x = v % Wo, y = (v / Wo) % Ho, z = v / (Wo * Ho);
  const float* __restrict__ p = in + c * (size_t)(D * H * W);
  float s = 0.0f;
  for (int d = 0; d < 8; ++d) s += p[(2 * x + (d & 1)) + W * ((2 * y + ((d >> 1) & 1)) + H * (2 * z + (d >> 2)))];
  out[i] = 0.125f * s;
}

// pooling backward: dx[fine] = dy[coarse] / 8 (added to dx when accumulate)
__global__ void avgpool2_grad_kernel(const float* __restrict__ dy, float* __restrict__ dx, long long C, int D, int H,
                                     int W, int accumulate) {
  const int V = D * H * W, Ho = H / 2, Wo = W / 2, Vo = V / 8;
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= (size_t)C * V) return;
  const size_t c = i / V;
  const int v = (int)(i - c * V);
  const int x = v % W, y = (v / W) % H, z = v / (W * H);
  const float d = 0.125f * dy[c * Vo + (x / 2) + Wo * ((y / 2) + Ho * (z / 2))];
  dx[i] = accumulate ? dx[i] + d : d;
}

// nearest 2x upsampling of C coarse volumes to (D, H, W)
__global__ void upsample2_kernel(const float* __restrict__ in, float* __restrict__ out, long long C, int D, int H,
                                 int W) {
  const int V = D * H * W, Ho = H / 2, Wo = W / 2, Vo = V / 8;
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= (size_t)C * V) return;
  const size_t c = i / V;
  const int v = (int)(i - c * V);
  const int x = v % W, y = (v / W) % H, z = v / (W * H);
  out[i] = in[c * Vo + (x / 2) + Wo * ((y / 2) + Ho * (z / 2))];
}

// upsampling backward: dcoarse = sum of the 8 fine values
__global__ void upsample2_grad_kernel(const float* __restrict__ dy, float* __restrict__ dx, long long C, int D, int H,
                                      int W) {
  const int Ho = H / 2, Wo = W / 2, Vo = (D / 2) * Ho * Wo;
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= (size_t)C * Vo) return;
  const size_t c = i / Vo;
  const int v = (int)(i - c * Vo);
  const int x = v % Wo, y = (v / Wo) % Ho, z = v / (Wo * Ho);
  const float* __restrict__ p = dy + c * (size_t)(D * H * W);
  float s = 0.0f;
  for (int d = 0; d < 8; ++d) s += p[(2 * x + (d & 1)) + W * ((2 * y + ((d >> 1) & 1)) + H * (2 * z + (d >> 2)))];
  dx[i] = s;
}

constexpr int MSE_BLOCKS = 256;

// squared-error partial sums per block; grad = 2 (y - t) / n
__global__ void __launch_bounds__(256) mse_kernel(const float* __restrict__ y, const float* __restrict__ t, size_t n,
                                                  float* __restrict__ grad, float* __restrict__ part) {
  __shared__ float red[256];
  float c = 0.0f;
  for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n; i += (size_t)256 * gridDim.x) {
    const float d = y[i] - t[i];
    if (grad) grad[i] = 2.0f * d / (float)n;
    c += d * d;
  }
  red[threadIdx.x] = c;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// loss = sum of the block partials (in block order) / n
__global__ void loss_finish_kernel(const float* __restrict__ part, int nb, size_t n, float* __restrict__ loss) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    float c = 0.0f;
    for (int b = 0; b < nb; ++b) c += part[b];
    loss[0] = c / (float)n;
  }
}

// Adam step on FP32 master weights, then their BF16 copy (wb null: biases)
__global__ void adam_kernel(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, unsigned short* __restrict__ wb, size_t n, float lr, float b1,
                            float b2, float eps, float c1, float c2) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  if (wb) wb[i] = (unsigned short)to_bf16(wi);
}

__global__ void to_bf16_kernel(const float* __restrict__ w, unsigned short* __restrict__ wb, size_t n) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) wb[i] = (unsigned short)to_bf16(w[i]);
}

inline unsigned nblocks(size_t n, unsigned b = 256) { return (unsigned)((n + b - 1) / b); }

bool bad_vol(int D, int H, int W, int NB) { return D < 1 || H < 1 || W < 1 || NB < 1; }
bool odd(int D, int H, int W) { return (D | H | W) & 1; }

}  // namespace

// ------------------------------------------------------------------ C ABI
// Every entry point queues its kernels on `stream` (a hipStream_t, 0 for the
// null stream) and returns 0 or a nonzero error code; pointers are device
// pointers.  Volumes are NB samples of (D, H, W), x fastest.
extern "C" {

int hr_nn_conv3d(const float* in, const unsigned short* w, const float* bias, float* out, int cin, int cout, int NB,
                 int D, int H, int W, int accumulate, void* stream) {
  if (!in || !w || !out || cin < 1 || cout < 1 || bad_vol(D, H, W, NB)) return 1;
  const Vol g{D, H, W, NB};
  const long long ncol = (long long)g.n() * NB;
  if (ncol >= (1ll << 31) - 64) return 1;
  const dim3 grid((unsigned)((ncol + 63) / 64), (unsigned)((cout + 15) / 16));
  hipLaunchKernelGGL(conv3d_kernel, grid, dim3(256), 0, (hipStream_t)stream, in, w, bias, out, cin, cout, g,
                     accumulate);
  return (int)hipGetLastError();
}

// part: nslice * cout * cin * 27 floats of scratch; db may be null
int hr_nn_conv3d_wgrad(const float* in, const float* dy, float* part, int nslice, float* dw, float* db, int cin,
                       int cout, int NB, int D, int H, int W, void* stream) {
  if (!in || !dy || !part || !dw || nslice < 1 || cin < 1 || cout < 1 || bad_vol(D, H, W, NB)) return 1;
  const Vol g{D, H, W, NB};
  const long long ncol = (long long)g.n() * NB;
  if (ncol >= (1ll << 31) - 64) return 1;
  const int K = cin * 27;
  const int cslice = (int)(((ncol + nslice - 1) / nslice + 15) / 16 * 16);
  const int ns = (int)((ncol + cslice - 1) / cslice);
  const dim3 grid((unsigned)((K + 63) / 64), (unsigned)((cout + 15) / 16), (unsigned)ns);
  hipLaunchKernelGGL(conv3d_wgrad_kernel, grid, dim3(256), 0, (hipStream_t)stream, in, dy, part, cin, cout, g, cslice);
  hipLaunchKernelGGL(reduce_slices_kernel, dim3(nblocks((size_t)cout * K)), dim3(256), 0, (hipStream_t)stream, part,
                     ns, cout * K, dw);
  if (db) hipLaunchKernelGGL(bias_grad_kernel, dim3(cout), dim3(256), 0, (hipStream_t)stream, dy, cout, g, db);
  return (int)hipGetLastError();
}

int hr_nn_transpose_weights(const unsigned short* w, unsigned short* wt, int cin, int cout, void* stream) {
  if (!w || !wt || cin < 1 || cout < 1) return 1;
  hipLaunchKernelGGL(transpose_weights_kernel, dim3(nblocks((size_t)cin * cout * 27)), dim3(256), 0,
                     (hipStream_t)stream, w, wt, cin, cout);
  return (int)hipGetLastError();
}

int hr_nn_leaky_relu(float* x, long long n, float slope, void* stream) {
  if (!x || n < 0) return 1;
  if (n == 0) return 0;
  hipLaunchKernelGGL(leaky_relu_kernel, dim3(nblocks((size_t)n)), dim3(256), 0, (hipStream_t)stream, x, (size_t)n,
                     slope);
  return (int)hipGetLastError();
}

int hr_nn_leaky_relu_grad(const float* y, float* dy, long long n, float slope, void* stream) {
  if (!y || !dy || n < 0) return 1;
  if (n == 0) return 0;
  hipLaunchKernelGGL(leaky_relu_grad_kernel, dim3(nblocks((size_t)n)), dim3(256), 0, (hipStream_t)stream, y, dy,
                     (size_t)n, slope);
  return (int)hipGetLastError();
}

// C = samples x channels volumes of the fine size (D, H, W)
int hr_nn_avgpool2(const float* in, float* out, long long C, int D, int H, int W, void* stream) {
  if (!in || !out || C < 1 || bad_vol(D, H, W, 1) || odd(D, H, W)) return 1;
  hipLaunchKernelGGL(avgpool2_kernel, dim3(nblocks((size_t)C * D * H * W / 8)), dim3(256), 0, (hipStream_t)stream, in,
                     out, C, D, H, W);
  return (int)hipGetLastError();
}

int hr_nn_avgpool2_grad(const float* dy, float* dx, long long C, int D, int H, int W, int accumulate, void* stream) {
  if (!dy || !dx || C < 1 || bad_vol(D, H, W, 1) || odd(D, H, W)) return 1;
  hipLaunchKernelGGL(avgpool2_grad_kernel, dim3(nblocks((size_t)C * D * H * W)), dim3(256), 0, (hipStream_t)stream,
                     dy, dx, C, D, H, W, accumulate);
  return (int)hipGetLastError();
}

int hr_nn_upsample2(const float* in, float* out, long long C, int D, int H, int W, void* stream) {
  if (!in || !out || C < 1 || bad_vol(D, H, W, 1) || odd(D, H, W)) return 1;
  hipLaunchKernelGGL(upsample2_kernel, dim3(nblocks((size_t)C * D * H * W)), dim3(256), 0, (hipStream_t)stream, in,
                     out, C, D, H, W);
  return (int)hipGetLastError();
}

int hr_nn_upsample2_grad(const float* dy, float* dx, long long C, int D, int H, int W, void* stream) {
  if (!dy || !dx || C < 1 || bad_vol(D, H, W, 1) || odd(D, H, W)) return 1;
  hipLaunchKernelGGL(upsample2_grad_kernel, dim3(nblocks((size_t)C * D * H * W / 8)), dim3(256), 0,
                     (hipStream_t)stream, dy, dx, C, D, H, W);
  return (int)hipGetLastError();
}

// loss[0] = mean (y - t)^2 over n; grad (may be null) = 2 (y - t) / n; part: 256 floats of scratch
int hr_nn_mse(const float* y, const float* t, long long n, float* grad, float* part, float* loss, void* stream) {
  if (!y || !t || !part || !loss || n < 1) return 1;
  hipLaunchKernelGGL(mse_kernel, dim3(MSE_BLOCKS), dim3(256), 0, (hipStream_t)stream, y, t, (size_t)n, grad, part);
  hipLaunchKernelGGL(loss_finish_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, part, MSE_BLOCKS, (size_t)n, loss);
  return (int)hipGetLastError();
}

// step >= 1: bias corrections 1 - b1^step, 1 - b2^step; wb (may be null) gets the BF16 copy
int hr_nn_adam(float* w, const float* g, float* m, float* v, unsigned short* wb, long long n, float lr, float b1,
               float b2, float eps, int step, void* stream) {
  if (!w || !g || !m || !v || n < 1 || step < 1) return 1;
  const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
  hipLaunchKernelGGL(adam_kernel, dim3(nblocks((size_t)n)), dim3(256), 0, (hipStream_t)stream, w, g, m, v, wb,
                     (size_t)n, lr, b1, b2, eps, c1, c2);
  return (int)hipGetLastError();
}

int hr_nn_to_bf16(const float* w, unsigned short* wb, long long n, void* stream) {
  if (!w || !wb || n < 1) return 1;
  hipLaunchKernelGGL(to_bf16_kernel, dim3(nblocks((size_t)n)), dim3(256), 0, (hipStream_t)stream, w, wb, (size_t)n);
  return (int)hipGetLastError();
}

}  // extern "C"
