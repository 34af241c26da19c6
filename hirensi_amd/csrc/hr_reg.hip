// Gradient post-processing on the regular-domain box: the gradient mass
// (Assem_R_Mg_3D, Assemble_PETSc_3D.F90:670-748, inverted as in
// Control_Problem_Operators.F90:12-24), the TV and TN regularisation vectors
// (ELEMENT_TV_TN_Reg_3D / ELEMENT_R_Mg_3D, Element_Matrices_3D.F90:1555-1718,
// 1899-2075, assembled by Assemble_PETSc_3D.F90:670-860) and the update of the
// misfit gradient of Program_Main.F90:431-455.
//
// All elements of the box share one affine geometry, so at a GLL point q =
// (i, j, k) of an element the derivative of the shape function of local node
// (a, b, c) is dxi_x L_a'(xi_i) [b == j][c == k] (and likewise along y, z):
// the gradient of a nodal field at q is three 3-term sums along the lattice
// lines through q, and the element vector scatters back along the same lines.
//
// Assembly is by element colour: the elements of one colour (ex, ey, ez
// parities) share no node, so one launch per colour adds its element vectors
// into the nodal output without atomics, in a fixed order (deterministic).
#include <hip/hip_runtime.h>

#include "hr_device.h"

namespace hr {
namespace {

// L_a'(xi_q) of the quadratic Lagrange basis on the GLL points (-1, 0, 1)
__device__ __forceinline__ double dl(int q, int a) {
  constexpr double D[3][3] = {{-1.5, 2.0, -0.5}, {-0.5, 0.0, 0.5}, {0.5, -2.0, 1.5}};
  return D[q][a];
}

// Fortran REAL -> default INTEGER as flang compiles it (truncation, saturation
// at the int32 range, NaN -> 0): the reference's TV cost takes its
// `numerator` through an implicitly INTEGER variable.
__device__ __forceinline__ double f90_int4(double x) {
  if (x != x) return 0.0;
  const double t = trunc(x);
  return t < -2147483648.0 ? -2147483648.0 : (t > 2147483647.0 ? 2147483647.0 : t);
}

struct RegArgs {
  Geo g;
  const double* zc;          // lattice z coordinates
  int px, py, pz;            // colour (element parities relative to the box origin)
  int cx, cy, cz;            // elements of this colour per axis
  int weighted;              // Weighted_Regularization
  double w_top, w_bot, w_len;
  const double* m[2];        // lambda, mu (nodal)
  double* r[2];              // regularisation vectors (nodal, accumulated)
  double* cost_el;           // per box element: [2 * e + field]
};

// One thread per element of the colour.  TV: reg_I += sum_q W_q F_q
// (grad N_I . grad m) / sqrt(|grad m|^2 + 0.01), cost += W_q F_q
// int(|grad m|^2) / sqrt(...); TN: reg = ER m, cost = 1/2 m.ER m.
template <bool TV>
__global__ __launch_bounds__(128) void reg_color_kernel(RegArgs A) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= A.cx * A.cy * A.cz) return;
  const Geo& g = A.g;
  const int ia = t % A.cx, ib = (t / A.cx) % A.cy, ic = t / (A.cx * A.cy);
  const int rx = A.px + 2 * ia, ry = A.py + 2 * ib, rz = A.pz + 2 * ic;  // element inside the box
  const int ex = g.rlo[0] + rx, ey = g.rlo[1] + ry, ez = g.rlo[2] + rz;
  const long long NX = g.NX, NXY = (long long)g.NX * g.NY;
  const long long n0 = 2 * ex + NX * (2 * ey) + NXY * (2 * ez);
  const long long re = rx + (long long)(g.rhi[0] - g.rlo[0]) * (ry + (long long)(g.rhi[1] - g.rlo[1]) * rz);
  double fac[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    if (A.weighted) {
      const double z = A.zc[2 * ez + k];
      const double f2 = fabs(z / A.w_len);
      fac[k] = (1.0 - f2) * A.w_top + f2 * A.w_bot;
    } else {
      fac[k] = 1.0;
    }
  }
  const double d0 = g.dxi[0], d1 = g.dxi[1], d2 = g.dxi[2];
  for (int f = 0; f < 2; ++f) {
    const double* __restrict__ M = A.m[f];
    double m[27], r[27];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          m[a + 3 * b + 9 * c] = M[n0 + a + NX * b + NXY * c];
          r[a + 3 * b + 9 * c] = 0.0;
        }
    double cost = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          const int q = i + 3 * j + 9 * k;
          double gx = 0.0, gy = 0.0, gz = 0.0;
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            gx += dl(i, a) * m[a + 3 * j + 9 * k];
            gy += dl(j, a) * m[i + 3 * a + 9 * k];
            gz += dl(k, a) * m[i + 3 * j + 9 * a];
          }
          gx *= d0;
          gy *= d1;
          gz *= d2;
          const double s = (gx * gx + gy * gy) + gz * gz;
          double w = g.wq[q] * fac[k];
          if (TV) {
            const double den = sqrt(s + 0.01);
            cost += (f90_int4(s) / den) * w;
            w = w / den;
          } else {
            cost += s * w;
          }
          const double wx = w * gx * d0, wy = w * gy * d1, wz = w * gz * d2;
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            r[a + 3 * j + 9 * k] += dl(i, a) * wx;
            r[i + 3 * a + 9 * k] += dl(j, a) * wy;
            r[i + 3 * j + 9 * a] += dl(k, a) * wz;
          }
        }
    double* __restrict__ R = A.r[f];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int b = 0; b < 3; ++b)
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          const long long n = n0 + a + NX * b + NXY * c;
          R[n] = R[n] + r[a + 3 * b + 9 * c];
        }
    A.cost_el[2 * re + f] = TV ? cost : 0.5 * cost;
  }
}

// Gradient mass at a lattice node: the GLL weights x det J of the box elements
// that hold it, summed in element order; 1/Mg where Mg > 1e-10 (else 0) and
// the 0/1 restriction.
__global__ __launch_bounds__(256) void gradient_mass_kernel(Geo g, double* __restrict__ img,
                                                            double* __restrict__ restr) {
  const long long NN = g.NN;
  for (long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x; n < NN;
       n += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(n % g.NX), j = (int)((n / g.NX) % g.NY), k = (int)(n / ((long long)g.NX * g.NY));
    const int idx[3] = {i, j, k};
    int e_lo[3], e_hi[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      // elements along axis d holding lattice index idx[d]: (idx-1)/2 .. idx/2
      e_lo[d] = idx[d] >= 1 ? (idx[d] - 1) / 2 : 0;
      e_hi[d] = idx[d] / 2;
      const int ne = d == 0 ? g.nex : (d == 1 ? g.ney : g.nez);
      if (e_hi[d] > ne - 1) e_hi[d] = ne - 1;
      if (e_lo[d] < g.rlo[d]) e_lo[d] = g.rlo[d];
      if (e_hi[d] > g.rhi[d] - 1) e_hi[d] = g.rhi[d] - 1;
    }
    double mg = 0.0;
    for (int ez = e_lo[2]; ez <= e_hi[2]; ++ez)
      for (int ey = e_lo[1]; ey <= e_hi[1]; ++ey)
        for (int ex = e_lo[0]; ex <= e_hi[0]; ++ex)
          mg += g.wq[(i - 2 * ex) + 3 * (j - 2 * ey) + 9 * (k - 2 * ez)];
    const bool on = mg > 1.0e-10;
    img[n] = on ? 1.0 / mg : 0.0;
    if (restr) restr[n] = on ? 1.0 : 0.0;
  }
}

__global__ __launch_bounds__(256) void fill2_kernel(long long n, double* __restrict__ a, double* __restrict__ b) {
  for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x) {
    a[q] = 0.0;
    b[q] = 0.0;
  }
}

// Deterministic dot products: part[4 * blockIdx + v] = partial sums of
// x_v . y_v over a fixed grid-stride split; the tail kernel adds the partials
// in block order.
constexpr int RED_BLOCKS = 256;
constexpr int RED_THREADS = 256;

struct DotArgs {
  int nd;                    // number of dot products (<= 4)
  long long n;
  long long stride;          // element stride inside x and y
  const double* x[4];
  const double* y[4];        // nullptr: plain sum of x
};

__global__ __launch_bounds__(RED_THREADS) void dot_partial_kernel(DotArgs A, double* __restrict__ part) {
  __shared__ double sh[RED_THREADS];
  for (int v = 0; v < A.nd; ++v) {
    double s = 0.0;
    for (long long q = blockIdx.x * (long long)RED_THREADS + threadIdx.x; q < A.n; q += (long long)RED_BLOCKS * RED_THREADS)
      s += A.y[v] ? A.x[v][q * A.stride] * A.y[v][q * A.stride] : A.x[v][q * A.stride];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = RED_THREADS / 2; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x * 4 + v] = sh[0];
    __syncthreads();
  }
}

__global__ __launch_bounds__(64) void dot_tail_kernel(int nd, const double* __restrict__ part, double* __restrict__ out) {
  if (threadIdx.x < nd) {
    double s = 0.0;
    for (int b = 0; b < RED_BLOCKS; ++b) s += part[b * 4 + threadIdx.x];
    out[threadIdx.x] = s;
  }
}

// g <- (g + f * reg) * img  (add = 1), g <- g * img (add = 0)
__global__ __launch_bounds__(256) void reg_update_kernel(long long n, int add, double fl, double fm,
                                                         const double* __restrict__ rl, const double* __restrict__ rm,
                                                         const double* __restrict__ img, double* __restrict__ gl,
                                                         double* __restrict__ gm) {
  for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x) {
    double a = gl[q], b = gm[q];
    if (add) {
      a = a + fl * rl[q];
      b = b + fm * rm[q];
    }
    gl[q] = a * img[q];
    gm[q] = b * img[q];
  }
}

int grid_for(long long n, int threads) {
  long long b = (n + threads - 1) / threads;
  if (b > 8192) b = 8192;
  if (b < 1) b = 1;
  return (int)b;
}

}  // namespace

long long reg_box_elements(const Geo& g) {
  return (long long)(g.rhi[0] - g.rlo[0]) * (g.rhi[1] - g.rlo[1]) * (g.rhi[2] - g.rlo[2]);
}

hipError_t launch_gradient_mass(const Geo& g, double* img, double* restr, hipStream_t s) {
  hipLaunchKernelGGL(gradient_mass_kernel, dim3(grid_for(g.NN, 256)), dim3(256), 0, s, g, img, restr);
  return hipGetLastError();
}

hipError_t launch_regularization(const Geo& g, const double* zc, int tv, const double* wreg, const double* lam,
                                 const double* mu, double* rl, double* rm, double* cost_el, hipStream_t s) {
  hipLaunchKernelGGL(fill2_kernel, dim3(grid_for(g.NN, 256)), dim3(256), 0, s, g.NN, rl, rm);
  RegArgs A{};
  A.g = g;
  A.zc = zc;
  A.weighted = wreg != nullptr;
  if (wreg) {
    A.w_top = wreg[0];
    A.w_bot = wreg[1];
    A.w_len = wreg[2];
  }
  A.m[0] = lam;
  A.m[1] = mu;
  A.r[0] = rl;
  A.r[1] = rm;
  A.cost_el = cost_el;
  const int R[3] = {g.rhi[0] - g.rlo[0], g.rhi[1] - g.rlo[1], g.rhi[2] - g.rlo[2]};
  for (int col = 0; col < 8; ++col) {
    A.px = col & 1;
    A.py = (col >> 1) & 1;
    A.pz = (col >> 2) & 1;
    A.cx = (R[0] - A.px + 1) / 2;
    A.cy = (R[1] - A.py + 1) / 2;
    A.cz = (R[2] - A.pz + 1) / 2;
    const long long ne = (long long)A.cx * A.cy * A.cz;
    if (ne <= 0) continue;
    const int blocks = (int)((ne + 127) / 128);
    if (tv)
      hipLaunchKernelGGL(reg_color_kernel<true>, dim3(blocks), dim3(128), 0, s, A);
    else
      hipLaunchKernelGGL(reg_color_kernel<false>, dim3(blocks), dim3(128), 0, s, A);
  }
  return hipGetLastError();
}

hipError_t launch_dots(int nd, long long n, long long stride, const double* const* x, const double* const* y,
                       double* part, double* out, hipStream_t s) {
  DotArgs A{};
  A.nd = nd;
  A.n = n;
  A.stride = stride;
  for (int v = 0; v < nd; ++v) {
    A.x[v] = x[v];
    A.y[v] = y[v];
  }
  hipLaunchKernelGGL(dot_partial_kernel, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s, A, part);
  hipLaunchKernelGGL(dot_tail_kernel, dim3(1), dim3(64), 0, s, nd, part, out);
  return hipGetLastError();
}

int dot_partials_size() { return RED_BLOCKS * 4; }

hipError_t launch_reg_update(long long n, int add, double fl, double fm, const double* rl, const double* rm,
                             const double* img, double* gl, double* gm, hipStream_t s) {
  hipLaunchKernelGGL(reg_update_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, n, add, fl, fm, rl, rm, img, gl, gm);
  return hipGetLastError();
}

}  // namespace hr
