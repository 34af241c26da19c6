// Small HIP kernels around the stage kernel (hr_stage3.hip): reference
// numbering <-> lattice slots for the operator apply, receiver traces,
// regular-domain history slices, the per-step lambda/mu gradient, misfit.
//
// State layout: per state vector, the displacement block [3 * n + c]
// (c = u_x, u_y, u_z) followed by the stress block [SB + 6 * n + m]
// (m = s_xx s_yy s_zz s_xy s_xz s_yz; SB = 3 NN rounded up to an even count,
// so that a node's stress components are 16-byte aligned),
// n = i + NX * (j + NY * k): the components of a node are contiguous.
// Inactive entries are zero and never written.

#include <hip/hip_runtime.h>
#include "hr_device.h"

#pragma clang fp contract(off)

namespace hr {

// 1D GLL derivative matrix: DM[a][b] = l_b'(xi_a), xi = -1, 0, 1
__constant__ double c_dm[9] = {-1.5, 2.0, -0.5, -0.5, 0.0, 0.5, 0.5, -2.0, 1.5};

__device__ __forceinline__ double DM(int a, int b) { return c_dm[3 * a + b]; }

// ----------------------------------------------------- small kernels
// lattice <- reference numbering (optionally scaled): out[slot[e]] = x[e] * r[e]
__global__ void gather_eq_kernel(long long neq, const long long* __restrict__ slot, const double* __restrict__ x,
                                 const double* __restrict__ r, double* __restrict__ out) {
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < neq; e += (long long)gridDim.x * blockDim.x) {
    const double v = x[e];
    out[slot[e]] = r ? v * r[e] : v;
  }
}

// reference numbering <- lattice: z[e] = y[e] + s * (l[e] * raw[slot[e]])
__global__ void scatter_eq_kernel(long long neq, const long long* __restrict__ slot, const double* __restrict__ raw,
                                  const double* __restrict__ y, double s, const double* __restrict__ l,
                                  double* __restrict__ z) {
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < neq; e += (long long)gridDim.x * blockDim.x) {
    double a = raw[slot[e]];
    if (l) a = l[e] * a;
    z[e] = y[e] + s * a;
  }
}

// receiver traces: row[r] = x2[slot[r]]
__global__ void record_kernel(int nrec, const long long* __restrict__ slot, const double* __restrict__ x2,
                              double* __restrict__ row, const int* __restrict__ step, long long stride) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (step) row += (long long)(*step) * stride;
  if (r < nrec) row[r] = x2[slot[r]];
}

// device-side step index of a captured time loop
__global__ void step_set_kernel(int* __restrict__ step, int value) {
  if (threadIdx.x == 0) step[0] = value;
}
__global__ void step_advance_kernel(int* __restrict__ step, int delta) {
  if (threadIdx.x == 0) step[0] = step[0] + delta;
}

// displacement of the regular-domain node box -> history slice (3 x nbox,
// component-major): a thread reads one node's three contiguous components and
// writes them to the three component planes (each plane coalesced across lanes)
__global__ void store_box_kernel(Geo g, int bi0, int bj0, int bk0, int bnx, int bny, int bnz,
                                 const double* __restrict__ x2, double* __restrict__ hist,
                                 const int* __restrict__ step, long long stride) {
  // 32-bit indices: 9 NN < 2^31 (checked by the host)
  const int nbox = bnx * bny * bnz;
  if (step) hist += (long long)(*step) * stride;
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < nbox; r += gridDim.x * blockDim.x) {
    const int bi = r % bnx, bj = (r / bnx) % bny, bk = r / (bnx * bny);
    const int n = (bi0 + bi) + g.NX * ((bj0 + bj) + g.NY * (bk0 + bk));
    const double u0 = x2[3 * n], u1 = x2[3 * n + 1], u2 = x2[3 * n + 2];
    hist[r] = u0;
    hist[nbox + r] = u1;
    hist[2 * nbox + r] = u2;
  }
}

// Per-step gradient accumulation (Update_g + ELEMENT_g_3D, Assemble_PETSc_3D.F90:556-661,
// Element_Matrices_3D.F90:1725-1890): for every node of the regular-domain box,
// sum over the regular-domain elements containing it of W * div u * div p and
// W * [2 eps(u):eps(p) in the reference's component form], times dt.
// u comes from the history slice (3 x nbox), p from the adjoint x2 (lattice).
__global__ void gradient_kernel(Geo g, int bi0, int bj0, int bk0, int bnx, int bny, int bnz,
                                const double* __restrict__ uh, const double* __restrict__ p2, double dt,
                                double* __restrict__ glam, double* __restrict__ gmu) {
  // 32-bit indices: 9 NN < 2^31 (checked by the host)
  const int nbox = bnx * bny * bnz;
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < nbox; r += gridDim.x * blockDim.x) {
    const int bi = r % bnx, bj = (r / bnx) % bny, bk = r / (bnx * bny);
    const int pos[3] = {bi0 + bi, bj0 + bj, bk0 + bk};
    const int n = pos[0] + g.NX * (pos[1] + g.NY * pos[2]);
    const int stride[3] = {1, bnx, bnx * bny};
    const int lstride[3] = {1, g.NX, g.NX * g.NY};
    // per axis two element sides: s = 0 the element below (local 2; local 1 for an
    // odd node), s = 1 the element above (local 0).  Only sides whose element is
    // in the regular-domain range of the axis are used: their stencils stay in
    // the regular-domain node box.
    bool has[3][2];
    int el[3][2], lo[3][2];
    double du[3][2][3], dp[3][2][3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const int q = pos[a];
      const bool odd = q & 1;
      el[a][0] = odd ? (q - 1) >> 1 : (q >> 1) - 1;
      lo[a][0] = odd ? 1 : 2;
      el[a][1] = q >> 1;
      lo[a][1] = 0;
      has[a][0] = el[a][0] >= g.rlo[a] && el[a][0] < g.rhi[a];
      has[a][1] = !odd && el[a][1] >= g.rlo[a] && el[a][1] < g.rhi[a];
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        const int l = lo[a][s];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          double su = 0.0, sp = 0.0;
          if (has[a][s]) {
#pragma unroll
            for (int m = 0; m < 3; ++m) {
              const int off = m - l;
              su += DM(l, m) * uh[c * nbox + r + off * stride[a]];
              sp += DM(l, m) * p2[3 * (n + off * lstride[a]) + c];
            }
          }
          du[a][s][c] = g.dxi[a] * su;
          dp[a][s][c] = g.dxi[a] * sp;
        }
      }
    }
    double sl = 0.0, sm = 0.0;
#pragma unroll
    for (int sx = 0; sx < 2; ++sx)
#pragma unroll
      for (int sy = 0; sy < 2; ++sy)
#pragma unroll
        for (int sz = 0; sz < 2; ++sz) {
          if (!(has[0][sx] && has[1][sy] && has[2][sz])) continue;
          const double u1x = du[0][sx][0], u2x = du[0][sx][1], u3x = du[0][sx][2];
          const double u1y = du[1][sy][0], u2y = du[1][sy][1], u3y = du[1][sy][2];
          const double u1z = du[2][sz][0], u2z = du[2][sz][1], u3z = du[2][sz][2];
          const double p1x = dp[0][sx][0], p2x = dp[0][sx][1], p3x = dp[0][sx][2];
          const double p1y = dp[1][sy][0], p2y = dp[1][sy][1], p3y = dp[1][sy][2];
          const double p1z = dp[2][sz][0], p2z = dp[2][sz][1], p3z = dp[2][sz][2];
          const double sL = (u1x + u2y + u3z) * (p1x + p2y + p3z);
          const double sM = 2.0 * (u1x * p1x + u2y * p2y + u3z * p3z) + (p1y + p2x) * (u1y + u2x) +
                            (p1z + p3x) * (u1z + u3x) + (p3y + p2z) * (u3y + u2z);
          const double W = g.wq[lo[0][sx] + 3 * lo[1][sy] + 9 * lo[2][sz]];
          sl += sL * W;
          sm += sM * W;
        }
    glam[r] += sl * dt;
    gmu[r] += sm * dt;
  }
}

// The same per-step gradient, tiled: a block owns GTX x GTY box nodes in x-y
// and marches up a chunk of GZC planes.  Each value of u and p is read from
// memory once per block (plus a halo of two nodes in x and y and two planes in z):
// the thread of a column keeps the planes k-2..k+2 of its own u and p in
// registers (z stencils), the plane k of the block's window sits in LDS (x and y
// stencils).  Sums and their order are those of gradient_kernel, so the results
// are identical.
constexpr int GTX = 32, GTY = 8, GH = 2, GZC = 4;
constexpr int GWX = GTX + 2 * GH, GWY = GTY + 2 * GH;
constexpr int GNT = GTX * GTY;
// halo positions read by the stencils: rows 0, 1 and GWY-2, GWY-1 over the owned
// columns (y stencils), columns 0, 1 and GWX-2, GWX-1 over the owned rows (x stencils)
constexpr int GHALO = 2 * 2 * GTX + 2 * 2 * GTY;

__global__ void __launch_bounds__(GNT) gradient_tile_kernel(Geo g, int bi0, int bj0, int bk0, int bnx, int bny,
                                                            int bnz, const double* __restrict__ uh,
                                                            const double* __restrict__ p2, double dt,
                                                            double* __restrict__ glam, double* __restrict__ gmu,
                                                            const int* __restrict__ step, long long stride) {
  __shared__ double sU[3][GWY][GWX], sP[3][GWY][GWX];
  if (step) uh += (long long)(*step) * stride;
  const int tid = threadIdx.x;
  const int tx = tid % GTX, ty = tid / GTX;
  const int ti0 = blockIdx.x * GTX, tj0 = blockIdx.y * GTY;
  const int kc0 = blockIdx.z * GZC, kc1 = min(kc0 + GZC, bnz);
  const int bi = ti0 + tx, bj = tj0 + ty;
  const bool own = bi < bnx && bj < bny;
  const int nbox = bnx * bny * bnz;              // 32-bit: 9 NN < 2^31 (checked by the host)
  // u and p of box node (i, j, k) of this block's window (zero outside the box)
  auto fetch = [&](int i, int j, int k, double* u, double* p) {
    if (i >= 0 && j >= 0 && k >= 0 && i < bnx && j < bny && k < bnz) {
      const int r = i + bnx * (j + bny * k);
      const int n = (bi0 + i) + g.NX * ((bj0 + j) + g.NY * (bk0 + k));
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        u[c] = uh[c * nbox + r];
        p[c] = p2[3 * n + c];
      }
    } else {
#pragma unroll
      for (int c = 0; c < 3; ++c) { u[c] = 0.0; p[c] = 0.0; }
    }
  };
  // own column: planes k-2 .. k+2 (index 0..4)
  double qu[5][3], qp[5][3];
#pragma unroll
  for (int d = 0; d < 4; ++d) fetch(bi, bj, kc0 - 2 + d, qu[d], qp[d]);
  const int wx = GH + tx, wy = GH + ty;
  for (int k = kc0; k < kc1; ++k) {
    // this plane's accumulators, read with the plane's inputs
    const int r = bi + bnx * (bj + bny * k);
    double gl0 = 0.0, gm0 = 0.0;
    if (own) { gl0 = glam[r]; gm0 = gmu[r]; }
    fetch(bi, bj, k + 2, qu[4], qp[4]);
#pragma unroll
    for (int c = 0; c < 3; ++c) { sU[c][wy][wx] = qu[2][c]; sP[c][wy][wx] = qp[2][c]; }
    for (int h = tid; h < GHALO; h += GNT) {
      int hx, hy;
      if (h < 4 * GTX) { hy = h / GTX; hx = GH + (h - hy * GTX); if (hy >= 2) hy += GTY; }
      else { const int q = h - 4 * GTX; hy = GH + (q >> 2); hx = q & 3; if (hx >= 2) hx += GTX; }
      double u[3], p[3];
      fetch(ti0 + hx - GH, tj0 + hy - GH, k, u, p);
#pragma unroll
      for (int c = 0; c < 3; ++c) { sU[c][hy][hx] = u[c]; sP[c][hy][hx] = p[c]; }
    }
    __syncthreads();
    if (own) {
      const int pos[3] = {bi, bj, k};
      const int bn[3] = {bnx, bny, bnz};
      // per axis two element sides: s = 0 the element below (local 2; local 1 for an
      // odd node), s = 1 the element above (local 0), within the box
      bool has[3][2];
      int lo[3][2];
      double du[3][2][3], dp[3][2][3];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const int q = pos[a];
        const bool odd = q & 1;
        lo[a][0] = odd ? 1 : 2;
        lo[a][1] = 0;
        has[a][0] = odd ? true : q >= 2;
        has[a][1] = !odd && q + 2 <= bn[a] - 1;
#pragma unroll
        for (int s = 0; s < 2; ++s) {
          const int l = lo[a][s];
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            double su = 0.0, sp = 0.0;
            if (has[a][s]) {
#pragma unroll
              for (int m = 0; m < 3; ++m) {
                const int off = m - l;
                double uv, pv;
                if (a == 0) { uv = sU[c][wy][wx + off]; pv = sP[c][wy][wx + off]; }
                else if (a == 1) { uv = sU[c][wy + off][wx]; pv = sP[c][wy + off][wx]; }
                else { uv = qu[2 + off][c]; pv = qp[2 + off][c]; }
                su += DM(l, m) * uv;
                sp += DM(l, m) * pv;
              }
            }
            du[a][s][c] = g.dxi[a] * su;
            dp[a][s][c] = g.dxi[a] * sp;
          }
        }
      }
      double sl = 0.0, sm = 0.0;
#pragma unroll
      for (int sx = 0; sx < 2; ++sx)
#pragma unroll
        for (int sy = 0; sy < 2; ++sy)
#pragma unroll
          for (int sz = 0; sz < 2; ++sz) {
            if (!(has[0][sx] && has[1][sy] && has[2][sz])) continue;
            const double u1x = du[0][sx][0], u2x = du[0][sx][1], u3x = du[0][sx][2];
            const double u1y = du[1][sy][0], u2y = du[1][sy][1], u3y = du[1][sy][2];
            const double u1z = du[2][sz][0], u2z = du[2][sz][1], u3z = du[2][sz][2];
            const double p1x = dp[0][sx][0], p2x = dp[0][sx][1], p3x = dp[0][sx][2];
            const double p1y = dp[1][sy][0], p2y = dp[1][sy][1], p3y = dp[1][sy][2];
            const double p1z = dp[2][sz][0], p2z = dp[2][sz][1], p3z = dp[2][sz][2];
            const double sL = (u1x + u2y + u3z) * (p1x + p2y + p3z);
            const double sM = 2.0 * (u1x * p1x + u2y * p2y + u3z * p3z) + (p1y + p2x) * (u1y + u2x) +
                              (p1z + p3x) * (u1z + u3x) + (p3y + p2z) * (u3y + u2z);
            const double W = g.wq[lo[0][sx] + 3 * lo[1][sy] + 9 * lo[2][sz]];
            sl += sL * W;
            sm += sM * W;
          }
      glam[r] = gl0 + sl * dt;
      gmu[r] = gm0 + sm * dt;
    }
    __syncthreads();
#pragma unroll
    for (int d = 0; d < 4; ++d)
#pragma unroll
      for (int c = 0; c < 3; ++c) { qu[d][c] = qu[d + 1][c]; qp[d][c] = qp[d + 1][c]; }
  }
}

// dst[n][q] = src[n][perm[q]]: receiver columns in source-entry order
__global__ void permute_cols_kernel(long long nrows, int ncols, const long long* __restrict__ perm,
                                    const double* __restrict__ src, double* __restrict__ dst) {
  const long long tot = nrows * ncols;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < tot; t += (long long)gridDim.x * blockDim.x) {
    const long long n = t / ncols;
    const int q = (int)(t - n * ncols);
    dst[t] = src[n * ncols + perm[q]];
  }
}

// Misfit partial sums per step (Compute_Misfit, Solver_PETSc_Inverse.F90:633-650):
// the residuals of receivers that share an equation are added first (VecSetValues
// with ADD_VALUES), then the squares are summed in equation order (VecDot).
// Group g holds the receivers gidx[gptr[g] .. gptr[g+1]), groups in ascending
// equation order.
__global__ void __launch_bounds__(256) misfit_rows_kernel(int nstep, int ngrp, const int* __restrict__ gptr,
                                                          const int* __restrict__ gidx, int nrec,
                                                          const double* __restrict__ tr,
                                                          const double* __restrict__ ob, double* __restrict__ part) {
  // one wavefront per time row: lane l takes groups l, l + 64, ... (neighbouring
  // receivers of the row, coalesced), then a fixed butterfly sums the lanes
  const int lane = threadIdx.x & 63;
  const int nstp = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (nstp > nstep) return;                       // whole wavefront (uniform)
  const long long row = (long long)nstp * nrec;
  double c = 0.0;
  for (int g = lane; g < ngrp; g += 64) {
    double b = 0.0;
    for (int j = gptr[g]; j < gptr[g + 1]; ++j) {
      const int r = gidx[j];
      b += tr[row + r] - ob[row + r];
    }
    c += b * b;
  }
  for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
  if (lane == 0) part[nstp] = c;
}

// J = sum_{n = NSTEP..1} 0.5 * part[n] * dt  (Compute_Misfit order)
__global__ void misfit_sum_kernel(int nstep, const double* __restrict__ part, double dt, double* __restrict__ out) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    double cost = 0.0;
    for (int n = nstep; n >= 1; --n) cost = cost + 0.5 * part[n] * dt;
    out[0] = cost;
  }
}

// out = in[0] + in[1] + ... + in[n-1], in that order (per-shot misfits)
__global__ void sum_kernel(int n, const double* __restrict__ in, double* __restrict__ out) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    double c = 0.0;
    for (int q = 0; q < n; ++q) c = c + in[q];
    out[0] = c;
  }
}

// g <- -g for the regular-domain box, scattered to the nodal output
__global__ void gradient_finish_kernel(Geo g, int bi0, int bj0, int bk0, int bnx, int bny, int bnz,
                                       const double* __restrict__ gl_box, const double* __restrict__ gm_box,
                                       double* __restrict__ gl, double* __restrict__ gm) {
  const long long nbox = (long long)bnx * bny * bnz;
  for (long long r = blockIdx.x * (long long)blockDim.x + threadIdx.x; r < nbox; r += (long long)gridDim.x * blockDim.x) {
    const int bi = (int)(r % bnx), bj = (int)((r / bnx) % bny), bk = (int)(r / ((long long)bnx * bny));
    const long long n = (long long)(bi0 + bi) + (long long)g.NX * ((long long)(bj0 + bj) + (long long)g.NY * (bk0 + bk));
    gl[n] = gl_box[r] * -1.0;
    gm[n] = gm_box[r] * -1.0;
  }
}

// g[n] += box[r] over the regular-domain node box (one Update_g step into the
// caller's nodal gradients)
__global__ void box_add_kernel(Geo g, int bi0, int bj0, int bk0, int bnx, int bny, int bnz,
                               const double* __restrict__ gl_box, const double* __restrict__ gm_box,
                               double* __restrict__ gl, double* __restrict__ gm) {
  const long long nbox = (long long)bnx * bny * bnz;
  for (long long r = blockIdx.x * (long long)blockDim.x + threadIdx.x; r < nbox; r += (long long)gridDim.x * blockDim.x) {
    const int bi = (int)(r % bnx), bj = (int)((r / bnx) % bny), bk = (int)(r / ((long long)bnx * bny));
    const long long n = (long long)(bi0 + bi) + (long long)g.NX * ((long long)(bj0 + bj) + (long long)g.NY * (bk0 + bk));
    gl[n] = gl[n] + gl_box[r];
    gm[n] = gm[n] + gm_box[r];
  }
}

// Regular-domain energy (TOTAL_ENERGY_RD_PETSC_EXPLICIT, Obsolete.F90:108-150):
// 0.5 (u . AK_RD u + (M_RD x3) . x3); AK_RD is never assembled (the insertion is
// commented out, Assemble_PETSc_3D.F90:163), so only the kinetic term remains.
// Fixed grid of ENERGY_BLOCKS blocks, partial sums in block order: deterministic.
constexpr int ENERGY_BLOCKS = 256;
__global__ void __launch_bounds__(256) energy_partial_kernel(long long nn, const double* __restrict__ mrd,
                                                             const double* __restrict__ x3,
                                                             double* __restrict__ part) {
  __shared__ double red[256];
  double acc = 0.0;
  for (long long n = blockIdx.x * 256LL + threadIdx.x; n < nn; n += 256LL * gridDim.x) {
    const double m = mrd[n];
    const double a = x3[3 * n], b = x3[3 * n + 1], c = x3[3 * n + 2];
    acc += (m * a) * a + (m * b) * b + (m * c) * c;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}
__global__ void energy_finish_kernel(const double* __restrict__ part, double* __restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double c = 0.0;
    for (int b = 0; b < ENERGY_BLOCKS; ++b) c += part[b];
    out[0] = 0.50 * (0.0 + c);
  }
}

}  // namespace hr

// ------------------------------------------------------------ launchers
namespace hr {
namespace {
int grid_for(long long n) {
  long long b = (n + 255) / 256;
  if (b < 1) b = 1;
  return (int)(b < 65535 ? b : 65535);
}
}  // namespace

hipError_t launch_gather_eq(long long neq, const long long* slot, const double* x, const double* r, double* out,
                            hipStream_t s) {
  hipLaunchKernelGGL(gather_eq_kernel, dim3(grid_for(neq)), dim3(256), 0, s, neq, slot, x, r, out);
  return hipGetLastError();
}
hipError_t launch_scatter_eq(long long neq, const long long* slot, const double* raw, const double* y, double sc,
                             const double* l, double* z, hipStream_t s) {
  hipLaunchKernelGGL(scatter_eq_kernel, dim3(grid_for(neq)), dim3(256), 0, s, neq, slot, raw, y, sc, l, z);
  return hipGetLastError();
}
hipError_t launch_record(int nrec, const long long* slot, const double* x2, double* row, hipStream_t s, StepRef sr) {
  hipLaunchKernelGGL(record_kernel, dim3((nrec + 255) / 256), dim3(256), 0, s, nrec, slot, x2, row, sr.step, sr.stride);
  return hipGetLastError();
}
hipError_t launch_step_set(int* step, int value, hipStream_t s) {
  hipLaunchKernelGGL(step_set_kernel, dim3(1), dim3(64), 0, s, step, value);
  return hipGetLastError();
}
hipError_t launch_step_advance(int* step, int delta, hipStream_t s) {
  hipLaunchKernelGGL(step_advance_kernel, dim3(1), dim3(64), 0, s, step, delta);
  return hipGetLastError();
}
hipError_t launch_store_box(const Geo& g, const int* b0, const int* bn, const double* x2, double* hist, hipStream_t s,
                            StepRef sr) {
  const long long tot = (long long)bn[0] * bn[1] * bn[2];
  hipLaunchKernelGGL(store_box_kernel, dim3(grid_for(tot)), dim3(256), 0, s, g, b0[0], b0[1], b0[2], bn[0], bn[1],
                     bn[2], x2, hist, sr.step, sr.stride);
  return hipGetLastError();
}
hipError_t launch_gradient(const Geo& g, const int* b0, const int* bn, const double* uh, const double* p2, double dt,
                           double* glam, double* gmu, hipStream_t s, StepRef sr) {
  const dim3 grid((bn[0] + GTX - 1) / GTX, (bn[1] + GTY - 1) / GTY, (bn[2] + GZC - 1) / GZC);
  hipLaunchKernelGGL(gradient_tile_kernel, grid, dim3(GNT), 0, s, g, b0[0], b0[1], b0[2], bn[0], bn[1], bn[2], uh, p2,
                     dt, glam, gmu, sr.step, sr.stride);
  return hipGetLastError();
}

// the node-per-thread form (kept as the reference for the tiled kernel's tests)
hipError_t launch_gradient_pointwise(const Geo& g, const int* b0, const int* bn, const double* uh, const double* p2,
                                     double dt, double* glam, double* gmu, hipStream_t s) {
  const long long tot = (long long)bn[0] * bn[1] * bn[2];
  hipLaunchKernelGGL(gradient_kernel, dim3(grid_for(tot)), dim3(256), 0, s, g, b0[0], b0[1], b0[2], bn[0], bn[1],
                     bn[2], uh, p2, dt, glam, gmu);
  return hipGetLastError();
}
hipError_t launch_gradient_finish(const Geo& g, const int* b0, const int* bn, const double* glb, const double* gmb,
                                  double* gl, double* gm, hipStream_t s) {
  const long long tot = (long long)bn[0] * bn[1] * bn[2];
  hipLaunchKernelGGL(gradient_finish_kernel, dim3(grid_for(tot)), dim3(256), 0, s, g, b0[0], b0[1], b0[2], bn[0],
                     bn[1], bn[2], glb, gmb, gl, gm);
  return hipGetLastError();
}
hipError_t launch_box_add(const Geo& g, const int* b0, const int* bn, const double* glb, const double* gmb, double* gl,
                          double* gm, hipStream_t s) {
  const long long tot = (long long)bn[0] * bn[1] * bn[2];
  hipLaunchKernelGGL(box_add_kernel, dim3(grid_for(tot)), dim3(256), 0, s, g, b0[0], b0[1], b0[2], bn[0], bn[1],
                     bn[2], glb, gmb, gl, gm);
  return hipGetLastError();
}
int energy_partials_size() { return ENERGY_BLOCKS; }
hipError_t launch_energy(long long nn, const double* mrd, const double* x3, double* part, double* out, hipStream_t s) {
  hipLaunchKernelGGL(energy_partial_kernel, dim3(ENERGY_BLOCKS), dim3(256), 0, s, nn, mrd, x3, part);
  hipLaunchKernelGGL(energy_finish_kernel, dim3(1), dim3(64), 0, s, part, out);
  return hipGetLastError();
}
hipError_t launch_permute_cols(long long nrows, int ncols, const long long* perm, const double* src, double* dst,
                               hipStream_t s) {
  hipLaunchKernelGGL(permute_cols_kernel, dim3(grid_for(nrows * ncols)), dim3(256), 0, s, nrows, ncols, perm, src, dst);
  return hipGetLastError();
}
hipError_t launch_sum(int n, const double* in, double* out, hipStream_t s) {
  hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(64), 0, s, n, in, out);
  return hipGetLastError();
}
hipError_t launch_misfit(int nstep, int ngrp, const int* gptr, const int* gidx, int nrec, const double* tr,
                         const double* ob, double dt, double* part, double* out, hipStream_t s) {
  hipLaunchKernelGGL(misfit_rows_kernel, dim3((nstep + 1 + 3) / 4), dim3(256), 0, s, nstep, ngrp, gptr, gidx, nrec,
                     tr, ob, part);
  hipLaunchKernelGGL(misfit_sum_kernel, dim3(1), dim3(64), 0, s, nstep, part, dt, out);
  return hipGetLastError();
}
}  // namespace hr
