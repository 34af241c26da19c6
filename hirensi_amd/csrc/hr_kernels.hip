// HIP kernels for the matrix-free explicit elastic operator with the hybrid
// displacement-stress PML on a structured lattice of 27-node GLL hexahedra.
//
// What is computed is the operator the reference assembles into PETSc AIJ
// matrices (Assemble_PETSc_3D.F90:6-436, element blocks of
// Element_Matrices_3D.F90:7-882) and applies 12 times per RK4 step
// (Solver_PETSc_Explicit.F90:1001-1046, transposed at :1319-1364).
//
// Layout: every field is SoA over the full lattice of GLL nodes,
// f[c * NN + n], n = i + NX * (j + NY * k), c = 0..8 (u_x u_y u_z s_xx s_yy
// s_zz s_xy s_xz s_yz).  Constrained entries are zero and never written.
//
// Stage kernel: one workgroup owns a tile of TX x TY elements in x-y and a
// chunk of g.tz element layers in z, and streams slab by slab up the chunk.
// A slab's lattice window (owned nodes plus a low-side halo of one element)
// is staged in LDS; regular-domain elements store W*sigma per GLL point in
// LDS; the PML u-rows use a nodal flux G (the PML stress coupling is pointwise
// at GLL points) and the PML s-rows use element-side derivative stencils.
// Each owned node pulls its contributions, so every output is written once,
// without atomics, in a fixed order.

#include <hip/hip_runtime.h>
#include "hr_device.h"

#pragma clang fp contract(off)

namespace hr {

// 1D GLL derivative matrix: DM[a][b] = l_b'(xi_a), xi = -1, 0, 1
__constant__ double c_dm[9] = {-1.5, 2.0, -0.5, -0.5, 0.0, 0.5, 0.5, -2.0, 1.5};

__device__ __forceinline__ double DM(int a, int b) { return c_dm[3 * a + b]; }

// ------------------------------------------------------------------ PML
// Stretch factors at a node, as PML_FINDER_3D (PML.F90:325-891) with
// polynomial degree 2 and zero cross ratio.
struct Stretch {
  double ax, ay, az, bx, by, bz;
};

__device__ __forceinline__ Stretch pml_stretch(const double* P, double x, double y, double z) {
  Stretch s{1.0, 1.0, 1.0, 0.0, 0.0, 0.0};
  if (x >= P[18]) {
    const double f = (x - P[18]) * 1.0 / P[12];
    s.ax = 1.0 + P[0] * (f * f);
    s.bx = 0.0 + P[6] * (f * f);
  } else if (x <= P[19]) {
    const double f = (x - P[19]) * -1.0 / P[13];
    s.ax = 1.0 + P[1] * (f * f);
    s.bx = 0.0 + P[7] * (f * f);
  }
  if (y >= P[20]) {
    const double f = (y - P[20]) * 1.0 / P[14];
    s.ay = 1.0 + P[2] * (f * f);
    s.by = 0.0 + P[8] * (f * f);
  } else if (y <= P[21]) {
    const double f = (y - P[21]) * -1.0 / P[15];
    s.ay = 1.0 + P[3] * (f * f);
    s.by = 0.0 + P[9] * (f * f);
  }
  if (z <= P[23]) {
    const double f = (z - P[23]) * -1.0 / P[17];
    s.az = 1.0 + P[5] * (f * f);
    s.bz = 0.0 + P[11] * (f * f);
  }
  return s;
}

// a, b, c, d of PML_FINDER_3D items 7..10
struct Abcd {
  double a, b, c, d;
};
__device__ __forceinline__ Abcd pml_abcd(const Stretch& s) {
  Abcd r;
  r.a = s.ax * s.ay * s.az;
  r.b = s.ax * s.ay * s.bz + s.ax * s.az * s.by + s.ay * s.az * s.bx;
  r.c = s.ax * s.by * s.bz + s.ay * s.bx * s.bz + s.az * s.bx * s.by;
  r.d = s.bx * s.by * s.bz;
  return r;
}

// F[set][k]: stretch weight of the derivative along k in the coupling blocks
// (ELEMENT_EA_3D FAC3/FAC2/FAC1 for k = x/y/z); set 0 = w (G, v1),
// 1 = p (K, v2), 2 = e (C, v3).
__device__ __forceinline__ void pml_f(const Stretch& s, double F[3][3]) {
  F[2][0] = s.ay * s.az;
  F[2][1] = s.ax * s.az;
  F[2][2] = s.ax * s.ay;
  F[1][0] = s.ay * s.bz + s.az * s.by;
  F[1][1] = s.ax * s.bz + s.az * s.bx;
  F[1][2] = s.ax * s.by + s.ay * s.bx;
  F[0][0] = s.by * s.bz;
  F[0][1] = s.bx * s.bz;
  F[0][2] = s.bx * s.by;
}

// ------------------------------------------------------------- DOF rules
__device__ __forceinline__ bool in_domain(const Geo& g, int i, int j, int k) {
  return i >= 0 && j >= 0 && k >= 0 && i < g.NX && j < g.NY && k < g.NZ;
}
__device__ __forceinline__ bool u_active(const Geo& g, int i, int j, int k) {
  return !(i == 0 || j == 0 || k == 0 || i == g.NX - 1 || j == g.NY - 1);
}
// node touched by at least one PML element
__device__ __forceinline__ bool pml_node(const Geo& g, int i, int j, int k) {
  return !(i > 2 * g.rlo[0] && i < 2 * g.rhi[0] && j > 2 * g.rlo[1] && j < 2 * g.rhi[1] && k > 2 * g.rlo[2]);
}
// active stress components (bit m for s component m = 0..5) of a PML node
__device__ __forceinline__ int s_mask(const Geo& g, int i, int j, int k) {
  if (!pml_node(g, i, j, k)) return 0;
  if (g.top_rule && k == g.NZ - 1 && !((i & 1) && (j & 1))) return 0x3f & ~((1 << 2) | (1 << 4) | (1 << 5));
  return 0x3f;
}
__device__ __forceinline__ bool rd_elem(const Geo& g, int ex, int ey, int ez) {
  return ex >= g.rlo[0] && ex < g.rhi[0] && ey >= g.rlo[1] && ey < g.rhi[1] && ez >= g.rlo[2] && ez < g.rhi[2];
}

// symmetric index of (i, k): xx yy zz xy xz yz
__device__ __forceinline__ int sym(int i, int k) { return i == k ? i : i + k + 2; }

// --------------------------------------------------------- stage inputs
// Stage-s inputs v1, v2, v3 for one field component, in the reference's
// nested VecWAXPY order (Solver_PETSc_Explicit.F90:1001-1034).
struct In6 {
  const double* p[6];
};

__device__ __forceinline__ void stage_inputs(const StageArgs& A, long long idx, bool n1, bool n2, bool n3,
                                             double& v1, double& v2, double& v3) {
  v1 = v2 = v3 = 0.0;
  const double hh = A.hh, h = A.h;
  switch (A.stage) {
    case 0: {
      const double x = A.in[0][idx];
      if (A.set_mask & 1) v1 = x;
      if (A.set_mask & 2) v2 = x;
      if (A.set_mask & 4) v3 = x;
      return;
    }
    case 1: {
      if (n1) v1 = A.in[0][idx];
      if (n2) v2 = A.in[1][idx];
      if (n3) v3 = A.in[2][idx];
      break;
    }
    case 2: {
      const double x2 = (n1 || n2) ? A.in[1][idx] : 0.0;
      const double x3 = (n2 || n3) ? A.in[2][idx] : 0.0;
      if (n1) v1 = A.in[0][idx] + hh * x2;
      if (n2) v2 = x2 + hh * x3;
      if (n3) v3 = x3 + hh * A.in[3][idx];
      break;
    }
    case 3: {
      const double x2 = (n1 || n2) ? A.in[1][idx] : 0.0;
      const double x3 = A.in[2][idx];
      if (n1) v1 = A.in[0][idx] + hh * (x2 + hh * x3);
      if (n2) v2 = x2 + hh * (x3 + hh * A.in[3][idx]);
      if (n3) v3 = x3 + hh * A.in[4][idx];
      break;
    }
    default: {
      const double x2 = (n1 || n2) ? A.in[1][idx] : 0.0;
      const double x3 = A.in[2][idx];
      if (n1) v1 = A.in[0][idx] + h * (x2 + hh * (x3 + hh * A.in[3][idx]));
      if (n2) v2 = x2 + h * (x3 + hh * A.in[4][idx]);
      if (n3) v3 = x3 + h * A.in[5][idx];
      break;
    }
  }
  v1 = v1 * A.sgn[0];
  v2 = v2 * A.sgn[1];
  v3 = v3 * A.sgn[2];
}

// ----------------------------------------------------------- stage kernel
template <int TX, int TY>
struct TileShape {
  static constexpr int WX = 2 * TX + 3;        // window lattice width (low-side halo element)
  static constexpr int WY = 2 * TY + 3;
  static constexpr int WP = WX * WY;           // window nodes per plane
  static constexpr int NEW = (TX + 1) * (TY + 1);
  static constexpr int OX = 2 * TX + 1;        // owner columns (last one only in the last tile)
  static constexpr int OY = 2 * TY + 1;
  static constexpr int OC = OX * OY;
};

constexpr int STAGE_THREADS = 256;

template <int TX, int TY>
__global__ void __launch_bounds__(STAGE_THREADS) stage_kernel(const StageArgs A) {
  using TS = TileShape<TX, TY>;
  static_assert(3 * TS::OC <= STAGE_THREADS, "owner mapping needs 3 planes x columns threads");
  __shared__ double sU[3][9][TS::WP];      // v1_u v2_u v3_u, ring of 3 planes (slot = k mod 3)
  __shared__ double sG[3][9][TS::WP];      // nodal PML flux G_ik, index 3*i + k
  __shared__ double sSig[TS::NEW][6][27];  // regular-domain W * sigma per element point
  __shared__ double sCarry[2][9][TS::OC];  // partial sums of the shared top plane of a slab
  __shared__ int sType[TS::NEW];           // 0 outside, 1 regular domain, 2 PML

  const Geo& g = A.g;
  const int tid = threadIdx.x;
  const int b = blockIdx.x;
  const int tx = b % g.tiles[0];
  const int ty = (b / g.tiles[0]) % g.tiles[1];
  const int tz = b / (g.tiles[0] * g.tiles[1]);
  const int ex0 = tx * TX, ey0 = ty * TY, ez0 = tz * g.tz;
  const int ez_end = min(ez0 + g.tz, g.nez);
  const int gi0 = 2 * ex0 - 2, gj0 = 2 * ey0 - 2;
  const int own_i1 = (ex0 + TX >= g.nex) ? g.NX : 2 * (ex0 + TX);
  const int own_j1 = (ey0 + TY >= g.ney) ? g.NY : 2 * (ey0 + TY);
  const int ez_first = max(ez0 - 1, 0);
  const bool useG = (A.set_mask & 1), useK = (A.set_mask & 2), useC = (A.set_mask & 4);
  const long long NN = g.NN;

  for (int ez = ez_first; ez < ez_end; ++ez) {
    const bool halo = ez < ez0;
    // ---------------------------------------------------------------- load
    const int kfirst = (ez == ez_first) ? 2 * ez : 2 * ez + 1;
    const int nload = (2 * ez + 2 - kfirst + 1) * TS::WP;
    for (int t = tid; t < nload; t += STAGE_THREADS) {
      const int pl = t / TS::WP, r = t - pl * TS::WP;
      const int li = r % TS::WX, lj = r / TS::WX;
      const int i = gi0 + li, j = gj0 + lj, k = kfirst + pl;
      const int slot = k % 3;
      double U9[9], G9[9];
#pragma unroll
      for (int c = 0; c < 9; ++c) { U9[c] = 0.0; G9[c] = 0.0; }
      if (in_domain(g, i, j, k)) {
        const long long n = (long long)i + (long long)g.NX * ((long long)j + (long long)g.NY * k);
        const bool pn = pml_node(g, i, j, k);
        if (u_active(g, i, j, k)) {
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            double v1, v2, v3;
            stage_inputs(A, c * NN + n, pn && useG, useK, pn && useC, v1, v2, v3);
            U9[c] = v1; U9[3 + c] = v2; U9[6 + c] = v3;
          }
        }
        if (pn) {
          const int sm = s_mask(g, i, j, k);
          double s1[6], s2[6], s3[6];
#pragma unroll
          for (int m = 0; m < 6; ++m) {
            s1[m] = s2[m] = s3[m] = 0.0;
            if (sm & (1 << m)) stage_inputs(A, (3 + m) * NN + n, useG, useK, useC, s1[m], s2[m], s3[m]);
          }
          const Stretch st = pml_stretch(g.pml, A.nd.xc[i], A.nd.yc[j], A.nd.zc[k]);
          double F[3][3];
          pml_f(st, F);
          const double* sv[3] = {s1, s2, s3};
          if (!A.transpose) {
            const double lam = A.nd.lam[n], mu = A.nd.mu[n];
            const double l2m = lam + 2.0 * mu, m2 = 2.0 * mu;
#pragma unroll
            for (int X = 0; X < 3; ++X) {
              if (!(A.set_mask & (1 << X))) continue;
              const double* s = sv[X];
              // T(s): stress from the auxiliary strain (ELEMENT_EA_3D lower block)
              const double Txx = l2m * s[0] + lam * s[1] + lam * s[2];
              const double Tyy = lam * s[0] + l2m * s[1] + lam * s[2];
              const double Tzz = lam * s[0] + lam * s[1] + l2m * s[2];
              const double Txy = m2 * s[3], Txz = m2 * s[4], Tyz = m2 * s[5];
              const double Tm[9] = {Txx, Txy, Txz, Txy, Tyy, Tyz, Txz, Tyz, Tzz};
#pragma unroll
              for (int ii = 0; ii < 3; ++ii)
#pragma unroll
                for (int kk = 0; kk < 3; ++kk) G9[3 * ii + kk] -= F[X][kk] * Tm[3 * ii + kk];
            }
          } else {
#pragma unroll
            for (int X = 0; X < 3; ++X) {
              if (!(A.set_mask & (1 << X))) continue;
              const double* s = sv[X];
              const double Ym[9] = {s[0], s[3], s[4], s[3], s[1], s[5], s[4], s[5], s[2]};
#pragma unroll
              for (int ii = 0; ii < 3; ++ii)
#pragma unroll
                for (int kk = 0; kk < 3; ++kk) G9[3 * ii + kk] += F[X][kk] * Ym[3 * ii + kk];
            }
          }
        }
      }
#pragma unroll
      for (int c = 0; c < 9; ++c) { sU[slot][c][r] = U9[c]; sG[slot][c][r] = G9[c]; }
    }
    for (int t = tid; t < TS::NEW; t += STAGE_THREADS) {
      const int ex = ex0 - 1 + t % (TX + 1), ey = ey0 - 1 + t / (TX + 1);
      int ty_ = 0;
      if (ex >= 0 && ey >= 0 && ex < g.nex && ey < g.ney) ty_ = rd_elem(g, ex, ey, ez) ? 1 : 2;
      sType[t] = ty_;
    }
    __syncthreads();

    // ------------------------------------------- regular-domain W * sigma
    if (useK) {
      for (int t = tid; t < TS::NEW * 27; t += STAGE_THREADS) {
        const int e = t / 27, q = t - 27 * e;
        if (sType[e] != 1) continue;
        const int lx = q % 3, ly = (q / 3) % 3, lz = q / 9;
        const int li0 = 2 * (e % (TX + 1)), lj0 = 2 * (e / (TX + 1));
        const int rq = (li0 + lx) + TS::WX * (lj0 + ly);
        const int sq = (2 * ez + lz) % 3;
        double d[3][3];   // d[i][k] = d u_i / d x_k of v2
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const double* Uc0 = &sU[sq][3 + c][0];
          d[c][0] = g.dxi[0] * (DM(lx, 0) * Uc0[rq - lx] + DM(lx, 1) * Uc0[rq - lx + 1] + DM(lx, 2) * Uc0[rq - lx + 2]);
          d[c][1] = g.dxi[1] * (DM(ly, 0) * Uc0[rq - TS::WX * ly] + DM(ly, 1) * Uc0[rq - TS::WX * ly + TS::WX] +
                                DM(ly, 2) * Uc0[rq - TS::WX * ly + 2 * TS::WX]);
          double acc = 0.0;
#pragma unroll
          for (int bb = 0; bb < 3; ++bb) acc += DM(lz, bb) * sU[(2 * ez + bb) % 3][3 + c][rq];
          d[c][2] = g.dxi[2] * acc;
        }
        const int i = gi0 + li0 + lx, j = gj0 + lj0 + ly, k = 2 * ez + lz;
        const long long n = (long long)i + (long long)g.NX * ((long long)j + (long long)g.NY * k);
        const double lam = A.nd.lam[n], mu = A.nd.mu[n];
        const double W = g.wq[q];
        const double tr = d[0][0] + d[1][1] + d[2][2];
        sSig[e][0][q] = W * (lam * tr + 2.0 * mu * d[0][0]);
        sSig[e][1][q] = W * (lam * tr + 2.0 * mu * d[1][1]);
        sSig[e][2][q] = W * (lam * tr + 2.0 * mu * d[2][2]);
        sSig[e][3][q] = W * (mu * (d[0][1] + d[1][0]));
        sSig[e][4][q] = W * (mu * (d[0][2] + d[2][0]));
        sSig[e][5][q] = W * (mu * (d[1][2] + d[2][1]));
      }
    }
    __syncthreads();

    // ------------------------------------------------------- owner phase
    if (tid < 3 * TS::OC) {
      const int p = tid / TS::OC, c = tid - p * TS::OC;
      const int ix = c % TS::OX, iy = c / TS::OX;
      const int i = 2 * ex0 + ix, j = 2 * ey0 + iy, k = 2 * ez + p;
      const bool owned = i < own_i1 && j < own_j1;
      const bool top_plane = (ez == g.nez - 1) && p == 2;
      if (owned && (!halo || p == 2)) {
        double au[3] = {0.0, 0.0, 0.0};
        double aE[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        const int lz = p;
        const int sP = k % 3;
        // element candidates along x and y: a = 0 the element below (local 2, or
        // local 1 for an odd node), a = 1 the element above (local 0)
        const bool oddx = i & 1, oddy = j & 1;
        const bool pn = pml_node(g, i, j, k);
        double F[3][3];
        if (pn) {
          const Stretch st = pml_stretch(g.pml, A.nd.xc[i], A.nd.yc[j], A.nd.zc[k]);
          pml_f(st, F);
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int bb = 0; bb < 2; ++bb) {
            int ex, lx, ey, ly;
            if (oddx) { if (a) continue; ex = (i - 1) >> 1; lx = 1; }
            else if (a == 0) { if (i < 2) continue; ex = (i >> 1) - 1; lx = 2; }
            else { if ((i >> 1) >= g.nex) continue; ex = i >> 1; lx = 0; }
            if (oddy) { if (bb) continue; ey = (j - 1) >> 1; ly = 1; }
            else if (bb == 0) { if (j < 2) continue; ey = (j >> 1) - 1; ly = 2; }
            else { if ((j >> 1) >= g.ney) continue; ey = j >> 1; ly = 0; }
            const int e = (ex - ex0 + 1) + (TX + 1) * (ey - ey0 + 1);
            const int et = sType[e];
            const int li0 = 2 * (ex - ex0 + 1), lj0 = 2 * (ey - ey0 + 1);
            if (et == 1) {
              if (!useK) continue;
#pragma unroll
              for (int ii = 0; ii < 3; ++ii) {
                double s = 0.0, t2 = 0.0, t3 = 0.0;
#pragma unroll
                for (int qq = 0; qq < 3; ++qq) {
                  s += DM(qq, lx) * sSig[e][sym(ii, 0)][qq + 3 * ly + 9 * lz];
                  t2 += DM(qq, ly) * sSig[e][sym(ii, 1)][lx + 3 * qq + 9 * lz];
                  t3 += DM(qq, lz) * sSig[e][sym(ii, 2)][lx + 3 * ly + 9 * qq];
                }
                au[ii] += g.dxi[0] * s + g.dxi[1] * t2 + g.dxi[2] * t3;
              }
            } else if (et == 2) {
              // u-rows: divergence of the nodal PML flux
              const int rP = (li0 + lx) + TS::WX * (lj0 + ly);
#pragma unroll
              for (int ii = 0; ii < 3; ++ii) {
                double s = 0.0, t2 = 0.0, t3 = 0.0;
#pragma unroll
                for (int qq = 0; qq < 3; ++qq) {
                  s += DM(qq, lx) * (g.wq[qq + 3 * ly + 9 * lz] * sG[sP][3 * ii + 0][rP + (qq - lx)]);
                  t2 += DM(qq, ly) * (g.wq[lx + 3 * qq + 9 * lz] * sG[sP][3 * ii + 1][rP + TS::WX * (qq - ly)]);
                  t3 += DM(qq, lz) * (g.wq[lx + 3 * ly + 9 * qq] * sG[(2 * ez + qq) % 3][3 * ii + 2][rP]);
                }
                au[ii] += g.dxi[0] * s + g.dxi[1] * t2 + g.dxi[2] * t3;
              }
              // s-rows: F-weighted element-local derivatives at the node
              double dX[3] = {0, 0, 0}, dY[3] = {0, 0, 0}, dZ[3] = {0, 0, 0};
#pragma unroll
              for (int X = 0; X < 3; ++X) {
                if (!(A.set_mask & (1 << X))) continue;
#pragma unroll
                for (int ii = 0; ii < 3; ++ii) {
                  const double* Uc = &sU[sP][3 * X + ii][0];
                  const double ddx = g.dxi[0] * (DM(lx, 0) * Uc[rP - lx] + DM(lx, 1) * Uc[rP - lx + 1] +
                                                 DM(lx, 2) * Uc[rP - lx + 2]);
                  const double ddy = g.dxi[1] * (DM(ly, 0) * Uc[rP - TS::WX * ly] +
                                                 DM(ly, 1) * Uc[rP - TS::WX * ly + TS::WX] +
                                                 DM(ly, 2) * Uc[rP - TS::WX * ly + 2 * TS::WX]);
                  double acc = 0.0;
#pragma unroll
                  for (int qq = 0; qq < 3; ++qq) acc += DM(lz, qq) * sU[(2 * ez + qq) % 3][3 * X + ii][rP];
                  const double ddz = g.dxi[2] * acc;
                  dX[ii] += F[X][0] * ddx;
                  dY[ii] += F[X][1] * ddy;
                  dZ[ii] += F[X][2] * ddz;
                }
              }
              const double W = g.wq[lx + 3 * ly + 9 * lz];
              aE[0] += W * dX[0];
              aE[1] += W * dY[1];
              aE[2] += W * dZ[2];
              aE[3] += W * (dY[0] + dX[1]);
              aE[4] += W * (dZ[0] + dX[2]);
              aE[5] += W * (dZ[1] + dY[2]);
            }
          }
        double acc9[9] = {au[0], au[1], au[2], 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        const long long n = (long long)i + (long long)g.NX * ((long long)j + (long long)g.NY * k);
        if (pn) {
          if (!A.transpose) {
#pragma unroll
            for (int m = 0; m < 6; ++m) acc9[3 + m] = aE[m];
          } else {
            const double lam = A.nd.lam[n], mu = A.nd.mu[n];
            const double l2m = lam + 2.0 * mu, m2 = 2.0 * mu;
            acc9[3] = -(l2m * aE[0] + lam * aE[1] + lam * aE[2]);
            acc9[4] = -(lam * aE[0] + l2m * aE[1] + lam * aE[2]);
            acc9[5] = -(lam * aE[0] + lam * aE[1] + l2m * aE[2]);
            acc9[6] = -(m2 * aE[3]);
            acc9[7] = -(m2 * aE[4]);
            acc9[8] = -(m2 * aE[5]);
          }
        }
        const int cur = ez & 1, prev = cur ^ 1;
        if (p == 0 && ez > ez_first) {
#pragma unroll
          for (int m = 0; m < 9; ++m) acc9[m] += sCarry[prev][m][c];
        }
        if (p == 2 && !top_plane) {
#pragma unroll
          for (int m = 0; m < 9; ++m) sCarry[cur][m][c] = acc9[m];
        } else if (!halo) {
          // --------------------------------------------- node complete
          const bool ua = u_active(g, i, j, k);
          const int sm = s_mask(g, i, j, k);
          double v1[9], v2[9], v3[9];
#pragma unroll
          for (int m = 0; m < 9; ++m) {
            v1[m] = v2[m] = v3[m] = 0.0;
            const bool act = m < 3 ? ua : ((sm >> (m - 3)) & 1);
            if (act && pn) stage_inputs(A, m * NN + n, useG, useK, useC, v1[m], v2[m], v3[m]);
          }
          double raw[9];
          double ab = 0.0, bb2 = 0.0, cc = 0.0, dd = 0.0;
          if (pn) {
            const Stretch st = pml_stretch(g.pml, A.nd.xc[i], A.nd.yc[j], A.nd.zc[k]);
            const Abcd q = pml_abcd(st);
            ab = q.a; bb2 = q.b; cc = q.c; dd = q.d;
          }
          const double rwp = pn ? A.nd.rwp[n] : 0.0;
          const double wps = pn ? A.nd.wps[n] : 0.0;
#pragma unroll
          for (int m = 0; m < 9; ++m) {
            double dg = 0.0;
            if (pn) {
              const double wm = m < 3 ? rwp : (m < 6 ? wps : 2.0 * wps);
              double s = 0.0;
              if (useC) s += bb2 * v3[m];
              if (useK) s += cc * v2[m];
              if (useG) s += dd * v1[m];
              dg = wm * s;
            }
            raw[m] = acc9[m] + dg;
          }
          if (A.stage == 0) {
#pragma unroll
            for (int m = 0; m < 9; ++m) {
              const bool act = m < 3 ? ua : ((sm >> (m - 3)) & 1);
              if (act) A.out[m * NN + n] = raw[m];
            }
          } else {
            double gout[9];
            const double minv_u = A.nd.minv_u[n];
#pragma unroll
            for (int m = 0; m < 9; ++m) {
              double src = 0.0;
              if (A.nsrc_tiles && m < 3) {
                const Source& S = A.src;
                for (int jj = S.tile_ptr[b]; jj < S.tile_ptr[b + 1]; ++jj) {
                  if (S.node[jj] == n && S.comp[jj] == m) {
                    double va = S.a[jj];
                    if (S.a2) va = va - S.a2[jj];
                    double val = S.ca * (va * S.m[jj]);
                    if (S.b) {
                      double vb = S.b[jj];
                      if (S.b2) vb = vb - S.b2[jj];
                      val = val + S.cb * (vb * S.m[jj]);
                    }
                    src += val;
                  }
                }
              }
              double minv;
              if (m < 3) minv = minv_u;
              else minv = 1.0 / ((m < 6 ? 1.0 : 2.0) * ab * wps);
              gout[m] = src - minv * raw[m];
            }
            if (A.stage < 4) {
#pragma unroll
              for (int m = 0; m < 9; ++m) {
                const bool act = m < 3 ? ua : ((sm >> (m - 3)) & 1);
                if (act) A.out[m * NN + n] = gout[m];
              }
            } else {
              // fused RK4 update (Solver_PETSc_Explicit.F90:1038-1046)
              const double hh = A.hh, h = A.h, h6 = A.h6;
#pragma unroll
              for (int m = 0; m < 9; ++m) {
                const bool act = m < 3 ? ua : ((sm >> (m - 3)) & 1);
                if (!act) continue;
                const long long id = m * NN + n;
                const double x1 = A.in[0][id], x2 = A.in[1][id], x3 = A.in[2][id];
                const double g1 = A.in[3][id], g2 = A.in[4][id], g3 = A.in[5][id], g4 = gout[m];
                const double k2t = x2 + hh * x3;
                const double k3t = x2 + hh * (x3 + hh * g1);
                const double k4t = x2 + h * (x3 + hh * g2);
                const double k2m = x3 + hh * g1;
                const double k3m = x3 + hh * g2;
                const double k4m = x3 + h * g3;
                const double ht = 0.0 + (((1.0 * x2 + 2.0 * k2t) + 2.0 * k3t) + 1.0 * k4t);
                const double hm = 0.0 + (((1.0 * x3 + 2.0 * k2m) + 2.0 * k3m) + 1.0 * k4m);
                const double hb = 0.0 + (((1.0 * g1 + 2.0 * g2) + 2.0 * g3) + 1.0 * g4);
                A.xnew[0][id] = x1 + h6 * ht;
                A.xnew[1][id] = x2 + h6 * hm;
                A.xnew[2][id] = x3 + h6 * hb;
              }
            }
          }
        }
      }
    }
    __syncthreads();
  }
}

// Explicit instantiation used by the host launcher.
template __global__ void stage_kernel<4, 4>(const StageArgs A);

// ============================================================ stage kernel v2
// Same operator as stage_kernel; differences:
//  * owner threads are grouped by lattice plane k mod 3, so the plane shared
//    by consecutive slabs stays in the registers of the thread that owns it;
//  * tile classes: HAS_RD / HAS_PML select the LDS arrays a tile needs
//    (regular-domain-only tiles stage v2 only and no PML flux);
//  * blocks walk a tile list; list position is assigned so that each XCD
//    (blocks b, b+8, b+16, ...) processes a contiguous run of neighbouring tiles.

// stage inputs and PML nodal flux of one lattice node
template <bool HAS_PML>
__device__ __forceinline__ void node_inputs(const StageArgs& A, int i, int j, int k, double U9[9], double G9[9]) {
  const Geo& g = A.g;
#pragma unroll
  for (int c = 0; c < 9; ++c) { U9[c] = 0.0; G9[c] = 0.0; }
  if (!in_domain(g, i, j, k)) return;
  const long long NN = g.NN;
  const long long n = (long long)i + (long long)g.NX * ((long long)j + (long long)g.NY * k);
  const bool useG = (A.set_mask & 1), useK = (A.set_mask & 2), useC = (A.set_mask & 4);
  const bool pn = HAS_PML && pml_node(g, i, j, k);
  if (u_active(g, i, j, k)) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      double v1, v2, v3;
      stage_inputs(A, c * NN + n, pn && useG, useK, pn && useC, v1, v2, v3);
      U9[c] = v1; U9[3 + c] = v2; U9[6 + c] = v3;
    }
  }
  if (!pn) return;
  const int sm = s_mask(g, i, j, k);
  double s1[6], s2[6], s3[6];
#pragma unroll
  for (int m = 0; m < 6; ++m) {
    s1[m] = s2[m] = s3[m] = 0.0;
    if (sm & (1 << m)) stage_inputs(A, (3 + m) * NN + n, useG, useK, useC, s1[m], s2[m], s3[m]);
  }
  const Stretch st = pml_stretch(g.pml, A.nd.xc[i], A.nd.yc[j], A.nd.zc[k]);
  double F[3][3];
  pml_f(st, F);
  if (!A.transpose) {
    const double lam = A.nd.lam[n], mu = A.nd.mu[n];
    const double l2m = lam + 2.0 * mu, m2 = 2.0 * mu;
#pragma unroll
    for (int X = 0; X < 3; ++X) {
      if (!(A.set_mask & (1 << X))) continue;
      const double* s = X == 0 ? s1 : (X == 1 ? s2 : s3);
      const double Txx = l2m * s[0] + lam * s[1] + lam * s[2];
      const double Tyy = lam * s[0] + l2m * s[1] + lam * s[2];
      const double Tzz = lam * s[0] + lam * s[1] + l2m * s[2];
      const double Txy = m2 * s[3], Txz = m2 * s[4], Tyz = m2 * s[5];
      const double Tm[9] = {Txx, Txy, Txz, Txy, Tyy, Tyz, Txz, Tyz, Tzz};
#pragma unroll
      for (int ii = 0; ii < 3; ++ii)
#pragma unroll
        for (int kk = 0; kk < 3; ++kk) G9[3 * ii + kk] -= F[X][kk] * Tm[3 * ii + kk];
    }
  } else {
#pragma unroll
    for (int X = 0; X < 3; ++X) {
      if (!(A.set_mask & (1 << X))) continue;
      const double* s = X == 0 ? s1 : (X == 1 ? s2 : s3);
      const double Ym[9] = {s[0], s[3], s[4], s[3], s[1], s[5], s[4], s[5], s[2]};
#pragma unroll
      for (int ii = 0; ii < 3; ++ii)
#pragma unroll
        for (int kk = 0; kk < 3; ++kk) G9[3 * ii + kk] += F[X][kk] * Ym[3 * ii + kk];
    }
  }
}

// Complete node: diagonal terms, M^-1, sources, output or fused RK4 update.
__device__ __forceinline__ void finish_node(const StageArgs& A, int tile, int i, int j, int k, const double acc9[9]) {
  const Geo& g = A.g;
  const long long NN = g.NN;
  const long long n = (long long)i + (long long)g.NX * ((long long)j + (long long)g.NY * k);
  const bool useG = (A.set_mask & 1), useK = (A.set_mask & 2), useC = (A.set_mask & 4);
  const bool pn = pml_node(g, i, j, k);
  const bool ua = u_active(g, i, j, k);
  const int sm = s_mask(g, i, j, k);
  double raw[9];
  double ab = 1.0, wps = 0.0;
  if (pn) {
    const Stretch st = pml_stretch(g.pml, A.nd.xc[i], A.nd.yc[j], A.nd.zc[k]);
    const Abcd q = pml_abcd(st);
    ab = q.a;
    const double rwp = A.nd.rwp[n];
    wps = A.nd.wps[n];
#pragma unroll
    for (int m = 0; m < 9; ++m) {
      const bool act = m < 3 ? ua : ((sm >> (m - 3)) & 1);
      double v1 = 0.0, v2 = 0.0, v3 = 0.0;
      if (act) stage_inputs(A, m * NN + n, useG, useK, useC, v1, v2, v3);
      const double wm = m < 3 ? rwp : (m < 6 ? wps : 2.0 * wps);
      double s = 0.0;
      if (useC) s += q.b * v3;
      if (useK) s += q.c * v2;
      if (useG) s += q.d * v1;
      raw[m] = acc9[m] + wm * s;
    }
  } else {
#pragma unroll
    for (int m = 0; m < 9; ++m) raw[m] = acc9[m];
  }
  if (A.stage == 0) {
#pragma unroll
    for (int m = 0; m < 9; ++m) {
      const bool act = m < 3 ? ua : ((sm >> (m - 3)) & 1);
      if (act) A.out[m * NN + n] = raw[m];
    }
    return;
  }
  const double minv_u = ua ? A.nd.minv_u[n] : 0.0;
  double gout[9];
#pragma unroll
  for (int m = 0; m < 9; ++m) {
    double src = 0.0;
    if (A.nsrc_tiles && m < 3) {
      const Source& S = A.src;
      for (int jj = S.tile_ptr[tile]; jj < S.tile_ptr[tile + 1]; ++jj) {
        if (S.node[jj] == n && S.comp[jj] == m) {
          double va = S.a[jj];
          if (S.a2) va = va - S.a2[jj];
          double val = S.ca * (va * S.m[jj]);
          if (S.b) {
            double vb = S.b[jj];
            if (S.b2) vb = vb - S.b2[jj];
            val = val + S.cb * (vb * S.m[jj]);
          }
          src += val;
        }
      }
    }
    double minv;
    if (m < 3) minv = minv_u;
    else minv = ((sm >> (m - 3)) & 1) ? 1.0 / ((m < 6 ? 1.0 : 2.0) * ab * wps) : 0.0;
    gout[m] = src - minv * raw[m];
  }
  if (A.stage < 4) {
#pragma unroll
    for (int m = 0; m < 9; ++m) {
      const bool act = m < 3 ? ua : ((sm >> (m - 3)) & 1);
      if (act) A.out[m * NN + n] = gout[m];
    }
    return;
  }
  const double hh = A.hh, h = A.h, h6 = A.h6;
#pragma unroll
  for (int m = 0; m < 9; ++m) {
    const bool act = m < 3 ? ua : ((sm >> (m - 3)) & 1);
    if (!act) continue;
    const long long id = m * NN + n;
    const double x1 = A.in[0][id], x2 = A.in[1][id], x3 = A.in[2][id];
    const double g1 = A.in[3][id], g2 = A.in[4][id], g3 = A.in[5][id], g4 = gout[m];
    const double k2t = x2 + hh * x3;
    const double k3t = x2 + hh * (x3 + hh * g1);
    const double k4t = x2 + h * (x3 + hh * g2);
    const double k2m = x3 + hh * g1;
    const double k3m = x3 + hh * g2;
    const double k4m = x3 + h * g3;
    const double ht = 0.0 + (((1.0 * x2 + 2.0 * k2t) + 2.0 * k3t) + 1.0 * k4t);
    const double hm = 0.0 + (((1.0 * x3 + 2.0 * k2m) + 2.0 * k3m) + 1.0 * k4m);
    const double hb = 0.0 + (((1.0 * g1 + 2.0 * g2) + 2.0 * g3) + 1.0 * g4);
    A.xnew[0][id] = x1 + h6 * ht;
    A.xnew[1][id] = x2 + h6 * hm;
    A.xnew[2][id] = x3 + h6 * hb;
  }
}

template <int TX, int TY, bool HAS_RD, bool HAS_PML>
__global__ void __launch_bounds__(STAGE_THREADS, 2) stage_kernel2(const StageArgs A, const int* __restrict__ tiles,
                                                               int ntl) {
  using TS = TileShape<TX, TY>;
  static_assert(3 * TS::OC <= STAGE_THREADS, "owner mapping needs 3 planes x columns threads");
  constexpr int NU = HAS_PML ? 9 : 3;       // staged u components: v1 v2 v3, or v2 only
  constexpr int V2 = HAS_PML ? 3 : 0;       // offset of v2 in sU
  __shared__ double sU[3][NU][TS::WP];
  __shared__ double sG[HAS_PML ? 3 : 1][HAS_PML ? 9 : 1][HAS_PML ? TS::WP : 1];
  __shared__ double sSig[HAS_RD ? TS::NEW : 1][6][27];
  __shared__ int sType[TS::NEW];

  const Geo& g = A.g;
  const int tid = threadIdx.x;
  // XCD-aware list position: XCD x = blockIdx % 8 takes list entries [x*C, (x+1)*C)
  const int C = (ntl + 7) / 8;
  const int li = (blockIdx.x & 7) * C + (blockIdx.x >> 3);
  if (li >= ntl) return;
  const int b = tiles[li];
  const int tx = b % g.tiles[0];
  const int ty = (b / g.tiles[0]) % g.tiles[1];
  const int tz = b / (g.tiles[0] * g.tiles[1]);
  const int ex0 = tx * TX, ey0 = ty * TY, ez0 = tz * g.tz;
  const int ez_end = min(ez0 + g.tz, g.nez);
  const int gi0 = 2 * ex0 - 2, gj0 = 2 * ey0 - 2;
  const int own_i1 = (ex0 + TX >= g.nex) ? g.NX : 2 * (ex0 + TX);
  const int own_j1 = (ey0 + TY >= g.ney) ? g.NY : 2 * (ey0 + TY);
  const int ez_first = max(ez0 - 1, 0);
  const bool useK = (A.set_mask & 2);

  // owner mapping: group P owns lattice planes with k % 3 == P
  const bool owner = tid < 3 * TS::OC;
  const int P = tid / TS::OC, c = tid - P * TS::OC;
  const int ix = c % TS::OX, iy = c / TS::OX;
  const int i = 2 * ex0 + ix, j = 2 * ey0 + iy;
  const bool owned = owner && i < own_i1 && j < own_j1;
  const bool oddx = i & 1, oddy = j & 1;
  double acc[9];
#pragma unroll
  for (int m = 0; m < 9; ++m) acc[m] = 0.0;

  for (int ez = ez_first; ez < ez_end; ++ez) {
    const bool halo = ez < ez0;
    // -------------------------------------------------------------- load
    const int kfirst = (ez == ez_first) ? 2 * ez : 2 * ez + 1;
    const int nload = (2 * ez + 2 - kfirst + 1) * TS::WP;
    for (int t = tid; t < nload; t += STAGE_THREADS) {
      const int pl = t / TS::WP, r = t - pl * TS::WP;
      const int li_ = r % TS::WX, lj = r / TS::WX;
      const int k = kfirst + pl;
      const int slot = k % 3;
      double U9[9], G9[9];
      node_inputs<HAS_PML>(A, gi0 + li_, gj0 + lj, k, U9, G9);
      if constexpr (HAS_PML) {
#pragma unroll
        for (int cc = 0; cc < 9; ++cc) { sU[slot][cc][r] = U9[cc]; sG[slot][cc][r] = G9[cc]; }
      } else {
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) sU[slot][cc][r] = U9[3 + cc];
      }
    }
    for (int t = tid; t < TS::NEW; t += STAGE_THREADS) {
      const int ex = ex0 - 1 + t % (TX + 1), ey = ey0 - 1 + t / (TX + 1);
      int ty_ = 0;
      if (ex >= 0 && ey >= 0 && ex < g.nex && ey < g.ney) ty_ = rd_elem(g, ex, ey, ez) ? 1 : 2;
      sType[t] = ty_;
    }
    __syncthreads();

    // ------------------------------------------- regular-domain W * sigma
    if constexpr (HAS_RD) {
      if (useK) {
        for (int t = tid; t < TS::NEW * 27; t += STAGE_THREADS) {
          const int e = t / 27, q = t - 27 * e;
          if (sType[e] != 1) continue;
          const int lx = q % 3, ly = (q / 3) % 3, lz = q / 9;
          const int li0 = 2 * (e % (TX + 1)), lj0 = 2 * (e / (TX + 1));
          const int rq = (li0 + lx) + TS::WX * (lj0 + ly);
          const int sq = (2 * ez + lz) % 3;
          double d[3][3];
#pragma unroll
          for (int cc = 0; cc < 3; ++cc) {
            const double* Uc0 = &sU[sq][V2 + cc][0];
            d[cc][0] = g.dxi[0] * (DM(lx, 0) * Uc0[rq - lx] + DM(lx, 1) * Uc0[rq - lx + 1] + DM(lx, 2) * Uc0[rq - lx + 2]);
            d[cc][1] = g.dxi[1] * (DM(ly, 0) * Uc0[rq - TS::WX * ly] + DM(ly, 1) * Uc0[rq - TS::WX * ly + TS::WX] +
                                   DM(ly, 2) * Uc0[rq - TS::WX * ly + 2 * TS::WX]);
            double a3 = 0.0;
#pragma unroll
            for (int bb = 0; bb < 3; ++bb) a3 += DM(lz, bb) * sU[(2 * ez + bb) % 3][V2 + cc][rq];
            d[cc][2] = g.dxi[2] * a3;
          }
          const long long n = (long long)(gi0 + li0 + lx) +
                              (long long)g.NX * ((long long)(gj0 + lj0 + ly) + (long long)g.NY * (2 * ez + lz));
          const double lam = A.nd.lam[n], mu = A.nd.mu[n];
          const double W = g.wq[q];
          const double tr = d[0][0] + d[1][1] + d[2][2];
          sSig[e][0][q] = W * (lam * tr + 2.0 * mu * d[0][0]);
          sSig[e][1][q] = W * (lam * tr + 2.0 * mu * d[1][1]);
          sSig[e][2][q] = W * (lam * tr + 2.0 * mu * d[2][2]);
          sSig[e][3][q] = W * (mu * (d[0][1] + d[1][0]));
          sSig[e][4][q] = W * (mu * (d[0][2] + d[2][0]));
          sSig[e][5][q] = W * (mu * (d[1][2] + d[2][1]));
        }
      }
    }
    __syncthreads();

    // ------------------------------------------------------- owner phase
    const int p = ((P - 2 * ez) % 3 + 3) % 3;       // this group's plane within the slab
    const int k = 2 * ez + p;
    if (p != 0 || ez == ez_first) {
#pragma unroll
      for (int m = 0; m < 9; ++m) acc[m] = 0.0;
    }
    const bool top_plane = (ez == g.nez - 1) && p == 2;
    if (owned && (!halo || p == 2)) {
      const int lz = p;
      const int sP = k % 3;
      const bool pn = HAS_PML && pml_node(g, i, j, k);
      double F[3][3];
      if (pn) {
        const Stretch st = pml_stretch(g.pml, A.nd.xc[i], A.nd.yc[j], A.nd.zc[k]);
        pml_f(st, F);
      }
      double aE[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int bb = 0; bb < 2; ++bb) {
          int ex, lx, ey, ly;
          if (oddx) { if (a) continue; ex = (i - 1) >> 1; lx = 1; }
          else if (a == 0) { if (i < 2) continue; ex = (i >> 1) - 1; lx = 2; }
          else { if ((i >> 1) >= g.nex) continue; ex = i >> 1; lx = 0; }
          if (oddy) { if (bb) continue; ey = (j - 1) >> 1; ly = 1; }
          else if (bb == 0) { if (j < 2) continue; ey = (j >> 1) - 1; ly = 2; }
          else { if ((j >> 1) >= g.ney) continue; ey = j >> 1; ly = 0; }
          const int e = (ex - ex0 + 1) + (TX + 1) * (ey - ey0 + 1);
          const int et = sType[e];
          const int li0 = 2 * (ex - ex0 + 1), lj0 = 2 * (ey - ey0 + 1);
          const int rP = (li0 + lx) + TS::WX * (lj0 + ly);
          if constexpr (HAS_RD) {
            if (et == 1 && useK) {
#pragma unroll
              for (int ii = 0; ii < 3; ++ii) {
                double s = 0.0, t2 = 0.0, t3 = 0.0;
#pragma unroll
                for (int qq = 0; qq < 3; ++qq) {
                  s += DM(qq, lx) * sSig[e][sym(ii, 0)][qq + 3 * ly + 9 * lz];
                  t2 += DM(qq, ly) * sSig[e][sym(ii, 1)][lx + 3 * qq + 9 * lz];
                  t3 += DM(qq, lz) * sSig[e][sym(ii, 2)][lx + 3 * ly + 9 * qq];
                }
                acc[ii] += g.dxi[0] * s + g.dxi[1] * t2 + g.dxi[2] * t3;
              }
            }
          }
          if constexpr (HAS_PML) {
            if (et == 2) {
#pragma unroll
              for (int ii = 0; ii < 3; ++ii) {
                double s = 0.0, t2 = 0.0, t3 = 0.0;
#pragma unroll
                for (int qq = 0; qq < 3; ++qq) {
                  s += DM(qq, lx) * (g.wq[qq + 3 * ly + 9 * lz] * sG[sP][3 * ii + 0][rP + (qq - lx)]);
                  t2 += DM(qq, ly) * (g.wq[lx + 3 * qq + 9 * lz] * sG[sP][3 * ii + 1][rP + TS::WX * (qq - ly)]);
                  t3 += DM(qq, lz) * (g.wq[lx + 3 * ly + 9 * qq] * sG[(2 * ez + qq) % 3][3 * ii + 2][rP]);
                }
                acc[ii] += g.dxi[0] * s + g.dxi[1] * t2 + g.dxi[2] * t3;
              }
              double dX[3] = {0, 0, 0}, dY[3] = {0, 0, 0}, dZ[3] = {0, 0, 0};
#pragma unroll
              for (int X = 0; X < 3; ++X) {
                if (!(A.set_mask & (1 << X))) continue;
#pragma unroll
                for (int ii = 0; ii < 3; ++ii) {
                  const double* Uc = &sU[sP][3 * X + ii][0];
                  const double ddx = g.dxi[0] * (DM(lx, 0) * Uc[rP - lx] + DM(lx, 1) * Uc[rP - lx + 1] +
                                                 DM(lx, 2) * Uc[rP - lx + 2]);
                  const double ddy = g.dxi[1] * (DM(ly, 0) * Uc[rP - TS::WX * ly] +
                                                 DM(ly, 1) * Uc[rP - TS::WX * ly + TS::WX] +
                                                 DM(ly, 2) * Uc[rP - TS::WX * ly + 2 * TS::WX]);
                  double a3 = 0.0;
#pragma unroll
                  for (int qq = 0; qq < 3; ++qq) a3 += DM(lz, qq) * sU[(2 * ez + qq) % 3][3 * X + ii][rP];
                  const double ddz = g.dxi[2] * a3;
                  dX[ii] += F[X][0] * ddx;
                  dY[ii] += F[X][1] * ddy;
                  dZ[ii] += F[X][2] * ddz;
                }
              }
              const double W = g.wq[lx + 3 * ly + 9 * lz];
              aE[0] += W * dX[0];
              aE[1] += W * dY[1];
              aE[2] += W * dZ[2];
              aE[3] += W * (dY[0] + dX[1]);
              aE[4] += W * (dZ[0] + dX[2]);
              aE[5] += W * (dZ[1] + dY[2]);
            }
          }
        }
      if (pn) {
        const long long n = (long long)i + (long long)g.NX * ((long long)j + (long long)g.NY * k);
        if (!A.transpose) {
#pragma unroll
          for (int m = 0; m < 6; ++m) acc[3 + m] += aE[m];
        } else {
          const double lam = A.nd.lam[n], mu = A.nd.mu[n];
          const double l2m = lam + 2.0 * mu, m2 = 2.0 * mu;
          acc[3] += -(l2m * aE[0] + lam * aE[1] + lam * aE[2]);
          acc[4] += -(lam * aE[0] + l2m * aE[1] + lam * aE[2]);
          acc[5] += -(lam * aE[0] + lam * aE[1] + l2m * aE[2]);
          acc[6] += -(m2 * aE[3]);
          acc[7] += -(m2 * aE[4]);
          acc[8] += -(m2 * aE[5]);
        }
      }
      if (!(p == 2 && !top_plane) && !halo) finish_node(A, b, i, j, k, acc);
    }
    __syncthreads();
  }
}

#define HR_INST(TX, TY)                                                                                    \
  template __global__ void stage_kernel2<TX, TY, true, true>(const StageArgs, const int* __restrict__, int);  \
  template __global__ void stage_kernel2<TX, TY, true, false>(const StageArgs, const int* __restrict__, int); \
  template __global__ void stage_kernel2<TX, TY, false, true>(const StageArgs, const int* __restrict__, int);
HR_INST(4, 4)
HR_INST(4, 3)
#undef HR_INST

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
                              double* __restrict__ row) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < nrec) row[r] = x2[slot[r]];
}

// displacement of the regular-domain node box -> history slice (3 x nbox)
__global__ void store_box_kernel(Geo g, int bi0, int bj0, int bk0, int bnx, int bny, int bnz,
                                 const double* __restrict__ x2, double* __restrict__ hist) {
  const long long nbox = (long long)bnx * bny * bnz;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < 3 * nbox; t += (long long)gridDim.x * blockDim.x) {
    const int c = (int)(t / nbox);
    const long long r = t - c * nbox;
    const int bi = (int)(r % bnx), bj = (int)((r / bnx) % bny), bk = (int)(r / ((long long)bnx * bny));
    const long long n = (long long)(bi0 + bi) + (long long)g.NX * ((long long)(bj0 + bj) + (long long)g.NY * (bk0 + bk));
    hist[t] = x2[c * g.NN + n];
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
  const long long nbox = (long long)bnx * bny * bnz;
  for (long long r = blockIdx.x * (long long)blockDim.x + threadIdx.x; r < nbox; r += (long long)gridDim.x * blockDim.x) {
    const int bi = (int)(r % bnx), bj = (int)((r / bnx) % bny), bk = (int)(r / ((long long)bnx * bny));
    const int pos[3] = {bi0 + bi, bj0 + bj, bk0 + bk};
    const long long n = (long long)pos[0] + (long long)g.NX * ((long long)pos[1] + (long long)g.NY * pos[2]);
    const long long stride[3] = {1, bnx, (long long)bnx * bny};
    const long long lstride[3] = {1, g.NX, (long long)g.NX * g.NY};
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
              const long long off = (long long)(m - l);
              su += DM(l, m) * uh[c * nbox + r + off * stride[a]];
              sp += DM(l, m) * p2[c * g.NN + n + off * lstride[a]];
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

// Misfit partial sums per step: part[n] = sum_r (trace[n][r] - obs[n][r])^2
__global__ void misfit_rows_kernel(int nstep, int nrec, const double* __restrict__ tr, const double* __restrict__ ob,
                                   double* __restrict__ part) {
  const int nstp = blockIdx.x * blockDim.x + threadIdx.x;
  if (nstp > nstep) return;
  double c = 0.0;
  for (int r = 0; r < nrec; ++r) {
    const double d = tr[(long long)nstp * nrec + r] - ob[(long long)nstp * nrec + r];
    c += d * d;
  }
  part[nstp] = c;
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

hipError_t launch_stage(const StageArgs& A, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL(HIP_KERNEL_NAME(stage_kernel<4, 4>), dim3(nblocks), dim3(STAGE_THREADS), 0, s, A);
  return hipGetLastError();
}
hipError_t launch_stage2(const StageArgs& A, int cls, const int* tiles, int ntl, hipStream_t s) {
  if (ntl <= 0) return hipSuccess;
  const int nb = 8 * ((ntl + 7) / 8);
  const dim3 grid(nb), block(STAGE_THREADS);
#define HR_GO(TX, TY, R, PM) hipLaunchKernelGGL(HIP_KERNEL_NAME(stage_kernel2<TX, TY, R, PM>), grid, block, 0, s, A, tiles, ntl)
  if (A.g.tx != 4 || (A.g.ty != 4 && A.g.ty != 3)) return hipErrorInvalidValue;
  if (A.g.ty == 4) {
    if (cls == 0) HR_GO(4, 4, true, true);
    else if (cls == 1) HR_GO(4, 4, true, false);
    else HR_GO(4, 4, false, true);
  } else {
    if (cls == 0) HR_GO(4, 3, true, true);
    else if (cls == 1) HR_GO(4, 3, true, false);
    else HR_GO(4, 3, false, true);
  }
#undef HR_GO
  return hipGetLastError();
}
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
hipError_t launch_record(int nrec, const long long* slot, const double* x2, double* row, hipStream_t s) {
  hipLaunchKernelGGL(record_kernel, dim3((nrec + 255) / 256), dim3(256), 0, s, nrec, slot, x2, row);
  return hipGetLastError();
}
hipError_t launch_store_box(const Geo& g, const int* b0, const int* bn, const double* x2, double* hist, hipStream_t s) {
  const long long tot = 3LL * bn[0] * bn[1] * bn[2];
  hipLaunchKernelGGL(store_box_kernel, dim3(grid_for(tot)), dim3(256), 0, s, g, b0[0], b0[1], b0[2], bn[0], bn[1],
                     bn[2], x2, hist);
  return hipGetLastError();
}
hipError_t launch_gradient(const Geo& g, const int* b0, const int* bn, const double* uh, const double* p2, double dt,
                           double* glam, double* gmu, hipStream_t s) {
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
hipError_t launch_permute_cols(long long nrows, int ncols, const long long* perm, const double* src, double* dst,
                               hipStream_t s) {
  hipLaunchKernelGGL(permute_cols_kernel, dim3(grid_for(nrows * ncols)), dim3(256), 0, s, nrows, ncols, perm, src, dst);
  return hipGetLastError();
}
hipError_t launch_sum(int n, const double* in, double* out, hipStream_t s) {
  hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(64), 0, s, n, in, out);
  return hipGetLastError();
}
hipError_t launch_misfit(int nstep, int nrec, const double* tr, const double* ob, double dt, double* part, double* out,
                         hipStream_t s) {
  hipLaunchKernelGGL(misfit_rows_kernel, dim3((nstep + 1 + 127) / 128), dim3(128), 0, s, nstep, nrec, tr, ob, part);
  hipLaunchKernelGGL(misfit_sum_kernel, dim3(1), dim3(64), 0, s, nstep, part, dt, out);
  return hipGetLastError();
}
}  // namespace hr
