// Stage kernel v3: one RK4 stage (or one operator apply) of the explicit
// elastic system with the hybrid displacement-stress PML, matrix-free on the
// structured lattice of 27-node GLL hexahedra.
//
// The operator is the one the reference assembles into AK, AC, AG
// (Assemble_PETSc_3D.F90:6-436, element blocks Element_Matrices_3D.F90:7-882)
// and the stage is the VecWAXPY/MatMultAdd sequence of
// Solver_PETSc_Explicit.F90:1001-1046 (forward) and :1319-1364 (adjoint,
// transposed with the signs -G, +K, -C).
//
// Work decomposition (same tiles as the v2 kernel): a workgroup owns TX x TY
// elements in x-y and a chunk of tz element layers, and streams up the chunk
// one element layer (slab) at a time.  A slab's lattice window (owned nodes
// plus the low-side halo element) sits in a three-plane LDS ring.
//
// Differences from v2, all aimed at instruction count:
//  * the stage, the transpose and the tile class are template parameters;
//  * PML stretch factors come from per-axis tables (a_x, b_x depend on x
//    only, ...), the stress-row mass inverse from a node table: no division;
//  * the PML strain rows use F-weighted inputs w_d = sum_X F[X][d] v_X, which
//    is exact algebra because F[X][d] does not vary along the d-stencil
//    (F[X][x] depends on y and z only, ...): 9 derivatives per element side
//    instead of 27;
//  * each owner thread loads its own node, keeps the diagonal terms (and for
//    stage 4 x3 and the partial g-sum) in registers, and writes the x1, x2
//    updates of stage 4 at load time; helpers load the low-side halo;
//  * state vectors keep a node's components contiguous (displacement block
//    [3 n + c], then stress block [SB + 6 n + m], SB = 3 NN rounded up to
//    an even count: a node's six stress components are three aligned 16-byte
//    accesses), so a window row touches
//    about half the cache lines of a component-major layout; 32-bit byte
//    offsets on uniform bases; small constant tables (GLL derivative matrix,
//    weights) in LDS.

#include <hip/hip_runtime.h>
#include <type_traits>
#include "hr_device.h"
#include "hr_split.h"

// Fused multiply-add is allowed here: the parity bar is a tolerance, and FMA
// halves the FP64 instruction count of the stencils.
#pragma clang fp contract(fast)

namespace hr {
namespace {

constexpr int NT3 = 256;
constexpr int SRC_LDS = 64;   // source entries of a tile staged in LDS (more: read from memory)

__device__ __forceinline__ void sto(double* p, unsigned off, double v) {
  *reinterpret_cast<double*>(reinterpret_cast<char*>(p) + off) = v;
}
// two consecutive doubles at a 16-byte aligned offset
__device__ __forceinline__ void ld2(const double* p, unsigned off, double& a, double& b) {
  const double2 t = *reinterpret_cast<const double2*>(reinterpret_cast<const char*>(p) + off);
  a = t.x;
  b = t.y;
}
// three consecutive components of a node (8-byte aligned): one 16-byte and one
// 8-byte access instead of three 8-byte ones
struct alignas(8) D3 {
  double v[3];
};
__device__ __forceinline__ void ld3(const double* p, unsigned off, double& a, double& b, double& c) {
  const D3 t = *reinterpret_cast<const D3*>(reinterpret_cast<const char*>(p) + off);
  a = t.v[0];
  b = t.v[1];
  c = t.v[2];
}
__device__ __forceinline__ void st3(double* p, unsigned off, double a, double b, double c) {
  D3 t;
  t.v[0] = a;
  t.v[1] = b;
  t.v[2] = c;
  *reinterpret_cast<D3*>(reinterpret_cast<char*>(p) + off) = t;
}
// six consecutive components at a 16-byte aligned offset: three 16-byte stores
__device__ __forceinline__ void st6(double* p, unsigned off, const double* v) {
  char* q = reinterpret_cast<char*>(p) + off;
#pragma unroll
  for (int a = 0; a < 3; ++a) *reinterpret_cast<double2*>(q + 16 * a) = make_double2(v[2 * a], v[2 * a + 1]);
}
// stores of three consecutive components, each only where its bit of act is set
__device__ __forceinline__ void st3m(double* p, unsigned off, int act, double a, double b, double c) {
  if (act == 7) {
    st3(p, off, a, b, c);
  } else {
    if (act & 1) sto(p, off, a);
    if (act & 2) sto(p, off + 8u, b);
    if (act & 4) sto(p, off + 16u, c);
  }
}

__device__ __forceinline__ bool in_dom3(const Geo3& g, int i, int j, int k) {
  return i >= 0 && j >= 0 && k >= 0 && i < g.NX && j < g.NY && k < g.NZ;
}
__device__ __forceinline__ bool u_act3(const Geo3& g, int i, int j, int k) {
  return !(i == 0 || j == 0 || k == 0 || i == g.NX - 1 || j == g.NY - 1);
}
// node touched by at least one PML element
__device__ __forceinline__ bool pml3(const Geo3& g, int i, int j, int k) {
  return !(i > 2 * g.rlo[0] && i < 2 * g.rhi[0] && j > 2 * g.rlo[1] && j < 2 * g.rhi[1] && k > 2 * g.rlo[2]);
}
// active stress components (bit m) of a PML node
__device__ __forceinline__ int smask3(const Geo3& g, int i, int j, int k) {
  if (g.top_rule && k == g.NZ - 1 && !((i & 1) && (j & 1))) return 0x3f & ~((1 << 2) | (1 << 4) | (1 << 5));
  return 0x3f;
}
__device__ __forceinline__ bool rd3(const Geo3& g, int ex, int ey, int ez) {
  return ex >= g.rlo[0] && ex < g.rhi[0] && ey >= g.rlo[1] && ey < g.rhi[1] && ez >= g.rlo[2];
}
__device__ __forceinline__ int sym3(int i, int k) { return i == k ? i : i + k + 2; }

// 3-point GLL rule on xi = -1, 0, 1: weights and DM(a, c) = l_c'(xi_a)
__device__ __forceinline__ constexpr double gll_w(int a) { return a == 1 ? 4.0 / 3.0 : 1.0 / 3.0; }
__device__ __forceinline__ constexpr double gll_dm(int a, int c) {
  return c == 0 ? (a - 1) - 0.5 : (c == 1 ? -2.0 * (a - 1) : (a - 1) + 0.5);
}

// One axis of the separable PML pull at lattice index i (nel elements along
// the axis; node i is local point l_e of the one or two elements e that hold
// it).  c[d] = sum_e w(a) DM(a, l_e) with a = l_e + d, d = -2..2 (index d + 2):
// the coefficient of the value at i + d in the transposed derivative
// sum_a DM(a, l) w(a) v(a).  By the summation-by-parts identity of the GLL rule
// (w_a DM(a, b) + w_b DM(b, a) = 0 unless a = b = 0 or 2), the forward
// derivative weighted by w(l_e), sum_e w(l_e) sum_c DM(l_e, c) v(c), has the
// coefficients -c[d] for d != 0 and c[0] at d = 0.  m = sum_e w(l_e);
// cov[d]: node i + d belongs to one of the elements.
__device__ __forceinline__ void axis_stencil(int i, int nel, double c[5], double& m, bool cov[5]) {
  m = 0.0;
#pragma unroll
  for (int d = 0; d < 5; ++d) { c[d] = 0.0; cov[d] = false; }
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    int l;
    if (i & 1) { if (s) continue; l = 1; }
    else if (s == 0) { if (i < 2) continue; l = 2; }
    else { if ((i >> 1) >= nel) continue; l = 0; }
    m += gll_w(l);
#pragma unroll
    for (int d = 0; d < 5; ++d) {
      const int a = l + d - 2;
      if (a >= 0 && a <= 2) { c[d] += gll_w(a) * gll_dm(a, l); cov[d] = true; }
    }
  }
}

// Stage-S inputs (v1, v2, v3) of one field component from its raw inputs X
// (x1 x2 x3 g1 g2 g3; apply: x) in the reference's nested VecWAXPY order.
// NEG: adjoint signs on v1, v3.  S = 5 is one step of the central-difference
// variant (X = u_n, u_{n-1}, u_his; hh = dt^2, h = dt, h6 = dt^3): the G, K
// and C inputs whose operator sum, scaled by -M^-1, is that step's update.
template <int S, bool NEG>
__device__ __forceinline__ void combine(const double* X, double hh, double h, double h6, int mask, double& v1,
                                        double& v2, double& v3) {
  if constexpr (S == 0) {
    v1 = (mask & 1) ? X[0] : 0.0;
    v2 = (mask & 2) ? X[0] : 0.0;
    v3 = (mask & 4) ? X[0] : 0.0;
    return;
  } else if constexpr (S == 1) {
    v1 = X[0]; v2 = X[1]; v3 = X[2];
  } else if constexpr (S == 2) {
    v1 = X[0] + hh * X[1]; v2 = X[1] + hh * X[2]; v3 = X[2] + hh * X[3];
  } else if constexpr (S == 3) {
    v1 = X[0] + hh * (X[1] + hh * X[2]); v2 = X[1] + hh * (X[2] + hh * X[3]); v3 = X[2] + hh * X[4];
  } else if constexpr (S == 5) {
    v1 = h6 * X[0] + hh * X[2]; v2 = hh * X[0]; v3 = h * (X[0] - X[1]);
  } else {
    v1 = X[0] + h * (X[1] + hh * (X[2] + hh * X[3])); v2 = X[1] + h * (X[2] + hh * X[4]); v3 = X[2] + h * X[5];
  }
  if constexpr (NEG) { v1 = -v1; v3 = -v3; }
}

// raw input q enters the K-set input v2 of stage S
template <int S>
__device__ __forceinline__ constexpr bool v2_input(int q) {
  return S == 0 ? q == 0 : (S == 1 ? q == 1 : (S == 2 ? (q == 1 || q == 2) : (S == 3 ? (q >= 1 && q <= 3) : (S == 5 ? q == 0 : (q == 1 || q == 2 || q == 4)))));
}

// Stage-4 update of x1 and x2 (no g4 needed), and the x3 carry
// c3 = x3 + dt/6 (g1 + 2 g2 + 2 g3), so that x3_new = c3 + dt/6 g4
// (Solver_PETSc_Explicit.F90:1036-1046)
__device__ __forceinline__ void rk4_update12(const double* X, double hh, double h, double h6, double& xn0, double& xn1,
                                             double& c3) {
  const double x1 = X[0], x2 = X[1], x3v = X[2], g1 = X[3], g2 = X[4], g3 = X[5];
  const double k2t = x2 + hh * x3v, k3t = x2 + hh * (x3v + hh * g1), k4t = x2 + h * (x3v + hh * g2);
  const double k2m = x3v + hh * g1, k3m = x3v + hh * g2, k4m = x3v + h * g3;
  const double ht = 0.0 + (((1.0 * x2 + 2.0 * k2t) + 2.0 * k3t) + 1.0 * k4t);
  const double hm = 0.0 + (((1.0 * x3v + 2.0 * k2m) + 2.0 * k3m) + 1.0 * k4m);
  xn0 = x1 + h6 * ht;
  xn1 = x2 + h6 * hm;
  c3 = x3v + h6 * ((1.0 * g1 + 2.0 * g2) + 2.0 * g3);
}

template <int TX, int TY>
struct Shape3 {
  static constexpr int WX = 2 * TX + 3, WY = 2 * TY + 3, WP = WX * WY;   // window (low-side halo element)
  static constexpr int NEW = (TX + 1) * (TY + 1);                         // window elements
  static constexpr int OX = 2 * TX + 1, OY = 2 * TY + 1, OC = OX * OY;    // owner columns
  static constexpr int HP = WP - OC;                                      // halo positions per plane
};

// Occupancy hints: LDS allows two blocks per CU of the mixed and PML-only
// classes and four of the regular-domain class, but asking for one more (the
// mixed and regular-domain classes; the PML-only class at stages 1-3) gives
// register allocations that measured 6 %, 1.5 % and 2 % faster in the whole
// gradient (the compiler reports the unmet targets).  At stage 4 the PML-only
// class would then need more than 256 registers.
// SPLIT (regular-domain class only): the kernel also runs the tiles that hold
// PML elements, skips them, and leaves the PML nodes to the nodal kernels: their
// stage-4 x1, x2 updates are not written here and their regular-domain K sums go
// to A.E instead of being finished.
template <int TX, int TY, bool HAS_RD, bool HAS_PML, int S, bool TR, bool SPLIT = false>
__global__ void __launch_bounds__(NT3, HAS_RD ? (HAS_PML ? 3 : 5) : (S >= 1 && S <= 3 ? 3 : 2)) stage_kernel3(const Stage3Args A, const int* __restrict__ tiles, int ntl) {
  static_assert(!SPLIT || (HAS_RD && !HAS_PML && S >= 1 && S <= 4), "split form: regular-domain RK4 stages");
  using SH = Shape3<TX, TY>;
  constexpr int WX = SH::WX, WP = SH::WP, NEW = SH::NEW, OX = SH::OX, OC = SH::OC, HP = SH::HP;
  static_assert(OC <= 64 && NT3 == 256, "one owner plane group per wavefront, plus one helper wavefront");
  constexpr bool NEG = (S > 0) && TR;
  constexpr bool UPD = S == 4 || S == 5;                   // the stage writes the new state
  constexpr int NS = UPD ? (HAS_PML ? 9 : 3) : 1;          // register carry width of the update
  // LDS ring of lattice planes.  With five slots the two planes that the next
  // slab loads never hold a plane of the current slab, so a wavefront that has
  // finished the current slab can start loading the next one without waiting
  // for the others (no end-of-slab barrier).  Only the PML class uses five: the
  // mixed class would not fit twice per CU, and the regular-domain class runs
  // faster at four blocks per CU with three slots than at three with five.
  constexpr int RING = (HAS_PML && !HAS_RD) ? 5 : 3;

  __shared__ double sV[HAS_RD ? RING : 1][HAS_RD ? 3 : 1][HAS_RD ? WP : 1];     // K input of u (v2)
  __shared__ double sW[HAS_PML ? RING : 1][HAS_PML ? 9 : 1][HAS_PML ? WP : 1];  // F-weighted u inputs, 3*d + i
  __shared__ double sG[HAS_PML ? RING : 1][HAS_PML ? 9 : 1][HAS_PML ? WP : 1];  // nodal PML flux, 3*i + k
  __shared__ double sLM[HAS_RD ? RING : 1][2][HAS_RD ? WP : 1];                  // lam, mu of the window nodes
  __shared__ double sSig[HAS_RD ? NEW : 1][6][27];                            // regular domain W * sigma
  __shared__ double sWq[27];
  __shared__ double sDM[9];
  __shared__ long long sSrcNode[SRC_LDS];
  __shared__ int sSrcComp[SRC_LDS];

  const Geo3& g = A.g;
  const int tid = threadIdx.x;
  // XCD-aware position in the (shot, tile) list: XCD x = blockIdx % 8 takes entries [x*C, (x+1)*C)
  const int ntot = ntl * A.nshot;
  const int C = (ntot + 7) / 8;
  const int li = (blockIdx.x & 7) * C + (blockIdx.x >> 3);
  if (li >= ntot) return;
  const int shot = li / ntl;
  const int b = tiles[li - shot * ntl];
  // this shot's state vectors (uniform pointers)
  const long long so = (long long)shot * A.shot_stride;
  const double* inp[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) inp[q] = A.in[q] ? A.in[q] + so : nullptr;
  double* const outp = A.out ? A.out + so : nullptr;
  double* xnp[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) xnp[q] = A.xnew[q] ? A.xnew[q] + so : nullptr;
  const int tx = b % g.tiles0;
  const int ty = (b / g.tiles0) % g.tiles1;
  const int tz = b / (g.tiles0 * g.tiles1);
  const int ex0 = tx * TX, ey0 = ty * TY, ez0 = g.zb[tz];
  // this (shot, tile)'s source entries [src_j0, src_j1), read once per block (uniform)
  int src_j0 = 0, src_j1 = 0, sdyn = 0;
  if (A.src_mode) {
    const __attribute__((address_space(4))) int* tptr = (const __attribute__((address_space(4))) int*)A.src.tile_ptr;
    const int tp = shot * A.src.ntiles + b;
    src_j0 = tptr[tp];
    src_j1 = tptr[tp + 1];
    // captured time loops: the step (rows) are offsets from a device-side step index
    if (A.step_dev) sdyn = *A.step_dev;
  }
  const int ez_end = g.zb[tz + 1];
  const int gi0 = 2 * ex0 - 2, gj0 = 2 * ey0 - 2;
  const int own_i1 = (ex0 + TX >= g.nex) ? g.NX : 2 * (ex0 + TX);
  const int own_j1 = (ey0 + TY >= g.ney) ? g.NY : 2 * (ey0 + TY);
  // split form: no regular-domain element lies below layer rlo[2], so the slabs
  // under it hold no work for this kernel (their nodes are PML nodes)
  const int ez_first = SPLIT ? max(max(ez0 - 1, 0), g.rlo[2]) : max(ez0 - 1, 0);
  const int mask = (S == 0) ? A.set_mask : 7;
  const bool useG = mask & 1, useK = mask & 2, useC = mask & 4;
  const unsigned NN8 = (unsigned)g.NN * 8u;   // bytes of one double per lattice node
  // byte offset of the stress block of a state vector: 3 NN doubles rounded up to an
  // even count, so that a node's six stress components are 16-byte aligned
  const unsigned SB = 8u * (3u * (unsigned)g.NN + ((unsigned)g.NN & 1u));
  const double* __restrict__ tab = A.tab;
  const double* __restrict__ nodet = A.node;

  // the tile's source entries (lattice node, component) in LDS, so that the
  // finish of every owned node scans them without global loads (the other
  // fields are read for matching entries only)
  const int nsrc = src_j1 - src_j0;
  const bool src_lds = nsrc <= SRC_LDS;
  if (src_lds) {
    for (int t = tid; t < nsrc; t += NT3) {
      sSrcNode[t] = A.src.node[src_j0 + t];
      sSrcComp[t] = A.src.comp[src_j0 + t];
    }
  }
  if (tid < 27) sWq[tid] = tab[tid];
  if (tid < 9) {
    const int a = tid / 3, c = tid - 3 * a;   // DM(a, c) = l_c'(xi_a)
    sDM[tid] = (c == 0) ? (a == 0 ? -1.5 : (a == 1 ? -0.5 : 0.5))
                        : (c == 1 ? (a == 0 ? 2.0 : (a == 1 ? 0.0 : -2.0)) : (a == 0 ? -0.5 : (a == 1 ? 0.5 : 1.5)));
  }

  // owner mapping: wavefront P < 3 is the group that owns the lattice planes with
  // k % 3 == P (lanes oc < OC); wavefront 3 loads halo positions.  Keeping a
  // group inside one wavefront means that a wavefront runs one load path per slab.
  const int P = tid >> 6, oc = tid & 63;
  const bool owner = P < 3 && oc < OC;
  const int ix = oc % OX, iy = oc / OX;
  const int i = 2 * ex0 + ix, j = 2 * ey0 + iy;
  const bool owned = owner && i < own_i1 && j < own_j1;
  const int rown = (ix + 2) + WX * (iy + 2);
  const bool oddx = i & 1, oddy = j & 1;

  // PML-only tiles: every element around an owned node is a PML element, whose
  // weights W(a, b, c) = det J w(a) w(b) w(c) factor per axis and whose inputs
  // (nodal flux G, F-weighted w) are nodal.  The pull over the up to four x-y
  // elements of a slab then separates: the x terms of the elements that share a
  // y position are equal up to w(l_y), so they combine into one 5-point x
  // stencil scaled by m_y = sum w(l_y), and likewise for y; the z terms of all
  // of them are equal up to w(l_x) w(l_y).  Each owner keeps its stencil
  // coefficients (scaled by det J dxi m) and LDS positions in registers;
  // positions outside the elements read the node itself with coefficient 0.
  double cxs[HAS_PML && !HAS_RD ? 5 : 1], cys[HAS_PML && !HAS_RD ? 5 : 1], azs = 0.0;
  int pxs[HAS_PML && !HAS_RD ? 5 : 1], pys[HAS_PML && !HAS_RD ? 5 : 1];
  if constexpr (HAS_PML && !HAS_RD) {
    double cx[5], cy[5], mx, my;
    bool vx[5], vy[5];
    axis_stencil(i, g.nex, cx, mx, vx);
    axis_stencil(j, g.ney, cy, my, vy);
    const double fx = g.detj * g.dxi[0] * my, fy = g.detj * g.dxi[1] * mx;
#pragma unroll
    for (int d = 0; d < 5; ++d) {
      cxs[d] = fx * cx[d];
      cys[d] = fy * cy[d];
      pxs[d] = rown + (vx[d] ? d - 2 : 0);
      pys[d] = rown + (vy[d] ? WX * (d - 2) : 0);
    }
    azs = g.detj * g.dxi[2] * mx * my;
  }

  // PML stretch factors of the owner column (x and y depend on i and j only)
  double col_ax = 0.0, col_bx = 0.0, col_ay = 0.0, col_by = 0.0;
  if constexpr (HAS_PML) {
    if (owner && i < g.NX && j < g.NY) {
      ld2(tab, 8u * (unsigned)(TAB_STRETCH + 2 * i), col_ax, col_bx);
      ld2(tab, 8u * (unsigned)(TAB_STRETCH + 2 * g.NX + 2 * j), col_ay, col_by);
    }
  }

  // owner registers: accumulators, diagonal terms, stage-4 carry
  // c3o: stage 4 x3 + dt/6 (g1 + 2 g2 + 2 g3); stage 5 2 u_n - u_{n-1}.  chs: stage 5 u_his
  // acc starts at the node's diagonal terms (set when the owner loads its node)
  double acc[9], c3o[NS], chs[S == 5 ? NS : 1];
  double lam_o = 0.0, mu_o = 0.0, minvu_o = 0.0, minvs_o = 0.0;
#pragma unroll
  for (int m = 0; m < 9; ++m) acc[m] = 0.0;
#pragma unroll
  for (int m = 0; m < NS; ++m) c3o[m] = 0.0;
#pragma unroll
  for (int m = 0; m < (S == 5 ? NS : 1); ++m) chs[m] = 0.0;

  // Load one lattice node into the window ring; OWN: also the owner registers.
  // The loads of a node class are issued unconditionally, so that they leave
  // in one batch: inactive lattice entries are zero in every state vector and
  // never written, and contribute nothing.  Activity only gates the stores.
  constexpr int NIN = (S == 0) ? 1 : (S == 5 ? 3 : 2 + S);   // x1 x2 x3 g1 g2 g3 prefix (apply: x)
  // COL: the node is in the lane's owner column (ii, jj = i, j) and its plane kk is
  // the same for the whole wavefront
  auto load = [&](auto own_tag, auto col_tag, int ii, int jj, int kk, int r, int slot, bool own_plane) {
    constexpr bool OWN = decltype(own_tag)::value;
    constexpr bool COL = decltype(col_tag)::value;
    constexpr bool ALL_U = HAS_PML || (OWN && UPD);      // v1 and v3 of u, or the state update
    if constexpr (OWN) {
      // a newly loaded owned node starts its sums at its diagonal terms (zero off the PML)
#pragma unroll
      for (int m = 0; m < 9; ++m) acc[m] = 0.0;
    }
    double vu2[3] = {0.0, 0.0, 0.0};
    double lmv[2] = {0.0, 0.0};
    double wv[9], Gv[9];
#pragma unroll
    for (int m = 0; m < 9; ++m) { wv[m] = 0.0; Gv[m] = 0.0; }
    if (in_dom3(g, ii, jj, kk)) {
      const unsigned n = (unsigned)ii + (unsigned)g.NX * ((unsigned)jj + (unsigned)g.NY * (unsigned)kk);
      const unsigned offp = n * 16u;               // node table: (lam, mu)
      const unsigned offq = 2u * NN8 + n * 32u;    // node table: (minv_u, minv_s, rwp, wps)
      const unsigned offu = n * 24u;               // displacement block of the state vectors
      const unsigned offs = SB + n * 48u;          // stress block
      const bool ua = u_act3(g, ii, jj, kk);
      const bool pn = HAS_PML && pml3(g, ii, jj, kk);
      // ---- u components: raw inputs
      double XU[3][NIN];
#pragma unroll
      for (int q = 0; q < NIN; ++q) {
        if (ALL_U || v2_input<S>(q)) ld3(inp[q], offu, XU[0][q], XU[1][q], XU[2][q]);
        else XU[0][q] = XU[1][q] = XU[2][q] = 0.0;
      }
      if constexpr (HAS_RD) {
        ld2(nodet, offp, lmv[0], lmv[1]);
      }
      double minvs_l = 0.0;
      if constexpr (OWN && S > 0) ld2(nodet, offq, minvu_o, minvs_l);
      double v1u[3], v3u[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) combine<S, NEG>(XU[c], A.hh, A.h, A.h6, mask, v1u[c], vu2[c], v3u[c]);
      if constexpr (OWN && S == 5) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          c3o[c] = 2.0 * XU[c][0] + (-1.0) * XU[c][1];
          chs[c] = XU[c][2];
        }
      }
      if constexpr (OWN && S == 4) {
        if (own_plane && ua && !(SPLIT && pml3(g, ii, jj, kk))) {
          double xn0[3], xn1[3];
#pragma unroll
          for (int c = 0; c < 3; ++c) rk4_update12(XU[c], A.hh, A.h, A.h6, xn0[c], xn1[c], c3o[c]);
          st3(xnp[0], offu, xn0[0], xn0[1], xn0[2]);
          st3(xnp[1], offu, xn1[0], xn1[1], xn1[2]);
        }
      }
      if constexpr (HAS_PML) {
        // Only PML nodes (all window nodes of a PML-class tile): at any other node
        // the stress components are zero (so G = 0), the diagonal terms vanish
        // (b = 0) and w is never read, as PML elements touch PML nodes only.  The
        // owner registers set here are read for PML nodes only.
        if (HAS_RD ? pn : true) {
          // ---- stretch factors, material
          double ax, bx, ay, by, az, bz;
          if constexpr (COL) {
            // x, y factors of the owner column from registers; z factors of the
            // wavefront's plane with one uniform load
            ax = col_ax; bx = col_bx; ay = col_ay; by = col_by;
            const int ku = __builtin_amdgcn_readfirstlane(kk);
            const __attribute__((address_space(4))) double* tz_ =
                (const __attribute__((address_space(4))) double*)tab + (TAB_STRETCH + 2 * g.NX + 2 * g.NY + 2 * ku);
            az = tz_[0];
            bz = tz_[1];
          } else {
            ld2(tab, 8u * (unsigned)(TAB_STRETCH + 2 * ii), ax, bx);
            ld2(tab, 8u * (unsigned)(TAB_STRETCH + 2 * g.NX + 2 * jj), ay, by);
            ld2(tab, 8u * (unsigned)(TAB_STRETCH + 2 * g.NX + 2 * g.NY + 2 * kk), az, bz);
          }
          double lam, mu;
          if constexpr (HAS_RD) { lam = lmv[0]; mu = lmv[1]; }
          else ld2(nodet, offp, lam, mu);
          double rwp = 0.0, wps = 0.0;
          if constexpr (OWN) {
            ld2(nodet, offq + 16u, rwp, wps);
            if constexpr (S > 0) minvs_o = pn ? minvs_l : 0.0;
          }
          // F[set][k], set 0 = w (G, v1), 1 = p (K, v2), 2 = e (C, v3)
          double F[3][3];
          F[2][0] = ay * az; F[2][1] = ax * az; F[2][2] = ax * ay;
          F[1][0] = ay * bz + az * by; F[1][1] = ax * bz + az * bx; F[1][2] = ax * by + ay * bx;
          F[0][0] = by * bz; F[0][1] = bx * bz; F[0][2] = bx * by;
#pragma unroll
          for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
              double sacc = 0.0;
              if (useG) sacc += F[0][d] * v1u[c];
              if (useK) sacc += F[1][d] * vu2[c];
              if (useC) sacc += F[2][d] * v3u[c];
              wv[3 * d + c] = sacc;
            }
          // diagonal terms rho W (b v3 + c v2 + d v1) of ELEMENT_EM/EN (PML.F90 items 7..10)
          const double qb = ax * ay * bz + ax * az * by + ay * az * bx;
          const double qc = ax * by * bz + ay * bx * bz + az * bx * by;
          const double qd = bx * by * bz;
          auto diag = [&](double wm, double a1, double a2, double a3) {
            double sacc = 0.0;
            if (useC) sacc += qb * a3;
            if (useK) sacc += qc * a2;
            if (useG) sacc += qd * a1;
            return wm * sacc;
          };
          if constexpr (OWN) {
#pragma unroll
            for (int c = 0; c < 3; ++c) acc[c] = diag(rwp, v1u[c], vu2[c], v3u[c]);
            lam_o = lam;
            mu_o = mu;
          }
          const double l2m = lam + 2.0 * mu, m2 = 2.0 * mu;
          const int sm = smask3(g, ii, jj, kk);
          // ---- stress components in two halves (normal, shear): G splits the same way.
          // The node's six stress components of each input: three 16-byte loads
          double XS6[6][NIN];
#pragma unroll
          for (int q = 0; q < NIN; ++q) {
            ld2(inp[q], offs, XS6[0][q], XS6[1][q]);
            ld2(inp[q], offs + 16u, XS6[2][q], XS6[3][q]);
            ld2(inp[q], offs + 32u, XS6[4][q], XS6[5][q]);
          }
#pragma unroll
          for (int hf = 0; hf < 2; ++hf) {
            const double(*XS)[NIN] = &XS6[3 * hf];
            double s1[3], s2[3], s3[3];
#pragma unroll
            for (int t = 0; t < 3; ++t) combine<S, NEG>(XS[t], A.hh, A.h, A.h6, mask, s1[t], s2[t], s3[t]);
#pragma unroll
            for (int X = 0; X < 3; ++X) {
              if (!(mask & (1 << X))) continue;
              const double* sv = X == 0 ? s1 : (X == 1 ? s2 : s3);
              if (hf == 0) {
                if constexpr (!TR) {
                  // T(s): stress of the auxiliary strain (ELEMENT_EA_3D lower block), normal part
                  const double Txx = l2m * sv[0] + lam * sv[1] + lam * sv[2];
                  const double Tyy = lam * sv[0] + l2m * sv[1] + lam * sv[2];
                  const double Tzz = lam * sv[0] + lam * sv[1] + l2m * sv[2];
                  Gv[0] -= F[X][0] * Txx; Gv[4] -= F[X][1] * Tyy; Gv[8] -= F[X][2] * Tzz;
                } else {
                  Gv[0] += F[X][0] * sv[0]; Gv[4] += F[X][1] * sv[1]; Gv[8] += F[X][2] * sv[2];
                }
              } else {
                if constexpr (!TR) {
                  const double Txy = m2 * sv[0], Txz = m2 * sv[1], Tyz = m2 * sv[2];
                  Gv[1] -= F[X][1] * Txy; Gv[2] -= F[X][2] * Txz; Gv[3] -= F[X][0] * Txy;
                  Gv[5] -= F[X][2] * Tyz; Gv[6] -= F[X][0] * Txz; Gv[7] -= F[X][1] * Tyz;
                } else {
                  Gv[1] += F[X][1] * sv[0]; Gv[2] += F[X][2] * sv[1]; Gv[3] += F[X][0] * sv[0];
                  Gv[5] += F[X][2] * sv[2]; Gv[6] += F[X][0] * sv[1]; Gv[7] += F[X][1] * sv[2];
                }
              }
            }
            if constexpr (OWN) {
              const double wm = hf == 0 ? wps : 2.0 * wps;
#pragma unroll
              for (int t = 0; t < 3; ++t) acc[3 + 3 * hf + t] = diag(wm, s1[t], s2[t], s3[t]);
              if constexpr (S == 5 && NS == 9) {
#pragma unroll
                for (int t = 0; t < 3; ++t) {
                  c3o[3 + 3 * hf + t] = 2.0 * XS[t][0] + (-1.0) * XS[t][1];
                  chs[3 + 3 * hf + t] = XS[t][2];
                }
              }
              if constexpr (S == 4) {
                if (own_plane) {
                  double xn0[3], xn1[3];
#pragma unroll
                  for (int t = 0; t < 3; ++t) {
                    double c3v;
                    rk4_update12(XS[t], A.hh, A.h, A.h6, xn0[t], xn1[t], c3v);
                    if constexpr (NS == 9) c3o[3 + 3 * hf + t] = c3v;
                  }
                  const int act = (sm >> (3 * hf)) & 7;
                  if (act) {
                    st3m(xnp[0], offs + 24u * hf, act, xn0[0], xn0[1], xn0[2]);
                    st3m(xnp[1], offs + 24u * hf, act, xn1[0], xn1[1], xn1[2]);
                  }
                }
              }
            }
          }
        }
      }
    }
    if constexpr (HAS_RD) {
#pragma unroll
      for (int c = 0; c < 3; ++c) sV[slot][c][r] = vu2[c];
      sLM[slot][0][r] = lmv[0];
      sLM[slot][1][r] = lmv[1];
    }
    if constexpr (HAS_PML) {
#pragma unroll
      for (int m = 0; m < 9; ++m) { sW[slot][m][r] = wv[m]; sG[slot][m][r] = Gv[m]; }
    }
  };

  for (int ez = ez_first; ez < ez_end; ++ez) {
    const bool halo = ez < ez0;
    const bool first = ez == ez_first;
    const int kfirst = first ? 2 * ez : 2 * ez + 1;
    const int p = ((P - 2 * ez) % 3 + 3) % 3;       // this group's plane within the slab
    const int k = 2 * ez + p;
    // ------------------------------------------------------------ loads
    if (owner && k >= kfirst) {
      const bool own_plane = k >= 2 * ez0 && (k < 2 * ez_end || k == g.NZ - 1);
      if (owned) load(std::true_type{}, std::true_type{}, i, j, k, rown, k % RING, own_plane);
      else load(std::false_type{}, std::true_type{}, i, j, k, rown, k % RING, false);
    }
    {
      // Halo positions of a plane: rows 0, 1 of the window, then columns 0, 1 of
      // the rows above.  The separable pull of PML-only tiles reads only the row
      // and the column of an owned node, so there the four corner positions
      // (rows 0, 1 x columns 0, 1) are skipped, and of the first slab of a chunk
      // above another (a halo slab, whose only owners are those of its top plane)
      // only the top plane's ring is loaded.
      constexpr bool SEP = HAS_PML && !HAS_RD;
      constexpr int HPL = SEP ? HP - 4 : HP;          // halo positions loaded per plane
      const int pl0 = (SEP && halo) ? 2 : 0;          // first plane with its ring loaded
      const int nnew = 2 * ez + 3 - kfirst - pl0;
      int hr_, nh;
      if (first) {
        // all three groups load a plane: wavefront 3 and the spare lanes take the halo
        hr_ = P == 3 ? oc : (oc >= OC ? 64 + P * (64 - OC) + (oc - OC) : -1);
        nh = 64 + 3 * (64 - OC);
      } else {
        // the group whose plane is kept from the last slab helps wavefront 3
        const int idleP = (2 * ez) % 3;
        hr_ = P == 3 ? oc : (P == idleP ? 64 + oc : -1);
        nh = 128;
      }
      if (hr_ >= 0) {
        for (int hx = hr_; hx < nnew * HPL; hx += nh) {
          const int pl = hx / HPL, q = hx - pl * HPL;
          int lx_, ly_;
          if constexpr (SEP) {
            if (q < 2 * (WX - 2)) { ly_ = q / (WX - 2); lx_ = 2 + q - ly_ * (WX - 2); }
            else { const int q2 = q - 2 * (WX - 2); ly_ = 2 + (q2 >> 1); lx_ = q2 & 1; }
          } else {
            if (q < 2 * WX) { ly_ = q / WX; lx_ = q - ly_ * WX; }
            else { const int q2 = q - 2 * WX; ly_ = 2 + (q2 >> 1); lx_ = q2 & 1; }
          }
          const int kk = kfirst + pl0 + pl;
          load(std::false_type{}, std::false_type{}, gi0 + lx_, gj0 + ly_, kk, lx_ + WX * ly_, kk % RING, false);
        }
      }
    }
    __syncthreads();

    // ------------------------------------------- regular-domain W * sigma
    if constexpr (HAS_RD) {
      if (useK) {
        for (int t = tid; t < NEW * 27; t += NT3) {
          const int e = t / 27, q = t - 27 * e;
          {
            const int ex = ex0 - 1 + e % (TX + 1), ey = ey0 - 1 + e / (TX + 1);
            if (ex < 0 || ey < 0 || ex >= g.nex || ey >= g.ney || !rd3(g, ex, ey, ez)) continue;
          }
          const int lx = q % 3, ly = (q / 3) % 3, lz = q / 9;
          const int li0 = 2 * (e % (TX + 1)), lj0 = 2 * (e / (TX + 1));
          const int rq = (li0 + lx) + WX * (lj0 + ly);
          const double fx0 = sDM[3 * lx], fx1 = sDM[3 * lx + 1], fx2 = sDM[3 * lx + 2];
          const double fy0 = sDM[3 * ly], fy1 = sDM[3 * ly + 1], fy2 = sDM[3 * ly + 2];
          const double fz0 = sDM[3 * lz], fz1 = sDM[3 * lz + 1], fz2 = sDM[3 * lz + 2];
          const int sq = (2 * ez + lz) % RING, z0 = (2 * ez) % RING, z1 = (2 * ez + 1) % RING, z2 = (2 * ez + 2) % RING;
          double d[3][3];
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            const double* U = &sV[sq][c][0];
            d[c][0] = g.dxi[0] * (fx0 * U[rq - lx] + fx1 * U[rq - lx + 1] + fx2 * U[rq - lx + 2]);
            d[c][1] = g.dxi[1] * (fy0 * U[rq - WX * ly] + fy1 * U[rq - WX * ly + WX] + fy2 * U[rq - WX * ly + 2 * WX]);
            d[c][2] = g.dxi[2] * (((0.0 + fz0 * sV[z0][c][rq]) + fz1 * sV[z1][c][rq]) + fz2 * sV[z2][c][rq]);
          }
          const double lam = sLM[sq][0][rq], mu = sLM[sq][1][rq];
          const double W = sWq[q];
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
    }

    // ------------------------------------------------------- owner phase
    const bool top_plane = (ez == g.nez - 1) && p == 2;
    if (owned && (!halo || p == 2)) {
      const int lz = p;
      const int sP = k % RING;
      const int z0 = (2 * ez) % RING, z1 = (2 * ez + 1) % RING, z2 = (2 * ez + 2) % RING;
      const bool pn = HAS_PML && pml3(g, i, j, k);
      // z coefficients: forward DM(lz, .), transposed DM(., lz)
      const double fz0 = sDM[3 * lz], fz1 = sDM[3 * lz + 1], fz2 = sDM[3 * lz + 2];
      const double tz0 = sDM[lz], tz1 = sDM[3 + lz], tz2 = sDM[6 + lz];
      double aE[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
      if constexpr (HAS_PML && !HAS_RD) {
        // separable pull (PML-only tiles): w(l_z) and the z stencil are uniform per slab
        const double wl = gll_w(lz);
        const double tzw0 = tz0 * gll_w(0), tzw1 = tz1 * gll_w(1), tzw2 = tz2 * gll_w(2);
#pragma unroll
        for (int ii = 0; ii < 3; ++ii) {
          const double* Gx = &sG[sP][3 * ii + 0][0];
          const double* Gy = &sG[sP][3 * ii + 1][0];
          double sx = 0.0, sy = 0.0;
#pragma unroll
          for (int d = 0; d < 5; ++d) {
            sx += cxs[d] * Gx[pxs[d]];
            sy += cys[d] * Gy[pys[d]];
          }
          const double sz =
              ((0.0 + tzw0 * sG[z0][3 * ii + 2][rown]) + tzw1 * sG[z1][3 * ii + 2][rown]) + tzw2 * sG[z2][3 * ii + 2][rown];
          acc[ii] += wl * (sx + sy) + azs * sz;
        }
        double dX[3], dY[3], dZ[3];
#pragma unroll
        for (int ii = 0; ii < 3; ++ii) {
          const double* Wx = &sW[sP][ii][0];
          const double* Wy = &sW[sP][3 + ii][0];
          double sx = cxs[2] * Wx[rown], sy = cys[2] * Wy[rown];
#pragma unroll
          for (int d = 0; d < 5; ++d) {
            if (d == 2) continue;
            sx -= cxs[d] * Wx[pxs[d]];
            sy -= cys[d] * Wy[pys[d]];
          }
          dX[ii] = sx;
          dY[ii] = sy;
          dZ[ii] = azs * (((0.0 + fz0 * sW[z0][6 + ii][rown]) + fz1 * sW[z1][6 + ii][rown]) + fz2 * sW[z2][6 + ii][rown]);
        }
        aE[0] += wl * dX[0];
        aE[1] += wl * dY[1];
        aE[2] += wl * dZ[2];
        aE[3] += wl * (dY[0] + dX[1]);
        aE[4] += wl * (dZ[0] + dX[2]);
        aE[5] += wl * (dZ[1] + dY[2]);
      } else
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
          const int et = rd3(g, ex, ey, ez) ? 1 : 2;
          const int rP = rown;                        // window position of (i, j)
          const double tx0 = sDM[lx], tx1 = sDM[3 + lx], tx2 = sDM[6 + lx];
          const double ty0 = sDM[ly], ty1 = sDM[3 + ly], ty2 = sDM[6 + ly];
          if constexpr (HAS_RD) {
            if (et == 1 && useK) {
              const double* Se = &sSig[e][0][0];
              const int bx_ = 3 * ly + 9 * lz, by_ = lx + 9 * lz, bz_ = lx + 3 * ly;
#pragma unroll
              for (int ii = 0; ii < 3; ++ii) {
                const double* Sx = Se + 27 * sym3(ii, 0);
                const double* Sy = Se + 27 * sym3(ii, 1);
                const double* Sz = Se + 27 * sym3(ii, 2);
                const double s = ((0.0 + tx0 * Sx[bx_]) + tx1 * Sx[bx_ + 1]) + tx2 * Sx[bx_ + 2];
                const double t2 = ((0.0 + ty0 * Sy[by_]) + ty1 * Sy[by_ + 3]) + ty2 * Sy[by_ + 6];
                const double t3 = ((0.0 + tz0 * Sz[bz_]) + tz1 * Sz[bz_ + 9]) + tz2 * Sz[bz_ + 18];
                acc[ii] += g.dxi[0] * s + g.dxi[1] * t2 + g.dxi[2] * t3;
              }
            }
          }
          if constexpr (HAS_PML) {
            if (et == 2) {
              // u rows: transposed derivatives of the nodal flux, weights of the element points
              const double wx0 = sWq[0 + 3 * ly + 9 * lz], wx1 = sWq[1 + 3 * ly + 9 * lz], wx2 = sWq[2 + 3 * ly + 9 * lz];
              const double wy0 = sWq[lx + 0 + 9 * lz], wy1 = sWq[lx + 3 + 9 * lz], wy2 = sWq[lx + 6 + 9 * lz];
              const double wz0 = sWq[lx + 3 * ly], wz1 = sWq[lx + 3 * ly + 9], wz2 = sWq[lx + 3 * ly + 18];
              const int rx = rP - lx, ry = rP - WX * ly;
#pragma unroll
              for (int ii = 0; ii < 3; ++ii) {
                const double* Gx = &sG[sP][3 * ii + 0][0];
                const double* Gy = &sG[sP][3 * ii + 1][0];
                const double s = ((0.0 + tx0 * (wx0 * Gx[rx])) + tx1 * (wx1 * Gx[rx + 1])) + tx2 * (wx2 * Gx[rx + 2]);
                const double t2 =
                    ((0.0 + ty0 * (wy0 * Gy[ry])) + ty1 * (wy1 * Gy[ry + WX])) + ty2 * (wy2 * Gy[ry + 2 * WX]);
                const double t3 = ((0.0 + tz0 * (wz0 * sG[z0][3 * ii + 2][rP])) + tz1 * (wz1 * sG[z1][3 * ii + 2][rP])) +
                                  tz2 * (wz2 * sG[z2][3 * ii + 2][rP]);
                acc[ii] += g.dxi[0] * s + g.dxi[1] * t2 + g.dxi[2] * t3;
              }
              // s rows: F-weighted strains of the element at this point
              const double fx0 = sDM[3 * lx], fx1 = sDM[3 * lx + 1], fx2 = sDM[3 * lx + 2];
              const double fy0 = sDM[3 * ly], fy1 = sDM[3 * ly + 1], fy2 = sDM[3 * ly + 2];
              double dX[3], dY[3], dZ[3];
#pragma unroll
              for (int ii = 0; ii < 3; ++ii) {
                const double* Wx = &sW[sP][ii][0];
                const double* Wy = &sW[sP][3 + ii][0];
                dX[ii] = g.dxi[0] * (fx0 * Wx[rx] + fx1 * Wx[rx + 1] + fx2 * Wx[rx + 2]);
                dY[ii] = g.dxi[1] * (fy0 * Wy[ry] + fy1 * Wy[ry + WX] + fy2 * Wy[ry + 2 * WX]);
                dZ[ii] = g.dxi[2] * (((0.0 + fz0 * sW[z0][6 + ii][rP]) + fz1 * sW[z1][6 + ii][rP]) + fz2 * sW[z2][6 + ii][rP]);
              }
              const double W = sWq[lx + 3 * ly + 9 * lz];
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
        if constexpr (!TR) {
#pragma unroll
          for (int m = 0; m < 6; ++m) acc[3 + m] += aE[m];
        } else {
          const double lam = lam_o, mu = mu_o;
          const double l2m = lam + 2.0 * mu, m2 = 2.0 * mu;
          acc[3] += -(l2m * aE[0] + lam * aE[1] + lam * aE[2]);
          acc[4] += -(lam * aE[0] + l2m * aE[1] + lam * aE[2]);
          acc[5] += -(lam * aE[0] + lam * aE[1] + l2m * aE[2]);
          acc[6] += -(m2 * aE[3]);
          acc[7] += -(m2 * aE[4]);
          acc[8] += -(m2 * aE[5]);
        }
      }
      // ------------------------------------------------------- finish
      if (!(p == 2 && !top_plane) && !halo) {
        const unsigned n = (unsigned)i + (unsigned)g.NX * ((unsigned)j + (unsigned)g.NY * (unsigned)k);
        const unsigned offu = n * 24u, offs = SB + n * 48u;
        const bool ua = u_act3(g, i, j, k);
        if constexpr (SPLIT) {
          if (pml3(g, i, j, k)) {
            // interface node: its regular-domain K sums, finished by the nodal pull
            if (ua) st3(A.E + (long long)shot * 3 * g.NN, offu, acc[0], acc[1], acc[2]);
            goto next_slab;
          }
        }
        const int sm = pn ? smask3(g, i, j, k) : 0;
        double raw[9];
#pragma unroll
        for (int m = 0; m < 9; ++m) raw[m] = acc[m];
        double out[9], his[S == 5 ? 9 : 1];
        if constexpr (S == 0) {
#pragma unroll
          for (int m = 0; m < 9; ++m) out[m] = raw[m];
        } else {
          const double minv_u = minvu_o, minv_s = minvs_o;
          // this node's source terms per component, the entries in list order
          double srcv[3] = {0.0, 0.0, 0.0};
          if (src_j1 > src_j0 && ua) {
            const Src3* sp = &A.src;
            for (int jj = src_j0; jj < src_j1; ++jj) {
              const long long nd = src_lds ? sSrcNode[jj - src_j0] : sp->node[jj];
              if (nd != (long long)n) continue;
              const int cpt = src_lds ? sSrcComp[jj - src_j0] : sp->comp[jj];
              const double mm = sp->m[jj];
              double val;
              if (A.src_mode == 1) {
                const double f = sp->ftab[((size_t)shot * sp->nstep + (A.step + sdyn)) * 3 + A.fcol];
                val = f * (sp->val[jj] * mm);
              } else {
                const size_t nr = (size_t)sp->nrec;
                const size_t base = (size_t)shot * ((size_t)sp->nstep + 1) * nr + ((size_t)jj - (size_t)shot * nr);
                const size_t ra = (size_t)(A.ra + sdyn);
                const double va = sp->tr[base + ra * nr] - sp->obs[base + ra * nr];
                val = A.ca * (va * mm);
                if (A.rb_on) {
                  const size_t rb = (size_t)(A.rb + sdyn);
                  const double vb = sp->tr[base + rb * nr] - sp->obs[base + rb * nr];
                  val = val + A.cb * (vb * mm);
                }
              }
              if (cpt == 0) srcv[0] += val;
              else if (cpt == 1) srcv[1] += val;
              else srcv[2] += val;
            }
          }
#pragma unroll
          for (int m = 0; m < 9; ++m) {
            const double src = m < 3 ? srcv[m] : 0.0;
            const double minv = m < 3 ? minv_u : (m < 6 ? minv_s : 0.5 * minv_s);
            const double gout = src - minv * raw[m];
            if constexpr (!UPD) {
              out[m] = gout;
            } else if constexpr (S == 4) {
              const int cm = (NS == 9) ? m : (m < 3 ? m : 0);
              out[m] = c3o[cm] + A.h6 * gout;
            } else {
              const int cm = (NS == 9) ? m : (m < 3 ? m : 0);
              out[m] = c3o[cm] + gout;                        // u_{n+1}
              his[m] = chs[cm] + A.h * out[m];                // u_his + dt u_{n+1}
            }
          }
        }
        double* const dst = (S == 4) ? xnp[2] : (S == 5 ? xnp[0] : outp);
        if (ua) st3(dst, offu, out[0], out[1], out[2]);
        if constexpr (HAS_PML) {
          if (sm == 0x3f) st6(dst, offs, &out[3]);
          else {
            if (sm & 7) st3m(dst, offs, sm & 7, out[3], out[4], out[5]);
            if (sm & 56) st3m(dst, offs + 24u, (sm >> 3) & 7, out[6], out[7], out[8]);
          }
        }
        if constexpr (S == 5) {
          if (ua) st3(xnp[1], offu, his[0], his[1], his[2]);
          if constexpr (HAS_PML) {
            if (sm == 0x3f) st6(xnp[1], offs, &his[3]);
            else {
              if (sm & 7) st3m(xnp[1], offs, sm & 7, his[3], his[4], his[5]);
              if (sm & 56) st3m(xnp[1], offs + 24u, (sm >> 3) & 7, his[6], his[7], his[8]);
            }
          }
        }
      }
    }
  next_slab:
    if constexpr (RING == 3) __syncthreads();
  }
}

// ------------------------------------------------------------ split stage path
// Blocks of consecutive PML nodes in lattice order (sp_pml_node), one node per
// thread (nodal pull: SPLIT_NT nodes; nodal pass: SPLIT_OWN nodes and two on each
// side).  XCD x = blockIdx % 8 takes the
// logical blocks [x C, (x + 1) C) of the launch, so that each XCD's L2 holds a
// contiguous run of the list (the neighbouring rows and planes of its nodes); a
// launch has 8 C blocks and starts at logical block A.pb0.
__device__ __forceinline__ int split_block(int nblk) {
  const int C = (nblk + 7) / 8;
  return (blockIdx.x & 7) * C + (blockIdx.x >> 3);
}
// the same runs, each walked from its end: a nodal pull whose sources include
// PML nodes dispatches the top planes (surface sources and receivers, and with
// them the blocks that also add source terms) first instead of last
__device__ __forceinline__ int split_block_rev(int nblk) {
  const int C = (nblk + 7) / 8;
  return (blockIdx.x & 7) * C + (C - 1 - (int)(blockIdx.x >> 3));
}

__device__ __forceinline__ void st2(double* p, unsigned off, double a, double b) {
  *reinterpret_cast<double2*>(reinterpret_cast<char*>(p) + off) = make_double2(a, b);
}

// Flux records D of the split path, per shot 12 NN doubles: two arrays (ax =
// y, z) of six doubles per lattice node, (G_{x,ax}, G_{y,ax}, G_{z,ax},
// w_{ax,x}, w_{ax,y}, w_{ax,z}) with G_{i,ax} the nodal PML flux and w_{ax,i}
// the F-weighted inputs: a line along ax reads three 16-byte pairs per node.
// The x-lines are summed by the nodal pass itself, from LDS.
__device__ __forceinline__ unsigned d_rec(int ax, unsigned NN, unsigned n) {
  return 48u * ((unsigned)(ax - 1) * NN + n);
}

// Node-local part of RK4 stage S at one PML node: its stage inputs from the raw
// inputs (the fused kernel's PML window load, once per node), the flux records
// of its y- and z-lines, and into the stage output the node-local terms: stages
// 1-3 -M^-1 (diagonal terms + x-line sums), stage 4 the carry x3 + dt/6 (g1 +
// 2 g2 + 2 g3) minus dt/6 M^-1 (diagonal terms + x-line sums), with the x1, x2
// updates written here.  The x-line sums read the x records of the block's
// entries from LDS: consecutive list entries are x-neighbours.
template <int S, bool TR>
__global__ void __launch_bounds__(SPLIT_NT, S == 4 ? 2 : 3) pml_nodal_kernel(const Stage3Args A) {
  constexpr bool NEG = TR;
  constexpr int NIN = 2 + S;
  __shared__ double sXr[6][SPLIT_NT];     // x flux records of the block's entries
  __shared__ unsigned sXn[SPLIT_NT];      // their lattice nodes (~0: no entry)
  const Geo3& g = A.g;
  const int lb = split_block(A.pnbn);
  if (lb >= A.pnbn) return;
  const int tid = threadIdx.x;
  // list entry of this thread: the block owns [lb SPLIT_OWN, (lb + 1) SPLIT_OWN) and
  // computes two more on each side for their x records; a thread without an
  // entry runs node 0 (in bounds) and stores nothing
  const long long t = (long long)lb * SPLIT_OWN - 2 + tid;
  int i = 0, j = 0, k = 0;
  const bool valid = t >= 0 && sp_pml_node((unsigned)t, g.NX, g.NY, g.NZ, g.rlo, g.rhi, i, j, k);
  if (!valid) i = j = k = 0;
  const bool own = valid && tid >= 2 && tid < SPLIT_NT - 2;
  const int shot = blockIdx.y;
  const unsigned n = (unsigned)i + (unsigned)g.NX * ((unsigned)j + (unsigned)g.NY * (unsigned)k);
  const long long so = (long long)shot * A.shot_stride;
  const unsigned NN = (unsigned)g.NN;
  const unsigned SB = 8u * (3u * NN + (NN & 1u));
  const unsigned offu = n * 24u, offs = SB + n * 48u;
  const double* __restrict__ tab = A.tab;
  const double* __restrict__ nodet = A.node;

  double XU[3][NIN], XS6[6][NIN];
#pragma unroll
  for (int q = 0; q < NIN; ++q) {
    const double* ip = A.in[q] + so;
    ld3(ip, offu, XU[0][q], XU[1][q], XU[2][q]);
    ld2(ip, offs, XS6[0][q], XS6[1][q]);
    ld2(ip, offs + 16u, XS6[2][q], XS6[3][q]);
    ld2(ip, offs + 32u, XS6[4][q], XS6[5][q]);
  }
  double lam, mu, ax, bx, ay, by, az, bz;
  ld2(nodet, n * 16u, lam, mu);
  const double minv_u = A.minvu[n], rwp = A.rwp[n];
  ld2(tab, 8u * (unsigned)(TAB_STRETCH + 2 * i), ax, bx);
  ld2(tab, 8u * (unsigned)(TAB_STRETCH + 2 * g.NX + 2 * j), ay, by);
  ld2(tab, 8u * (unsigned)(TAB_STRETCH + 2 * g.NX + 2 * g.NY + 2 * k), az, bz);
  // the node's sides per axis; its PML weight sum and the stress rows' M^-1
  const SpAxis AX = sp_axis(i, g.nex, g.rlo[0], g.rhi[0]), AY = sp_axis(j, g.ney, g.rlo[1], g.rhi[1]),
               AZ = sp_axis(k, g.nez, g.rlo[2], g.nez);
  const double wps = g.detj * sp_pml_weight(AX, AY, AZ);
  const double minv_s = 1.0 / (((ax * ay) * az) * wps);
  const bool ua = u_act3(g, i, j, k);
  const int sm = smask3(g, i, j, k);

  double v1u[3], vu2[3], v3u[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) combine<S, NEG>(XU[c], A.hh, A.h, A.h6, 7, v1u[c], vu2[c], v3u[c]);
  // F[set][k], set 0 = w (G, v1), 1 = p (K, v2), 2 = e (C, v3)
  double F[3][3];
  F[2][0] = ay * az; F[2][1] = ax * az; F[2][2] = ax * ay;
  F[1][0] = ay * bz + az * by; F[1][1] = ax * bz + az * bx; F[1][2] = ax * by + ay * bx;
  F[0][0] = by * bz; F[0][1] = bx * bz; F[0][2] = bx * by;
  double wv[9];
#pragma unroll
  for (int d = 0; d < 3; ++d)
#pragma unroll
    for (int c = 0; c < 3; ++c) wv[3 * d + c] = ((0.0 + F[0][d] * v1u[c]) + F[1][d] * vu2[c]) + F[2][d] * v3u[c];
  // diagonal terms rho W (b v3 + c v2 + d v1) of ELEMENT_EM/EN (PML.F90 items 7..10)
  const double qb = ax * ay * bz + ax * az * by + ay * az * bx;
  const double qc = ax * by * bz + ay * bx * bz + az * bx * by;
  const double qd = bx * by * bz;
  auto diag = [&](double wm, double a1, double a2, double a3) { return wm * (((0.0 + qb * a3) + qc * a2) + qd * a1); };
  double dg[9], c3[9], Gv[9];
#pragma unroll
  for (int m = 0; m < 9; ++m) { c3[m] = 0.0; Gv[m] = 0.0; }
#pragma unroll
  for (int c = 0; c < 3; ++c) dg[c] = diag(rwp, v1u[c], vu2[c], v3u[c]);
  double* const x1n = A.xnew[0] ? A.xnew[0] + so : nullptr;
  double* const x2n = A.xnew[1] ? A.xnew[1] + so : nullptr;
  if constexpr (S == 4) {
    double xn0[3], xn1[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) rk4_update12(XU[c], A.hh, A.h, A.h6, xn0[c], xn1[c], c3[c]);
    if (own && ua) {
      st3(x1n, offu, xn0[0], xn0[1], xn0[2]);
      st3(x2n, offu, xn1[0], xn1[1], xn1[2]);
    }
  }
  const double l2m = lam + 2.0 * mu, m2 = 2.0 * mu;
#pragma unroll
  for (int hf = 0; hf < 2; ++hf) {
    const double(*XS)[NIN] = &XS6[3 * hf];
    double s1[3], s2[3], s3[3];
#pragma unroll
    for (int tt = 0; tt < 3; ++tt) combine<S, NEG>(XS[tt], A.hh, A.h, A.h6, 7, s1[tt], s2[tt], s3[tt]);
#pragma unroll
    for (int X = 0; X < 3; ++X) {
      const double* sv = X == 0 ? s1 : (X == 1 ? s2 : s3);
      if (hf == 0) {
        if constexpr (!TR) {
          // T(s): stress of the auxiliary strain (ELEMENT_EA_3D lower block), normal part
          const double Txx = l2m * sv[0] + lam * sv[1] + lam * sv[2];
          const double Tyy = lam * sv[0] + l2m * sv[1] + lam * sv[2];
          const double Tzz = lam * sv[0] + lam * sv[1] + l2m * sv[2];
          Gv[0] -= F[X][0] * Txx; Gv[4] -= F[X][1] * Tyy; Gv[8] -= F[X][2] * Tzz;
        } else {
          Gv[0] += F[X][0] * sv[0]; Gv[4] += F[X][1] * sv[1]; Gv[8] += F[X][2] * sv[2];
        }
      } else {
        if constexpr (!TR) {
          const double Txy = m2 * sv[0], Txz = m2 * sv[1], Tyz = m2 * sv[2];
          Gv[1] -= F[X][1] * Txy; Gv[2] -= F[X][2] * Txz; Gv[3] -= F[X][0] * Txy;
          Gv[5] -= F[X][2] * Tyz; Gv[6] -= F[X][0] * Txz; Gv[7] -= F[X][1] * Tyz;
        } else {
          Gv[1] += F[X][1] * sv[0]; Gv[2] += F[X][2] * sv[1]; Gv[3] += F[X][0] * sv[0];
          Gv[5] += F[X][2] * sv[2]; Gv[6] += F[X][0] * sv[1]; Gv[7] += F[X][1] * sv[2];
        }
      }
    }
    const double wm = hf == 0 ? wps : 2.0 * wps;
#pragma unroll
    for (int tt = 0; tt < 3; ++tt) dg[3 + 3 * hf + tt] = diag(wm, s1[tt], s2[tt], s3[tt]);
    if constexpr (S == 4) {
      double xn0[3], xn1[3];
#pragma unroll
      for (int tt = 0; tt < 3; ++tt) rk4_update12(XS[tt], A.hh, A.h, A.h6, xn0[tt], xn1[tt], c3[3 + 3 * hf + tt]);
      const int act = own ? (sm >> (3 * hf)) & 7 : 0;
      if (act) {
        st3m(x1n, offs + 24u * hf, act, xn0[0], xn0[1], xn0[2]);
        st3m(x2n, offs + 24u * hf, act, xn1[0], xn1[1], xn1[2]);
      }
    }
  }
  // node-local terms of the stage output, without the x-line sums
  const double sc = S == 4 ? A.h6 : 1.0;
  double pre[9];
#pragma unroll
  for (int m = 0; m < 9; ++m) {
    const double minv = m < 3 ? minv_u : (m < 6 ? minv_s : 0.5 * minv_s);
    pre[m] = c3[m] - sc * (minv * dg[m]);
  }
  // x records to LDS; y and z records to memory
  sXn[tid] = valid ? n : ~0u;
  sXr[0][tid] = Gv[0]; sXr[1][tid] = Gv[3]; sXr[2][tid] = Gv[6];
  sXr[3][tid] = wv[0]; sXr[4][tid] = wv[1]; sXr[5][tid] = wv[2];
  // flux records: (G_{x,ax}, G_{y,ax}, G_{z,ax}, w_{ax,x}, w_{ax,y}, w_{ax,z}), Gv[3 i + ax], wv[3 ax + i]
  double* const Dp = A.D + (long long)shot * 12 * g.NN;
  if (own) {
#pragma unroll
    for (int a = 1; a < 3; ++a) {
      const unsigned o = d_rec(a, NN, n);
      st2(Dp, o, Gv[a], Gv[3 + a]);
      st2(Dp, o + 16u, Gv[6 + a], wv[3 * a]);
      st2(Dp, o + 32u, wv[3 * a + 1], wv[3 * a + 2]);
    }
  }
  __syncthreads();
  if (!own) return;
  // the x-line of the pull: side-scaled stencil over the x-neighbours' records.
  // The neighbour at lattice offset di (|di| <= 2) is at most |di| entries away
  // in the list (no PML node between, or one); one that is not a PML node (a
  // regular-domain node) has no record, and offsets outside the lattice have zero
  // coefficients.
  double cs0, cs1, cu[5];
  sp_side_scales(AX, AY, AZ, cs0, cs1);
  sp_stencil(AX.odd, cs0, cs1, cu);
  double sux[3] = {0.0, 0.0, 0.0}, sdx[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int d = 0; d < 5; ++d) {
    const int di = d - 2;
    int slot = -1;
    if (di == 0) {
      slot = tid;
    } else if (i + di >= 0 && i + di < g.NX) {
      const unsigned ne = (unsigned)((int)n + di);
      const int near = tid + (di > 0 ? 1 : -1);
      if (di == 2 || di == -2) slot = sXn[tid + di] == ne ? tid + di : (sXn[near] == ne ? near : -1);
      else slot = sXn[near] == ne ? near : -1;
    }
    // u rows with cu, s rows with every offset but 0 negated
    const double cw = d == 2 ? cu[d] : -cu[d];
    if (slot >= 0) {
      sux[0] += cu[d] * sXr[0][slot]; sux[1] += cu[d] * sXr[1][slot]; sux[2] += cu[d] * sXr[2][slot];
      sdx[0] += cw * sXr[3][slot]; sdx[1] += cw * sXr[4][slot]; sdx[2] += cw * sXr[5][slot];
    }
  }
  // its terms of the sums, taken into the node-local terms
  double rx[9];
#pragma unroll
  for (int c = 0; c < 3; ++c) rx[c] = g.detj * (g.dxi[0] * sux[c]);
  const double cx = g.detj * g.dxi[0];
  const double ddx0 = cx * sdx[0], ddx1 = cx * sdx[1], ddx2 = cx * sdx[2];
  if constexpr (!TR) {
    rx[3] = ddx0; rx[4] = 0.0; rx[5] = 0.0; rx[6] = ddx1; rx[7] = ddx2; rx[8] = 0.0;
  } else {
    rx[3] = -(l2m * ddx0); rx[4] = -(lam * ddx0); rx[5] = -(lam * ddx0);
    rx[6] = -(m2 * ddx1); rx[7] = -(m2 * ddx2); rx[8] = 0.0;
  }
#pragma unroll
  for (int m = 0; m < 9; ++m) {
    const double minv = m < 3 ? minv_u : (m < 6 ? minv_s : 0.5 * minv_s);
    pre[m] = pre[m] - sc * (minv * rx[m]);
  }
  double* const dst = (S == 4 ? A.xnew[2] : A.out) + so;
  if (ua) st3(dst, offu, pre[0], pre[1], pre[2]);
  if (sm == 0x3f) st6(dst, offs, &pre[3]);
  else {
    if (sm & 7) st3m(dst, offs, sm & 7, pre[3], pre[4], pre[5]);
    if (sm & 56) st3m(dst, offs + 24u, (sm >> 3) & 7, pre[6], pre[7], pre[8]);
  }
}

// Pull of RK4 stage S at one PML node: the PML elements' terms from the flux
// records along the y- and z-lines through the node (side-scaled 5-point
// stencils; the x-line terms are in the node-local terms), the regular-domain K
// sums of an interface node from E, the sources, then out = (node-local terms)
// + sc (src - M^-1 sums), sc = 1 or dt/6.
template <int S, bool TR>
__global__ void __launch_bounds__(SPLIT_NT, 2) pml_pull_kernel(const Stage3Args A) {
  __shared__ int sPn[SRC_LDS], sPj[SRC_LDS], sPc[SRC_LDS];   // the block's source entries
  const Geo3& g = A.g;
  const int lb = (A.src_mode && A.src.ptop) ? split_block_rev(A.pnb) : split_block(A.pnb);
  if (lb >= A.pnb) return;
  const int blk = A.pb0 + lb;
  const int shot = blockIdx.y;
  // this block's source entries (a uniform range), staged in LDS: lattice node,
  // entry and component, so that each node scans them without global loads
  int e0 = 0, e1 = 0;
  if (A.src_mode) {
    const int pb = shot * A.src.pnblk + blk;
    e0 = A.src.pblk_ptr[pb];
    e1 = A.src.pblk_ptr[pb + 1];
  }
  const bool src_lds = e1 - e0 <= SRC_LDS;
  if (e1 > e0 && src_lds) {
    for (int t = threadIdx.x; t < e1 - e0; t += SPLIT_NT) {
      const int jj = A.src.pblk_idx[e0 + t];
      sPj[t] = jj;
      sPn[t] = (int)A.src.node[jj];
      sPc[t] = A.src.comp[jj];
    }
    __syncthreads();
  }
  int i, j, k;
  if (!sp_pml_node((unsigned)(blk * SPLIT_NT) + threadIdx.x, g.NX, g.NY, g.NZ, g.rlo, g.rhi, i, j, k)) return;
  const unsigned n = (unsigned)i + (unsigned)g.NX * ((unsigned)j + (unsigned)g.NY * (unsigned)k);
  const long long so = (long long)shot * A.shot_stride;
  const unsigned NN = (unsigned)g.NN;
  const unsigned SB = 8u * (3u * NN + (NN & 1u));
  const unsigned offu = n * 24u, offs = SB + n * 48u;
  const double* __restrict__ nodet = A.node;

  // node-local terms the nodal pass wrote, and M^-1, issued with the records
  double* const dst = (S == 4 ? A.xnew[2] : A.out) + so;
  const double minv_u = A.minvu[n];
  const double* __restrict__ tab = A.tab;
  const double sax = tab[TAB_STRETCH + 2 * i], say = tab[TAB_STRETCH + 2 * g.NX + 2 * j],
               saz = tab[TAB_STRETCH + 2 * g.NX + 2 * g.NY + 2 * k];
  double pre[9];
  ld3(dst, offu, pre[0], pre[1], pre[2]);
  ld2(dst, offs, pre[3], pre[4]);
  ld2(dst, offs + 16u, pre[5], pre[6]);
  ld2(dst, offs + 32u, pre[7], pre[8]);
  double lam = 0.0, mu = 0.0;
  if constexpr (TR) ld2(nodet, n * 16u, lam, mu);
  // interface sums, read with the first batch where a regular-domain element
  // holds the node; other lanes read the shot's first entry (one shared line)
  // and drop it
  const SpAxis ax3[3] = {sp_axis(i, g.nex, g.rlo[0], g.rhi[0]), sp_axis(j, g.ney, g.rlo[1], g.rhi[1]),
                         sp_axis(k, g.nez, g.rlo[2], g.nez)};
  const bool iface = sp_iface(ax3[0], ax3[1], ax3[2]);
  double e3[3];
  ld3(A.E + (long long)shot * 3 * g.NN, iface ? offu : 0u, e3[0], e3[1], e3[2]);
  // M^-1 of the stress rows: 1 / (a W), a the stretch-factor product, W the PML weight sum
  const double minv_s = 1.0 / (((sax * say) * saz) * (g.detj * sp_pml_weight(ax3[0], ax3[1], ax3[2])));
  const double* const Dp = A.D + (long long)shot * 12 * g.NN;
  const int cc[3] = {i, j, k};
  const int nax[3] = {g.NX, g.NY, g.NZ};
  const int str[3] = {1, g.NX, g.NX * g.NY};
  // y- and z-lines (the x-line terms are in the node-local terms): per axis
  // su = sum_d cu[d] G_{.,ax}(n + (d - 2) str), sd the same with the s-row
  // stencil (every offset but 0 negated) on w_{ax,.}
  double su[3][3], sd[3][3];
#pragma unroll
  for (int ax = 1; ax < 3; ++ax) {
    double s0, s1, cu[5];
    sp_side_scales(ax3[ax], ax3[ax == 0 ? 1 : 0], ax3[ax == 2 ? 1 : 2], s0, s1);
    sp_stencil(ax3[ax].odd, s0, s1, cu);
#pragma unroll
    for (int c = 0; c < 3; ++c) { su[ax][c] = 0.0; sd[ax][c] = 0.0; }
#pragma unroll
    for (int d = 0; d < 5; ++d) {
      // odd coordinates have no +-2 terms: skipped (a row of consecutive list
      // entries shares the y and z coordinates)
      if ((d == 0 || d == 4) && ax3[ax].odd) continue;
      // offsets outside the lattice have zero coefficients: read the node itself
      const int p = cc[ax] + d - 2;
      const unsigned m = (p >= 0 && p < nax[ax]) ? (unsigned)((int)n + (d - 2) * str[ax]) : n;
      const unsigned o = d_rec(ax, NN, m);
      double G0, G1, G2, W0, W1, W2;
      ld2(Dp, o, G0, G1);
      ld2(Dp, o + 16u, G2, W0);
      ld2(Dp, o + 32u, W1, W2);
      const double cw = d == 2 ? cu[d] : -cu[d];
      su[ax][0] += cu[d] * G0; su[ax][1] += cu[d] * G1; su[ax][2] += cu[d] * G2;
      sd[ax][0] += cw * W0; sd[ax][1] += cw * W1; sd[ax][2] += cw * W2;
    }
  }
  double raw[9];
#pragma unroll
  for (int c = 0; c < 3; ++c) raw[c] = g.detj * ((0.0 + g.dxi[1] * su[1][c]) + g.dxi[2] * su[2][c]);
#pragma unroll
  for (int c = 0; c < 3; ++c) raw[c] += iface ? e3[c] : 0.0;
  double dd[3][3];
#pragma unroll
  for (int ax = 1; ax < 3; ++ax)
#pragma unroll
    for (int c = 0; c < 3; ++c) dd[ax][c] = (g.detj * g.dxi[ax]) * sd[ax][c];
  const double aE[6] = {0.0, dd[1][1], dd[2][2], dd[1][0], dd[2][0], dd[2][1] + dd[1][2]};
  if constexpr (!TR) {
#pragma unroll
    for (int m = 0; m < 6; ++m) raw[3 + m] = aE[m];
  } else {
    const double l2m = lam + 2.0 * mu, m2 = 2.0 * mu;
    raw[3] = -(l2m * aE[0] + lam * aE[1] + lam * aE[2]);
    raw[4] = -(lam * aE[0] + l2m * aE[1] + lam * aE[2]);
    raw[5] = -(lam * aE[0] + lam * aE[1] + l2m * aE[2]);
    raw[6] = -(m2 * aE[3]);
    raw[7] = -(m2 * aE[4]);
    raw[8] = -(m2 * aE[5]);
  }
  // sources of this node (the entries of this block, a uniform range)
  const bool ua = u_act3(g, i, j, k);
  double src[3] = {0.0, 0.0, 0.0};
  if (e1 > e0 && ua) {
    const Src3& sp = A.src;
    const int sdyn = A.step_dev ? *A.step_dev : 0;
    for (int e = e0; e < e1; ++e) {
      const int jj = src_lds ? sPj[e - e0] : sp.pblk_idx[e];
      if ((src_lds ? (long long)sPn[e - e0] : sp.node[jj]) != (long long)n) continue;
      const double mm = sp.m[jj];
      double val;
      if (A.src_mode == 1) {
        const double f = sp.ftab[((size_t)shot * sp.nstep + (A.step + sdyn)) * 3 + A.fcol];
        val = f * (sp.val[jj] * mm);
      } else {
        const size_t nr = (size_t)sp.nrec;
        const size_t base = (size_t)shot * ((size_t)sp.nstep + 1) * nr + ((size_t)jj - (size_t)shot * nr);
        const size_t ra = (size_t)(A.ra + sdyn);
        const double va = sp.tr[base + ra * nr] - sp.obs[base + ra * nr];
        val = A.ca * (va * mm);
        if (A.rb_on) {
          const size_t rb = (size_t)(A.rb + sdyn);
          const double vb = sp.tr[base + rb * nr] - sp.obs[base + rb * nr];
          val = val + A.cb * (vb * mm);
        }
      }
      const int cpt = src_lds ? sPc[e - e0] : sp.comp[jj];
      src[0] += cpt == 0 ? val : 0.0;
      src[1] += cpt == 1 ? val : 0.0;
      src[2] += cpt == 2 ? val : 0.0;
    }
  }
  // finish on the node-local terms
  const int sm = smask3(g, i, j, k);
  const double sc = S == 4 ? A.h6 : 1.0;
  double out[9];
#pragma unroll
  for (int m = 0; m < 9; ++m) {
    const double minv = m < 3 ? minv_u : (m < 6 ? minv_s : 0.5 * minv_s);
    const double sv = m < 3 ? src[m] : 0.0;
    out[m] = pre[m] + sc * (sv - minv * raw[m]);
  }
  if (ua) st3(dst, offu, out[0], out[1], out[2]);
  if (sm == 0x3f) st6(dst, offs, &out[3]);
  else {
    if (sm & 7) st3m(dst, offs, sm & 7, out[3], out[4], out[5]);
    if (sm & 56) st3m(dst, offs + 24u, (sm >> 3) & 7, out[6], out[7], out[8]);
  }
}

template <int S, bool TR>
hipError_t launch_split(const Stage3Args& A, int part, const int* tiles, int ntl, hipStream_t s) {
  if (A.nshot <= 0) return hipSuccess;
  if (part == 1) {
    if (ntl <= 0) return hipSuccess;
    const int nb = 8 * ((ntl * A.nshot + 7) / 8);
    stage_kernel3<4, 3, true, false, S, TR, true><<<nb, NT3, 0, s>>>(A, tiles, ntl);
  } else {
    const int nb = part == 0 ? A.pnbn : A.pnb;
    if (nb <= 0) return hipSuccess;
    const dim3 grid(8 * ((nb + 7) / 8), A.nshot);
    if (part == 0) pml_nodal_kernel<S, TR><<<grid, SPLIT_NT, 0, s>>>(A);
    else pml_pull_kernel<S, TR><<<grid, SPLIT_NT, 0, s>>>(A);
  }
  return hipGetLastError();
}

hipError_t launch_split_any(const Stage3Args& A, int stage, int transpose, int part, const int* tiles, int ntl,
                            hipStream_t s) {
  switch (stage * 2 + (transpose ? 1 : 0)) {
    case 2: return launch_split<1, false>(A, part, tiles, ntl, s);
    case 3: return launch_split<1, true>(A, part, tiles, ntl, s);
    case 4: return launch_split<2, false>(A, part, tiles, ntl, s);
    case 5: return launch_split<2, true>(A, part, tiles, ntl, s);
    case 6: return launch_split<3, false>(A, part, tiles, ntl, s);
    case 7: return launch_split<3, true>(A, part, tiles, ntl, s);
    case 8: return launch_split<4, false>(A, part, tiles, ntl, s);
    case 9: return launch_split<4, true>(A, part, tiles, ntl, s);
    default: return hipErrorInvalidValue;
  }
}

template <int S, bool TR>
hipError_t launch3(const Stage3Args& A, int cls, const int* tiles, int ntl, hipStream_t s) {
  if (ntl <= 0 || A.nshot <= 0) return hipSuccess;
  const int nb = 8 * ((ntl * A.nshot + 7) / 8);
  switch (cls) {
    case 0: stage_kernel3<4, 3, true, true, S, TR><<<nb, NT3, 0, s>>>(A, tiles, ntl); break;
    case 1: stage_kernel3<4, 3, true, false, S, TR><<<nb, NT3, 0, s>>>(A, tiles, ntl); break;
    default: stage_kernel3<4, 3, false, true, S, TR><<<nb, NT3, 0, s>>>(A, tiles, ntl); break;
  }
  return hipGetLastError();
}

}  // namespace

hipError_t launch_stage3(const Stage3Args& A, int stage, int transpose, int cls, const int* tiles, int ntl,
                         hipStream_t s) {
  switch (stage * 2 + (transpose ? 1 : 0)) {
    case 0: return launch3<0, false>(A, cls, tiles, ntl, s);
    case 1: return launch3<0, true>(A, cls, tiles, ntl, s);
    case 2: return launch3<1, false>(A, cls, tiles, ntl, s);
    case 3: return launch3<1, true>(A, cls, tiles, ntl, s);
    case 4: return launch3<2, false>(A, cls, tiles, ntl, s);
    case 5: return launch3<2, true>(A, cls, tiles, ntl, s);
    case 6: return launch3<3, false>(A, cls, tiles, ntl, s);
    case 7: return launch3<3, true>(A, cls, tiles, ntl, s);
    case 8: return launch3<4, false>(A, cls, tiles, ntl, s);
    case 9: return launch3<4, true>(A, cls, tiles, ntl, s);
    case 10: return launch3<5, false>(A, cls, tiles, ntl, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_split_nodal(const Stage3Args& A, int stage, int transpose, hipStream_t s) {
  return launch_split_any(A, stage, transpose, 0, nullptr, 0, s);
}
hipError_t launch_split_rd(const Stage3Args& A, int stage, int transpose, const int* tiles, int ntl, hipStream_t s) {
  return launch_split_any(A, stage, transpose, 1, tiles, ntl, s);
}
hipError_t launch_split_pull(const Stage3Args& A, int stage, int transpose, hipStream_t s) {
  return launch_split_any(A, stage, transpose, 2, nullptr, 0, s);
}

}  // namespace hr
