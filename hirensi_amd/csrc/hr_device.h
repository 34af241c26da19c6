// Device-side data structures shared by the HIP kernels and the host API.
#pragma once
#include <cstdint>

namespace hr {

constexpr int NCOMP = 9;           // u_x u_y u_z s_xx s_yy s_zz s_xy s_xz s_yz

// Structured box geometry and the per-element constants (all elements of the
// box share one geometry: edges hx, hy, hz).
struct Geo {
  int nex, ney, nez;               // elements per axis (PML included)
  int NX, NY, NZ;                  // lattice nodes per axis (2*ne + 1)
  long long NN;                    // lattice nodes
  int rlo[3], rhi[3];              // regular-domain element box [lo, hi)
  int top_rule;                    // 1: original top PML nodes lose s_zz, s_xz, s_yz
  int tiles[3];                    // stage-kernel tiles per axis
  int tx, ty;                      // elements per tile in x and y
  int tz;                          // elements per z-chunk of a tile
  double dxi[3];                   // dxi/dx = DJI(k,k) of the affine element
  double detj;                     // det J of the affine element
  double wq[27];                   // GLL weight * det J at local point a + 3b + 9c
  double pml[24];                  // PML_PARAM(6,4), column-major
};

struct StageArgs {
  Geo g;
  const double* in[6];             // stage: x1 x2 x3 g1 g2 g3; apply: in[0] = x
  int stage;                       // 0 = apply (raw operator), 1..4 = RK4 stage, 5 = central difference
  int set_mask;                    // bit0: G on v1, bit1: K on v2, bit2: C on v3
  int transpose;                   // 1: apply the transposed operator
  double sgn[3];                   // multipliers on v1, v2, v3
  double hh, h, h6;                // 0.5 dt, dt, dt / 6 (signed); stage 5: dt^2, dt, dt^3
  double* out;                     // stage 1-3: g_s; apply: raw result
  double* xnew[3];                 // stage 4: updated x1, x2, x3; stage 5: u_{n+1}, u_his
  int nsrc_tiles;                  // 0: no source
};

}  // namespace hr

#include <hip/hip_runtime.h>
namespace hr {
// Host launchers (defined with the kernels).
hipError_t launch_gather_eq(long long neq, const long long* slot, const double* x, const double* r, double* out,
                            hipStream_t s);
hipError_t launch_scatter_eq(long long neq, const long long* slot, const double* raw, const double* y, double sc,
                             const double* l, double* z, hipStream_t s);
// Device-side step index of a captured time loop: the kernel reads *step once and
// offsets its per-step pointer by (*step) * stride elements (step null: no offset).
struct StepRef {
  const int* step = nullptr;
  long long stride = 0;
};
hipError_t launch_record(int nrec, const long long* slot, const double* x2, double* row, hipStream_t s,
                         StepRef sr = StepRef());
hipError_t launch_store_box(const Geo& g, const int* box0, const int* boxn, const double* x2, double* hist,
                            hipStream_t s, StepRef sr = StepRef());
hipError_t launch_gradient(const Geo& g, const int* box0, const int* boxn, const double* uh, const double* p2,
                           double dt, double* glam, double* gmu, hipStream_t s, StepRef sr = StepRef());
// *step = value; *step += delta (one lane, a vector store)
hipError_t launch_step_set(int* step, int value, hipStream_t s);
hipError_t launch_step_advance(int* step, int delta, hipStream_t s);
hipError_t launch_gradient_finish(const Geo& g, const int* box0, const int* boxn, const double* glb,
                                  const double* gmb, double* gl, double* gm, hipStream_t s);
hipError_t launch_permute_cols(long long nrows, int ncols, const long long* perm, const double* src, double* dst,
                               hipStream_t s);
// g += box sums on the regular-domain node box
hipError_t launch_box_add(const Geo& g, const int* box0, const int* boxn, const double* glb, const double* gmb,
                          double* gl, double* gm, hipStream_t s);
// out = 0.5 sum_n mrd[n] |x3 of node n|^2 (displacement block); part: energy_partials_size() doubles
int energy_partials_size();
hipError_t launch_energy(long long nn, const double* mrd, const double* x3, double* part, double* out, hipStream_t s);
hipError_t launch_sum(int n, const double* in, double* out, hipStream_t s);
// receivers grouped by equation: group g = receivers gidx[gptr[g] .. gptr[g+1])
hipError_t launch_misfit(int nstep, int ngrp, const int* gptr, const int* gidx, int nrec, const double* tr,
                         const double* ob, double dt, double* part, double* out, hipStream_t s);
// gradient post-processing (hr_reg.hip)
long long reg_box_elements(const Geo& g);
hipError_t launch_gradient_mass(const Geo& g, double* img, double* restr, hipStream_t s);
// tv: 1 total variation, 0 TN; wreg: host (w_top, w_bot, w_Length) or nullptr;
// cost_el: 2 x reg_box_elements(g)
hipError_t launch_regularization(const Geo& g, const double* zc, int tv, const double* wreg, const double* lam,
                                 const double* mu, double* rl, double* rm, double* cost_el, hipStream_t s);
// out[v] = x[v] . y[v] (y[v] == nullptr: sum of x[v]), v < nd <= 4, over n
// entries spaced by stride; part: dot_partials_size() doubles
hipError_t launch_dots(int nd, long long n, long long stride, const double* const* x, const double* const* y,
                       double* part, double* out, hipStream_t s);
int dot_partials_size();
hipError_t launch_reg_update(long long n, int add, double fl, double fm, const double* rl, const double* rm,
                             const double* img, double* gl, double* gm, hipStream_t s);
}  // namespace hr

namespace hr {
// ----------------------------------------------------------- stage kernel v3
constexpr int HR_MAX_ZCHUNKS = 64;
// Compact launch arguments: per-node and per-axis coefficients live in device
// tables, source lists behind a device pointer.
struct Geo3 {
  int NX, NY, NZ, NN;              // lattice (9 * NN < 2^31)
  int nex, ney, nez;
  int rlo[3], rhi[2];              // regular-domain element box; it extends to the top in z
  int top_rule;
  int tiles0, tiles1, tz;
  double dxi[3];
  double detj;
  int zb[HR_MAX_ZCHUNKS + 1];      // z-chunk c spans element layers [zb[c], zb[c+1])
};

// Node table: pairs (lam, mu) per node (2 NN doubles), then quadruples
// (minv_u, minv_s, rwp, wps) per node (4 NN doubles); each node's entries are
// one or two 16-byte loads.
constexpr int NT_PAIR = 2, NT_QUAD = 4, NT_COUNT = 6;
// Small table: [wq (27) | pad | (a_x, b_x) pairs (NX) | (a_y, b_y) (NY) | (a_z, b_z) (NZ)].
constexpr int TAB_STRETCH = 28;

struct Src3 {                      // device-resident source entries of one solve, grouped by shot and tile
  const int* tile_ptr;             // (nshot * ntiles + 1): entries of (shot, tile) at [ptr[s*ntiles+t], ptr[..+1])
  const long long* node;           // lattice node
  const int* comp;                 // component 0..2
  const double* m;                 // M^-1 of that DOF
  const double* val;               // forward: B entries
  const double* ftab;              // forward: time function f(T), f(T+dt/2), f(T+dt): nshot x nstep x 3
  const double* tr;                // adjoint: traces per shot, (nstep+1) rows of nrec, columns in entry order
  const double* obs;               // adjoint: observations, same layout
  int nrec;                        // adjoint: entries per shot (shared receivers)
  int nstep;
  int ntiles;
  // the same entries at PML nodes, by block of the nodal PML pull: entries of
  // (shot, block) are pblk_idx[pblk_ptr[s*pnblk+b] .. pblk_ptr[s*pnblk+b+1])
  const int* pblk_ptr;
  const int* pblk_idx;
  int pnblk;
  int ptop;                        // some entry is at a PML node: the pull walks its block runs top-first
};

struct Stage3Args {
  Geo3 g;
  const double* in[6];             // stage: x1 x2 x3 g1 g2 g3; apply: in[0] = x
  double* out;                     // stage 1-3: g_s; apply: raw result
  double* xnew[3];                 // stage 4: updated x1, x2, x3; stage 5: u_{n+1}, u_his
  const double* node;              // node table
  const double* tab;               // small table
  const double* minvu;             // split path: M^-1 of the displacement rows per node (NN)
  const double* rwp;               // split path: sum of rho W over the PML elements per node (NN)
  Src3 src;                        // source entries, used when src_mode != 0
  int set_mask;                    // apply: bit0 G, bit1 K, bit2 C
  int src_mode;                    // 0 none, 1 forward: f * B * m, 2 adjoint misfit
  int step, fcol;                  // forward: f = ftab[(shot * nstep + step) * 3 + fcol]
  int ra, rb, rb_on;               // adjoint source rows (rb used when rb_on)
  double ca, cb;                   // adjoint row weights
  const int* step_dev;             // captured loops: step, ra, rb are offsets from *step_dev
  int nshot;                       // shots in the launch; state vectors are shot-major
  long long shot_stride;           // elements per shot of a state vector (9 NN)
  double hh, h, h6;                // 0.5 dt, dt, dt / 6 (signed); stage 5: dt^2, dt, dt^3
  // split stage path (RK4 stages): the nodal pull covers the SPLIT_NT-node
  // blocks [pb0, pb0 + pnb) of the PML nodes in lattice order, the nodal pass its
  // pnbn blocks of SPLIT_OWN nodes; the y- and z-line flux records D (12 NN
  // doubles per shot) and the regular-domain K sums of interface nodes E (3 NN
  // doubles per shot)
  int pb0, pnb;
  int pnbn;                        // blocks of the nodal pass (SPLIT_OWN list entries each)
  double* D;
  double* E;
};

// stage 0 = apply; transposed stages 1..4 carry the adjoint signs (-G, +K, -C)
hipError_t launch_stage3(const Stage3Args& A, int stage, int transpose, int tile_class, const int* tiles, int ntiles,
                         hipStream_t s);
// split stage path, RK4 stages 1..4: the nodal pass over the PML nodes (flux
// records, node-local terms, stage-4 x1 x2), the regular-domain kernel over the
// tiles with regular-domain elements (finishes regular-domain-only nodes, K
// contribution of interface nodes into E) and the nodal pull over the PML nodes
constexpr int SPLIT_NT = 256;
// list entries a nodal-pass block owns: it also computes two entries on each
// side, for the x flux records of its own entries' x-lines
constexpr int SPLIT_OWN = SPLIT_NT - 4;      // nodes (threads) per block of the nodal kernels
hipError_t launch_split_nodal(const Stage3Args& A, int stage, int transpose, hipStream_t s);
hipError_t launch_split_rd(const Stage3Args& A, int stage, int transpose, const int* tiles, int ntiles, hipStream_t s);
hipError_t launch_split_pull(const Stage3Args& A, int stage, int transpose, hipStream_t s);
}  // namespace hr
