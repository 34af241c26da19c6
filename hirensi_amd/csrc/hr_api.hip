// Host side of the C ABI declared in include/hirensi.h: context, mesh and
// DOF model, node coefficients, stage-kernel launches and the RK4 drivers.

#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdio>
#include <cstdarg>
#include <cstring>
#include <string>
#include <vector>
#include <algorithm>
#include <functional>
#include <unordered_map>
#include "hr_device.h"
#include "hr_revolve.h"
#include "hr_internal.h"
#include "hr_split.h"
#include "../../include/hirensi.h"

#pragma clang fp contract(off)


using namespace hr;

namespace {
// element layers per z-chunk: fused tile classes (sweep of 3..8 at 64x64x32: 7 first) and
// split path (chunks from the regular domain's bottom; sweep of 5..26 at 64x64x32, PML 6:
// 10 first, 9 and 13 3 % and 14 % slower)
constexpr int DEFAULT_TZ = 7;
constexpr int DEFAULT_TZ_SPLIT = 10;
int default_tz(int version) { return version == 4 ? DEFAULT_TZ_SPLIT : DEFAULT_TZ; }
constexpr int DEFAULT_TY = 3;
constexpr int DEFAULT_VERSION = 4;   // 4 = split stage path, 3 = fused tile classes

struct HipError {
  hipError_t e;
};

// per-launch source parameters of the v3 stage kernel
struct Src3Spec {
  const Src3* lists = nullptr;   // host copy of the solve's source lists
  int mode = 0, ra = 0, rb = -1, step = 0, fcol = 0;
  bool rb_on = false;            // the second adjoint row rb is used
  double ca = 0.0, cb = 0.0;
};

struct Shell {
  double s = 1.0;
  double* l = nullptr;   // device, neq (nullptr = identity)
  double* r = nullptr;
};

}  // namespace

struct hr_ctx {
  std::string err;
  int flags = 0;
  Geo g{};
  double hel[3] = {1.0, 1.0, 1.0};          // element edges
  long long neq = 0;
  std::vector<long long> eq_slot;          // host copy of the numbering
  long long* d_eq_slot = nullptr;
  // regular-domain node box
  int box0[3] = {0, 0, 0}, boxn[3] = {0, 0, 0};
  long long nbox = 0;
  // node data (device)
  double *d_xc = nullptr, *d_yc = nullptr, *d_zc = nullptr;
  double* d_node = nullptr;                 // v3 node table (NT_COUNT x NN)
  double* d_tab = nullptr;                  // v3 small table (weights, per-axis stretch)
  Src3 src3_fwd{}, src3_adj{};              // source lists of the last forward / adjoint (device arrays)
  std::vector<double> h_minv_u, h_wps, h_a;  // host copies for M^-1 of sources, diag M
  bool have_material = false, have_pml = false;
  // state (device, 9*NN each)
  double* x[3] = {nullptr, nullptr, nullptr};
  double* y[3] = {nullptr, nullptr, nullptr};
  double* gs[3] = {nullptr, nullptr, nullptr};
  double* tmp = nullptr;                    // lattice scratch for hr_apply
  double* tmp2 = nullptr;
  Shell shell[3];
  // shots held in the state buffers (shot-major, 9 NN per shot and vector)
  int nshot = 1, shot_cap = 1;
  double* d_ftab = nullptr;                 // time-function table of the last forward (nshot x nstep x 3)
  size_t ftab_cap = 0;
  // forward record
  int f_nshot = 1, f_nstep = 0, f_nrec = 0, f_history = 0, f_segment = 0;
  double f_dt = 0.0;
  double* d_traces = nullptr;
  size_t traces_cap = 0;
  std::vector<long long> rec_slot_h;
  long long* d_rec_slot = nullptr;
  double* d_hist = nullptr;                 // store mode: (nstep+1) x 3 x nbox
  size_t hist_cap = 0;
  // checkpointed history: snapshot pool (3 x field x nshot each, kept across
  // solves), the slice buffer of one segment, the schedule of the last forward
  std::vector<double*> snap;
  size_t snap_elems = 0;
  double* d_seg = nullptr;
  int h_seglen = 0, h_nseg = 0, h_nsnap = 0;  // segment length, segments, snapshot slots
  bool h_store = false;                       // the plan keeps every slice (store mode)
  RevolvePlan h_plan;
  double h_budget = 0.0;                      // HBM bytes for the revolve history (0: automatic)
  // adjoint state buffers used while the forward state is recomputed (6 fields)
  double* adj_buf[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  // receivers grouped by equation for the misfit (receivers sharing one add first)
  int mis_ngrp = 0;
  int* d_mis_ptr = nullptr;
  int* d_mis_idx = nullptr;
  // staging of host vectors for hr_apply (neq each)
  double *d_ax = nullptr, *d_ay = nullptr;
  // regular-domain lumped mass per node (DIAG_M_RD of each displacement DOF)
  double* d_mrd = nullptr;
  double* d_minvu = nullptr;                  // split path node tables: M^-1 of u rows, rho W over PML elements
  double* d_rwp = nullptr;
  // regular-domain energy every e_every steps of the forward (0: off)
  int e_every = 0, e_rows = 0;
  double *d_energy = nullptr, *d_epart = nullptr;
  // one history slice for hr_gradient_accumulate
  double* d_uslice = nullptr;
  // capacities of the grow-only device allocations, by the address of their pointer field
  std::unordered_map<void*, size_t> caps;
  // forward sources for recomputation
  bool have_src = false;
  std::vector<int> src_tile_ptr_h;
  int* d_src_tile_ptr = nullptr;
  long long* d_src_node = nullptr;
  int* d_src_comp = nullptr;
  double* d_src_m = nullptr;
  double* d_src_val = nullptr;
  std::vector<double> ftab;
  // adjoint sources
  int* d_rec_tile_ptr = nullptr;
  long long* d_rec_node = nullptr;
  int* d_rec_comp = nullptr;
  double* d_rec_m = nullptr;
  long long* d_rec_perm = nullptr;          // tile-sorted entry -> receiver index
  double *d_obs = nullptr, *d_gl = nullptr, *d_gm = nullptr, *d_part = nullptr, *d_scalar = nullptr;
  double *d_tr_sorted = nullptr, *d_obs_sorted = nullptr;
  size_t obs_cap = 0, sorted_cap = 0;
  // gradient post-processing: 1/Mg and restriction (geometry only, built once),
  // regularisation vectors, per-element costs, reduction partials, staging of
  // host material / gradient vectors
  double *d_img = nullptr, *d_restr = nullptr, *d_reg = nullptr, *d_cost_el = nullptr, *d_red = nullptr;
  double *d_pm = nullptr, *d_pg = nullptr;
  double recompute = 1.0;
  hipStream_t stream = nullptr;
  // side streams (default priority) that run tile classes 1 and 2 of a stage
  // concurrently with class 0 on the main stream; fork/join through events
  hipStream_t cstream[2] = {nullptr, nullptr};
  hipEvent_t ev_fork = nullptr, ev_join[2] = {nullptr, nullptr};
  int concurrent = 1;
  // Per-step work (receiver traces, history slices, the per-step gradient) that
  // only reads the state a stage-1 launch reads and does not write: queued beside
  // the next stage launch (fused classes: on the main stream after the fork; split
  // path: on the second side stream), so that it runs beside the stage kernels
  // instead of between stages.  The next launch_stage consumes it.
  std::vector<std::function<int()>> deferred;
  // split stage path (version 4): the PML nodes' positions in lattice order (host;
  // the kernels walk that order without a list), the tile list of its
  // regular-domain kernel (the mixed and regular-domain tiles), the flux records
  // and interface sums of split_shots shots, source entries by pull block
  std::vector<std::pair<long long, int>> pnode_pos;   // (lattice index, position in lattice order), by lattice index
  bool split_ok = false;                      // the lattice fits the split path
  int pnblk = 0;                              // SPLIT_NT-node blocks of the PML nodes (nodal pull)
  int pnblk_nodal = 0;                        // SPLIT_OWN-node blocks of the PML nodes (nodal pass)
  int* d_tl_split = nullptr;
  int ntl_split = 0;
  double* d_D = nullptr;
  double* d_E = nullptr;
  int split_shots = 0;
  int *d_psrc_ptr = nullptr, *d_psrc_idx = nullptr, *d_prec_ptr = nullptr, *d_prec_idx = nullptr;
  // stage-kernel version and tile-class lists
  int version = DEFAULT_VERSION;
  std::vector<int> zb;                      // z-chunk boundaries (element layers)
  int* d_tl[3] = {nullptr, nullptr, nullptr};
  int ntl[3] = {0, 0, 0};
  // stage-kernel timing
  int profiling = 0;
  // captured time loops: graph_k steps per hipGraph (0: eager launches), the
  // device-side step index they read, and whether launches are being captured
  int graph_k = 0;
  int* d_step = nullptr;
  int dyn = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  long long graph_launches = 0;
  std::vector<hipEvent_t> ev_pool;
  // one launch in PROF_PERIOD is timed: (start, end) events around the kernel on
  // each stream it uses; its span is the first start to the last end
  struct StageEv {
    int stage, n;
    hipEvent_t start[3], end[3];
  };
  std::vector<StageEv> ev_open;
  long long prof_tick = 0;
  long long prof_all[5] = {0, 0, 0, 0, 0};    // launches while profiling
  long long prof_n[5] = {0, 0, 0, 0, 0};      // timed launches
  double prof_ms[5] = {0, 0, 0, 0, 0};        // their summed spans
};

namespace {

// State vectors: displacement block [3 n + c], then the stress block from
// stress_base = 3 NN rounded up to an even count [stress_base + 6 n + (c - 3)]: a
// node's components are contiguous and its six stress components 16-byte aligned
// (the field size 9 NN + (NN odd) is even, so every shot's vectors are too).
long long stress_base(const Geo& g) { return 3 * g.NN + (g.NN & 1); }
size_t field_size(const hr_ctx* ctx) { return (size_t)(stress_base(ctx->g) + 6 * ctx->g.NN); }

// State-vector slot of component c of lattice node n
long long slot_of(const Geo& g, int c, long long n) { return c < 3 ? 3 * n + c : stress_base(g) + 6 * n + (c - 3); }
void slot_decode(const Geo& g, long long slot, int& c, long long& n) {
  if (slot < 3 * g.NN) { n = slot / 3; c = (int)(slot - 3 * n); }
  else { const long long r = slot - stress_base(g); n = r / 6; c = 3 + (int)(r - 6 * n); }
}

int fail(hr_ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

#define HIPC(call)                                                                                   \
  do {                                                                                               \
    hipError_t _e = (call);                                                                          \
    if (_e != hipSuccess) return fail(ctx, (int)_e, "%s failed: %s", #call, hipGetErrorString(_e)); \
  } while (0)

// Grow-only device allocation: a pointer field keeps its buffer while the
// request fits, so repeated solves allocate nothing (contents are not kept on
// growth).
template <class T>
int dalloc(hr_ctx* ctx, T** p, size_t n) {
  if (n == 0) n = 1;
  auto it = ctx->caps.find((void*)p);
  if (*p && it != ctx->caps.end() && it->second >= n * sizeof(T)) return 0;
  if (*p) { (void)hipFree(*p); *p = nullptr; }
  hipError_t e = hipMalloc((void**)p, n * sizeof(T));
  if (e != hipSuccess) {
    ctx->caps.erase((void*)p);
    return fail(ctx, (int)e, "hipMalloc of %zu bytes failed: %s", n * sizeof(T), hipGetErrorString(e));
  }
  ctx->caps[(void*)p] = n * sizeof(T);
  return 0;
}

template <class T>
void dfree(hr_ctx* ctx, T** p) {
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  ctx->caps.erase((void*)p);
}

// host <-> device copies honouring HR_DEVICE_PTRS
int to_dev(hr_ctx* ctx, double* dst, const double* src, size_t n) {
  HIPC(hipMemcpyAsync(dst, src, n * sizeof(double),
                      (ctx->flags & HR_DEVICE_PTRS) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, ctx->stream));
  return 0;
}
int from_dev(hr_ctx* ctx, double* dst, const double* src, size_t n) {
  HIPC(hipMemcpyAsync(dst, src, n * sizeof(double),
                      (ctx->flags & HR_DEVICE_PTRS) ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, ctx->stream));
  if (!(ctx->flags & HR_DEVICE_PTRS)) HIPC(hipStreamSynchronize(ctx->stream));
  return 0;
}

// ----------------------------------------------------------- DOF rules (host)
bool h_u_active(const Geo& g, int i, int j, int k) {
  return !(i == 0 || j == 0 || k == 0 || i == g.NX - 1 || j == g.NY - 1);
}
bool h_pml_node(const Geo& g, int i, int j, int k) {
  return !(i > 2 * g.rlo[0] && i < 2 * g.rhi[0] && j > 2 * g.rlo[1] && j < 2 * g.rhi[1] && k > 2 * g.rlo[2]);
}
int h_s_mask(const Geo& g, int i, int j, int k) {
  if (!h_pml_node(g, i, j, k)) return 0;
  if (g.top_rule && k == g.NZ - 1 && !((i & 1) && (j & 1))) return 0x3f & ~((1 << 2) | (1 << 4) | (1 << 5));
  return 0x3f;
}
bool h_active(const Geo& g, int i, int j, int k, int c) {
  return c < 3 ? h_u_active(g, i, j, k) : ((h_s_mask(g, i, j, k) >> (c - 3)) & 1);
}

void h_stretch(const double* P, double x, double y, double z, double s[6]) {
  s[0] = s[1] = s[2] = 1.0;
  s[3] = s[4] = s[5] = 0.0;
  if (x >= P[18]) { double f = (x - P[18]) * 1.0 / P[12]; s[0] = 1.0 + P[0] * (f * f); s[3] = 0.0 + P[6] * (f * f); }
  else if (x <= P[19]) { double f = (x - P[19]) * -1.0 / P[13]; s[0] = 1.0 + P[1] * (f * f); s[3] = 0.0 + P[7] * (f * f); }
  if (y >= P[20]) { double f = (y - P[20]) * 1.0 / P[14]; s[1] = 1.0 + P[2] * (f * f); s[4] = 0.0 + P[8] * (f * f); }
  else if (y <= P[21]) { double f = (y - P[21]) * -1.0 / P[15]; s[1] = 1.0 + P[3] * (f * f); s[4] = 0.0 + P[9] * (f * f); }
  if (z <= P[23]) { double f = (z - P[23]) * -1.0 / P[17]; s[2] = 1.0 + P[5] * (f * f); s[5] = 0.0 + P[11] * (f * f); }
}

long long lat(const Geo& g, int i, int j, int k) {
  return (long long)i + (long long)g.NX * ((long long)j + (long long)g.NY * k);
}

// tile owning lattice node (i, j, k): chunk c owns planes [2 zb[c], 2 zb[c+1]), the last one also the top
int tile_of(const hr_ctx* ctx, int i, int j, int k) {
  const Geo& g = ctx->g;
  const int tx = std::min(i / (2 * g.tx), g.tiles[0] - 1);
  const int ty = std::min(j / (2 * g.ty), g.tiles[1] - 1);
  int tz = 0;
  while (tz + 1 < g.tiles[2] && k >= 2 * ctx->zb[tz + 1]) ++tz;
  return tx + g.tiles[0] * (ty + g.tiles[1] * tz);
}

int ntiles(const Geo& g) { return g.tiles[0] * g.tiles[1] * g.tiles[2]; }

template <class T>
int upload(hr_ctx* ctx, T** d, const std::vector<T>& h) {
  if (int rc = dalloc(ctx, d, h.size())) return rc;
  if (!h.empty()) HIPC(hipMemcpy(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

// Source entries at PML nodes by block of the split path's nodal pull: the
// (shot, block) lists of entry indices, in entry order (ptr: the (shot, tile)
// lists of nshot shots over nt tiles)
int build_pull_blocks(hr_ctx* ctx, const std::vector<int>& ptr, const std::vector<long long>& node, int nshot, int nt,
                      int** d_ptr, int** d_idx, Src3& h3) {
  const int nblk = ctx->pnblk;
  std::vector<std::pair<long long, int>> kv;
  for (int q = 0; q < nshot; ++q)
    for (int jj = ptr[(size_t)q * nt]; jj < ptr[(size_t)(q + 1) * nt]; ++jj) {
      auto it = std::lower_bound(ctx->pnode_pos.begin(), ctx->pnode_pos.end(), std::make_pair(node[jj], -1));
      if (it == ctx->pnode_pos.end() || it->first != node[jj]) continue;
      kv.push_back({(long long)q * nblk + it->second / SPLIT_NT, jj});
    }
  std::sort(kv.begin(), kv.end());
  std::vector<int> bp((size_t)nshot * nblk + 1, 0), idx;
  for (auto& e : kv) { bp[e.first + 1]++; idx.push_back(e.second); }
  for (size_t b = 0; b + 1 < bp.size(); ++b) bp[b + 1] += bp[b];
  if (int rc = upload(ctx, d_ptr, bp)) return rc;
  if (int rc = upload(ctx, d_idx, idx)) return rc;
  h3.pblk_ptr = *d_ptr;
  h3.pblk_idx = *d_idx;
  h3.pnblk = nblk;
  h3.ptop = idx.empty() ? 0 : 1;
  return 0;
}

// Sparse source grouped by owning tile.  entries: (lattice node, comp, M^-1)
struct SrcBuild {
  std::vector<int> ptr;
  std::vector<long long> node;
  std::vector<int> comp;
  std::vector<double> m;
  std::vector<long long> perm;   // sorted position -> input index
};

SrcBuild build_src(hr_ctx* ctx, int n, const long long* eq) {
  const Geo& g = ctx->g;
  std::vector<std::pair<int, int>> key(n);
  for (int q = 0; q < n; ++q) {
    const long long slot = ctx->eq_slot[eq[q]];
    int c;
    long long nd;
    slot_decode(g, slot, c, nd);
    const int i = (int)(nd % g.NX), j = (int)((nd / g.NX) % g.NY), k = (int)(nd / ((long long)g.NX * g.NY));
    key[q] = {tile_of(ctx, i, j, k), q};
  }
  std::stable_sort(key.begin(), key.end());
  SrcBuild s;
  const int nt = ntiles(g);
  s.ptr.assign(nt + 1, 0);
  for (auto& kq : key) s.ptr[kq.first + 1]++;
  for (int t = 0; t < nt; ++t) s.ptr[t + 1] += s.ptr[t];
  for (auto& kq : key) {
    const long long slot = ctx->eq_slot[eq[kq.second]];
    int c;
    long long nd;
    slot_decode(g, slot, c, nd);
    s.node.push_back(nd);
    s.comp.push_back(c);
    s.m.push_back(ctx->h_minv_u[nd]);
    s.perm.push_back(kq.second);
  }
  return s;
}


hipEvent_t take_event(hr_ctx* ctx) {
  if (!ctx->ev_pool.empty()) {
    hipEvent_t e = ctx->ev_pool.back();
    ctx->ev_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// timed launches: one in PROF_PERIOD, a period prime to the four RK4 stages
constexpr long long PROF_PERIOD = 7;

// fold finished event sets into the per-stage totals
int drain_events(hr_ctx* ctx) {
  if (ctx->ev_open.empty()) return 0;
  HIPC(hipStreamSynchronize(ctx->stream));
  for (auto& p : ctx->ev_open) {
    float lo = 0.0f, hi = 0.0f;
    for (int q = 0; q < p.n; ++q) {
      float a = 0.0f, b = 0.0f;
      HIPC(hipEventElapsedTime(&a, p.start[0], p.start[q]));
      HIPC(hipEventElapsedTime(&b, p.start[0], p.end[q]));
      lo = std::min(lo, a);
      hi = std::max(hi, b);
    }
    ctx->prof_n[p.stage] += 1;
    ctx->prof_ms[p.stage] += hi - lo;
    for (int q = 0; q < p.n; ++q) {
      ctx->ev_pool.push_back(p.start[q]);
      ctx->ev_pool.push_back(p.end[q]);
    }
  }
  ctx->ev_open.clear();
  return 0;
}

Stage3Args args3(hr_ctx* ctx, const StageArgs& A, const Src3Spec* sp) {
  Stage3Args B;
  std::memset(&B, 0, sizeof(B));
  const Geo& g = ctx->g;
  B.g.NX = g.NX; B.g.NY = g.NY; B.g.NZ = g.NZ; B.g.NN = (int)g.NN;
  B.g.nex = g.nex; B.g.ney = g.ney; B.g.nez = g.nez;
  for (int a = 0; a < 3; ++a) B.g.rlo[a] = g.rlo[a];
  B.g.rhi[0] = g.rhi[0]; B.g.rhi[1] = g.rhi[1];
  B.g.top_rule = g.top_rule;
  B.g.tiles0 = g.tiles[0]; B.g.tiles1 = g.tiles[1]; B.g.tz = g.tz;
  for (size_t c = 0; c < ctx->zb.size() && c <= (size_t)HR_MAX_ZCHUNKS; ++c) B.g.zb[c] = ctx->zb[c];
  for (int a = 0; a < 3; ++a) B.g.dxi[a] = g.dxi[a];
  B.g.detj = g.detj;
  for (int q = 0; q < 6; ++q) B.in[q] = A.in[q];
  B.out = A.out;
  for (int q = 0; q < 3; ++q) B.xnew[q] = A.xnew[q];
  B.node = ctx->d_node;
  B.tab = ctx->d_tab;
  B.minvu = ctx->d_minvu;
  B.rwp = ctx->d_rwp;
  B.set_mask = A.set_mask;
  B.hh = A.hh; B.h = A.h; B.h6 = A.h6;
  if (sp && sp->mode) {
    B.src = *sp->lists; B.src_mode = sp->mode; B.ra = sp->ra; B.rb = sp->rb; B.ca = sp->ca; B.cb = sp->cb;
    B.rb_on = sp->rb_on ? 1 : 0;
    B.step = sp->step; B.fcol = sp->fcol;
    B.step_dev = ctx->dyn ? ctx->d_step : nullptr;
  }
  // stage launches over the state buffers cover every shot; hr_apply works on one field
  B.nshot = A.stage > 0 ? ctx->nshot : 1;
  B.shot_stride = (long long)field_size(ctx);
  return B;
}

// queue the deferred per-step work on the main stream (in the order it was deferred)
int run_deferred(hr_ctx* ctx) {
  std::vector<std::function<int()>> work;
  work.swap(ctx->deferred);
  for (auto& f : work)
    if (int rc = f()) return rc;
  return 0;
}

int launch_stage(hr_ctx* ctx, StageArgs& A, const Src3Spec* sp = nullptr) {
  // per-kind statistics cover the apply and RK4 stages
  const bool counted = ctx->profiling && A.stage <= 4;
  if (counted) ctx->prof_all[A.stage] += 1;
  const bool prof = counted && (ctx->prof_tick++ % PROF_PERIOD) == 0;
  const bool split = ctx->version == 4 && A.stage >= 1 && A.stage <= 4;
  hr_ctx::StageEv ev{A.stage, ctx->concurrent ? (split ? 2 : 3) : 1, {nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
  if (prof) {
    if (ctx->ev_open.size() >= 4096) if (int rc = drain_events(ctx)) return rc;
    for (int q = 0; q < ev.n; ++q) {
      ev.start[q] = take_event(ctx);
      ev.end[q] = take_event(ctx);
      if (!ev.start[q] || !ev.end[q]) return fail(ctx, 3, "hipEventCreate failed");
    }
  }
  {
    // 32-bit byte offsets: the node table (6 NN) and the 9 components must stay below 4 GiB
    if ((unsigned long long)field_size(ctx) * 8ull >= (1ull << 32))
      return fail(ctx, 3, "lattice of %lld nodes is too large for the v3 stage kernel", ctx->g.NN);
    // the v3 kernel fixes the adjoint signs by the transpose flag of stages 1-4
    const bool adj_signs = A.sgn[0] == -1.0 && A.sgn[1] == 1.0 && A.sgn[2] == -1.0;
    const bool unit_signs = A.sgn[0] == 1.0 && A.sgn[1] == 1.0 && A.sgn[2] == 1.0;
    if (!((A.stage == 0 && unit_signs) || (A.stage > 0 && (A.transpose ? adj_signs : unit_signs))))
      return fail(ctx, 3, "stage sign pattern not supported by the v3 kernel");
    if (A.nsrc_tiles && !(sp && sp->mode)) return fail(ctx, 3, "v3 stage launch without its source lists");
    if (split) {
      // nodal pass (main stream) beside the regular-domain kernel (side stream),
      // then the nodal pull, which reads the interface sums of the latter
      Stage3Args B = args3(ctx, A, sp);
      if (ctx->split_shots < B.nshot || !ctx->d_D || !ctx->split_ok)
        return fail(ctx, 3, "split stage path without its buffers for %d shots", B.nshot);
      B.D = ctx->d_D;
      B.E = ctx->d_E;
      // all node blocks
      Stage3Args Ball = B;
      Ball.pb0 = 0;
      Ball.pnb = ctx->pnblk;
      Ball.pnbn = ctx->pnblk_nodal;
      if (ctx->concurrent) {
        // the nodal pass on a side stream, the deferred per-step kernels on the
        // other, the regular-domain kernel (the longer one) on the main stream,
        // which then runs the pull once both side streams have joined
        HIPC(hipEventRecord(ctx->ev_fork, ctx->stream));
        HIPC(hipStreamWaitEvent(ctx->cstream[0], ctx->ev_fork, 0));
        if (prof) HIPC(hipEventRecord(ev.start[1], ctx->cstream[0]));
        HIPC(launch_split_nodal(Ball, A.stage, A.transpose, ctx->cstream[0]));
        if (prof) HIPC(hipEventRecord(ev.end[1], ctx->cstream[0]));
        HIPC(hipEventRecord(ctx->ev_join[0], ctx->cstream[0]));
        const bool side_work = !ctx->deferred.empty();
        if (side_work) {
          // the deferred kernels read the state the stage reads and write their
          // own outputs only; they queue on the context's stream, here the second side stream
          HIPC(hipStreamWaitEvent(ctx->cstream[1], ctx->ev_fork, 0));
          std::swap(ctx->stream, ctx->cstream[1]);
          const int rc = run_deferred(ctx);
          std::swap(ctx->stream, ctx->cstream[1]);
          if (rc) return rc;
          HIPC(hipEventRecord(ctx->ev_join[1], ctx->cstream[1]));
        }
        if (prof) HIPC(hipEventRecord(ev.start[0], ctx->stream));
        HIPC(launch_split_rd(B, A.stage, A.transpose, ctx->d_tl_split, ctx->ntl_split, ctx->stream));
        HIPC(hipStreamWaitEvent(ctx->stream, ctx->ev_join[0], 0));
        if (side_work) HIPC(hipStreamWaitEvent(ctx->stream, ctx->ev_join[1], 0));
        HIPC(launch_split_pull(Ball, A.stage, A.transpose, ctx->stream));
        if (prof) HIPC(hipEventRecord(ev.end[0], ctx->stream));
      } else {
        if (int rc = run_deferred(ctx)) return rc;
        if (prof) HIPC(hipEventRecord(ev.start[0], ctx->stream));
        HIPC(launch_split_nodal(Ball, A.stage, A.transpose, ctx->stream));
        HIPC(launch_split_rd(B, A.stage, A.transpose, ctx->d_tl_split, ctx->ntl_split, ctx->stream));
        HIPC(launch_split_pull(Ball, A.stage, A.transpose, ctx->stream));
        if (prof) HIPC(hipEventRecord(ev.end[0], ctx->stream));
      }
      if (prof) ctx->ev_open.push_back(ev);
      return 0;
    }
    const Stage3Args B = args3(ctx, A, sp);
    if (ctx->concurrent) {
      // classes own disjoint nodes: classes 1, 2 run on the side streams
      HIPC(hipEventRecord(ctx->ev_fork, ctx->stream));
      for (int q = 0; q < 2; ++q) {
        HIPC(hipStreamWaitEvent(ctx->cstream[q], ctx->ev_fork, 0));
        if (prof) HIPC(hipEventRecord(ev.start[q + 1], ctx->cstream[q]));
        HIPC(launch_stage3(B, A.stage, A.transpose, q + 1, ctx->d_tl[q + 1], ctx->ntl[q + 1], ctx->cstream[q]));
        if (prof) HIPC(hipEventRecord(ev.end[q + 1], ctx->cstream[q]));
        HIPC(hipEventRecord(ctx->ev_join[q], ctx->cstream[q]));
      }
      if (prof) HIPC(hipEventRecord(ev.start[0], ctx->stream));
      HIPC(launch_stage3(B, A.stage, A.transpose, 0, ctx->d_tl[0], ctx->ntl[0], ctx->stream));
      if (prof) HIPC(hipEventRecord(ev.end[0], ctx->stream));
      if (int rc = run_deferred(ctx)) return rc;
      for (int q = 0; q < 2; ++q) HIPC(hipStreamWaitEvent(ctx->stream, ctx->ev_join[q], 0));
    } else {
      if (int rc = run_deferred(ctx)) return rc;
      if (prof) HIPC(hipEventRecord(ev.start[0], ctx->stream));
      for (int cls = 0; cls < 3; ++cls)
        HIPC(launch_stage3(B, A.stage, A.transpose, cls, ctx->d_tl[cls], ctx->ntl[cls], ctx->stream));
      if (prof) HIPC(hipEventRecord(ev.end[0], ctx->stream));
    }
  }
  if (prof) ctx->ev_open.push_back(ev);
  return 0;
}

StageArgs base_args(hr_ctx* ctx) {
  StageArgs A;
  std::memset(&A, 0, sizeof(A));
  A.g = ctx->g;
  A.sgn[0] = A.sgn[1] = A.sgn[2] = 1.0;
  return A;
}

// Tile grid and the three tile-class lists of the stage kernel.  A tile's
// window is its elements plus the low-side halo element in x, y and z.
int setup_tiles(hr_ctx* ctx) {
  Geo& g = ctx->g;
  g.tiles[0] = (g.nex + g.tx - 1) / g.tx;
  g.tiles[1] = (g.ney + g.ty - 1) / g.ty;
  // z-chunks of tz element layers (the v3 kernel reads the boundaries, so chunks could differ;
  // for the fused tile classes, aligning one with the bottom PML interface measured slower:
  // PML-only slabs cost as much).  The split path's tile kernel covers regular-domain
  // elements only: its chunks start at the regular domain's bottom layer (the layers
  // below are one chunk, run by the fused kernels of single-operator applies only)
  std::vector<int>& zb = ctx->zb;
  zb.assign(1, 0);
  const int z0 = (ctx->version == 4 && g.rlo[2] > 0 && g.rlo[2] < g.nez) ? g.rlo[2] : 0;
  if (z0 > 0) zb.push_back(z0);
  for (int z = z0 + g.tz; z < g.nez; z += g.tz) zb.push_back(z);
  zb.push_back(g.nez);
  if ((int)zb.size() - 1 > HR_MAX_ZCHUNKS) return fail(ctx, 1, "too many z-chunks (%d); raise tz", (int)zb.size() - 1);
  g.tiles[2] = (int)zb.size() - 1;
  std::vector<int> lists[3], rd_any;
  const int nt = ntiles(g);
  // list order: x fastest, then z, then y, so that the z-neighbour whose top slab is a
  // chunk's halo slab runs on the same XCD soon after (measured 2 % faster than x, y, z)
  for (int bb = 0; bb < nt; ++bb) {
    const int tx = bb % g.tiles[0], tz = (bb / g.tiles[0]) % g.tiles[2], ty = bb / (g.tiles[0] * g.tiles[2]);
    const int b = tx + g.tiles[0] * (ty + g.tiles[1] * tz);
    bool has_rd = false, has_pml = false;
    for (int ez = std::max(zb[tz] - 1, 0); ez < zb[tz + 1]; ++ez)
      for (int ey = std::max(ty * g.ty - 1, 0); ey < std::min((ty + 1) * g.ty, g.ney); ++ey)
        for (int ex = std::max(tx * g.tx - 1, 0); ex < std::min((tx + 1) * g.tx, g.nex); ++ex) {
          const bool rd = ex >= g.rlo[0] && ex < g.rhi[0] && ey >= g.rlo[1] && ey < g.rhi[1] && ez >= g.rlo[2];
          if (rd) has_rd = true; else has_pml = true;
        }
    lists[(has_rd && has_pml) ? 0 : (has_rd ? 1 : 2)].push_back(b);
    if (has_rd) rd_any.push_back(b);
  }
  // Each XCD (blockIdx mod 8) takes a contiguous run of a class list.  A Morton order of the
  // runs measured more fabric traffic.
  for (int c = 0; c < 3; ++c) {
    ctx->ntl[c] = (int)lists[c].size();
    if (int rc = upload(ctx, &ctx->d_tl[c], lists[c])) return rc;
  }
  // the split path's regular-domain kernel runs every tile with a regular-domain
  // element, in the same traversal order, so that each XCD's contiguous run mixes
  // interface and interior tiles as the domain does
  ctx->ntl_split = (int)rd_any.size();
  if (int rc = upload(ctx, &ctx->d_tl_split, rd_any)) return rc;
  // per-tile source lists of an earlier forward no longer match the tiles
  ctx->have_src = false;
  ctx->f_nstep = 0;
  ctx->f_history = HR_HISTORY_NONE;
  return 0;
}

int check_ready(hr_ctx* ctx) {
  if (!ctx) return 1;
  if (!ctx->have_pml) return fail(ctx, 2, "hr_set_pml has not been called");
  if (!ctx->have_material) return fail(ctx, 2, "hr_set_material has not been called");
  return 0;
}

// state buffers for nshot shots (shot-major); contents are not kept on growth
// flux records (y- and z-lines) and interface sums of the split path for
// shot_cap shots; the records of regular-domain-only nodes are never written and
// stay zero (the pull reads them)
int ensure_split(hr_ctx* ctx) {
  if (ctx->version != 4 || ctx->split_shots >= ctx->shot_cap) return 0;
  const size_t nd = 12 * (size_t)ctx->g.NN * ctx->shot_cap, ne = 3 * (size_t)ctx->g.NN * ctx->shot_cap;
  if (int rc = dalloc(ctx, &ctx->d_D, nd)) return rc;
  if (int rc = dalloc(ctx, &ctx->d_E, ne)) return rc;
  HIPC(hipMemsetAsync(ctx->d_D, 0, nd * sizeof(double), ctx->stream));
  ctx->split_shots = ctx->shot_cap;
  return 0;
}

int ensure_shots(hr_ctx* ctx, int nshot) {
  const size_t nf = field_size(ctx);
  if (nshot > ctx->shot_cap) {
    for (int a = 0; a < 3; ++a) {
      if (int rc = dalloc(ctx, &ctx->x[a], nf * nshot)) return rc;
      if (int rc = dalloc(ctx, &ctx->y[a], nf * nshot)) return rc;
      if (int rc = dalloc(ctx, &ctx->gs[a], nf * nshot)) return rc;
    }
    ctx->shot_cap = nshot;
  }
  ctx->nshot = nshot;
  return ensure_split(ctx);
}

int zero_state(hr_ctx* ctx) {
  const size_t n = field_size(ctx) * ctx->nshot;
  for (int a = 0; a < 3; ++a) {
    HIPC(hipMemsetAsync(ctx->x[a], 0, n * sizeof(double), ctx->stream));
    HIPC(hipMemsetAsync(ctx->y[a], 0, n * sizeof(double), ctx->stream));
    HIPC(hipMemsetAsync(ctx->gs[a], 0, n * sizeof(double), ctx->stream));
  }
  return 0;
}

// One RK4 step of every shot on (x) -> (y), then swap.  dir = +1 forward, -1 adjoint.
int rk4_step(hr_ctx* ctx, double dt, int dir, int transpose, const double sgn[3], const Src3Spec* sp4) {
  const double hh = dir * (0.5 * dt), h = dir * dt, h6 = dir * (dt / 6.0);
  for (int s = 1; s <= 4; ++s) {
    StageArgs A = base_args(ctx);
    A.stage = s;
    A.set_mask = 7;
    A.transpose = transpose;
    for (int a = 0; a < 3; ++a) A.sgn[a] = sgn[a];
    A.hh = hh; A.h = h; A.h6 = h6;
    A.in[0] = ctx->x[0]; A.in[1] = ctx->x[1]; A.in[2] = ctx->x[2];
    A.in[3] = ctx->gs[0]; A.in[4] = ctx->gs[1]; A.in[5] = ctx->gs[2];
    if (s < 4) A.out = ctx->gs[s - 1];
    else { A.xnew[0] = ctx->y[0]; A.xnew[1] = ctx->y[1]; A.xnew[2] = ctx->y[2]; }
    if (sp4 && sp4[s - 1].mode) A.nsrc_tiles = 1;
    if (int rc = launch_stage(ctx, A, sp4 ? &sp4[s - 1] : nullptr)) return rc;
  }
  std::swap(ctx->x, ctx->y);
  return 0;
}

// receiver traces of every shot at the start of step n
int record_step(hr_ctx* ctx, int n) {
  if (ctx->f_nrec <= 0) return 0;
  const size_t nf = field_size(ctx), rows = (size_t)ctx->f_nstep + 1;
  const StepRef sr{ctx->dyn ? ctx->d_step : nullptr, ctx->f_nrec};
  for (int q = 0; q < ctx->nshot; ++q)
    HIPC(launch_record(ctx->f_nrec, ctx->d_rec_slot, ctx->x[1] + q * nf,
                       ctx->d_traces + ((size_t)q * rows + n) * ctx->f_nrec, ctx->stream, sr));
  return 0;
}

// regular-domain displacement of every shot into history row `row` of blocks of `rows` rows per shot
int store_slice(hr_ctx* ctx, double* base, size_t rows, size_t row) {
  const size_t nf = field_size(ctx), slice = 3 * (size_t)ctx->nbox;
  const StepRef sr{ctx->dyn ? ctx->d_step : nullptr, (long long)slice};
  for (int q = 0; q < ctx->nshot; ++q)
    HIPC(launch_store_box(ctx->g, ctx->box0, ctx->boxn, ctx->x[1] + q * nf, base + ((size_t)q * rows + row) * slice,
                          ctx->stream, sr));
  return 0;
}

// forward sources of step n: B values times f(T), f(T + dt/2), f(T + dt)
void forward_sources(hr_ctx* ctx, int n, Src3Spec sp4[4]) {
  for (int s = 0; s < 4; ++s) {
    sp4[s] = Src3Spec();
    if (!ctx->have_src) continue;
    sp4[s].lists = &ctx->src3_fwd;
    sp4[s].mode = 1;
    sp4[s].step = n;
    sp4[s].fcol = (s == 0) ? 0 : (s == 3 ? 2 : 1);
  }
}

// Steps start, start + dir, ... (count of them) of a time loop whose body(n)
// queues step n.  Eager, or (graph_k > 0) through a hipGraph of graph_k steps
// captured once and replayed: the per-step values (source row, trace and history
// rows) then come from a device-side step index that each captured step advances
// by dir, the body is captured with n = 0 (ctx->dyn: n is an offset from that
// index), and the steps that do not fill a graph run eagerly.  graph_k is even,
// so the double-buffered state has the same parity at every replay.
int run_steps(hr_ctx* ctx, int start, int count, int dir, const std::function<int(int)>& body) {
  const int K = ctx->graph_k;
  int done = 0;
  if (K > 0 && count >= K && !ctx->profiling) {
    if (ctx->gexec || ctx->graph) {
      HIPC(hipStreamSynchronize(ctx->stream));
      if (ctx->gexec) (void)hipGraphExecDestroy(ctx->gexec);
      if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
      ctx->gexec = nullptr;
      ctx->graph = nullptr;
    }
    if (!ctx->d_step) if (int rc = dalloc(ctx, &ctx->d_step, 1)) return rc;
    HIPC(launch_step_set(ctx->d_step, start, ctx->stream));
    HIPC(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    ctx->dyn = 1;
    int rc = 0;
    for (int k = 0; k < K && !rc; ++k) {
      rc = body(0);
      if (!rc && launch_step_advance(ctx->d_step, dir, ctx->stream) != hipSuccess)
        rc = fail(ctx, 3, "step index launch failed in capture");
    }
    ctx->dyn = 0;
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(ctx->stream, &g);
    if (rc || ec != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      ctx->deferred.clear();
      return rc ? rc : fail(ctx, 3, "hipStreamEndCapture failed: %s", hipGetErrorString(ec));
    }
    ctx->graph = g;
    HIPC(hipGraphInstantiate(&ctx->gexec, g, nullptr, nullptr, 0));
    const int reps = count / K;
    for (int r = 0; r < reps; ++r) {
      HIPC(hipGraphLaunch(ctx->gexec, ctx->stream));
      ctx->graph_launches += 1;
    }
    done = reps * K;
  }
  for (int i = done; i < count; ++i)
    if (int rc = body(start + dir * i)) return rc;
  return 0;
}

int run_forward_steps(hr_ctx* ctx, int n0, int n1, double* seg_hist, size_t seg_rows) {
  // steps n0 .. n1-1 from the current state; stores slices n0..n1 into seg_hist when given
  const double sgn[3] = {1.0, 1.0, 1.0};
  for (int n = n0; n < n1; ++n) {
    // the slice of the state at the start of step n, queued beside its stage 1
    if (seg_hist) ctx->deferred.push_back([ctx, seg_hist, seg_rows, n, n0]() {
      return store_slice(ctx, seg_hist, seg_rows, (size_t)(n - n0));
    });
    Src3Spec sp4[4];
    forward_sources(ctx, n, sp4);
    if (int rc = rk4_step(ctx, ctx->f_dt, +1, 0, sgn, sp4)) return rc;
  }
  if (seg_hist) if (int rc = store_slice(ctx, seg_hist, seg_rows, (size_t)(n1 - n0))) return rc;
  return 0;
}

}  // namespace

// =================================================================== C ABI
extern "C" {

const char* hr_last_error(const hr_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

double hr_time_function(double T, int lcase, double fr) {
  const double pi = 3.14159265358979323846264338327950288419716939937510;
  switch (lcase) {
    case 1: return std::sin(2.0 * pi * fr * T);
    case 2: {
      const double wr = 2.0 * pi * fr;
      const double t_max = 6.0 * std::sqrt(6.0) / wr;
      const double u = wr * T - 3.0 * std::sqrt(6.0);
      if (T < t_max) {
        const double f1 = (0.25 * (u * u) - 0.5) * std::exp(-0.25 * (u * u)) - 13.0 * std::exp(-13.5);
        const double f2 = 0.50 + 13.0 * std::exp(-13.5);
        return f1 / f2;
      }
      return 0.0;
    }
    case 3: return std::exp(-(T - 0.06) * (T - 0.06) / 0.00040);
    case 4:
    case 5: {
      // Chopra's cosine wave and its derivative: t_d = 10 s, w_f = 3 rad/s, an
      // integer number of cycles nc (truncated), period stretched to fit t_d
      const double t_d = 10.0;
      const int nc = (int)(t_d / (2.0 * pi / 3.00) - 0.50);
      const double T_f = t_d / ((double)nc + 0.50);
      const double tn = (double)nc * T_f;
      if (T < T_f / 2.0)
        return lcase == 4 ? 0.5 * (1.0 - std::cos(2.0 * pi * T / T_f)) : (pi / T_f) * std::sin(2.0 * pi * T / T_f);
      if (T < tn)
        return lcase == 4 ? std::cos(2.0 * pi * (T - T_f / 2.0) / T_f)
                          : (-2.0 * pi / T_f) * std::sin(2.0 * pi * (T - T_f / 2.0) / T_f);
      if (T <= t_d)
        return lcase == 4 ? 0.50 * (1.0 - std::cos(2.0 * pi * (T - tn) / T_f)) - 1.0
                          : (pi / T_f) * std::sin(2.0 * pi * (T - tn) / T_f);
      return 0.0;
    }
    case 6: {
      // Greenberg's delta: Cauchy density of scale 1/10 centred at 5e-4 s
      const double a = 10.0 * (T - 0.0005);
      return 10.0 / (pi * (1.0 + a * a));
    }
    case 7: {
      // time derivative of the Ricker pulse of case 2
      const double wr = 2.0 * pi * fr;
      const double t_max = 6.0 * std::sqrt(6.0) / wr;
      const double u = wr * T - 3.0 * std::sqrt(6.0);
      if (T < t_max) {
        const double e = std::exp(-0.25 * (u * u));
        const double f1 = 0.25 * 2.0 * wr * u * e + (0.25 * (u * u) - 0.5) * (-0.25 * 2.0 * wr * u) * e;
        return f1 / (0.50 + 13.0 * std::exp(-13.5));
      }
      return 0.0;
    }
    case 8: return 1.0;
    case 9: return 0.0;
    case 10: return T <= 0.0005 ? T / 0.0005 : 1.0;   // smooth Heaviside
    default: return std::nan("");                      // the reference stops here
  }
}

int hr_ctx_create(const hr_mesh_desc* d, hr_ctx** out) {
  if (!out) return 1;
  *out = nullptr;
  hr_ctx* ctx = new hr_ctx;
  *out = ctx;
  if (!d) return fail(ctx, 1, "null mesh descriptor");
  if (d->nex < 2 || d->ney < 2 || d->nez < 2) return fail(ctx, 1, "mesh must have at least 2 elements per axis");
  // PML_PARAM divides by the layer thickness (PML_FINDER_3D), so every side needs a layer
  for (int a = 0; a < 5; ++a)
    if (d->npml[a] < 1) return fail(ctx, 1, "every side needs at least one PML element layer");
  if (d->npml[0] + d->npml[1] >= d->nex || d->npml[2] + d->npml[3] >= d->ney || d->npml[4] >= d->nez)
    return fail(ctx, 1, "PML layers leave no regular domain");
  if (!(d->h[0] > 0 && d->h[1] > 0 && d->h[2] > 0)) return fail(ctx, 1, "element edges must be positive");
  ctx->flags = d->flags;
  for (int a = 0; a < 3; ++a) ctx->hel[a] = d->h[a];
  Geo& g = ctx->g;
  g.nex = d->nex; g.ney = d->ney; g.nez = d->nez;
  g.NX = 2 * g.nex + 1; g.NY = 2 * g.ney + 1; g.NZ = 2 * g.nez + 1;
  g.NN = (long long)g.NX * g.NY * g.NZ;
  g.rlo[0] = d->npml[0]; g.rhi[0] = g.nex - d->npml[1];
  g.rlo[1] = d->npml[2]; g.rhi[1] = g.ney - d->npml[3];
  g.rlo[2] = d->npml[4]; g.rhi[2] = g.nez;
  g.top_rule = d->top_stress_bc ? 1 : 0;
  g.tx = 4;
  g.ty = DEFAULT_TY;
  g.tz = default_tz(ctx->version);
  // affine element: DJ = diag(h/2); DETJ and DJI as in ELEMENT_K_3D
  const double j11 = d->h[0] * 0.5, j22 = d->h[1] * 0.5, j33 = d->h[2] * 0.5;
  const double detj = j11 * j22 * j33 + 0.0 * 0.0 * 0.0 + 0.0 * 0.0 * 0.0 - 0.0 * j22 * 0.0 - 0.0 * 0.0 * j11 - j33 * 0.0 * 0.0;
  g.dxi[0] = (+j33 * j22 - 0.0 * 0.0) / detj;
  g.dxi[1] = (+j33 * j11 - 0.0 * 0.0) / detj;
  g.dxi[2] = (+j22 * j11 - 0.0 * 0.0) / detj;
  g.detj = detj;
  const double w[3] = {0.33333333333333333333, 1.33333333333333333333, 0.33333333333333333333};
  for (int c = 0; c < 3; ++c)
    for (int b = 0; b < 3; ++b)
      for (int a = 0; a < 3; ++a) g.wq[a + 3 * b + 9 * c] = (w[a] * w[b] * w[c]) * detj;
  // regular-domain node box
  for (int a = 0; a < 3; ++a) {
    ctx->box0[a] = 2 * g.rlo[a];
    ctx->boxn[a] = 2 * (g.rhi[a] - g.rlo[a]) + 1;
  }
  ctx->nbox = (long long)ctx->boxn[0] * ctx->boxn[1] * ctx->boxn[2];
  // equation numbering: lattice order, active components interleaved
  ctx->eq_slot.clear();
  for (int k = 0; k < g.NZ; ++k)
    for (int j = 0; j < g.NY; ++j)
      for (int i = 0; i < g.NX; ++i)
        for (int c = 0; c < NCOMP; ++c)
          if (h_active(g, i, j, k, c)) ctx->eq_slot.push_back(slot_of(g, c, lat(g, i, j, k)));
  ctx->neq = (long long)ctx->eq_slot.size();
  HIPC(hipStreamCreate(&ctx->stream));
  for (int q = 0; q < 2; ++q) {
    HIPC(hipStreamCreateWithFlags(&ctx->cstream[q], hipStreamNonBlocking));
    HIPC(hipEventCreateWithFlags(&ctx->ev_join[q], hipEventDisableTiming));
  }
  HIPC(hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming));
  // PML nodes of the split stage path; lattices whose flux records pass 32-bit
  // byte offsets use the fused tile classes
  if (96.0 * (double)g.NN < 4294967296.0) {
    // the nodal kernels walk the PML nodes in lattice order (sp_pml_node; measured
    // 20 % faster per gradient than with the interface nodes last); their
    // positions in that order, for the pull's per-block source lists
    int pos = 0;
    for (int k = 0; k < g.NZ; ++k)
      for (int j = 0; j < g.NY; ++j)
        for (int i = 0; i < g.NX; ++i)
          if (h_pml_node(g, i, j, k)) ctx->pnode_pos.push_back({lat(g, i, j, k), pos++});
    int gl[3] = {g.rlo[0], g.rlo[1], g.rlo[2]}, gh[2] = {g.rhi[0], g.rhi[1]};
    if ((unsigned)pos != sp_pml_count(g.NX, g.NY, g.NZ, gl, gh)) return fail(ctx, 3, "PML node count mismatch");
    ctx->pnblk = (pos + SPLIT_NT - 1) / SPLIT_NT;
    ctx->pnblk_nodal = (pos + SPLIT_OWN - 1) / SPLIT_OWN;
    ctx->split_ok = true;
  } else {
    ctx->version = 3;
    g.tz = default_tz(ctx->version);
  }
  if (int rc = setup_tiles(ctx)) return rc;
  if (int rc = upload(ctx, &ctx->d_eq_slot, ctx->eq_slot)) return rc;
  // coordinates (origin: centred in x, y; top surface at z = 0)
  std::vector<double> xc(g.NX), yc(g.NY), zc(g.NZ);
  const double xmin = -0.5 * g.nex * d->h[0], ymin = -0.5 * g.ney * d->h[1], zmin = -(double)g.nez * d->h[2];
  for (int i = 0; i < g.NX; ++i) xc[i] = xmin + (double)i * (0.5 * d->h[0]);
  for (int j = 0; j < g.NY; ++j) yc[j] = ymin + (double)j * (0.5 * d->h[1]);
  for (int k = 0; k < g.NZ; ++k) zc[k] = zmin + (double)k * (0.5 * d->h[2]);
  if (int rc = upload(ctx, &ctx->d_xc, xc)) return rc;
  if (int rc = upload(ctx, &ctx->d_yc, yc)) return rc;
  if (int rc = upload(ctx, &ctx->d_zc, zc)) return rc;
  const size_t nf = field_size(ctx);
  for (int a = 0; a < 3; ++a) {
    if (int rc = dalloc(ctx, &ctx->x[a], nf)) return rc;
    if (int rc = dalloc(ctx, &ctx->y[a], nf)) return rc;
    if (int rc = dalloc(ctx, &ctx->gs[a], nf)) return rc;
  }
  if (int rc = dalloc(ctx, &ctx->tmp, nf)) return rc;
  if (int rc = dalloc(ctx, &ctx->tmp2, nf)) return rc;
  if (!(ctx->flags & HR_DEVICE_PTRS)) {
    // staging of host vectors for hr_apply, so that the product allocates nothing per call
    if (int rc = dalloc(ctx, &ctx->d_ax, (size_t)ctx->neq)) return rc;
    if (int rc = dalloc(ctx, &ctx->d_ay, (size_t)ctx->neq)) return rc;
  }
  HIPC(hipMemset(ctx->tmp, 0, nf * sizeof(double)));
  HIPC(hipMemset(ctx->tmp2, 0, nf * sizeof(double)));
  if (int rc = ensure_split(ctx)) return rc;
  if (int rc = zero_state(ctx)) return rc;
  HIPC(hipStreamSynchronize(ctx->stream));
  return 0;
}

int hr_ctx_destroy(hr_ctx* ctx) {
  if (!ctx) return 0;
  auto F = [](void* p) { if (p) (void)hipFree(p); };
  F(ctx->d_eq_slot); F(ctx->d_xc); F(ctx->d_yc); F(ctx->d_zc);
  F(ctx->d_node); F(ctx->d_tab); F(ctx->d_ftab);
  for (int a = 0; a < 3; ++a) { F(ctx->x[a]); F(ctx->y[a]); F(ctx->gs[a]); F(ctx->shell[a].l); F(ctx->shell[a].r); }
  F(ctx->tmp); F(ctx->tmp2); F(ctx->d_traces); F(ctx->d_rec_slot); F(ctx->d_hist); F(ctx->d_seg);
  for (double* p : ctx->snap) F(p);
  for (int a = 0; a < 6; ++a) F(ctx->adj_buf[a]);
  F(ctx->d_mis_ptr); F(ctx->d_mis_idx); F(ctx->d_ax); F(ctx->d_ay);
  F(ctx->d_mrd); F(ctx->d_minvu); F(ctx->d_rwp); F(ctx->d_energy); F(ctx->d_epart); F(ctx->d_uslice);
  F(ctx->d_src_tile_ptr); F(ctx->d_src_node); F(ctx->d_src_comp); F(ctx->d_src_m); F(ctx->d_src_val);
  F(ctx->d_rec_tile_ptr); F(ctx->d_rec_node); F(ctx->d_rec_comp); F(ctx->d_rec_m); F(ctx->d_rec_perm);
  F(ctx->d_obs); F(ctx->d_gl); F(ctx->d_gm); F(ctx->d_part); F(ctx->d_scalar); F(ctx->d_tr_sorted);
  F(ctx->d_obs_sorted);
  F(ctx->d_img); F(ctx->d_restr); F(ctx->d_reg); F(ctx->d_cost_el); F(ctx->d_red); F(ctx->d_pm); F(ctx->d_pg);
  for (int c = 0; c < 3; ++c) F(ctx->d_tl[c]);
  F(ctx->d_tl_split); F(ctx->d_D); F(ctx->d_E);
  F(ctx->d_psrc_ptr); F(ctx->d_psrc_idx); F(ctx->d_prec_ptr); F(ctx->d_prec_idx);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->gexec) (void)hipGraphExecDestroy(ctx->gexec);
  if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
  F(ctx->d_step);
  for (auto& p : ctx->ev_open)
    for (int q = 0; q < p.n; ++q) { (void)hipEventDestroy(p.start[q]); (void)hipEventDestroy(p.end[q]); }
  for (hipEvent_t e : ctx->ev_pool) (void)hipEventDestroy(e);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  for (int q = 0; q < 2; ++q) {
    if (ctx->cstream[q]) (void)hipStreamDestroy(ctx->cstream[q]);
    if (ctx->ev_join[q]) (void)hipEventDestroy(ctx->ev_join[q]);
  }
  if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
  delete ctx;
  return 0;
}

int hr_get_sizes(hr_ctx* ctx, long long* neq, long long* nnode, long long* nstore) {
  if (!ctx) return 1;
  if (neq) *neq = ctx->neq;
  if (nnode) *nnode = ctx->g.NN;
  if (nstore) *nstore = 3 * ctx->nbox;
  return 0;
}

int hr_get_store_eq(hr_ctx* ctx, long long* store_eq) {
  if (!ctx || !store_eq) return fail(ctx, 1, "null argument");
  // inverse numbering for the box
  const Geo& g = ctx->g;
  std::vector<long long> inv(field_size(ctx), -1);
  for (long long e = 0; e < ctx->neq; ++e) inv[ctx->eq_slot[e]] = e;
  long long t = 0;
  for (int c = 0; c < 3; ++c)
    for (int bk = 0; bk < ctx->boxn[2]; ++bk)
      for (int bj = 0; bj < ctx->boxn[1]; ++bj)
        for (int bi = 0; bi < ctx->boxn[0]; ++bi)
          store_eq[t++] = inv[slot_of(g, c, lat(g, ctx->box0[0] + bi, ctx->box0[1] + bj, ctx->box0[2] + bk))];
  return 0;
}

int hr_set_pml(hr_ctx* ctx, const double* p) {
  if (!ctx || !p) return fail(ctx, 1, "null argument");
  for (int f = 0; f < 6; ++f)
    if (f != 4 && !(p[12 + f] > 0.0)) return fail(ctx, 1, "PML length of face %d must be positive", f + 1);
  // the reference stops in a top PML zone (PML.F90:629-633): no node may lie above z0 of face 5
  if (!(p[22] >= 0.0)) return fail(ctx, 1, "top PML zone (face 5) would contain nodes");
  std::memcpy(ctx->g.pml, p, sizeof(double) * 24);
  ctx->have_pml = true;
  // the mass inverses and stretch tables depend on PML_PARAM: hr_set_material must follow
  ctx->have_material = false;
  return 0;
}

int hr_set_material(hr_ctx* ctx, const double* lam, const double* mu, const double* rho_el) {
  if (!ctx || !lam || !mu || !rho_el) return fail(ctx, 1, "null argument");
  if (!ctx->have_pml) return fail(ctx, 2, "call hr_set_pml before hr_set_material");
  const Geo& g = ctx->g;
  const size_t NN = (size_t)g.NN, nel = (size_t)g.nex * g.ney * g.nez;
  std::vector<double> hl(NN), hm(NN), hr(nel);
  if (ctx->flags & HR_DEVICE_PTRS) {
    HIPC(hipMemcpy(hl.data(), lam, NN * sizeof(double), hipMemcpyDeviceToHost));
    HIPC(hipMemcpy(hm.data(), mu, NN * sizeof(double), hipMemcpyDeviceToHost));
    HIPC(hipMemcpy(hr.data(), rho_el, nel * sizeof(double), hipMemcpyDeviceToHost));
  } else {
    std::memcpy(hl.data(), lam, NN * sizeof(double));
    std::memcpy(hm.data(), mu, NN * sizeof(double));
    std::memcpy(hr.data(), rho_el, nel * sizeof(double));
  }
  for (size_t e = 0; e < nel; ++e)
    if (!(hr[e] > 0.0)) return fail(ctx, 1, "element %zu has non-positive density", e);
  std::vector<double> mrd(NN, 0.0), rwp(NN, 0.0), wps(NN, 0.0);
  for (int ez = 0; ez < g.nez; ++ez)
    for (int ey = 0; ey < g.ney; ++ey)
      for (int ex = 0; ex < g.nex; ++ex) {
        const size_t e = (size_t)ex + (size_t)g.nex * ((size_t)ey + (size_t)g.ney * ez);
        const bool rd = ex >= g.rlo[0] && ex < g.rhi[0] && ey >= g.rlo[1] && ey < g.rhi[1] && ez >= g.rlo[2];
        for (int c = 0; c < 3; ++c)
          for (int b = 0; b < 3; ++b)
            for (int a = 0; a < 3; ++a) {
              const long long n = lat(g, 2 * ex + a, 2 * ey + b, 2 * ez + c);
              const double W = g.wq[a + 3 * b + 9 * c];
              if (rd) mrd[n] += hr[e] * W;
              else { rwp[n] += hr[e] * W; wps[n] += W; }
            }
      }
  // coordinates
  std::vector<double> xc(g.NX), yc(g.NY), zc(g.NZ);
  HIPC(hipMemcpy(xc.data(), ctx->d_xc, g.NX * sizeof(double), hipMemcpyDeviceToHost));
  HIPC(hipMemcpy(yc.data(), ctx->d_yc, g.NY * sizeof(double), hipMemcpyDeviceToHost));
  HIPC(hipMemcpy(zc.data(), ctx->d_zc, g.NZ * sizeof(double), hipMemcpyDeviceToHost));
  ctx->h_minv_u.assign(NN, 0.0);
  ctx->h_a.assign(NN, 1.0);
  for (int k = 0; k < g.NZ; ++k)
    for (int j = 0; j < g.NY; ++j)
      for (int i = 0; i < g.NX; ++i) {
        const long long n = lat(g, i, j, k);
        double s[6];
        h_stretch(g.pml, xc[i], yc[j], zc[k], s);
        const double a = s[0] * s[1] * s[2];
        ctx->h_a[n] = a;
        const double M = mrd[n] + a * rwp[n];
        ctx->h_minv_u[n] = M > 0.0 ? 1.0 / M : 0.0;
      }
  ctx->h_wps = wps;
  // v3 tables: node table [(lam, mu) pairs | (minv_u, minv_s, rwp, wps) quadruples]
  // and [wq | per-axis stretch (a, b) pairs]
  {
    std::vector<double> nt((size_t)NT_COUNT * NN, 0.0);
    double* quad = nt.data() + (size_t)NT_PAIR * NN;
    for (size_t n = 0; n < NN; ++n) {
      nt[2 * n] = hl[n];
      nt[2 * n + 1] = hm[n];
      quad[4 * n] = ctx->h_minv_u[n];
      quad[4 * n + 1] = wps[n] > 0.0 ? 1.0 / (ctx->h_a[n] * wps[n]) : 0.0;
      quad[4 * n + 2] = rwp[n];
      quad[4 * n + 3] = wps[n];
    }
    if (int rc = upload(ctx, &ctx->d_node, nt)) return rc;
    std::vector<double> tb(TAB_STRETCH + 2 * (size_t)(g.NX + g.NY + g.NZ), 0.0);
    for (int q = 0; q < 27; ++q) tb[q] = g.wq[q];
    double* ax = tb.data() + TAB_STRETCH;
    double* ay = ax + 2 * g.NX;
    double* az = ay + 2 * g.NY;
    for (int i = 0; i < g.NX; ++i) { double s[6]; h_stretch(g.pml, xc[i], yc[0], zc[g.NZ - 1], s); ax[2 * i] = s[0]; ax[2 * i + 1] = s[3]; }
    for (int j = 0; j < g.NY; ++j) { double s[6]; h_stretch(g.pml, xc[0], yc[j], zc[g.NZ - 1], s); ay[2 * j] = s[1]; ay[2 * j + 1] = s[4]; }
    for (int k = 0; k < g.NZ; ++k) { double s[6]; h_stretch(g.pml, xc[0], yc[0], zc[k], s); az[2 * k] = s[2]; az[2 * k + 1] = s[5]; }
    if (int rc = upload(ctx, &ctx->d_tab, tb)) return rc;
  }
  if (int rc = upload(ctx, &ctx->d_mrd, mrd)) return rc;
  // the split path's nodal kernels read M^-1 of the u rows and the PML rho W sum
  // from arrays of their own (8 bytes a node each); the stress rows' M^-1 and
  // weight sum follow from the stretch factors and the element count
  if (int rc = upload(ctx, &ctx->d_minvu, ctx->h_minv_u)) return rc;
  if (int rc = upload(ctx, &ctx->d_rwp, rwp)) return rc;
  ctx->have_material = true;
  // the history and traces of an earlier forward belong to the old operator
  ctx->f_nstep = 0;
  ctx->f_history = HR_HISTORY_NONE;
  return 0;
}

int hr_get_diag_m(hr_ctx* ctx, double* m) {
  if (int rc = check_ready(ctx)) return rc;
  const Geo& g = ctx->g;
  std::vector<double> hm(ctx->neq);
  for (long long e = 0; e < ctx->neq; ++e) {
    const long long slot = ctx->eq_slot[e];
    int c;
    long long n;
    slot_decode(g, slot, c, n);
    if (c < 3) hm[e] = 1.0 / ctx->h_minv_u[n];
    else hm[e] = (c < 6 ? 1.0 : 2.0) * ctx->h_a[n] * ctx->h_wps[n];
  }
  if (ctx->flags & HR_DEVICE_PTRS) HIPC(hipMemcpy(m, hm.data(), hm.size() * sizeof(double), hipMemcpyHostToDevice));
  else std::memcpy(m, hm.data(), hm.size() * sizeof(double));
  return 0;
}

int hr_scale(hr_ctx* ctx, int which, double s) {
  if (!ctx || which < 0 || which > 2) return fail(ctx, 1, "bad operator index");
  ctx->shell[which].s *= s;
  return 0;
}

int hr_reset_scaling(hr_ctx* ctx, int which) {
  if (!ctx || which < 0 || which > 2) return fail(ctx, 1, "bad operator index");
  Shell& S = ctx->shell[which];
  S.s = 1.0;
  dfree(ctx, &S.l);
  dfree(ctx, &S.r);
  return 0;
}

int hr_diagonal_scale(hr_ctx* ctx, int which, const double* l, const double* r) {
  if (!ctx || which < 0 || which > 2) return fail(ctx, 1, "bad operator index");
  Shell& S = ctx->shell[which];
  const size_t n = (size_t)ctx->neq;
  auto apply = [&](double** dst, const double* v) -> int {
    std::vector<double> cur(n, 1.0), in(n);
    if (*dst) HIPC(hipMemcpy(cur.data(), *dst, n * sizeof(double), hipMemcpyDeviceToHost));
    if (ctx->flags & HR_DEVICE_PTRS) HIPC(hipMemcpy(in.data(), v, n * sizeof(double), hipMemcpyDeviceToHost));
    else std::memcpy(in.data(), v, n * sizeof(double));
    for (size_t e = 0; e < n; ++e) cur[e] = cur[e] * in[e];
    if (!*dst) if (int rc = dalloc(ctx, dst, n)) return rc;
    HIPC(hipMemcpy(*dst, cur.data(), n * sizeof(double), hipMemcpyHostToDevice));
    return 0;
  };
  if (l) if (int rc = apply(&S.l, l)) return rc;
  if (r) if (int rc = apply(&S.r, r)) return rc;
  return 0;
}

int hr_apply(hr_ctx* ctx, int which, int transpose, const double* x, const double* y, double* z) {
  if (int rc = check_ready(ctx)) return rc;
  if (which < 0 || which > 2) return fail(ctx, 1, "bad operator index %d", which);
  if (!x || !y || !z) return fail(ctx, 1, "null vector");
  const Shell& S = ctx->shell[which];
  const long long neq = ctx->neq;
  const bool dev = ctx->flags & HR_DEVICE_PTRS;
  const double *dx = x, *dy = y;
  double* dz = z;
  if (!dev) {
    // host vectors through the staging buffers allocated with the context
    HIPC(hipMemcpyAsync(ctx->d_ax, x, neq * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HIPC(hipMemcpyAsync(ctx->d_ay, y, neq * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    dx = ctx->d_ax;
    dy = ctx->d_ay;
    dz = ctx->d_ay;
  }
  // A_eff x = s l (A (r x)) ; A_eff^T x = s r (A^T (l x))
  const double* pre = transpose ? S.l : S.r;
  const double* post = transpose ? S.r : S.l;
  HIPC(launch_gather_eq(neq, ctx->d_eq_slot, dx, pre, ctx->tmp, ctx->stream));
  StageArgs A = base_args(ctx);
  A.stage = 0;
  A.set_mask = which == HR_G ? 1 : (which == HR_K ? 2 : 4);
  A.transpose = transpose;
  A.in[0] = ctx->tmp;
  A.out = ctx->tmp2;
  if (int rc = launch_stage(ctx, A)) return rc;
  HIPC(launch_scatter_eq(neq, ctx->d_eq_slot, ctx->tmp2, dy, S.s, post, dz, ctx->stream));
  if (!dev) {
    HIPC(hipMemcpyAsync(z, dz, neq * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPC(hipStreamSynchronize(ctx->stream));
  }
  return 0;
}

int hr_synchronize(hr_ctx* ctx) {
  if (!ctx) return 1;
  HIPC(hipStreamSynchronize(ctx->stream));
  return 0;
}

int hr_set_tiling(hr_ctx* ctx, int ty, int tz, int version) {
  if (!ctx) return 1;
  const int serial = (version & HR_TILING_SERIAL_CLASSES) ? 1 : 0;
  version &= ~HR_TILING_SERIAL_CLASSES;
  if (ty < 0 || tz < 0 || (version != 0 && version != 3 && version != 4))
    return fail(ctx, 1, "bad tiling (stage kernel version 3 or 4)");
  if (version == 4 && !ctx->split_ok) return fail(ctx, 1, "lattice too large for the split stage path (version 4)");
  Geo& g = ctx->g;
  ctx->concurrent = serial ? 0 : 1;
  const int t = ty ? ty : DEFAULT_TY;
  if (t != 3) return fail(ctx, 1, "stage kernel tiles are 4 x 3 elements");
  if (version) ctx->version = version;
  g.tx = 4;
  g.ty = t;
  g.tz = tz ? tz : default_tz(ctx->version);
  if (int rc = ensure_split(ctx)) return rc;
  return setup_tiles(ctx);
}

int hr_get_tiling(hr_ctx* ctx, int* ty, int* tz, int* version) {
  if (!ctx) return 1;
  if (ty) *ty = ctx->g.ty;
  if (tz) *tz = ctx->g.tz;
  if (version) *version = ctx->version;
  return 0;
}

int hr_set_graph(hr_ctx* ctx, int steps) {
  if (!ctx) return 1;
  if (steps < 0 || (steps & 1) || steps > 4096) return fail(ctx, 1, "steps per graph must be 0 or an even count <= 4096");
  ctx->graph_k = steps;
  return 0;
}

int hr_get_graph_launches(hr_ctx* ctx, long long* launches) {
  if (!ctx || !launches) return 1;
  *launches = ctx->graph_launches;
  return 0;
}

}  // extern "C"

namespace {

// One step of the central-difference variant (EXPLICIT_Central_Difference_3,
// Obsolete.F90:815-995) on the lattice state: x[1] = u_n, x[0] = u_{n-1},
// x[2] = u_his; writes u_{n+1} to y[1] and u_his + dt u_{n+1} to y[2], then
// rotates (u_{n-1}, u_n, u_his) <- (u_n, u_{n+1}, new u_his).
int cd_step(hr_ctx* ctx, double dt, int n) {
  StageArgs A = base_args(ctx);
  A.stage = 5;
  A.set_mask = 7;
  A.hh = dt * dt; A.h = dt; A.h6 = dt * dt * dt;
  A.in[0] = ctx->x[1]; A.in[1] = ctx->x[0]; A.in[2] = ctx->x[2];
  A.xnew[0] = ctx->y[1]; A.xnew[1] = ctx->y[2];
  Src3Spec sp;
  if (ctx->have_src) {
    A.nsrc_tiles = 1;
    sp.lists = &ctx->src3_fwd;
    sp.mode = 1;
    sp.step = n;
    sp.fcol = 0;
  }
  if (int rc = launch_stage(ctx, A, ctx->have_src ? &sp : nullptr)) return rc;
  double* nm = ctx->x[0];
  double* his = ctx->x[2];
  ctx->x[0] = ctx->x[1];
  ctx->x[1] = ctx->y[1];
  ctx->x[2] = ctx->y[2];
  ctx->y[1] = nm;
  ctx->y[2] = his;
  return 0;
}

// Bytes of one history slice (regular-domain displacement of every shot) and
// of one snapshot of the RK4 state (x1, x2, x3 of every shot).
size_t slice_bytes(const hr_ctx* ctx, int nshot) { return 3 * (size_t)ctx->nbox * nshot * sizeof(double); }
size_t state_bytes(const hr_ctx* ctx, int nshot) { return 3 * field_size(ctx) * nshot * sizeof(double); }

// Plan the forward history for the adjoint: every slice (store), uniform
// segment checkpoints, or the binomial schedule that fits the HBM budget with
// the fewest recomputed steps.  Sets h_store, h_seglen, h_nseg, h_nsnap, h_plan.
int plan_history(hr_ctx* ctx, int history, int segment, int nstep, int nshot) {
  ctx->h_store = false;
  ctx->h_seglen = ctx->h_nseg = ctx->h_nsnap = 0;
  if (history == HR_HISTORY_NONE) return 0;
  if (history == HR_HISTORY_STORE) {
    ctx->h_store = true;
    return 0;
  }
  const double sl = (double)slice_bytes(ctx, nshot), st = (double)state_bytes(ctx, nshot);
  if (history == HR_HISTORY_CHECKPOINT) {
    const int L = std::min(segment, nstep);
    const int m = (nstep + L - 1) / L;
    ctx->h_seglen = L;
    ctx->h_nseg = m;
    ctx->h_plan.build(m, m - 1);          // a snapshot at every segment start
    ctx->h_nsnap = (int)ctx->h_plan.chain().size();
    return 0;
  }
  // HR_HISTORY_REVOLVE
  double budget = ctx->h_budget;
  if (!(budget > 0.0)) {
    size_t fr = 0, tot = 0;
    HIPC(hipMemGetInfo(&fr, &tot));
    double held = 0.0;
    for (auto& kv : ctx->caps)
      if (kv.first == (void*)&ctx->d_hist || kv.first == (void*)&ctx->d_seg) held += (double)kv.second;
    held += (double)ctx->snap.size() * (double)ctx->snap_elems * sizeof(double);
    budget = 0.5 * ((double)fr + held);
  }
  if ((double)(nstep + 1) * sl <= budget) {
    ctx->h_store = true;                  // the whole history fits: no recomputation
    return 0;
  }
  long long best = -1;
  int bestL = 0, bests = 0;
  for (int L = 4; L <= 4096; L *= 2) {
    if (L >= nstep) break;
    const int m = (nstep + L - 1) / L;
    if (m > 2048) continue;
    const double rem = budget - (double)(L + 1) * sl;
    if (rem < 0.0) continue;
    const int s = (int)std::min<double>(std::floor(rem / st), (double)(m - 1));
    RevolvePlan P;
    P.build(m, s);
    const long long steps = (long long)nstep + (long long)L * P.cost();
    if (best < 0 || steps < best) { best = steps; bestL = L; bests = s; }
  }
  if (best < 0)
    return fail(ctx, 1, "history budget of %.3g bytes is too small (one slice is %.3g bytes, one snapshot %.3g)",
                budget, sl, st);
  ctx->h_seglen = bestL;
  ctx->h_nseg = (nstep + bestL - 1) / bestL;
  ctx->h_plan.build(ctx->h_nseg, bests);
  ctx->h_nsnap = (int)ctx->h_plan.chain().size();
  return 0;
}

// snapshot pool and slice buffer of the planned history (allocated before the time loop)
int ensure_history_buffers(hr_ctx* ctx, int nshot) {
  if (ctx->h_store || ctx->h_seglen == 0) return 0;
  const size_t elems = 3 * field_size(ctx) * nshot;
  if (elems > ctx->snap_elems) {
    for (double* p : ctx->snap) (void)hipFree(p);
    ctx->snap.clear();
    ctx->snap_elems = elems;
  }
  // the schedule never holds more snapshots than its slots
  const size_t need = (size_t)std::max(ctx->h_plan.s, ctx->h_nsnap);
  while (ctx->snap.size() < need) {
    double* p = nullptr;
    hipError_t e = hipMalloc((void**)&p, elems * sizeof(double));
    if (e != hipSuccess) return fail(ctx, (int)e, "snapshot allocation failed: %s", hipGetErrorString(e));
    ctx->snap.push_back(p);
  }
  const size_t seg = (size_t)nshot * (ctx->h_seglen + 1) * 3 * (size_t)ctx->nbox;
  if (int rc = dalloc(ctx, &ctx->d_seg, seg)) return rc;
  return 0;
}

// snapshot <-> the three state fields of every shot
int snap_save(hr_ctx* ctx, double* const* fields, double* snap, int nshot) {
  const size_t nfs = field_size(ctx) * nshot;
  for (int a = 0; a < 3; ++a)
    HIPC(hipMemcpyAsync(snap + a * nfs, fields[a], nfs * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
  return 0;
}
int snap_load(hr_ctx* ctx, double* const* fields, const double* snap, int nshot) {
  const size_t nfs = field_size(ctx) * nshot;
  for (int a = 0; a < 3; ++a)
    HIPC(hipMemcpyAsync(fields[a], snap + a * nfs, nfs * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
  return 0;
}

// scheme 0: RK4 (hr_forward_shots); scheme 1: central difference (hr_forward_cd,
// one shot, ftab already expanded to dt^2 f(n dt) in column 0 on the host)
int forward_impl(hr_ctx* ctx, int nshot, int nstep, double dt, const double* ftab, int nsrc, const int* src_shot,
                 const long long* src_eq, const double* src_val, int nrec, const long long* rec_eq, double* traces,
                 int history, int segment, int scheme) {
  if (int rc = check_ready(ctx)) return rc;
  ctx->deferred.clear();   // work left by a solve that stopped on an error
  if (nshot < 1) return fail(ctx, 1, "nshot must be >= 1");
  if (nstep < 1 || !(dt > 0.0)) return fail(ctx, 1, "nstep must be >= 1 and dt > 0");
  if (!ftab) return fail(ctx, 1, "null time-function table");
  if (nsrc < 0 || nrec < 0) return fail(ctx, 1, "negative counts");
  if (nsrc > 0 && (!src_eq || !src_val || (nshot > 1 && !src_shot))) return fail(ctx, 1, "null source arrays");
  if (nrec > 0 && !rec_eq) return fail(ctx, 1, "null receiver array");
  if (history < 0 || history > 3) return fail(ctx, 1, "bad history mode");
  if (history == HR_HISTORY_CHECKPOINT && segment < 1) return fail(ctx, 1, "checkpoint segment must be >= 1");
  const bool dev = ctx->flags & HR_DEVICE_PTRS;
  // host copies of the small index arrays
  std::vector<long long> seq(nsrc), req(nrec);
  std::vector<double> sval(nsrc);
  std::vector<int> sshot(nsrc, 0);
  if (nsrc) {
    if (dev) {
      HIPC(hipMemcpy(seq.data(), src_eq, nsrc * sizeof(long long), hipMemcpyDeviceToHost));
      HIPC(hipMemcpy(sval.data(), src_val, nsrc * sizeof(double), hipMemcpyDeviceToHost));
      if (src_shot) HIPC(hipMemcpy(sshot.data(), src_shot, nsrc * sizeof(int), hipMemcpyDeviceToHost));
    } else {
      std::memcpy(seq.data(), src_eq, nsrc * sizeof(long long));
      std::memcpy(sval.data(), src_val, nsrc * sizeof(double));
      if (src_shot) std::memcpy(sshot.data(), src_shot, nsrc * sizeof(int));
    }
  }
  if (nrec) {
    if (dev) HIPC(hipMemcpy(req.data(), rec_eq, nrec * sizeof(long long), hipMemcpyDeviceToHost));
    else std::memcpy(req.data(), rec_eq, nrec * sizeof(long long));
  }
  for (int q = 0; q < nsrc; ++q) {
    if (seq[q] < 0 || seq[q] >= ctx->neq) return fail(ctx, 1, "source equation %lld out of range", seq[q]);
    if (ctx->eq_slot[seq[q]] >= 3 * ctx->g.NN) return fail(ctx, 1, "source equation %lld is not a displacement", seq[q]);
    if (sshot[q] < 0 || sshot[q] >= nshot) return fail(ctx, 1, "source shot %d out of range", sshot[q]);
  }
  for (int q = 0; q < nrec; ++q) {
    if (req[q] < 0 || req[q] >= ctx->neq) return fail(ctx, 1, "receiver equation %lld out of range", req[q]);
    if (ctx->eq_slot[req[q]] >= 3 * ctx->g.NN) return fail(ctx, 1, "receiver equation %lld is not a displacement", req[q]);
  }
  const size_t nft = 3 * (size_t)nstep * nshot;
  ctx->ftab.resize(nft);
  if (dev && scheme == 0) HIPC(hipMemcpy(ctx->ftab.data(), ftab, nft * sizeof(double), hipMemcpyDeviceToHost));
  else std::memcpy(ctx->ftab.data(), ftab, nft * sizeof(double));
  if (ctx->ftab_cap < nft) {
    if (int rc = dalloc(ctx, &ctx->d_ftab, nft)) return rc;
    ctx->ftab_cap = nft;
  }
  HIPC(hipMemcpy(ctx->d_ftab, ctx->ftab.data(), nft * sizeof(double), hipMemcpyHostToDevice));
  if (int rc = ensure_shots(ctx, nshot)) return rc;
  // sources grouped by shot, then tile
  if (nsrc) {
    const int nt = ntiles(ctx->g);
    std::vector<int> ptr(1, 0);
    std::vector<long long> node;
    std::vector<int> comp;
    std::vector<double> m, v;
    for (int q = 0; q < nshot; ++q) {
      std::vector<long long> eq;
      std::vector<double> val;
      for (int e = 0; e < nsrc; ++e)
        if (sshot[e] == q) { eq.push_back(seq[e]); val.push_back(sval[e]); }
      SrcBuild sb = build_src(ctx, (int)eq.size(), eq.data());
      const int base = (int)node.size();
      for (int t = 0; t < nt; ++t) ptr.push_back(base + sb.ptr[t + 1]);
      for (size_t e = 0; e < eq.size(); ++e) {
        node.push_back(sb.node[e]);
        comp.push_back(sb.comp[e]);
        m.push_back(sb.m[e]);
        v.push_back(val[sb.perm[e]]);
      }
    }
    if (int rc = upload(ctx, &ctx->d_src_tile_ptr, ptr)) return rc;
    if (int rc = upload(ctx, &ctx->d_src_node, node)) return rc;
    if (int rc = upload(ctx, &ctx->d_src_comp, comp)) return rc;
    if (int rc = upload(ctx, &ctx->d_src_m, m)) return rc;
    if (int rc = upload(ctx, &ctx->d_src_val, v)) return rc;
    Src3 h3;
    std::memset(&h3, 0, sizeof(h3));
    h3.tile_ptr = ctx->d_src_tile_ptr; h3.node = ctx->d_src_node; h3.comp = ctx->d_src_comp;
    h3.m = ctx->d_src_m; h3.val = ctx->d_src_val; h3.ftab = ctx->d_ftab; h3.nstep = nstep; h3.ntiles = nt;
    if (ctx->split_ok)
      if (int rc = build_pull_blocks(ctx, ptr, node, nshot, nt, &ctx->d_psrc_ptr, &ctx->d_psrc_idx, h3)) return rc;
    ctx->src3_fwd = h3;
    ctx->have_src = true;
  } else {
    ctx->have_src = false;
  }
  // receivers
  ctx->f_nshot = nshot; ctx->f_nstep = nstep; ctx->f_dt = dt; ctx->f_nrec = nrec; ctx->f_history = history;
  ctx->f_segment = segment;
  ctx->rec_slot_h.resize(nrec);
  for (int q = 0; q < nrec; ++q) ctx->rec_slot_h[q] = ctx->eq_slot[req[q]];
  if (int rc = upload(ctx, &ctx->d_rec_slot, ctx->rec_slot_h)) return rc;
  {
    // receivers grouped by equation, groups in ascending equation order (misfit)
    std::vector<int> ord(nrec), gptr(1, 0);
    for (int q = 0; q < nrec; ++q) ord[q] = q;
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return req[a] < req[b]; });
    for (int q = 0; q < nrec; ++q)
      if (q + 1 == nrec || req[ord[q + 1]] != req[ord[q]]) gptr.push_back(q + 1);
    ctx->mis_ngrp = (int)gptr.size() - 1;
    if (int rc = upload(ctx, &ctx->d_mis_ptr, gptr)) return rc;
    if (int rc = upload(ctx, &ctx->d_mis_idx, ord)) return rc;
  }
  const size_t rows = (size_t)nstep + 1;
  const size_t ntr = (size_t)nshot * rows * (nrec > 0 ? nrec : 1);
  if (ctx->traces_cap < ntr) {
    if (int rc = dalloc(ctx, &ctx->d_traces, ntr)) return rc;
    ctx->traces_cap = ntr;
  }
  // history storage, all of it allocated before the time loop
  const size_t slice = 3 * (size_t)ctx->nbox;
  if (scheme == 0) {
    if (int rc = plan_history(ctx, history, segment, nstep, nshot)) return rc;
    if (ctx->h_store) {
      const size_t need = (size_t)nshot * rows * slice;
      if (ctx->hist_cap < need) {
        if (int rc = dalloc(ctx, &ctx->d_hist, need)) return rc;
        ctx->hist_cap = need;
      }
    }
    if (int rc = ensure_history_buffers(ctx, nshot)) return rc;
  }
  if (int rc = zero_state(ctx)) return rc;
  if (scheme == 1) {
    ctx->e_rows = 0;
    // traces row n = u at the start of step n (n >= 1; row 0 is u_0 = 0), as U_Store(:, ISTEP)
    if (int rc = record_step(ctx, 0)) return rc;
    for (int n = 1; n <= nstep; ++n) {
      if (int rc = record_step(ctx, n)) return rc;
      if (int rc = cd_step(ctx, dt, n - 1)) return rc;
    }
    if (traces && nrec)
      HIPC(hipMemcpyAsync(traces, ctx->d_traces, rows * nrec * sizeof(double),
                          dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, ctx->stream));
    HIPC(hipStreamSynchronize(ctx->stream));
    ctx->f_nstep = 0;                 // no history for an adjoint
    return 0;
  }
  const double sgn[3] = {1.0, 1.0, 1.0};
  // checkpointed history: the forward takes the first chain of snapshots and
  // keeps the slices of the last segment
  const int L = ctx->h_seglen;
  std::vector<int> chain_slot;
  if (L > 0) {
    chain_slot.assign(ctx->h_nseg, -1);
    const std::vector<int> ch = ctx->h_plan.chain();
    for (size_t i = 0; i < ch.size(); ++i) chain_slot[ch[i]] = (int)i;
  }
  const int last0 = L > 0 ? (ctx->h_nseg - 1) * L : 0;
  const size_t seg_rows = (size_t)L + 1;
  // regular-domain energy of every shot at steps 0, e_every, 2 e_every, ... <= nstep
  ctx->e_rows = ctx->e_every > 0 ? nstep / ctx->e_every + 1 : 0;
  if (ctx->e_rows) {
    if (int rc = dalloc(ctx, &ctx->d_energy, (size_t)nshot * ctx->e_rows)) return rc;
    if (int rc = dalloc(ctx, &ctx->d_epart, (size_t)energy_partials_size())) return rc;
  }
  auto energy_at = [ctx, nshot](int n) -> int {
    const size_t nf = field_size(ctx);
    for (int q = 0; q < nshot; ++q)
      HIPC(launch_energy(ctx->g.NN, ctx->d_mrd, ctx->x[2] + q * nf, ctx->d_epart,
                         ctx->d_energy + (size_t)q * ctx->e_rows + n / ctx->e_every, ctx->stream));
    return 0;
  };
  auto step_body = [&](int n) -> int {
    // traces and history slice of the state at the start of step n, queued beside its stage 1
    ctx->deferred.push_back([ctx, n]() { return record_step(ctx, n); });
    if (ctx->e_rows && n % ctx->e_every == 0) ctx->deferred.push_back([energy_at, n]() { return energy_at(n); });
    if (ctx->h_store)
      ctx->deferred.push_back([ctx, rows, n]() { return store_slice(ctx, ctx->d_hist, rows, (size_t)n); });
    if (L > 0) {
      if (n % L == 0 && chain_slot[n / L] >= 0)
        if (int rc = snap_save(ctx, ctx->x, ctx->snap[chain_slot[n / L]], nshot)) return rc;
      if (n >= last0)
        ctx->deferred.push_back([ctx, seg_rows, n, last0]() {
          return store_slice(ctx, ctx->d_seg, seg_rows, (size_t)(n - last0));
        });
    }
    Src3Spec sp4[4];
    forward_sources(ctx, n, sp4);
    return rk4_step(ctx, dt, +1, 0, sgn, sp4);
  };
  if (L == 0 && ctx->e_rows == 0) {
    // per-step work depends on n only through the device step index: capturable
    if (int rc = run_steps(ctx, 0, nstep, +1, step_body)) return rc;
  } else {
    for (int n = 0; n < nstep; ++n)
      if (int rc = step_body(n)) return rc;
  }
  if (L > 0) if (int rc = store_slice(ctx, ctx->d_seg, seg_rows, (size_t)(nstep - last0))) return rc;
  if (int rc = record_step(ctx, nstep)) return rc;
  if (ctx->e_rows && nstep % ctx->e_every == 0) if (int rc = energy_at(nstep)) return rc;
  if (ctx->h_store) if (int rc = store_slice(ctx, ctx->d_hist, rows, (size_t)nstep)) return rc;
  if (traces && nrec) {
    HIPC(hipMemcpyAsync(traces, ctx->d_traces, (size_t)nshot * rows * nrec * sizeof(double),
                        dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, ctx->stream));
  }
  HIPC(hipStreamSynchronize(ctx->stream));
  return 0;
}

}  // namespace

extern "C" {

int hr_forward_shots(hr_ctx* ctx, int nshot, int nstep, double dt, const double* ftab, int nsrc, const int* src_shot,
                     const long long* src_eq, const double* src_val, int nrec, const long long* rec_eq, double* traces,
                     int history, int segment) {
  return forward_impl(ctx, nshot, nstep, dt, ftab, nsrc, src_shot, src_eq, src_val, nrec, rec_eq, traces, history,
                      segment, 0);
}

int hr_forward_cd(hr_ctx* ctx, int nstep, double dt, const double* ftab, int nsrc, const long long* src_eq,
                  const double* src_val, int nrec, const long long* rec_eq, double* traces) {
  if (!ctx) return 1;
  if (nstep < 1 || !(dt > 0.0)) return fail(ctx, 1, "nstep must be >= 1 and dt > 0");
  if (!ftab) return fail(ctx, 1, "null time-function table");
  // f(n dt), n = 1..nstep, into column 0 of step n - 1, scaled by dt^2 (B' = dt^2 M^-1 B)
  std::vector<double> f(nstep), t3(3 * (size_t)nstep, 0.0);
  if (ctx->flags & HR_DEVICE_PTRS) HIPC(hipMemcpy(f.data(), ftab, nstep * sizeof(double), hipMemcpyDeviceToHost));
  else std::memcpy(f.data(), ftab, nstep * sizeof(double));
  for (int n = 0; n < nstep; ++n) t3[3 * (size_t)n] = (dt * dt) * f[n];
  return forward_impl(ctx, 1, nstep, dt, t3.data(), nsrc, nullptr, src_eq, src_val, nrec, rec_eq, traces,
                      HR_HISTORY_NONE, 1, 1);
}

int hr_forward(hr_ctx* ctx, int nstep, double dt, const double* ftab, int nsrc, const long long* src_eq,
               const double* src_val, int nrec, const long long* rec_eq, double* traces, int history, int segment) {
  return hr_forward_shots(ctx, 1, nstep, dt, ftab, nsrc, nullptr, src_eq, src_val, nrec, rec_eq, traces, history,
                          segment);
}

}  // extern "C"

namespace hr {
size_t misfit_scratch_size(const hr_ctx* ctx) { return (size_t)ctx->f_nshot * ((size_t)ctx->f_nstep + 1) + ctx->f_nshot; }

int misfit_check(hr_ctx* ctx, int nstep, int nrec, const void* dis_meas, const void* misfit) {
  if (int rc = check_ready(ctx)) return rc;
  if (nstep != ctx->f_nstep || nrec != ctx->f_nrec) return fail(ctx, 1, "nstep/nrec differ from the last forward");
  if (!dis_meas || !misfit) return fail(ctx, 1, "null argument");
  return 0;
}

// Compute_Misfit of every shot of the last forward against observations on the
// device, summed over shots in shot order into *out (device); part holds
// misfit_scratch_size(ctx) doubles, all written before they are read
int misfit_launch(hr_ctx* ctx, double dt, const double* dobs, double* part, double* out) {
  const int nshot = ctx->f_nshot, nstep = ctx->f_nstep, nrec = ctx->f_nrec;
  const size_t rows = (size_t)nstep + 1;
  double* per_shot = part + (size_t)nshot * rows;
  for (int q = 0; q < nshot; ++q)
    HIPC(launch_misfit(nstep, ctx->mis_ngrp, ctx->d_mis_ptr, ctx->d_mis_idx, nrec,
                       ctx->d_traces + (size_t)q * rows * nrec, dobs + (size_t)q * rows * nrec, dt,
                       part + (size_t)q * rows, per_shot + q, ctx->stream));
  HIPC(launch_sum(nshot, per_shot, out, ctx->stream));
  return 0;
}

bool ctx_device_ptrs(const hr_ctx* ctx) { return ctx->flags & HR_DEVICE_PTRS; }
hipStream_t ctx_stream(const hr_ctx* ctx) { return ctx->stream; }
int ctx_fail(hr_ctx* ctx, int code, const char* msg) { return fail(ctx, code, "%s", msg); }
}  // namespace hr

namespace {
int misfit_all(hr_ctx* ctx, double dt, const double* dobs, double* out) {
  if (int rc = dalloc(ctx, &ctx->d_part, hr::misfit_scratch_size(ctx))) return rc;
  return hr::misfit_launch(ctx, dt, dobs, ctx->d_part, out);
}
}  // namespace

extern "C" {

int hr_misfit(hr_ctx* ctx, int nstep, double dt, int nrec, const double* dis_meas, double* misfit) {
  if (int rc = hr::misfit_check(ctx, nstep, nrec, dis_meas, misfit)) return rc;
  const bool dev = ctx->flags & HR_DEVICE_PTRS;
  const size_t nobs = (size_t)ctx->f_nshot * (nstep + 1) * nrec;
  const double* dobs = dis_meas;
  if (!dev) {
    if (ctx->obs_cap < nobs) {
      if (int rc = dalloc(ctx, &ctx->d_obs, nobs)) return rc;
      ctx->obs_cap = nobs;
    }
    HIPC(hipMemcpyAsync(ctx->d_obs, dis_meas, nobs * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    dobs = ctx->d_obs;
  }
  if (!ctx->d_scalar) if (int rc = dalloc(ctx, &ctx->d_scalar, 1)) return rc;
  if (int rc = misfit_all(ctx, dt, dobs, ctx->d_scalar)) return rc;
  HIPC(hipMemcpyAsync(misfit, ctx->d_scalar, sizeof(double), dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                      ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  return 0;
}

int hr_adjoint_gradient(hr_ctx* ctx, int nstep, double dt, int nrec, const long long* rec_eq, const double* dis_meas,
                        double* g_lambda, double* g_mu, double* misfit) {
  if (int rc = check_ready(ctx)) return rc;
  ctx->deferred.clear();   // work left by a solve that stopped on an error
  if (nstep != ctx->f_nstep || nrec != ctx->f_nrec || dt != ctx->f_dt)
    return fail(ctx, 1, "nstep/dt/nrec differ from the last hr_forward");
  if (ctx->f_history == HR_HISTORY_NONE) return fail(ctx, 2, "the last hr_forward kept no history");
  if (!dis_meas || !g_lambda || !g_mu) return fail(ctx, 1, "null argument");
  if (nrec > 0 && !rec_eq) return fail(ctx, 1, "null receiver array");
  const bool dev = ctx->flags & HR_DEVICE_PTRS;
  const int nshot = ctx->f_nshot;
  std::vector<long long> req(nrec);
  if (nrec) {
    if (dev) HIPC(hipMemcpy(req.data(), rec_eq, nrec * sizeof(long long), hipMemcpyDeviceToHost));
    else std::memcpy(req.data(), rec_eq, nrec * sizeof(long long));
  }
  for (int q = 0; q < nrec; ++q)
    if (req[q] < 0 || req[q] >= ctx->neq || ctx->eq_slot[req[q]] != ctx->rec_slot_h[q])
      return fail(ctx, 1, "receivers differ from the last hr_forward");
  const size_t rows = (size_t)nstep + 1;
  const size_t nobs = (size_t)nshot * rows * nrec;
  if (ctx->obs_cap < nobs) {
    if (int rc = dalloc(ctx, &ctx->d_obs, nobs)) return rc;
    ctx->obs_cap = nobs;
  }
  HIPC(hipMemcpyAsync(ctx->d_obs, dis_meas, nobs * sizeof(double), dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                      ctx->stream));
  // adjoint source entries: the shared receivers, grouped by tile, repeated per shot;
  // traces and observations permuted into entry order (device gather)
  SrcBuild sb = build_src(ctx, nrec, req.data());
  std::vector<int> rptr(1, 0);
  std::vector<long long> rnode;
  {
    const int nt = ntiles(ctx->g);
    std::vector<int>& ptr = rptr;
    std::vector<long long>& node = rnode;
    std::vector<int> comp;
    std::vector<double> m;
    for (int q = 0; q < nshot; ++q) {
      for (int t = 0; t < nt; ++t) ptr.push_back(q * nrec + sb.ptr[t + 1]);
      node.insert(node.end(), sb.node.begin(), sb.node.end());
      comp.insert(comp.end(), sb.comp.begin(), sb.comp.end());
      m.insert(m.end(), sb.m.begin(), sb.m.end());
    }
    if (int rc = upload(ctx, &ctx->d_rec_tile_ptr, ptr)) return rc;
    if (int rc = upload(ctx, &ctx->d_rec_node, node)) return rc;
    if (int rc = upload(ctx, &ctx->d_rec_comp, comp)) return rc;
    if (int rc = upload(ctx, &ctx->d_rec_m, m)) return rc;
  }
  if (int rc = upload(ctx, &ctx->d_rec_perm, sb.perm)) return rc;
  if (ctx->sorted_cap < nobs) {
    if (int rc = dalloc(ctx, &ctx->d_tr_sorted, nobs)) return rc;
    if (int rc = dalloc(ctx, &ctx->d_obs_sorted, nobs)) return rc;
    ctx->sorted_cap = nobs;
  }
  {
    Src3 h3;
    std::memset(&h3, 0, sizeof(h3));
    h3.tile_ptr = ctx->d_rec_tile_ptr; h3.node = ctx->d_rec_node; h3.comp = ctx->d_rec_comp;
    h3.m = ctx->d_rec_m; h3.tr = ctx->d_tr_sorted; h3.obs = ctx->d_obs_sorted; h3.nrec = nrec;
    h3.nstep = nstep; h3.ntiles = ntiles(ctx->g);
    if (ctx->split_ok)
      if (int rc = build_pull_blocks(ctx, rptr, rnode, nshot, h3.ntiles, &ctx->d_prec_ptr, &ctx->d_prec_idx, h3))
        return rc;
    ctx->src3_adj = h3;
  }
  if (nrec) {
    HIPC(launch_permute_cols((long long)nshot * rows, nrec, ctx->d_rec_perm, ctx->d_traces, ctx->d_tr_sorted,
                             ctx->stream));
    HIPC(launch_permute_cols((long long)nshot * rows, nrec, ctx->d_rec_perm, ctx->d_obs, ctx->d_obs_sorted,
                             ctx->stream));
  }
  // gradient accumulators on the RD box (sum over shots)
  if (int rc = dalloc(ctx, &ctx->d_gl, ctx->nbox)) return rc;
  if (int rc = dalloc(ctx, &ctx->d_gm, ctx->nbox)) return rc;
  HIPC(hipMemsetAsync(ctx->d_gl, 0, ctx->nbox * sizeof(double), ctx->stream));
  HIPC(hipMemsetAsync(ctx->d_gm, 0, ctx->nbox * sizeof(double), ctx->stream));
  const size_t slice = 3 * (size_t)ctx->nbox;
  const double sgn[3] = {-1.0, 1.0, -1.0};
  const size_t nf = field_size(ctx);
  const size_t nfs = nf * nshot;
  const int L = ctx->h_seglen;
  const size_t seg_rows = (size_t)L + 1;
  // adjoint state starts from rest
  if (int rc = zero_state(ctx)) return rc;
  // the recomputed forward state (F, FY) lives in the state buffers; the adjoint
  // state (P, PY) in buffers of its own while the forward is recomputed
  double *F[3], *FY[3], *P[3], *PY[3];
  for (int a = 0; a < 3; ++a) { F[a] = ctx->x[a]; FY[a] = ctx->y[a]; }
  if (ctx->h_store) {
    for (int a = 0; a < 3; ++a) { P[a] = F[a]; PY[a] = FY[a]; }
  } else {
    for (int a = 0; a < 6; ++a) {
      if (int rc = dalloc(ctx, &ctx->adj_buf[a], nfs)) return rc;
      HIPC(hipMemsetAsync(ctx->adj_buf[a], 0, nfs * sizeof(double), ctx->stream));
    }
    for (int a = 0; a < 3; ++a) { P[a] = ctx->adj_buf[a]; PY[a] = ctx->adj_buf[3 + a]; }
  }
  auto use = [ctx](double* const* x, double* const* y) {
    for (int a = 0; a < 3; ++a) { ctx->x[a] = x[a]; ctx->y[a] = y[a]; }
  };
  auto keep = [ctx](double** x, double** y) {
    for (int a = 0; a < 3; ++a) { x[a] = ctx->x[a]; y[a] = ctx->y[a]; }
  };
  // adjoint step n, with the gradient of step n from history slice uh_row of uh
  auto adjoint_step = [&](int n, const double* uh, size_t uh_rows, size_t uh_row) -> int {
    if (n <= nstep - 1) {
      // queued beside stage 1 of the adjoint step, which reads but does not write p
      double* const a1 = P[1];
      const StepRef sr{ctx->dyn ? ctx->d_step : nullptr, (long long)slice};
      ctx->deferred.push_back([ctx, nshot, uh, uh_rows, uh_row, slice, a1, nf, dt, sr]() -> int {
        for (int q = 0; q < nshot; ++q)
          HIPC(launch_gradient(ctx->g, ctx->box0, ctx->boxn, uh + ((size_t)q * uh_rows + uh_row) * slice,
                               a1 + q * nf, dt, ctx->d_gl, ctx->d_gm, ctx->stream, sr));
        return 0;
      });
    }
    // misfit sources at t_n (stage 1), t_{n-1} (stage 4), mean (stages 2, 3)
    Src3Spec sp4[4];
    for (int st = 0; st < 4; ++st) {
      sp4[st] = Src3Spec();
      if (!nrec) continue;
      sp4[st].lists = &ctx->src3_adj;
      sp4[st].mode = 2;
      if (st == 0) { sp4[st].ra = n; sp4[st].ca = 1.0; }
      else if (st == 3) { sp4[st].ra = n - 1; sp4[st].ca = 1.0; }
      else { sp4[st].ra = n; sp4[st].ca = 0.5; sp4[st].rb = n - 1; sp4[st].cb = 0.5; sp4[st].rb_on = true; }
    }
    use(P, PY);
    const int rc = rk4_step(ctx, dt, -1, 1, sgn, sp4);
    keep(P, PY);
    return rc;
  };
  long long recomputed = 0;
  int status = 0;
  if (ctx->h_store) {
    // step nstep has no gradient term; steps nstep-1 .. 1 are alike (capturable):
    // in a captured step n is the offset 0 from the device step index
    status = adjoint_step(nstep, ctx->d_hist, rows, (size_t)nstep);
    if (!status && nstep > 1)
      status = run_steps(ctx, nstep - 1, nstep - 1, -1, [&](int n) -> int {
        return adjoint_step(ctx->dyn ? 0 : n, ctx->d_hist, rows, (size_t)(ctx->dyn ? 0 : n));
      });
  } else {
    // the reverse sweep of the binomial schedule; the forward kept the slices of the last segment
    const std::vector<RevAction> acts = ctx->h_plan.actions(true);
    const int m = ctx->h_nseg;
    auto forward = [&](int n0, int n1, double* hist) -> int {
      use(F, FY);
      const int rc = run_forward_steps(ctx, n0, n1, hist, seg_rows);
      keep(F, FY);
      recomputed += n1 - n0;
      return rc;
    };
    for (const RevAction& act : acts) {
      if (status) break;
      switch (act.op) {
        case REV_ZERO:
          for (int a = 0; a < 3; ++a)
            if (hipMemsetAsync(F[a], 0, nfs * sizeof(double), ctx->stream) != hipSuccess) status = fail(ctx, 3, "memset failed");
          break;
        case REV_RESTORE: status = snap_load(ctx, F, ctx->snap[act.a], nshot); break;
        case REV_SNAP: status = snap_save(ctx, F, ctx->snap[act.a], nshot); break;
        case REV_ADVANCE: status = forward(act.a * L, act.b * L, nullptr); break;
        case REV_RECORD: status = forward(act.a * L, std::min((act.a + 1) * L, nstep), ctx->d_seg); break;
        case REV_CONSUME: {
          const int lo = std::max(act.a * L, 1), hi = act.a == m - 1 ? nstep : (act.a + 1) * L - 1;
          for (int n = hi; n >= lo && !status; --n)
            status = adjoint_step(n, ctx->d_seg, seg_rows, (size_t)(n - act.a * L));
          break;
        }
        default: status = fail(ctx, 3, "bad schedule action");
      }
    }
  }
  // the state buffers go back to the forward triple (the adjoint ones are context-owned)
  use(F, FY);
  if (status) {
    ctx->deferred.clear();
    return status;
  }
  ctx->recompute = 1.0 + (double)recomputed / (double)nstep;
  // nodal outputs: zero, then -g on the box
  double *dgl = g_lambda, *dgm = g_mu;
  const size_t NN = (size_t)ctx->g.NN;
  if (!dev) {
    dgl = ctx->tmp;
    dgm = ctx->tmp2;
  }
  HIPC(hipMemsetAsync(dgl, 0, NN * sizeof(double), ctx->stream));
  HIPC(hipMemsetAsync(dgm, 0, NN * sizeof(double), ctx->stream));
  HIPC(launch_gradient_finish(ctx->g, ctx->box0, ctx->boxn, ctx->d_gl, ctx->d_gm, dgl, dgm, ctx->stream));
  if (!dev) {
    HIPC(hipMemcpyAsync(g_lambda, dgl, NN * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPC(hipMemcpyAsync(g_mu, dgm, NN * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPC(hipStreamSynchronize(ctx->stream));
    HIPC(hipMemsetAsync(ctx->tmp, 0, field_size(ctx) * sizeof(double), ctx->stream));
    HIPC(hipMemsetAsync(ctx->tmp2, 0, field_size(ctx) * sizeof(double), ctx->stream));
  }
  if (misfit) {
    if (!ctx->d_scalar) if (int rc = dalloc(ctx, &ctx->d_scalar, 1)) return rc;
    if (int rc = misfit_all(ctx, dt, ctx->d_obs, ctx->d_scalar)) return rc;
    HIPC(hipMemcpyAsync(misfit, ctx->d_scalar, sizeof(double), dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                        ctx->stream));
  }
  HIPC(hipStreamSynchronize(ctx->stream));
  return 0;
}

double hr_recompute_ratio(const hr_ctx* ctx) { return ctx ? ctx->recompute : 0.0; }

int hr_set_history_budget(hr_ctx* ctx, double bytes) {
  if (!ctx) return 1;
  if (!(bytes >= 0.0)) return fail(ctx, 1, "history budget must be >= 0");
  ctx->h_budget = bytes;
  return 0;
}

int hr_get_history_plan(hr_ctx* ctx, int nstep, int nshot, int history, int segment, int* seglen, int* nsnap,
                        double* bytes, double* ratio) {
  if (!ctx) return 1;
  if (nstep < 1 || nshot < 1 || history < 0 || history > 3) return fail(ctx, 1, "bad history arguments");
  if (history == HR_HISTORY_CHECKPOINT && segment < 1) return fail(ctx, 1, "checkpoint segment must be >= 1");
  // plan into a scratch copy of the fields the planner sets
  const bool st0 = ctx->h_store;
  const int l0 = ctx->h_seglen, m0 = ctx->h_nseg, s0 = ctx->h_nsnap;
  RevolvePlan p0 = ctx->h_plan;
  int rc = plan_history(ctx, history, segment, nstep, nshot);
  if (rc == 0) {
    const double sl = (double)slice_bytes(ctx, nshot), st = (double)state_bytes(ctx, nshot);
    double b = 0.0, r = 1.0;
    int L = 0, ns = 0;
    if (ctx->h_store) {
      b = (nstep + 1) * sl;
    } else if (ctx->h_seglen > 0) {
      L = ctx->h_seglen;
      ns = std::max(ctx->h_plan.s, ctx->h_nsnap);
      b = ns * st + (L + 1) * sl;
      // recomputation: every segment but the last once, plus the schedule's advances
      const double rec = (double)std::max(0, nstep - (ctx->h_nseg - 1) * L);
      r = 1.0 + ((double)nstep - rec + (double)L * (double)ctx->h_plan.cost()) / (double)nstep;
    }
    if (seglen) *seglen = L;
    if (nsnap) *nsnap = ns;
    if (bytes) *bytes = b;
    if (ratio) *ratio = r;
  }
  ctx->h_store = st0;
  ctx->h_seglen = l0;
  ctx->h_nseg = m0;
  ctx->h_nsnap = s0;
  ctx->h_plan = p0;
  return rc;
}

int hr_autotune(hr_ctx* ctx, int nshot, int steps, int* tz_best, double* ms_per_step) {
  if (int rc = check_ready(ctx)) return rc;
  if (nshot < 1 || steps < 1) return fail(ctx, 1, "autotune needs nshot >= 1 and steps >= 1");
  if (int rc = ensure_shots(ctx, nshot)) return rc;
  ctx->deferred.clear();
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIPC(hipEventCreate(&e0));
  HIPC(hipEventCreate(&e1));
  const int tz0 = ctx->g.tz;
  const double sgn[3] = {1.0, 1.0, 1.0};
  double best = -1.0;
  int bt = tz0;
  int status = 0;
  // z-chunk candidates: element layers per chunk of the stage-kernel tiles
  for (int tz : {3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 14}) {
    if (tz > ctx->g.nez) continue;
    ctx->g.tz = tz;
    if (setup_tiles(ctx) != 0) continue;      // too many chunks for this mesh
    // timed RK4 steps on the zero state without sources: the same launches and traffic
    if ((status = zero_state(ctx))) break;
    if ((status = rk4_step(ctx, 1.0e-4, +1, 0, sgn, nullptr))) break;
    if (hipEventRecord(e0, ctx->stream) != hipSuccess) { status = fail(ctx, 3, "hipEventRecord failed"); break; }
    for (int k = 0; k < steps && !status; ++k) status = rk4_step(ctx, 1.0e-4, +1, 0, sgn, nullptr);
    if (status) break;
    float ms = 0.0f;
    if (hipEventRecord(e1, ctx->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
      status = fail(ctx, 3, "autotune timing failed");
      break;
    }
    const double per = (double)ms / steps;
    if (best < 0.0 || per < best) { best = per; bt = tz; }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  ctx->g.tz = status ? tz0 : bt;
  if (int rc = setup_tiles(ctx)) return rc;
  if (int rc = zero_state(ctx)) return rc;
  if (status) return status;
  if (tz_best) *tz_best = bt;
  if (ms_per_step) *ms_per_step = best;
  return 0;
}

int hr_set_energy_interval(hr_ctx* ctx, int every) {
  if (!ctx) return 1;
  if (every < 0) return fail(ctx, 1, "energy interval must be >= 0");
  ctx->e_every = every;
  return 0;
}

int hr_get_energy(hr_ctx* ctx, double* energy, int* nrow) {
  if (!ctx || !nrow) return fail(ctx, 1, "null argument");
  *nrow = ctx->e_rows;
  if (!energy || ctx->e_rows == 0) return 0;
  const size_t n = (size_t)ctx->f_nshot * ctx->e_rows;
  HIPC(hipMemcpyAsync(energy, ctx->d_energy, n * sizeof(double),
                      (ctx->flags & HR_DEVICE_PTRS) ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  return 0;
}

int hr_gradient_accumulate(hr_ctx* ctx, const double* u, const double* p, double scale, double* g_lambda,
                           double* g_mu) {
  if (int rc = check_ready(ctx)) return rc;
  if (!u || !p || !g_lambda || !g_mu) return fail(ctx, 1, "null argument");
  const bool dev = ctx->flags & HR_DEVICE_PTRS;
  const long long neq = ctx->neq;
  const size_t NN = (size_t)ctx->g.NN;
  const double *du = u, *dp = p;
  if (!dev) {
    HIPC(hipMemcpyAsync(ctx->d_ax, u, neq * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HIPC(hipMemcpyAsync(ctx->d_ay, p, neq * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    du = ctx->d_ax;
    dp = ctx->d_ay;
  }
  // u and p on the lattice (the scratch fields hold zeros at inactive slots), u's
  // regular-domain displacement as one history slice
  HIPC(launch_gather_eq(neq, ctx->d_eq_slot, du, nullptr, ctx->tmp, ctx->stream));
  HIPC(launch_gather_eq(neq, ctx->d_eq_slot, dp, nullptr, ctx->tmp2, ctx->stream));
  if (int rc = dalloc(ctx, &ctx->d_uslice, 3 * (size_t)ctx->nbox)) return rc;
  if (int rc = dalloc(ctx, &ctx->d_gl, ctx->nbox)) return rc;
  if (int rc = dalloc(ctx, &ctx->d_gm, ctx->nbox)) return rc;
  HIPC(launch_store_box(ctx->g, ctx->box0, ctx->boxn, ctx->tmp, ctx->d_uslice, ctx->stream));
  HIPC(hipMemsetAsync(ctx->d_gl, 0, ctx->nbox * sizeof(double), ctx->stream));
  HIPC(hipMemsetAsync(ctx->d_gm, 0, ctx->nbox * sizeof(double), ctx->stream));
  HIPC(launch_gradient(ctx->g, ctx->box0, ctx->boxn, ctx->d_uslice, ctx->tmp2, scale, ctx->d_gl, ctx->d_gm,
                       ctx->stream));
  double *gl = g_lambda, *gm = g_mu;
  if (!dev) {
    if (int rc = dalloc(ctx, &ctx->d_pg, 2 * NN)) return rc;
    gl = ctx->d_pg;
    gm = ctx->d_pg + NN;
    HIPC(hipMemcpyAsync(gl, g_lambda, NN * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HIPC(hipMemcpyAsync(gm, g_mu, NN * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  }
  HIPC(launch_box_add(ctx->g, ctx->box0, ctx->boxn, ctx->d_gl, ctx->d_gm, gl, gm, ctx->stream));
  if (!dev) {
    HIPC(hipMemcpyAsync(g_lambda, gl, NN * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPC(hipMemcpyAsync(g_mu, gm, NN * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  }
  HIPC(hipStreamSynchronize(ctx->stream));
  return 0;
}

int hr_set_numbering(hr_ctx* ctx, const long long* id) {
  if (!ctx || !id) return fail(ctx, 1, "null argument");
  const Geo& g = ctx->g;
  const size_t n9 = (size_t)NCOMP * (size_t)g.NN;
  std::vector<long long> h(n9);
  if (ctx->flags & HR_DEVICE_PTRS) HIPC(hipMemcpy(h.data(), id, n9 * sizeof(long long), hipMemcpyDeviceToHost));
  else std::memcpy(h.data(), id, n9 * sizeof(long long));
  std::vector<long long> slot((size_t)ctx->neq, -1);
  for (int k = 0; k < g.NZ; ++k)
    for (int j = 0; j < g.NY; ++j)
      for (int i = 0; i < g.NX; ++i) {
        const long long n = lat(g, i, j, k);
        for (int c = 0; c < NCOMP; ++c) {
          const long long e = h[(size_t)NCOMP * n + c];
          const bool act = h_active(g, i, j, k, c);
          if (act != (e >= 0))
            return fail(ctx, 1, "numbering: component %d of node %lld is %s here", c, n, act ? "active" : "constrained");
          if (!act) continue;
          if (e >= ctx->neq || slot[e] >= 0) return fail(ctx, 1, "numbering: equation %lld out of range or repeated", e);
          slot[e] = slot_of(g, c, n);
        }
      }
  ctx->eq_slot.swap(slot);
  if (int rc = upload(ctx, &ctx->d_eq_slot, ctx->eq_slot)) return rc;
  // MatShell scalings and the last forward were expressed in the old numbering
  for (int w = 0; w < 3; ++w) {
    ctx->shell[w].s = 1.0;
    dfree(ctx, &ctx->shell[w].l);
    dfree(ctx, &ctx->shell[w].r);
  }
  ctx->f_nstep = 0;
  ctx->f_history = HR_HISTORY_NONE;
  return 0;
}

int hr_surface_force(hr_ctx* ctx, const int* nface, long long* eq, double* val, long long* nnz) {
  if (!ctx || !nface || !nnz) return fail(ctx, 1, "null argument");
  const Geo& g = ctx->g;
  // LFACE, FVAL, IPERM of ELEMENT_FS_3D: the face is xi_N = FVAL with in-face axes L, M
  const int LFACE[6] = {2, 2, 3, 3, 1, 1};
  const double FVAL[6] = {+1.0, -1.0, +1.0, -1.0, +1.0, -1.0};
  const int IPERM[3] = {2, 3, 1};
  const double w[3] = {1.0 / 3.0, 4.0 / 3.0, 1.0 / 3.0};
  const double* h = ctx->hel;
  std::vector<double> B((size_t)ctx->neq, 0.0);
  std::vector<long long> inv(field_size(ctx), -1);
  for (long long e = 0; e < ctx->neq; ++e) inv[ctx->eq_slot[e]] = e;
  for (int ez = 0; ez < g.nez; ++ez)
    for (int ey = 0; ey < g.ney; ++ey)
      for (int ex = 0; ex < g.nex; ++ex) {
        const bool rd = ex >= g.rlo[0] && ex < g.rhi[0] && ey >= g.rlo[1] && ey < g.rhi[1] && ez >= g.rlo[2];
        int f = nface[(size_t)ex + (size_t)g.nex * ((size_t)ey + (size_t)g.ney * ez)];
        if (!rd || f == 0) continue;                  // MTYPE 1 only; NFACE 0: no traction
        if (f < 0 || f > 6) return fail(ctx, 1, "face code %d of element (%d, %d, %d) not in 0..6", f, ex, ey, ez);
        if (f == 6) f = 5;                            // as ELEMENT_FS_3D maps it
        const int L = LFACE[f - 1], M = IPERM[L - 1], N = IPERM[M - 1];
        // affine element: |d x / d xi_L  x  d x / d xi_M| = h_L h_M / 4
        const double detj_surf = (0.5 * h[L - 1]) * (0.5 * h[M - 1]);
        for (int lj = 0; lj < 3; ++lj)
          for (int li = 0; li < 3; ++li) {
            int loc[3];
            loc[N - 1] = FVAL[f - 1] > 0.0 ? 2 : 0;
            loc[L - 1] = li;
            loc[M - 1] = lj;
            const long long n = lat(g, 2 * ex + loc[0], 2 * ey + loc[1], 2 * ez + loc[2]);
            const long long q = inv[slot_of(g, 2, n)];  // unit traction in z
            if (q >= 0) B[q] += (1.0 * 1.0) * ((w[li] * w[lj]) * detj_surf);
          }
      }
  long long k = 0;
  for (long long q = 0; q < ctx->neq; ++q)
    if (B[q] != 0.0) {
      if (eq && val) { eq[k] = q; val[k] = B[q]; }
      ++k;
    }
  *nnz = k;
  return 0;
}

int hr_revolve_schedule(int nseg, int nsnap, int max_actions, int* actions, int* nactions, long long* advances) {
  if (nseg < 1 || nsnap < 0 || !nactions) return 1;
  RevolvePlan P;
  P.build(nseg, nsnap);
  const std::vector<RevAction> a = P.actions(true);
  const std::vector<int> ch = P.chain();
  *nactions = (int)(a.size() + ch.size());
  if (advances) *advances = P.cost();
  if (actions) {
    if (max_actions < *nactions) return 2;
    int k = 0;
    // the forward's snapshots first (REV_SNAP, slot, segment), then the reverse sweep
    for (size_t i = 0; i < ch.size(); ++i, ++k) {
      actions[3 * k] = REV_SNAP; actions[3 * k + 1] = (int)i; actions[3 * k + 2] = ch[i];
    }
    for (const RevAction& x : a) {
      actions[3 * k] = x.op; actions[3 * k + 1] = x.a; actions[3 * k + 2] = x.b;
      ++k;
    }
  }
  return 0;
}

}  // extern "C"

namespace {

int ensure_gradient_mass(hr_ctx* ctx) {
  if (ctx->d_img) return 0;
  const size_t NN = (size_t)ctx->g.NN;
  if (int rc = dalloc(ctx, &ctx->d_restr, NN)) return rc;
  if (int rc = dalloc(ctx, &ctx->d_img, NN)) return rc;
  HIPC(launch_gradient_mass(ctx->g, ctx->d_img, ctx->d_restr, ctx->stream));
  return 0;
}

int check_wreg(hr_ctx* ctx, const double* wreg) {
  if (wreg && !(wreg[2] != 0.0)) return fail(ctx, 1, "w_Length must be nonzero");
  return 0;
}

// Regularisation vectors of lam, mu (device) into ctx->d_reg (lambda, then mu)
// and the two unscaled costs into ctx->d_red[RED_OUT + {0, 1}].
constexpr int RED_OUT_COST = 4;
int regularize(hr_ctx* ctx, int method, const double* wreg, const double* dl, const double* dm) {
  const size_t NN = (size_t)ctx->g.NN;
  const long long nbe = reg_box_elements(ctx->g);
  if (!ctx->d_reg) if (int rc = dalloc(ctx, &ctx->d_reg, 2 * NN)) return rc;
  if (!ctx->d_cost_el) if (int rc = dalloc(ctx, &ctx->d_cost_el, 2 * (size_t)nbe)) return rc;
  if (!ctx->d_red) if (int rc = dalloc(ctx, &ctx->d_red, (size_t)dot_partials_size() + 8)) return rc;
  HIPC(launch_regularization(ctx->g, ctx->d_zc, method == 0, wreg, dl, dm, ctx->d_reg, ctx->d_reg + NN,
                             ctx->d_cost_el, ctx->stream));
  const double* xs[2] = {ctx->d_cost_el, ctx->d_cost_el + 1};
  const double* ys[2] = {nullptr, nullptr};
  HIPC(launch_dots(2, nbe, 2, xs, ys, ctx->d_red, ctx->d_red + dot_partials_size() + RED_OUT_COST, ctx->stream));
  return 0;
}

// device pointers to lam and mu: the caller's, or staged copies of host arrays
int stage_material(hr_ctx* ctx, const double* lam, const double* mu, const double** dl, const double** dm) {
  const size_t NN = (size_t)ctx->g.NN;
  if (ctx->flags & HR_DEVICE_PTRS) {
    *dl = lam;
    *dm = mu;
    return 0;
  }
  if (!ctx->d_pm) if (int rc = dalloc(ctx, &ctx->d_pm, 2 * NN)) return rc;
  if (int rc = to_dev(ctx, ctx->d_pm, lam, NN)) return rc;
  if (int rc = to_dev(ctx, ctx->d_pm + NN, mu, NN)) return rc;
  *dl = ctx->d_pm;
  *dm = ctx->d_pm + NN;
  return 0;
}

int copy_scalars(hr_ctx* ctx, double* dst, const double* src, int n) {
  HIPC(hipMemcpyAsync(dst, src, n * sizeof(double),
                      (ctx->flags & HR_DEVICE_PTRS) ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, ctx->stream));
  return 0;
}

}  // namespace

extern "C" {

int hr_gradient_mass(hr_ctx* ctx, double* imgrad, double* restriction) {
  if (!ctx || !imgrad) return fail(ctx, 1, "null argument");
  if (int rc = ensure_gradient_mass(ctx)) return rc;
  const size_t NN = (size_t)ctx->g.NN;
  if (restriction) if (int rc = from_dev(ctx, restriction, ctx->d_restr, NN)) return rc;
  if (int rc = from_dev(ctx, imgrad, ctx->d_img, NN)) return rc;
  HIPC(hipStreamSynchronize(ctx->stream));
  return 0;
}

int hr_regularization(hr_ctx* ctx, int method, const double* wreg, const double* lam, const double* mu,
                      double* reg_lam, double* reg_mu, double* cost) {
  if (!ctx || !lam || !mu || !reg_lam || !reg_mu) return fail(ctx, 1, "null argument");
  if (method != HR_REG_TV && method != HR_REG_TN) return fail(ctx, 1, "method must be HR_REG_TV or HR_REG_TN");
  if (int rc = check_wreg(ctx, wreg)) return rc;
  const double *dl, *dm;
  if (int rc = stage_material(ctx, lam, mu, &dl, &dm)) return rc;
  if (int rc = regularize(ctx, method, wreg, dl, dm)) return rc;
  const size_t NN = (size_t)ctx->g.NN;
  if (cost) if (int rc = copy_scalars(ctx, cost, ctx->d_red + dot_partials_size() + RED_OUT_COST, 2)) return rc;
  if (int rc = from_dev(ctx, reg_lam, ctx->d_reg, NN)) return rc;
  if (int rc = from_dev(ctx, reg_mu, ctx->d_reg + NN, NN)) return rc;
  HIPC(hipStreamSynchronize(ctx->stream));
  return 0;
}

int hr_regularize_gradient(hr_ctx* ctx, int iter, int method, double factor, const double* wreg, const double* lam,
                           const double* mu, double* g_lambda, double* g_mu, double* cost_reg, double* reg_factor) {
  if (!ctx || !lam || !mu || !g_lambda || !g_mu) return fail(ctx, 1, "null argument");
  if (method != HR_REG_TV && method != HR_REG_TN) return fail(ctx, 1, "method must be HR_REG_TV or HR_REG_TN");
  if (int rc = check_wreg(ctx, wreg)) return rc;
  const bool dev = ctx->flags & HR_DEVICE_PTRS;
  const size_t NN = (size_t)ctx->g.NN;
  if (int rc = ensure_gradient_mass(ctx)) return rc;
  const double *dl, *dm;
  if (int rc = stage_material(ctx, lam, mu, &dl, &dm)) return rc;
  double *gl = g_lambda, *gm = g_mu;
  if (!dev) {
    if (!ctx->d_pg) if (int rc = dalloc(ctx, &ctx->d_pg, 2 * NN)) return rc;
    gl = ctx->d_pg;
    gm = ctx->d_pg + NN;
    if (int rc = to_dev(ctx, gl, g_lambda, NN)) return rc;
    if (int rc = to_dev(ctx, gm, g_mu, NN)) return rc;
  }
  if (int rc = regularize(ctx, method, wreg, dl, dm)) return rc;
  // ||g||^2 and ||reg||^2 of both fields (VecNorm, Program_Main.F90:433-441)
  const double* rl = ctx->d_reg;
  const double* rm = ctx->d_reg + NN;
  const double* xs[4] = {gl, rl, gm, rm};
  double* out = ctx->d_red + dot_partials_size();
  HIPC(launch_dots(4, (long long)NN, 1, xs, xs, ctx->d_red, out, ctx->stream));
  double h[6];
  HIPC(hipMemcpyAsync(h, out, 6 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  // adaptive factor; a zero regularisation vector (homogeneous field) gets
  // factor 0 where the reference divides by zero
  const double fl = h[1] > 0.0 ? factor * std::sqrt(h[0]) / std::sqrt(h[1]) : 0.0;
  const double fm = h[3] > 0.0 ? factor * std::sqrt(h[2]) / std::sqrt(h[3]) : 0.0;
  const int add = iter != 1;
  HIPC(launch_reg_update((long long)NN, add, fl, fm, rl, rm, ctx->d_img, gl, gm, ctx->stream));
  if (!dev) {
    HIPC(hipMemcpyAsync(g_lambda, gl, NN * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPC(hipMemcpyAsync(g_mu, gm, NN * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  }
  const double c[2] = {add ? h[4] * fl : 0.0, add ? h[5] * fm : 0.0};
  const double f2[2] = {fl, fm};
  if (dev) {
    if (cost_reg) HIPC(hipMemcpyAsync(cost_reg, c, 2 * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    if (reg_factor) HIPC(hipMemcpyAsync(reg_factor, f2, 2 * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  } else {
    if (cost_reg) { cost_reg[0] = c[0]; cost_reg[1] = c[1]; }
    if (reg_factor) { reg_factor[0] = f2[0]; reg_factor[1] = f2[1]; }
  }
  HIPC(hipStreamSynchronize(ctx->stream));
  return 0;
}

int hr_set_profiling(hr_ctx* ctx, int mode) {
  if (!ctx) return 1;
  if (int rc = drain_events(ctx)) return rc;
  if (mode == 2) {
    for (int s = 0; s < 5; ++s) { ctx->prof_all[s] = ctx->prof_n[s] = 0; ctx->prof_ms[s] = 0.0; }
    ctx->prof_tick = 0;
    return 0;
  }
  ctx->profiling = mode ? 1 : 0;
  return 0;
}

int hr_get_kernel_stats(hr_ctx* ctx, long long* launches, double* ms) {
  if (!ctx || !launches || !ms) return fail(ctx, 1, "null argument");
  if (int rc = drain_events(ctx)) return rc;
  for (int s = 0; s < 5; ++s) {
    launches[s] = ctx->prof_all[s];
    ms[s] = ctx->prof_n[s] ? ctx->prof_ms[s] * (double)ctx->prof_all[s] / (double)ctx->prof_n[s] : 0.0;
  }
  return 0;
}

}  // extern "C"
