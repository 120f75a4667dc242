// gfx950 reactor library for one mechanism: host side and the C-ABI of include/pelec_amd.h.
// Built once per mechanism with -DPELEC_MECH=<ident> -DPELEC_MECH_HEADER="gen/mech_<name>.h".
//
// prc.form 0 (plain): one launch of react_plain per box; prc.form 1 (batched): k_start, then a
// captured graph of a fixed number of iterations replayed until every started cell is done, with a
// compaction of the live slots whenever fewer than half remain alive. Each iteration forks into
// two branches that run concurrently on disjoint cells: k_jac -> k_lu (set-up list) and k_newton
// (Newton-phase cells). prc.lu_fp32 keeps the Newton-matrix factors in FP32 with FP64 refinement.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "pelec_amd.h"
#include PELEC_MECH_HEADER

using Mech = pelec::mech::PELEC_MECH;

#include "reactor_kernels.h"

namespace {
// Newton-matrix factorisation kernel: one row per lane, a 32-lane group per matrix when n <= 32,
// a whole wave when n <= 64. (Larger systems would need factors streamed through HBM; no
// mechanism of this build has them.)
template <typename LUT>
constexpr auto lu_kernel() {
  static_assert(NV <= 64, "the batched form supports n = N + 1 <= 64");
  if constexpr (NV <= 32)
    return &k_lu_rows<LUT, 32>;
  else
    return &k_lu_rows<LUT, 64>;
}
// launch shape of the batched form (prc.* knobs; prc_autotune picks them by measurement)
struct Shape {
  int iters = 15;           // iterations per captured graph, a multiple of NLISTS
  int coop = 0;             // bit 0: k_newton_g, bit 1: k_jac_g (one cell per 32-lane group)
  int bpc_newton = 0;       // blocks per CU of k_newton / k_jac / the LU kernel (0: occupancy)
  int bpc_jac = 0;
  int bpc_lu = 0;
  int compact_pct = 50;     // compact when fewer than this percentage of the slots are alive
  // tail: once at most tail_slots slots are in use, replay a second graph whose Newton kernel is
  // k_newton_g (bpc_tail blocks per CU): with few live cells a cell's Newton iteration is bound by
  // one lane's latency, which the group form divides among 32 lanes. 0: off.
  int tail_slots = 0;
  int bpc_tail = 0;
  // FP64 refinement steps after each FP32-LU solve (prc.lu_refine). With the equilibrated factors
  // one step makes the Newton corrections FP64-equivalent on every drm19 64^3 cell (each within
  // 7.3e-5 K of the oracle, as with FP64 factors); with none, 1333 of its 262144 cells took
  // other step decisions and ended up to 0.056 K off.
  int refine = 1;
  // prc.compact_every: the graph checks for a compaction after every this many iterations (with
  // the host's rule, compact_pct; only while no cell waits for a slot), so compactions need not
  // wait for the end of a replay. 0: only the host compacts, between replays.
  int compact_every = 0;
  bool operator==(const Shape& o) const {
    return iters == o.iters && coop == o.coop && bpc_newton == o.bpc_newton && bpc_jac == o.bpc_jac &&
           bpc_lu == o.bpc_lu && compact_pct == o.compact_pct && tail_slots == o.tail_slots &&
           bpc_tail == o.bpc_tail && refine == o.refine && compact_every == o.compact_every;
  }
};

// The Newton and Jacobian kernels of one batched iteration: one cell per lane, or one cell per
// 32-lane group (two per 64-lane block)
using PhaseKernel = void (*)(const Args*, int);
struct PhaseKernels {
  PhaseKernel newton, jac;
};
template <typename LUT, int NREF>
PhaseKernels phase_kernels_r(const Shape& sh, bool tail) {
  if constexpr (COOP64) {
    // 53-species mechanisms: one cell per 64-lane wave, the only Newton and Jacobian kernels built
    using GM64 = GroupModel<Mech, 64>;
    return PhaseKernels{&k_newton_g<LUT, 64, NREF, GM64>, &k_jac_g<64, GM64>};
  } else {
    PhaseKernels k{&k_newton<LUT, NREF>, &k_jac<>};
    if constexpr (COOP) {
      if ((sh.coop & 1) || tail) k.newton = &k_newton_g<LUT, 64, NREF>;
      if (sh.coop & 2) k.jac = &k_jac_g<64>;
    }
    return k;
  }
}
template <typename LUT>
PhaseKernels phase_kernels(const Shape& sh, bool tail = false) {
  if constexpr (sizeof(LUT) == sizeof(double))
    return phase_kernels_r<LUT, 1>(sh, tail);
  else
    return sh.refine <= 0   ? phase_kernels_r<LUT, 0>(sh, tail)
           : sh.refine == 1 ? phase_kernels_r<LUT, 1>(sh, tail)
                            : phase_kernels_r<LUT, 2>(sh, tail);
}
}

struct prc_reactor {
  std::string name;
  int integrator = 0;  // 0 ReactorCvode (BDF), 1 ReactorRK64, 2 ReactorNull
  int rk_guess = 10, rk_min = 5, rk_max = 100000;  // rk64.nsubsteps_{guess,min,max}
  bool initialised = false;
  double rtol = 1e-6, atol = 1e-10;
  int max_steps = 100000;
  int max_order = 5;
  int form = 0;
  int lu_fp32 = 0;
  int max_slots = 0;  // prc.max_slots: slots of the batched form (0: one per cell, memory allowing)
  double typ[NV];
  bool have_typ = false;
  double* d_atol = nullptr;
  unsigned long long* d_ctr = nullptr;
  unsigned long long* h_ctr = nullptr;  // pinned
  Args* d_args = nullptr;
  BoxDev* d_boxes = nullptr;
  int d_boxes_cap = 0;
  int32_t* diag = nullptr;
  int64_t diag_cap = 0;
  // batched workspace
  void* ws = nullptr;
  size_t ws_bytes = 0;
  int64_t ws_cap = 0;   // slots
  int64_t ws_req = 0;   // slots asked for when it was allocated
  int ws_lu_fp32 = -1;
  int ws_src = -1;
  Args ws_layout{};
  hipStream_t cap_stream = nullptr;
  hipStream_t cap_stream2 = nullptr;  // second branch of each captured iteration
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  hipGraphExec_t graph_tail = nullptr;  // Shape::tail_slots: the graph with k_newton_g
  int graph_lu_fp32 = -1;
  Shape shape, graph_shape;
  int ncu = 0;
  int occ_newton = 1, occ_jac = 1, occ_lu = 1, occ_tail = 1;
  int64_t last_iterations = 0;
  // PELEC_AMD_TRACE=1: one stderr line per graph replay of the batched form (live-cell counts)
  bool trace = false;
  // adjoint: checkpoint segment length, checkpoints and tape records per cell; workspace
  int adj_seg = 8, adj_kmax = 128, adj_tcap = 256;
  double* aws = nullptr;
  size_t aws_bytes = 0;
  // plain form above PLAIN_PRIVATE_N: matrices of the box's cells
  double* pws = nullptr;
  int64_t pws_cells = 0;
  // react_state workspace: rhoY (N), T, rho e, rho e forcing per cell of the box
  double* sws = nullptr;
  int64_t sws_cap = 0;
  // prc.kernel_timing: batched iterations launched eagerly with HIP events around every
  // k_newton / k_jac / k_lu launch (on the stream each runs on); totals of the last call
  int kernel_timing = 0;
  // prc.test_lu_fail: the batched form's factorisation kernel reports every Newton matrix as
  // singular, so tests can drive the convergence-failure handling from the host
  int test_lu_fail = 0;
  double kt_ms[3] = {0.0, 0.0, 0.0};  // newton, jac, lu
  int64_t kt_launches = 0;
  std::vector<hipEvent_t> kt_ev;
};

namespace {

#define HIP_OK(expr)                           \
  do {                                         \
    if ((expr) != hipSuccess) return PRC_EHIP; \
  } while (0)

A4 mk(const prc_array4& a) {
  return A4{a.p, a.jstride, a.kstride, a.nstride, a.begin[0], a.begin[1], a.begin[2]};
}
A4I mki(const prc_iarray4& a) {
  return A4I{a.p, a.jstride, a.kstride, a.nstride, a.begin[0], a.begin[1], a.begin[2]};
}

int upload_atol(prc_reactor* r) {
  double h[NV];
  for (int i = 0; i < NV; ++i) h[i] = r->have_typ ? r->atol * r->typ[i] : r->atol;
  HIP_OK(hipMemcpy(r->d_atol, h, sizeof(h), hipMemcpyHostToDevice));
  return PRC_OK;
}

BoxDev make_box(const int32_t lo[3], const int32_t hi[3], const prc_array4& rY,
                const prc_array4& rYsrc, const prc_array4& T, const prc_array4& rE,
                const prc_array4& rEsrc, const prc_array4& FC, const prc_iarray4& mask,
                int64_t first) {
  BoxDev b;
  b.lo0 = lo[0];
  b.lo1 = lo[1];
  b.lo2 = lo[2];
  b.len0 = hi[0] - lo[0] + 1;
  b.len1 = hi[1] - lo[1] + 1;
  const int len2 = hi[2] - lo[2] + 1;
  b.ncells = (b.len0 > 0 && b.len1 > 0 && len2 > 0) ? (int64_t)b.len0 * b.len1 * len2 : 0;
  b.first = first;
  b.rY = mk(rY);
  b.rYsrc = mk(rYsrc);
  b.T = mk(T);
  b.rE = mk(rE);
  b.rEsrc = mk(rEsrc);
  b.FC = mk(FC);
  b.mask = mki(mask);
  return b;
}

bool box_ok(const BoxDev& b) { return b.ncells == 0 || (b.rY.p && b.T.p && b.rE.p); }

size_t align_up(size_t x) { return (x + 255) & ~size_t(255); }

// Layout of the batched workspace for C slots (and C matrix blocks) at base: two slot-state
// buffers, per-block J / LU / pivots, the set-up lists, the compaction remap and the free ring of
// matrix blocks. Returns the bytes it spans; with base == nullptr only the size is computed.
size_t workspace_layout(Args& L, char* base, int64_t C, bool lu_fp32, bool has_src) {
  const size_t lub = lu_fp32 ? sizeof(float) : sizeof(double);
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off += align_up(bytes);
    return base + o;
  };
  for (int b = 0; b < 2; ++b) {
    StateBuf& st = L.st[b];
    st.zn = reinterpret_cast<double*>(take(sizeof(double) * (BDF_QMAX + 1) * NV * C));
    st.acor = reinterpret_cast<double*>(take(sizeof(double) * NV * C));
    st.ewt = reinterpret_cast<double*>(take(sizeof(double) * NV * C));
    st.sc = reinterpret_cast<uint64_t*>(take(sizeof(uint64_t) * BDF_SCALAR_WORDS * C));
    st.rhoe = reinterpret_cast<double*>(take(sizeof(double) * 2 * C));
    st.src = has_src ? reinterpret_cast<double*>(take(sizeof(double) * NSP * C)) : nullptr;
    st.cell = reinterpret_cast<int32_t*>(take(sizeof(int32_t) * C));
    st.phase = reinterpret_cast<int32_t*>(take(sizeof(int32_t) * C));
    st.mb = reinterpret_cast<int32_t*>(take(sizeof(int32_t) * C));
  }
  L.C = C;
  L.J = reinterpret_cast<double*>(take(sizeof(double) * NV * NV * C));
  L.LU = take(lub * NV * NV * C);
  L.perm = reinterpret_cast<int32_t*>(take(sizeof(int32_t) * NV * C));
  L.dexp = lu_fp32 ? reinterpret_cast<int32_t*>(take(sizeof(int32_t) * NV * C)) : nullptr;
  for (int b = 0; b < NLISTS; ++b) L.list[b] = reinterpret_cast<int32_t*>(take(sizeof(int32_t) * C));
  for (int b = 0; b < NLISTS; ++b) L.jlist[b] = reinterpret_cast<int32_t*>(take(sizeof(int32_t) * C));
  L.remap = reinterpret_cast<int32_t*>(take(sizeof(int32_t) * C));
  L.ring = reinterpret_cast<int32_t*>(take(sizeof(int32_t) * C));
  return off;
}

// Batched workspace for a call of ncells cells: one slot per cell up to prc.max_slots (0: no
// limit) and up to half of the device memory free at the time; a call with more cells than slots
// refills freed slots from the cell queue (run_batched).
int ensure_workspace(prc_reactor* r, int64_t ncells, bool has_src) {
  const int want_fp32 = r->lu_fp32;
  int64_t want = std::max<int64_t>(ncells, 64);
  if (r->max_slots > 0) want = std::min<int64_t>(want, std::max(r->max_slots, 64));
  // an existing workspace serves when it holds every cell, or when it was already as large as the
  // free memory allowed (the call then refills its slots)
  const bool enough = want <= r->ws_cap || (r->ws_cap < r->ws_req && want >= r->ws_req);
  if (r->ws && enough && r->ws_lu_fp32 == want_fp32 && r->ws_src >= (int)has_src) return PRC_OK;
  if (r->ws) {
    (void)hipFree(r->ws);
    r->ws = nullptr;
    r->ws_cap = 0;
  }
  Args L{};
  const size_t per_slot = workspace_layout(L, nullptr, 1024, want_fp32, has_src) / 1024 + 1;
  size_t free_b = 0, total_b = 0;
  HIP_OK(hipMemGetInfo(&free_b, &total_b));
  const int64_t fit = (int64_t)(free_b / 2 / per_slot) & ~int64_t(63);
  const int64_t cap = std::max<int64_t>(std::min(want, fit), 64);
  const size_t bytes = workspace_layout(L, nullptr, cap, want_fp32, has_src);
  void* p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) return PRC_ENOMEM;
  workspace_layout(r->ws_layout, static_cast<char*>(p), cap, want_fp32, has_src);
  r->ws = p;
  r->ws_bytes = bytes;
  r->ws_cap = cap;
  r->ws_req = want;
  r->ws_lu_fp32 = want_fp32;
  r->ws_src = has_src;
  return PRC_OK;
}

// grid of the Newton kernel (main or tail graph)
int newton_grid(const prc_reactor* r, bool tail) {
  if (tail) return r->ncu * (r->shape.bpc_tail > 0 ? r->shape.bpc_tail : r->occ_tail);
  return r->ncu * (r->shape.bpc_newton > 0 ? r->shape.bpc_newton : r->occ_newton);
}

// One graph's worth of batched iterations enqueued on s (set-up branch) and s2 (Newton branch),
// forked and joined every iteration; the kernels read the call's arguments at `args`.
// The graph's own compaction check after iteration `it` (prc.compact_every): the gate kernel
// takes the host's decision on the device, the three compaction kernels follow it. The set-up
// list pending at that point is the one the iteration appended to.
bool compact_here(const Shape& sh, int it) {
  return sh.compact_every > 0 && (it + 1) % sh.compact_every == 0;
}
void enqueue_compaction(prc_reactor* r, hipStream_t s, const Args* args, int pending_list,
                        int grid) {
  hipLaunchKernelGGL(k_compact_gate, dim3(1), dim3(64), 0, s, args, r->shape.compact_pct);
  hipLaunchKernelGGL(k_compact, dim3(grid), dim3(64), 0, s, args, 1);
  hipLaunchKernelGGL(k_remap, dim3(grid), dim3(64), 0, s, args, pending_list, 1);
  hipLaunchKernelGGL(k_compact_done, dim3(1), dim3(64), 0, s, args, 1);
}

template <typename LUT>
bool enqueue_iterations(prc_reactor* r, hipStream_t s, hipStream_t s2, const Args* args,
                        bool tail) {
  const int gj = r->ncu * (r->shape.bpc_jac > 0 ? r->shape.bpc_jac : r->occ_jac);
  const int gl = r->ncu * (r->shape.bpc_lu > 0 ? r->shape.bpc_lu : r->occ_lu);
  const int gn = newton_grid(r, tail);
  const PhaseKernels pk = phase_kernels<LUT>(r->shape, tail);
  bool ok = true;
  for (int it = 0; it < r->shape.iters && ok; ++it) {
    const int par = it % NLISTS;
    // fork: Newton branch on s2, set-up branch on s; join before the next iteration
    ok = hipEventRecord(r->ev_fork, s) == hipSuccess &&
         hipStreamWaitEvent(s2, r->ev_fork, 0) == hipSuccess;
    if (!ok) break;
    hipLaunchKernelGGL(pk.newton, dim3(gn), dim3(64), 0, s2, args, par);
    hipLaunchKernelGGL(pk.jac, dim3(gj), dim3(64), 0, s, args, par);
    hipLaunchKernelGGL(lu_kernel<LUT>(), dim3(gl), dim3(64), 0, s, args, par);
    ok = hipEventRecord(r->ev_join, s2) == hipSuccess &&
         hipStreamWaitEvent(s, r->ev_join, 0) == hipSuccess;
    if (ok && compact_here(r->shape, it)) enqueue_compaction(r, s, args, list_next(par), gn);
  }
  return ok && hipGetLastError() == hipSuccess;
}

template <typename LUT>
int capture_one(prc_reactor* r, bool tail, hipGraphExec_t* out) {
  hipStream_t s = r->cap_stream, s2 = r->cap_stream2;
  HIP_OK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  const bool ok = enqueue_iterations<LUT>(r, s, s2, r->d_args, tail);
  hipGraph_t graph = nullptr;
  if (hipStreamEndCapture(s, &graph) != hipSuccess || !ok) {
    if (graph) (void)hipGraphDestroy(graph);
    return PRC_EHIP;
  }
  const hipError_t e = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  return e == hipSuccess ? PRC_OK : PRC_EHIP;
}

template <typename LUT>
int capture_graph(prc_reactor* r) {
  if (r->graph_exec && r->graph_lu_fp32 == r->lu_fp32 && r->graph_shape == r->shape) return PRC_OK;
  if (r->graph_exec) (void)hipGraphExecDestroy(r->graph_exec);
  if (r->graph_tail) (void)hipGraphExecDestroy(r->graph_tail);
  r->graph_exec = r->graph_tail = nullptr;
  r->graph_lu_fp32 = -1;
  int e = capture_one<LUT>(r, false, &r->graph_exec);
  if (!e && r->shape.tail_slots > 0) e = capture_one<LUT>(r, true, &r->graph_tail);
  if (e) return e;
  r->graph_lu_fp32 = r->lu_fp32;
  r->graph_shape = r->shape;
  return PRC_OK;
}

template <typename LUT>
int occupancy_setup(prc_reactor* r) {
  int dev = 0;
  HIP_OK(hipGetDevice(&dev));
  HIP_OK(hipDeviceGetAttribute(&r->ncu, hipDeviceAttributeMultiprocessorCount, dev));
  int o = 0;
  const PhaseKernels pk = phase_kernels<LUT>(r->shape);
  HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, pk.newton, 64, 0));
  r->occ_newton = std::max(o, 1);
  HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, pk.jac, 64, 0));
  r->occ_jac = std::max(o, 1);
  HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, lu_kernel<LUT>(), 64, 0));
  r->occ_lu = std::max(o, 1);
  HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, phase_kernels<LUT>(r->shape, true).newton,
                                                      64, 0));  // = k_newton when !COOP
  r->occ_tail = std::max(o, 1);
  return PRC_OK;
}

// One graph's worth of iterations launched eagerly with the graph's fork/join structure and timing
// events around each kernel; the elapsed times are added to r->kt_ms after the replay's sync.
template <typename LUT>
int eager_iterations(prc_reactor* r, hipStream_t s, bool tail) {
  constexpr int EV = 5;  // newton begin/end on s2; jac begin, jac end = lu begin, lu end on s
  const size_t need = (size_t)EV * r->shape.iters;
  while (r->kt_ev.size() < need) {
    hipEvent_t e = nullptr;
    HIP_OK(hipEventCreate(&e));
    r->kt_ev.push_back(e);
  }
  hipStream_t s2 = r->cap_stream2;
  const int gj = r->ncu * (r->shape.bpc_jac > 0 ? r->shape.bpc_jac : r->occ_jac);
  const int gl = r->ncu * (r->shape.bpc_lu > 0 ? r->shape.bpc_lu : r->occ_lu);
  const int gn = newton_grid(r, tail);
  const PhaseKernels pk = phase_kernels<LUT>(r->shape, tail);
  for (int it = 0; it < r->shape.iters; ++it) {
    const int par = it % NLISTS;
    hipEvent_t* e = r->kt_ev.data() + EV * it;
    HIP_OK(hipEventRecord(r->ev_fork, s));
    HIP_OK(hipStreamWaitEvent(s2, r->ev_fork, 0));
    HIP_OK(hipEventRecord(e[0], s2));
    hipLaunchKernelGGL(pk.newton, dim3(gn), dim3(64), 0, s2, r->d_args, par);
    HIP_OK(hipEventRecord(e[1], s2));
    HIP_OK(hipEventRecord(e[2], s));
    hipLaunchKernelGGL(pk.jac, dim3(gj), dim3(64), 0, s, r->d_args, par);
    HIP_OK(hipEventRecord(e[3], s));
    hipLaunchKernelGGL(lu_kernel<LUT>(), dim3(gl), dim3(64), 0, s, r->d_args, par);
    HIP_OK(hipEventRecord(e[4], s));
    HIP_OK(hipEventRecord(r->ev_join, s2));
    HIP_OK(hipStreamWaitEvent(s, r->ev_join, 0));
    if (compact_here(r->shape, it)) enqueue_compaction(r, s, r->d_args, list_next(par), gn);
  }
  HIP_OK(hipGetLastError());
  return PRC_OK;
}

int eager_accumulate(prc_reactor* r) {
  constexpr int EV = 5;
  for (int it = 0; it < r->shape.iters; ++it) {
    hipEvent_t* e = r->kt_ev.data() + EV * it;
    float t = 0.0f;
    HIP_OK(hipEventElapsedTime(&t, e[0], e[1]));
    r->kt_ms[0] += t;
    HIP_OK(hipEventElapsedTime(&t, e[2], e[3]));
    r->kt_ms[1] += t;
    HIP_OK(hipEventElapsedTime(&t, e[3], e[4]));
    r->kt_ms[2] += t;
  }
  r->kt_launches += r->shape.iters;
  return PRC_OK;
}

// The batched form of one call. `dargs` / `dctr` are the device argument block and counters
// (the handle's, or the no-allocation entry point's scratch). With `graphs`, iterations run as
// replays of the handle's captured graphs, which read the handle's argument block; without, they
// are enqueued directly (no graph is instantiated).
template <typename LUT>
int run_batched(prc_reactor* r, Args& a, hipStream_t s, Args* dargs, unsigned long long* dctr,
                bool graphs) {
  if (r->graph_lu_fp32 != r->lu_fp32 || !r->ncu || !(r->graph_shape == r->shape)) {
    const int e = occupancy_setup<LUT>(r);
    if (e) return e;
  }
  // every workspace array the kernels of this shape index must be in the argument block
  if (!a.J || !a.LU || !a.perm || !a.ring || !a.remap || (r->lu_fp32 && !a.dexp))
    return PRC_ESTATE;
  for (int b = 0; b < NLISTS; ++b)
    if (!a.list[b] || !a.jlist[b]) return PRC_ESTATE;
  HIP_OK(hipMemcpyAsync(dargs, &a, sizeof(Args), hipMemcpyHostToDevice, s));
  // the first fill: cells 0 .. count0-1 into slots 0 .. count0-1, each slot its own matrix block
  const int64_t C = a.C;
  const int64_t count0 = std::min(C, a.ncells);
  // compactions inside the graph while no cell of the call waits for a slot
  hipLaunchKernelGGL(k_init_counters, dim3(1), dim3(64), 0, s, dctr, count0,
                     count0 == a.ncells ? 1 : 0);
  hipLaunchKernelGGL(k_start, dim3((unsigned)((count0 + 63) / 64)), dim3(64), 0, s, dargs,
                     (int64_t)0, count0, (int64_t)0, (int64_t)-1);
  HIP_OK(hipGetLastError());
  int64_t next = count0;  // next cell of the call to start
  int64_t head = 0;       // free-ring entries taken so far
  // the argument block is read by every kernel that follows: `a` (host) stays untouched until the
  // copy has completed
  HIP_OK(hipStreamSynchronize(s));
  int e = PRC_OK;
  if (graphs) {
    // the captured graphs only read r->d_args, so they are independent of the call
    e = capture_graph<LUT>(r);
    if (e) return e;
  }
  const int gn = r->ncu * (r->shape.bpc_newton > 0 ? r->shape.bpc_newton : r->occ_newton);
  int64_t iters = 0;
  // safety net: every cell ends within max_steps steps of a bounded number of attempts, and the
  // slots are refilled at most ncells / C times over
  const int64_t max_iters = ((int64_t)(r->max_steps + 1) * 64 + 64) * (1 + a.ncells / C);
  int64_t nactive = count0;
  r->kt_ms[0] = r->kt_ms[1] = r->kt_ms[2] = 0.0;
  r->kt_launches = 0;
  for (;;) {
    const bool tail = r->shape.tail_slots > 0 && nactive <= r->shape.tail_slots;
    if (!graphs) {
      if (!enqueue_iterations<LUT>(r, s, r->cap_stream2, dargs, tail)) return PRC_EHIP;
    } else if (r->kernel_timing) {
      e = eager_iterations<LUT>(r, s, tail);
      if (e) return e;
    } else {
      HIP_OK(hipGraphLaunch(tail ? r->graph_tail : r->graph_exec, s));
    }
    iters += r->shape.iters;
    HIP_OK(hipMemcpyAsync(r->h_ctr, dctr, CT_COUNT * sizeof(unsigned long long),
                          hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    if (graphs && r->kernel_timing) {
      e = eager_accumulate(r);
      if (e) return e;
    }
    const int64_t started = (int64_t)r->h_ctr[CT_STARTED];
    const int64_t done = (int64_t)r->h_ctr[CT_DONE];
    nactive = (int64_t)r->h_ctr[CT_NACTIVE];  // the graph may have compacted
    const bool pending = next < a.ncells;
    if (done >= started && !pending) break;
    if (iters > max_iters) return PRC_ESTATE;
    const int64_t alive = started - done;
    if (r->trace)
      std::fprintf(stderr, "[pelec_amd] iter %lld started %lld done %lld alive %lld slots %lld "
                   "queued %lld\n", (long long)iters, (long long)started, (long long)done,
                   (long long)alive, (long long)nactive, (long long)(a.ncells - next));
    if (alive * 100 < nactive * r->shape.compact_pct && (nactive > 4096 || pending)) {
      hipLaunchKernelGGL(k_compact, dim3(gn), dim3(64), 0, s, dargs, 0);
      hipLaunchKernelGGL(k_remap, dim3(gn), dim3(64), 0, s, dargs, 0, 0);
      hipLaunchKernelGGL(k_compact_done, dim3(1), dim3(64), 0, s, dargs, 0);
      HIP_OK(hipGetLastError());
      nactive = alive;
      if (pending) {
        // refill the freed slots [alive, C) with the next cells; their matrix blocks come from
        // the free ring, which holds C - alive blocks (every slot filled so far either holds a
        // live cell or has returned its block)
        const int64_t k = std::min(C - alive, a.ncells - next);
        hipLaunchKernelGGL(k_start, dim3((unsigned)((k + 63) / 64)), dim3(64), 0, s, dargs,
                           alive, k, next, head);
        HIP_OK(hipGetLastError());
        next += k;
        head += k;
        nactive = alive + k;
        if (next == a.ncells)  // the queue is empty: the graph may compact from now on
          hipLaunchKernelGGL(k_set_counter, dim3(1), dim3(64), 0, s, dctr, (int)CT_INGRAPH, 1ull);
      }
    }
  }
  r->last_iterations = iters;
  return PRC_OK;
}

// Integrator parameters, diagnostics and counters of the handle into the call's argument block.
void fill_params(const prc_reactor* r, Args& a, unsigned long long* ctr) {
  a.prm.rtol = r->rtol;
  a.prm.atolv = r->d_atol;
  a.prm.qmax = std::min(std::max(r->max_order, 1), BDF_QMAX);
  a.prm.max_steps = r->max_steps;
  a.rk = Rk64Params{r->rtol, r->d_atol, r->rk_guess, r->rk_min, r->rk_max, r->max_steps};
  a.diag = r->diag;
  a.diag_cap = r->diag_cap;
  a.ctr = ctr;
}

// One react call: a.box (single box) or a.boxes / a.nboxes (device table), a.ncells cells.
int launch(prc_reactor* r, Args& a, bool has_src, hipStream_t s, bool reset_counters = true) {
  if (r->integrator == 1 && r->rk_min > r->rk_max) return PRC_EINVAL;
  fill_params(r, a, r->d_ctr);
  if (a.ncells == 0 || r->form == 0 || r->integrator == 1) {
    if (reset_counters) {
      hipLaunchKernelGGL(k_init_counters, dim3(1), dim3(64), 0, s, r->d_ctr, (int64_t)0, 0);
      HIP_OK(hipGetLastError());
    }
    if (a.ncells == 0) return PRC_OK;
    a.pws = nullptr;
    if (PLAIN_EXT && r->integrator == 0) {
      if (a.box.ncells > r->pws_cells) {
        if (r->pws) (void)hipFree(r->pws);
        r->pws = nullptr;
        r->pws_cells = 0;
        if (hipMalloc(&r->pws, sizeof(double) * 2 * NV * NV * a.box.ncells) != hipSuccess)
          return PRC_ENOMEM;
        r->pws_cells = a.box.ncells;
      }
      a.pws = r->pws;
    }
    HIP_OK(hipMemcpyAsync(r->d_args, &a, sizeof(Args), hipMemcpyHostToDevice, s));
    const int64_t nb = (a.box.ncells + 63) / 64;
    if (r->integrator == 1)
      hipLaunchKernelGGL(react_rk64, dim3((unsigned)nb), dim3(64), 0, s, r->d_args);
    else if (r->lu_fp32 && r->shape.refine >= 2)
      hipLaunchKernelGGL((react_plain<float, 2>), dim3((unsigned)nb), dim3(64), 0, s, r->d_args);
    else if (r->lu_fp32)
      hipLaunchKernelGGL((react_plain<float, 1>), dim3((unsigned)nb), dim3(64), 0, s, r->d_args);
    else
      hipLaunchKernelGGL((react_plain<double, 1>), dim3((unsigned)nb), dim3(64), 0, s, r->d_args);
    HIP_OK(hipGetLastError());
    // the argument block is reused by the next call: keep it until this launch has read it
    HIP_OK(hipStreamSynchronize(s));
    return PRC_OK;
  }
  int e = ensure_workspace(r, a.ncells, has_src);
  if (e) return e;
  const Args& L = r->ws_layout;
  a.st[0] = L.st[0];
  a.st[1] = L.st[1];
  a.C = L.C;
  a.J = L.J;
  a.LU = L.LU;
  a.perm = L.perm;
  a.dexp = L.dexp;
  for (int b = 0; b < NLISTS; ++b) a.list[b] = L.list[b];
  for (int b = 0; b < NLISTS; ++b) a.jlist[b] = L.jlist[b];
  a.remap = L.remap;
  a.ring = L.ring;
  a.has_src = has_src ? 1 : 0;
  a.test_lu_fail = r->test_lu_fail;
  return r->lu_fp32 ? run_batched<float>(r, a, s, r->d_args, r->d_ctr, true)
                    : run_batched<double>(r, a, s, r->d_args, r->d_ctr, true);
}

int ensure_init(prc_reactor* r) {
  if (!r) return PRC_EINVAL;
  if (!r->initialised) return PRC_ESTATE;
  return PRC_OK;
}

template <typename K, typename... Xs>
int launch_1d(K kern, int64_t nc, void* stream, Xs... xs) {
  if (nc <= 0) return PRC_OK;
  const int64_t nb = (nc + 255) / 256;
  hipLaunchKernelGGL(kern, dim3((unsigned)nb), dim3(256), 0, (hipStream_t)stream, nc, xs...);
  HIP_OK(hipGetLastError());
  return PRC_OK;
}

void release(prc_reactor* r) {
  if (r->graph_exec) (void)hipGraphExecDestroy(r->graph_exec);
  if (r->graph_tail) (void)hipGraphExecDestroy(r->graph_tail);
  if (r->cap_stream) (void)hipStreamDestroy(r->cap_stream);
  if (r->cap_stream2) (void)hipStreamDestroy(r->cap_stream2);
  if (r->ev_fork) (void)hipEventDestroy(r->ev_fork);
  if (r->ev_join) (void)hipEventDestroy(r->ev_join);
  if (r->ws) (void)hipFree(r->ws);
  if (r->d_atol) (void)hipFree(r->d_atol);
  if (r->d_ctr) (void)hipFree(r->d_ctr);
  if (r->h_ctr) (void)hipHostFree(r->h_ctr);
  if (r->d_args) (void)hipFree(r->d_args);
  if (r->d_boxes) (void)hipFree(r->d_boxes);
  r->graph_exec = r->graph_tail = nullptr;
  for (hipEvent_t e : r->kt_ev) (void)hipEventDestroy(e);
  r->kt_ev.clear();
  r->cap_stream = nullptr;
  r->cap_stream2 = nullptr;
  r->ev_fork = r->ev_join = nullptr;
  r->ws = nullptr;
  r->ws_cap = 0;
  r->ws_lu_fp32 = r->ws_src = -1;
  r->d_atol = nullptr;
  r->d_ctr = nullptr;
  r->h_ctr = nullptr;
  r->d_args = nullptr;
  r->d_boxes = nullptr;
  r->d_boxes_cap = 0;
  if (r->sws) (void)hipFree(r->sws);
  r->sws = nullptr;
  r->sws_cap = 0;
  if (r->aws) (void)hipFree(r->aws);
  r->aws = nullptr;
  r->aws_bytes = 0;
  if (r->pws) (void)hipFree(r->pws);
  r->pws = nullptr;
  r->pws_cells = 0;
  r->graph_lu_fp32 = -1;
  r->ncu = 0;
  r->initialised = false;
}

// ---- the no-allocation entry point's scratch (prc_react_flat_noalloc) ------------------------
// One array per GPU the library runs on, held until the process ends: the argument block, the
// work-queue counters and the batched workspace of prc_react_flat_noalloc. In host code `scratch`
// names a host-side stand-in; its device address on each GPU is looked up once per device at
// prc_init (hipGetSymbolAddress) and only that address is passed to kernels or copies.
constexpr size_t NOALLOC_SCRATCH_BYTES = PRC_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];
constexpr int NOALLOC_MAX_DEVICES = 64;
void* g_scratch[NOALLOC_MAX_DEVICES] = {};
std::mutex g_scratch_mu;

// Device address of this library's scratch on the current device (set-up, under the lock).
int scratch_setup() {
  int dev = 0;
  HIP_OK(hipGetDevice(&dev));
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return PRC_EHIP;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch[dev]) return PRC_OK;
  void* ptr = nullptr;
  if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return PRC_EHIP;
  g_scratch[dev] = ptr;
  return PRC_OK;
}

void* scratch_ptr() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch[dev];
}

}  // namespace

extern "C" {

int prc_num_species(void) { return NSP; }
int prc_num_reactions(void) { return Mech::NRXN; }
const char* prc_species_name(int k) { return (k >= 0 && k < NSP) ? Mech::SPECIES[k] : nullptr; }
const char* prc_mechanism_name(void) { return Mech::NAME; }
const char* prc_mechanism_sha256(void) { return Mech::SHA256; }

int prc_create(const char* name, prc_reactor** out) {
  if (!name || !out) return PRC_EINVAL;
  int kind;
  if (!std::strcmp(name, "ReactorCvode"))
    kind = 0;
  else if (!std::strcmp(name, "ReactorRK64"))
    kind = 1;
  else if (!std::strcmp(name, "ReactorNull"))
    kind = 2;
  else
    return PRC_EINVAL;
  prc_reactor* r = new (std::nothrow) prc_reactor;
  if (!r) return PRC_ENOMEM;
  const char* tr = std::getenv("PELEC_AMD_TRACE");
  r->trace = tr && tr[0] && tr[0] != '0';
  r->name = name;
  r->integrator = kind;
  *out = r;
  return PRC_OK;
}

int prc_init(prc_reactor* r, int reactor_type, int ncells) {
  (void)ncells;
  if (!r || reactor_type != 1) return PRC_EINVAL;
  if (!r->initialised) {
    if (hipMalloc(&r->d_atol, NV * sizeof(double)) != hipSuccess ||
        hipMalloc(&r->d_ctr, CT_COUNT * sizeof(unsigned long long)) != hipSuccess ||
        hipMalloc(&r->d_args, sizeof(Args)) != hipSuccess ||
        hipHostMalloc(&r->h_ctr, CT_COUNT * sizeof(unsigned long long), 0) != hipSuccess) {
      release(r);
      return PRC_ENOMEM;
    }
    if (hipStreamCreateWithFlags(&r->cap_stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&r->cap_stream2, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&r->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&r->ev_join, hipEventDisableTiming) != hipSuccess ||
        scratch_setup() != PRC_OK) {
      release(r);
      return PRC_EHIP;
    }
    r->initialised = true;
  }
  return upload_atol(r);
}

int prc_set_typ_vals_ode(prc_reactor* r, const double* typ, int n) {
  if (!r) return PRC_EINVAL;
  if (n == 0) {
    r->have_typ = false;
  } else {
    if (n != NV || !typ) return PRC_EINVAL;
    for (int i = 0; i < NV; ++i) r->typ[i] = typ[i];
    r->have_typ = true;
  }
  return r->initialised ? upload_atol(r) : PRC_OK;
}

int prc_set_param_real(prc_reactor* r, const char* key, double v) {
  if (!r || !key) return PRC_EINVAL;
  if (!std::strcmp(key, "ode.rtol")) {
    if (!(v > 0)) return PRC_EINVAL;
    r->rtol = v;
  } else if (!std::strcmp(key, "ode.atol")) {
    if (!(v > 0)) return PRC_EINVAL;
    r->atol = v;
  } else {
    return PRC_EINVAL;
  }
  return r->initialised ? upload_atol(r) : PRC_OK;
}

int prc_set_param_int(prc_reactor* r, const char* key, int v) {
  if (!r || !key) return PRC_EINVAL;
  if (!std::strcmp(key, "ode.max_steps")) {
    if (v < 1) return PRC_EINVAL;
    r->max_steps = v;
  } else if (!std::strcmp(key, "cvode.max_order")) {
    if (v < 1 || v > BDF_QMAX) return PRC_EINVAL;
    r->max_order = v;
  } else if (!std::strcmp(key, "prc.form")) {
    if (v < 0 || v > 1) return PRC_EINVAL;
    r->form = v;
  } else if (!std::strcmp(key, "prc.lu_fp32")) {
    r->lu_fp32 = v ? 1 : 0;
  } else if (!std::strcmp(key, "rk64.nsubsteps_guess")) {
    if (v < 1) return PRC_EINVAL;
    r->rk_guess = v;
  } else if (!std::strcmp(key, "rk64.nsubsteps_min")) {
    if (v < 1) return PRC_EINVAL;
    r->rk_min = v;
  } else if (!std::strcmp(key, "rk64.nsubsteps_max")) {
    if (v < 1) return PRC_EINVAL;
    r->rk_max = v;
  } else if (!std::strcmp(key, "adj.checkpoint_steps")) {
    if (v < 1) return PRC_EINVAL;
    r->adj_seg = v;
  } else if (!std::strcmp(key, "adj.max_checkpoints")) {
    if (v < 1) return PRC_EINVAL;
    r->adj_kmax = v;
  } else if (!std::strcmp(key, "adj.segment_records")) {
    if (v < 8) return PRC_EINVAL;
    r->adj_tcap = v;
  } else if (!std::strcmp(key, "prc.iters_per_graph")) {
    if (v < NLISTS || v > 300 || v % NLISTS) return PRC_EINVAL;
    r->shape.iters = v;
  } else if (!std::strcmp(key, "prc.lu_refine")) {
    if (v < 0 || v > 2) return PRC_EINVAL;
    r->shape.refine = v;
  } else if (!std::strcmp(key, "prc.cooperative")) {
    if (v < 0 || v > 3 || (v && !COOP)) return PRC_EINVAL;
    r->shape.coop = v;
  } else if (!std::strcmp(key, "prc.coop_tail_slots")) {
    if (v < 0 || (v && !COOP)) return PRC_EINVAL;
    r->shape.tail_slots = v;

  } else if (!std::strcmp(key, "prc.blocks_tail")) {
    if (v < 0 || v > 64) return PRC_EINVAL;
    r->shape.bpc_tail = v;
  } else if (!std::strcmp(key, "prc.blocks_newton")) {
    if (v < 0 || v > 64) return PRC_EINVAL;
    r->shape.bpc_newton = v;
  } else if (!std::strcmp(key, "prc.blocks_jac")) {
    if (v < 0 || v > 64) return PRC_EINVAL;
    r->shape.bpc_jac = v;
  } else if (!std::strcmp(key, "prc.blocks_lu")) {
    if (v < 0 || v > 64) return PRC_EINVAL;
    r->shape.bpc_lu = v;
  } else if (!std::strcmp(key, "prc.compact_pct")) {
    if (v < 1 || v > 100) return PRC_EINVAL;
    r->shape.compact_pct = v;
  } else if (!std::strcmp(key, "prc.compact_every")) {
    if (v < 0 || v > 300) return PRC_EINVAL;
    r->shape.compact_every = v;
  } else if (!std::strcmp(key, "prc.max_slots")) {
    if (v < 0) return PRC_EINVAL;
    r->max_slots = v;
  } else if (!std::strcmp(key, "prc.kernel_timing")) {
    r->kernel_timing = v ? 1 : 0;
  } else if (!std::strcmp(key, "prc.test_lu_fail")) {
    r->test_lu_fail = v ? 1 : 0;
  } else {
    return PRC_EINVAL;
  }
  return PRC_OK;
}

int prc_get_param_int(prc_reactor* r, const char* key, int* v) {
  if (!r || !key || !v) return PRC_EINVAL;
  if (!std::strcmp(key, "prc.iters_per_graph"))
    *v = r->shape.iters;
  else if (!std::strcmp(key, "prc.blocks_newton"))
    *v = r->shape.bpc_newton > 0 ? r->shape.bpc_newton : r->occ_newton;
  else if (!std::strcmp(key, "prc.blocks_jac"))
    *v = r->shape.bpc_jac > 0 ? r->shape.bpc_jac : r->occ_jac;
  else if (!std::strcmp(key, "prc.blocks_lu"))
    *v = r->shape.bpc_lu > 0 ? r->shape.bpc_lu : r->occ_lu;
  else if (!std::strcmp(key, "prc.compact_pct"))
    *v = r->shape.compact_pct;
  else if (!std::strcmp(key, "prc.compact_every"))
    *v = r->shape.compact_every;
  else if (!std::strcmp(key, "prc.cooperative"))
    *v = r->shape.coop;
  else if (!std::strcmp(key, "prc.coop_tail_slots"))
    *v = r->shape.tail_slots;

  else if (!std::strcmp(key, "prc.blocks_tail"))
    *v = r->shape.bpc_tail > 0 ? r->shape.bpc_tail : r->occ_tail;
  else if (!std::strcmp(key, "prc.form"))
    *v = r->form;
  else if (!std::strcmp(key, "prc.lu_fp32"))
    *v = r->lu_fp32;
  else if (!std::strcmp(key, "prc.lu_refine"))
    *v = r->shape.refine;
  else if (!std::strcmp(key, "prc.max_slots"))
    *v = r->max_slots;
  else if (!std::strcmp(key, "prc.slots"))  // slots of the current batched workspace
    *v = (int)r->ws_cap;
  else
    return PRC_EINVAL;
  return PRC_OK;
}

int prc_set_diagnostics(prc_reactor* r, int32_t* diag, int64_t capacity) {
  if (!r || capacity < 0) return PRC_EINVAL;
  r->diag = diag;
  r->diag_cap = diag ? capacity : 0;
  return PRC_OK;
}

int prc_close(prc_reactor* r) {
  if (!r) return PRC_EINVAL;
  release(r);
  return PRC_OK;
}

int prc_destroy(prc_reactor* r) {
  if (!r) return PRC_EINVAL;
  release(r);
  delete r;
  return PRC_OK;
}

int prc_react_box(prc_reactor* r, const int32_t lo[3], const int32_t hi[3], prc_array4 rY,
                  prc_array4 rYsrc, prc_array4 T, prc_array4 rE, prc_array4 rEsrc,
                  prc_array4 FC, prc_iarray4 mask, double* dt_react, double* time,
                  void* stream) {
  int e = ensure_init(r);
  if (e) return e;
  if (!lo || !hi || !dt_react || !time) return PRC_EINVAL;
  if (rY.ncomp != 0 && rY.ncomp < NSP) return PRC_EINVAL;
  Args a{};
  a.box = make_box(lo, hi, rY, rYsrc, T, rE, rEsrc, FC, mask, 0);
  if (!box_ok(a.box)) return PRC_EINVAL;
  a.boxes = nullptr;
  a.nboxes = 1;
  a.ncells = a.box.ncells;
  a.dt = *dt_react;
  if (r->integrator != 2) {
    e = launch(r, a, rYsrc.p != nullptr, (hipStream_t)stream);
    if (e) return e;
  }
  *time += *dt_react;
  return PRC_OK;
}

int prc_react_flat(prc_reactor* r, double* rY, const double* rYsrc, double* T, double* rE,
                   const double* rEsrc, double* FC, const int32_t* mask, int64_t ncells,
                   double* dt_react, double* time, void* stream) {
  int e = ensure_init(r);
  if (e) return e;
  if (ncells < 0 || ncells > INT32_MAX || !dt_react || !time) return PRC_EINVAL;
  if (ncells == 0) {
    *time += *dt_react;
    return PRC_OK;
  }
  const int32_t lo[3] = {0, 0, 0};
  const int32_t hi[3] = {(int32_t)ncells - 1, 0, 0};
  auto view = [&](const double* p, int nc) {
    prc_array4 v{};
    v.p = const_cast<double*>(p);
    v.jstride = v.kstride = v.nstride = ncells;
    v.end[0] = (int32_t)ncells;
    v.end[1] = v.end[2] = 1;
    v.ncomp = nc;
    return v;
  };
  prc_iarray4 m{};
  m.p = const_cast<int32_t*>(mask);
  m.jstride = m.kstride = m.nstride = ncells;
  m.end[0] = (int32_t)ncells;
  m.end[1] = m.end[2] = 1;
  m.ncomp = 1;
  return prc_react_box(r, lo, hi, view(rY, NSP), view(rYsrc, NSP), view(T, 1), view(rE, 1),
                       view(rEsrc, 1), view(FC, 1), m, dt_react, time, stream);
}

int prc_react_flat_noalloc(prc_reactor* r, double* rY, const double* rYsrc, double* T, double* rE,
                           const double* rEsrc, double* FC, const int32_t* mask, int64_t ncells,
                           double* dt_react, double* time, void* stream) {
  int e = ensure_init(r);
  if (e) return e;
  if (ncells < 0 || ncells > INT32_MAX || !dt_react || !time) return PRC_EINVAL;
  if (r->integrator == 1) return PRC_EINVAL;  // ReactorRK64 has no batched form
  if (ncells == 0 || r->integrator == 2) {
    *time += *dt_react;
    return PRC_OK;
  }
  if (!rY || !T || !rE) return PRC_EINVAL;
  char* base = static_cast<char*>(scratch_ptr());
  if (!base) return PRC_ESTATE;  // prc_init ran on another device
  hipStream_t s = (hipStream_t)stream;
  // scratch: [argument block][counters][batched workspace of as many slots as fit]
  const size_t o_ctr = align_up(sizeof(Args));
  const size_t o_ws = o_ctr + align_up(CT_COUNT * sizeof(unsigned long long));
  const bool has_src = rYsrc != nullptr;
  Args a{};
  const size_t per64 = workspace_layout(a, nullptr, 64, r->lu_fp32, has_src);
  if (o_ws + per64 > NOALLOC_SCRATCH_BYTES) return PRC_ENOMEM;  // not even 64 slots fit
  int64_t C = 64;
  while (o_ws + workspace_layout(a, nullptr, C + 64, r->lu_fp32, has_src) <= NOALLOC_SCRATCH_BYTES &&
         C < ncells)
    C += 64;
  workspace_layout(a, base + o_ws, C, r->lu_fp32, has_src);
  Args* dargs = reinterpret_cast<Args*>(base);
  unsigned long long* dctr = reinterpret_cast<unsigned long long*>(base + o_ctr);
  fill_params(r, a, dctr);
  auto view = [&](const double* p) {
    A4 v{};
    v.p = const_cast<double*>(p);
    v.js = v.ks = v.ns = ncells;
    return v;
  };
  a.box.lo0 = a.box.lo1 = a.box.lo2 = 0;
  a.box.len0 = (int)ncells;
  a.box.len1 = 1;
  a.box.ncells = ncells;
  a.box.first = 0;
  a.box.rY = view(rY);
  a.box.rYsrc = view(rYsrc);
  a.box.T = view(T);
  a.box.rE = view(rE);
  a.box.rEsrc = view(rEsrc);
  a.box.FC = view(FC);
  a.box.mask = A4I{const_cast<int32_t*>(mask), ncells, ncells, ncells, 0, 0, 0};
  a.boxes = nullptr;
  a.nboxes = 1;
  a.ncells = ncells;
  a.dt = *dt_react;
  a.has_src = has_src ? 1 : 0;
  a.test_lu_fail = r->test_lu_fail;
  e = r->lu_fp32 ? run_batched<float>(r, a, s, dargs, dctr, false)
                 : run_batched<double>(r, a, s, dargs, dctr, false);
  if (e) return e;
  // prc_last_failures reads the handle's counters
  HIP_OK(hipMemcpyAsync(r->d_ctr, dctr, CT_COUNT * sizeof(unsigned long long),
                        hipMemcpyDeviceToDevice, s));
  *time += *dt_react;
  return PRC_OK;
}

int prc_react_boxes(prc_reactor* r, int nboxes, const prc_box_args* boxes, double* dt_react,
                    double* time, void* stream) {
  int e = ensure_init(r);
  if (e) return e;
  if (nboxes < 0 || (nboxes > 0 && !boxes) || !dt_react || !time) return PRC_EINVAL;
  if (r->integrator == 2 || nboxes == 0) {
    *time += *dt_react;
    return PRC_OK;
  }
  hipStream_t s = (hipStream_t)stream;
  if (r->form == 0 || r->integrator == 1) {
    // plain form (and RK64): box by box, as the reference loops over MFIter boxes
    // (React.cpp:137-228)
    int64_t first = 0;
    for (int ib = 0; ib < nboxes; ++ib) {
      const prc_box_args& x = boxes[ib];
      Args a{};
      a.box = make_box(x.lo, x.hi, x.rY, x.rYsrc_ext, x.T, x.rEner, x.rEner_src_ext, x.FC,
                       x.mask, first);
      if (!box_ok(a.box)) return PRC_EINVAL;
      a.nboxes = 1;
      a.ncells = a.box.ncells;
      a.dt = *dt_react;
      first += a.box.ncells;
      if (a.box.ncells == 0 && ib > 0) continue;
      e = launch(r, a, x.rYsrc_ext.p != nullptr, s, ib == 0);
      if (e) return e;
    }
    *time += *dt_react;
    return PRC_OK;
  }
  // batched form: one work queue over the concatenation of all boxes
  BoxDev* hb = new (std::nothrow) BoxDev[nboxes];
  if (!hb) return PRC_ENOMEM;
  int64_t first = 0;
  bool has_src = false;
  for (int ib = 0; ib < nboxes; ++ib) {
    const prc_box_args& x = boxes[ib];
    hb[ib] = make_box(x.lo, x.hi, x.rY, x.rYsrc_ext, x.T, x.rEner, x.rEner_src_ext, x.FC,
                      x.mask, first);
    has_src = has_src || x.rYsrc_ext.p != nullptr;
    if (!box_ok(hb[ib])) {
      delete[] hb;
      return PRC_EINVAL;
    }
    first += hb[ib].ncells;
  }
  if (nboxes > r->d_boxes_cap) {
    if (r->d_boxes) (void)hipFree(r->d_boxes);
    r->d_boxes = nullptr;
    r->d_boxes_cap = 0;
    if (hipMalloc(&r->d_boxes, sizeof(BoxDev) * nboxes) != hipSuccess) {
      delete[] hb;
      return PRC_ENOMEM;
    }
    r->d_boxes_cap = nboxes;
  }
  const hipError_t ce = hipMemcpy(r->d_boxes, hb, sizeof(BoxDev) * nboxes, hipMemcpyHostToDevice);
  delete[] hb;
  if (ce != hipSuccess) return PRC_EHIP;
  Args a{};
  a.boxes = r->d_boxes;
  a.nboxes = nboxes;
  a.ncells = first;
  a.dt = *dt_react;
  e = launch(r, a, has_src, s);
  if (e) return e;
  *time += *dt_react;
  return PRC_OK;
}

int64_t prc_last_failures(prc_reactor* r, void* stream) {
  if (!r || !r->initialised) return PRC_EINVAL;
  if (hipMemcpyAsync(r->h_ctr, r->d_ctr, CT_COUNT * sizeof(unsigned long long),
                     hipMemcpyDeviceToHost, (hipStream_t)stream) != hipSuccess)
    return PRC_EHIP;
  if (hipStreamSynchronize((hipStream_t)stream) != hipSuccess) return PRC_EHIP;
  return (int64_t)r->h_ctr[CT_FAILED];
}

int prc_autotune(prc_reactor* r, int64_t ncells, const double* rY, const double* T,
                 const double* rE, double dt, double* best_ms, void* stream) {
  int e = ensure_init(r);
  if (e) return e;
  if (r->form != 1 || r->integrator != 0 || ncells <= 0 || !rY || !T || !rE || !(dt > 0.0))
    return PRC_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  const size_t bytes = sizeof(double) * (size_t)(NSP + 2) * ncells;
  double* work = nullptr;
  if (hipMalloc(&work, bytes) != hipSuccess) return PRC_ENOMEM;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    (void)hipFree(work);
    if (e0) (void)hipEventDestroy(e0);
    return PRC_EHIP;
  }
  // the caller's diagnostics buffer is left alone while tuning
  int32_t* diag = r->diag;
  const int64_t diag_cap = r->diag_cap;
  r->diag = nullptr;
  r->diag_cap = 0;
  double* wY = work;
  double* wT = work + (size_t)NSP * ncells;
  double* wE = wT + ncells;
  // one timed react call of the sample with the current shape (best of two)
  auto measure = [&](double& ms) -> int {
    ms = 1e300;
    for (int rep = 0; rep < 2; ++rep) {
      HIP_OK(hipMemcpyAsync(wY, rY, sizeof(double) * NSP * ncells, hipMemcpyDeviceToDevice, s));
      HIP_OK(hipMemcpyAsync(wT, T, sizeof(double) * ncells, hipMemcpyDeviceToDevice, s));
      HIP_OK(hipMemcpyAsync(wE, rE, sizeof(double) * ncells, hipMemcpyDeviceToDevice, s));
      double d = dt, tm = 0.0;
      HIP_OK(hipEventRecord(e0, s));
      const int rc = prc_react_flat(r, wY, nullptr, wT, wE, nullptr, nullptr, nullptr, ncells, &d,
                                    &tm, s);
      if (rc) return rc;
      HIP_OK(hipEventRecord(e1, s));
      HIP_OK(hipEventSynchronize(e1));
      float t = 0.0f;
      HIP_OK(hipEventElapsedTime(&t, e0, e1));
      ms = std::min(ms, (double)t);
    }
    return PRC_OK;
  };
  int rc = PRC_OK;
  double best = 0.0;
  {
    // occupancy-derived defaults are the starting point
    Shape base = r->shape;
    if (!r->ncu) rc = r->lu_fp32 ? occupancy_setup<float>(r) : occupancy_setup<double>(r);
    if (!rc) rc = measure(best);
    Shape cur = r->shape;
    // kernel pair first (each one cell per lane or per group), at occupancy-derived block counts
    for (int c = 0; c <= (COOP ? 3 : 0) && !rc; ++c) {
      Shape trial = cur;
      trial.coop = c;
      trial.bpc_newton = trial.bpc_jac = trial.bpc_lu = 0;
      if (trial == cur) continue;
      r->shape = trial;
      double ms = 0.0;
      rc = measure(ms);
      if (!rc && ms < best) {
        best = ms;
        cur = trial;
      }
    }
    r->shape = cur;
    if (!rc) rc = r->lu_fp32 ? occupancy_setup<float>(r) : occupancy_setup<double>(r);
    const int on = r->occ_newton, oj = r->occ_jac, ol = r->occ_lu, ot = r->occ_tail;
    // tail graph thresholds relative to the sample (slots in use when the tail graph takes over)
    const int nt = (int)std::min<int64_t>(ncells, 1 << 30);
    struct Axis {
      int Shape::*field;
      int cand[4];
    };
    const Axis axes[] = {{&Shape::iters, {6, 15, 30, 30}},
                         {&Shape::bpc_newton, {on, 2 * on, 4 * on, 4 * on}},
                         {&Shape::bpc_jac, {oj, 2 * oj, 4 * oj, 4 * oj}},
                         {&Shape::bpc_lu, {ol, 2 * ol, 4 * ol, 4 * ol}},
                         {&Shape::compact_pct, {25, 50, 75, 75}},
                         {&Shape::compact_every, {0, 3, 6, 6}},
                         {&Shape::tail_slots, {0, nt / 64, nt / 16, nt / 4}},
                         {&Shape::bpc_tail, {ot, 4 * ot, 8 * ot, 16 * ot}}};
    for (const Axis& ax : axes) {
      if (rc) break;
      if (ax.field == &Shape::bpc_tail && cur.tail_slots == 0) continue;
      if (ax.field == &Shape::tail_slots && !COOP) continue;
      for (int c : ax.cand) {
        Shape trial = cur;
        const bool blocks = ax.field != &Shape::iters && ax.field != &Shape::compact_pct &&
                            ax.field != &Shape::tail_slots && ax.field != &Shape::compact_every;
        trial.*ax.field = blocks ? std::min(c, 64) : c;
        if (trial == cur) continue;
        r->shape = trial;
        double ms = 0.0;
        rc = measure(ms);
        if (rc) break;
        if (ms < best) {
          best = ms;
          cur = trial;
        }
      }
    }
    r->shape = rc ? base : cur;
  }
  r->diag = diag;
  r->diag_cap = diag_cap;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(work);
  if (best_ms) *best_ms = best;
  return rc;
}

int prc_kernel_times(prc_reactor* r, double* ms, int64_t* launches) {
  if (!r || !ms) return PRC_EINVAL;
  for (int k = 0; k < 3; ++k) ms[k] = r->kt_ms[k];
  if (launches) *launches = r->kt_launches;
  return PRC_OK;
}

int prc_react_state_box(prc_reactor* r, const int32_t lo[3], const int32_t hi[3],
                        prc_array4 S_old, prc_array4 S_new, prc_array4 nonreact_src,
                        prc_array4 I_R, prc_iarray4 mask, prc_array4 rot_energy, prc_array4 cost,
                        int ufs, int react_init, double* dt_react, double* time, void* stream) {
  int e = ensure_init(r);
  if (e) return e;
  if (!lo || !hi || !dt_react || !time || ufs < UTEMP + 1 || !(*dt_react > 0.0)) return PRC_EINVAL;
  if (!S_new.p || !I_R.p || !nonreact_src.p || (!react_init && !S_old.p)) return PRC_EINVAL;
  if ((S_new.ncomp && S_new.ncomp < ufs + NSP) || (I_R.ncomp && I_R.ncomp < NSP + 2))
    return PRC_EINVAL;
  const int len0 = hi[0] - lo[0] + 1, len1 = hi[1] - lo[1] + 1, len2 = hi[2] - lo[2] + 1;
  if (len0 <= 0 || len1 <= 0 || len2 <= 0) {
    *time += *dt_react;
    return PRC_OK;
  }
  const int64_t nc = (int64_t)len0 * len1 * len2;
  hipStream_t s = (hipStream_t)stream;
  if (nc > r->sws_cap) {
    if (r->sws) (void)hipFree(r->sws);
    r->sws = nullptr;
    r->sws_cap = 0;
    if (hipMalloc(&r->sws, sizeof(double) * (NSP + 3) * nc) != hipSuccess) return PRC_ENOMEM;
    r->sws_cap = nc;
  }
  StateArgs sa{};
  sa.lo0 = lo[0];
  sa.lo1 = lo[1];
  sa.lo2 = lo[2];
  sa.len0 = len0;
  sa.len1 = len1;
  sa.ncells = nc;
  sa.sold = mk(react_init ? S_new : S_old);
  sa.snew = mk(S_new);
  sa.nonrs = mk(nonreact_src);
  sa.IR = mk(I_R);
  sa.rot = mk(rot_energy);
  sa.mask = mki(mask);
  sa.ufs = ufs;
  sa.react_init = react_init ? 1 : 0;
  sa.dt = *dt_react;
  sa.rY = r->sws;
  sa.T = r->sws + NSP * nc;
  sa.rE = r->sws + (NSP + 1) * nc;
  sa.rEsrc = r->sws + (NSP + 2) * nc;
  const unsigned nb = (unsigned)((nc + 255) / 256);
  hipLaunchKernelGGL(k_state_pack, dim3(nb), dim3(256), 0, s, sa);
  HIP_OK(hipGetLastError());
  // the reactor integrates the packed box; species sources are read in place from S_nr
  auto box_view = [&](double* p, int ncomp) {
    prc_array4 v{};
    v.p = p;
    v.jstride = len0;
    v.kstride = (int64_t)len0 * len1;
    v.nstride = nc;
    for (int d = 0; d < 3; ++d) {
      v.begin[d] = lo[d];
      v.end[d] = hi[d] + 1;
    }
    v.ncomp = ncomp;
    return v;
  };
  prc_array4 src = nonreact_src;
  src.p = nonreact_src.p + (int64_t)ufs * nonreact_src.nstride;
  src.ncomp = NSP;
  Args a{};
  a.box = make_box(lo, hi, box_view(sa.rY, NSP), src, box_view(sa.T, 1), box_view(sa.rE, 1),
                   box_view(sa.rEsrc, 1), cost, mask, 0);
  a.boxes = nullptr;
  a.nboxes = 1;
  a.ncells = nc;
  a.dt = *dt_react;
  if (r->integrator != 2) {
    e = launch(r, a, true, s);
    if (e) return e;
  }
  hipLaunchKernelGGL(k_state_unpack, dim3(nb), dim3(256), 0, s, sa);
  HIP_OK(hipGetLastError());
  *time += *dt_react;
  return PRC_OK;
}

int prc_adjoint(prc_reactor* r, int64_t ncells, const double* rY0, const double* T0,
                const double* rE0, const double* rYsrc, const double* rEsrc, double dt,
                const double* logA, const double* ct_rY, const double* ct_T, double* rY_out,
                double* T_out, double* g_rY0, double* g_rE0, double* g_logA, int32_t* status,
                double* tape, int64_t tape_cap, int32_t* tape_count, void* stream) {
  int e = ensure_init(r);
  if (e) return e;
  if (ncells < 0 || !(dt > 0.0) || r->integrator != 0) return PRC_EINVAL;
  if (!ADJ_OK) return PRC_EINVAL;  // n above ADJ_MAX_N: no adjoint kernel for this mechanism
  if (ncells == 0) return PRC_OK;
  if (!rY0 || !T0 || !rE0 || !rY_out || !T_out || !g_rY0 || !g_rE0 || !g_logA || !status)
    return PRC_EINVAL;
  if (tape && (tape_cap < 1 || !tape_count)) return PRC_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  const size_t per_cell =
      (size_t)r->adj_kmax * CKW + (size_t)r->adj_tcap * TRS + (size_t)ADJ_MATS * NV * NV;
  const size_t bytes = sizeof(double) * ((size_t)ncells * per_cell + (size_t)Mech::NRXN);
  if (bytes > r->aws_bytes) {
    if (r->aws) (void)hipFree(r->aws);
    r->aws = nullptr;
    r->aws_bytes = 0;
    if (hipMalloc(&r->aws, bytes) != hipSuccess) return PRC_ENOMEM;
    r->aws_bytes = bytes;
  }
  AdjArgs a{};
  a.ncells = ncells;
  a.rY0 = rY0;
  a.T0 = T0;
  a.rE0 = rE0;
  a.rYsrc = rYsrc;
  a.rEsrc = rEsrc;
  a.dt = dt;
  double* rmul = r->aws;
  a.rmul = rmul;
  a.ct_rY = ct_rY;
  a.ct_T = ct_T;
  a.rY_out = rY_out;
  a.T_out = T_out;
  a.g_rY0 = g_rY0;
  a.g_rE0 = g_rE0;
  a.g_p = g_logA;
  a.status = status;
  a.prm.rtol = r->rtol;
  a.prm.atolv = r->d_atol;
  a.prm.qmax = std::min(std::max(r->max_order, 1), BDF_QMAX);
  a.prm.max_steps = r->max_steps;
  a.seg = r->adj_seg;
  a.kmax = r->adj_kmax;
  a.ckpt = r->aws + Mech::NRXN;
  a.tcap = r->adj_tcap;
  a.tape = a.ckpt + (size_t)ncells * r->adj_kmax * CKW;
  a.mats = a.tape + (size_t)ncells * r->adj_tcap * TRS;
  a.full_cap = tape ? tape_cap : 0;
  a.full_tape = tape;
  a.full_count = tape_count;
  hipLaunchKernelGGL(k_exp, dim3((Mech::NRXN + 63) / 64), dim3(64), 0, s, (int)Mech::NRXN, logA,
                     rmul);
  hipLaunchKernelGGL(k_adjoint<NV>, dim3((unsigned)((ncells + 63) / 64)), dim3(64), 0, s, a);
  HIP_OK(hipGetLastError());
  return PRC_OK;
}

int prc_compute_temp(const int32_t lo[3], const int32_t hi[3], prc_array4 S, prc_iarray4 mask,
                     int ufs, void* stream) {
  if (!lo || !hi || !S.p || ufs < UTEMP + 1 || (S.ncomp && S.ncomp < ufs + NSP)) return PRC_EINVAL;
  const int len0 = hi[0] - lo[0] + 1, len1 = hi[1] - lo[1] + 1, len2 = hi[2] - lo[2] + 1;
  if (len0 <= 0 || len1 <= 0 || len2 <= 0) return PRC_OK;
  const int64_t nc = (int64_t)len0 * len1 * len2;
  hipLaunchKernelGGL(k_cmp_temp, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, lo[0], lo[1], lo[2], len0, len1, nc, mk(S), mki(mask),
                     ufs);
  HIP_OK(hipGetLastError());
  return PRC_OK;
}

int prc_eos_pyt2re(int64_t nc, const double* p, const double* T, const double* Y, double* rho,
                   double* e, void* stream) {
  if (nc < 0 || (nc > 0 && (!p || !T || !Y || !rho || !e))) return PRC_EINVAL;
  return launch_1d(eos_pyt2re, nc, stream, p, T, Y, rho, e);
}
int prc_eos_rey2t(int64_t nc, const double* rho, const double* e, const double* Y, double* T,
                  void* stream) {
  (void)rho;
  if (nc < 0 || (nc > 0 && (!e || !Y || !T))) return PRC_EINVAL;
  return launch_1d(eos_rey2t, nc, stream, e, Y, T);
}
int prc_eos_rty2hi(int64_t nc, const double* T, double* hi, void* stream) {
  if (nc < 0 || (nc > 0 && (!T || !hi))) return PRC_EINVAL;
  return launch_1d(eos_rty2hi, nc, stream, T, hi);
}
int prc_eos_rty2wdot(int64_t nc, const double* rho, const double* T, const double* Y,
                     double* wdot, void* stream) {
  if (nc < 0 || (nc > 0 && (!rho || !T || !Y || !wdot))) return PRC_EINVAL;
  return launch_1d(eos_rty2wdot, nc, stream, rho, T, Y, wdot);
}

int prc_model_eval(int64_t nc, double t, const double* y, const double* rhoe0,
                   const double* rhoesrc, const double* rYsrc, double* ydot, double* J,
                   int cooperative, void* stream) {
  if (nc < 0 || (nc > 0 && (!y || !rhoe0 || !ydot))) return PRC_EINVAL;
  if (nc == 0) return PRC_OK;
  hipStream_t s = (hipStream_t)stream;
  if (cooperative) {
    if constexpr (COOP)
      hipLaunchKernelGGL(k_model_group<64>, dim3((unsigned)((nc + 1) / 2)), dim3(64), 0, s, nc, t,
                         y, rhoe0, rhoesrc, rYsrc, ydot, J);
    else if constexpr (COOP64)
      hipLaunchKernelGGL((k_model_group<64, GroupModel<Mech, 64>>), dim3((unsigned)nc), dim3(64), 0,
                         s, nc, t, y, rhoe0, rhoesrc, rYsrc, ydot, J);
    else
      return PRC_EINVAL;  // this mechanism has no cooperative kernels
  } else
    hipLaunchKernelGGL(k_model_lane, dim3((unsigned)((nc + 63) / 64)), dim3(64), 0, s, nc, t, y,
                       rhoe0, rhoesrc, rYsrc, ydot, J);
  HIP_OK(hipGetLastError());
  return PRC_OK;
}

}  // extern "C"
