// Device kernels of the gfx950 reactor for one mechanism (included by reactor.hip, which defines
// Mech). Forms:
//   plain    react_plain: one cell per lane over a static grid, the whole integration of the cell
//            in one call, working set in private memory -- the straightforward drop-in for
//            ReactorCvode::react (Source/React.cpp:222-228) and the baseline.
//   batched  device work queue of per-cell BDF phases (batched.h): k_start, then iterations of
//            k_jac / k_lu / k_newton over live slots, compaction when half the slots are done.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

#include "batched.h"
#include "bdf_cell.h"
#include "rk64.h"
#include "adjoint.h"

namespace {

using namespace pelec;

constexpr int NSP = Mech::NSP;
constexpr int NV = NSP + 1;
using Model = CellModel<Mech>;

// device view of an Array4
struct A4 {
  double* p;
  int64_t js, ks, ns;
  int b0, b1, b2;
  __device__ double& at(int i, int j, int k, int c) const {
    return p[(i - b0) + (int64_t)(j - b1) * js + (int64_t)(k - b2) * ks + (int64_t)c * ns];
  }
};
struct A4I {
  int32_t* p;
  int64_t js, ks, ns;
  int b0, b1, b2;
  __device__ int32_t at(int i, int j, int k) const {
    return p[(i - b0) + (int64_t)(j - b1) * js + (int64_t)(k - b2) * ks];
  }
};

struct BoxDev {
  int lo0, lo1, lo2, len0, len1;
  int64_t ncells;
  int64_t first;  // global index of the box's first cell (concatenation of the call's boxes)
  A4 rY, rYsrc, T, rE, rEsrc, FC;
  A4I mask;
};

// Everything a kernel needs for one react call, kept in device memory so the captured graph of
// the batched form does not depend on the call.
struct Args {
  const BoxDev* boxes;  // device table (multi-box) or null
  BoxDev box;           // single box
  int nboxes;
  int64_t ncells;       // cells of the call
  double dt;
  BdfParams prm;
  Rk64Params rk;  // ReactorRK64 knobs (integrator 1)
  int32_t* diag;
  int64_t diag_cap;
  unsigned long long* ctr;  // counters (CT_*)
  // batched workspace: C slots and C matrix blocks
  StateBuf st[2];
  int64_t C;
  double* J;       // per matrix block n*n (FP64)
  void* LU;        // per matrix block n*n (LUT)
  int32_t* perm;   // per matrix block n
  int32_t* dexp;   // per matrix block n: equilibration exponents of FP32 factors (scale_exp)
  int32_t* list[NLISTS];
  int32_t* jlist[NLISTS];  // the entries of list[i] whose set-up evaluates a Jacobian
  int32_t* remap;
  int32_t* ring;   // C: free matrix blocks (tail counter CT_FTAIL, head kept by the host)
  int has_src;
  int test_lu_fail;  // prc.test_lu_fail: every factorisation reports a singular matrix (tests)
  // plain form above PLAIN_PRIVATE_N: per cell of the box, its saved Jacobian and Newton matrix
  // (2 n^2 doubles), which would not fit the private segment
  double* pws;
};

__device__ inline void cell_ijk(const BoxDev& b, int64_t c, int& i, int& j, int& k) {
  i = b.lo0 + (int)(c % b.len0);
  const int64_t r = c / b.len0;
  j = b.lo1 + (int)(r % b.len1);
  k = b.lo2 + (int)(r / b.len1);
}

__device__ inline const BoxDev& box_of(const Args& a, int64_t g, int64_t& local) {
  if (!a.boxes) {
    local = g - a.box.first;
    return a.box;
  }
  int lo = 0, hi = a.nboxes - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (a.boxes[mid].first <= g)
      lo = mid;
    else
      hi = mid - 1;
  }
  local = g - a.boxes[lo].first;
  return a.boxes[lo];
}

// Per-cell integration counters reported in the diagnostics buffer (RK64 reports its rejected
// steps as error-test failures and has no Jacobian, set-up or Newton counts).
struct Counters {
  int nst, nfe, nje, nsetups, nni, netf, ncfn;
};
__device__ inline Counters counters(const BdfScalars& S) {
  return Counters{S.nst, S.nfe, S.nje, S.nsetups, S.nni, S.netf, S.ncfn};
}

__device__ inline void write_diag(const Args& a, int64_t g, int status, const Counters* S) {
  if (!a.diag || g >= a.diag_cap) return;
  int32_t* d = a.diag + g;
  const int64_t s = a.diag_cap;
  d[0] = status;
  d[s] = S ? S->nst : 0;
  d[2 * s] = S ? S->nfe : 0;
  d[3 * s] = S ? S->nje : 0;
  d[4 * s] = S ? S->nsetups : 0;
  d[5 * s] = S ? S->nni : 0;
  d[6 * s] = S ? S->netf : 0;
  d[7 * s] = S ? S->ncfn : 0;
}

// Final write-back of a cell: rhoY, T = T(e_end, Y) with the integrated T as guess, and
// rho*e = rho*e0 + dt rho*e_src (Source/React.cpp:292 convention). Failed cells keep their inputs.
template <class ZN>
__device__ inline void store_cell(const Args& a, int64_t g, int status, const Counters& S,
                                  ZN zn0, double rhoe0, double rhoesrc) {
  int64_t loc;
  const BoxDev& b = box_of(a, g, loc);
  int i, j, k;
  cell_ijk(b, loc, i, j, k);
  if (status == CELL_OK) {
    double y[NV], Y[NSP];
    double rho = 0.0;
#pragma unroll
    for (int s = 0; s < NV; ++s) y[s] = zn0(s);
#pragma unroll
    for (int s = 0; s < NSP; ++s) rho += y[s];
    const double irho = 1.0 / rho;
#pragma unroll
    for (int s = 0; s < NSP; ++s) Y[s] = y[s] * irho;
    const double rEn = rhoe0 + a.dt * rhoesrc;
    const double Tn = Model::T_from_e(rEn * irho, Y, y[NSP]);
#pragma unroll
    for (int s = 0; s < NSP; ++s) b.rY.at(i, j, k, s) = y[s];
    b.T.at(i, j, k, 0) = Tn;
    b.rE.at(i, j, k, 0) = rEn;
  }
  if (b.FC.p) b.FC.at(i, j, k, 0) = (double)S.nfe;
  write_diag(a, g, status, &S);
}

// ------------------------------------------------------------------ plain FP64 drop-in
// Up to this ODE size the plain kernel keeps a cell's n x n matrices in private memory; above it
// (the 53-species mechanisms: 2 n^2 doubles = 46 KB at n = 54) the runtime cannot back the
// private segment, and they live in a device workspace instead (Args::pws).
constexpr int PLAIN_PRIVATE_N = 32;
constexpr bool PLAIN_EXT = NV > PLAIN_PRIVATE_N;

// NREF: FP64 refinement steps after each FP32-LU solve (prc.lu_refine; FP64 factors: unused)
template <typename LUT, int NREF>
__global__ void __launch_bounds__(64) react_plain(const Args* __restrict__ ap) {
  const Args& a = *ap;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const BoxDev& b = a.box;
  if (c >= b.ncells) return;
  int i, j, k;
  cell_ijk(b, c, i, j, k);
  const int64_t g = b.first + c;
  if (b.mask.p && b.mask.at(i, j, k) == 0) {
    if (b.FC.p) b.FC.at(i, j, k, 0) = 0.0;
    write_diag(a, g, CELL_OK, nullptr);
    return;
  }
  BdfCell<Mech, LUT, NoTapeRec, false, false, PLAIN_EXT, NREF> cell;
  if constexpr (PLAIN_EXT) {
    double* w = a.pws + c * (int64_t)(2 * NV * NV);
    cell.savedJ = w;
    cell.M = reinterpret_cast<LUT*>(w + NV * NV);
    cell.Jx = nullptr;
  }
  double y0[NV];
  double rho = 0.0;
#pragma unroll
  for (int s = 0; s < NSP; ++s) {
    y0[s] = b.rY.at(i, j, k, s);
    rho += y0[s];
  }
  y0[NSP] = b.T.at(i, j, k, 0);
  cell.rhoe0 = b.rE.at(i, j, k, 0);
  cell.rhoesrc = b.rEsrc.p ? b.rEsrc.at(i, j, k, 0) : 0.0;
  cell.src = b.rYsrc.p ? &b.rYsrc.at(i, j, k, 0) : nullptr;
  cell.sstride = b.rYsrc.ns;
  cell.prm = a.prm;
  if (!(rho > 0.0) || !(y0[NSP] > 0.0) || !(a.dt > 0.0)) {
    atomicAdd(&a.ctr[CT_FAILED], 1ull);
    if (b.FC.p) b.FC.at(i, j, k, 0) = 0.0;
    write_diag(a, g, CELL_BADINPUT, nullptr);
    return;
  }
  cell.start(y0, a.dt);
  int s;
  do {
    s = cell.advance();
  } while (s < 0);
  if (s != CELL_OK) atomicAdd(&a.ctr[CT_FAILED], 1ull);
  store_cell(a, g, s, counters(cell.S), [&](int e) { return cell.zn[0][e]; }, cell.rhoe0, cell.rhoesrc);
}

// ------------------------------------------------------------------ ReactorRK64 drop-in
// One cell per lane over a static grid of the box, explicit RK64 with error control (rk64.h).
__global__ void __launch_bounds__(64) react_rk64(const Args* __restrict__ ap) {
  const Args& a = *ap;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const BoxDev& b = a.box;
  if (c >= b.ncells) return;
  int i, j, k;
  cell_ijk(b, c, i, j, k);
  const int64_t g = b.first + c;
  if (b.mask.p && b.mask.at(i, j, k) == 0) {
    if (b.FC.p) b.FC.at(i, j, k, 0) = 0.0;
    write_diag(a, g, CELL_OK, nullptr);
    return;
  }
  Rk64Cell<Mech> cell;
  double y[NV];
  double rho = 0.0;
#pragma unroll
  for (int s = 0; s < NSP; ++s) {
    y[s] = b.rY.at(i, j, k, s);
    rho += y[s];
  }
  y[NSP] = b.T.at(i, j, k, 0);
  cell.rhoe0 = b.rE.at(i, j, k, 0);
  cell.rhoesrc = b.rEsrc.p ? b.rEsrc.at(i, j, k, 0) : 0.0;
  cell.src = b.rYsrc.p ? &b.rYsrc.at(i, j, k, 0) : nullptr;
  cell.sstride = b.rYsrc.ns;
  if (!(rho > 0.0) || !(y[NSP] > 0.0) || !(a.dt > 0.0)) {
    atomicAdd(&a.ctr[CT_FAILED], 1ull);
    if (b.FC.p) b.FC.at(i, j, k, 0) = 0.0;
    write_diag(a, g, CELL_BADINPUT, nullptr);
    return;
  }
  const int st = cell.integrate(y, a.dt, a.rk);
  if (st != CELL_OK) atomicAdd(&a.ctr[CT_FAILED], 1ull);
  const Counters cn{cell.nst, cell.nfe, 0, 0, 0, cell.nrej, 0};
  store_cell(a, g, st, cn, [&](int e) { return y[e]; }, cell.rhoe0, cell.rhoesrc);
}

// ------------------------------------------------------------------ batched work-queue form
__device__ inline SoAV<NV> soa(const StateBuf& st, int64_t C, int64_t s) {
  return SoAV<NV>{{}, {}, st.zn, st.acor, st.ewt, C, s};
}

__device__ inline void rhs_slot(const Args& a, const StateBuf& st, int64_t s, double t,
                                const double* y, double* out) {
  const double* src = a.has_src ? st.src + s : nullptr;
  Model::rhs(t, y, st.rhoe[s], st.rhoe[a.C + s], src, a.C, out);
}

// Cells g0 .. g0+count-1 into slots slot0 .. slot0+count-1 of the current state buffer:
// inputs -> state, f(t0, y0), cvHin, first attempt (always a set-up: nst == 0). Matrix block of
// slot slot0+t: the slot itself when head < 0 (the first fill), else free-ring entry head+t. A
// cell that does not start (masked, bad input) returns its block at once.
__global__ void __launch_bounds__(64) k_start(const Args* __restrict__ ap, int64_t slot0,
                                              int64_t count, int64_t g0, int64_t head) {
  const Args& a = *ap;
  const int64_t t = (int64_t)blockIdx.x * 64 + threadIdx.x;
  const StateBuf& st = a.st[a.ctr[CT_CUR]];
  if (blockIdx.x == 0 && threadIdx.x == 0) a.ctr[CT_NACTIVE] = (unsigned long long)(slot0 + count);
  bool started = false, unused = false;
  const int64_t g = g0 + t;
  const int64_t slot = slot0 + t;
  int32_t mblk = 0;
  if (t < count) {
    mblk = head < 0 ? (int32_t)slot : a.ring[(head + t) % a.C];
    int64_t loc;
    const BoxDev& b = box_of(a, g, loc);
    int i, j, k;
    cell_ijk(b, loc, i, j, k);
    st.cell[slot] = (int32_t)g;
    st.phase[slot] = PH_DONE;
    st.mb[slot] = mblk;
    unused = true;
    if (b.mask.p && b.mask.at(i, j, k) == 0) {
      if (b.FC.p) b.FC.at(i, j, k, 0) = 0.0;
      write_diag(a, g, CELL_OK, nullptr);
    } else {
      double y0[NV];
      double rho = 0.0;
#pragma unroll
      for (int s = 0; s < NSP; ++s) {
        y0[s] = b.rY.at(i, j, k, s);
        rho += y0[s];
      }
      y0[NSP] = b.T.at(i, j, k, 0);
      if (!(rho > 0.0) || !(y0[NSP] > 0.0) || !(a.dt > 0.0)) {
        atomicAdd(&a.ctr[CT_FAILED], 1ull);
        if (b.FC.p) b.FC.at(i, j, k, 0) = 0.0;
        write_diag(a, g, CELL_BADINPUT, nullptr);
      } else {
        started = true;
        unused = false;
        st.rhoe[slot] = b.rE.at(i, j, k, 0);
        st.rhoe[a.C + slot] = b.rEsrc.p ? b.rEsrc.at(i, j, k, 0) : 0.0;
        if (a.has_src) {
#pragma unroll
          for (int s = 0; s < NSP; ++s)
            st.src[s * a.C + slot] = b.rYsrc.p ? b.rYsrc.at(i, j, k, s) : 0.0;
        }
        const SoAV<NV> v = soa(st, a.C, slot);
        BdfScalars S;
        scalars_reset(S, a.prm.qmax);
#pragma unroll
        for (int e = 0; e < NV; ++e) v.Z(0, e) = y0[e];
        set_ewt<NV>(v, a.prm);
        auto f = [&](double tt, const double* yy, double* out) {
          S.nfe++;
          rhs_slot(a, st, slot, tt, yy, out);
        };
        double f0[NV];
        f(S.tn, y0, f0);
#pragma unroll
        for (int e = 0; e < NV; ++e) v.Z(1, e) = f0[e];
        S.h = hin<NV>(S, v, a.dt, f);
        if ((S.tn + S.h - a.dt) * S.h > 0.0) S.h = (a.dt - S.tn) * (1.0 - 4.0 * cv::UROUND);
        S.hscale = S.h;
        S.hprime = S.h;
#pragma unroll
        for (int e = 0; e < NV; ++e) v.Z(1, e) *= S.h;
        step_prologue<NV>(S, v, a.prm.qmax);
        begin_attempt<NV>(S, v, a.dt);  // nst == 0: set-up required
        jac_decide(S);                  // nst == 0: with a Jacobian
        store_scalars(S, st.sc, a.C, slot);
        st.phase[slot] = PH_SETUP;
      }
    }
  }
  wave_count(&a.ctr[CT_STARTED], started);
  wave_append(a.list[0], &a.ctr[CT_SETUP0], started, (int32_t)slot);
  wave_append(a.jlist[0], &a.ctr[CT_JAC0], started, (int32_t)slot);
  wave_push_ring(a.ring, &a.ctr[CT_FTAIL], a.C, unused, mblk);
}

// The tn word of a slot's integrator scalars.
__device__ inline double slot_tn(const StateBuf& st, int64_t C, int64_t s) {
  constexpr int W_TN = offsetof(BdfScalars, tn) / 8;
  double tn;
  const uint64_t w = st.sc[W_TN * C + s];
  __builtin_memcpy(&tn, &w, 8);
  return tn;
}

// Jacobians of the set-up list entries that need one: the Jacobian list, which the kernel that
// requested the set-up filled after cvDlsSetup's decision (jac_decide). A template so that
// mechanisms whose Jacobian runs in the cooperative kernel do not build it.
template <int UNUSED = 0>
__global__ void __launch_bounds__(64) k_jac(const Args* __restrict__ ap, int par) {
  const Args& a = *ap;
  const int64_t cnt = (int64_t)a.ctr[CT_JAC0 + par];
  const StateBuf& st = a.st[a.ctr[CT_CUR]];
  for (int64_t base = (int64_t)blockIdx.x * 64; base < cnt; base += (int64_t)gridDim.x * 64) {
    const int64_t idx = base + threadIdx.x;
    if (idx >= cnt) continue;
    const int64_t s = a.jlist[par][idx];
    const int64_t mb = st.mb[s];
    const SoAV<NV> v = soa(st, a.C, s);
    double y[NV];
#pragma unroll
    for (int e = 0; e < NV; ++e) y[e] = v.Z(0, e);
    const double* src = a.has_src ? st.src + s : nullptr;
    Model::template jac<double>(slot_tn(st, a.C, s), y, st.rhoe[s], st.rhoe[a.C + s], src, a.C,
                                a.J + mb * (int64_t)(NV * NV));
  }
}

// Newton matrices of the set-up list factorised by G-lane groups, one matrix per group and one
// row per lane in registers (lu_rows): two matrices per wave with G = 32 (n <= 32), one with
// G = 64 (n <= 64, the 53-species mechanisms). The per-cell bookkeeping is done by the group's
// first lane.
template <typename LUT, int G>
__global__ void __launch_bounds__(64) k_lu_rows(const Args* __restrict__ ap, int par) {
  static_assert(NV <= G && (G == 32 || G == 64), "k_lu_rows: one row per lane of a 32/64 group");
  constexpr int MPW = 64 / G;  // matrices per wave
  const Args& a = *ap;
  const int64_t cnt = (int64_t)a.ctr[CT_SETUP0 + par];
  const StateBuf& st = a.st[a.ctr[CT_CUR]];
  const int lane = threadIdx.x;
  const int r = lane & (G - 1);
  const bool leader = r == 0;
  for (int64_t base = (int64_t)blockIdx.x * MPW; base < cnt; base += (int64_t)gridDim.x * MPW) {
    const int64_t idx = base + lane / G;
    const bool valid = idx < cnt;
    const int64_t s = valid ? a.list[par][idx] : 0;
    const int64_t g = valid ? st.mb[s] : 0;  // the cell's matrix block
    constexpr int W_GAMMA = offsetof(BdfScalars, gamma) / 8;
    double gamma = 0.0;
    if (valid) {
      const uint64_t w = st.sc[W_GAMMA * a.C + s];
      __builtin_memcpy(&gamma, &w, 8);
    }
    LUT row[NV];
    constexpr bool EQ = sizeof(LUT) == sizeof(float);  // FP32 factors are equilibrated
    int kr = 0;
    if (valid && r < NV) {
      const double* Jr = a.J + g * (int64_t)(NV * NV) + r * NV;
      if constexpr (EQ) {
        // D (I - gamma J) D^-1 with d_c = 2^k_c the power of two of the cell's error weight ewt_c
        const double* ew = st.ewt + s;
        kr = scale_exp(ew[(int64_t)r * a.C]);
#pragma unroll
        for (int c = 0; c < NV; ++c) {
          const double m = (c == r ? 1.0 : 0.0) - gamma * Jr[c];
          row[c] = LUT(ldexp(m, kr - scale_exp(ew[(int64_t)c * a.C])));
        }
      } else {
#pragma unroll
        for (int c = 0; c < NV; ++c) {
          const LUT m = LUT(-gamma * Jr[c]);
          row[c] = (c == r) ? m + LUT(1) : m;  // I - gamma J
        }
      }
    } else {
#pragma unroll
      for (int c = 0; c < NV; ++c) row[c] = LUT(0);
    }
    int pos = r;
    int ier = 0;
    if (valid) ier = lu_rows<NV, LUT, G>(row, r, pos);
    if (a.test_lu_fail && ier == 0) ier = 1;
    if (valid && r < NV && ier == 0) {
      LUT* LUg = reinterpret_cast<LUT*>(a.LU) + g * (int64_t)(NV * NV) + pos * NV;
#pragma unroll
      for (int c = 0; c < NV; ++c) LUg[c] = row[c];
      a.perm[g * NV + pos] = r;
      if constexpr (EQ) a.dexp[g * NV + r] = kr;
    }
    if (valid && leader) {
      after_setup_slot(st.sc, a.C, s);
      // Newton from the next iteration on; the Newton kernel handles a singular matrix there
      st.phase[s] = (ier == 0 ? PH_READY0 : PH_LUFAIL0) + par;
    }
  }
}

// b <- M^{-1} b with the factors of matrix block mb (one lane's cell): FP64 factors, or FP32
// factors of the equilibrated D M D^-1 followed by NREF FP64 refinement steps against the FP64
// Jacobian (residual b - (x - gammap J x)). gbl: the lane's LDS column (stride 64).
template <typename LUT, int NREF>
__device__ inline void newton_solve(const Args& a, int64_t mb, double gammap, double* b,
                                    double* gbl) {
  const LUT* lu = reinterpret_cast<const LUT*>(a.LU) + mb * (int64_t)(NV * NV);
  const int32_t* pm = a.perm + mb * NV;
  if constexpr (sizeof(LUT) == sizeof(double)) {
    lu_solve_perm<NV, LUT>(lu, pm, b, gbl);
  } else if constexpr (NREF == 0) {
    lu_solve_perm<NV, LUT>(lu, pm, b, gbl, a.dexp + mb * NV);
  } else {
    const double* Jg = a.J + mb * (int64_t)(NV * NV);
    const int32_t* dexp = a.dexp + mb * NV;
    double x[NV], res[NV];
#pragma unroll
    for (int e = 0; e < NV; ++e) x[e] = b[e];
    lu_solve_perm<NV, LUT>(lu, pm, x, gbl, dexp);
#pragma unroll
    for (int it = 0; it < NREF; ++it) {
#pragma unroll
      for (int e = 0; e < NV; ++e) {
        double jx = 0.0;
#pragma unroll
        for (int f2 = 0; f2 < NV; ++f2) jx += Jg[e * NV + f2] * x[f2];
        res[e] = b[e] - (x[e] - gammap * jx);
        PELEC_MEMORY_FENCE();
      }
      lu_solve_perm<NV, LUT>(lu, pm, res, gbl, dexp);
#pragma unroll
      for (int e = 0; e < NV; ++e) x[e] += res[e];
    }
#pragma unroll
    for (int e = 0; e < NV; ++e) b[e] = x[e];
  }
}

// One Newton correction per Newton-phase slot; on convergence the error test, step completion,
// order/step choice and the next attempt's prediction.
// NREF: FP64 refinement steps after each FP32-LU solve (FP64 factors: none)
template <typename LUT, int NREF>
__global__ void __launch_bounds__(64) k_newton(const Args* __restrict__ ap, int par) {
  __shared__ double gb[NV * 64];
  // each lane's integrator scalars in LDS rather than registers: k_newton runs at the register
  // limit, and the record is touched a few times per correction
  __shared__ BdfScalars scl[64];
  const Args& a = *ap;
  // list (par+2)%3 was consumed by the previous iteration and is first appended to by the next
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.ctr[CT_SETUP0 + list_prev(par)] = 0;
    a.ctr[CT_JAC0 + list_prev(par)] = 0;
  }
  const StateBuf& st = a.st[a.ctr[CT_CUR]];
  const int64_t nact = (int64_t)a.ctr[CT_NACTIVE];
  const int lane = threadIdx.x;
  constexpr int W_TN = offsetof(BdfScalars, tn) / 8;
  constexpr int W_M = offsetof(BdfScalars, m) / 8;
  constexpr int SH_M = (offsetof(BdfScalars, m) % 8) * 8;
  for (int64_t base = (int64_t)blockIdx.x * 64; base < nact; base += (int64_t)gridDim.x * 64) {
    const int64_t s = base + lane;
    const int ph = s < nact ? st.phase[s] : PH_DONE;
    const bool act = ph == PH_NEWTON || ph == PH_READY0 + list_prev(par);
    // singular Newton matrix in the previous iteration's set-up: CVODE's convergence-failure
    // handling, as for a Newton failure with a current Jacobian (r = 3 below)
    const bool lufail = ph == PH_LUFAIL0 + list_prev(par);
    bool need_setup = false, need_jac = false, done = false;
    if (act || lufail) {
      const SoAV<NV> v = soa(st, a.C, s);
      const int64_t g = st.cell[s];
      const int64_t mb = st.mb[s];
      BdfScalars& S = scl[lane];
      int r = 3;
      if (lufail) {
        load_scalars(S, st.sc, a.C, s);
      } else {
        // f(tn, zn[0] + acor), loading only tn and m first (keeps the RHS register budget)
        double tn;
        {
          const uint64_t w = st.sc[W_TN * a.C + s];
          __builtin_memcpy(&tn, &w, 8);
        }
        const int m0 = (int)(uint32_t)(st.sc[W_M * a.C + s] >> SH_M);
        double ftemp[NV];
        {
          double y[NV];
#pragma unroll
          for (int e = 0; e < NV; ++e) y[e] = v.Z(0, e) + (m0 > 0 ? v.A(e) : 0.0);
          rhs_slot(a, st, s, tn, y, ftemp);
        }
        load_scalars(S, st.sc, a.C, s);
        S.nfe++;
        if (S.m == 0) {
#pragma unroll
          for (int e = 0; e < NV; ++e) v.A(e) = 0.0;
        }
        const double gammap = S.gammap;
        double* gbl = gb + lane;
        auto solve = [&](double* b) { newton_solve<LUT, NREF>(a, mb, gammap, b, gbl); };
        r = newton_correct<NV>(S, v, ftemp, solve);
      }
      int status = -1;  // <0 running, else final status
      if (r == 2) {
        S.convfail = 2;  // stale Jacobian: rebuild and restart the Newton iteration
        S.m = 0;
        need_setup = true;
      } else if (r != 1) {
        bool attempt = false;
        if (r == 3) {
          const int e = conv_fail<NV>(S, v);
          if (e >= 0)
            status = e;
          else
            attempt = true;
        } else {
          const double dsm = S.acnrm * S.tq[2];
          if (dsm <= 1.0) {
            finish_step<NV>(S, v, dsm, a.prm.qmax);
            if (reached_tstop(S, a.dt)) {
              status = CELL_OK;
            } else if (S.nst >= a.prm.max_steps) {
              status = CELL_MAXSTEPS;
            } else {
              need_setup = next_attempt<NV>(S, v, a.prm, a.dt);
            }
          } else {
            const int e = error_test_fail<NV>(S, v, dsm, a.prm.qmax);
            if (e >= 0) {
              status = e;
            } else {
              if (e == -2) {
                double y[NV], t[NV];
#pragma unroll
                for (int q = 0; q < NV; ++q) y[q] = v.Z(0, q);
                S.nfe++;
                rhs_slot(a, st, s, S.tn, y, t);
#pragma unroll
                for (int q = 0; q < NV; ++q) v.Z(1, q) = S.h * t[q];
              }
              attempt = true;
            }
          }
        }
        if (attempt) need_setup = begin_attempt<NV>(S, v, a.dt);
      }
      if (status >= 0) {
        done = true;
        st.phase[s] = PH_DONE;
        if (status != CELL_OK) atomicAdd(&a.ctr[CT_FAILED], 1ull);
        store_cell(a, g, status, counters(S), [&](int e) { return v.Z(0, e); }, st.rhoe[s],
                   st.rhoe[a.C + s]);
      } else if (need_setup) {
        st.phase[s] = PH_SETUP;
        need_jac = jac_decide(S);
      } else if (ph != PH_NEWTON) {
        st.phase[s] = PH_NEWTON;
      }
      store_scalars(S, st.sc, a.C, s);
    }
    wave_append(a.list[list_next(par)], &a.ctr[CT_SETUP0 + list_next(par)], need_setup,
                (int32_t)s);
    wave_append(a.jlist[list_next(par)], &a.ctr[CT_JAC0 + list_next(par)], need_jac, (int32_t)s);
    wave_count(&a.ctr[CT_DONE], done);
    wave_push_ring(a.ring, &a.ctr[CT_FTAIL], a.C, done, done ? st.mb[s] : 0);
  }
}

// ------------------------------------------------------------------ cooperative kernels
// One cell per G-lane group: lane k holds state component k, the species and reaction work of the
// RHS and Jacobian is spread over the group (group_model.h), and the CVODE step logic runs on the
// group accessor (component k on lane k, the integrator scalars replicated, norms as group
// reductions). Same phases, lists, slot storage and counters as the one-cell-per-lane kernels,
// which they replace one for one.
//   n <= 32: G = 32, two cells per wave, chosen by prc.cooperative / the tail graph;
//   32 < n <= 64 (the 53-species mechanisms): G = 64, one cell per wave, always -- their
//   one-cell-per-lane Newton and Jacobian kernels would hold 54-component vectors and factor rows
//   per lane (registers spill to scratch, and the unrolled n^2 solves took the compiler tens of
//   minutes), so those are not built.
constexpr int GRP = 32;
using GModel = GroupModel<Mech, GRP>;
// The cooperative kernels exist for mechanisms whose data tables were generated and whose state
// fits one lane per component of a 32- or 64-lane group; for others only the one-cell-per-lane
// kernels are built.
constexpr bool COOP = Mech::T_OK && NV <= GRP;
constexpr bool COOP64 = Mech::T_OK && NV > GRP && NV <= 64;

// store_cell for a group: lane k < N writes rhoY_k, T(e_end, Y) by the group's T(e).
template <class GM = GModel>
__device__ inline void store_cell_g(const Args& a, int64_t g, int status, const Counters& S,
                                    double yl, int l, double rhoe0, double rhoesrc, double* sh) {
  int64_t loc;
  const BoxDev& b = box_of(a, g, loc);
  int i, j, k;
  cell_ijk(b, loc, i, j, k);
  if (status == CELL_OK) {
    const double rho = GM::Grp::sum(l < NSP ? yl : 0.0);
    const double irho = 1.0 / rho;
    const double rEn = rhoe0 + a.dt * rhoesrc;
    const double Tg = GM::Grp::bcast(yl, NSP);
    const double Tn = GM::T_from_e(rEn * irho, l < NSP ? yl * irho : 0.0, l, Tg, sh);
    if (l < NSP) b.rY.at(i, j, k, l) = yl;
    if (l == 0) {
      b.T.at(i, j, k, 0) = Tn;
      b.rE.at(i, j, k, 0) = rEn;
    }
  }
  if (l == 0) {
    if (b.FC.p) b.FC.at(i, j, k, 0) = (double)S.nfe;
    write_diag(a, g, status, &S);
  }
}

// Jacobians of the set-up list entries that need one, one cell per group (k_jac's work).
template <int BS, class GM = GModel>
__global__ void __launch_bounds__(BS) k_jac_g(const Args* __restrict__ ap, int par) {
  constexpr int G = GM::GS;
  constexpr int CPB = BS / G;
  __shared__ double lds[CPB * GM::JAC_WORDS];
  const Args& a = *ap;
  const int64_t cnt = (int64_t)a.ctr[CT_JAC0 + par];
  const StateBuf& st = a.st[a.ctr[CT_CUR]];
  const int l = threadIdx.x & (G - 1);
  double* sh = lds + (threadIdx.x / G) * GM::JAC_WORDS;
  for (int64_t base = (int64_t)blockIdx.x * CPB; base < cnt; base += (int64_t)gridDim.x * CPB) {
    const int64_t idx = base + threadIdx.x / G;
    if (idx >= cnt) continue;
    const int64_t s = a.jlist[par][idx];
    const int64_t mb = st.mb[s];
    const double yl = l < NV ? st.zn[(int64_t)l * a.C + s] : 0.0;
    const double srcl = (a.has_src && l < NSP) ? st.src[(int64_t)l * a.C + s] : 0.0;
    GM::jac(slot_tn(st, a.C, s), yl, l, st.rhoe[s], st.rhoe[a.C + s], srcl, sh,
            a.J + mb * (int64_t)(NV * NV));
  }
}

// One Newton correction per Newton-phase slot, one cell per group (k_newton's work).
template <typename LUT, int BS, int NREF, class GM = GModel>
__global__ void __launch_bounds__(BS) k_newton_g(const Args* __restrict__ ap, int par) {
  constexpr int G = GM::GS;
  constexpr int CPB = BS / G;
  __shared__ double lds[CPB * GM::RHS_WORDS];
  // the integrator scalars of each cell live in LDS, one copy per group (in registers they would
  // be replicated on every lane of the group and more than halve the occupancy)
  __shared__ BdfScalars scl[CPB];
  const Args& a = *ap;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.ctr[CT_SETUP0 + list_prev(par)] = 0;
    a.ctr[CT_JAC0 + list_prev(par)] = 0;
  }
  const StateBuf& st = a.st[a.ctr[CT_CUR]];
  const int64_t nact = (int64_t)a.ctr[CT_NACTIVE];
  const int l = threadIdx.x & (G - 1);
  double* sh = lds + (threadIdx.x / G) * GM::RHS_WORDS;
  for (int64_t base = (int64_t)blockIdx.x * CPB; base < nact; base += (int64_t)gridDim.x * CPB) {
    const int64_t s = base + threadIdx.x / G;
    const int ph = s < nact ? st.phase[s] : PH_DONE;
    const bool act = ph == PH_NEWTON || ph == PH_READY0 + list_prev(par);
    const bool lufail = ph == PH_LUFAIL0 + list_prev(par);  // as in k_newton
    bool need_setup = false, need_jac = false, done = false;
    if (act || lufail) {
      const GrpV<NV, G> v{{}, st.zn, st.acor, st.ewt, a.C, s, l};
      const int64_t g = st.cell[s];
      const int64_t mb = st.mb[s];
      const double rhoe0 = st.rhoe[s], rhoesrc = st.rhoe[a.C + s];
      const double srcl = (a.has_src && l < NSP) ? st.src[(int64_t)l * a.C + s] : 0.0;
      BdfScalars& S = scl[threadIdx.x / G];
      int r = 3;
      if (lufail) {
        load_scalars_group<G>(S, st.sc, a.C, s, l);
      } else {
        constexpr int W_TN = offsetof(BdfScalars, tn) / 8;
        constexpr int W_M = offsetof(BdfScalars, m) / 8;
        constexpr int SH_M = (offsetof(BdfScalars, m) % 8) * 8;
        LaneVal ftemp;
        {
          // f(tn, zn[0] + acor) before the integrator scalars are loaded (register budget)
          double tn;
          const uint64_t w = st.sc[W_TN * a.C + s];
          __builtin_memcpy(&tn, &w, 8);
          const int m0 = (int)(uint32_t)(st.sc[W_M * a.C + s] >> SH_M);
          const double yl = l < NV ? v.Z(0, l) + (m0 > 0 ? v.A(l) : 0.0) : 0.0;
          ftemp.x = GM::rhs(tn, yl, l, rhoe0, rhoesrc, srcl, sh);
        }
        load_scalars_group<G>(S, st.sc, a.C, s, l);
        S.nfe++;
        if (S.m == 0 && l < NV) v.A(l) = 0.0;
        const double gammap = S.gammap;
        auto solve = [&](LaneVal& b) {
          LUT row[NV];
          int pl = l;
          int el = 0;  // this lane's equilibration exponent (FP32 factors)
          if (l < NV) {
            const LUT* lr = reinterpret_cast<const LUT*>(a.LU) + mb * (int64_t)(NV * NV) + l * NV;
#pragma unroll
            for (int c = 0; c < NV; ++c) row[c] = lr[c];
            pl = a.perm[mb * NV + l];
            if constexpr (sizeof(LUT) != sizeof(double)) el = a.dexp[mb * NV + l];
          } else {
#pragma unroll
            for (int c = 0; c < NV; ++c) row[c] = LUT(0);
          }
          if constexpr (sizeof(LUT) == sizeof(double)) {
            b.x = lu_solve_group<NV, LUT, G>(row, pl, b.x, l);
          } else {
            // FP32 factors of D M D^-1, then NREF FP64 refinement steps against the FP64
            // Jacobian (k_newton's)
            double x = ldexp(lu_solve_group<NV, LUT, G>(row, pl, ldexp(b.x, el), l), -el);
            const double* Jr = a.J + mb * (int64_t)(NV * NV) + (l < NV ? l : 0) * NV;
#pragma unroll
            for (int it = 0; it < NREF; ++it) {
              double jx = 0.0;
#pragma unroll
              for (int c = 0; c < NV; ++c) {
                const double xc = GM::Grp::bcast(x, c);
                if (l < NV) jx += Jr[c] * xc;
              }
              const double rr = l < NV ? b.x - (x - gammap * jx) : 0.0;
              x += ldexp(lu_solve_group<NV, LUT, G>(row, pl, ldexp(rr, el), l), -el);
            }
            b.x = x;
          }
        };
        r = newton_correct<NV>(S, v, ftemp, solve);
      }
      int status = -1;
      if (r == 2) {
        S.convfail = 2;
        S.m = 0;
        need_setup = true;
      } else if (r != 1) {
        bool attempt = false;
        if (r == 3) {
          const int e = conv_fail<NV>(S, v);
          if (e >= 0)
            status = e;
          else
            attempt = true;
        } else {
          const double dsm = S.acnrm * S.tq[2];
          if (dsm <= 1.0) {
            finish_step<NV>(S, v, dsm, a.prm.qmax);
            if (reached_tstop(S, a.dt)) {
              status = CELL_OK;
            } else if (S.nst >= a.prm.max_steps) {
              status = CELL_MAXSTEPS;
            } else {
              need_setup = next_attempt<NV>(S, v, a.prm, a.dt);
            }
          } else {
            const int e = error_test_fail<NV>(S, v, dsm, a.prm.qmax);
            if (e >= 0) {
              status = e;
            } else {
              if (e == -2) {
                const double yl = l < NV ? v.Z(0, l) : 0.0;
                S.nfe++;
                const double fl = GM::rhs(S.tn, yl, l, rhoe0, rhoesrc, srcl, sh);
                if (l < NV) v.Z(1, l) = S.h * fl;
              }
              attempt = true;
            }
          }
        }
        if (attempt) need_setup = begin_attempt<NV>(S, v, a.dt);
      }
      if (status >= 0) {
        done = true;
        if (l == 0) {
          st.phase[s] = PH_DONE;
          if (status != CELL_OK) atomicAdd(&a.ctr[CT_FAILED], 1ull);
        }
        store_cell_g<GM>(a, g, status, counters(S), l < NV ? v.Z(0, l) : 0.0, l, rhoe0, rhoesrc, sh);
      } else {
        if (need_setup) need_jac = jac_decide(S);  // every lane, on the group's one record
        if (l == 0) {
          if (need_setup)
            st.phase[s] = PH_SETUP;
          else if (ph != PH_NEWTON)
            st.phase[s] = PH_NEWTON;
        }
      }
      store_scalars_group<G>(S, st.sc, a.C, s, l);
    }
    wave_append(a.list[list_next(par)], &a.ctr[CT_SETUP0 + list_next(par)], need_setup && l == 0,
                (int32_t)s);
    wave_append(a.jlist[list_next(par)], &a.ctr[CT_JAC0 + list_next(par)], need_jac && l == 0,
                (int32_t)s);
    wave_count(&a.ctr[CT_DONE], done && l == 0);
    wave_push_ring(a.ring, &a.ctr[CT_FTAIL], a.C, done && l == 0, done ? st.mb[s] : 0);
  }
}

// The compaction check a graph makes between two iterations (the host makes the same check between
// replays): compact when fewer than pct % of the slots in use are alive, more than 4096 slots are
// in use, and no cell of the call waits for a slot (refills stay with the host).
__global__ void k_compact_gate(const Args* __restrict__ ap, int pct) {
  const Args& a = *ap;
  if (threadIdx.x == 0) {
    const long long alive = (long long)a.ctr[CT_STARTED] - (long long)a.ctr[CT_DONE];
    const long long nact = (long long)a.ctr[CT_NACTIVE];
    a.ctr[CT_GATE] = (a.ctr[CT_INGRAPH] && nact > 4096 && alive * 100 < nact * pct) ? 1ull : 0ull;
  }
}

// Repack live slots (phase != DONE) densely into the other state buffer, lane order preserved.
// gated: only when the graph's compaction check decided so (k_compact_gate).
__global__ void __launch_bounds__(64) k_compact(const Args* __restrict__ ap, int gated) {
  const Args& a = *ap;
  if (gated && !a.ctr[CT_GATE]) return;
  const int cur = (int)a.ctr[CT_CUR];
  const StateBuf& src = a.st[cur];
  const StateBuf& dst = a.st[cur ^ 1];
  const int64_t nact = (int64_t)a.ctr[CT_NACTIVE];
  const int lane = threadIdx.x;
  for (int64_t base = (int64_t)blockIdx.x * 64; base < nact; base += (int64_t)gridDim.x * 64) {
    const int64_t s = base + lane;
    const bool alive = s < nact && src.phase[s] != PH_DONE;
    const unsigned long long mask = __ballot(alive);
    int nbase = 0;
    if (mask) {
      const int leader = __ffsll((long long)mask) - 1;
      if (lane == leader)
        nbase = (int)atomicAdd(&a.ctr[CT_NEWACTIVE], (unsigned long long)__popcll(mask));
      nbase = __shfl(nbase, leader, 64);
    }
    if (s < nact) a.remap[s] = -1;
    if (alive) {
      const int64_t d = nbase + __popcll(mask & ((1ull << lane) - 1ull));
      a.remap[s] = (int32_t)d;
      const int64_t C = a.C;
      for (int e = 0; e < (BDF_QMAX + 1) * NV; ++e) dst.zn[e * C + d] = src.zn[e * C + s];
#pragma unroll
      for (int e = 0; e < NV; ++e) {
        dst.acor[e * C + d] = src.acor[e * C + s];
        dst.ewt[e * C + d] = src.ewt[e * C + s];
      }
#pragma unroll
      for (int e = 0; e < BDF_SCALAR_WORDS; ++e) dst.sc[e * C + d] = src.sc[e * C + s];
      dst.rhoe[d] = src.rhoe[s];
      dst.rhoe[C + d] = src.rhoe[C + s];
      if (a.has_src) {
#pragma unroll
        for (int e = 0; e < NSP; ++e) dst.src[e * C + d] = src.src[e * C + s];
      }
      dst.cell[d] = src.cell[s];
      dst.phase[d] = src.phase[s];
      dst.mb[d] = src.mb[s];
    }
  }
}

// Renumber a pending set-up list after a compaction.
__global__ void __launch_bounds__(64) k_remap(const Args* __restrict__ ap, int par, int gated) {
  const Args& a = *ap;
  if (gated && !a.ctr[CT_GATE]) return;
  const int64_t cnt = (int64_t)a.ctr[CT_SETUP0 + par];
  for (int64_t idx = (int64_t)blockIdx.x * 64 + threadIdx.x; idx < cnt;
       idx += (int64_t)gridDim.x * 64)
    a.list[par][idx] = a.remap[a.list[par][idx]];
  const int64_t cj = (int64_t)a.ctr[CT_JAC0 + par];
  for (int64_t idx = (int64_t)blockIdx.x * 64 + threadIdx.x; idx < cj;
       idx += (int64_t)gridDim.x * 64)
    a.jlist[par][idx] = a.remap[a.jlist[par][idx]];
}

__global__ void k_compact_done(const Args* __restrict__ ap, int gated) {
  const Args& a = *ap;
  if (threadIdx.x == 0 && !(gated && !a.ctr[CT_GATE])) {
    a.ctr[CT_CUR] ^= 1ull;
    a.ctr[CT_NACTIVE] = a.ctr[CT_NEWACTIVE];
    a.ctr[CT_NEWACTIVE] = 0ull;
  }
}

// counters are cleared by a kernel rather than a memset node (graph-capturable)
__global__ void k_init_counters(unsigned long long* c, int64_t nactive, int ingraph) {
  if (threadIdx.x < CT_COUNT)
    c[threadIdx.x] = (threadIdx.x == CT_NACTIVE)   ? (unsigned long long)nactive
                     : (threadIdx.x == CT_INGRAPH) ? (unsigned long long)ingraph
                                                   : 0ull;
}

__global__ void k_set_counter(unsigned long long* c, int idx, unsigned long long v) {
  if (threadIdx.x == 0) c[idx] = v;
}

// ------------------------------------------------------------------ discrete adjoint
struct AdjArgs {
  int64_t ncells;
  const double *rY0, *T0, *rE0, *rYsrc, *rEsrc;  // component-major, ncells per component
  double dt;
  const double* rmul;                             // NRXN rate multipliers exp(log A_j)
  const double *ct_rY, *ct_T;                     // output cotangents (may be null)
  double *rY_out, *T_out;                         // end state
  double *g_rY0, *g_rE0, *g_p;                    // gradients; g_p is NRXN x ncells
  int32_t* status;
  BdfParams prm;
  int seg;          // accepted steps per checkpoint segment
  int kmax;         // checkpoints per cell
  double* ckpt;     // ncells x kmax x CKW
  int tcap;         // records per segment tape
  double* tape;     // ncells x tcap x RS
  int64_t full_cap; // optional complete forward tape (verification): records per cell
  double* full_tape;
  int32_t* full_count;
  double* mats;     // ncells x 5 x n x n: the integrator's and the adjoint's matrices
};

__global__ void k_exp(int n, const double* x, double* y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = x ? exp(x[i]) : 1.0;
}

using AdjCell = BdfCell<Mech, double, GTape<NV>, true, true, true>;
constexpr int ADJ_MATS = 5;  // savedJ, M, Jx (forward) and J, Mf (backward) per cell
constexpr int CKW = (BDF_QMAX + 1) * NV + BDF_SCALAR_WORDS + NV * NV;  // checkpoint words
constexpr int TRS = GTape<NV>::RS;

__device__ inline void ckpt_save(const AdjCell& cell, double* p) {
  for (int j = 0; j <= BDF_QMAX; ++j)
    for (int i = 0; i < NV; ++i) p[j * NV + i] = cell.zn[j][i];
  uint64_t w[BDF_SCALAR_WORDS];
  __builtin_memcpy(w, &cell.S, sizeof(BdfScalars));
  double* q = p + (BDF_QMAX + 1) * NV;
  for (int k = 0; k < BDF_SCALAR_WORDS; ++k) __builtin_memcpy(q + k, w + k, 8);
  q += BDF_SCALAR_WORDS;
  for (int k = 0; k < NV * NV; ++k) q[k] = cell.savedJ[k];
}

__device__ inline void ckpt_load(AdjCell& cell, const double* p) {
  for (int j = 0; j <= BDF_QMAX; ++j)
    for (int i = 0; i < NV; ++i) cell.zn[j][i] = p[j * NV + i];
  uint64_t w[BDF_SCALAR_WORDS];
  const double* q = p + (BDF_QMAX + 1) * NV;
  for (int k = 0; k < BDF_SCALAR_WORDS; ++k) __builtin_memcpy(w + k, q + k, 8);
  __builtin_memcpy(&cell.S, w, sizeof(BdfScalars));
  q += BDF_SCALAR_WORDS;
  for (int k = 0; k < NV * NV; ++k) cell.savedJ[k] = q[k];
  // the Newton matrix in use at the checkpoint: I - gammap J, refactorised (deterministic)
  for (int k = 0; k < NV * NV; ++k) cell.M[k] = -cell.S.gammap * cell.savedJ[k];
  for (int k = 0; k < NV; ++k) cell.M[k * NV + k] += 1.0;
  (void)lu_factor<NV, double>(cell.M, cell.piv);
}

// The adjoint keeps one cell's Nordsieck history, tape records and adjoint state in private memory:
// 24 KB per lane at n = 22 (drm19), 38 KB at n = 41, more than the runtime was shown to back.
// Mechanisms above this size build the kernel with an empty body and prc_adjoint rejects them.
constexpr int ADJ_MAX_N = 32;
constexpr bool ADJ_OK = NV <= ADJ_MAX_N;

// Forward integration with checkpoints, then the backward sweep segment by segment (re-running
// each segment from its checkpoint with the tape on). One cell per lane.
template <int NN>
__global__ void __launch_bounds__(64) k_adjoint(const AdjArgs a) {
  if constexpr (NN <= ADJ_MAX_N) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.ncells) return;
  const int64_t nc = a.ncells;
  double* mats = a.mats + c * (int64_t)ADJ_MATS * NV * NV;
  AdjCell cell;
  cell.savedJ = mats;
  cell.M = mats + NV * NV;
  cell.Jx = mats + 2 * NV * NV;
  cell.rhoe0 = a.rE0[c];
  cell.rhoesrc = a.rEsrc ? a.rEsrc[c] : 0.0;
  cell.src = a.rYsrc ? a.rYsrc + c : nullptr;
  cell.sstride = nc;
  cell.prm = a.prm;
  cell.rmul = a.rmul;
  double y0[NV];
  double rho = 0.0;
  for (int k = 0; k < NSP; ++k) {
    y0[k] = a.rY0[k * nc + c];
    rho += y0[k];
  }
  y0[NSP] = a.T0[c];
  auto fail = [&](int st) {
    a.status[c] = st;
    for (int k = 0; k < NSP; ++k) {
      a.rY_out[k * nc + c] = y0[k];
      a.g_rY0[k * nc + c] = 0.0;
    }
    a.T_out[c] = y0[NSP];
    a.g_rE0[c] = 0.0;
    for (int j = 0; j < Mech::NRXN; ++j) a.g_p[j * nc + c] = 0.0;
  };
  if (!(rho > 0.0) || !(y0[NSP] > 0.0) || !(a.dt > 0.0)) {
    fail(CELL_BADINPUT);
    return;
  }
  double* ck = a.ckpt + c * (int64_t)a.kmax * CKW;
  // forward pass
  if (a.full_tape) {
    cell.tape.base = a.full_tape + c * a.full_cap * TRS;
    cell.tape.cap = (int)a.full_cap;
  }
  cell.start(y0, a.dt);
  int st;
  do {
    st = cell.advance();
    if (st < 0 && cell.S.nst % a.seg == 0) {
      const int k = cell.S.nst / a.seg;
      if (k >= a.kmax) {
        st = CELL_ADJ_STORAGE;
        break;
      }
      ckpt_save(cell, ck + (int64_t)k * CKW);
    }
  } while (st < 0);
  if (a.full_tape) a.full_count[c] = cell.tape.overflow ? -1 : cell.tape.count;
  if (st != CELL_OK) {
    fail(st);
    return;
  }
  const int nst = cell.S.nst;
  double yend[NV];
  for (int i = 0; i < NV; ++i) yend[i] = cell.zn[0][i];
  {
    double Y[NSP];
    double r2 = 0.0;
    for (int k = 0; k < NSP; ++k) r2 += yend[k];
    for (int k = 0; k < NSP; ++k) Y[k] = yend[k] / r2;
    a.T_out[c] = Model::T_from_e((cell.rhoe0 + a.dt * cell.rhoesrc) / r2, Y, yend[NSP]);
    for (int k = 0; k < NSP; ++k) a.rY_out[k * nc + c] = yend[k];
  }
  if (!a.ct_rY && !a.ct_T) {  // no cotangents: forward only
    a.status[c] = CELL_OK;
    for (int k = 0; k < NSP; ++k) a.g_rY0[k * nc + c] = 0.0;
    a.g_rE0[c] = 0.0;
    for (int j = 0; j < Mech::NRXN; ++j) a.g_p[j * nc + c] = 0.0;
    return;
  }
  // backward pass
  Adjoint<Mech> adj;
  adj.rhoe0 = cell.rhoe0;
  adj.rhoesrc = cell.rhoesrc;
  adj.src = cell.src;
  adj.sstride = nc;
  adj.rmul = a.rmul;
  adj.J = mats + 3 * NV * NV;
  adj.Mf = mats + 4 * NV * NV;
  adj.reset();
  adj.seed(yend, a.dt, a.ct_rY ? a.ct_rY + c : nullptr, nc, a.ct_T ? a.ct_T[c] : 0.0);
  double* tp = a.tape + c * (int64_t)a.tcap * TRS;
  const int nseg = (nst + a.seg - 1) / a.seg;
  for (int s = nseg - 1; s >= 0; --s) {
    cell.tape.base = tp;
    cell.tape.cap = a.tcap;
    cell.tape.count = 0;
    cell.tape.overflow = false;
    if (s == 0) {
      cell.start(y0, a.dt);  // records the initial z_1 = h0 f(0, y0)
    } else {
      ckpt_load(cell, ck + (int64_t)s * CKW);
      cell.tstop = a.dt;
    }
    const int target = min((s + 1) * a.seg, nst);
    while (cell.S.nst < target) {
      const int r = cell.advance();
      if (r >= 0 && r != CELL_OK) break;
      if (r == CELL_OK) break;
    }
    if (cell.tape.overflow || cell.S.nst != target) {
      fail(CELL_ADJ_STORAGE);
      return;
    }
    for (int r = cell.tape.count - 1; r >= 0; --r) adj.back(tp + (int64_t)r * TRS);
  }
  a.status[c] = CELL_OK;
  for (int k = 0; k < NSP; ++k) a.g_rY0[k * nc + c] = adj.gy0[k];
  a.g_rE0[c] = adj.ge;
  for (int j = 0; j < Mech::NRXN; ++j) a.g_p[j * nc + c] = adj.gp[j];
  }
}

// ------------------------------------------------------------------ react_state pack / unpack
// PeleC::react_state (Source/React.cpp:29-362) around the reactor, fused into one kernel on each
// side. Conservative state layout of Source/IndexDefines.H:6-13,60-69: URHO 0, UMX..UMZ 1..3,
// UEDEN 4, UEINT 5, UTEMP 6, species at ufs = 7 + NUM_ADV.
constexpr int URHO = 0, UMX = 1, UMY = 2, UMZ = 3, UEDEN = 4, UEINT = 5, UTEMP = 6;
constexpr double R_ERG = 8.31446261815324e7;

struct StateArgs {
  int lo0, lo1, lo2, len0, len1;
  int64_t ncells;
  A4 sold, snew, nonrs, IR;  // NVAR, NVAR, NVAR, NSP + 2 components
  A4 rot;                    // optional 0.5 omega^2 r^2 per cell (rotating frame), p may be null
  A4I mask;                  // optional covered-cell mask (0: skipped), p may be null
  int ufs;
  int react_init;
  double dt;
  // reactor inputs/outputs on the box (workspace, component-major, ncells per component)
  double* rY;     // NSP
  double* T;      // 1
  double* rE;     // 1
  double* rEsrc;  // 1
};

__device__ inline void state_ijk(const StateArgs& a, int64_t c, int& i, int& j, int& k) {
  i = a.lo0 + (int)(c % a.len0);
  const int64_t r = c / a.len0;
  j = a.lo1 + (int)(r % a.len1);
  k = a.lo2 + (int)(r / a.len1);
}

// Pack (React.cpp:111-121, 181-220): STemp = [rhoY, T, rho e] of the old state (the new state
// at initialisation) and the energy forcing rhoedot_ext = (rho E_new - KE_new + rho_new rot -
// rho_old e_old) / dt with e_old = (rho E_old - KE_old) / rho_old + rot.
__global__ void k_state_pack(const StateArgs a) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.ncells) return;
  int i, j, k;
  state_ijk(a, c, i, j, k);
  if (a.mask.p && a.mask.at(i, j, k) == 0) return;  // covered: not integrated
  const A4& so = a.react_init ? a.snew : a.sold;
  const A4& sn = a.snew;
  const double rot = a.rot.p ? a.rot.at(i, j, k, 0) : 0.0;
  const double rho_old = so.at(i, j, k, URHO);
  double mu = so.at(i, j, k, UMX), mv = so.at(i, j, k, UMY), mw = so.at(i, j, k, UMZ);
  double rinv = 1.0 / rho_old;
  const double e_old = (so.at(i, j, k, UEDEN) - 0.5 * (mu * mu + mv * mv + mw * mw) * rinv) * rinv + rot;
  mu = sn.at(i, j, k, UMX);
  mv = sn.at(i, j, k, UMY);
  mw = sn.at(i, j, k, UMZ);
  rinv = 1.0 / sn.at(i, j, k, URHO);
  const double redot = (sn.at(i, j, k, UEDEN) - 0.5 * (mu * mu + mv * mv + mw * mw) * rinv +
                        sn.at(i, j, k, URHO) * rot - rho_old * e_old) /
                       a.dt;
#pragma unroll
  for (int s = 0; s < NSP; ++s) a.rY[s * a.ncells + c] = so.at(i, j, k, a.ufs + s);
  a.T[c] = so.at(i, j, k, UTEMP);
  a.rE[c] = so.at(i, j, k, UEINT);
  a.rEsrc[c] = redot;
}

// Unpack (React.cpp:233-314) and heat release (React.cpp:325-343): new rho, momenta, rhoY, T,
// rho e, rho E (unless initialising), I_R = [(rhoY_new - rhoY_old)/dt - S_nr, d(rho E)/dt - S_nr,
// -sum h_k(T) I_R,k] with h_k from the updated state.
__global__ void k_state_unpack(const StateArgs a) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.ncells) return;
  int i, j, k;
  state_ijk(a, c, i, j, k);
  if (a.mask.p && a.mask.at(i, j, k) == 0) {
    // covered: state untouched, I_R = 0 (React.cpp:96-97,157-164)
#pragma unroll
    for (int s = 0; s < NSP + 2; ++s) a.IR.at(i, j, k, s) = 0.0;
    return;
  }
  const A4& so = a.react_init ? a.snew : a.sold;
  const A4& sn = a.snew;
  const A4& nr = a.nonrs;
  const double rot = a.rot.p ? a.rot.at(i, j, k, 0) : 0.0;
  const double rho_old = so.at(i, j, k, URHO);
  double mu = so.at(i, j, k, UMX), mv = so.at(i, j, k, UMY), mw = so.at(i, j, k, UMZ);
  const double rinv = 1.0 / rho_old;
  const double e_old = (so.at(i, j, k, UEDEN) - 0.5 * (mu * mu + mv * mv + mw * mw) * rinv) * rinv + rot;
  const double redot = a.rEsrc[c];
  const double um = mu + a.dt * nr.at(i, j, k, UMX);
  const double vm = mv + a.dt * nr.at(i, j, k, UMY);
  const double wm = mw + a.dt * nr.at(i, j, k, UMZ);
  double rhonew = 0.0;
#pragma unroll
  for (int s = 0; s < NSP; ++s) rhonew += a.rY[s * a.ncells + c];
  const double reint = rho_old * e_old + a.dt * redot;
  const double ke_new = 0.5 * (um * um + vm * vm + wm * wm) / rhonew - rhonew * rot;
  const double Eold = so.at(i, j, k, UEDEN);
  double IRk[NSP];
#pragma unroll
  for (int s = 0; s < NSP; ++s)
    IRk[s] = (a.rY[s * a.ncells + c] - so.at(i, j, k, a.ufs + s)) / a.dt - nr.at(i, j, k, a.ufs + s);
  const double IRe = (reint + ke_new - Eold) / a.dt - nr.at(i, j, k, UEDEN);
  if (!a.react_init) {
    sn.at(i, j, k, URHO) = rhonew;
    sn.at(i, j, k, UMX) = um;
    sn.at(i, j, k, UMY) = vm;
    sn.at(i, j, k, UMZ) = wm;
#pragma unroll
    for (int s = 0; s < NSP; ++s) sn.at(i, j, k, a.ufs + s) = a.rY[s * a.ncells + c];
    sn.at(i, j, k, UTEMP) = a.T[c];
    sn.at(i, j, k, UEINT) = reint;
    sn.at(i, j, k, UEDEN) = reint + ke_new;
  }
#pragma unroll
  for (int s = 0; s < NSP; ++s) a.IR.at(i, j, k, s) = IRk[s];
  a.IR.at(i, j, k, NSP) = IRe;
  // heat release from the (updated) new state: h_k = e_k + R T / W_k
  const double Tn = sn.at(i, j, k, UTEMP);
  double ek[NSP], cvk[NSP];
  Mech::e_cv(Tn, ek, cvk);
  double hr = 0.0;
#pragma unroll
  for (int s = 0; s < NSP; ++s) hr -= (ek[s] + R_ERG * Tn * Mech::IW[s]) * IRk[s];
  a.IR.at(i, j, k, NSP + 1) = hr;
}

// computeTemp's pc_cmpTemp (Source/Utilities.H:68-85, called from Source/PeleC.cpp:2254-2272):
// T = REY2T(rho, e = rho e / rho, Y) over the full state, the stored T as the Newton guess;
// covered cells (mask == 0) are skipped.
__global__ void k_cmp_temp(int lo0, int lo1, int lo2, int len0, int len1, int64_t ncells, A4 S,
                           A4I mask, int ufs) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  const int i = lo0 + (int)(c % len0);
  const int64_t r = c / len0;
  const int j = lo1 + (int)(r % len1);
  const int k = lo2 + (int)(r / len1);
  if (mask.p && mask.at(i, j, k) == 0) return;
  const double rinv = 1.0 / S.at(i, j, k, URHO);
  double Y[NSP];
#pragma unroll
  for (int s = 0; s < NSP; ++s) Y[s] = S.at(i, j, k, ufs + s) * rinv;
  S.at(i, j, k, UTEMP) = Model::T_from_e(S.at(i, j, k, UEINT) * rinv, Y, S.at(i, j, k, UTEMP));
}

// ------------------------------------------------------------------ EOS helper kernels
__global__ void eos_pyt2re(int64_t nc, const double* p, const double* T, const double* Y,
                           double* rho, double* e) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  double ek[NSP], cv[NSP];
  Mech::e_cv(T[c], ek, cv);
  double iW = 0.0, es = 0.0;
#pragma unroll
  for (int s = 0; s < NSP; ++s) {
    const double y = Y[s * nc + c];
    iW += y * Mech::IW[s];
    es += y * ek[s];
  }
  rho[c] = p[c] / (8.31446261815324e7 * T[c] * iW);
  e[c] = es;
}

__global__ void eos_rey2t(int64_t nc, const double* e, const double* Y, double* T) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  double y[NSP];
#pragma unroll
  for (int s = 0; s < NSP; ++s) y[s] = Y[s * nc + c];
  T[c] = Model::T_from_e(e[c], y, T[c]);
}

__global__ void eos_rty2hi(int64_t nc, const double* T, double* hi) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  double ek[NSP], cv[NSP];
  const double t = T[c];
  Mech::e_cv(t, ek, cv);
#pragma unroll
  for (int s = 0; s < NSP; ++s) hi[s * nc + c] = ek[s] + 8.31446261815324e7 * t * Mech::IW[s];
}

__global__ void eos_rty2wdot(int64_t nc, const double* rho, const double* T, const double* Y,
                             double* wdot) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  double C[NSP], wd[NSP];
#pragma unroll
  for (int s = 0; s < NSP; ++s) C[s] = rho[c] * Y[s * nc + c] * Mech::IW[s];
  if constexpr (NV > KIN_INLINE_MAX_N)
    wdot_outline<Mech, false>(C, T[c], wd, nullptr);
  else
    Mech::wdot(C, T[c], wd);
#pragma unroll
  for (int s = 0; s < NSP; ++s) wdot[s * nc + c] = wd[s];
}

// Reactor model at given states (prc_model_eval): y and ydot component-major (n x nc), J per cell
// n x n row-major. One cell per lane (CellModel) ...
__global__ void __launch_bounds__(64) k_model_lane(int64_t nc, double t, const double* y,
                                                    const double* rhoe0, const double* rhoesrc,
                                                    const double* src, double* ydot, double* J) {
  const int64_t c = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (c >= nc) return;
  double yy[NV], f[NV];
#pragma unroll
  for (int e = 0; e < NV; ++e) yy[e] = y[e * nc + c];
  const double r1 = rhoesrc ? rhoesrc[c] : 0.0;
  Model::rhs(t, yy, rhoe0[c], r1, src ? src + c : nullptr, nc, f);
#pragma unroll
  for (int e = 0; e < NV; ++e) ydot[e * nc + c] = f[e];
  if (J) Model::template jac<double>(t, yy, rhoe0[c], r1, src ? src + c : nullptr, nc, J + c * (int64_t)(NV * NV));
}

// ... or one cell per G-lane group (GroupModel), BS / G per block.
template <int BS, class GM = GModel>
__global__ void __launch_bounds__(BS) k_model_group(int64_t nc, double t, const double* y,
                                                     const double* rhoe0, const double* rhoesrc,
                                                     const double* src, double* ydot, double* J) {
  constexpr int G = GM::GS;
  constexpr int CPB = BS / G;
  __shared__ double lds[CPB * GM::JAC_WORDS];
  const int64_t c = (int64_t)blockIdx.x * CPB + threadIdx.x / G;
  const int l = threadIdx.x & (G - 1);
  if (c >= nc) return;
  double* sh = lds + (threadIdx.x / G) * GM::JAC_WORDS;
  const double yl = l < NV ? y[l * nc + c] : 0.0;
  const double srcl = (src && l < NSP) ? src[l * nc + c] : 0.0;
  const double r1 = rhoesrc ? rhoesrc[c] : 0.0;
  const double f = GM::rhs(t, yl, l, rhoe0[c], r1, srcl, sh);
  if (l < NV) ydot[l * nc + c] = f;
  if (J) GM::jac(t, yl, l, rhoe0[c], r1, srcl, sh, J + c * (int64_t)(NV * NV));
}

}  // namespace
