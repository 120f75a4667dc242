// gfx950 reactor library: kernels and the C-ABI of include/pelec_amd.h for one mechanism.
// Built once per mechanism with -DPELEC_MECH=<ident> -DPELEC_MECH_HEADER="gen/mech_<name>.h".
//
// Kernels
//   react_plain    one cell per lane over a static grid (the plain FP64 drop-in for
//                  ReactorCvode::react, Source/React.cpp:222-228); per-lane working set
//                  (Nordsieck array, Jacobian, LU) in private memory.
//   react_queue    persistent grid, one cell per lane, cells taken from a device work queue:
//                  a lane that finishes its cell writes it back and pulls the next cell index,
//                  so cells with 50x different step counts no longer hold a wave hostage.
//   eos_*          boundary EOS helpers (PYT2RE, REY2T, RTY2Hi, RTY2WDOT).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>

#include "bdf_cell.h"
#include "pelec_amd.h"
#include PELEC_MECH_HEADER

using Mech = pelec::mech::PELEC_MECH;
using namespace pelec;

namespace {

constexpr int NSP = Mech::NSP;
constexpr int NV = NSP + 1;

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

A4 mk(const prc_array4& a) {
  return A4{a.p, a.jstride, a.kstride, a.nstride, a.begin[0], a.begin[1], a.begin[2]};
}
A4I mki(const prc_iarray4& a) {
  return A4I{a.p, a.jstride, a.kstride, a.nstride, a.begin[0], a.begin[1], a.begin[2]};
}

struct BoxDev {
  int lo0, lo1, lo2, len0, len1;
  int64_t ncells;
  int64_t first;  // global index of the box's first cell (batched concatenation)
  A4 rY, rYsrc, T, rE, rEsrc, FC;
  A4I mask;
};

struct LaunchArgs {
  const BoxDev* boxes;  // device array (batched) or nullptr
  BoxDev box;           // single box (plain)
  int nboxes;
  int64_t ncells_total;
  double dt;
  BdfParams prm;
  int32_t* diag;
  int64_t diag_cap;
  unsigned long long* queue;     // work-queue head (batched)
  unsigned long long* nfailed;   // failure counter
};

__device__ inline void cell_ijk(const BoxDev& b, int64_t c, int& i, int& j, int& k) {
  i = b.lo0 + (int)(c % b.len0);
  const int64_t r = c / b.len0;
  j = b.lo1 + (int)(r % b.len1);
  k = b.lo2 + (int)(r / b.len1);
}

__device__ inline void write_diag(const LaunchArgs& a, int64_t gidx, int status, int nst, int nfe,
                                  int nje, int nsetups, int nni, int netf, int ncfn) {
  if (!a.diag || gidx >= a.diag_cap) return;
  int32_t* d = a.diag + gidx;
  const int64_t s = a.diag_cap;
  d[0] = status;
  d[s] = nst;
  d[2 * s] = nfe;
  d[3 * s] = nje;
  d[4 * s] = nsetups;
  d[5 * s] = nni;
  d[6 * s] = netf;
  d[7 * s] = ncfn;
}

// Load cell (i,j,k) of box b into the integrator and start it. Returns false on bad input.
template <typename LUT>
__device__ inline bool load_start(BdfCell<Mech, LUT>& cell, const BoxDev& b, int i, int j, int k,
                                  double dt, const BdfParams& prm) {
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
  cell.prm = prm;
  if (!(rho > 0.0) || !(y0[NSP] > 0.0) || !(dt > 0.0)) return false;
  cell.start(y0, dt);
  return true;
}

// Write the integrated state back: rhoY, T = T(e_end, Y) with the integrated T as guess,
// rho*e = rho*e0 + dt * rho*e_src.
template <typename LUT>
__device__ inline void store(const BdfCell<Mech, LUT>& cell, const BoxDev& b, int i, int j, int k,
                             double dt) {
  double rho = 0.0;
#pragma unroll
  for (int s = 0; s < NSP; ++s) rho += cell.zn[0][s];
  const double irho = 1.0 / rho;
  double Y[NSP];
#pragma unroll
  for (int s = 0; s < NSP; ++s) Y[s] = cell.zn[0][s] * irho;
  const double rEn = cell.rhoe0 + dt * cell.rhoesrc;
  const double Tn = CellModel<Mech>::T_from_e(rEn * irho, Y, cell.zn[0][NSP]);
#pragma unroll
  for (int s = 0; s < NSP; ++s) b.rY.at(i, j, k, s) = cell.zn[0][s];
  b.T.at(i, j, k, 0) = Tn;
  b.rE.at(i, j, k, 0) = rEn;
}

// ------------------------------------------------------------------ plain FP64 drop-in
template <typename LUT>
__global__ void __launch_bounds__(64) react_plain(LaunchArgs a) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const BoxDev& b = a.box;
  if (c >= b.ncells) return;
  int i, j, k;
  cell_ijk(b, c, i, j, k);
  const int64_t g = b.first + c;
  if (b.mask.p && b.mask.at(i, j, k) == 0) {
    if (b.FC.p) b.FC.at(i, j, k, 0) = 0.0;
    write_diag(a, g, CELL_OK, 0, 0, 0, 0, 0, 0, 0);
    return;
  }
  BdfCell<Mech, LUT> cell;
  int s;
  if (!load_start(cell, b, i, j, k, a.dt, a.prm)) {
    s = CELL_BADINPUT;
    cell.nst = cell.nfe = cell.nje = cell.nsetups = cell.nni = cell.netf = cell.ncfn = 0;
  } else {
    do {
      s = cell.advance();
    } while (s < 0);
    if (s == CELL_OK) store(cell, b, i, j, k, a.dt);
  }
  if (s != CELL_OK) atomicAdd(a.nfailed, 1ull);
  if (b.FC.p) b.FC.at(i, j, k, 0) = (double)cell.nfe;
  write_diag(a, g, s, cell.nst, cell.nfe, cell.nje, cell.nsetups, cell.nni, cell.netf, cell.ncfn);
}

// ------------------------------------------------------------------ batched work-queue form
__device__ inline int find_box(const LaunchArgs& a, int64_t g) {
  int lo = 0, hi = a.nboxes - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (a.boxes[mid].first <= g)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

template <typename LUT>
__global__ void __launch_bounds__(64) react_queue(LaunchArgs a) {
  BdfCell<Mech, LUT> cell;
  int64_t g = -1;  // global cell index held by this lane
  int bi = 0, ci = 0, cj = 0, ck = 0;
  bool need = true;
  for (;;) {
    if (need) {
      // pull the next unmasked cell; masked cells are completed on the way
      for (;;) {
        g = (int64_t)atomicAdd(a.queue, 1ull);
        if (g >= a.ncells_total) {
          g = -1;
          break;
        }
        bi = a.boxes ? find_box(a, g) : 0;
        const BoxDev& b = a.boxes ? a.boxes[bi] : a.box;
        cell_ijk(b, g - b.first, ci, cj, ck);
        if (b.mask.p && b.mask.at(ci, cj, ck) == 0) {
          if (b.FC.p) b.FC.at(ci, cj, ck, 0) = 0.0;
          write_diag(a, g, CELL_OK, 0, 0, 0, 0, 0, 0, 0);
          continue;
        }
        if (!load_start(cell, b, ci, cj, ck, a.dt, a.prm)) {
          if (b.FC.p) b.FC.at(ci, cj, ck, 0) = 0.0;
          write_diag(a, g, CELL_BADINPUT, 0, 0, 0, 0, 0, 0, 0);
          atomicAdd(a.nfailed, 1ull);
          continue;
        }
        break;
      }
      need = false;
    }
    if (g < 0) break;  // queue drained for this lane
    const int s = cell.advance();
    if (s >= 0) {
      const BoxDev& b = a.boxes ? a.boxes[bi] : a.box;
      if (s == CELL_OK)
        store(cell, b, ci, cj, ck, a.dt);
      else
        atomicAdd(a.nfailed, 1ull);
      if (b.FC.p) b.FC.at(ci, cj, ck, 0) = (double)cell.nfe;
      write_diag(a, g, s, cell.nst, cell.nfe, cell.nje, cell.nsetups, cell.nni, cell.netf,
                 cell.ncfn);
      need = true;
    }
  }
}

// counters are cleared by a kernel rather than a memset node so the launch can be captured
__global__ void zero_counters(unsigned long long* c) {
  if (threadIdx.x < 2) c[threadIdx.x] = 0ull;
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

__global__ void eos_rey2t(int64_t nc, const double* rho, const double* e, const double* Y,
                          double* T) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  double y[NSP];
#pragma unroll
  for (int s = 0; s < NSP; ++s) y[s] = Y[s * nc + c];
  (void)rho;
  T[c] = CellModel<Mech>::T_from_e(e[c], y, T[c]);
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
  Mech::wdot(C, T[c], wd);
#pragma unroll
  for (int s = 0; s < NSP; ++s) wdot[s * nc + c] = wd[s];
}

}  // namespace

// ====================================================================== host side
struct prc_reactor {
  std::string name;
  int integrator = 0;  // 0 ReactorCvode (BDF), 1 ReactorRK64, 2 ReactorNull
  int reactor_type = 1;
  bool initialised = false;
  double rtol = 1e-6, atol = 1e-10;
  int max_steps = 100000;
  int max_order = 5;
  int form = 0;      // 0 plain, 1 batched
  int lu_fp32 = 0;   // Newton-matrix LU precision
  double typ[NV];
  bool have_typ = false;
  double* d_atol = nullptr;
  unsigned long long* d_counters = nullptr;  // [0] queue head, [1] failures
  BoxDev* d_boxes = nullptr;
  int d_boxes_cap = 0;
  int32_t* diag = nullptr;
  int64_t diag_cap = 0;
  int device = -1;
};

namespace {

#define HIP_OK(expr)                      \
  do {                                    \
    if ((expr) != hipSuccess) return PRC_EHIP; \
  } while (0)

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

bool box_ok(const BoxDev& b) {
  return b.rY.p && b.T.p && b.rE.p && b.len0 >= 0 && b.len1 >= 0;
}

int launch(prc_reactor* r, LaunchArgs& a, hipStream_t s) {
  a.prm.rtol = r->rtol;
  a.prm.atolv = r->d_atol;
  a.prm.qmax = std::min(std::max(r->max_order, 1), BDF_QMAX);
  a.prm.max_steps = r->max_steps;
  a.diag = r->diag;
  a.diag_cap = r->diag_cap;
  a.queue = r->d_counters;
  a.nfailed = r->d_counters + 1;
  hipLaunchKernelGGL(zero_counters, dim3(1), dim3(64), 0, s, r->d_counters);
  if (a.ncells_total == 0) return PRC_OK;
  if (r->form == 0 && a.boxes == nullptr) {
    const int64_t nb = (a.box.ncells + 63) / 64;
    if (r->lu_fp32)
      hipLaunchKernelGGL(react_plain<float>, dim3((unsigned)nb), dim3(64), 0, s, a);
    else
      hipLaunchKernelGGL(react_plain<double>, dim3((unsigned)nb), dim3(64), 0, s, a);
  } else {
    int dev = 0, ncu = 0;
    HIP_OK(hipGetDevice(&dev));
    HIP_OK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    int per_cu = 0;
    if (r->lu_fp32)
      HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, react_queue<float>, 64, 0));
    else
      HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, react_queue<double>, 64, 0));
    per_cu = std::max(per_cu, 1);
    int64_t nb = (int64_t)ncu * per_cu;
    nb = std::min<int64_t>(nb, (a.ncells_total + 63) / 64);
    if (r->lu_fp32)
      hipLaunchKernelGGL(react_queue<float>, dim3((unsigned)nb), dim3(64), 0, s, a);
    else
      hipLaunchKernelGGL(react_queue<double>, dim3((unsigned)nb), dim3(64), 0, s, a);
  }
  HIP_OK(hipGetLastError());
  return PRC_OK;
}

int ensure_init(prc_reactor* r) {
  if (!r) return PRC_EINVAL;
  if (!r->initialised) return PRC_ESTATE;
  return PRC_OK;
}

template <typename K, typename... Args>
int launch_1d(K kern, int64_t nc, void* stream, Args... args) {
  if (nc <= 0) return PRC_OK;
  const int64_t nb = (nc + 255) / 256;
  hipLaunchKernelGGL(kern, dim3((unsigned)nb), dim3(256), 0, (hipStream_t)stream, nc, args...);
  HIP_OK(hipGetLastError());
  return PRC_OK;
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
  else if (!std::strcmp(name, "ReactorNull"))
    kind = 2;
  else
    return PRC_EINVAL;
  prc_reactor* r = new (std::nothrow) prc_reactor;
  if (!r) return PRC_ENOMEM;
  r->name = name;
  r->integrator = kind;
  *out = r;
  return PRC_OK;
}

int prc_init(prc_reactor* r, int reactor_type, int ncells) {
  (void)ncells;
  if (!r || reactor_type != 1) return PRC_EINVAL;
  if (!r->initialised) {
    HIP_OK(hipGetDevice(&r->device));
    if (hipMalloc(&r->d_atol, NV * sizeof(double)) != hipSuccess) return PRC_ENOMEM;
    if (hipMalloc(&r->d_counters, 4 * sizeof(unsigned long long)) != hipSuccess) return PRC_ENOMEM;
    r->initialised = true;
  }
  r->reactor_type = reactor_type;
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
  } else {
    return PRC_EINVAL;
  }
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
  if (r->d_atol) (void)hipFree(r->d_atol);
  if (r->d_counters) (void)hipFree(r->d_counters);
  if (r->d_boxes) (void)hipFree(r->d_boxes);
  r->d_atol = nullptr;
  r->d_counters = nullptr;
  r->d_boxes = nullptr;
  r->d_boxes_cap = 0;
  r->initialised = false;
  return PRC_OK;
}

int prc_destroy(prc_reactor* r) {
  if (!r) return PRC_EINVAL;
  prc_close(r);
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
  LaunchArgs a{};
  a.box = make_box(lo, hi, rY, rYsrc, T, rE, rEsrc, FC, mask, 0);
  if (!box_ok(a.box)) return PRC_EINVAL;
  if (rY.ncomp < NSP && rY.ncomp != 0) return PRC_EINVAL;
  a.boxes = nullptr;
  a.nboxes = 1;
  a.ncells_total = a.box.ncells;
  a.dt = *dt_react;
  if (r->integrator == 2) {
    *time += *dt_react;
    return PRC_OK;
  }
  e = launch(r, a, (hipStream_t)stream);
  if (e) return e;
  *time += *dt_react;
  return PRC_OK;
}

int prc_react_flat(prc_reactor* r, double* rY, const double* rYsrc, double* T, double* rE,
                   const double* rEsrc, double* FC, const int32_t* mask, int64_t ncells,
                   double* dt_react, double* time, void* stream) {
  if (ncells < 0 || ncells > INT32_MAX) return PRC_EINVAL;
  const int32_t lo[3] = {0, 0, 0};
  const int32_t hi[3] = {(int32_t)ncells - 1, 0, 0};
  auto view = [&](const double* p, int nc) {
    prc_array4 v{};
    v.p = const_cast<double*>(p);
    v.jstride = ncells;
    v.kstride = ncells;
    v.nstride = ncells;
    v.end[0] = (int32_t)ncells;
    v.end[1] = 1;
    v.end[2] = 1;
    v.ncomp = nc;
    return v;
  };
  prc_iarray4 m{};
  m.p = const_cast<int32_t*>(mask);
  m.jstride = m.kstride = m.nstride = ncells;
  m.end[0] = (int32_t)ncells;
  m.end[1] = m.end[2] = 1;
  m.ncomp = 1;
  if (ncells == 0) {
    if (!r) return PRC_EINVAL;
    if (dt_react && time) *time += *dt_react;
    return ensure_init(r);
  }
  return prc_react_box(r, lo, hi, view(rY, NSP), view(rYsrc, NSP), view(T, 1), view(rE, 1),
                       view(rEsrc, 1), view(FC, 1), m, dt_react, time, stream);
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
  if (r->form == 0) {
    // plain form: box by box, as the reference loops over MFIter boxes (React.cpp:137-228)
    int64_t first = 0;
    for (int ib = 0; ib < nboxes; ++ib) {
      const prc_box_args& x = boxes[ib];
      LaunchArgs a{};
      a.box = make_box(x.lo, x.hi, x.rY, x.rYsrc_ext, x.T, x.rEner, x.rEner_src_ext, x.FC,
                       x.mask, first);
      if (!box_ok(a.box)) return PRC_EINVAL;
      a.nboxes = 1;
      a.ncells_total = a.box.ncells;
      a.dt = *dt_react;
      first += a.box.ncells;
      if (a.box.ncells == 0) continue;
      e = launch(r, a, s);
      if (e) return e;
    }
    *time += *dt_react;
    return PRC_OK;
  }
  // batched form: one launch over the concatenation of all boxes
  BoxDev* hb = new (std::nothrow) BoxDev[nboxes];
  if (!hb) return PRC_ENOMEM;
  int64_t first = 0;
  for (int ib = 0; ib < nboxes; ++ib) {
    const prc_box_args& x = boxes[ib];
    hb[ib] = make_box(x.lo, x.hi, x.rY, x.rYsrc_ext, x.T, x.rEner, x.rEner_src_ext, x.FC, x.mask,
                      first);
    if (!box_ok(hb[ib])) {
      delete[] hb;
      return PRC_EINVAL;
    }
    first += hb[ib].ncells;
  }
  if (nboxes > r->d_boxes_cap) {
    if (r->d_boxes) HIP_OK(hipFree(r->d_boxes));
    r->d_boxes = nullptr;
    r->d_boxes_cap = 0;
    if (hipMalloc(&r->d_boxes, sizeof(BoxDev) * nboxes) != hipSuccess) {
      delete[] hb;
      return PRC_ENOMEM;
    }
    r->d_boxes_cap = nboxes;
  }
  // the box table is read by the kernel: copy it in stream order, then release the host copy
  if (hipMemcpyAsync(r->d_boxes, hb, sizeof(BoxDev) * nboxes, hipMemcpyHostToDevice, s) !=
      hipSuccess) {
    delete[] hb;
    return PRC_EHIP;
  }
  if (hipStreamSynchronize(s) != hipSuccess) {
    delete[] hb;
    return PRC_EHIP;
  }
  delete[] hb;
  LaunchArgs a{};
  a.boxes = r->d_boxes;
  a.nboxes = nboxes;
  a.ncells_total = first;
  a.dt = *dt_react;
  e = launch(r, a, s);
  if (e) return e;
  *time += *dt_react;
  return PRC_OK;
}

int64_t prc_last_failures(prc_reactor* r, void* stream) {
  if (!r || !r->initialised) return PRC_EINVAL;
  unsigned long long h = 0;
  if (hipMemcpyAsync(&h, r->d_counters + 1, sizeof(h), hipMemcpyDeviceToHost,
                     (hipStream_t)stream) != hipSuccess)
    return PRC_EHIP;
  if (hipStreamSynchronize((hipStream_t)stream) != hipSuccess) return PRC_EHIP;
  return (int64_t)h;
}

int prc_eos_pyt2re(int64_t nc, const double* p, const double* T, const double* Y, double* rho,
                   double* e, void* stream) {
  if (nc < 0 || (nc > 0 && (!p || !T || !Y || !rho || !e))) return PRC_EINVAL;
  return launch_1d(eos_pyt2re, nc, stream, p, T, Y, rho, e);
}
int prc_eos_rey2t(int64_t nc, const double* rho, const double* e, const double* Y, double* T,
                  void* stream) {
  if (nc < 0 || (nc > 0 && (!e || !Y || !T))) return PRC_EINVAL;
  return launch_1d(eos_rey2t, nc, stream, rho, e, Y, T);
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

}  // extern "C"
