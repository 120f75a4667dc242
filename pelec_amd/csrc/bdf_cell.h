// Plain FP64 drop-in: one cell per lane, the whole CVODE integration of the cell in one call,
// working set (Nordsieck array, Jacobian, Newton-matrix LU) in private memory. This is the
// straightforward GPU form of ReactorCvode::react (Source/React.cpp:222-228) that the batched
// form is measured against; both use the step logic of bdf_core.h.
#pragma once
#include <type_traits>

#include "bdf_core.h"

namespace pelec {

// Dense LU with partial pivoting (row interchanges), in place, row-major n x n.
template <int n, typename T>
__device__ inline int lu_factor(T* a, int* piv) {
  for (int k = 0; k < n; ++k) {
    int p = k;
    T amax = fabs(a[k * n + k]);
    for (int i = k + 1; i < n; ++i) {
      const T v = fabs(a[i * n + k]);
      if (v > amax) {
        amax = v;
        p = i;
      }
    }
    piv[k] = p;
    if (a[p * n + k] == T(0)) return k + 1;
    if (p != k) {
      for (int j = 0; j < n; ++j) {
        const T tmp = a[k * n + j];
        a[k * n + j] = a[p * n + j];
        a[p * n + j] = tmp;
      }
    }
    const T inv = T(1) / a[k * n + k];
    for (int i = k + 1; i < n; ++i) {
      const T l = a[i * n + k] * inv;
      a[i * n + k] = l;
      for (int j = k + 1; j < n; ++j) a[i * n + j] -= l * a[k * n + j];
    }
  }
  return 0;
}

// Solve (LU) x = b in place (interchanges applied in factorisation order).
template <int n, typename T>
__device__ inline void lu_solve(const T* a, const int* piv, double* b) {
  T x[n];
#pragma unroll
  for (int i = 0; i < n; ++i) x[i] = T(b[i]);
  for (int k = 0; k < n; ++k) {
    const int p = piv[k];
    const T tmp = x[k];
    x[k] = x[p];
    x[p] = tmp;
  }
  for (int i = 1; i < n; ++i) {
    T s = x[i];
    for (int j = 0; j < i; ++j) s -= a[i * n + j] * x[j];
    x[i] = s;
  }
  for (int i = n - 1; i >= 0; --i) {
    T s = x[i];
    for (int j = i + 1; j < n; ++j) s -= a[i * n + j] * x[j];
    x[i] = s / a[i * n + i];
  }
#pragma unroll
  for (int i = 0; i < n; ++i) b[i] = double(x[i]);
}

// Tape (default NoTapeRec: nothing recorded) receives the Nordsieck operations of the step
// sequence for the discrete adjoint; MUL scales reaction j's rate constant by rmul[j].
struct NoTapeRec {
  template <class V> __device__ void rec_rescale(const V&, int, double) {}
  template <class V> __device__ void rec_pascal(const V&, int, double) {}
  template <class V> __device__ void rec_order_up(const V&, int, int, int, double, const double*) {}
  template <class V> __device__ void rec_order_down(const V&, int, const double*) {}
  template <class V> __device__ void rec_correct(const V&, const BdfScalars&, bool, int) {}
  template <class V> __device__ void rec_restart(const V&, double, double) {}
  template <class V> __device__ void rec_init(const V&, double) {}
  __device__ void rec_mark_save() {}
};

// POLISH: after the modified Newton iteration converges, the corrector equation is solved to
// machine precision with the exact Jacobian (the discrete adjoint linearises the exactly solved
// step map, so its forward trajectory must be that map's). EXTMAT: the n x n matrices live in
// caller-provided device memory (savedJ, M, Jx point there) instead of private memory. NREF: FP64
// refinement steps after each solve with FP32 factors (unused with FP64 factors).
template <class Mech, typename LUT, class Tape = NoTapeRec, bool MUL = false, bool POLISH = false,
          bool EXTMAT = false, int NREF = 1>
struct BdfCell {
  using Model = CellModel<Mech>;
  static constexpr int N = Mech::NSP;
  static constexpr int n = N + 1;

  // problem
  double rhoe0, rhoesrc;
  const double* src;
  int64_t sstride;
  BdfParams prm;
  double tstop;
  // state
  BdfScalars S;
  double zn[BDF_QMAX + 1][n];
  double ewt[n], acor[n];
  std::conditional_t<EXTMAT, double*, double[n * n]> savedJ;
  std::conditional_t<EXTMAT, LUT*, LUT[n * n]> M;
  int piv[n];
  Tape tape;
  const double* rmul = nullptr;  // per-reaction rate multipliers (MUL)
  // exact Jacobian / Newton matrix of the polish
  std::conditional_t<EXTMAT, double*, double[POLISH ? n * n : 1]> Jx;
  int pivx[POLISH ? n : 1];

  struct Vec : LaneComps {
    static constexpr bool kMemory = false;
    BdfCell* c;
    __device__ double& Z(int j, int i) const { return c->zn[j][i]; }
    __device__ double& A(int i) const { return c->acor[i]; }
    __device__ double& E(int i) const { return c->ewt[i]; }
    __device__ void rec_rescale(int q, double eta) const { c->tape.rec_rescale(*this, q, eta); }
    __device__ void rec_pascal(int q, double sign) const { c->tape.rec_pascal(*this, q, sign); }
    __device__ void rec_order_up(int q, int L, int ia, double A1, const double* l) const {
      c->tape.rec_order_up(*this, q, L, ia, A1, l);
    }
    __device__ void rec_order_down(int q, const double* l) const {
      c->tape.rec_order_down(*this, q, l);
    }
    __device__ void rec_correct(const BdfScalars& S, bool save, int qmax) const {
      c->tape.rec_correct(*this, S, save, qmax);
    }
    __device__ void rec_restart(double h, double t) const { c->tape.rec_restart(*this, h, t); }
    __device__ void rec_init(double h) const { c->tape.rec_init(*this, h); }
    __device__ void rec_mark_save() const { c->tape.rec_mark_save(); }
  };
  __device__ Vec vec() { return Vec{{}, this}; }

  __device__ void f(double t, const double* yy, double* out) {
    S.nfe++;
    Model::template rhs<MUL>(t, yy, rhoe0, rhoesrc, src, sstride, out, rmul);
  }

  __device__ int lsetup() {
    if (jac_needed(S)) {
      S.nje++;
      S.nstlj = S.nst;
      S.jcur = 1;
      Model::template jac<double, MUL>(S.tn, zn[0], rhoe0, rhoesrc, src, sstride, savedJ, nullptr,
                                       rmul);
    } else {
      S.jcur = 0;
    }
    for (int i = 0; i < n * n; ++i) M[i] = LUT(-S.gamma * savedJ[i]);
#pragma unroll
    for (int i = 0; i < n; ++i) M[i * n + i] += LUT(1);
    return lu_factor<n, LUT>(M, piv);
  }

  // Solve (I - gammap J) x = b in place; FP32 factors get NREF FP64 refinement steps
  // (r = b - (x - gammap J x) with the FP64 Jacobian, x += M^-1 r).
  __device__ void solve(double* b) {
    if constexpr (sizeof(LUT) == sizeof(double)) {
      lu_solve<n, LUT>(M, piv, b);
    } else {
      double x[n], r[n];
#pragma unroll
      for (int i = 0; i < n; ++i) x[i] = b[i];
      lu_solve<n, LUT>(M, piv, x);
      for (int it = 0; it < NREF; ++it) {
        for (int i = 0; i < n; ++i) {
          double jx = 0.0;
          for (int j = 0; j < n; ++j) jx += savedJ[i * n + j] * x[j];
          r[i] = b[i] - (x[i] - S.gammap * jx);
        }
        lu_solve<n, LUT>(M, piv, r);
#pragma unroll
        for (int i = 0; i < n; ++i) x[i] += r[i];
      }
#pragma unroll
      for (int i = 0; i < n; ++i) b[i] = x[i];
    }
  }

  // cvNls with the dense direct linear solver: 0 converged, 3 failed
  __device__ int nls(bool callSetup) {
    for (;;) {
      double ftemp[n];
      f(S.tn, zn[0], ftemp);
      if (callSetup) {
        const int ier = lsetup();
        after_setup(S);
        callSetup = false;
        if (ier != 0) return 3;
      }
#pragma unroll
      for (int i = 0; i < n; ++i) acor[i] = 0.0;
      S.m = 0;
      int r;
      for (;;) {
        r = newton_correct<n>(S, vec(), ftemp, [&](double* b) { solve(b); });
        if (r != 1) break;
        double y[n];
#pragma unroll
        for (int i = 0; i < n; ++i) y[i] = zn[0][i] + acor[i];
        f(S.tn, y, ftemp);
      }
      if (r == 0) {
        if constexpr (POLISH) polish();
        return 0;
      }
      if (r == 3) return 3;
      callSetup = true;  // stale Jacobian: rebuild and retry
      S.convfail = 2;
    }
  }

  // acor = gamma f(tn, zn0 + acor) - zn1 / l1 solved to convergence (exact Jacobian each iteration)
  __device__ void polish() {
    const double rl1 = 1.0 / S.l[1];
    for (int it = 0; it < 8; ++it) {
      double y[n], g[n];
#pragma unroll
      for (int i = 0; i < n; ++i) y[i] = zn[0][i] + acor[i];
      f(S.tn, y, g);
      for (int i = 0; i < n; ++i) g[i] = S.gamma * g[i] - rl1 * zn[1][i] - acor[i];
      Model::template jac<double, MUL>(S.tn, y, rhoe0, rhoesrc, src, sstride, Jx, nullptr, rmul);
      for (int k = 0; k < n * n; ++k) Jx[k] = -S.gamma * Jx[k];
      for (int k = 0; k < n; ++k) Jx[k * n + k] += 1.0;
      if (lu_factor<n, double>(Jx, pivx) != 0) break;
      lu_solve<n, double>(Jx, pivx, g);
      bool conv = true;
      for (int i = 0; i < n; ++i) {
        acor[i] += g[i];
        conv = conv && fabs(g[i]) <= 1e-14 * fabs(zn[0][i] + acor[i]) + 1e-300;
      }
      if (conv) break;
    }
    S.acnrm = wrms_acor<n>(vec());
  }

  // cvStep
  __device__ int step() {
    step_prologue<n>(S, vec(), prm.qmax);
    for (;;) {
      const bool cs = begin_attempt<n>(S, vec(), tstop);
      if (nls(cs) != 0) {
        const int s = conv_fail<n>(S, vec());
        if (s >= 0) return s;
        continue;
      }
      const double dsm = S.acnrm * S.tq[2];
      if (dsm <= 1.0) {
        finish_step<n>(S, vec(), dsm, prm.qmax);
        return CELL_OK;
      }
      const int s = error_test_fail<n>(S, vec(), dsm, prm.qmax);
      if (s >= 0) return s;
      if (s == -2) {
        vec().rec_restart(S.h, S.tn);
        double t[n];
        f(S.tn, zn[0], t);
#pragma unroll
        for (int i = 0; i < n; ++i) zn[1][i] = S.h * t[i];
      }
    }
  }

  __device__ void start(const double* y0, double tout) {
    scalars_reset(S, prm.qmax);
    tstop = tout;
#pragma unroll
    for (int i = 0; i < n; ++i) zn[0][i] = y0[i];
    set_ewt<n>(vec(), prm);
    f(S.tn, zn[0], zn[1]);
    S.h = hin<n>(S, vec(), tout, [&](double t, const double* yy, double* out) { f(t, yy, out); });
    if ((S.tn + S.h - tout) * S.h > 0.0) S.h = (tout - S.tn) * (1.0 - 4.0 * cv::UROUND);
    S.hscale = S.h;
    S.hprime = S.h;
#pragma unroll
    for (int i = 0; i < n; ++i) zn[1][i] *= S.h;
    vec().rec_init(S.h);
  }

  // Advance one step. Returns <0 while running, CELL_OK at tout, or a failure status.
  __device__ int advance() {
    if (S.nst >= prm.max_steps) return CELL_MAXSTEPS;
    set_ewt<n>(vec(), prm);
    const int s = step();
    if (s != CELL_OK) return s;
    return reached_tstop(S, tstop) ? CELL_OK : -1;
  }
};

}  // namespace pelec
