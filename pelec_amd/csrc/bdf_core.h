// CVODE BDF step logic shared by the plain one-cell-per-lane kernel and the batched phase kernels.
//
// The integrator is CVODE's published variable-order (1..qmax) variable-step BDF in Nordsieck
// form, as the reference runs it through ReactorCvode (Source/React.cpp:222-228; solver knobs
// cvode.solve_type / cvode.max_order at Exec/Production/JetFlame/inputs.inp:30-31):
//   cvSetBDF / cvSetTqBDF   -> set_bdf          cvPredict / cvRestore -> predict / restore
//   cvRescale               -> rescale          cvIncreaseBDF / cvDecreaseBDF -> adjust_order
//   cvCompleteStep          -> complete_step    cvPrepareNextStep / cvChooseEta / cvSetEta
//   cvDoErrorTest           -> error_test_fail  cvHandleNFlag (conv. failure) -> conv_fail
//   cvHin / cvUpperBoundH0 / cvYddNorm -> hin
// with the step clipped to land exactly on t = dt_react (CVODE's tstop handling).
//
// Storage is abstracted: BdfScalars holds every per-cell scalar of the integrator, and a vector
// accessor V provides Z(j,i) (Nordsieck column j), A(i) (accumulated correction) and E(i) (error
// weights) as double references -- private arrays in the plain kernel, structure-of-arrays HBM in
// the batched kernels. Both forms therefore take identical decisions for a cell.
#pragma once
#include "cell_model.h"

namespace pelec {

constexpr int BDF_QMAX = 5;

enum CellStatus : int {
  CELL_OK = 0,
  CELL_MAXSTEPS = 1,
  CELL_ERRFAIL = 2,
  CELL_CONVFAIL = 3,
  CELL_BADINPUT = 4,
  CELL_ADJ_STORAGE = 5,  // adjoint: checkpoint or tape capacity exceeded
};

struct BdfParams {
  double rtol;
  const double* atolv;  // n entries (device, shared by all cells)
  int qmax;
  int max_steps;
};

namespace cv {
constexpr double ETAMX1 = 10000.0, ETAMX2 = 10.0, ETAMX3 = 10.0, ETAMXF = 0.2, ETAMIN = 0.1,
                 ETACF = 0.25, ADDON = 1e-6, BIAS1 = 6.0, BIAS2 = 6.0, BIAS3 = 10.0,
                 THRESH = 1.5, CRDOWN = 0.3, DGMAX = 0.3, RDIV = 2.0, NLSCOEF = 0.1,
                 DGMAX_LS = 0.2, UROUND = 2.220446049250313e-16;
constexpr int MXNEF1 = 3, SMALL_NEF = 2, LONG_WAIT = 10, MAXCOR = 3, MXNCF = 10, MXNEF = 7,
              MSBP = 20, MSBJ = 50, SMALL_NST = 10;
}  // namespace cv

// Every per-cell scalar of the integrator (kept as whole 8-byte words so the batched form can
// stream it to and from structure-of-arrays storage).
struct BdfScalars {
  double tn, h, hscale, hprime, eta, etamax, saved_tq5;
  double gamma, gammap, gamrat, crate, acnrm, delp, saved_t;
  double tau[BDF_QMAX + 2], tq[6], l[BDF_QMAX + 2];
  int q, qprime, qwait, L, nstlp, nstlj, indx_acor, jcur;
  int nst, nfe, nje, nsetups, nni, netf, ncfn;
  int m, convfail, nflag, ncf, nef;
};
static_assert(sizeof(BdfScalars) % 8 == 0, "BdfScalars must be whole words");
constexpr int BDF_SCALAR_WORDS = sizeof(BdfScalars) / 8;

// a[idx] for a small array with a run-time index, as selects: every loop over the coefficient
// arrays below has a static trip count (predicated on q), so the integrator scalars stay in
// registers instead of private memory.
template <int K>
__device__ inline double sel(const double (&a)[K], int idx) {
  // The empty asm makes every element an opaque register value, so the compiler cannot fold the
  // select chain back into one load at a computed offset (which would pin the array in scratch).
  double r = a[0];
#pragma unroll
  for (int i = 1; i < K; ++i) {
    double v = a[i];
    __asm__("" : "+v"(v));
    r = (i == idx) ? v : r;
  }
  return r;
}

// Operations on the Nordsieck array are reported to the accessor (rec_* members), so a recording
// accessor can tape the step sequence for the discrete adjoint. NoTape is the empty default.
struct NoTape {
  __device__ void rec_rescale(int, double) const {}
  __device__ void rec_pascal(int, double) const {}
  __device__ void rec_order_up(int, int, int, double, const double*) const {}
  __device__ void rec_order_down(int, const double*) const {}
  __device__ void rec_correct(const struct BdfScalars&, bool, int) const {}
  __device__ void rec_restart(double, double) const {}
  __device__ void rec_init(double) const {}
  __device__ void rec_mark_save() const {}
};

// Component distribution of a vector accessor. Per-lane accessors (LaneComps) hold all n
// components of a cell in one lane: component loops run i = 0..n-1, sums are plain, and work
// vectors are private arrays. A group accessor holds component i on lane i of a group of lanes:
// c0() is the lane, kStep the group size, sum()/max() are group reductions returning the same
// value on every lane, and Arr<n> is one register per lane.
struct LaneComps {
  static constexpr int kStep = 1;
  __device__ static constexpr int c0() { return 0; }
  __device__ static double sum(double x) { return x; }
  __device__ static double max(double x) { return x; }
  template <int m>
  using Arr = double[m];
};

// Vector accessors declare kMemory = true when Z/A/E live in device memory (the batched form's
// structure-of-arrays state). The Nordsieck updates below then run one component at a time and
// end each component with a compiler memory fence: the compiler keeps one component's <= 6
// history values in registers instead of forwarding the whole (q+1) x n array between passes
// (which spills). For private-array accessors the fence is omitted.
template <class V>
__device__ inline void component_fence() {
  if constexpr (V::kMemory) PELEC_MEMORY_FENCE();
}

// Components per chunk of a Nordsieck pass. With the batched form's structure-of-arrays state
// (one cell per lane, kMemory) a pass loads a chunk's history together, updates it in registers,
// stores it and fences. One component at a time would cost a memory round trip per component and
// pass: a component's loads cannot move above the previous component's stores, because the
// compiler cannot prove that rows C slots apart do not alias. Private-array and group accessors
// keep one component per iteration.
constexpr int BDF_KCHUNK = 6;
template <class V>
__device__ constexpr int chunk_of() {
  return (V::kMemory && V::kStep == 1) ? BDF_KCHUNK : 1;
}

// Column j of the Nordsieck array for component i, loaded only where j <= q.
template <class V>
__device__ inline void load_hist(const V& v, int i, int q, double (&z)[BDF_QMAX + 1]) {
#pragma unroll
  for (int j = 0; j <= BDF_QMAX; ++j) z[j] = (j <= q) ? v.Z(j, i) : 0.0;
}
template <class V>
__device__ inline void store_hist(const V& v, int i, int q, const double (&z)[BDF_QMAX + 1]) {
#pragma unroll
  for (int j = 0; j <= BDF_QMAX; ++j)
    if (j <= q) v.Z(j, i) = z[j];
}

__device__ inline void scalars_reset(BdfScalars& S, int qmax) {
  S.tn = S.h = S.hscale = S.hprime = 0.0;
  S.eta = 1.0;
  S.etamax = cv::ETAMX1;
  S.saved_tq5 = 0.0;
  S.gamma = S.gammap = 0.0;
  S.gamrat = S.crate = 1.0;
  S.acnrm = S.delp = S.saved_t = 0.0;
#pragma unroll
  for (int i = 0; i < BDF_QMAX + 2; ++i) S.tau[i] = 0.0;
#pragma unroll
  for (int i = 0; i < 6; ++i) S.tq[i] = 0.0;
#pragma unroll
  for (int i = 0; i < BDF_QMAX + 2; ++i) S.l[i] = 0.0;
  S.q = 1;
  S.qprime = 1;
  S.L = 2;
  S.qwait = 2;
  S.nstlp = S.nstlj = 0;
  S.indx_acor = qmax;
  S.jcur = 0;
  S.nst = S.nfe = S.nje = S.nsetups = S.nni = S.netf = S.ncfn = 0;
  S.m = S.convfail = S.nflag = S.ncf = S.nef = 0;
}

template <int n, class V>
__device__ inline double wrms_acor(const V& v) {
  double s = 0.0;
#pragma unroll
  for (int i = v.c0(); i < n; i += V::kStep) {
    const double w = v.A(i) * v.E(i);
    s += w * w;
  }
  return sqrt(v.sum(s) / n);
}

template <int n, class V, class X>
__device__ inline double wrms_reg(const V& v, const X& x) {
  double s = 0.0;
#pragma unroll
  for (int i = v.c0(); i < n; i += V::kStep) {
    const double w = x[i] * v.E(i);
    s += w * w;
  }
  return sqrt(v.sum(s) / n);
}

template <int n, class V>
__device__ inline void set_ewt(const V& v, const BdfParams& prm) {
  constexpr int K = chunk_of<V>();
#pragma unroll
  for (int i0 = v.c0(); i0 < n; i0 += V::kStep * K) {
    double z0[K];
#pragma unroll
    for (int c = 0; c < K; ++c) {
      const int i = i0 + c * V::kStep;
      z0[c] = i < n ? v.Z(0, i) : 0.0;
    }
#pragma unroll
    for (int c = 0; c < K; ++c) {
      const int i = i0 + c * V::kStep;
      if (i < n) v.E(i) = 1.0 / (prm.rtol * fabs(z0[c]) + prm.atolv[i]);
    }
    component_fence<V>();
  }
}

template <int n, class V>
__device__ inline void rescale(BdfScalars& S, const V& v) {
  double f[BDF_QMAX + 1];
  f[0] = 1.0;
#pragma unroll
  for (int j = 1; j <= BDF_QMAX; ++j) f[j] = f[j - 1] * S.eta;
  const int q = S.q;
  v.rec_rescale(q, S.eta);
  constexpr int K = chunk_of<V>();
#pragma unroll
  for (int i0 = v.c0(); i0 < n; i0 += V::kStep * K) {
    double z[K][BDF_QMAX + 1];
#pragma unroll
    for (int c = 0; c < K; ++c) {
      const int i = i0 + c * V::kStep;
#pragma unroll
      for (int j = 1; j <= BDF_QMAX; ++j) z[c][j] = (i < n && j <= q) ? v.Z(j, i) : 0.0;
    }
#pragma unroll
    for (int c = 0; c < K; ++c) {
      const int i = i0 + c * V::kStep;
#pragma unroll
      for (int j = 1; j <= BDF_QMAX; ++j)
        if (i < n && j <= q) v.Z(j, i) = z[c][j] * f[j];
    }
    component_fence<V>();
  }
  S.h = S.hscale * S.eta;
  S.hscale = S.h;
}

// Pascal-triangle update of one component's history in registers (rows 0..q).
__device__ inline void pascal_apply(double (&z)[BDF_QMAX + 1], int q, double sign) {
#pragma unroll
  for (int k = 1; k <= BDF_QMAX; ++k)
#pragma unroll
    for (int j = BDF_QMAX; j >= k; --j)
      z[j - 1] = (k <= q && j <= q) ? z[j - 1] + sign * z[j] : z[j - 1];
}

// Pascal-triangle update of the Nordsieck array (sign +1: predict, -1: undo the prediction),
// k = 1..q, j = q..k: z_{j-1} += sign z_j -- per component, in registers.
template <int n, class V>
__device__ inline void pascal(const V& v, int q, double sign) {
  v.rec_pascal(q, sign);
  constexpr int K = chunk_of<V>();
#pragma unroll
  for (int i0 = v.c0(); i0 < n; i0 += V::kStep * K) {
    double z[K][BDF_QMAX + 1];
#pragma unroll
    for (int c = 0; c < K; ++c) {
      const int i = i0 + c * V::kStep;
      if (i < n) {
        load_hist(v, i, q, z[c]);
      } else {
#pragma unroll
        for (int j = 0; j <= BDF_QMAX; ++j) z[c][j] = 0.0;
      }
    }
#pragma unroll
    for (int c = 0; c < K; ++c) pascal_apply(z[c], q, sign);
#pragma unroll
    for (int c = 0; c < K; ++c) {
      const int i = i0 + c * V::kStep;
      if (i < n) store_hist(v, i, q - 1, z[c]);
    }
    component_fence<V>();
  }
}

template <int n, class V>
__device__ inline void predict(BdfScalars& S, const V& v, double tstop) {
  S.tn += S.h;
  if ((S.tn - tstop) * S.h > 0.0) S.tn = tstop;
  pascal<n>(v, S.q, 1.0);
}

template <int n, class V>
__device__ inline void restore(BdfScalars& S, const V& v) {
  S.tn = S.saved_t;
  pascal<n>(v, S.q, -1.0);
}

__device__ inline void set_bdf(BdfScalars& S) {
  double alpha0, alpha0_hat, xi_inv, xistar_inv, hsum;
  const int q = S.q;
  S.l[0] = S.l[1] = xi_inv = xistar_inv = 1.0;
#pragma unroll
  for (int i = 2; i < BDF_QMAX + 2; ++i) S.l[i] = 0.0;
  alpha0 = alpha0_hat = -1.0;
  hsum = S.h;
  if (q > 1) {
#pragma unroll
    for (int j = 2; j < BDF_QMAX; ++j) {
      const bool on = j < q;
      hsum = on ? hsum + S.tau[j - 1] : hsum;
      const double xij = S.h / hsum;
      xi_inv = on ? xij : xi_inv;
      alpha0 = on ? alpha0 - 1.0 / j : alpha0;
#pragma unroll
      for (int i = j; i >= 1; --i) S.l[i] = on ? S.l[i] + S.l[i - 1] * xij : S.l[i];
    }
    alpha0 -= 1.0 / q;
    xistar_inv = -S.l[1] - alpha0;
    hsum += sel(S.tau, q - 1);
    xi_inv = S.h / hsum;
    alpha0_hat = -S.l[1] - xi_inv;
#pragma unroll
    for (int i = BDF_QMAX; i >= 1; --i) S.l[i] = (i <= q) ? S.l[i] + S.l[i - 1] * xistar_inv : S.l[i];
  }
  const double lq = sel(S.l, q);
  const double A1 = 1.0 - alpha0_hat + alpha0;
  const double A2 = 1.0 + q * A1;
  S.tq[2] = fabs(A1 / (alpha0 * A2));
  S.tq[5] = fabs(A2 * xistar_inv / (lq * xi_inv));
  if (S.qwait == 1) {
    if (q > 1) {
      const double Cc = xistar_inv / lq;
      const double A3 = alpha0 + 1.0 / q;
      const double A4 = alpha0_hat + xi_inv;
      const double Cpinv = (1.0 - A4 + A3) / A3;
      S.tq[1] = fabs(Cc * Cpinv);
    } else {
      S.tq[1] = 1.0;
    }
    hsum += sel(S.tau, q);
    xi_inv = S.h / hsum;
    const double A5 = alpha0 - 1.0 / (q + 1);
    const double A6 = alpha0_hat - xi_inv;
    const double Cppinv = (1.0 - A6 + A5) / A2;
    S.tq[3] = fabs(Cppinv / (xi_inv * (q + 2) * A5));
  }
  S.tq[4] = cv::NLSCOEF / S.tq[2];
}

// Scalar part of an order change (cvIncreaseBDF / cvDecreaseBDF): the coefficients in S.l, and
// A1 for an increase. Returns 1 (increase) or -1 (decrease) when the Nordsieck array needs the
// component update (order_apply), 0 when it does not (a decrease from order 2). The operation is
// recorded here.
template <class V>
__device__ inline int order_coeffs(BdfScalars& S, const V& v, int dq, double& A1) {
  if (S.q == 2 && dq != 1) return 0;
  if (dq == 1) {
#pragma unroll
    for (int i = 0; i < BDF_QMAX + 2; ++i) S.l[i] = 0.0;
    double alpha1 = 1.0, prod = 1.0, xiold = 1.0, alpha0 = -1.0, hsum = S.hscale;
    S.l[2] = 1.0;
    if (S.q > 1) {
#pragma unroll
      for (int j = 1; j < BDF_QMAX; ++j) {
        const bool on = j < S.q;
        const double hs = hsum + S.tau[j + 1];
        const double xi = hs / S.hscale;
        hsum = on ? hs : hsum;
        prod = on ? prod * xi : prod;
        alpha0 = on ? alpha0 - 1.0 / (j + 1) : alpha0;
        alpha1 = on ? alpha1 + 1.0 / xi : alpha1;
#pragma unroll
        for (int i = j + 2; i >= 2; --i) S.l[i] = on ? S.l[i] * xiold + S.l[i - 1] : S.l[i];
        xiold = on ? xi : xiold;
      }
    }
    A1 = (-alpha0 - alpha1) / prod;
    v.rec_order_up(S.q, S.L, S.indx_acor, A1, S.l);
    return 1;
  }
#pragma unroll
  for (int i = 0; i < BDF_QMAX + 2; ++i) S.l[i] = 0.0;
  S.l[2] = 1.0;
  double hsum = 0.0;
#pragma unroll
  for (int j = 1; j <= BDF_QMAX - 2; ++j) {
    const bool on = j <= S.q - 2;
    hsum = on ? hsum + S.tau[j] : hsum;
    const double xi = hsum / S.hscale;
#pragma unroll
    for (int i = j + 2; i >= 2; --i) S.l[i] = on ? S.l[i] * xi + S.l[i - 1] : S.l[i];
  }
  v.rec_order_down(S.q, S.l);
  return -1;
}

// Component part of an order change on one component's history z (rows 0..QMAX), with the
// coefficients l of order_coeffs and the order q and L (= q + 1) before the change. Increase
// (kind 1): z_L = A1 z_ia (za: z_ia as loaded), then z_j += l_j z_L for 2 <= j <= q.
// Decrease (kind -1): z_j -= l_j z_q for 2 <= j < q.
__device__ inline void order_apply(int kind, const double* l, double A1, int q, int L, double za,
                                   double (&z)[BDF_QMAX + 1]) {
  if (kind == 1) {
    const double zl = A1 * za;
#pragma unroll
    for (int j = 2; j <= BDF_QMAX; ++j)
      if (j <= q) z[j] = z[j] + l[j] * zl;
#pragma unroll
    for (int j = 1; j <= BDF_QMAX; ++j)
      if (j == L) z[j] = zl;
  } else if (kind == -1) {
    const double zq = sel(z, q);
#pragma unroll
    for (int j = 2; j < BDF_QMAX; ++j)
      if (j < q) z[j] = z[j] - l[j] * zq;
  }
}

template <int n, class V>
__device__ inline void adjust_order(BdfScalars& S, const V& v, int dq, int qmax) {
  (void)qmax;
  double A1 = 0.0;
  const int kind = order_coeffs(S, v, dq, A1);
  if (!kind) return;
  const int q = S.q, L = S.L, ia = S.indx_acor;
  // rows read: 2..q and ia (increase), 2..q (decrease); written: 2..q and L / 2..q-1
  const int qr = kind == 1 ? (L > q ? L : q) : q;
  constexpr int K = chunk_of<V>();
#pragma unroll
  for (int i0 = v.c0(); i0 < n; i0 += V::kStep * K) {
    double za[K], z[K][BDF_QMAX + 1];
#pragma unroll
    for (int c = 0; c < K; ++c) {
      const int i = i0 + c * V::kStep;
      za[c] = (i < n && kind == 1) ? v.Z(ia, i) : 0.0;
      z[c][0] = z[c][1] = 0.0;
#pragma unroll
      for (int j = 2; j <= BDF_QMAX; ++j) z[c][j] = (i < n && j <= q) ? v.Z(j, i) : 0.0;
    }
#pragma unroll
    for (int c = 0; c < K; ++c) {
      const int i = i0 + c * V::kStep;
      order_apply(kind, S.l, A1, q, L, za[c], z[c]);
#pragma unroll
      for (int j = 2; j <= BDF_QMAX; ++j)
        if (i < n && j <= qr) v.Z(j, i) = z[c][j];
    }
    component_fence<V>();
  }
}

// Start of cvStep: order/step change decided at the end of the previous step.
template <int n, class V>
__device__ inline void step_prologue(BdfScalars& S, const V& v, int qmax) {
  S.saved_t = S.tn;
  S.ncf = S.nef = 0;
  S.nflag = 0;
  if (S.nst > 0 && S.hprime != S.h) {
    if (S.qprime != S.q) {
      adjust_order<n>(S, v, S.qprime - S.q, qmax);
      S.q = S.qprime;
      S.L = S.q + 1;
      S.qwait = S.L;
    }
    rescale<n>(S, v);
  }
}

// Predict, compute BDF coefficients and gamma, decide whether the Newton matrix must be
// (re)built (cvNls). Returns true when a set-up is needed; S.convfail is its argument.
// begin_attempt after the prediction: coefficients, gamma and the set-up decision.
__device__ inline bool attempt_setup(BdfScalars& S) {
  set_bdf(S);
  S.gamma = S.h / S.l[1];
  if (S.nst == 0) S.gammap = S.gamma;
  S.gamrat = (S.nst > 0) ? S.gamma / S.gammap : 1.0;
  S.convfail = (S.nflag == 0 || S.nflag == 2) ? 0 : 1;
  S.m = 0;
  return (S.nflag == 1) || (S.nflag == 2) || (S.nst == 0) || (S.nst >= S.nstlp + cv::MSBP) ||
         (fabs(S.gamrat - 1.0) > cv::DGMAX);
}

template <int n, class V>
__device__ inline bool begin_attempt(BdfScalars& S, const V& v, double tstop) {
  predict<n>(S, v, tstop);
  return attempt_setup(S);
}

// set_ewt, step_prologue and begin_attempt after a successful step, in one pass over the
// components: a component's error weight from z_0, then its order change, rescaling and
// prediction -- the operations of the three calls, in their order, on one component's history
// in registers. Returns begin_attempt's set-up decision.
template <int n, class V>
__device__ inline bool next_attempt(BdfScalars& S, const V& v, const BdfParams& prm,
                                    double tstop) {
  // step_prologue
  S.saved_t = S.tn;
  S.ncf = S.nef = 0;
  S.nflag = 0;
  const int q0 = S.q, L0 = S.L, ia = S.indx_acor;
  int kind = 0;  // order change on the components (order_coeffs)
  double A1 = 0.0;
  const bool resc = S.nst > 0 && S.hprime != S.h;
  if (resc) {
    if (S.qprime != S.q) {
      kind = order_coeffs(S, v, S.qprime - S.q, A1);
      S.q = S.qprime;
      S.L = S.q + 1;
      S.qwait = S.L;
    }
  }
  const int q = S.q;
  double f[BDF_QMAX + 1];
  f[0] = 1.0;
#pragma unroll
  for (int j = 1; j <= BDF_QMAX; ++j) f[j] = f[j - 1] * S.eta;
  if (resc) {
    v.rec_rescale(q, S.eta);
    S.h = S.hscale * S.eta;
    S.hscale = S.h;
  }
  // begin_attempt: predict
  S.tn += S.h;
  if ((S.tn - tstop) * S.h > 0.0) S.tn = tstop;
  v.rec_pascal(q, 1.0);
  const int ql = q0 > q ? q0 : q;  // rows read (an order decrease reads row q0)
  constexpr int K = chunk_of<V>();
#pragma unroll
  for (int i0 = v.c0(); i0 < n; i0 += V::kStep * K) {
    double za[K], z[K][BDF_QMAX + 1];
#pragma unroll
    for (int c = 0; c < K; ++c) {
      const int i = i0 + c * V::kStep;
      za[c] = (i < n && kind == 1) ? v.Z(ia, i) : 0.0;
#pragma unroll
      for (int j = 0; j <= BDF_QMAX; ++j) z[c][j] = (i < n && j <= ql) ? v.Z(j, i) : 0.0;
    }
#pragma unroll
    for (int c = 0; c < K; ++c) {
      const int i = i0 + c * V::kStep;
      const double ew = i < n ? 1.0 / (prm.rtol * fabs(z[c][0]) + prm.atolv[i]) : 0.0;
      order_apply(kind, S.l, A1, q0, L0, za[c], z[c]);
      if (resc) {
#pragma unroll
        for (int j = 1; j <= BDF_QMAX; ++j)
          if (j <= q) z[c][j] = z[c][j] * f[j];
      }
      pascal_apply(z[c], q, 1.0);
      if (i < n) {
        v.E(i) = ew;
#pragma unroll
        for (int j = 0; j <= BDF_QMAX; ++j)
          if (j <= q) v.Z(j, i) = z[c][j];
      }
    }
    component_fence<V>();
  }
  return attempt_setup(S);
}

// Jacobian re-evaluation decision of the dense linear solver set-up (cvDlsSetup).
__device__ inline bool jac_needed(const BdfScalars& S) {
  const double dgamma = fabs(S.gamma / S.gammap - 1.0);
  return (S.nst == 0) || (S.nst > S.nstlj + cv::MSBJ) ||
         (S.convfail == 2 && dgamma < cv::DGMAX_LS) || (S.convfail == 1);
}

// The Jacobian decision of a requested set-up with its bookkeeping (cvDlsSetup: nje, nstlj,
// jcur), taken when the set-up is requested; nothing it reads changes before the set-up runs.
__device__ inline bool jac_decide(BdfScalars& S) {
  const bool need = jac_needed(S);
  if (need) {
    S.nje++;
    S.nstlj = S.nst;
    S.jcur = 1;
  } else {
    S.jcur = 0;
  }
  return need;
}

// Bookkeeping after the Newton matrix was rebuilt.
__device__ inline void after_setup(BdfScalars& S) {
  S.nsetups++;
  S.gamrat = S.crate = 1.0;
  S.gammap = S.gamma;
  S.nstlp = S.nst;
}

// Unrecoverable-in-this-attempt Newton failure (cvHandleNFlag): returns CELL_CONVFAIL or -1 when
// the step is retried with h * ETACF.
template <int n, class V>
__device__ inline int conv_fail(BdfScalars& S, const V& v) {
  S.ncfn++;
  restore<n>(S, v);
  S.ncf++;
  S.etamax = 1.0;
  if (S.ncf == cv::MXNCF) return CELL_CONVFAIL;
  S.eta = cv::ETACF;
  S.nflag = 1;
  rescale<n>(S, v);
  return -1;
}

// Failed local error test (cvDoErrorTest). Returns CELL_ERRFAIL, -1 (retry), or -2 (retry after
// an order-1 restart: the caller must set zn[1] = h * f(tn, zn[0])).
template <int n, class V>
__device__ inline int error_test_fail(BdfScalars& S, const V& v, double dsm, int qmax) {
  S.nef++;
  S.netf++;
  S.nflag = 2;
  restore<n>(S, v);
  if (S.nef == cv::MXNEF) return CELL_ERRFAIL;
  S.etamax = 1.0;
  if (S.nef <= cv::MXNEF1) {
    S.eta = 1.0 / (pow(cv::BIAS2 * dsm, 1.0 / S.L) + cv::ADDON);
    S.eta = fmax(cv::ETAMIN, S.eta);
    if (S.nef >= cv::SMALL_NEF) S.eta = fmin(S.eta, cv::ETAMXF);
    rescale<n>(S, v);
    return -1;
  }
  if (S.q > 1) {
    S.eta = cv::ETAMIN;
    adjust_order<n>(S, v, -1, qmax);
    S.L = S.q;
    S.q--;
    S.qwait = S.L;
    rescale<n>(S, v);
    return -1;
  }
  S.eta = cv::ETAMIN;
  S.h *= S.eta;
  S.hscale = S.h;
  S.qwait = cv::LONG_WAIT;
  return -2;
}

__device__ inline void set_eta(BdfScalars& S) {
  if (S.eta < cv::THRESH) {
    S.eta = 1.0;
    S.hprime = S.h;
  } else {
    S.eta = fmin(S.eta, S.etamax);
    S.hprime = S.h * S.eta;
  }
}

// cvCompleteStep + cvPrepareNextStep + etamax update after a successful step. The Nordsieck
// update and the two norms of the order-change test (taken when qwait reaches 0: z_q after the
// update, and acor against the saved correction in column qmax) are one pass over the components;
// each norm is accumulated in component order, as a separate loop would.
template <int n, class V>
__device__ inline void finish_step(BdfScalars& S, const V& v, double dsm, int qmax) {
  S.nst++;
#pragma unroll
  for (int i = BDF_QMAX; i >= 2; --i) S.tau[i] = (i <= S.q) ? S.tau[i - 1] : S.tau[i];
  if (S.q == 1 && S.nst > 1) S.tau[2] = S.tau[1];
  S.tau[1] = S.h;
  S.qwait--;
  const bool save_acor = S.qwait == 1 && S.q != qmax;
  v.rec_correct(S, save_acor, qmax);
  const int q = S.q;
  // cvPrepareNextStep's order-change test, when it runs (saved_tq5 is not touched in between:
  // save_acor needs qwait == 1, the test qwait == 0)
  const bool test = S.etamax != 1.0 && S.qwait == 0;
  const bool down = test && q > 1;
  const bool up = test && q != qmax && S.saved_tq5 != 0.0;
  const double cquot = up ? (S.tq[5] / S.saved_tq5) * pow(S.h / S.tau[2], (double)S.L) : 0.0;
  double sdn = 0.0, sup = 0.0;
  {
    constexpr int K = chunk_of<V>();
#pragma unroll
    for (int i0 = v.c0(); i0 < n; i0 += V::kStep * K) {
      double a[K], e[K], zm[K], z[K][BDF_QMAX + 1];
#pragma unroll
      for (int c = 0; c < K; ++c) {
        const int i = i0 + c * V::kStep;
        a[c] = i < n ? v.A(i) : 0.0;
        e[c] = (i < n && test) ? v.E(i) : 0.0;
        zm[c] = (i < n && up) ? v.Z(qmax, i) : 0.0;
        if (i < n) {
          load_hist(v, i, q, z[c]);
        } else {
#pragma unroll
          for (int j = 0; j <= BDF_QMAX; ++j) z[c][j] = 0.0;
        }
      }
#pragma unroll
      for (int c = 0; c < K; ++c) {
        const int i = i0 + c * V::kStep;
        if (i < n) {
#pragma unroll
          for (int j = 0; j <= BDF_QMAX; ++j)
            if (j <= q) z[c][j] = z[c][j] + S.l[j] * a[c];
          if (down) {
            const double w = sel(z[c], q) * e[c];
            sdn += w * w;
          }
          if (up) {
            const double w = (a[c] - cquot * zm[c]) * e[c];
            sup += w * w;
          }
#pragma unroll
          for (int j = 0; j <= BDF_QMAX; ++j)
            if (j <= q) v.Z(j, i) = z[c][j];
          if (save_acor) v.Z(qmax, i) = a[c];
        }
      }
      component_fence<V>();
    }
  }
  if (save_acor) {
    S.saved_tq5 = S.tq[5];
    S.indx_acor = qmax;
  }
  // cvPrepareNextStep
  if (S.etamax == 1.0) {
    S.qwait = S.qwait > 2 ? S.qwait : 2;
    S.qprime = S.q;
    S.hprime = S.h;
    S.eta = 1.0;
  } else {
    const double etaq = 1.0 / (pow(cv::BIAS2 * dsm, 1.0 / S.L) + cv::ADDON);
    if (S.qwait != 0) {
      S.eta = etaq;
      S.qprime = S.q;
      set_eta(S);
    } else {
      S.qwait = 2;
      double etaqm1 = 0.0;
      if (down) {
        const double ddn = sqrt(v.sum(sdn) / n) * S.tq[1];
        etaqm1 = 1.0 / (pow(cv::BIAS1 * ddn, 1.0 / S.q) + cv::ADDON);
      }
      double etaqp1 = 0.0;
      if (up) {
        const double dup = sqrt(v.sum(sup) / n) * S.tq[3];
        etaqp1 = 1.0 / (pow(cv::BIAS3 * dup, 1.0 / (S.L + 1)) + cv::ADDON);
      }
      const double etam = fmax(etaqm1, fmax(etaq, etaqp1));
      if (etam < cv::THRESH) {
        S.eta = 1.0;
        S.qprime = S.q;
      } else if (etam == etaq) {
        S.eta = etaq;
        S.qprime = S.q;
      } else if (etam == etaqm1) {
        S.eta = etaqm1;
        S.qprime = S.q - 1;
      } else {
        S.eta = etaqp1;
        S.qprime = S.q + 1;
        v.rec_mark_save();  // the correction just applied is also saved in column qmax
        constexpr int K = chunk_of<V>();
#pragma unroll
        for (int i0 = v.c0(); i0 < n; i0 += V::kStep * K) {
          double a[K];
#pragma unroll
          for (int c = 0; c < K; ++c) {
            const int i = i0 + c * V::kStep;
            a[c] = i < n ? v.A(i) : 0.0;
          }
#pragma unroll
          for (int c = 0; c < K; ++c) {
            const int i = i0 + c * V::kStep;
            if (i < n) v.Z(qmax, i) = a[c];
          }
          component_fence<V>();
        }
      }
      set_eta(S);
    }
  }
  S.etamax = (S.nst <= cv::SMALL_NST) ? cv::ETAMX2 : cv::ETAMX3;
}

// After a successful step: true when tn has reached tstop; otherwise clips the next step so it
// does not pass tstop.
__device__ inline bool reached_tstop(BdfScalars& S, double tstop) {
  const double troundoff = 100.0 * cv::UROUND * (fabs(S.tn) + fabs(S.h));
  if (fabs(S.tn - tstop) <= troundoff) return true;
  if ((S.tn + S.hprime - tstop) * S.h > 0.0) {
    S.hprime = (tstop - S.tn) * (1.0 - 4.0 * cv::UROUND);
    S.eta = S.hprime / S.h;
  }
  return false;
}

// cvHin: initial step size from an estimate of ||y''||; zn[1] holds f(t0, y0) on entry.
// f(t, y, out) evaluates the RHS (and counts it).
template <int n, class V, class F>
__device__ inline double hin(BdfScalars& S, const V& v, double tout, F&& f) {
  const double tdist = tout - S.tn;
  const double tround = cv::UROUND * fmax(fabs(S.tn), fabs(tout));
  const double hlb = 100.0 * tround;
  double hub_inv = 0.0;
#pragma unroll
  for (int i = v.c0(); i < n; i += V::kStep) {
    const double d = 0.1 * fabs(v.Z(0, i)) + 1.0 / v.E(i);
    hub_inv = fmax(hub_inv, fabs(v.Z(1, i)) / d);
  }
  hub_inv = v.max(hub_inv);
  double hub = 0.1 * tdist;
  if (hub * hub_inv > 1.0) hub = 1.0 / hub_inv;
  double hg = sqrt(hlb * hub);
  if (hub < hlb) return hg;
  double hnew = hg;
  typename V::template Arr<n> y, t;
  for (int count1 = 1; count1 <= 4; ++count1) {
#pragma unroll
    for (int i = v.c0(); i < n; i += V::kStep) y[i] = v.Z(0, i) + hg * v.Z(1, i);
    f(S.tn + hg, y, t);
#pragma unroll
    for (int i = v.c0(); i < n; i += V::kStep) t[i] = (t[i] - v.Z(1, i)) / hg;
    const double yddnrm = wrms_reg<n>(v, t);
    hnew = (yddnrm * hub * hub > 2.0) ? sqrt(2.0 / yddnrm) : sqrt(hg * hub);
    if (count1 == 4) break;
    const double hrat = hnew / hg;
    if (hrat > 0.5 && hrat < 2.0) break;
    if (count1 > 1 && hrat > 2.0) {
      hnew = hg;
      break;
    }
    hg = hnew;
  }
  return fmin(fmax(0.5 * hnew, hlb), hub);
}

// One Newton correction (after f(tn, y) is in ftemp): tempv = M^{-1}(gamma f - rl1 zn1 - acor),
// scaled by 2/(1+gamrat); acor += tempv. Returns 0 converged, 1 continue, 2 failed (retry with a
// fresh Jacobian when it is stale), 3 failed with a current Jacobian. The solve is a callable.
template <int n, class V, class FT, class Solve>
__device__ inline int newton_correct(BdfScalars& S, const V& v, const FT& ftemp, Solve&& solve) {
  typename V::template Arr<n> tempv;
  const double rl1 = 1.0 / S.l[1];
#pragma unroll
  for (int i = v.c0(); i < n; i += V::kStep) tempv[i] = S.gamma * ftemp[i] - (rl1 * v.Z(1, i) + v.A(i));
  solve(tempv);
  if (S.gamrat != 1.0) {
    const double sc = 2.0 / (1.0 + S.gamrat);
#pragma unroll
    for (int i = v.c0(); i < n; i += V::kStep) tempv[i] *= sc;
  }
  S.nni++;
  // one pass: the WRMS norm of the correction (del), acor += correction, and the WRMS norm of the
  // updated acor (cvNlsNewton's acnrm), each sum in component order
  double sdel = 0.0, sacor = 0.0;
  {
    constexpr int K = chunk_of<V>();
#pragma unroll
    for (int i0 = v.c0(); i0 < n; i0 += V::kStep * K) {
      double a[K], e[K];
#pragma unroll
      for (int c = 0; c < K; ++c) {
        const int i = i0 + c * V::kStep;
        a[c] = i < n ? v.A(i) : 0.0;
        e[c] = i < n ? v.E(i) : 0.0;
      }
#pragma unroll
      for (int c = 0; c < K; ++c) {
        const int i = i0 + c * V::kStep;
        if (i < n) {
          const double w = tempv[i] * e[c];
          sdel += w * w;
          const double an = a[c] + tempv[i];
          v.A(i) = an;
          const double wa = an * e[c];
          sacor += wa * wa;
        }
      }
      component_fence<V>();
    }
  }
  const double del = sqrt(v.sum(sdel) / n);
  const double acn = sqrt(v.sum(sacor) / n);
  if (S.m > 0) S.crate = fmax(cv::CRDOWN * S.crate, del / S.delp);
  const double dcon = del * fmin(1.0, S.crate) / S.tq[4];
  if (dcon <= 1.0) {
    S.acnrm = (S.m == 0) ? del : acn;
    S.jcur = 0;
    return 0;
  }
  S.m++;
  if (S.m == cv::MAXCOR || (S.m >= 2 && del > cv::RDIV * S.delp)) return S.jcur ? 3 : 2;
  S.delp = del;
  return 1;
}

}  // namespace pelec
