// Per-cell variable-order BDF integrator (device code), the integration core of ReactorCvode
// (reference: Source/React.cpp:222-228 -> ReactorCvode::react; SUNDIALS CVODE BDF with the dense
// analytic-Jacobian linear solver, cvode.solve_type = denseAJ_direct, Exec/Production/JetFlame/
// inputs.inp:30). The algorithm is CVODE's published one:
//   * Nordsieck history zn[0..q], orders 1..qmax (cvode.max_order), fixed-leading-coefficient
//     BDF coefficients l[], error-test constants tq[] (cvSetBDF / cvSetTqBDF);
//   * WRMS norm with ewt_i = 1/(rtol |y_i| + atol_i), atol_i = atol * typ_vals_i;
//   * modified Newton, at most 3 corrections, convergence rate test with CRDOWN = 0.3, divergence
//     test RDIV = 2, Newton matrix I - gamma J rebuilt when |gamma/gammap - 1| > 0.3, every 20
//     steps, or after failures, Jacobian re-evaluated every 50 steps or on a bad-J failure;
//   * local error test dsm = ||acor|| tq[2] <= 1 with order/step selection (etaq, etaqm1, etaqp1,
//     THRESH 1.5, ETAMX 10 / 10000 on the first step, error-failure and convergence-failure
//     step reductions, order reduction after 3 error failures);
//   * initial step from the y'' estimate (cvHin), steps clipped to land exactly on dt_react.
// Every decision is made per cell: cells of uneven stiffness take their own step sequences.
//
// LUT is the element type of the saved Jacobian and of the LU factors of I - gamma J: double for
// the plain FP64 form, float for the mixed-precision form (residuals, rates, T(e) and the error
// control stay in double, so the inexact Newton iteration converges to the FP64 solution).
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
};

struct BdfParams {
  double rtol;
  const double* atolv;  // n entries (device, shared by all cells)
  int qmax;
  int max_steps;
};

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

// Solve (LU) x = b in place; b is double, arithmetic in T.
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

template <class Mech, typename LUT>
struct BdfCell {
  using Model = CellModel<Mech>;
  static constexpr int N = Mech::NSP;
  static constexpr int n = N + 1;
  static constexpr double ETAMX1 = 10000.0, ETAMX2 = 10.0, ETAMX3 = 10.0, ETAMXF = 0.2,
                          ETAMIN = 0.1, ETACF = 0.25, ADDON = 1e-6, BIAS1 = 6.0, BIAS2 = 6.0,
                          BIAS3 = 10.0, THRESH = 1.5, CRDOWN = 0.3, DGMAX = 0.3, RDIV = 2.0,
                          NLSCOEF = 0.1, DGMAX_LS = 0.2, UROUND = 2.220446049250313e-16;
  static constexpr int MXNEF1 = 3, SMALL_NEF = 2, LONG_WAIT = 10, MAXCOR = 3, MXNCF = 10,
                       MXNEF = 7, MSBP = 20, MSBJ = 50, SMALL_NST = 10;

  // problem
  double rhoe0, rhoesrc;
  const double* src;
  int64_t sstride;
  BdfParams prm;
  double tstop;
  // state
  double zn[BDF_QMAX + 1][n];
  double ewt[n], acor[n], y[n], ftemp[n], tempv[n];
  LUT savedJ[n * n];
  LUT M[n * n];
  int piv[n];
  double tn, h, hscale, hprime, eta, etamax, saved_tq5;
  double gamma, gammap, gamrat, crate, acnrm;
  double tau[BDF_QMAX + 2], tq[6], l[BDF_QMAX + 1];
  int q, qprime, qwait, L, nstlp, nstlj, indx_acor;
  bool jcur;
  // counters
  int nst, nfe, nje, nsetups, nni, netf, ncfn;

  __device__ double wrms(const double* v) const {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < n; ++i) {
      const double w = v[i] * ewt[i];
      s += w * w;
    }
    return sqrt(s / n);
  }
  __device__ void set_ewt() {
#pragma unroll
    for (int i = 0; i < n; ++i) ewt[i] = 1.0 / (prm.rtol * fabs(zn[0][i]) + prm.atolv[i]);
  }
  __device__ void f(double t, const double* yy, double* out) {
    nfe++;
    Model::rhs(t, yy, rhoe0, rhoesrc, src, sstride, out);
  }

  __device__ double hin(double tout) {
    const double tdist = tout - tn;
    const double tround = UROUND * fmax(fabs(tn), fabs(tout));
    const double hlb = 100.0 * tround;
    double hub_inv = 0.0;
#pragma unroll
    for (int i = 0; i < n; ++i) {
      const double d = 0.1 * fabs(zn[0][i]) + 1.0 / ewt[i];
      hub_inv = fmax(hub_inv, fabs(zn[1][i]) / d);
    }
    double hub = 0.1 * tdist;
    if (hub * hub_inv > 1.0) hub = 1.0 / hub_inv;
    double hg = sqrt(hlb * hub);
    if (hub < hlb) return hg;
    double hnew = hg;
    for (int count1 = 1; count1 <= 4; ++count1) {
#pragma unroll
      for (int i = 0; i < n; ++i) y[i] = zn[0][i] + hg * zn[1][i];
      f(tn + hg, y, tempv);
#pragma unroll
      for (int i = 0; i < n; ++i) tempv[i] = (tempv[i] - zn[1][i]) / hg;
      const double yddnrm = wrms(tempv);
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
    double h0 = 0.5 * hnew;
    h0 = fmin(fmax(h0, hlb), hub);
    return h0;
  }

  __device__ void rescale() {
    double factor = eta;
    for (int j = 1; j <= q; ++j) {
#pragma unroll
      for (int i = 0; i < n; ++i) zn[j][i] *= factor;
      factor *= eta;
    }
    h = hscale * eta;
    hscale = h;
  }
  __device__ void predict() {
    tn += h;
    if ((tn - tstop) * h > 0.0) tn = tstop;
    for (int k = 1; k <= q; ++k)
      for (int j = q; j >= k; --j)
#pragma unroll
        for (int i = 0; i < n; ++i) zn[j - 1][i] += zn[j][i];
  }
  __device__ void restore(double saved_t) {
    tn = saved_t;
    for (int k = 1; k <= q; ++k)
      for (int j = q; j >= k; --j)
#pragma unroll
        for (int i = 0; i < n; ++i) zn[j - 1][i] -= zn[j][i];
  }
  __device__ void set_bdf() {
    double alpha0, alpha0_hat, xi_inv, xistar_inv, hsum;
    l[0] = l[1] = xi_inv = xistar_inv = 1.0;
    for (int i = 2; i <= q; ++i) l[i] = 0.0;
    alpha0 = alpha0_hat = -1.0;
    hsum = h;
    if (q > 1) {
      for (int j = 2; j < q; ++j) {
        hsum += tau[j - 1];
        xi_inv = h / hsum;
        alpha0 -= 1.0 / j;
        for (int i = j; i >= 1; --i) l[i] += l[i - 1] * xi_inv;
      }
      alpha0 -= 1.0 / q;
      xistar_inv = -l[1] - alpha0;
      hsum += tau[q - 1];
      xi_inv = h / hsum;
      alpha0_hat = -l[1] - xi_inv;
      for (int i = q; i >= 1; --i) l[i] += l[i - 1] * xistar_inv;
    }
    const double A1 = 1.0 - alpha0_hat + alpha0;
    const double A2 = 1.0 + q * A1;
    tq[2] = fabs(A1 / (alpha0 * A2));
    tq[5] = fabs(A2 * xistar_inv / (l[q] * xi_inv));
    if (qwait == 1) {
      if (q > 1) {
        const double Cc = xistar_inv / l[q];
        const double A3 = alpha0 + 1.0 / q;
        const double A4 = alpha0_hat + xi_inv;
        const double Cpinv = (1.0 - A4 + A3) / A3;
        tq[1] = fabs(Cc * Cpinv);
      } else {
        tq[1] = 1.0;
      }
      hsum += tau[q];
      xi_inv = h / hsum;
      const double A5 = alpha0 - 1.0 / (q + 1);
      const double A6 = alpha0_hat - xi_inv;
      const double Cppinv = (1.0 - A6 + A5) / A2;
      tq[3] = fabs(Cppinv / (xi_inv * (q + 2) * A5));
    }
    tq[4] = NLSCOEF / tq[2];
  }
  __device__ void increase_bdf() {
    for (int i = 0; i <= prm.qmax; ++i) l[i] = 0.0;
    double alpha1 = 1.0, prod = 1.0, xiold = 1.0, alpha0 = -1.0, hsum = hscale;
    l[2] = 1.0;
    if (q > 1) {
      for (int j = 1; j < q; ++j) {
        hsum += tau[j + 1];
        const double xi = hsum / hscale;
        prod *= xi;
        alpha0 -= 1.0 / (j + 1);
        alpha1 += 1.0 / xi;
        for (int i = j + 2; i >= 2; --i) l[i] = l[i] * xiold + l[i - 1];
        xiold = xi;
      }
    }
    const double A1 = (-alpha0 - alpha1) / prod;
#pragma unroll
    for (int i = 0; i < n; ++i) zn[L][i] = A1 * zn[indx_acor][i];
    for (int j = 2; j <= q; ++j)
#pragma unroll
      for (int i = 0; i < n; ++i) zn[j][i] += l[j] * zn[L][i];
  }
  __device__ void decrease_bdf() {
    for (int i = 0; i <= prm.qmax; ++i) l[i] = 0.0;
    l[2] = 1.0;
    double hsum = 0.0;
    for (int j = 1; j <= q - 2; ++j) {
      hsum += tau[j];
      const double xi = hsum / hscale;
      for (int i = j + 2; i >= 2; --i) l[i] = l[i] * xi + l[i - 1];
    }
    for (int j = 2; j < q; ++j)
#pragma unroll
      for (int i = 0; i < n; ++i) zn[j][i] -= l[j] * zn[q][i];
  }
  __device__ void adjust_order(int dq) {
    if (q == 2 && dq != 1) return;
    if (dq == 1)
      increase_bdf();
    else
      decrease_bdf();
  }

  __device__ int lsetup(int convfail) {
    const double dgamma = fabs(gamma / gammap - 1.0);
    const bool jbad = (nst == 0) || (nst > nstlj + MSBJ) || (convfail == 2 && dgamma < DGMAX_LS) ||
                      (convfail == 1);
    if (!jbad) {
      jcur = false;
    } else {
      nje++;
      nstlj = nst;
      jcur = true;
      Model::template jac<LUT>(tn, zn[0], rhoe0, rhoesrc, src, sstride, savedJ);
    }
    const LUT mg = LUT(-gamma);
    for (int i = 0; i < n * n; ++i) M[i] = mg * savedJ[i];
#pragma unroll
    for (int i = 0; i < n; ++i) M[i * n + i] += LUT(1);
    return lu_factor<n, LUT>(M, piv);
  }

  // returns 0 converged, 1 recoverable failure
  __device__ int nls(int nflag) {
    int convfail = (nflag == 0 || nflag == 2) ? 0 : 1;
    bool callSetup = (nflag == 1) || (nflag == 2) || (nst == 0) || (nst >= nstlp + MSBP) ||
                     (fabs(gamrat - 1.0) > DGMAX);
    for (;;) {
      f(tn, zn[0], ftemp);
      if (callSetup) {
        const int ier = lsetup(convfail);
        nsetups++;
        callSetup = false;
        gamrat = crate = 1.0;
        gammap = gamma;
        nstlp = nst;
        if (ier != 0) return 1;
      }
#pragma unroll
      for (int i = 0; i < n; ++i) {
        acor[i] = 0.0;
        y[i] = zn[0][i];
      }
      int m = 0;
      double del = 0.0, delp = 0.0;
      const double rl1 = 1.0 / l[1];
      int ret;
      for (;;) {
#pragma unroll
        for (int i = 0; i < n; ++i) tempv[i] = gamma * ftemp[i] - (rl1 * zn[1][i] + acor[i]);
        lu_solve<n, LUT>(M, piv, tempv);
        if (gamrat != 1.0) {
          const double sc = 2.0 / (1.0 + gamrat);
#pragma unroll
          for (int i = 0; i < n; ++i) tempv[i] *= sc;
        }
        nni++;
        del = wrms(tempv);
#pragma unroll
        for (int i = 0; i < n; ++i) {
          acor[i] += tempv[i];
          y[i] = zn[0][i] + acor[i];
        }
        if (m > 0) crate = fmax(CRDOWN * crate, del / delp);
        const double dcon = del * fmin(1.0, crate) / tq[4];
        if (dcon <= 1.0) {
          acnrm = (m == 0) ? del : wrms(acor);
          jcur = false;
          ret = 0;
          break;
        }
        ++m;
        if (m == MAXCOR || (m >= 2 && del > RDIV * delp)) {
          ret = jcur ? 1 : 2;
          break;
        }
        delp = del;
        f(tn, y, ftemp);
      }
      if (ret != 2) return ret;
      callSetup = true;
      convfail = 2;
    }
  }

  __device__ void complete_step() {
    nst++;
    for (int i = q; i >= 2; --i) tau[i] = tau[i - 1];
    if (q == 1 && nst > 1) tau[2] = tau[1];
    tau[1] = h;
    for (int j = 0; j <= q; ++j)
#pragma unroll
      for (int i = 0; i < n; ++i) zn[j][i] += l[j] * acor[i];
    qwait--;
    if (qwait == 1 && q != prm.qmax) {
#pragma unroll
      for (int i = 0; i < n; ++i) zn[prm.qmax][i] = acor[i];
      saved_tq5 = tq[5];
      indx_acor = prm.qmax;
    }
  }
  __device__ void set_eta() {
    if (eta < THRESH) {
      eta = 1.0;
      hprime = h;
    } else {
      eta = fmin(eta, etamax);
      hprime = h * eta;
    }
  }
  __device__ void prepare_next_step(double dsm) {
    if (etamax == 1.0) {
      qwait = qwait > 2 ? qwait : 2;
      qprime = q;
      hprime = h;
      eta = 1.0;
      return;
    }
    const double etaq = 1.0 / (pow(BIAS2 * dsm, 1.0 / L) + ADDON);
    if (qwait != 0) {
      eta = etaq;
      qprime = q;
      set_eta();
      return;
    }
    qwait = 2;
    double etaqm1 = 0.0;
    if (q > 1) {
      const double ddn = wrms(zn[q]) * tq[1];
      etaqm1 = 1.0 / (pow(BIAS1 * ddn, 1.0 / q) + ADDON);
    }
    double etaqp1 = 0.0;
    if (q != prm.qmax && saved_tq5 != 0.0) {
      const double cquot = (tq[5] / saved_tq5) * pow(h / tau[2], (double)L);
#pragma unroll
      for (int i = 0; i < n; ++i) tempv[i] = acor[i] - cquot * zn[prm.qmax][i];
      const double dup = wrms(tempv) * tq[3];
      etaqp1 = 1.0 / (pow(BIAS3 * dup, 1.0 / (L + 1)) + ADDON);
    }
    const double etam = fmax(etaqm1, fmax(etaq, etaqp1));
    if (etam < THRESH) {
      eta = 1.0;
      qprime = q;
    } else if (etam == etaq) {
      eta = etaq;
      qprime = q;
    } else if (etam == etaqm1) {
      eta = etaqm1;
      qprime = q - 1;
    } else {
      eta = etaqp1;
      qprime = q + 1;
#pragma unroll
      for (int i = 0; i < n; ++i) zn[prm.qmax][i] = acor[i];
    }
    set_eta();
  }

  // One CVODE step (cvStep). Returns CELL_OK or a failure status.
  __device__ int step() {
    const double saved_t = tn;
    int ncf = 0, nef = 0, nflag = 0;
    if (nst > 0 && hprime != h) {
      if (qprime != q) {
        adjust_order(qprime - q);
        q = qprime;
        L = q + 1;
        qwait = L;
      }
      rescale();
    }
    double dsm = 0.0;
    for (;;) {
      predict();
      set_bdf();
      gamma = h / l[1];
      if (nst == 0) gammap = gamma;
      gamrat = (nst > 0) ? gamma / gammap : 1.0;
      const int nret = nls(nflag);
      if (nret != 0) {
        ncfn++;
        restore(saved_t);
        ncf++;
        etamax = 1.0;
        if (ncf == MXNCF) return CELL_CONVFAIL;
        eta = ETACF;
        nflag = 1;
        rescale();
        continue;
      }
      dsm = acnrm * tq[2];
      if (dsm <= 1.0) break;
      nef++;
      netf++;
      nflag = 2;
      restore(saved_t);
      if (nef == MXNEF) return CELL_ERRFAIL;
      etamax = 1.0;
      if (nef <= MXNEF1) {
        eta = 1.0 / (pow(BIAS2 * dsm, 1.0 / L) + ADDON);
        eta = fmax(ETAMIN, eta);
        if (nef >= SMALL_NEF) eta = fmin(eta, ETAMXF);
        rescale();
        continue;
      }
      if (q > 1) {
        eta = ETAMIN;
        adjust_order(-1);
        L = q;
        q--;
        qwait = L;
        rescale();
        continue;
      }
      eta = ETAMIN;
      h *= eta;
      hscale = h;
      qwait = LONG_WAIT;
      f(tn, zn[0], tempv);
#pragma unroll
      for (int i = 0; i < n; ++i) zn[1][i] = h * tempv[i];
    }
    complete_step();
    prepare_next_step(dsm);
    etamax = (nst <= SMALL_NST) ? ETAMX2 : ETAMX3;
    return CELL_OK;
  }

  // Set up the integration of y0 over [0, tout] (first RHS, initial step, Nordsieck array).
  __device__ void start(const double* y0, double tout) {
    nst = nfe = nje = nsetups = nni = netf = ncfn = 0;
    tn = 0.0;
    tstop = tout;
    q = 1;
    L = 2;
    qwait = L;
    etamax = ETAMX1;
    indx_acor = prm.qmax;
    saved_tq5 = 0.0;
    nstlp = nstlj = 0;
    jcur = false;
    gamma = gammap = 0.0;
    gamrat = crate = 1.0;
    acnrm = 0.0;
    eta = 1.0;
#pragma unroll
    for (int i = 0; i < BDF_QMAX + 2; ++i) tau[i] = 0.0;
#pragma unroll
    for (int i = 0; i < 6; ++i) tq[i] = 0.0;
#pragma unroll
    for (int i = 0; i < n; ++i) zn[0][i] = y0[i];
    set_ewt();
    f(tn, zn[0], zn[1]);
    h = hin(tout);
    if ((tn + h - tout) * h > 0.0) h = (tout - tn) * (1.0 - 4.0 * UROUND);
    hscale = h;
    hprime = h;
#pragma unroll
    for (int i = 0; i < n; ++i) zn[1][i] *= h;
  }

  // Advance one step toward tout. Returns <0 while running, CELL_OK when tn == tout, or a
  // failure status.
  __device__ int advance() {
    if (nst >= prm.max_steps) return CELL_MAXSTEPS;
    set_ewt();
    const int s = step();
    if (s != CELL_OK) return s;
    const double troundoff = 100.0 * UROUND * (fabs(tn) + fabs(h));
    if (fabs(tn - tstop) <= troundoff) return CELL_OK;
    if ((tn + hprime - tstop) * h > 0.0) {
      hprime = (tstop - tn) * (1.0 - 4.0 * UROUND);
      eta = hprime / h;
    }
    return -1;
  }
};

}  // namespace pelec
