// ORACLE -- test infrastructure only. CPU FP64 restatement of the reference hot path
// (PeleC::react_state -> ReactorBase::react, Source/React.cpp:222-228) used by tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker. It is never linked
// into, loaded by, or called from the product library.
//
// What it restates (the bodies are in the absent PelePhysics/SUNDIALS submodules; file list at
// CMake/BuildPelePhysicsLib.cmake:96-142):
//   * NASA-7 thermo, e_k = h_k - R T / W_k (Docs/sphinx/Equations.rst:36-37)
//   * T(e, Y) Newton inversion (EOS REY2T, used at Source/Utilities.H:81-82)
//   * mass-action production rates with third-body, Lindemann and Troe falloff terms and reverse
//     rates from equilibrium constants (mechanism productionRate / CKWC)
//   * the exact analytic Jacobian of the rho*e reactor RHS (T eliminated through e(T) = e(t))
//   * CVODE's variable-order (1..5) variable-step fixed-leading-coefficient BDF in Nordsieck
//     form with WRMS error control, modified Newton (maxcor 3, crate test) and a dense
//     partial-pivot LU (SUNDIALS cvode.c / cvode_ls.c / sunlinsol_dense published algorithm),
//     with the step clipped to land on t = dt_react (CVODE's tstop handling).
//   * the pack/unpack conventions of React.cpp: rho*e(t) = rho*e0 + t * rhoe_src (React.cpp:292),
//     masked cells (mask == 0) returned untouched with FC = 0.
//
// Every function here is pinned by a test in tests/ (golden Cantera trajectories, scipy Radau
// solutions, invariants, torch autograd and finite differences).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double R_ERG = 8.31446261815324e7;
constexpr double P_ATM = 1013250.0;
constexpr double LN10 = 2.302585092994046;

struct Rxn {
  int kind = 0;  // 0 elementary, 1 three-body, 2 falloff
  bool rev = true;
  double A = 0, b = 0, EaR = 0;
  std::vector<std::pair<int, double>> re, pr;
  std::vector<double> eff;  // N entries for kind 1/2
  double A0 = 0, b0 = 0, E0 = 0;
  int ntroe = 0;
  double troe[4] = {0, 0, 0, 0};
  double dnu = 0;  // sum of product minus reactant coefficients
};

struct Mech {
  int N = 0, R = 0;
  std::vector<double> W, Tmid, low, high;
  std::vector<Rxn> rx;
  std::vector<double> mult;  // optional per-reaction rate multipliers (empty: 1)
};

// ---------------------------------------------------------------- thermo
inline const double* coef(const Mech& m, int k, double T) {
  return (T < m.Tmid[k]) ? &m.low[7 * k] : &m.high[7 * k];
}

void thermo(const Mech& m, double T, double* cpR, double* hRT, double* sR, double* dcpR) {
  const double lnT = std::log(T);
  for (int k = 0; k < m.N; ++k) {
    const double* a = coef(m, k, T);
    if (cpR) cpR[k] = a[0] + T * (a[1] + T * (a[2] + T * (a[3] + T * a[4])));
    if (hRT)
      hRT[k] = a[0] + T * (a[1] / 2 + T * (a[2] / 3 + T * (a[3] / 4 + T * a[4] / 5))) + a[5] / T;
    if (sR) sR[k] = a[0] * lnT + T * (a[1] + T * (a[2] / 2 + T * (a[3] / 3 + T * a[4] / 4))) + a[6];
    if (dcpR) dcpR[k] = a[1] + T * (2 * a[2] + T * (3 * a[3] + T * 4 * a[4]));
  }
}

// mass-specific internal energy and cv of each species
void e_cv(const Mech& m, double T, double* e, double* cv) {
  for (int k = 0; k < m.N; ++k) {
    const double* a = coef(m, k, T);
    const double cpR = a[0] + T * (a[1] + T * (a[2] + T * (a[3] + T * a[4])));
    const double hRT = a[0] + T * (a[1] / 2 + T * (a[2] / 3 + T * (a[3] / 4 + T * a[4] / 5))) + a[5] / T;
    e[k] = (hRT - 1.0) * R_ERG * T / m.W[k];
    cv[k] = (cpR - 1.0) * R_ERG / m.W[k];
  }
}

// Newton inversion of sum_k Y_k e_k(T) = e; |dT| <= 1e-12 T, at most 20 iterations.
double T_from_e(const Mech& m, double e, const double* Y, double T, int* iters) {
  std::vector<double> ek(m.N), cvk(m.N);
  int it = 0;
  for (; it < 20; ++it) {
    e_cv(m, T, ek.data(), cvk.data());
    double f = -e, d = 0;
    for (int k = 0; k < m.N; ++k) {
      f += Y[k] * ek[k];
      d += Y[k] * cvk[k];
    }
    const double dT = -f / d;
    T += dT;
    if (std::fabs(dT) <= 1e-12 * T) {
      ++it;
      break;
    }
  }
  if (iters) *iters = it;
  return T;
}

// ---------------------------------------------------------------- kinetics
// Rates of progress q_j and, when dq is given, dq_j/dC_m (R x N, row-major) and dq_j/dT.
void progress(const Mech& m, const double* C, double T, double* q, double* dqdC, double* dqdT) {
  const int N = m.N;
  std::vector<double> hRT(N), sR(N);
  thermo(m, T, nullptr, hRT.data(), sR.data(), nullptr);
  const double lnT = std::log(T), invT = 1.0 / T;
  const double lnPRT = std::log(P_ATM / (R_ERG * T));
  if (dqdC) std::fill(dqdC, dqdC + (size_t)m.R * N, 0.0);
  for (int j = 0; j < m.R; ++j) {
    const Rxn& r = m.rx[j];
    double kf = r.A * std::exp(r.b * lnT - r.EaR * invT);
    double dlnkf = (r.b + r.EaR * invT) * invT;
    double sumg = 0, sumh = 0;
    for (auto& p : r.re) {
      sumg -= p.second * (hRT[p.first] - sR[p.first]);
      sumh -= p.second * hRT[p.first];
    }
    for (auto& p : r.pr) {
      sumg += p.second * (hRT[p.first] - sR[p.first]);
      sumh += p.second * hRT[p.first];
    }
    double Kc = 0, dlnKc = 0, invKc = 0;
    if (r.rev) {
      const double lnKc = -sumg + r.dnu * lnPRT;
      Kc = std::exp(lnKc);
      invKc = 1.0 / Kc;
      dlnKc = (sumh - r.dnu) * invT;
    }
    double Pf = 1.0, Pr = 1.0;
    for (auto& p : r.re) Pf *= std::pow(C[p.first], p.second);
    for (auto& p : r.pr) Pr *= std::pow(C[p.first], p.second);
    if (!r.rev) Pr = 0.0;
    double M = 0;
    if (r.kind > 0)
      for (int k = 0; k < N; ++k) M += r.eff[k] * C[k];
    // k: effective forward coefficient; dk_dM and dlnk_dT for the Jacobian
    double keff = kf, dkdM = 0, dlnkdT = dlnkf;
    if (r.kind == 1) {
      keff = kf * M;
      dkdM = kf;
    } else if (r.kind == 2) {
      const double k0 = r.A0 * std::exp(r.b0 * lnT - r.E0 * invT);
      const double dlnk0 = (r.b0 + r.E0 * invT) * invT;
      const double Prd = std::max(k0 * M / kf, 1e-300);
      double F = 1.0, g = 0.0, dlnF_dT_fixedPr = 0.0;
      if (r.ntroe > 0) {
        const double a = r.troe[0], T3 = r.troe[1], T1 = r.troe[2];
        double Fc = (1 - a) * std::exp(-T / T3) + a * std::exp(-T / T1);
        double dFc = -(1 - a) / T3 * std::exp(-T / T3) - a / T1 * std::exp(-T / T1);
        if (r.ntroe > 3) {
          Fc += std::exp(-r.troe[3] * invT);
          dFc += r.troe[3] * invT * invT * std::exp(-r.troe[3] * invT);
        }
        const double lFc = std::log10(Fc);
        const double c = -0.4 - 0.67 * lFc, n = 0.75 - 1.27 * lFc;
        const double x = std::log10(Prd) + c;
        const double D = n - 0.14 * x;
        const double f1 = x / D;
        const double den = 1.0 + f1 * f1;
        const double lF = lFc / den;
        F = std::pow(10.0, lF);
        g = -lFc * 2.0 * f1 * (n / (D * D)) / (den * den);  // dlog10F/dlog10Pr
        const double df1_dlFc = (-0.67 * n + 1.27 * x) / (D * D);
        const double dlF_dlFc = 1.0 / den - lFc * 2.0 * f1 * df1_dlFc / (den * den);
        const double dlFc_dT = dFc / (Fc * LN10);
        dlnF_dT_fixedPr = LN10 * dlF_dlFc * dlFc_dT;
      }
      keff = kf * Prd / (1.0 + Prd) * F;
      dkdM = k0 * F * (1.0 / ((1.0 + Prd) * (1.0 + Prd)) + g / (1.0 + Prd));
      const double dlnPr = dlnk0 - dlnkf;
      dlnkdT = dlnkf + dlnPr / (1.0 + Prd) + g * dlnPr + dlnF_dT_fixedPr;
    }
    if (!m.mult.empty()) {
      keff *= m.mult[j];
      dkdM *= m.mult[j];
    }
    const double net = Pf - Pr * invKc;
    q[j] = keff * net;
    if (dqdC) {
      double* row = dqdC + (size_t)j * N;
      for (size_t a = 0; a < r.re.size(); ++a) {
        const int k = r.re[a].first;
        const double nu = r.re[a].second;
        double d = nu * std::pow(C[k], nu - 1.0);
        for (size_t b = 0; b < r.re.size(); ++b)
          if (b != a) d *= std::pow(C[r.re[b].first], r.re[b].second);
        row[k] += keff * d;
      }
      if (r.rev) {
        for (size_t a = 0; a < r.pr.size(); ++a) {
          const int k = r.pr[a].first;
          const double nu = r.pr[a].second;
          double d = nu * std::pow(C[k], nu - 1.0);
          for (size_t b = 0; b < r.pr.size(); ++b)
            if (b != a) d *= std::pow(C[r.pr[b].first], r.pr[b].second);
          row[k] -= keff * invKc * d;
        }
      }
      if (r.kind > 0)
        for (int k = 0; k < N; ++k) row[k] += dkdM * r.eff[k] * net;
      // dq/dT: keff * dlnk/dT * net + keff * (Pr/Kc) * dlnKc/dT
      dqdT[j] = keff * dlnkdT * net + keff * Pr * invKc * dlnKc;
    }
  }
}

void wdot_from_q(const Mech& m, const double* q, double* wdot) {
  std::fill(wdot, wdot + m.N, 0.0);
  for (int j = 0; j < m.R; ++j) {
    for (auto& p : m.rx[j].re) wdot[p.first] -= p.second * q[j];
    for (auto& p : m.rx[j].pr) wdot[p.first] += p.second * q[j];
  }
}

// ---------------------------------------------------------------- reactor RHS / Jacobian
struct CellProblem {
  const Mech* m;
  double rhoe0, rhoesrc;
  const double* rYsrc;  // N, contiguous
  long nrhs = 0;
};

// y = [rhoY_0..rhoY_{N-1}, T]; returns T used
double rhs(CellProblem& P, double t, const double* y, double* ydot) {
  const Mech& m = *P.m;
  const int N = m.N;
  P.nrhs++;
  std::vector<double> Y(N), C(N), q(m.R), wd(N), ek(N), cvk(N);
  double rho = 0;
  for (int k = 0; k < N; ++k) rho += y[k];
  const double irho = 1.0 / rho;
  for (int k = 0; k < N; ++k) {
    Y[k] = y[k] * irho;
    C[k] = y[k] / m.W[k];
  }
  const double e = (P.rhoe0 + t * P.rhoesrc) * irho;
  const double T = T_from_e(m, e, Y.data(), y[N], nullptr);
  progress(m, C.data(), T, q.data(), nullptr, nullptr);
  wdot_from_q(m, q.data(), wd.data());
  e_cv(m, T, ek.data(), cvk.data());
  double num = P.rhoesrc, D = 0;
  for (int k = 0; k < N; ++k) {
    ydot[k] = m.W[k] * wd[k] + P.rYsrc[k];
    num -= ek[k] * ydot[k];
    D += y[k] * cvk[k];
  }
  ydot[N] = num / D;
  return T;
}

// Exact Jacobian of rhs() (row-major n x n). T is a function of (rhoY, t) through the energy
// constraint, so the T column is zero and dT/drhoY_m = -e_m / (rho cv) enters every column.
void jac(CellProblem& P, double t, const double* y, double* J, double* fe = nullptr) {
  const Mech& m = *P.m;
  const int N = m.N, n = N + 1;
  std::vector<double> Y(N), C(N), q(m.R), dqdC((size_t)m.R * N), dqdT(m.R);
  std::vector<double> ek(N), cvk(N), cpR(N), dcpR(N), f(N), dwdC((size_t)N * N, 0.0), dwdT(N, 0.0);
  double rho = 0;
  for (int k = 0; k < N; ++k) rho += y[k];
  for (int k = 0; k < N; ++k) {
    Y[k] = y[k] / rho;
    C[k] = y[k] / m.W[k];
  }
  const double e = (P.rhoe0 + t * P.rhoesrc) / rho;
  const double T = T_from_e(m, e, Y.data(), y[N], nullptr);
  progress(m, C.data(), T, q.data(), dqdC.data(), dqdT.data());
  std::vector<double> wd(N);
  wdot_from_q(m, q.data(), wd.data());
  for (int j = 0; j < m.R; ++j) {
    const double* row = &dqdC[(size_t)j * N];
    auto acc = [&](int k, double nu) {
      for (int mm = 0; mm < N; ++mm) dwdC[(size_t)k * N + mm] += nu * row[mm];
      dwdT[k] += nu * dqdT[j];
    };
    for (auto& p : m.rx[j].re) acc(p.first, -p.second);
    for (auto& p : m.rx[j].pr) acc(p.first, p.second);
  }
  e_cv(m, T, ek.data(), cvk.data());
  thermo(m, T, cpR.data(), nullptr, nullptr, dcpR.data());
  double D = 0, dDdT = 0, num = P.rhoesrc;
  for (int k = 0; k < N; ++k) {
    f[k] = m.W[k] * wd[k] + P.rYsrc[k];
    D += y[k] * cvk[k];
    dDdT += y[k] * dcpR[k] * R_ERG / m.W[k];
    num -= ek[k] * f[k];
  }
  const double fT = num / D;
  // A_km = df_k/dx_m at fixed T, B_k = df_k/dT
  std::vector<double> A((size_t)N * N), B(N);
  for (int k = 0; k < N; ++k) {
    for (int mm = 0; mm < N; ++mm) A[(size_t)k * N + mm] = m.W[k] * dwdC[(size_t)k * N + mm] / m.W[mm];
    B[k] = m.W[k] * dwdT[k];
  }
  double dNdT = 0;
  for (int k = 0; k < N; ++k) dNdT -= cvk[k] * f[k] + ek[k] * B[k];
  const double dfTdT = (dNdT - fT * dDdT) / D;
  if (fe) {  // df/d(rho e0) = (df/dT at fixed rhoY) / (rho cv)
    for (int k = 0; k < N; ++k) fe[k] = B[k] / D;
    fe[N] = dfTdT / D;
  }
  for (int mm = 0; mm < N; ++mm) {
    const double tau = -ek[mm] / D;
    double dNdx = 0;
    for (int k = 0; k < N; ++k) dNdx -= ek[k] * A[(size_t)k * N + mm];
    const double dfTdx = (dNdx - fT * cvk[mm]) / D;
    for (int k = 0; k < N; ++k) J[(size_t)k * n + mm] = A[(size_t)k * N + mm] + B[k] * tau;
    J[(size_t)N * n + mm] = dfTdx + dfTdT * tau;
  }
  for (int k = 0; k < n; ++k) J[(size_t)k * n + N] = 0.0;
}

// ---------------------------------------------------------------- dense LU (partial pivoting)
int getrf(int n, double* a, int* piv) {
  for (int k = 0; k < n; ++k) {
    int p = k;
    double amax = std::fabs(a[(size_t)k * n + k]);
    for (int i = k + 1; i < n; ++i) {
      const double v = std::fabs(a[(size_t)i * n + k]);
      if (v > amax) {
        amax = v;
        p = i;
      }
    }
    piv[k] = p;
    if (a[(size_t)p * n + k] == 0.0) return k + 1;
    if (p != k)
      for (int j = 0; j < n; ++j) std::swap(a[(size_t)k * n + j], a[(size_t)p * n + j]);
    const double inv = 1.0 / a[(size_t)k * n + k];
    for (int i = k + 1; i < n; ++i) {
      const double l = a[(size_t)i * n + k] * inv;
      a[(size_t)i * n + k] = l;
      for (int j = k + 1; j < n; ++j) a[(size_t)i * n + j] -= l * a[(size_t)k * n + j];
    }
  }
  return 0;
}

void getrs(int n, const double* a, const int* piv, double* b) {
  for (int k = 0; k < n; ++k) std::swap(b[k], b[piv[k]]);
  for (int i = 1; i < n; ++i)
    for (int j = 0; j < i; ++j) b[i] -= a[(size_t)i * n + j] * b[j];
  for (int i = n - 1; i >= 0; --i) {
    for (int j = i + 1; j < n; ++j) b[i] -= a[(size_t)i * n + j] * b[j];
    b[i] /= a[(size_t)i * n + i];
  }
}

// ---------------------------------------------------------------- CVODE BDF (per cell)
constexpr int QMAX = 5;
constexpr double ETAMX1 = 10000.0, ETAMX2 = 10.0, ETAMX3 = 10.0, ETAMXF = 0.2, ETAMIN = 0.1,
                 ETACF = 0.25, ADDON = 1e-6, BIAS1 = 6.0, BIAS2 = 6.0, BIAS3 = 10.0,
                 THRESH = 1.5, CRDOWN = 0.3, DGMAX = 0.3, RDIV = 2.0, NLSCOEF = 0.1,
                 DGMAX_LS = 0.2, UROUND = 2.220446049250313e-16;
constexpr int MXNEF1 = 3, SMALL_NEF = 2, LONG_WAIT = 10, MAXCOR = 3, MXNCF = 10, MXNEF = 7,
              MSBP = 20, MSBJ = 50, SMALL_NST = 10;

struct Stats {
  int nst = 0, nfe = 0, nje = 0, nsetups = 0, nni = 0, netf = 0, ncfn = 0;
};

enum { ST_OK = 0, ST_MAXSTEPS = 1, ST_ERRFAIL = 2, ST_CONVFAIL = 3, ST_BADINPUT = 4, ST_LSETUP = 5 };

struct Bdf {
  CellProblem& P;
  int n;
  double rtol;
  const double* atol;  // n
  int max_steps;
  int qmax = QMAX;
  double tstop = 0;
  std::vector<double> zn[QMAX + 1];
  std::vector<double> ewt, acor, y, ftemp, tempv, savedJ, M;
  std::vector<int> piv;
  double tn = 0, h = 0, hscale = 0, hprime = 0, eta = 1, etamax = ETAMX1, saved_tq5 = 0;
  double gamma = 0, gammap = 0, gamrat = 1, crate = 1, acnrm = 0;
  double tau[QMAX + 2] = {0}, tq[6] = {0}, l[QMAX + 1] = {0};
  double etaq = 0, etaqm1 = 0, etaqp1 = 0;
  int q = 1, qprime = 1, qwait = 2, L = 2, nstlp = 0, nstlj = 0, indx_acor = QMAX;
  bool jcur = false;
  Stats st;

  Bdf(CellProblem& P_, int n_, double rtol_, const double* atol_, int ms)
      : P(P_), n(n_), rtol(rtol_), atol(atol_), max_steps(ms) {
    for (auto& z : zn) z.assign(n, 0.0);
    ewt.assign(n, 0);
    acor.assign(n, 0);
    y.assign(n, 0);
    ftemp.assign(n, 0);
    tempv.assign(n, 0);
    savedJ.assign((size_t)n * n, 0);
    M.assign((size_t)n * n, 0);
    piv.assign(n, 0);
  }

  double wrms(const double* v) const {
    double s = 0;
    for (int i = 0; i < n; ++i) {
      const double w = v[i] * ewt[i];
      s += w * w;
    }
    return std::sqrt(s / n);
  }
  void set_ewt(const double* yy) {
    for (int i = 0; i < n; ++i) ewt[i] = 1.0 / (rtol * std::fabs(yy[i]) + atol[i]);
  }
  double f(double t, const double* yy, double* out) {
    st.nfe++;
    return rhs(P, t, yy, out);
  }

  // cvHin: initial step size from an estimate of ||y''||
  double hin(double tout) {
    const double tdist = tout - tn;
    const double tround = UROUND * std::max(std::fabs(tn), std::fabs(tout));
    const double hlb = 100.0 * tround;
    // cvUpperBoundH0
    double hub_inv = 0;
    for (int i = 0; i < n; ++i) {
      const double d = 0.1 * std::fabs(zn[0][i]) + 1.0 / ewt[i];
      hub_inv = std::max(hub_inv, std::fabs(zn[1][i]) / d);
    }
    double hub = 0.1 * tdist;
    if (hub * hub_inv > 1.0) hub = 1.0 / hub_inv;
    double hg = std::sqrt(hlb * hub);
    if (hub < hlb) return hg;
    double hnew = hg;
    for (int count1 = 1; count1 <= 4; ++count1) {
      for (int i = 0; i < n; ++i) y[i] = zn[0][i] + hg * zn[1][i];
      f(tn + hg, y.data(), tempv.data());
      for (int i = 0; i < n; ++i) tempv[i] = (tempv[i] - zn[1][i]) / hg;
      const double yddnrm = wrms(tempv.data());
      hnew = (yddnrm * hub * hub > 2.0) ? std::sqrt(2.0 / yddnrm) : std::sqrt(hg * hub);
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
    h0 = std::min(std::max(h0, hlb), hub);
    return h0;
  }

  void rescale() {
    double factor = eta;
    for (int j = 1; j <= q; ++j) {
      for (int i = 0; i < n; ++i) zn[j][i] *= factor;
      factor *= eta;
    }
    h = hscale * eta;
    hscale = h;
  }
  void predict() {
    tn += h;
    if ((tn - tstop) * h > 0.0) tn = tstop;
    for (int k = 1; k <= q; ++k)
      for (int j = q; j >= k; --j)
        for (int i = 0; i < n; ++i) zn[j - 1][i] += zn[j][i];
  }
  void restore(double saved_t) {
    tn = saved_t;
    for (int k = 1; k <= q; ++k)
      for (int j = q; j >= k; --j)
        for (int i = 0; i < n; ++i) zn[j - 1][i] -= zn[j][i];
  }
  void set_bdf() {
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
    // cvSetTqBDF
    const double A1 = 1.0 - alpha0_hat + alpha0;
    const double A2 = 1.0 + q * A1;
    tq[2] = std::fabs(A1 / (alpha0 * A2));
    tq[5] = std::fabs(A2 * xistar_inv / (l[q] * xi_inv));
    if (qwait == 1) {
      if (q > 1) {
        const double C = xistar_inv / l[q];
        const double A3 = alpha0 + 1.0 / q;
        const double A4 = alpha0_hat + xi_inv;
        const double Cpinv = (1.0 - A4 + A3) / A3;
        tq[1] = std::fabs(C * Cpinv);
      } else {
        tq[1] = 1.0;
      }
      hsum += tau[q];
      xi_inv = h / hsum;
      const double A5 = alpha0 - 1.0 / (q + 1);
      const double A6 = alpha0_hat - xi_inv;
      const double Cppinv = (1.0 - A6 + A5) / A2;
      tq[3] = std::fabs(Cppinv / (xi_inv * (q + 2) * A5));
    }
    tq[4] = NLSCOEF / tq[2];
  }
  void increase_bdf() {
    for (int i = 0; i <= qmax; ++i) l[i] = 0.0;
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
    for (int i = 0; i < n; ++i) zn[L][i] = A1 * zn[indx_acor][i];
    for (int j = 2; j <= q; ++j)
      for (int i = 0; i < n; ++i) zn[j][i] += l[j] * zn[L][i];
  }
  void decrease_bdf() {
    for (int i = 0; i <= qmax; ++i) l[i] = 0.0;
    l[2] = 1.0;
    double hsum = 0.0;
    for (int j = 1; j <= q - 2; ++j) {
      hsum += tau[j];
      const double xi = hsum / hscale;
      for (int i = j + 2; i >= 2; --i) l[i] = l[i] * xi + l[i - 1];
    }
    for (int j = 2; j < q; ++j)
      for (int i = 0; i < n; ++i) zn[j][i] -= l[j] * zn[q][i];
  }
  void adjust_order(int dq) {
    if (q == 2 && dq != 1) return;
    if (dq == 1)
      increase_bdf();
    else
      decrease_bdf();
  }

  // Newton matrix set-up (cvDlsSetup): returns 0 ok, >0 singular
  int lsetup(int convfail) {
    const double dgamma = std::fabs(gamma / gammap - 1.0);
    const bool jbad = (st.nst == 0) || (st.nst > nstlj + MSBJ) ||
                      (convfail == 2 && dgamma < DGMAX_LS) || (convfail == 1);
    if (!jbad) {
      jcur = false;
    } else {
      st.nje++;
      nstlj = st.nst;
      jcur = true;
      jac(P, tn, zn[0].data(), savedJ.data());
    }
    for (size_t i = 0; i < M.size(); ++i) M[i] = -gamma * savedJ[i];
    for (int i = 0; i < n; ++i) M[(size_t)i * n + i] += 1.0;
    return getrf(n, M.data(), piv.data());
  }

  // cvNls: returns 0 converged, 1 recoverable convergence failure
  int nls(int nflag) {
    // nflag: 0 first call, 1 previous convergence failure, 2 previous error-test failure
    int convfail = (nflag == 0 || nflag == 2) ? 0 : 1;
    bool callSetup = (nflag == 1) || (nflag == 2) || (st.nst == 0) || (st.nst >= nstlp + MSBP) ||
                     (std::fabs(gamrat - 1.0) > DGMAX);
    for (;;) {
      f(tn, zn[0].data(), ftemp.data());
      if (callSetup) {
        const int ier = lsetup(convfail);
        st.nsetups++;
        callSetup = false;
        gamrat = crate = 1.0;
        gammap = gamma;
        nstlp = st.nst;
        if (ier != 0) return 1;
      }
      std::fill(acor.begin(), acor.end(), 0.0);
      for (int i = 0; i < n; ++i) y[i] = zn[0][i];
      // Newton iteration
      int m = 0;
      double del = 0, delp = 0;
      const double rl1 = 1.0 / l[1];
      int ret;
      for (;;) {
        for (int i = 0; i < n; ++i) tempv[i] = gamma * ftemp[i] - (rl1 * zn[1][i] + acor[i]);
        getrs(n, M.data(), piv.data(), tempv.data());
        if (gamrat != 1.0)
          for (int i = 0; i < n; ++i) tempv[i] *= 2.0 / (1.0 + gamrat);
        st.nni++;
        del = wrms(tempv.data());
        for (int i = 0; i < n; ++i) {
          acor[i] += tempv[i];
          y[i] = zn[0][i] + acor[i];
        }
        if (m > 0) crate = std::max(CRDOWN * crate, del / delp);
        const double dcon = del * std::min(1.0, crate) / tq[4];
        if (dcon <= 1.0) {
          acnrm = (m == 0) ? del : wrms(acor.data());
          jcur = false;
          ret = 0;
          break;
        }
        ++m;
        if (m == MAXCOR || (m >= 2 && del > RDIV * delp)) {
          ret = jcur ? 1 : 2;  // 2: try again with a fresh Jacobian
          break;
        }
        delp = del;
        f(tn, y.data(), ftemp.data());
      }
      if (ret != 2) return ret;
      callSetup = true;
      convfail = 2;
    }
  }

  void complete_step() {
    st.nst++;
    for (int i = q; i >= 2; --i) tau[i] = tau[i - 1];
    if (q == 1 && st.nst > 1) tau[2] = tau[1];
    tau[1] = h;
    for (int j = 0; j <= q; ++j)
      for (int i = 0; i < n; ++i) zn[j][i] += l[j] * acor[i];
    qwait--;
    if (qwait == 1 && q != qmax) {
      for (int i = 0; i < n; ++i) zn[qmax][i] = acor[i];
      saved_tq5 = tq[5];
      indx_acor = qmax;
    }
  }
  void set_eta() {
    if (eta < THRESH) {
      eta = 1.0;
      hprime = h;
    } else {
      eta = std::min(eta, etamax);
      hprime = h * eta;
    }
  }
  void prepare_next_step(double dsm) {
    if (etamax == 1.0) {
      qwait = std::max(qwait, 2);
      qprime = q;
      hprime = h;
      eta = 1.0;
      return;
    }
    etaq = 1.0 / (std::pow(BIAS2 * dsm, 1.0 / L) + ADDON);
    if (qwait != 0) {
      eta = etaq;
      qprime = q;
      set_eta();
      return;
    }
    qwait = 2;
    etaqm1 = 0.0;
    if (q > 1) {
      const double ddn = wrms(zn[q].data()) * tq[1];
      etaqm1 = 1.0 / (std::pow(BIAS1 * ddn, 1.0 / q) + ADDON);
    }
    etaqp1 = 0.0;
    if (q != qmax && saved_tq5 != 0.0) {
      const double cquot = (tq[5] / saved_tq5) * std::pow(h / tau[2], L);
      for (int i = 0; i < n; ++i) tempv[i] = acor[i] - cquot * zn[qmax][i];
      const double dup = wrms(tempv.data()) * tq[3];
      etaqp1 = 1.0 / (std::pow(BIAS3 * dup, 1.0 / (L + 1)) + ADDON);
    }
    // cvChooseEta
    const double etam = std::max(etaqm1, std::max(etaq, etaqp1));
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
      for (int i = 0; i < n; ++i) zn[qmax][i] = acor[i];
    }
    set_eta();
  }

  // one CVODE step (cvStep); returns 0 ok or a failure status
  int step() {
    const double saved_t = tn;
    int ncf = 0, nef = 0, nflag = 0;
    if (st.nst > 0 && hprime != h) {
      if (qprime != q) {
        adjust_order(qprime - q);
        q = qprime;
        L = q + 1;
        qwait = L;
      }
      rescale();
    }
    double dsm = 0;
    for (;;) {
      predict();
      set_bdf();
      gamma = h / l[1];
      if (st.nst == 0) gammap = gamma;
      gamrat = (st.nst > 0) ? gamma / gammap : 1.0;
      const int nret = nls(nflag);
      if (nret != 0) {
        st.ncfn++;
        restore(saved_t);
        ncf++;
        etamax = 1.0;
        if (ncf == MXNCF) return ST_CONVFAIL;
        eta = ETACF;
        nflag = 1;
        rescale();
        continue;
      }
      dsm = acnrm * tq[2];
      if (dsm <= 1.0) break;
      // error test failed
      nef++;
      st.netf++;
      nflag = 2;
      restore(saved_t);
      if (nef == MXNEF) return ST_ERRFAIL;
      etamax = 1.0;
      if (nef <= MXNEF1) {
        eta = 1.0 / (std::pow(BIAS2 * dsm, 1.0 / L) + ADDON);
        eta = std::max(ETAMIN, eta);
        if (nef >= SMALL_NEF) eta = std::min(eta, ETAMXF);
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
      f(tn, zn[0].data(), tempv.data());
      for (int i = 0; i < n; ++i) zn[1][i] = h * tempv[i];
    }
    complete_step();
    prepare_next_step(dsm);
    etamax = (st.nst <= SMALL_NST) ? ETAMX2 : ETAMX3;
    return ST_OK;
  }

  // integrate y0 over [0, tout]; y0 overwritten with the solution
  int integrate(double* yio, double tout) {
    tn = 0.0;
    tstop = tout;
    indx_acor = qmax;
    for (int i = 0; i < n; ++i) zn[0][i] = yio[i];
    set_ewt(zn[0].data());
    f(tn, zn[0].data(), zn[1].data());
    h = hin(tout);
    if ((tn + h - tout) * h > 0.0) h = (tout - tn) * (1.0 - 4.0 * UROUND);
    hscale = h;
    hprime = h;
    for (int i = 0; i < n; ++i) zn[1][i] *= h;
    q = 1;
    L = 2;
    qwait = L;
    etamax = ETAMX1;
    for (;;) {
      if (st.nst >= max_steps) return ST_MAXSTEPS;
      set_ewt(zn[0].data());
      const int s = step();
      if (s != ST_OK) return s;
      const double troundoff = 100.0 * UROUND * (std::fabs(tn) + std::fabs(h));
      if (std::fabs(tn - tout) <= troundoff) break;
      if ((tn + hprime - tout) * h > 0.0) {
        hprime = (tout - tn) * (1.0 - 4.0 * UROUND);
        eta = hprime / h;
      }
    }
    for (int i = 0; i < n; ++i) yio[i] = zn[0][i];
    return ST_OK;
  }
};

// ---------------------------------------------------------------------------------------------
// Explicit RK64 (the ReactorRK64 drop-in): the six-stage fourth-order two-register low-storage
// scheme RK4(3)6[2R+]C of Kennedy, Carpenter & Lewis (Appl. Numer. Math. 35 (2000) 177-219), with
// its embedded third-order error estimate. Register form, per stage s:
//   k = f(t + c_s h, S);  E += e_s h k;  S = Q + alpha_s h k;  Q = S + beta_s h k
// starting from Q = S = y_n; y_{n+1} = Q. The controller (the reference's is in the absent
// PelePhysics submodule) accepts a step when max_i |E_i| / (rtol |y_i| + atol_i) <= 1 or h is at
// its floor dt / nsub_max, scales h by 0.9 err^(-1/4) within [0.2, 5], keeps dt / nsub_max <= h <=
// dt / nsub_min, starts from h = dt / nsub_guess and clips the last step to land on dt (within
// 100 u dt, so accumulated round-off in t never adds a sliver of a step). A step
// whose stages leave the finite domain is rejected with the factor 0.2; at the floor the cell
// fails with status 2 (error-test failure).
constexpr double RK_ALPHA[6] = {3296351145737.0 / 15110423921029.0, 1879360555526.0 / 7321162733569.0,
                                10797097731880.0 / 20472212111779.0, 754636544611.0 / 15563872110659.0,
                                3260218886217.0 / 2618290685819.0, 5069185909380.0 / 12292927838509.0};
constexpr double RK_BETA[6] = {-1204558336989.0 / 10607789004752.0, -3028468927040.0 / 14078136890693.0,
                               -455570672869.0 / 8930094212428.0, -17275898420483.0 / 15997285579755.0,
                               -2453906524165.0 / 9868353053862.0, 0.0};
constexpr double RK_ERR[6] = {-530312978447.0 / 9560368366154.0, 473021958881.0 / 2984707536468.0,
                              -947229622805.0 / 10456009803779.0, -2921473878215.0 / 13334914072261.0,
                              1519535112975.0 / 9264196100452.0, 167623581683.0 / 3930932046784.0};

struct Rk64Stats {
  int nst = 0, nfe = 0, nrej = 0;
};

int rk64_integrate(CellProblem& P, int n, double* y, double dt, double rtol, const double* atolv,
                   int nsub_guess, int nsub_min, int nsub_max, int max_steps, Rk64Stats& st) {
  // stage times c_s (row sums of the Butcher matrix of the register form)
  double c[6];
  c[0] = 0.0;
  for (int s = 1; s < 6; ++s) {
    double cs = RK_ALPHA[s - 1];
    for (int j = 0; j + 1 < s; ++j) cs += RK_ALPHA[j] + RK_BETA[j];
    c[s] = cs;
  }
  const double hmin = dt / nsub_max, hmax = dt / nsub_min;
  double h = dt / nsub_guess;
  double t = 0.0;
  std::vector<double> Q(n), S(n), E(n), k(n);
  const double tend_tol = 100.0 * 2.220446049250313e-16 * dt;
  while (dt - t > tend_tol) {
    if (st.nst + st.nrej >= max_steps) return ST_MAXSTEPS;
    // the last step lands on dt (a remainder below the round-off tolerance is absorbed)
    const double hu = (t + h >= dt - tend_tol) ? dt - t : h;
    for (int i = 0; i < n; ++i) {
      Q[i] = y[i];
      S[i] = y[i];
      E[i] = 0.0;
    }
    for (int s = 0; s < 6; ++s) {
      rhs(P, t + c[s] * hu, S.data(), k.data());
      st.nfe++;
      for (int i = 0; i < n; ++i) {
        E[i] += RK_ERR[s] * hu * k[i];
        S[i] = Q[i] + RK_ALPHA[s] * hu * k[i];
        Q[i] = S[i] + RK_BETA[s] * hu * k[i];
      }
    }
    double en = 0.0;
    bool finite = true;
    for (int i = 0; i < n; ++i) {
      const double ei = std::fabs(E[i]) / (rtol * std::fabs(y[i]) + atolv[i]);
      finite = finite && std::isfinite(ei) && std::isfinite(Q[i]);
      en = std::max(en, ei);
    }
    if (!finite) {
      // a trial stage left the physical domain: shrink, or fail at the step floor
      if (hu <= hmin) return ST_ERRFAIL;
      st.nrej++;
      h = std::max(0.2 * hu, hmin);
      continue;
    }
    const double fac = en > 0.0 ? 0.9 * std::pow(en, -0.25) : 5.0;
    if (en <= 1.0 || hu <= hmin) {
      for (int i = 0; i < n; ++i) y[i] = Q[i];
      t += hu;
      st.nst++;
      h = std::min(std::max(hu * std::min(std::max(fac, 0.2), 5.0), hmin), hmax);
    } else {
      st.nrej++;
      h = std::max(hu * std::max(fac, 0.2), hmin);
    }
  }
  return ST_OK;
}

// ---------------------------------------------------------------------------------------------
// Replay of a recorded BDF step sequence (the tape the device adjoint records; record layout:
// [0] op [1] q [2] L|save [3] ia|qmax [4] eta|sign|A1|gamma|h [5..10] l_0..l_5 [11] 1/l_1 [12] t
// [16..16+n) vector). The Nordsieck operations are re-applied with the tape's step-control
// decisions; each corrector equation D = gamma f(t, z0 + D) - z1 / l_1 is solved to convergence
// with a fresh Jacobian per Newton iteration, at the given initial state, initial rho*e and rate
// multipliers exp(logA). This is the map the discrete adjoint differentiates: central differences
// of it check the adjoint, and its tangent-linear form (one direction) gives forward sensitivities.
enum { OP_RESCALE = 1, OP_PASCAL, OP_ORDER_UP, OP_ORDER_DOWN, OP_CORRECT, OP_RESTART, OP_INIT };

struct Replay {
  CellProblem& P;
  int N, n;
  bool tangent;
  std::vector<double> dp;  // tangent direction of logA (R)
  double de = 0.0;         // tangent direction of rho e0
  std::vector<double> z, dz, J, fe, M;
  std::vector<int> piv;

  Replay(CellProblem& P_, bool tan) : P(P_), N(P_.m->N), n(P_.m->N + 1), tangent(tan) {
    z.assign(6 * n, 0.0);
    dz.assign(6 * n, 0.0);
    J.resize(n * n);
    fe.resize(n);
    M.resize(n * n);
    piv.resize(n);
  }
  double& Z(int j, int i) { return z[j * n + i]; }
  double& dZ(int j, int i) { return dz[j * n + i]; }

  // out = J v + (df/dp) dp + fe de at (t, y) (J and fe must be current for (t, y))
  void tangent_rhs(double t, const double* y, const double* v, double* out) {
    const Mech& m = *P.m;
    for (int k = 0; k < n; ++k) {
      double a = fe[k] * de;
      for (int j = 0; j < n; ++j) a += J[k * n + j] * v[j];
      out[k] = a;
    }
    std::vector<double> Y(N), C(N), q(m.R), wd(N), ek(N), cvk(N);
    double rho = 0;
    for (int k = 0; k < N; ++k) rho += y[k];
    for (int k = 0; k < N; ++k) {
      Y[k] = y[k] / rho;
      C[k] = y[k] / m.W[k];
    }
    const double T = T_from_e(m, (P.rhoe0 + t * P.rhoesrc) / rho, Y.data(), y[N], nullptr);
    progress(m, C.data(), T, q.data(), nullptr, nullptr);
    for (int j = 0; j < m.R; ++j) q[j] *= dp[j];
    wdot_from_q(m, q.data(), wd.data());
    e_cv(m, T, ek.data(), cvk.data());
    double D = 0, s = 0;
    for (int k = 0; k < N; ++k) {
      D += y[k] * cvk[k];
      const double fk = m.W[k] * wd[k];
      out[k] += fk;
      s += ek[k] * fk;
    }
    out[N] -= s / D;
  }

  void op(const double* r) {
    const int o = (int)r[0], q = (int)r[1];
    switch (o) {
      case OP_RESCALE: {
        double f = 1.0;
        for (int j = 1; j <= 5; ++j) {
          f *= r[4];
          if (j <= q)
            for (int i = 0; i < n; ++i) {
              Z(j, i) *= f;
              dZ(j, i) *= f;
            }
        }
        break;
      }
      case OP_PASCAL:
        for (int k = 1; k <= q; ++k)
          for (int j = q; j >= k; --j)
            for (int i = 0; i < n; ++i) {
              Z(j - 1, i) += r[4] * Z(j, i);
              dZ(j - 1, i) += r[4] * dZ(j, i);
            }
        break;
      case OP_ORDER_UP: {
        const int L = (int)r[2], ia = (int)r[3];
        for (int i = 0; i < n; ++i) {
          const double zl = r[4] * Z(ia, i), dl = r[4] * dZ(ia, i);
          Z(L, i) = zl;
          dZ(L, i) = dl;
          for (int j = 2; j <= q; ++j) {
            Z(j, i) += r[5 + j] * zl;
            dZ(j, i) += r[5 + j] * dl;
          }
        }
        break;
      }
      case OP_ORDER_DOWN:
        for (int i = 0; i < n; ++i) {
          const double zq = Z(q, i), dq = dZ(q, i);
          for (int j = 2; j < q; ++j) {
            Z(j, i) -= r[5 + j] * zq;
            dZ(j, i) -= r[5 + j] * dq;
          }
        }
        break;
      case OP_CORRECT: {
        const bool save = r[2] != 0.0;
        const int qmax = (int)r[3];
        const double gamma = r[4], rl1 = r[11], t = r[12];
        std::vector<double> D(n), y(n), fv(n), g(n);
        for (int i = 0; i < n; ++i) D[i] = r[16 + i] - Z(0, i);
        for (int it = 0; it < 40; ++it) {
          for (int i = 0; i < n; ++i) y[i] = Z(0, i) + D[i];
          rhs(P, t, y.data(), fv.data());
          jac(P, t, y.data(), J.data());
          for (int i = 0; i < n; ++i) g[i] = -(D[i] - gamma * fv[i] + rl1 * Z(1, i));
          for (int k = 0; k < n * n; ++k) M[k] = -gamma * J[k];
          for (int k = 0; k < n; ++k) M[k * n + k] += 1.0;
          getrf(n, M.data(), piv.data());
          getrs(n, M.data(), piv.data(), g.data());
          bool conv = true;
          for (int i = 0; i < n; ++i) {
            D[i] += g[i];
            conv = conv && std::fabs(g[i]) <= 1e-14 * std::fabs(Z(0, i) + D[i]) + 1e-300;
          }
          if (conv) break;
        }
        for (int i = 0; i < n; ++i) y[i] = Z(0, i) + D[i];
        std::vector<double> dD(n);
        if (tangent) {
          jac(P, t, y.data(), J.data(), fe.data());
          std::vector<double> v(n), tr(n);
          for (int i = 0; i < n; ++i) v[i] = dZ(0, i);
          tangent_rhs(t, y.data(), v.data(), tr.data());
          for (int i = 0; i < n; ++i) dD[i] = gamma * tr[i] - rl1 * dZ(1, i);
          for (int k = 0; k < n * n; ++k) M[k] = -gamma * J[k];
          for (int k = 0; k < n; ++k) M[k * n + k] += 1.0;
          getrf(n, M.data(), piv.data());
          getrs(n, M.data(), piv.data(), dD.data());
        }
        for (int i = 0; i < n; ++i) {
          for (int j = 0; j <= q; ++j) {
            Z(j, i) += r[5 + j] * D[i];
            dZ(j, i) += r[5 + j] * dD[i];
          }
          if (save) {
            Z(qmax, i) = D[i];
            dZ(qmax, i) = dD[i];
          }
        }
        break;
      }
      case OP_RESTART:
      case OP_INIT: {
        const double h = r[4], t = r[12];
        std::vector<double> y(n), fv(n);
        for (int i = 0; i < n; ++i) y[i] = Z(0, i);
        rhs(P, t, y.data(), fv.data());
        if (tangent) {
          jac(P, t, y.data(), J.data(), fe.data());
          std::vector<double> v(n), tr(n);
          for (int i = 0; i < n; ++i) v[i] = dZ(0, i);
          tangent_rhs(t, y.data(), v.data(), tr.data());
          for (int i = 0; i < n; ++i) dZ(1, i) = h * tr[i];
        }
        for (int i = 0; i < n; ++i) Z(1, i) = h * fv[i];
        break;
      }
      default:
        break;
    }
  }
};

}  // namespace

// ====================================================================== C ABI (tests only)
extern "C" {

void* orc_mech_new(int N, int R, const double* W, const double* Tmid, const double* low,
                   const double* high, const int* kind, const int* rev, const double* A,
                   const double* b, const double* EaR, const double* nuf, const double* nur,
                   const double* eff, const double* lowp, const double* troe, const int* ntroe) {
  Mech* m = new Mech;
  m->N = N;
  m->R = R;
  m->W.assign(W, W + N);
  m->Tmid.assign(Tmid, Tmid + N);
  m->low.assign(low, low + 7 * N);
  m->high.assign(high, high + 7 * N);
  m->rx.resize(R);
  for (int j = 0; j < R; ++j) {
    Rxn& r = m->rx[j];
    r.kind = kind[j];
    r.rev = rev[j] != 0;
    r.A = A[j];
    r.b = b[j];
    r.EaR = EaR[j];
    for (int k = 0; k < N; ++k) {
      if (nuf[(size_t)j * N + k] != 0) r.re.push_back({k, nuf[(size_t)j * N + k]});
      if (nur[(size_t)j * N + k] != 0) r.pr.push_back({k, nur[(size_t)j * N + k]});
    }
    r.dnu = 0;
    for (auto& p : r.pr) r.dnu += p.second;
    for (auto& p : r.re) r.dnu -= p.second;
    if (r.kind > 0) r.eff.assign(eff + (size_t)j * N, eff + (size_t)(j + 1) * N);
    if (r.kind == 2) {
      r.A0 = lowp[3 * j];
      r.b0 = lowp[3 * j + 1];
      r.E0 = lowp[3 * j + 2];
      r.ntroe = ntroe[j];
      for (int i = 0; i < 4; ++i) r.troe[i] = troe[4 * j + i];
    }
  }
  return m;
}

void orc_mech_free(void* m) { delete static_cast<Mech*>(m); }

void orc_thermo(void* mp, double T, double* cpR, double* hRT, double* sR) {
  thermo(*static_cast<Mech*>(mp), T, cpR, hRT, sR, nullptr);
}

double orc_T_from_e(void* mp, double e, const double* Y, double Tguess, int* iters) {
  return T_from_e(*static_cast<Mech*>(mp), e, Y, Tguess, iters);
}

void orc_progress(void* mp, const double* C, double T, double* q, double* dqdC, double* dqdT) {
  progress(*static_cast<Mech*>(mp), C, T, q, dqdC, dqdT);
}

void orc_wdot(void* mp, const double* C, double T, double* wdot) {
  Mech& m = *static_cast<Mech*>(mp);
  std::vector<double> q(m.R);
  progress(m, C, T, q.data(), nullptr, nullptr);
  wdot_from_q(m, q.data(), wdot);
}

double orc_rhs(void* mp, double t, const double* y, double rhoe0, double rhoesrc,
               const double* rYsrc, double* ydot) {
  CellProblem P{static_cast<Mech*>(mp), rhoe0, rhoesrc, rYsrc};
  return rhs(P, t, y, ydot);
}

void orc_jac(void* mp, double t, const double* y, double rhoe0, double rhoesrc, const double* rYsrc,
             double* J) {
  CellProblem P{static_cast<Mech*>(mp), rhoe0, rhoesrc, rYsrc};
  jac(P, t, y, J);
}

int orc_getrf(int n, double* a, int* piv) { return getrf(n, a, piv); }
void orc_getrs(int n, const double* a, const int* piv, double* b) { getrs(n, a, piv, b); }

// Batched reactor over ncells cells in component-major (SoA) layout: rY[k*ncells + i].
// stats: 7 ints per cell (nst, nfe, nje, nsetups, nni, netf, ncfn), may be null.
// status: per-cell status (0 ok), may be null. Returns the number of failed cells.
int orc_react(void* mp, int ncells, double* rY, const double* rYsrc, double* T, double* rE,
              const double* rEsrc, double* FC, const int* mask, double dt, double rtol,
              double atol, const double* typ, int max_steps, int max_order, int nthreads,
              int* status, int* stats) {
  const Mech& m = *static_cast<Mech*>(mp);
  const int N = m.N, n = N + 1;
  std::vector<double> atolv(n);
  for (int i = 0; i < n; ++i) atolv[i] = typ ? atol * typ[i] : atol;
  int nfail = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : nfail)
#endif
  for (int c = 0; c < ncells; ++c) {
    if (mask && mask[c] == 0) {
      if (FC) FC[c] = 0;
      if (status) status[c] = 0;
      continue;
    }
    std::vector<double> y(n), src(N);
    for (int k = 0; k < N; ++k) {
      y[k] = rY[(size_t)k * ncells + c];
      src[k] = rYsrc ? rYsrc[(size_t)k * ncells + c] : 0.0;
    }
    y[N] = T[c];
    const double rhoe0 = rE[c];
    const double resrc = rEsrc ? rEsrc[c] : 0.0;
    CellProblem P{&m, rhoe0, resrc, src.data()};
    Bdf bdf(P, n, rtol, atolv.data(), max_steps);
    bdf.qmax = std::min(std::max(max_order, 1), QMAX);
    int s = ST_OK;
    double rho = 0;
    for (int k = 0; k < N; ++k) rho += y[k];
    if (!(rho > 0) || !(y[N] > 0) || !(dt > 0)) s = ST_BADINPUT;
    if (s == ST_OK) s = bdf.integrate(y.data(), dt);
    if (s == ST_OK) {
      double rhon = 0;
      for (int k = 0; k < N; ++k) rhon += y[k];
      std::vector<double> Y(N);
      for (int k = 0; k < N; ++k) Y[k] = y[k] / rhon;
      const double rEn = rhoe0 + dt * resrc;
      const double Tn = T_from_e(m, rEn / rhon, Y.data(), y[N], nullptr);
      for (int k = 0; k < N; ++k) rY[(size_t)k * ncells + c] = y[k];
      T[c] = Tn;
      rE[c] = rEn;
    } else {
      nfail++;
    }
    if (FC) FC[c] = bdf.st.nfe;
    if (status) status[c] = s;
    if (stats) {
      int* o = stats + (size_t)7 * c;
      o[0] = bdf.st.nst;
      o[1] = bdf.st.nfe;
      o[2] = bdf.st.nje;
      o[3] = bdf.st.nsetups;
      o[4] = bdf.st.nni;
      o[5] = bdf.st.netf;
      o[6] = bdf.st.ncfn;
    }
  }
  return nfail;
}

// ReactorRK64 restatement over ncells cells (same array conventions as orc_react);
// stats per cell: steps, RHS evaluations, rejected steps.
int orc_react_rk64(void* mp, int ncells, double* rY, const double* rYsrc, double* T, double* rE,
                   const double* rEsrc, double* FC, const int* mask, double dt, double rtol,
                   double atol, const double* typ, int nsub_guess, int nsub_min, int nsub_max,
                   int max_steps, int nthreads, int* status, int* stats) {
  const Mech& m = *static_cast<Mech*>(mp);
  const int N = m.N, n = N + 1;
  std::vector<double> atolv(n);
  for (int i = 0; i < n; ++i) atolv[i] = typ ? atol * typ[i] : atol;
  int nfail = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : nfail)
#endif
  for (int c = 0; c < ncells; ++c) {
    if (mask && mask[c] == 0) {
      if (FC) FC[c] = 0;
      if (status) status[c] = 0;
      continue;
    }
    std::vector<double> y(n), src(N);
    for (int k2 = 0; k2 < N; ++k2) {
      y[k2] = rY[(size_t)k2 * ncells + c];
      src[k2] = rYsrc ? rYsrc[(size_t)k2 * ncells + c] : 0.0;
    }
    y[N] = T[c];
    const double rhoe0 = rE[c];
    const double resrc = rEsrc ? rEsrc[c] : 0.0;
    CellProblem P{&m, rhoe0, resrc, src.data()};
    Rk64Stats st;
    int s = ST_OK;
    double rho = 0;
    for (int k2 = 0; k2 < N; ++k2) rho += y[k2];
    if (!(rho > 0) || !(y[N] > 0) || !(dt > 0)) s = ST_BADINPUT;
    if (s == ST_OK)
      s = rk64_integrate(P, n, y.data(), dt, rtol, atolv.data(), nsub_guess, nsub_min, nsub_max,
                         max_steps, st);
    if (s == ST_OK) {
      double rhon = 0;
      for (int k2 = 0; k2 < N; ++k2) rhon += y[k2];
      std::vector<double> Y(N);
      for (int k2 = 0; k2 < N; ++k2) Y[k2] = y[k2] / rhon;
      const double rEn = rhoe0 + dt * resrc;
      const double Tn = T_from_e(m, rEn / rhon, Y.data(), y[N], nullptr);
      for (int k2 = 0; k2 < N; ++k2) rY[(size_t)k2 * ncells + c] = y[k2];
      T[c] = Tn;
      rE[c] = rEn;
    } else {
      nfail++;
    }
    if (FC) FC[c] = st.nfe;
    if (status) status[c] = s;
    if (stats) {
      int* o = stats + (size_t)3 * c;
      o[0] = st.nst;
      o[1] = st.nfe;
      o[2] = st.nrej;
    }
  }
  return nfail;
}

// Replay one cell's tape (nrec records of rs doubles) from y0 (n = N + 1: rhoY, T guess), rho e0
// and logA (R values, null = 0); writes y_end (n) and T_end. With dy0 non-null also the tangent:
// dy_end / dT_end along (dy0, drhoe0, dlogA).
int orc_replay(void* mp, const double* tape, int nrec, int rs, const double* y0, double rhoe0,
               double rhoesrc, const double* rYsrc, const double* logA, double dt, double* yend,
               double* Tend, const double* dy0, double drhoe0, const double* dlogA, double* dyend,
               double* dTend) {
  Mech m = *static_cast<Mech*>(mp);
  const int N = m.N, n = N + 1;
  m.mult.assign(m.R, 1.0);
  if (logA)
    for (int j = 0; j < m.R; ++j) m.mult[j] = std::exp(logA[j]);
  std::vector<double> src(N, 0.0);
  if (rYsrc)
    for (int k = 0; k < N; ++k) src[k] = rYsrc[k];
  CellProblem P{&m, rhoe0, rhoesrc, src.data()};
  Replay rp(P, dy0 != nullptr);
  rp.dp.assign(m.R, 0.0);
  if (dlogA)
    for (int j = 0; j < m.R; ++j) rp.dp[j] = dlogA[j];
  rp.de = drhoe0;
  for (int r = 0; r < nrec; ++r) {
    const double* rec = tape + (size_t)r * rs;
    if ((int)rec[0] == OP_INIT)
      for (int i = 0; i < n; ++i) {
        rp.Z(0, i) = y0[i];
        rp.dZ(0, i) = dy0 ? dy0[i] : 0.0;
      }
    rp.op(rec);
  }
  for (int i = 0; i < n; ++i) yend[i] = rp.Z(0, i);
  double rho = 0;
  for (int k = 0; k < N; ++k) rho += yend[k];
  std::vector<double> Y(N), ek(N), cvk(N);
  for (int k = 0; k < N; ++k) Y[k] = yend[k] / rho;
  const double T = T_from_e(m, (rhoe0 + dt * rhoesrc) / rho, Y.data(), yend[N], nullptr);
  *Tend = T;
  if (dy0) {
    e_cv(m, T, ek.data(), cvk.data());
    double D = 0, dT = drhoe0;
    for (int k = 0; k < N; ++k) D += yend[k] * cvk[k];
    for (int k = 0; k < N; ++k) dT -= ek[k] * rp.dZ(0, k);
    for (int i = 0; i < n; ++i) dyend[i] = rp.dZ(0, i);
    *dTend = dT / D;
  }
  return 0;
}

}  // extern "C"
