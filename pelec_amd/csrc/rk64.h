// Explicit RK64 integrator for one cell (the ReactorRK64 drop-in, BuildPelePhysicsLib.cmake:136-137;
// selected by pelec.chem_integrator = "ReactorRK64", e.g. Exec/RegTests/zeroD/example.inp:20).
//
// Method: the six-stage fourth-order two-register low-storage scheme RK4(3)6[2R+]C of Kennedy,
// Carpenter & Lewis (Appl. Numer. Math. 35 (2000) 177-219) with its embedded third-order error
// estimate. Register form, per stage s:
//   k = f(t + c_s h, S);  E += e_s h k;  S = Q + alpha_s h k;  Q = S + beta_s h k
// from Q = S = y_n; y_{n+1} = Q. Only the two registers Q, S and the error accumulator E are kept.
// Step control (the reference's controller lives in the absent PelePhysics submodule, so this one is
// defined here): accept when max_i |E_i| / (rtol |y_i| + atol_i) <= 1 or h is at its floor
// dt / nsub_max; the next h is 0.9 err^(-1/4) h within [0.2 h, 5 h] and [dt / nsub_max,
// dt / nsub_min]; the first h is dt / nsub_guess; the last step lands on dt (within 100 u dt). A
// step whose stages leave the finite domain is rejected with factor 0.2, or fails at the floor.
#pragma once
#include "bdf_core.h"

namespace pelec {

struct Rk64Params {
  double rtol;
  const double* atolv;  // n entries
  int nsub_guess, nsub_min, nsub_max;
  int max_steps;
};

namespace rk {
constexpr double ALPHA[6] = {3296351145737.0 / 15110423921029.0, 1879360555526.0 / 7321162733569.0,
                             10797097731880.0 / 20472212111779.0, 754636544611.0 / 15563872110659.0,
                             3260218886217.0 / 2618290685819.0, 5069185909380.0 / 12292927838509.0};
constexpr double BETA[6] = {-1204558336989.0 / 10607789004752.0, -3028468927040.0 / 14078136890693.0,
                            -455570672869.0 / 8930094212428.0, -17275898420483.0 / 15997285579755.0,
                            -2453906524165.0 / 9868353053862.0, 0.0};
constexpr double ERR[6] = {-530312978447.0 / 9560368366154.0, 473021958881.0 / 2984707536468.0,
                           -947229622805.0 / 10456009803779.0, -2921473878215.0 / 13334914072261.0,
                           1519535112975.0 / 9264196100452.0, 167623581683.0 / 3930932046784.0};
// stage abscissae: c_0 = 0, c_s = alpha_{s-1} + sum_{j < s-1} (alpha_j + beta_j)
__host__ __device__ constexpr double stage_c(int s) {
  double c = 0.0;
  if (s > 0) {
    c = ALPHA[s - 1];
    for (int j = 0; j + 1 < s; ++j) c += ALPHA[j] + BETA[j];
  }
  return c;
}
}  // namespace rk

template <class Mech>
struct Rk64Cell {
  using Model = CellModel<Mech>;
  static constexpr int N = Mech::NSP;
  static constexpr int n = N + 1;

  double rhoe0 = 0.0, rhoesrc = 0.0;
  const double* src = nullptr;
  int64_t sstride = 0;
  int nst = 0, nfe = 0, nrej = 0;

  // Integrates y (n entries, in/out) over [0, dt]; returns a CellStatus.
  __device__ int integrate(double* y, double dt, const Rk64Params& p) {
    const double hmin = dt / p.nsub_max, hmax = dt / p.nsub_min;
    const double tend_tol = 100.0 * cv::UROUND * dt;
    double h = dt / p.nsub_guess;
    double t = 0.0;
    while (dt - t > tend_tol) {
      if (nst + nrej >= p.max_steps) return CELL_MAXSTEPS;
      const double hu = (t + h >= dt - tend_tol) ? dt - t : h;
      double Q[n], S[n], E[n];
#pragma unroll
      for (int i = 0; i < n; ++i) {
        Q[i] = y[i];
        S[i] = y[i];
        E[i] = 0.0;
      }
#pragma unroll 1
      for (int s = 0; s < 6; ++s) {
        double k[n];
        Model::rhs(t + rk::stage_c(s) * hu, S, rhoe0, rhoesrc, src, sstride, k);
        nfe++;
        const double ea = rk::ERR[s] * hu, aa = rk::ALPHA[s] * hu, ba = rk::BETA[s] * hu;
#pragma unroll
        for (int i = 0; i < n; ++i) {
          E[i] += ea * k[i];
          S[i] = Q[i] + aa * k[i];
          Q[i] = S[i] + ba * k[i];
        }
      }
      double en = 0.0;
      bool finite = true;
#pragma unroll
      for (int i = 0; i < n; ++i) {
        const double ei = fabs(E[i]) / (p.rtol * fabs(y[i]) + p.atolv[i]);
        finite = finite && isfinite(ei) && isfinite(Q[i]);
        en = fmax(en, ei);
      }
      if (!finite) {
        if (hu <= hmin) return CELL_ERRFAIL;
        nrej++;
        h = fmax(0.2 * hu, hmin);
        continue;
      }
      const double fac = en > 0.0 ? 0.9 * pow(en, -0.25) : 5.0;
      if (en <= 1.0 || hu <= hmin) {
#pragma unroll
        for (int i = 0; i < n; ++i) y[i] = Q[i];
        t += hu;
        nst++;
        h = fmin(fmax(hu * fmin(fmax(fac, 0.2), 5.0), hmin), hmax);
      } else {
        nrej++;
        h = fmax(hu * fmax(fac, 0.2), hmin);
      }
    }
    return CELL_OK;
  }
};

}  // namespace pelec
