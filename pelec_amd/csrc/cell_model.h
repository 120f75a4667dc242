// Per-cell constant-volume rho*e reactor model on gfx950: RHS, exact Jacobian and T(e) inversion.
//
// Replaces the RHS/Jacobian kernels of the reference reactors (ReactorCvodeUtils /
// ReactorCvodeJacobian / ReactorUtils, CMake/BuildPelePhysicsLib.cmake:132-141) and the Fuego
// EOS T(e) inversion (REY2T, Source/Utilities.H:81-82). State per cell: y = [rhoY_0..N-1, T].
//   rho = sum rhoY,  e(t) = (rhoe0 + t*rhoe_src)/rho          (Source/React.cpp:212-219, 292)
//   T   = T(e(t), Y)  -- the state T is only the Newton starting guess
//   d rhoY_k/dt = W_k wdot_k + rhoY_src_k,  dT/dt = (rhoe_src - sum e_k d rhoY_k/dt)/(rho cv)
// The Jacobian is exact for that RHS: its T column is zero and dT/d rhoY_m = -e_m/(rho cv)
// enters every species column.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>

#ifndef PELEC_MEMORY_FENCE
#define PELEC_MEMORY_FENCE() __asm__ __volatile__("" ::: "memory")
#endif

namespace pelec {

// Mechanisms above this ODE size (the 53-species ones) call their production rates and rate
// Jacobian as out-of-line device functions: the straight-line kinetics of 229-325 reactions,
// inlined into every kernel that evaluates the model, made one library take tens of minutes to
// compile. Each kernel now calls one compiled copy; the arrays cross the call in private memory.
constexpr int KIN_INLINE_MAX_N = 32;

template <class Mech, bool MUL>
__device__ __noinline__ void wdot_outline(const double* C, double T, double* wd,
                                          const double* rmul) {
  Mech::template wdot<MUL>(C, T, wd, rmul);
}

template <class Mech, int n, typename JT, bool MUL>
__device__ __noinline__ void wdot_jac_outline(const double* C, double T, double* wd, JT* J,
                                              double* dwdT, const double* rmul) {
  Mech::template wdot_jac<n, JT, MUL>(C, T, wd, J, dwdT, rmul);
}

template <class Mech>
struct CellModel {
  static constexpr int N = Mech::NSP;
  static constexpr int n = N + 1;

  // Newton on e(T) = sum_k Y_k e_k(T) using the mixture polynomial; |dT| <= 1e-12 T, <= 20 its.
  __device__ static double T_from_e(double e, const double* Y, double T) {
    double pe[Mech::NTMID * 12];
    Mech::mix_e_poly(Y, pe);
    return T_from_pe(e, pe, T);
  }

  // The Newton iteration itself, from the mixture polynomial pe[(g*2 + b)*6 + i] (Tmid group g,
  // branch b = 0 below / 1 above Tmid, coefficient i of e(T)).
  __device__ static double T_from_pe(double e, const double* pe, double T) {
    for (int it = 0; it < 20; ++it) {
      double fv = -e, dv = 0.0;
#pragma unroll
      for (int g = 0; g < Mech::NTMID; ++g) {
        const bool lo = T < Mech::tmid(g);
        double c[6];
#pragma unroll
        for (int i = 0; i < 6; ++i) c[i] = lo ? pe[g * 12 + i] : pe[g * 12 + 6 + i];
        fv += c[0] + T * (c[1] + T * (c[2] + T * (c[3] + T * (c[4] + T * c[5]))));
        dv += c[1] + T * (2.0 * c[2] + T * (3.0 * c[3] + T * (4.0 * c[4] + T * 5.0 * c[5])));
      }
      const double dT = -fv / dv;
      T += dT;
      if (fabs(dT) <= 1e-12 * T) break;
    }
    return T;
  }

  // ydot = f(t, y); src (may be null) is rhoY_src_k at src[k * sstride]. Returns the T used.
  // MUL: reaction j's rate constant is multiplied by rmul[j] (the sensitivity/adjoint paths).
  template <bool MUL = false>
  __device__ static double rhs(double t, const double* y, double rhoe0, double rhoesrc,
                               const double* src, int64_t sstride, double* ydot,
                               const double* rmul = nullptr) {
    double rho = 0.0;
#pragma unroll
    for (int k = 0; k < N; ++k) rho += y[k];
    const double irho = 1.0 / rho;
    double Y[N], C[N];
#pragma unroll
    for (int k = 0; k < N; ++k) {
      Y[k] = y[k] * irho;
      C[k] = y[k] * Mech::IW[k];
    }
    const double T = T_from_e((rhoe0 + t * rhoesrc) * irho, Y, y[N]);
    double wd[N];
    if constexpr (n > KIN_INLINE_MAX_N)
      wdot_outline<Mech, MUL>(C, T, wd, rmul);
    else
      Mech::template wdot<MUL>(C, T, wd, rmul);
    double ek[N], cvk[N];
    Mech::e_cv(T, ek, cvk);
    double num = rhoesrc, D = 0.0;
#pragma unroll
    for (int k = 0; k < N; ++k) {
      double fk = Mech::W[k] * wd[k];
      if (src) fk += src[k * sstride];
      ydot[k] = fk;
      num -= ek[k] * fk;
      D += y[k] * cvk[k];
    }
    ydot[N] = num / D;
    return T;
  }

  // Exact Jacobian of rhs() into J (n x n row-major, element type JT). fe (may be null) receives
  // df/d(rho e0) = (df/dT at fixed rhoY) / (rho c_v): the column through which the initial energy
  // (and the energy forcing) enters the RHS.
  template <typename JT, bool MUL = false>
  __device__ static void jac(double t, const double* y, double rhoe0, double rhoesrc,
                             const double* src, int64_t sstride, JT* J, double* fe = nullptr,
                             const double* rmul = nullptr) {
    double rho = 0.0;
#pragma unroll
    for (int k = 0; k < N; ++k) rho += y[k];
    const double irho = 1.0 / rho;
    double Y[N], C[N];
#pragma unroll
    for (int k = 0; k < N; ++k) {
      Y[k] = y[k] * irho;
      C[k] = y[k] * Mech::IW[k];
    }
    const double T = T_from_e((rhoe0 + t * rhoesrc) * irho, Y, y[N]);
    double wd[N], dwdT[N];
    if constexpr (n > KIN_INLINE_MAX_N)
      wdot_jac_outline<Mech, n, JT, MUL>(C, T, wd, J, dwdT, rmul);
    else
      Mech::template wdot_jac<n, JT, MUL>(C, T, wd, J, dwdT, rmul);
    double ek[N], cvk[N], dcv[N];
    Mech::e_cv(T, ek, cvk);
    Mech::dcv_dT(T, dcv);
    double D = 0.0, dDdT = 0.0, num = rhoesrc, dNdT = 0.0;
    double B[N];
#pragma unroll
    for (int k = 0; k < N; ++k) {
      double fk = Mech::W[k] * wd[k];
      if (src) fk += src[k * sstride];
      D += y[k] * cvk[k];
      dDdT += y[k] * dcv[k];
      num -= ek[k] * fk;
      B[k] = Mech::W[k] * dwdT[k];
      dNdT -= cvk[k] * fk + ek[k] * B[k];
    }
    const double iD = 1.0 / D;
    const double fT = num * iD;
    const double dfTdT = (dNdT - fT * dDdT) * iD;
#pragma unroll
    for (int m = 0; m < N; ++m) {
      const double tau = -ek[m] * iD;
      double dNdx = 0.0;
#pragma unroll
      for (int k = 0; k < N; ++k) {
        const double A = Mech::W[k] * Mech::IW[m] * double(J[k * n + m]);
        dNdx -= ek[k] * A;
        J[k * n + m] = JT(A + B[k] * tau);
      }
      const double dfTdx = (dNdx - fT * cvk[m]) * iD;
      J[N * n + m] = JT(dfTdx + dfTdT * tau);
      PELEC_MEMORY_FENCE();  // one column at a time
    }
#pragma unroll
    for (int k = 0; k < n; ++k) J[k * n + N] = JT(0);
    if (fe) {
#pragma unroll
      for (int k = 0; k < N; ++k) fe[k] = B[k] * iD;
      fe[N] = dfTdT * iD;
    }
  }
};

}  // namespace pelec
