// Discrete adjoint of the BDF reactor (one cell per lane): gradients of the end state with respect
// to the per-reaction log pre-exponential factors, the initial rhoY and the initial rho*e.
//
// The forward integration is the plain cell integrator (bdf_cell.h) with its step-control decisions
// (step sizes, orders, accepted and rejected attempts) held fixed. Every operation on the Nordsieck
// array z is then either linear with recorded coefficients or the converged corrector of a step:
//   rescale      z_j *= eta^j                      (j = 1..q)
//   pascal       z_{j-1} += s z_j                  (predict s = +1, undo s = -1)
//   order up     z_L = A1 z_ia; z_j += l_j z_L     (j = 2..q)
//   order down   z_j -= l_j z_q                    (j = 2..q-1)
//   correct      z_j += l_j D (j = 0..q) [z_qmax = D], D = gamma f(t, z0 + D) - z1 / l_1
//   restart      z_1 = h f(t, z_0)
//   init         z_0 = y0, z_1 = h0 f(0, y0)
// The adjoint applies the transposes in reverse. For a correction the implicit function theorem
// gives, with M = I - gamma J(t, y) at the converged y and mu = sum_j l_j lambda_j:
//   nu = M^-T mu;  lambda_0 += gamma J^T nu;  lambda_1 -= nu / l_1;  g_p += gamma (df/dp)^T nu
// with df/dp_j = W nu_j q_j (reaction j's rate of progress; p_j = log of its pre-exponential
// multiplier) and df/d(rho e0) = (df/dT) / (rho c_v). The forward integration solves every corrector
// to machine precision (BdfCell POLISH), so the trajectory is the exactly-solved step map's and the
// adjoint is its exact transpose. The forward pass stores checkpoints (z, the
// integrator scalars and the Jacobian the modified Newton iteration reuses) every `seg` accepted
// steps; the backward pass re-runs each segment from its checkpoint with the tape on (the same
// decisions, deterministically) and walks it backwards.
#pragma once
#include "bdf_cell.h"

namespace pelec {

enum TapeOp : int {
  OP_RESCALE = 1,
  OP_PASCAL = 2,
  OP_ORDER_UP = 3,
  OP_ORDER_DOWN = 4,
  OP_CORRECT = 5,
  OP_RESTART = 6,
  OP_INIT = 7
};

// Tape record: 16 header words + one state vector (n doubles):
//   [0] op  [1] q  [2] L / save  [3] ia / qmax  [4] eta | sign | A1 | gamma | h
//   [5..10] l_0..l_5  [11] 1/l_1  [12] t  [16..16+n) y (correct: converged y; restart, init: z_0)
template <int n>
struct GTape {
  static constexpr int RS = 16 + n;
  double* base = nullptr;  // this cell's records; null: not recording
  int cap = 0;
  int count = 0;
  bool overflow = false;
  double* last_correct = nullptr;

  __device__ double* next(int op) {
    if (!base) return nullptr;
    if (count >= cap) {
      overflow = true;
      return nullptr;
    }
    double* r = base + (int64_t)count * RS;
    ++count;
    r[0] = op;
    return r;
  }
  template <class V>
  __device__ void rec_rescale(const V&, int q, double eta) {
    if (double* r = next(OP_RESCALE)) {
      r[1] = q;
      r[4] = eta;
    }
  }
  template <class V>
  __device__ void rec_pascal(const V&, int q, double sign) {
    if (double* r = next(OP_PASCAL)) {
      r[1] = q;
      r[4] = sign;
    }
  }
  template <class V>
  __device__ void rec_order_up(const V&, int q, int L, int ia, double A1, const double* l) {
    if (double* r = next(OP_ORDER_UP)) {
      r[1] = q;
      r[2] = L;
      r[3] = ia;
      r[4] = A1;
      for (int j = 0; j <= BDF_QMAX; ++j) r[5 + j] = l[j];
    }
  }
  template <class V>
  __device__ void rec_order_down(const V&, int q, const double* l) {
    if (double* r = next(OP_ORDER_DOWN)) {
      r[1] = q;
      for (int j = 0; j <= BDF_QMAX; ++j) r[5 + j] = l[j];
    }
  }
  template <class V>
  __device__ void rec_correct(const V& v, const BdfScalars& S, bool save, int qmax) {
    last_correct = nullptr;
    if (double* r = next(OP_CORRECT)) {
      last_correct = r;
      r[1] = S.q;
      r[2] = save ? 1 : 0;
      r[3] = qmax;
      r[4] = S.gamma;
      for (int j = 0; j <= BDF_QMAX; ++j) r[5 + j] = S.l[j];
      r[11] = 1.0 / S.l[1];
      r[12] = S.tn;
      for (int i = 0; i < n; ++i) r[16 + i] = v.Z(0, i) + v.A(i);
    }
  }
  // the order-increase decision saves the last correction in column qmax as well
  __device__ void rec_mark_save() {
    if (last_correct) last_correct[2] = 1;
  }
  template <class V>
  __device__ void rec_restart(const V& v, double h, double t) {
    if (double* r = next(OP_RESTART)) {
      r[4] = h;
      r[12] = t;
      for (int i = 0; i < n; ++i) r[16 + i] = v.Z(0, i);
    }
  }
  template <class V>
  __device__ void rec_init(const V& v, double h) {
    if (double* r = next(OP_INIT)) {
      r[4] = h;
      r[12] = 0.0;
      for (int i = 0; i < n; ++i) r[16 + i] = v.Z(0, i);
    }
  }
};

// x = A^-T b for the factors of lu_factor (P A = L U, interchanges in factorisation order):
// U^T w = b, L^T x = w, then the interchanges undone in reverse order.
template <int n>
__device__ inline void lu_solve_transposed(const double* a, const int* piv, double* b) {
  for (int i = 0; i < n; ++i) {
    double s = b[i];
    for (int j = 0; j < i; ++j) s -= a[j * n + i] * b[j];
    b[i] = s / a[i * n + i];
  }
  for (int i = n - 1; i >= 0; --i) {
    double s = b[i];
    for (int j = i + 1; j < n; ++j) s -= a[j * n + i] * b[j];
    b[i] = s;
  }
  for (int k = n - 1; k >= 0; --k) {
    const int p = piv[k];
    const double t = b[k];
    b[k] = b[p];
    b[p] = t;
  }
}

template <class Mech>
struct Adjoint {
  using Model = CellModel<Mech>;
  static constexpr int N = Mech::NSP;
  static constexpr int n = N + 1;
  static constexpr int R = Mech::NRXN;

  // problem (as in the forward cell)
  double rhoe0, rhoesrc;
  const double* src;
  int64_t sstride;
  const double* rmul;
  // adjoint state
  double lam[BDF_QMAX + 1][n];
  double gp[R];
  double ge;
  double gy0[n];
  // scratch: J and Mf point to n x n blocks of device memory (per cell)
  double* J;
  double* Mf;
  double fe[n];
  int piv[n];

  __device__ void reset() {
    for (int j = 0; j <= BDF_QMAX; ++j)
      for (int i = 0; i < n; ++i) lam[j][i] = 0.0;
    for (int j = 0; j < R; ++j) gp[j] = 0.0;
    ge = 0.0;
  }

  __device__ void jacobian(double t, const double* y) {
    Model::template jac<double, true>(t, y, rhoe0, rhoesrc, src, sstride, J, fe, rmul);
  }

  // out += s J^T v;  gp += s (df/dp)^T v;  ge += s fe . v   (J, fe evaluated at (t, y))
  __device__ void vjp(double t, const double* y, const double* v, double s, double* out) {
    for (int m = 0; m < n; ++m) {
      double acc = 0.0;
      for (int k = 0; k < n; ++k) acc += J[k * n + m] * v[k];
      out[m] += s * acc;
    }
    double fev = 0.0;
    for (int k = 0; k < n; ++k) fev += fe[k] * v[k];
    ge += s * fev;
    // df_k/dp_j = W_k nu_kj q_j (species rows), -(sum_k e_k W_k nu_kj q_j) / (rho c_v) (T row)
    double rho = 0.0;
    for (int k = 0; k < N; ++k) rho += y[k];
    const double irho = 1.0 / rho;
    double Y[N], C[N];
    for (int k = 0; k < N; ++k) {
      Y[k] = y[k] * irho;
      C[k] = y[k] * Mech::IW[k];
    }
    const double T = Model::T_from_e((rhoe0 + t * rhoesrc) * irho, Y, y[N]);
    double ek[N], cvk[N];
    Mech::e_cv(T, ek, cvk);
    double D = 0.0;
    for (int k = 0; k < N; ++k) D += y[k] * cvk[k];
    const double vT = v[N] / D;
    double w[N];
    for (int k = 0; k < N; ++k) w[k] = Mech::W[k] * (v[k] - ek[k] * vT);
    double q[R], nw[R];
    Mech::template progress<true>(C, T, q, rmul);
    Mech::nu_dot(w, nw);
    for (int j = 0; j < R; ++j) gp[j] += s * q[j] * nw[j];
  }

  // Adjoint of one tape record (records are walked last to first).
  __device__ void back(const double* r) {
    const int op = (int)r[0];
    const int q = (int)r[1];
    switch (op) {
      case OP_RESCALE: {
        double f = 1.0;
        for (int j = 1; j <= BDF_QMAX; ++j) {
          f *= r[4];
          if (j <= q)
            for (int i = 0; i < n; ++i) lam[j][i] *= f;
        }
        break;
      }
      case OP_PASCAL: {
        // forward order: k = 1..q, j = q..k: z_{j-1} += s z_j; transpose in reverse order
        const double sg = r[4];
        for (int k = q; k >= 1; --k)
          for (int j = k; j <= q; ++j)
            for (int i = 0; i < n; ++i) lam[j][i] += sg * lam[j - 1][i];
        break;
      }
      case OP_ORDER_UP: {
        const int L = (int)r[2], ia = (int)r[3];
        const double A1 = r[4];
        for (int i = 0; i < n; ++i) {
          double lL = lam[L][i];
          for (int j = 2; j <= q; ++j) lL += r[5 + j] * lam[j][i];
          if (ia == L) {
            lam[L][i] = A1 * lL;
          } else {
            lam[L][i] = 0.0;
            lam[ia][i] += A1 * lL;
          }
        }
        break;
      }
      case OP_ORDER_DOWN: {
        for (int i = 0; i < n; ++i) {
          double s = 0.0;
          for (int j = 2; j < q; ++j) s += r[5 + j] * lam[j][i];
          lam[q][i] -= s;
        }
        break;
      }
      case OP_CORRECT: {
        const bool save = r[2] != 0.0;
        const int qmax = (int)r[3];
        const double gamma = r[4], rl1 = r[11], t = r[12];
        const double* y = r + 16;
        double nu[n];
        for (int i = 0; i < n; ++i) {
          double mu = 0.0;
          for (int j = 0; j <= q; ++j) mu += r[5 + j] * lam[j][i];
          if (save) {
            mu += lam[qmax][i];
            lam[qmax][i] = 0.0;
          }
          nu[i] = mu;
        }
        jacobian(t, y);
        for (int k = 0; k < n * n; ++k) Mf[k] = -gamma * J[k];
        for (int k = 0; k < n; ++k) Mf[k * n + k] += 1.0;
        (void)lu_factor<n, double>(Mf, piv);
        lu_solve_transposed<n>(Mf, piv, nu);
        vjp(t, y, nu, gamma, lam[0]);
        for (int i = 0; i < n; ++i) lam[1][i] -= rl1 * nu[i];
        break;
      }
      case OP_RESTART: {
        const double h = r[4], t = r[12];
        const double* z0 = r + 16;
        double l1[n];
        for (int i = 0; i < n; ++i) {
          l1[i] = lam[1][i];
          lam[1][i] = 0.0;
        }
        jacobian(t, z0);
        vjp(t, z0, l1, h, lam[0]);
        break;
      }
      case OP_INIT: {
        const double h = r[4];
        const double* y0 = r + 16;
        for (int i = 0; i < n; ++i) gy0[i] = lam[0][i];
        jacobian(0.0, y0);
        vjp(0.0, y0, lam[1], h, gy0);
        break;
      }
      default:
        break;
    }
  }

  // Seed from the cotangents of the outputs: rhoY_end (ct_rY[k * cs]) and
  // T_end = T(e_end, Y_end) (ct_T) with e_end = (rho e0 + dt rho e_src) / rho.
  __device__ void seed(const double* y, double dt, const double* ct_rY, int64_t cs, double ct_T) {
    double rho = 0.0;
    for (int k = 0; k < N; ++k) rho += y[k];
    const double irho = 1.0 / rho;
    double Y[N];
    for (int k = 0; k < N; ++k) Y[k] = y[k] * irho;
    const double T = Model::T_from_e((rhoe0 + dt * rhoesrc) * irho, Y, y[N]);
    double ek[N], cvk[N];
    Mech::e_cv(T, ek, cvk);
    double D = 0.0;
    for (int k = 0; k < N; ++k) D += y[k] * cvk[k];
    for (int k = 0; k < N; ++k) lam[0][k] = (ct_rY ? ct_rY[k * cs] : 0.0) - ct_T * ek[k] / D;
    lam[0][N] = 0.0;
    ge += ct_T / D;  // dT_end / d(rho e0) = 1 / (rho c_v)
  }
};

}  // namespace pelec
