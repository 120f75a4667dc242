// Cooperative form of the constant-volume rho*e reactor model: one group of G lanes evaluates one
// cell's RHS and exact Jacobian (G = 32, two cells per 64-lane wave, for n <= 32; G = 64, one cell
// per wave, for the 53-species mechanisms).
//
// Same mathematics as CellModel (cell_model.h: ReactorCvodeUtils / ReactorCvodeJacobian /
// ReactorUtils of the reference, CMake/BuildPelePhysicsLib.cmake:132-141), laid out for lanes:
//   - lane k of the group holds state component k (rhoY_k for k < N, T for k = N);
//   - species work (thermo, production rates, Jacobian row k) runs on lane k;
//   - reaction j runs on lane (position of j in Mech::T_RORD) mod G, in ceil(R/G) rounds;
//   - per-species and per-reaction results are exchanged through the cell's LDS block;
//   - sums over components are xor-butterfly reductions. IEEE addition is commutative, so every
//     lane of the group ends with the same bits and the group takes identical scalar decisions.
// The mechanism comes from the generated data tables (Mech::T_*), so every lane runs one code path.
#pragma once
#include <hip/hip_runtime.h>

#include "cell_model.h"

namespace pelec {

// Lane exchanges that avoid the LDS crossbar (ds_bpermute, which __shfl uses): v_readlane for a
// lane index that is the same on every lane of the wave, DPP row rotations inside rows of 16.

// x of lane src. src must be wave-uniform (a constant, a loop constant or a scalar value): the
// value is read into scalar registers, so every lane receives the same one.
__device__ inline int lane_read(int x, int src) { return __builtin_amdgcn_readlane(x, src); }
__device__ inline float lane_read(float x, int src) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), src));
}
__device__ inline double lane_read(double x, int src) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(x), src);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(x), src);
  return __hiloint2double(hi, lo);
}

// x of lane (i -+ k) mod 16 within lane i's row of 16 (DPP row_ror:k).
template <int k>
__device__ inline int row_ror(int x) {
  static_assert(k >= 1 && k <= 15, "row_ror:1..15");
  return __builtin_amdgcn_mov_dpp(x, 0x120 + k, 0xf, 0xf, false);
}
template <int k>
__device__ inline float row_ror(float x) {
  return __int_as_float(row_ror<k>(__float_as_int(x)));
}
template <int k>
__device__ inline double row_ror(double x) {
  return __hiloint2double(row_ror<k>(__double2hiint(x)), row_ror<k>(__double2loint(x)));
}

// Partner value of the step with offset `off` of an xor butterfly over a group of G >= 16 lanes
// run with decreasing offsets (G/2, ..., 2, 1). Offsets of 16 and more cross rows and use a
// shuffle. Below 16, row_ror:off reads lane (i -+ off) mod 16, which is i ^ off with some of the
// higher bits of the row flipped by the carry; after the steps with the larger offsets every lane
// holds the same value as each lane that differs from it only in those bits, so the value read is
// the one lane i ^ off holds: the butterfly combines the same operands in the same order.
template <int G, int off, typename T>
__device__ inline T xor_partner(T x) {
  static_assert(G >= 16 && G <= 64 && off < G, "DPP butterfly within a group of 16..64 lanes");
  if constexpr (off >= 16)
    return __shfl_xor(x, off, G);
  else
    return row_ror<off>(x);
}

// x of lane src of the caller's group (src: 0..G-1, wave-uniform), one v_readlane per group of
// the wave.
template <int G, typename T>
__device__ inline T group_read(T x, int src) {
  T r = lane_read(x, src);
#pragma unroll
  for (int g = 1; g < 64 / G; ++g) {
    const T v = lane_read(x, g * G + src);
    r = (int)__lane_id() >= g * G ? v : r;
  }
  return r;
}

template <int G>
struct Group {
  static_assert(G <= 64 && G >= 16 && (G & (G - 1)) == 0,
                "group of a power-of-two lanes (16..64) within a wave");
  // butterfly sums and maxima: offsets G/2 .. 1 (IEEE addition and fmax are commutative, so
  // every lane of the group ends with the same bits)
  __device__ static double sum(double x) {
    if constexpr (G > 32) x += xor_partner<G, 32>(x);
    if constexpr (G > 16) x += xor_partner<G, 16>(x);
    x += xor_partner<G, 8>(x);
    x += xor_partner<G, 4>(x);
    x += xor_partner<G, 2>(x);
    x += xor_partner<G, 1>(x);
    return x;
  }
  __device__ static double max(double x) {
    if constexpr (G > 32) x = fmax(x, xor_partner<G, 32>(x));
    if constexpr (G > 16) x = fmax(x, xor_partner<G, 16>(x));
    x = fmax(x, xor_partner<G, 8>(x));
    x = fmax(x, xor_partner<G, 4>(x));
    x = fmax(x, xor_partner<G, 2>(x));
    x = fmax(x, xor_partner<G, 1>(x));
    return x;
  }
  // x of lane src of the group; src is wave-uniform at every call site (a loop constant)
  __device__ static double bcast(double x, int src) { return group_read<G>(x, src); }
  __device__ static float bcast(float x, int src) { return group_read<G>(x, src); }
  // LDS written by one lane and read by another lane of the same wave: LDS operations of a wave
  // are performed in order, so ordering the compiler's accesses is enough.
  __device__ static void sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
};

// The cell's LDS block as an LDS-qualified pointer. The model functions below are compiled out of
// line (rhs, jac, T_from_e), and through a generic pointer they would address the block with flat
// instructions, which go through the vector memory path and wait on both memory counters. Every
// caller passes a __shared__ block, so the cast (the low 32 bits of the flat address, the LDS
// offset) is exact.
typedef __attribute__((address_space(3))) double lds_double;
__device__ inline lds_double* lds_ptr(double* p) { return (lds_double*)p; }

// x, made to depend on dep for the compiler: table loads indexed by the result cannot be hoisted
// above the computation of dep (the tables are constant, so the compiler would otherwise issue
// every record's loads up front and keep them live across the T(e) iteration).
template <class X, class D>
__device__ inline X after(X x, D dep) {
  __asm__("" : "+v"(x) : "v"(dep));
  return x;
}

template <class Mech, int G = 32>
struct GroupModel {
  static constexpr int GS = G;  // lanes per cell
  using Grp = Group<G>;
  using Model = CellModel<Mech>;
  static constexpr int N = Mech::NSP;
  static constexpr int n = N + 1;
  static constexpr int R = Mech::NRXN;
  static constexpr int NSETS = Mech::T_NSETS;
  static constexpr int NS = NSETS > 0 ? NSETS : 1;
  static constexpr int RROUNDS = (R + G - 1) / G;
  // species -> reaction gathers load their table entries in batches of CB (RHS) / CBJ (Jacobian)
  static constexpr int CB = 8, CBJ = 4;
  static_assert(Mech::T_CSR_PAD >= CB && Mech::T_CSR_PAD >= CBJ, "table padding covers a batch");
  static_assert(n <= G, "one state component per lane");
  static_assert(NSETS <= G, "one efficiency set per lane");

  // LDS block of one cell, in doubles: RHS part
  static constexpr int L_C = 0;             // C_k
  static constexpr int L_EG = L_C + N;      // exp(g_k/RT)
  static constexpr int L_M = L_EG + N;      // third-body concentration per efficiency set
  static constexpr int L_Q = L_M + NS;      // q_j
  static constexpr int NPE = Mech::NTMID * 12;
  static constexpr int L_PE = L_Q + R;      // mixture e(T) polynomial
  static constexpr int RHS_WORDS = L_PE + NPE;
  static_assert(NPE <= G, "one mixture-polynomial coefficient per lane");
  // Jacobian part
  static constexpr int L_HRT = RHS_WORDS;   // h_k/RT
  static constexpr int L_EK = L_HRT + N;    // e_k
  static constexpr int L_CV = L_EK + N;     // cv_k
  static constexpr int L_DQT = L_CV + N;    // dq_j/dT
  static constexpr int L_DQM = L_DQT + R;   // dkeff_j/dM * net_j
  static constexpr int L_DQ = L_DQM + R;    // dq_j/dC of the 3 reactant and 3 product slots
  static constexpr int JS = (n % 2 == 0) ? n + 1 : n;  // row stride of the matrix (odd)
  static constexpr int L_J = L_DQ + 6 * R;  // n rows of JS
  static constexpr int JAC_WORDS = L_J + n * JS;

  // ---- species thermo of lane k (folded NASA-7, the generated straight-line code's coefficients)
  __device__ static double horner(double T, double c0, double c1, double c2, double c3, double c4) {
    return fma(T, fma(T, fma(T, fma(T, c4, c3), c2), c1), c0);
  }
  __device__ static double g_RT(int k, bool lo, double T, double lnT, double invT) {
    const double* c = Mech::T_G[k][lo ? 0 : 1];
    return horner(T, c[0], c[1], c[2], c[3], c[4]) - c[5] * lnT + c[6] * invT;
  }
  __device__ static double h_RT(int k, bool lo, double T, double invT) {
    const double* c = Mech::T_H[k][lo ? 0 : 1];
    return horner(T, c[0], c[1], c[2], c[3], c[4]) + c[5] * invT;
  }
  __device__ static void e_cv(int k, bool lo, double T, double& e, double& cv) {
    const double* ce = Mech::T_E[k][lo ? 0 : 1];
    const double* cc = Mech::T_CV[k][lo ? 0 : 1];
    e = fma(T, horner(T, ce[1], ce[2], ce[3], ce[4], ce[5]), ce[0]);
    cv = horner(T, cc[0], cc[1], cc[2], cc[3], cc[4]);
  }
  __device__ static double dcv(int k, bool lo, double T) {
    const double* c = Mech::T_DCV[k][lo ? 0 : 1];
    return fma(T, fma(T, fma(T, c[3], c[2]), c[1]), c[0]);
  }

  // T(e, Y): mixture polynomial coefficient i on lane i (a sum over species in species order,
  // through LDS), then CellModel's Newton iteration, replicated on every lane with identical
  // inputs. Yl: the lane's mass fraction (0 on lanes >= N). Uses L_C (as Y) and L_PE of sh.
  // Not inlined: inlined, its iteration and the kinetics that follow were scheduled together
  // and the register count of the RHS doubled.
  __device__ __noinline__ static double T_from_e(double e, double Yl, int l, double Tguess,
                                                 double* shg) {
    lds_double* sh = lds_ptr(shg);
    if (l < N) sh[L_C + l] = Yl;
    Grp::sync();
    if (l < NPE) {
      const int g = l / 12, i = l - 12 * g;
      double acc = 0.0;
      // species in batches of CB: a batch's loads are issued together (every entry read is in
      // range, whatever its temperature group), then added in species order
#pragma unroll
      for (int k0 = 0; k0 < N; k0 += CB) {
        double yv[CB], ev[CB];
#pragma unroll
        for (int u = 0; u < CB; ++u) {
          const int k = k0 + u < N ? k0 + u : N - 1;
          yv[u] = sh[L_C + k];
          ev[u] = (&Mech::T_E[k][0][0])[i];
        }
#pragma unroll
        for (int u = 0; u < CB; ++u)
          if (k0 + u < N && Mech::T_TMIDG[k0 + u] == g) acc += yv[u] * ev[u];
      }
      sh[L_PE + l] = acc;
    }
    Grp::sync();
    double pe[NPE];
#pragma unroll
    for (int i = 0; i < NPE; ++i) pe[i] = sh[L_PE + i];
    Grp::sync();  // L_C is rewritten with concentrations next
    return Model::T_from_pe(e, pe, Tguess);
  }

  // ---- reactions: lane-level evaluation of reaction j from the tables. Reads C, exp(g/RT), M
  // (and h/RT when JAC) from the cell's LDS block; writes q_j (and its derivatives when JAC).
  template <bool JAC>
  __device__ static void reaction(int j, double T, double lnT, double invT, lds_double* sh) {
    const int kind = Mech::T_KIND[j];
    const double lnk = Mech::T_LNA[j] + Mech::T_B[j] * lnT + (-Mech::T_EAR[j]) * invT;
    const double karr = Mech::T_KCONST[j] ? Mech::T_A[j] : exp(lnk);
    double kf = karr, dkdM = 0.0, dlnk = 0.0;
    if (JAC) dlnk = (Mech::T_B[j] + Mech::T_EAR[j] * invT) * invT;
    if (kind == 1) {
      const double Mj = sh[L_M + Mech::T_EFF[j]];
      kf = karr * Mj;
      dkdM = karr;
    } else if (kind == 2) {
      const double Mj = sh[L_M + Mech::T_EFF[j]];
      const double* lw = Mech::T_LOW[j];
      const double k0r = exp(lw[0] + lw[1] * lnT + (-lw[2]) * invT);
      const double Pr = fmax(k0r * Mj, 1e-300);
      const double iPr1 = 1.0 / (1.0 + Pr);
      const int tk = Mech::T_TROEK[j];
      const double* tr = Mech::T_TROE[j];
      double F = 1.0, g = 0.0, dlnF_T = 0.0;
      if (tk == 1) {
        const double lFc = tr[0], cc = tr[1], nn = tr[2];
        const double xx = log10(Pr) + cc;
        const double DD = 1.0 / (nn - 0.14 * xx);
        const double f1 = xx * DD;
        const double iden = 1.0 / (1.0 + f1 * f1);
        F = exp(tr[3] * iden);
        if (JAC) g = -lFc * 2.0 * f1 * nn * DD * DD * iden * iden;
      } else if (tk == 2) {
        const double e3 = exp(tr[1] * T), e1 = exp(tr[3] * T);
        const double e2 = tr[5] != 0.0 ? exp(-tr[4] * invT) : 0.0;
        const double Fc = tr[0] * e3 + tr[2] * e1 + e2;
        const double lFc = log10(Fc);
        const double cc = -0.4 - 0.67 * lFc, nn = 0.75 - 1.27 * lFc;
        const double xx = log10(Pr) + cc;
        const double DD = 1.0 / (nn - 0.14 * xx);
        const double f1 = xx * DD;
        const double iden = 1.0 / (1.0 + f1 * f1);
        F = exp(Mech::T_LN10 * lFc * iden);
        if (JAC) {
          g = -lFc * 2.0 * f1 * nn * DD * DD * iden * iden;
          const double df1 = (-0.67 * nn + 1.27 * xx) * DD * DD;
          const double dlF = iden - lFc * 2.0 * f1 * df1 * iden * iden;
          const double dFc = tr[6] * e3 + tr[7] * e1 + tr[4] * invT * invT * e2;
          dlnF_T = dlF * dFc / Fc;
        }
      }
      const double kinf = karr;
      kf = kinf * Pr * iPr1 * F;
      if (JAC) {
        const double dlnPr = (lw[1] + lw[2] * invT) * invT;
        dkdM = kinf * k0r * F * (iPr1 * iPr1 + g * iPr1);
        dlnk = dlnk + dlnPr * iPr1 + g * dlnPr + dlnF_T;
      }
    }
    // mass action, 1/Kc
    const int* rs = Mech::T_RSP[j];
    const int* rn = Mech::T_RNU[j];
    const int* ps = Mech::T_PSP[j];
    const int* pn = Mech::T_PNU[j];
    double cr[3], cp[3];
#pragma unroll
    for (int s = 0; s < 3; ++s) {
      cr[s] = rs[s] >= 0 ? sh[L_C + rs[s]] : 1.0;
      cp[s] = ps[s] >= 0 ? sh[L_C + ps[s]] : 1.0;
    }
    double Pf = 1.0, Pb = 1.0;
#pragma unroll
    for (int s = 0; s < 3; ++s)
#pragma unroll
      for (int t = 0; t < 3; ++t) {
        if (t < rn[s]) Pf *= cr[s];
        if (t < pn[s]) Pb *= cp[s];
      }
    const bool rev = Mech::T_REV[j] != 0;
    double iKc = 0.0;
    if (rev) {
      double num = 1.0, den = 1.0;
#pragma unroll
      for (int s = 0; s < 3; ++s) {
        const double egp = ps[s] >= 0 ? sh[L_EG + ps[s]] : 1.0;
        const double egr = rs[s] >= 0 ? sh[L_EG + rs[s]] : 1.0;
#pragma unroll
        for (int t = 0; t < 3; ++t) {
          if (t < pn[s]) num *= egp;
          if (t < rn[s]) den *= egr;
        }
      }
      const int dnu = Mech::T_DNU[j];
      const double refC = Mech::T_REFC * invT, irefC = Mech::T_IREFC * T;
      const double rc = dnu > 0 ? irefC : refC;
      if (dnu != 0) num *= rc;
      if (dnu == 2 || dnu == -2) num *= rc;
      iKc = num / den;
    }
    const double net = rev ? Pf - iKc * Pb : Pf;
    sh[L_Q + j] = kf * net;
    if (JAC) {
      double dqdT;
      if (rev) {
        double sh_ = 0.0;
        const int* ns = Mech::T_NSPC[j];
        const double* nv = Mech::T_NNU[j];
#pragma unroll
        for (int s = 0; s < 6; ++s)
          if (ns[s] >= 0) sh_ += nv[s] * sh[L_HRT + ns[s]];
        const double dlnKc = (sh_ - double(Mech::T_DNU[j])) * invT;
        dqdT = kf * (dlnk * net + iKc * Pb * dlnKc);
      } else {
        dqdT = kf * dlnk * net;
      }
      sh[L_DQT + j] = dqdT;
      sh[L_DQM + j] = kind != 0 ? dkdM * net : 0.0;
      // d(prod)/dC of each slot: nu C^(nu-1) times the other slots' factors
      lds_double* dq = sh + L_DQ + 6 * j;
#pragma unroll
      for (int s = 0; s < 3; ++s) {
        double df = rn[s] == 1 ? 1.0 : (rn[s] == 2 ? 2.0 * cr[s] : 3.0 * cr[s] * cr[s]);
        double db = pn[s] == 1 ? 1.0 : (pn[s] == 2 ? 2.0 * cp[s] : 3.0 * cp[s] * cp[s]);
#pragma unroll
        for (int o = 0; o < 3; ++o) {
          if (o == s) continue;
#pragma unroll
          for (int t = 0; t < 3; ++t) {
            if (t < rn[o]) df *= cr[o];
            if (t < pn[o]) db *= cp[o];
          }
        }
        dq[s] = rs[s] >= 0 ? kf * df : 0.0;
        dq[3 + s] = (rev && ps[s] >= 0) ? -kf * iKc * db : 0.0;
      }
    }
  }

  // Species lane k's thermo into LDS, reactions, and the production rate of species k.
  template <bool JAC>
  __device__ static double kinetics(double T, double Cl, int l, lds_double* sh, double* eko,
                                    double* cvo) {
    const bool sp = l < N;
    const int k = after(sp ? l : 0, T);
    const double lnT = log(T), invT = 1.0 / T;
    const bool lo = T < Mech::T_TMID[k];
    if (sp) {
      sh[L_C + k] = Cl;
      sh[L_EG + k] = exp(g_RT(k, lo, T, lnT, invT));
      if (JAC) sh[L_HRT + k] = h_RT(k, lo, T, invT);
      e_cv(k, lo, T, *eko, *cvo);
      if (JAC) {
        sh[L_EK + k] = *eko;
        sh[L_CV + k] = *cvo;
      }
    }
    Grp::sync();
    // third-body concentrations: set s on lane s, in the straight-line code's order
    if (l < NSETS) {
      double ctot = 0.0;
      for (int m = 0; m < N; ++m) ctot += sh[L_C + m];
      double M = ctot;
      for (int m = 0; m < N; ++m) M += Mech::T_EFFD[l][m] * sh[L_C + m];
      sh[L_M + l] = M;
    }
    Grp::sync();
#pragma unroll 1
    for (int r = 0; r < RROUNDS; ++r) {
      const int idx = l + G * r;
      if (idx < R) reaction<JAC>(after(Mech::T_RORD[idx], lnT), T, lnT, invT, sh);
    }
    Grp::sync();
    double wd = 0.0;
    if (sp) {
      // in batches of CB entries: the table loads of a batch are independent, so a lane waits
      // for one batch at a time instead of one entry at a time; same summation order
      const int e = Mech::T_CSR_PTR[k + 1];
      for (int p0 = Mech::T_CSR_PTR[k]; p0 < e; p0 += CB) {
        int jj[CB];
        double nu[CB], q[CB];
#pragma unroll
        for (int u = 0; u < CB; ++u) {
          jj[u] = Mech::T_CSR_J[p0 + u];
          nu[u] = Mech::T_CSR_NU[p0 + u];
        }
        // entries past this species' range name valid reactions (the next species' entries or
        // the table's padding), so the batch's q_j are read together
#pragma unroll
        for (int u = 0; u < CB; ++u) q[u] = sh[L_Q + jj[u]];
#pragma unroll
        for (int u = 0; u < CB; ++u)
          if (p0 + u < e) wd += nu[u] * q[u];
      }
    }
    return wd;
  }

  // ydot for the lane's component (0 on lanes >= n). yl: the lane's state component; srcl: the
  // lane's rhoY source (0 when there is none); sh: the cell's LDS block (RHS_WORDS doubles).
  __device__ static double rhs(double t, double yl, int l, double rhoe0, double rhoesrc,
                               double srcl, double* shg) {
    lds_double* sh = lds_ptr(shg);
    const bool sp = l < N;
    const int k = sp ? l : 0;
    const double rho = Grp::sum(sp ? yl : 0.0);
    const double irho = 1.0 / rho;
    const double Tg = Grp::bcast(yl, N);
    const double T = T_from_e((rhoe0 + t * rhoesrc) * irho, sp ? yl * irho : 0.0, l, Tg, shg);
    double ek = 0.0, cvk = 0.0;
    const double wd = kinetics<false>(T, yl * Mech::IW[k], l, sh, &ek, &cvk);
    const double fk = sp ? Mech::W[k] * wd + srcl : 0.0;
    const double num = rhoesrc - Grp::sum(sp ? ek * fk : 0.0);
    const double D = Grp::sum(sp ? yl * cvk : 0.0);
    Grp::sync();  // the block's reads are done before the next use rewrites it
    return sp ? fk : (l == N ? num / D : 0.0);
  }

  // Exact Jacobian of rhs() (row-major n x n, the layout of CellModel::jac) written to Jg by the
  // group; sh: the cell's LDS block (JAC_WORDS doubles).
  __device__ static void jac(double t, double yl, int l, double rhoe0, double rhoesrc, double srcl,
                             double* shg, double* Jg) {
    lds_double* sh = lds_ptr(shg);
    const bool sp = l < N;
    const int k = sp ? l : 0;
    const double rho = Grp::sum(sp ? yl : 0.0);
    const double irho = 1.0 / rho;
    const double Tg = Grp::bcast(yl, N);
    const double T = T_from_e((rhoe0 + t * rhoesrc) * irho, sp ? yl * irho : 0.0, l, Tg, shg);
    double ek = 0.0, cvk = 0.0;
    const double Cl = yl * Mech::IW[k];
    // kinetics<true> also leaves the per-reaction derivatives in LDS
    const double wd = kinetics<true>(T, Cl, l, sh, &ek, &cvk);
    lds_double* row = sh + L_J + k * JS;
    double dwdT = 0.0;
    if (sp) {
      double tb[NS];
#pragma unroll
      for (int s = 0; s < NS; ++s) tb[s] = 0.0;
      for (int m = 0; m < N; ++m) row[m] = 0.0;
      const int e = Mech::T_CSR_PTR[k + 1];
      for (int p0 = Mech::T_CSR_PTR[k]; p0 < e; p0 += CBJ) {
        int jj[CBJ], sl[CBJ][6], ef[CBJ];
        double nu[CBJ];
#pragma unroll
        for (int u = 0; u < CBJ; ++u) {
          jj[u] = Mech::T_CSR_J[p0 + u];
          nu[u] = Mech::T_CSR_NU[p0 + u];
          ef[u] = Mech::T_CSR_EFF[p0 + u];
#pragma unroll
          for (int q = 0; q < 6; ++q) sl[u][q] = Mech::T_CSR_SLOT[p0 + u][q];
        }
#pragma unroll
        for (int u = 0; u < CBJ; ++u) {
          if (p0 + u >= e) break;
          const int j = jj[u];
          dwdT += nu[u] * sh[L_DQT + j];
          const lds_double* dq = sh + L_DQ + 6 * j;
#pragma unroll
          for (int q = 0; q < 6; ++q)
            if (sl[u][q] >= 0) row[sl[u][q]] += nu[u] * dq[q];
          if (ef[u] >= 0) {
            const double c = nu[u] * sh[L_DQM + j];
#pragma unroll
            for (int q = 0; q < NS; ++q) tb[q] += (q == ef[u]) ? c : 0.0;
          }
        }
      }
      if (NSETS > 0) {
        for (int m = 0; m < N; ++m) {
          double acc = row[m];
#pragma unroll
          for (int s = 0; s < NS; ++s) acc += tb[s] * Mech::T_EFFV[s][m];
          row[m] = acc;
        }
      }
    }
    // energy row and the T-elimination (CellModel::jac)
    const double fk = sp ? Mech::W[k] * wd + srcl : 0.0;
    const double Bk = sp ? Mech::W[k] * dwdT : 0.0;
    const bool lo = T < Mech::T_TMID[k];
    const double dcvk = sp ? dcv(k, lo, T) : 0.0;
    const double D = Grp::sum(sp ? yl * cvk : 0.0);
    const double dDdT = Grp::sum(sp ? yl * dcvk : 0.0);
    const double num = rhoesrc - Grp::sum(sp ? ek * fk : 0.0);
    const double dNdT = -Grp::sum(sp ? cvk * fk + ek * Bk : 0.0);
    const double iD = 1.0 / D;
    const double fT = num * iD;
    const double dfTdT = (dNdT - fT * dDdT) * iD;
    if (sp) {
      const double Wk = Mech::W[k];
      for (int m = 0; m < N; ++m) row[m] = Wk * Mech::IW[m] * row[m];
    }
    Grp::sync();
    if (sp) {  // T row, column m = k: dNdx_m = -sum_k' e_k' A_k'm
      const int m = k;
      double dNdx = 0.0;
      for (int kk = 0; kk < N; ++kk) dNdx -= sh[L_EK + kk] * sh[L_J + kk * JS + m];
      const double tau = -ek * iD;
      sh[L_J + N * JS + m] = (dNdx - fT * cvk) * iD + dfTdT * tau;
    }
    Grp::sync();
    if (sp) {
      for (int m = 0; m < N; ++m) row[m] = row[m] + Bk * (-sh[L_EK + m] * iD);
      row[N] = 0.0;
    }
    if (l == N) sh[L_J + N * JS + N] = 0.0;
    Grp::sync();
    for (int e = l; e < n * n; e += G) {
      const int r = e / n, c = e - r * n;
      Jg[e] = sh[L_J + r * JS + c];
    }
    Grp::sync();
  }
};

}  // namespace pelec
