// Batched form of the reactor: cells of uneven stiffness integrated as a device-side work queue of
// per-cell BDF phases, one cell per lane, every kernel running one uniform code path.
//
// Work is split by phase so a wave never executes another cell's rare branch:
//   k_start   inputs -> state, f(t0,y0), initial step (cvHin), first attempt      (once per cell)
//             for a range of free slots: all of them at first, then, whenever compaction frees
//             enough slots, the next cells of the call (the queue refills, so a call may hold
//             more cells than there are slots)
//   k_jac     analytic Jacobian of the cells in the set-up list that need one       (~2 per cell)
//   k_lu_rows Newton matrix I - gamma J factorised with one row per lane, one 32- or 64-lane
//             group per matrix                                                      (~10 per cell)
//   k_newton  one Newton correction for every cell in the Newton phase: RHS, LU solve, convergence
//             test, and -- when the correction converged -- the error test, step completion,
//             order/step choice and the next step's prediction                      (~60 per cell)
//   k_compact repacks live cells into a dense slot range when enough slots are done.
// Cells that need a Newton-matrix set-up are appended (wave-aggregated) to a rotating list consumed
// by k_jac/k_lu of the next iteration. Within an iteration the set-up cells and the Newton cells
// are disjoint, so k_jac -> k_lu and k_newton run concurrently (two branches of the graph). The
// host replays a captured graph of a fixed number of iterations and stops when every started cell
// is done.
//
// Storage: per-slot structure-of-arrays for vectors and scalars (coalesced: a wave's lanes hold
// consecutive slots); per-cell array-of-structures for the n x n Jacobian and LU factors, so the
// scattered cells of a set-up batch write whole contiguous blocks and the Newton lanes stream
// their own block. A cell takes one of the C matrix blocks when it starts and keeps it while
// compaction moves its slot; a finished cell returns its block to a free ring that the next
// refill draws from.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "bdf_core.h"
#include "group_model.h"

#ifndef PELEC_MEMORY_FENCE
#define PELEC_MEMORY_FENCE() __asm__ __volatile__("" ::: "memory")
#endif

namespace pelec {

// PH_READY0 + r: Newton matrix factorised during an iteration i with i % 3 == r. The Newton kernel
// of iteration i runs concurrently with the set-up kernels of the same iteration, so it takes
// cells tagged by iteration i - 1 only. PH_LUFAIL0 + r: the matrix factorised during such an
// iteration was singular; the Newton kernel of iteration i + 1 runs the convergence-failure
// handling for it (which needs every integrator scalar, so the factorisation kernel does not).
enum Phase : int {
  PH_NEWTON = 0,
  PH_SETUP = 1,
  PH_DONE = 2,
  PH_READY0 = 3,   // + list: factorised in the set-up branch of that list's iteration
  PH_LUFAIL0 = 6   // + list: that factorisation found the Newton matrix singular
};

// Set-up lists rotate over three buffers: iteration i consumes list i%3, appends to list (i+1)%3
// and clears list (i+2)%3 (consumed during i-1, first appended to during i+1).
constexpr int NLISTS = 3;

// device counters
enum Ctr : int {
  CT_SETUP0 = 0,   // lengths of set-up lists 0..2
  CT_STARTED = 3,  // cells whose integration started
  CT_DONE = 4,     // started cells finished (ok or failed)
  CT_FAILED = 5,   // cells with status != 0 (incl. bad input)
  CT_NACTIVE = 6,  // slots in use [0, nactive)
  CT_CUR = 7,      // state buffer holding the live slots (0/1)
  CT_NEWACTIVE = 8,
  CT_FTAIL = 9,    // matrix blocks returned to the free ring (its tail; the host tracks the head)
  CT_JAC0 = 10,    // lengths of Jacobian lists 0..2: the set-up entries that need a Jacobian
  CT_INGRAPH = 13, // 1: compactions inside the graph allowed (no cell of the call waits for a slot)
  CT_GATE = 14,    // the graph's compaction check of the current iteration decided to compact
  CT_COUNT = 15
};

__host__ __device__ constexpr int list_next(int r) { return r == NLISTS - 1 ? 0 : r + 1; }
__host__ __device__ constexpr int list_prev(int r) { return r == 0 ? NLISTS - 1 : r - 1; }

struct StateBuf {
  double* zn;       // (QMAX+1) * n * C
  double* acor;     // n * C
  double* ewt;      // n * C
  uint64_t* sc;     // BDF_SCALAR_WORDS * C
  double* rhoe;     // 2 * C  (rho*e0, rho*e source)
  double* src;      // N * C  (rhoY sources) or null
  int32_t* cell;    // C  slot -> global cell index
  int32_t* phase;   // C
  int32_t* mb;      // C  slot -> the cell's matrix block (J, LU factors, pivots), kept for the
                    //    cell's lifetime while compaction moves its slot
};

template <int n>
struct SoAV : NoTape, LaneComps {
  static constexpr bool kMemory = true;
  double* zn;
  double* ac;
  double* ew;
  int64_t C;
  int64_t s;
  __device__ double& Z(int j, int i) const { return zn[((int64_t)j * n + i) * C + s]; }
  __device__ double& A(int i) const { return ac[(int64_t)i * C + s]; }
  __device__ double& E(int i) const { return ew[(int64_t)i * C + s]; }
};

// One lane's share of a group-distributed vector (the work vectors of a group accessor).
struct LaneVal {
  double x = 0.0;
  __device__ double& operator[](int) { return x; }
  __device__ const double& operator[](int) const { return x; }
};

// Group accessor: component i of slot s on lane i of a G-lane group, in the structure-of-arrays
// storage of SoAV. The integrator scalars are replicated on every lane of the group.
template <int n, int G>
struct GrpV : NoTape {
  static constexpr bool kMemory = true;
  static constexpr int kStep = G;
  double* zn;
  double* ac;
  double* ew;
  int64_t C;
  int64_t s;
  int l;  // lane within the group = the component this lane holds
  __device__ int c0() const { return l; }
  __device__ static double sum(double x) { return Group<G>::sum(x); }
  __device__ static double max(double x) { return Group<G>::max(x); }
  template <int m>
  using Arr = LaneVal;
  __device__ double& Z(int j, int i) const { return zn[((int64_t)j * n + i) * C + s]; }
  __device__ double& A(int i) const { return ac[(int64_t)i * C + s]; }
  __device__ double& E(int i) const { return ew[(int64_t)i * C + s]; }
};

__device__ inline void load_scalars(BdfScalars& S, const uint64_t* sc, int64_t C, int64_t s) {
  uint64_t w[BDF_SCALAR_WORDS];
#pragma unroll
  for (int k = 0; k < BDF_SCALAR_WORDS; ++k) w[k] = sc[k * C + s];
  __builtin_memcpy(&S, w, sizeof(S));
}
__device__ inline void store_scalars(const BdfScalars& S, uint64_t* sc, int64_t C, int64_t s) {
  uint64_t w[BDF_SCALAR_WORDS];
  __builtin_memcpy(w, &S, sizeof(S));
#pragma unroll
  for (int k = 0; k < BDF_SCALAR_WORDS; ++k) sc[k * C + s] = w[k];
}

// after_setup on the scalar record of slot s, moving only the words it reads or writes.
__device__ inline void after_setup_slot(uint64_t* sc, int64_t C, int64_t s) {
  constexpr int K[] = {(int)(offsetof(BdfScalars, gamma) / 8), (int)(offsetof(BdfScalars, gammap) / 8),
                       (int)(offsetof(BdfScalars, gamrat) / 8), (int)(offsetof(BdfScalars, crate) / 8),
                       (int)(offsetof(BdfScalars, nstlp) / 8), (int)(offsetof(BdfScalars, nst) / 8),
                       (int)(offsetof(BdfScalars, nsetups) / 8)};
  constexpr int NK = sizeof(K) / sizeof(K[0]);
  uint64_t w[BDF_SCALAR_WORDS] = {};
#pragma unroll
  for (int q = 0; q < NK; ++q) w[K[q]] = sc[K[q] * C + s];
  BdfScalars S;
  __builtin_memcpy(&S, w, sizeof(S));
  after_setup(S);
  __builtin_memcpy(w, &S, sizeof(S));
#pragma unroll
  for (int q = 0; q < NK; ++q) sc[K[q] * C + s] = w[K[q]];
}

// The integrator scalars of slot s moved between structure-of-arrays storage and a copy shared by
// a G-lane group (in LDS): lane l moves words l, l + G, ... (the group's lanes together load
// or store the record once, instead of every lane holding the whole record in registers).
template <int G>
__device__ inline void load_scalars_group(BdfScalars& S, const uint64_t* sc, int64_t C, int64_t s,
                                          int l) {
  uint64_t* w = reinterpret_cast<uint64_t*>(&S);
  for (int k = l; k < BDF_SCALAR_WORDS; k += G) w[k] = sc[k * C + s];
  Group<G>::sync();
}
template <int G>
__device__ inline void store_scalars_group(const BdfScalars& S, uint64_t* sc, int64_t C,
                                           int64_t s, int l) {
  Group<G>::sync();
  const uint64_t* w = reinterpret_cast<const uint64_t*>(&S);
  for (int k = l; k < BDF_SCALAR_WORDS; k += G) sc[k * C + s] = w[k];
}

// Wave-aggregated append of `val` for the lanes with pred set: one atomic per wave, lane order
// preserved inside the wave.
__device__ inline void wave_append(int32_t* list, unsigned long long* cnt, bool pred, int32_t val) {
  const unsigned long long mask = __ballot(pred);
  if (mask == 0ull) return;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((long long)mask) - 1;
  int base = 0;
  if (lane == leader) base = (int)atomicAdd(cnt, (unsigned long long)__popcll(mask));
  base = __shfl(base, leader, 64);
  if (pred) list[base + __popcll(mask & ((1ull << lane) - 1ull))] = val;
}

// Wave-aggregated push of `val` onto a ring of C entries whose tail counter is *tail (monotonic;
// entry i lives at ring[i % C]). The ring never holds more than C entries: each entry is a matrix
// block, and there are C blocks.
__device__ inline void wave_push_ring(int32_t* ring, unsigned long long* tail, int64_t C, bool pred,
                                      int32_t val) {
  const unsigned long long mask = __ballot(pred);
  if (mask == 0ull) return;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((long long)mask) - 1;
  unsigned long long base = 0;
  if (lane == leader) base = atomicAdd(tail, (unsigned long long)__popcll(mask));
  base = __shfl(base, leader, 64);
  if (pred) ring[(base + (unsigned long long)__popcll(mask & ((1ull << lane) - 1ull))) % C] = val;
}

__device__ inline void wave_count(unsigned long long* cnt, bool pred) {
  const unsigned long long mask = __ballot(pred);
  if (mask == 0ull) return;
  const int lane = threadIdx.x & 63;
  if (lane == __ffsll((long long)mask) - 1) atomicAdd(cnt, (unsigned long long)__popcll(mask));
}

// Dense LU with partial pivoting of one n x n matrix held by a group of G lanes (G a power of two,
// n <= G, groups aligned within the wave), lane r of the group holding row r in registers. Rows
// are not moved: each lane tracks the position its row would occupy in the row-swapping algorithm
// (pos), so the pivot of step k is the max |a_ik| over positions >= k with ties to the lower
// position -- the same choice, and the same factors, as the row-swapping algorithm (lu_factor in
// bdf_cell.h, and the oracle's). The pivot search is a DPP butterfly and the pivot row is read
// by v_readlane (no LDS round trips). On return lane r holds row pos of the factors (unit-lower multipliers left
// of the diagonal, U right of it, the diagonal inverted) and perm[pos] = r. Returns 0, or k + 1
// when column k has no nonzero pivot (every lane of the group returns the same value).
template <int n, typename T, int G>
__device__ inline int lu_rows(T (&row)[n], int r, int& pos) {
  static_assert(n <= G && (G & (G - 1)) == 0 && G >= 16 && G <= 64, "one row per lane");
  const int lane = threadIdx.x & 63;
  pos = r;
  int ier = 0;
#pragma unroll
  for (int k = 0; k < n; ++k) {
    // pivot search over positions >= k: key = (|a_rk|, -pos)
    T v = (r < n && pos >= k) ? T(fabs(row[k])) : T(-1);
    int kp = pos;
    int src = lane;
    // butterfly with decreasing offsets: the key order is total (positions are distinct), so
    // every lane of the group ends with the same pivot
    auto step = [&](auto off) {
      constexpr int o = decltype(off)::value;
      if constexpr (o < G) {
        const T ov = xor_partner<G, o>(v);
        const int op = xor_partner<G, o>(kp);
        const int os = xor_partner<G, o>(src);
        if (ov > v || (ov == v && op < kp)) {
          v = ov;
          kp = op;
          src = os;
        }
      }
    };
    step(std::integral_constant<int, 32>());
    step(std::integral_constant<int, 16>());
    step(std::integral_constant<int, 8>());
    step(std::integral_constant<int, 4>());
    step(std::integral_constant<int, 2>());
    step(std::integral_constant<int, 1>());
    if (!(v > T(0))) {
      if (ier == 0) ier = k + 1;
      continue;  // singular: the factors are not used
    }
    // the row at position k moves to the pivot row's old position
    const int pivot_lane = src;
    const bool is_pivot = lane == pivot_lane;
    const int old_pos_pivot = kp;
    if (r < n && pos == k && !is_pivot) pos = old_pos_pivot;
    if (is_pivot) pos = k;
    // broadcast the pivot row's columns k..n-1: the pivot lane of each group of the wave, read
    // into scalar registers (it is uniform within a group), then one v_readlane per column and
    // group
    int pl[64 / G];
#pragma unroll
    for (int g = 0; g < 64 / G; ++g) pl[g] = lane_read(pivot_lane, g * G);
    T prow[n];
#pragma unroll
    for (int c = k; c < n; ++c) {
      T pv = lane_read(row[c], pl[0]);
#pragma unroll
      for (int g = 1; g < 64 / G; ++g) {
        const T o = lane_read(row[c], pl[g]);
        pv = lane >= g * G ? o : pv;
      }
      prow[c] = pv;
    }
    const T inv = T(1) / prow[k];
    if (is_pivot) row[k] = inv;
    if (r < n && pos > k) {
      const T l = row[k] * inv;
      row[k] = l;
#pragma unroll
      for (int c = k + 1; c < n; ++c) row[c] -= l * prow[c];
    }
  }
  return ier;
}

// x = (LU)^{-1} P b by a G-lane group: lane l holds b_l and row l of the factors in pivot order
// (as k_lu_rows stores them: unit-lower multipliers, U, inverted diagonal) and perm_l = perm[l].
// Forward and back substitution broadcast one solved component per step. Returns x_l (0 on lanes
// >= n).
template <int n, typename T, int G>
__device__ inline double lu_solve_group(const T (&row)[n], int perm_l, double b, int l) {
  T x = T(__shfl(b, perm_l, G));
#pragma unroll
  for (int j = 0; j < n - 1; ++j) {
    const T xj = Group<G>::bcast(x, j);
    if (l > j && l < n) x -= row[j] * xj;
  }
#pragma unroll
  for (int j = n - 1; j >= 0; --j) {
    if (l == j) x *= row[j];
    const T xj = Group<G>::bcast(x, j);
    if (l < j) x -= row[j] * xj;
  }
  return l < n ? double(x) : 0.0;
}

// Equilibration of the FP32 Newton matrix: component i is scaled by d_i = 2^k_i, the power of
// two at or below its error weight ewt_i (k_i clamped to -60..60), and the factors are those of
// D M D^-1. Powers of two make the scaling exact (one ldexp per element), so only the FP32 cast
// of D M D^-1 rounds. In the error-weighted variables the Newton convergence test measures, the
// scaled matrix is well conditioned where M itself (rho Y_k ~ 1e-4 beside T ~ 1e3) is not.
__device__ inline int scale_exp(double w) {
  const int k = ilogb(w);
  return k < -60 ? -60 : (k > 60 ? 60 : k);
}

// x = (LU)^{-1} P b with factors from lu_rows (AoS block `lu`, inverted diagonal). b in/out (FP64),
// arithmetic in T. `gb` is this lane's LDS scratch (stride 64) used to gather b by the permutation.
// With `dexp` (the equilibration exponents k_i of an FP32 matrix) the factors are those of
// D M D^-1: b is scaled by D on the way in and x by D^-1 on the way out.
template <int n, typename T>
__device__ inline void lu_solve_perm(const T* lu, const int32_t* perm, double* b, double* gb,
                                     const int32_t* dexp = nullptr) {
  if (dexp) {
#pragma unroll
    for (int i = 0; i < n; ++i) gb[i * 64] = ldexp(b[i], dexp[i]);
  } else {
#pragma unroll
    for (int i = 0; i < n; ++i) gb[i * 64] = b[i];
  }
  T x[n];
#pragma unroll
  for (int i = 0; i < n; ++i) x[i] = T(gb[perm[i] * 64]);
#pragma unroll
  for (int i = 1; i < n; ++i) {
    T s = x[i];
#pragma unroll
    for (int j = 0; j < n; ++j)
      if (j < i) s -= lu[i * n + j] * x[j];
    x[i] = s;
    PELEC_MEMORY_FENCE();  // bound the loads in flight to one row
  }
#pragma unroll
  for (int i = n - 1; i >= 0; --i) {
    T s = x[i];
#pragma unroll
    for (int j = 0; j < n; ++j)
      if (j > i) s -= lu[i * n + j] * x[j];
    x[i] = s * lu[i * n + i];
    PELEC_MEMORY_FENCE();
  }
  if (dexp) {
#pragma unroll
    for (int i = 0; i < n; ++i) b[i] = ldexp(double(x[i]), -dexp[i]);
  } else {
#pragma unroll
    for (int i = 0; i < n; ++i) b[i] = double(x[i]);
  }
}

}  // namespace pelec
