/* pelec_amd.h -- C-ABI of the MI355X (gfx950) cell-wise stiff-chemistry reactor.
 *
 * One shared library is built per mechanism (libpelec_amd_<mechanism>.so), the way the reference
 * compiles one PelePhysics mechanism per build (Chemistry_Model, Exec/Make.PeleC; mechanism sources
 * at CMake/BuildPelePhysicsLib.cmake:113-116). Every library exports the same symbols.
 *
 * The reference interface replaced here is pele::physics::reactions::ReactorBase as PeleC uses it:
 *   create/init/close         Source/PeleC.cpp:2032-2047 (owner Source/PeleC.H:700-702)
 *   set_typ_vals_ode          Source/React.cpp:7-27
 *   react(box, ...)           Source/React.cpp:222-228
 * (the ReactorBase header itself lives in the absent PelePhysics submodule).
 *
 * Conventions
 *  - All array pointers are device pointers (HBM) unless stated. Work is enqueued on the stream
 *    passed in, after the caller's earlier work on it.
 *  - Which calls block: prc_react_box, prc_react_flat, prc_react_boxes and prc_react_state_box
 *    return only once the reactor's kernels on the stream have finished (react_state's unpack
 *    kernel excepted, which is left enqueued). The plain form synchronises because its argument
 *    block is reused by the next call; the batched form reads its work-queue counters back after
 *    every replay of its captured graph to decide whether another is needed. PeleC synchronises
 *    the stream right after react anyway (React.cpp:230), so the drop-in keeps the reference's
 *    observable ordering. prc_autotune and prc_last_failures also block. prc_adjoint,
 *    prc_compute_temp, the EOS helpers and prc_model_eval enqueue and return.
 *  - A reactor handle owns per-call device workspaces (batched state, react_state packing,
 *    adjoint checkpoints): calls on one handle must be issued on one stream, or ordered by the
 *    caller, or they race on those workspaces. Separate handles are independent.
 *  - Units are CGS (g, cm, s, K, erg), as PeleC and PelePhysics use them.
 *  - Return value: PRC_OK (0) or a negative PRC_E* error code; per-cell integration failures are
 *    reported through the optional per-cell diagnostics buffer, never by aborting the batch.
 *  - A cell with mask == 0 is not integrated: its rY, T, rE are left bit-identical and FC = 0.
 */
#ifndef PELEC_AMD_H
#define PELEC_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PRC_OK 0
#define PRC_EINVAL (-1)     /* bad argument (null handle, unknown name/key, bad shape) */
#define PRC_EHIP (-2)       /* a HIP runtime call failed */
#define PRC_ESTATE (-3)     /* call not valid in the reactor's current state */
#define PRC_ENOMEM (-4)     /* device allocation failed */

/* Per-cell diagnostics record (int32, component-major, PRC_DIAG_NCOMP components):
 * status, steps, RHS evaluations, Jacobian evaluations, LU factorisations, Newton iterations,
 * error-test failures, convergence failures. status: 0 ok, 1 max steps, 2 repeated error-test
 * failures, 3 repeated convergence failures, 4 bad input (rho <= 0, T <= 0 or dt <= 0), 5 adjoint
 * checkpoint or tape capacity exceeded (prc_adjoint only). */
#define PRC_DIAG_NCOMP 8

/* Array4-shaped view of caller-owned device memory, mirroring amrex::Array4<Real>:
 * element (i,j,k,n) is p[(i-begin[0]) + (j-begin[1])*jstride + (k-begin[2])*kstride + n*nstride];
 * end[] is exclusive. */
typedef struct prc_array4 {
  double* p;
  int64_t jstride, kstride, nstride;
  int32_t begin[3];
  int32_t end[3];
  int32_t ncomp;
} prc_array4;

/* amrex::Array4<int> counterpart (the mask argument). */
typedef struct prc_iarray4 {
  int32_t* p;
  int64_t jstride, kstride, nstride;
  int32_t begin[3];
  int32_t end[3];
  int32_t ncomp;
} prc_iarray4;

typedef struct prc_reactor prc_reactor;

/* ---- mechanism (compile-time, one per library) ------------------------------------------ */
/* NUM_SPECIES of the mechanism (mechanism.H NUM_SPECIES). */
int prc_num_species(void);
/* Number of reactions (mechanism.H NUM_REACTIONS). */
int prc_num_reactions(void);
/* Name of species k in mechanism order (CKSYMS / the order pinned by the PMF files). */
const char* prc_species_name(int k);
/* Mechanism name and SHA-256 of the mechanism data file the library was generated from. */
const char* prc_mechanism_name(void);
const char* prc_mechanism_sha256(void);

/* ---- reactor life cycle (ReactorBase) ---------------------------------------------------- */
/* ReactorBase::create(chem_integrator), Source/PeleC.cpp:2034. Names: "ReactorCvode" (BDF),
 * "ReactorRK64" (explicit RK with error control) and "ReactorNull" (no-op). */
int prc_create(const char* chem_integrator, prc_reactor** out);
/* ReactorBase::init(reactor_type, ncells), Source/PeleC.cpp:2040. Only reactor_type 1
 * (constant-volume rho*e) exists in PeleC; ncells is a sizing hint. */
int prc_init(prc_reactor* r, int reactor_type, int ncells);
/* ReactorBase::set_typ_vals_ode(typ), Source/React.cpp:11,25: n = NUM_SPECIES + 1 typical values
 * (rhoY_k ..., T); absolute tolerances become atol * typ_i. Pass n = 0 to clear. */
int prc_set_typ_vals_ode(prc_reactor* r, const double* typ, int n);
/* ParmParse knobs the reactors read (ode.rtol, ode.atol, ode.max_steps, cvode.max_order; for
 * ReactorRK64 rk64.nsubsteps_guess / rk64.nsubsteps_min / rk64.nsubsteps_max, default 10 / 5 /
 * 100000: first step dt/guess, steps within [dt/max, dt/min]; the pelec.adaptrk_nsubsteps_* knobs
 * of Exec/Production/ChallengeProblem/challenge.inp:29-31) and the gfx950 execution form:
 * "prc.form" = 0 plain FP64 one-cell-per-lane drop-in, 1 batched persistent work-queue kernel
 * (ReactorCvode only); "prc.lu_fp32" = 1 runs the Newton-matrix LU in FP32, each solve followed
 * by "prc.lu_refine" (0, 1 or 2, default 1) FP64 refinement steps against the FP64 Jacobian. In
 * the batched form the FP32 factors are those of D M D^-1, D the powers of two of the cell's error
 * weights (exact scaling); the plain form factors M unscaled and always refines at least once. */
int prc_set_param_real(prc_reactor* r, const char* key, double value);
/* (launch shape of the batched form: prc.iters_per_graph (multiple of 3), prc.cooperative (bit 0:
 * the Newton kernel, bit 1: the Jacobian kernel run one cell per 32-lane group instead of one
 * cell per lane), prc.blocks_newton, prc.blocks_jac, prc.blocks_lu (blocks per CU, 0 =
 * occupancy-derived), prc.compact_pct (compact the work queue when fewer than this percentage
 * of its slots are alive), prc.compact_every (the graph itself checks for a compaction after
 * every this many iterations while no cell waits for a slot; 0 = only between replays),
 * prc.coop_tail_slots (once at most this many slots are in
 * use, the Newton kernel runs one cell per 32-lane group; 0 = never) and prc.blocks_tail (its
 * blocks per CU), prc.max_slots (slots of the batched form's work queue: 0 = one per cell while
 * half the free device memory holds them; a call with more cells than slots refills the slots
 * that finished cells free, from the call's cell queue); for tests, prc.test_lu_fail = 1 makes
 * the batched form's factorisation report every Newton matrix as singular, which drives CVODE's
 * convergence-failure handling) */
int prc_set_param_int(prc_reactor* r, const char* key, int value);
/* Read back an integer knob (the prc.* launch-shape knobs report the values in use; "prc.slots"
 * gives the slot count of the batched workspace now allocated). */
int prc_get_param_int(prc_reactor* r, const char* key, int* value);
/* Launch-shape autotuner of the batched form (prc.form 1, ReactorCvode): times react calls on the
 * given sample of cells (device arrays, component-major; copied, not modified) over candidate
 * shapes -- the kernel forms (prc.cooperative, prc.coop_tail_slots), iterations per captured
 * graph (prc.iters_per_graph), blocks per CU of the Newton,
 * Jacobian and LU kernels (prc.blocks_newton / _jac / _lu), the compaction threshold
 * and interval (prc.compact_pct, prc.compact_every) -- by coordinate descent from the
 * occupancy-derived defaults, keeps the
 * fastest in the reactor and returns its time (ms) in *best_ms. Synchronises the stream. No
 * counterpart in the reference (its GPU launch shapes are fixed at compile time). */
int prc_autotune(prc_reactor* r, int64_t ncells, const double* rY, const double* T,
                 const double* rE, double dt, double* best_ms, void* stream);
/* Optional per-cell diagnostics (device int32 buffer, PRC_DIAG_NCOMP x capacity, component
 * stride `capacity`); cell c of the next react call writes diag[comp*capacity + c]. Pass NULL to
 * disable. */
int prc_set_diagnostics(prc_reactor* r, int32_t* diag, int64_t capacity);
/* ReactorBase::close(), Source/PeleC.cpp:2046; releases device workspace. */
int prc_close(prc_reactor* r);
int prc_destroy(prc_reactor* r);

/* ---- reaction step ---------------------------------------------------------------------- */
/* ReactorBase::react(box, rY, rYsrc_ext, T, rEner, rEner_src_ext, FC, mask, dt_react, time,
 * stream), Source/React.cpp:222-228. Integrates every unmasked cell of the box [lo, hi]
 * (inclusive) over dt_react:
 *   d(rhoY_k)/dt = W_k wdot_k + rYsrc_k,  rho*e(t) = rEner + t * rEner_src.
 * rY (NUM_SPECIES comps), T and rEner are updated in place; FC receives the RHS evaluation count.
 * rYsrc_ext.p / rEner_src_ext.p / mask.p may be NULL (zero sources, all cells active). On return
 * *time has advanced by *dt_react. */
int prc_react_box(prc_reactor* r, const int32_t lo[3], const int32_t hi[3], prc_array4 rY,
                  prc_array4 rYsrc_ext, prc_array4 T, prc_array4 rEner,
                  prc_array4 rEner_src_ext, prc_array4 FC, prc_iarray4 mask, double* dt_react,
                  double* time, void* stream);

/* Flat (1-D) variant over ncells cells stored component-major: rY[k*ncells + c],
 * rYsrc[k*ncells + c]; T, rE, rEsrc, FC, mask have one entry per cell. NULL rYsrc/rEsrc/mask/FC
 * allowed. */
int prc_react_flat(prc_reactor* r, double* rY, const double* rYsrc, double* T, double* rE,
                   const double* rEsrc, double* FC, const int32_t* mask, int64_t ncells,
                   double* dt_react, double* time, void* stream);

/* Batched multi-box call: all boxes are integrated by one launch (the batched form walks one
 * global cell queue across boxes). Equivalent to calling prc_react_box per box. Per-cell
 * diagnostics are indexed by the cell's position in the concatenation of the boxes. */
typedef struct prc_box_args {
  int32_t lo[3], hi[3];
  prc_array4 rY, rYsrc_ext, T, rEner, rEner_src_ext, FC;
  prc_iarray4 mask;
} prc_box_args;
int prc_react_boxes(prc_reactor* r, int nboxes, const prc_box_args* boxes, double* dt_react,
                    double* time, void* stream);

/* No-allocation variant of prc_react_flat. No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Same arguments and results as prc_react_flat with the
 * batched form (ReactorCvode; ReactorNull only advances *time; ReactorRK64 is rejected with
 * PRC_EINVAL), whatever prc.form says. It calls no hipMalloc/hipFree: the argument block, the
 * work-queue counters and the batched workspace live in a static device array of
 * PRC_NOALLOC_SCRATCH_BYTES per GPU (held by the library until the process ends, located at
 * prc_init), and the work queue gets as many slots as that array holds (drm19 about 13,000,
 * GRI-Mech 3.0 about 3,000), refilled from the call's cells, so any ncells up to INT32_MAX is
 * accepted. PRC_ENOMEM if not even 64 slots fit; PRC_ESTATE if the handle was initialised on
 * another device. Iterations are enqueued directly (no graph is instantiated). Every call on a
 * GPU shares that GPU's array: calls that overlap on one GPU -- from any handle of this
 * mechanism's library -- race, and must be serialised by the caller. Blocks like prc_react_flat;
 * prc_last_failures reports the call's failures. */
#define PRC_NOALLOC_SCRATCH_BYTES ((size_t)128 << 20)
int prc_react_flat_noalloc(prc_reactor* r, double* rY, const double* rYsrc, double* T, double* rE,
                           const double* rEsrc, double* FC, const int32_t* mask, int64_t ncells,
                           double* dt_react, double* time, void* stream);

/* ---- PeleC-side passes around the reactor ----------------------------------------------- */
/* PeleC::react_state for one box (Source/React.cpp:29-362, the per-MFIter body), fused: one pack
 * kernel (STemp = [rhoY, T, rho e] of S_old -- of S_new when react_init -- and the energy forcing
 * rhoedot_ext, React.cpp:111-121,181-220), the reactor on the packed box with the species forcing
 * read in place from nonreact_src (React.cpp:121,222-228), one unpack kernel (new rho, momenta,
 * rhoY, T, rho e, rho E unless react_init; I_R = [rho wdot_k, rho edot, heat release],
 * React.cpp:233-314,325-343). Conservative layout of Source/IndexDefines.H: URHO 0, UMX..UMZ 1..3,
 * UEDEN 4, UEINT 5, UTEMP 6, species from ufs = 7 + NUM_ADV. S_old, S_new, nonreact_src: >= ufs +
 * NUM_SPECIES comps; I_R: NUM_SPECIES + 2 comps. Optional (p may be NULL): mask (0 = covered cell,
 * left untouched with I_R = 0, the per-cell form of React.cpp:157-164), rot_energy (0.5 omega^2
 * r^2 per cell for rotating frames, React.cpp:189-192) and cost (per-cell RHS evaluations of the
 * reactor: the device-side work estimate that replaces the per-box wall time of
 * React.cpp:168-169,316-322). On return *time has advanced by *dt_react. */
int prc_react_state_box(prc_reactor* r, const int32_t lo[3], const int32_t hi[3],
                        prc_array4 S_old, prc_array4 S_new, prc_array4 nonreact_src,
                        prc_array4 I_R, prc_iarray4 mask, prc_array4 rot_energy, prc_array4 cost,
                        int ufs, int react_init, double* dt_react, double* time, void* stream);

/* PeleC::computeTemp's per-cell pc_cmpTemp (Source/PeleC.cpp:2254-2272, Source/Utilities.H:68-85):
 * S(UTEMP) = REY2T(rho, S(UEINT)/rho, Y) over the box, the stored T as the Newton guess; cells
 * with mask == 0 (covered) are skipped. The internal-energy floor that computeTemp applies first
 * (reset_internal_energy) is a hydrodynamics fix-up and stays with the caller. */
int prc_compute_temp(const int32_t lo[3], const int32_t hi[3], prc_array4 S, prc_iarray4 mask,
                     int ufs, void* stream);

/* ---- discrete adjoint (no counterpart in the reference; SUNDIALS CVODES is not built upstream,
 * CMake/SetSundialsOptions.cmake:5) --------------------------------------------------------- */
/* Integrates ncells cells (flat, component-major: rY0[k*ncells + c]) over dt with the BDF reactor
 * (the ReactorCvode algorithm, plain form) at rate multipliers exp(logA[j]) on each reaction's
 * pre-exponential factor(s) (logA: NUM_REACTIONS device values, NULL = nominal), writes the end
 * state (rY_out, T_out) and the vector-Jacobian product of the end state with the cotangents
 * (ct_rY: NUM_SPECIES x ncells, ct_T: ncells; NULL = 0): g_rY0 (d/d rhoY0, NUM_SPECIES x ncells),
 * g_rE0 (d/d rho*e0, ncells), g_logA (d/d logA_j per cell, NUM_REACTIONS x ncells; the caller
 * sums over cells). The input T0 is only the Newton guess of T(e, Y): its gradient is zero. The
 * adjoint is the exact transpose of the discrete integrator with its step-control decisions held
 * fixed and every corrector equation solved to machine precision (the forward pass polishes each
 * converged Newton correction with the exact Jacobian, so its end state can differ from
 * prc_react's by the Newton tolerance), evaluated from checkpoints every "adj.checkpoint_steps" (8) accepted steps with
 * recomputation ("adj.max_checkpoints" 128 per cell, "adj.segment_records" 256 tape records per
 * segment; status 5 when exceeded). status: per-cell CellStatus. tape (optional, may be NULL):
 * the complete forward tape of each cell (tape_cap records of 16 + NUM_SPECIES + 1 doubles per
 * cell; tape_count[c] = records written, -1 if tape_cap was too small) for verification.
 * Mechanisms with NUM_SPECIES + 1 > 32 (the 53-species ones) have no adjoint kernel: its
 * per-lane private state would exceed what the runtime backs; the call returns PRC_EINVAL. */
int prc_adjoint(prc_reactor* r, int64_t ncells, const double* rY0, const double* T0,
                const double* rE0, const double* rYsrc, const double* rEsrc, double dt,
                const double* logA, const double* ct_rY, const double* ct_T, double* rY_out,
                double* T_out, double* g_rY0, double* g_rE0, double* g_logA, int32_t* status,
                double* tape, int64_t tape_cap, int32_t* tape_count, void* stream);

/* Synchronise the stream and return the number of cells whose integration failed in the
 * reactor's previous react call on it (status != 0). */
int64_t prc_last_failures(prc_reactor* r, void* stream);

/* Per-kernel device time of the previous batched react call, when the reactor was configured
 * with prc_set_param_int(r, "prc.kernel_timing", 1): ms[0..2] = total milliseconds of the
 * k_newton, k_jac and k_lu launches (HIP events on the stream each kernel ran on), *launches =
 * launches of each. Instrumentation with no counterpart in the reference (which only times whole
 * fabs, Source/React.cpp:168-169). */
int prc_kernel_times(prc_reactor* r, double* ms, int64_t* launches);

/* ---- EOS helpers at the boundary (Fuego EOS calls PeleC makes around the reactor) -------- */
/* PYT2RE: rho and e from p, T, Y (Exec/RegTests/zeroD/prob.cpp:34-39). Y component-major. */
int prc_eos_pyt2re(int64_t ncells, const double* p, const double* T, const double* Y,
                   double* rho, double* e, void* stream);
/* REY2T: T from rho, e, Y by Newton, T in/out as initial guess (pc_cmpTemp, Source/Utilities.H:
 * 68-85). */
int prc_eos_rey2t(int64_t ncells, const double* rho, const double* e, const double* Y, double* T,
                  void* stream);
/* RTY2Hi: species enthalpies h_k(T) [erg/g] (heat release, Source/React.cpp:337-338); hi
 * component-major. */
int prc_eos_rty2hi(int64_t ncells, const double* T, double* hi, void* stream);
/* Production rates (mechanism productionRate / RTY2WDOT): wdot_k [mol/(cm^3 s)] from rho, T, Y. */
int prc_eos_rty2wdot(int64_t ncells, const double* rho, const double* T, const double* Y,
                     double* wdot, void* stream);

/* ---- reactor model ------------------------------------------------------------------------ */
/* The constant-volume rho*e reactor's RHS f(t, y) and exact Jacobian df/dy at given states: the
 * CVODE user functions of ReactorCvode (cF_RHS and the analytic Jacobian of ReactorCvodeUtils /
 * ReactorCvodeJacobian, CMake/BuildPelePhysicsLib.cmake:132-141). y = [rhoY_0..N-1, T] and ydot
 * component-major (n x ncells, n = N+1); rhoe0 per cell; rhoesrc (rho e forcing) and rYsrc
 * (N x ncells) may be null; J (ncells x n x n, row-major per cell) may be null. cooperative = 0
 * evaluates one cell per lane, 1 one cell per 32-lane group (the kinetics of the cooperative
 * batched kernels). Device pointers. */
int prc_model_eval(int64_t ncells, double t, const double* y, const double* rhoe0,
                   const double* rhoesrc, const double* rYsrc, double* ydot, double* J,
                   int cooperative, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* PELEC_AMD_H */
