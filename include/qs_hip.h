/*
 * qs_hip.h -- C ABI of the MI355X (gfx950) implementation of QUICHOTTE's
 * self-consistent NEGF path (orders 0, 1 and 2; order 2 is the hot path).
 *
 * Plain pointers and sizes only.  Host arrays use the reference's conventions:
 * complex numbers are interleaved (re, im) doubles, matrices are column-major
 * (Fortran order), orbital indices are 0-based here where the reference's are
 * 1-based, packed vectors use pack_gf's layout (common/linear_algebra.f90:142-191).
 *
 * Errors: every entry point returns QS_OK (0) or a nonzero status;
 * qs_last_error() gives the message.  QS_ERR_SINGULAR is the reference's
 * `stop 1` on a failed factorisation (common/linear_algebra.f90:73-76,84-87)
 * and is fatal for the run.  A runaway SCF is not an error: it is reported in
 * the result (src/scf.f90:180-186).
 *
 * Threading: one handle per request.  Handles share no mutable state, so
 * different handles may be used from different host threads.  A handle itself
 * must not be used by two threads at once.
 */
#ifndef QS_HIP_H
#define QS_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QS_OK 0
#define QS_ERR_SINGULAR 1
#define QS_ERR_ARG 2
#define QS_ERR_HIP 3
#define QS_ERR_NOMEM 4
#define QS_ERR_UNSUPPORTED 5
#define QS_ERR_STATE 6

/* Everything solve_scf reads from module state (src/green_functions.f90:38-67,
 * src/parameters.f90:7-90, src/scf.f90:38-43).  Copied at qs_create. */
typedef struct qs_problem {
    int32_t norb;             /* spin-orbitals = 2 N_sites; basis 2i = (site i, up), 2i+1 = (site i, down) */
    int32_t n_omega;          /* energy points: omega_i = w_init + i*delta, i = 1..n_omega (src/quichotte.f90:85-93) */
    double w_init, delta;
    const double *H;          /* (norb,norb) complex, column-major */
    const double *gammaL;     /* (norb,norb) complex: left lead coupling (src/input_output.f90:310-346) */
    const double *gammaR;     /* (norb,norb) complex */
    const double *U_site;     /* (norb): Hubbard U on the spin-up index of each site (src/input_output.f90:432-462) */
    int32_t order;            /* 0, 1 or 2 */
    double beta, mu;          /* 1/(kB T), chemical potential */
    int32_t n_atoms;          /* atoms; 0 => one orbital per atom */
    const int32_t *atom_norb; /* (n_atoms) spatial orbitals of each atom, or NULL */
    double alpha_mix;         /* Broyden settings (src/scf.f90:38-43) */
    int32_t nhist, maxscf;
    double tolf, tolf_order1;
    int32_t single_history;   /* broyden_single_precision */
    int32_t device;           /* HIP device ordinal */
} qs_problem;

typedef struct qs_scf_result {
    int32_t converged;        /* Broyden converged and the final sweep is physical (src/scf.f90:142-150) */
    int32_t n_iter;           /* residual evaluations */
    double fmax;              /* max|F| of the last residual */
    int32_t runaway;          /* the SCF was aborted (src/scf.f90:180-186) */
    int32_t unphysical;       /* final_sanity (src/scf.f90:260-275) */
    int32_t i_worst;          /* worst state (1-based), its occupation and spectral weight */
    double n_worst, w_worst;
} qs_scf_result;

typedef struct qs_handle qs_handle;

/* setup_storage + scf_setup_sizes + sigma2_init (src/green_functions.f90:90-190,
 * src/scf.f90:61-68, src/self_energies.f90:91-106): active subspace, device
 * storage and FFT plan.  *out is NULL on failure. */
int qs_create(const qs_problem *p, qs_handle **out);
/* sigma2_finalize (src/self_energies.f90:108-113) and every device allocation */
void qs_destroy(qs_handle *h);
const char *qs_last_error(const qs_handle *h);

/* Point an existing handle at another system of the same sizes (norb, n_omega,
 * order, nhist, history precision, device) and the same active subspace, lead
 * contacts and atoms: H, Gamma_L/R, U, T, mu and the grid origin/step change, no
 * device memory is allocated and the Dyson launch shape is kept.  The next solve
 * must be fresh.  QS_ERR_STATE if the problem does not fit the handle.  (No
 * counterpart in the reference; a request pool reuses handles with it.) */
int qs_reset_problem(qs_handle *h, const qs_problem *p);

/* sizes of the active subspace and of the packed vector (n_vec, src/scf.f90:61-68);
 * nf is the padded FFT length used on the device (a power of two >= 2 n_omega - 1) */
int qs_sizes(const qs_handle *h, int32_t *nA, int32_t *nS, int32_t *nCL, int32_t *nAs, int64_t *n_vec,
             int32_t *nf);
/* bytes of device memory held by the handle */
int64_t qs_device_bytes(const qs_handle *h);
/* device bytes qs_create would allocate for p (admission control of a request
 * queue; no counterpart in the reference, whose memory estimate is printed by
 * setup_storage, src/green_functions.f90:178-184); -1 if p is invalid */
int64_t qs_footprint_bytes(const qs_problem *p);
/* free and total device memory of a device (hipMemGetInfo) */
int qs_device_memory(int32_t device, int64_t *free_bytes, int64_t *total_bytes);

/* solve_scf (src/scf.f90:78-151): converge at voltage V (fresh start from the
 * non-interacting solution, or continue from the last input), then the final
 * sweep.  log_fmax / log_rms (length >= maxscf, may be NULL) receive max|F|
 * and rms|F| per step, the lines of scf_V_<V>.dat (common/broyden_mixing.f90:92-95). */
int qs_solve_scf(qs_handle *h, double V, int32_t fresh, qs_scf_result *res, double *log_fmax, double *log_rms);

/* observables of the last final sweep */
int qs_current(const qs_handle *h, double *J_up, double *J_down);    /* Current, src/observables.f90:22-39 */
int qs_g_less_obs(const qs_handle *h, double *G);                    /* G_less_obs (norb,norb) complex */
int qs_dosd(const qs_handle *h, double *dosd);                       /* dosd (norb, n_omega) real */
int qs_trUD(const qs_handle *h, double *trU, double *trD);           /* current integrand (n_omega) each */

/* Precision of the order-2 SCF sweeps (no counterpart in the reference, which is
 * FP64 throughout; replaces the FP64 zgetrf/zgetri and matmul of
 * common/linear_algebra.f90:58-88 and src/scf.f90:318-344 in the early steps).
 * mode 0: FP64 (the default, the drop-in path).  mode 1, mixed: the Dyson inverse
 * and the two Keldysh products run in complex FP32 on MFMA with FP32 accumulation
 * while max|F| >= switch_fmax (and keeps falling), then FP64 steps take the
 * solution to tolf / 10 (a decade below the FP64 path's stopping point, so the
 * currents agree with it to ~tolf); the Sigma2 FFT, the integrals, Broyden and the final sweep
 * are FP64 throughout.  Mode 1 allocates the large-Norb chunk buffers if the
 * size has none.  qs_precision_steps gives the last solve's residual counts in
 * each precision. */
int qs_set_precision(qs_handle *h, int32_t mode, double switch_fmax);
int qs_precision_steps(const qs_handle *h, int32_t *fp32_steps, int32_t *fp64_steps);

/* How order-2 SCF steps run (no counterpart in the reference, whose loop is
 * host code; the steps are src/scf.f90:158-188 and common/broyden_mixing.f90:51-184).
 * mode 0: the host makes the step's scalar decisions (sanity, Broyden slot, the
 * Gram solve) between launches; 1 (default): a fixed launch sequence with those
 * decisions in device kernels, one small copy to the host per step; 2: as 1, with
 * steps from the third on replayed from two captured graphs (same results as 1,
 * bit for bit).  Mode 2 is for a process whose other threads make no HIP calls on
 * the device while a solve runs: a capture is disturbed by such calls (the solve
 * then goes on as mode 1), and they can fail.  Modes 1 and 2 apply to FP64 sweeps
 * with FP32 history of at most 16 columns (the fused update); other
 * configurations use mode 0.  qs_step_info: whether the device path applies, and
 * how many steps of the last solve were graph replays. */
int qs_set_step_mode(qs_handle *h, int32_t mode);
int qs_step_info(const qs_handle *h, int32_t *device_steps_active, int32_t *graph_steps);

/* --- groups: several solves stepped together ---------------------------------
 * No counterpart in the reference, whose voltage loop solves one bias point at a
 * time (src/quichotte.f90:125-143).  For callers with many independent bias
 * points (a sweep split over several handles, a serving loop): a group runs one
 * SCF step of every member per qs_group_step.  Each member's step is its own
 * solve_scf step (src/scf.f90:78-151; residual src/scf.f90:158-188,
 * broyden_solve common/broyden_mixing.f90:51-184) with device-controlled scalars
 * (step mode 1) on its handle's stream, except that the Dyson/Keldysh sweeps of
 * all members that share a kernel instance (same register tile and launch shape:
 * orbital counts may differ within a tile class, grid lengths may differ) run as
 * one launch on the group's stream, between the members' stages.  A member's
 * results equal those of qs_solve_scf on its handle bit for bit.  Solves join
 * (qs_group_add) and leave (when they end) between steps, so a loop can keep the
 * group full as requests arrive.  One thread drives a group; while a handle is a
 * member, no other call may use it (qs_destroy removes it first).  At most 64
 * members. */
typedef struct qs_group qs_group;
int qs_group_create(int32_t device, qs_group **out);
void qs_group_destroy(qs_group *g);
const char *qs_group_last_error(const qs_group *g);
/* start a solve at voltage V on h and make it a member: fresh = 1 starts from
 * the non-interacting solution, 0 continues from the handle's last input, as
 * qs_solve_scf.  Needs a handle on which qs_solve_scf takes device-controlled
 * steps (qs_step_info): order 2, FP64 sweeps, FP32 history of at most 16 columns,
 * and one energy's packed slice within the fused update's LDS (at most 82 active
 * spin-orbitals); QS_ERR_UNSUPPORTED otherwise. */
int qs_group_add(qs_group *g, qs_handle *h, double V, int32_t fresh);
int qs_group_members(const qs_group *g);
/* one SCF step of every member.  Members whose solve ended (converged, aborted
 * as a runaway, maxscf reached, or an error) get their final sweep and leave:
 * done[k], res[k] (as qs_solve_scf's) and status[k] (QS_OK, or the error; the
 * message in qs_last_error(done[k])) for k < *n_done.  The arrays must have room
 * for every member (max_done >= qs_group_members); the observables of a finished
 * handle are read as after qs_solve_scf. */
int qs_group_step(qs_group *g, int32_t max_done, qs_handle **done, qs_scf_result *res, int32_t *status,
                  int32_t *n_done);
/* qs_group_step in two halves: _begin enqueues the step and returns at once, _end
 * waits for it and reports as qs_group_step.  A caller with two groups can begin
 * one while the other's step runs, so that one group's Dyson sweep overlaps the
 * other's memory-bound stages.  No qs_group_add between the halves. */
int qs_group_step_begin(qs_group *g);
int qs_group_step_end(qs_group *g, int32_t max_done, qs_handle **done, qs_scf_result *res, int32_t *status,
                      int32_t *n_done);
/* measurement: Dyson launches of the last step and how many members ran in
 * grouped launches; HIP-event time of all steps' Dyson sweeps and the step count */
int qs_group_info(const qs_group *g, int32_t *launches, int32_t *fused_members, double *dyson_ms, int64_t *steps);

/* --- seams inside the SCF step (order 2) ------------------------------------ */
/* sweep(M_INIT2) + pack_gf (src/scf.f90:116-122): the non-interacting active
 * G^r, G^< packed into the handle's x; copied to x_out (n_vec) if not NULL */
int qs_init2(qs_handle *h, double V, double *x_out);
/* scf_residual2 (src/scf.f90:158-188), the residual_function contract of
 * common/broyden_mixing.f90:39-47: F = pack(G_out[unpack(x)]) - x, or F = 0 and
 * *abort = 1 on a runaway.  x and F are host arrays of length n (= n_vec). */
int qs_residual2(qs_handle *h, double V, int64_t n, const double *x, double *F, int32_t *abort_flag);
/* sigma2_compute (src/self_energies.f90:135-219) of the input of the last
 * residual: SigR2, SigL2 as (nA,nA,n_omega) complex column-major, and Sigma1 as
 * (norb,norb) complex (src/self_energies.f90:45-72); any pointer may be NULL */
int qs_get_sigma(const qs_handle *h, double *SigR2, double *SigL2, double *Sig1);
/* sweep(M_FINAL) + gather_onsite + gather_atoms + final_sanity with the
 * self-energies of the last residual (src/scf.f90:134-140) */
int qs_final_sweep(qs_handle *h, double V, qs_scf_result *res);

/* --- instrumentation (no counterpart in the reference, which only times the
 * whole run in runtime_datasheet.dat, src/quichotte.f90:50-58,149-155) ------ */
/* enable (1) or disable (0) per-stage HIP-event timing on the handle's stream;
 * resets the totals */
int qs_profile(qs_handle *h, int32_t enable);
/* total milliseconds and launch counts per stage; returns the number of stages */
int qs_profile_read(qs_handle *h, double *ms, int64_t *count, int32_t nstage);
const char *qs_stage_name(int32_t stage);
/* launch shape of the Dyson kernel picked for this size at qs_create (0: 256
 * threads, 1: 512 threads per energy with the matrix in registers and LDS, 2:
 * 16x16 or 8x8 threads with a per-energy scratch slab, 3: the large-Norb chunked
 * path on MFMA, above 96 orbitals) and the timed ms of each on a slice of the
 * grid (0 for a shape that does not apply to this size); returns the number of
 * shapes */
int qs_dyson_choice(const qs_handle *h, int32_t *variant, double *ms_per_variant, int32_t nmax);
/* a description of the launch shape the handle uses; *uses_mfma (may be NULL)
 * is 1 when its dense contractions issue MFMA instructions, 0 when they are
 * FP64 vector FMAs */
const char *qs_dyson_shape(const qs_handle *h, int32_t *uses_mfma);

/* ---- SANCHO: interaction self-energies of orders 2 and 3 ----------------
 * The work arrays of SANCHO's SelfEnergies module (SANCHO/src/self_energies.f90:23-35,
 * alloc_sigma_work :209-230) for na sites, n_omega energies and order 2 or 3 on
 * one device.  hbar = 1 (SANCHO/common/constants.f90:18). */
typedef struct qs_sancho qs_sancho;
int qs_sancho_create(int32_t na, int32_t n_omega, int32_t order, int32_t device, qs_sancho **out);
void qs_sancho_destroy(qs_sancho *s);
const char *qs_sancho_last_error(const qs_sancho *s);
int64_t qs_sancho_device_bytes(const qs_sancho *s);
/* complete_input_gf + build_self_energies (SANCHO/src/green_functions.f90:163-173,
 * SANCHO/src/self_energies.f90:38-46): G^a, G^> from G^r, G^<, then the bubbles,
 * Sig2R/L/G and, at order 3, Sig3R/L/G.  GR, GL and the Sig arrays are
 * (na, na, n_omega) complex128, column-major (G_in%r, G_in%L, Sig2R, ...);
 * U_site (na) and delta are host values.  on_device = 0: the arrays are host
 * memory; 1: device memory on the handle's device.  Synchronous: returns when
 * the outputs are written.  The work runs on the handle's own stream, which is
 * not ordered after any stream of the caller, so device inputs must be complete
 * (the caller's writes to them finished) before the call.  Calls on one handle
 * must not overlap. */
int qs_sancho_self_energies(qs_sancho *s, double delta, const double *U_site, const double *GR, const double *GL,
                            double *Sig2R, double *Sig2L, double *Sig2G, double *Sig3R, double *Sig3L,
                            double *Sig3G, int32_t on_device);
/* measurement hooks (no counterpart in the reference), as qs_profile above:
 * enable (1) or disable (0) HIP-event timing of the two summation stages on the
 * handle's stream (0: bubbles chi/phi, 1: self-energies; each with its
 * partial-sum reduction); resets the totals.  qs_sancho_profile_read fills
 * ms[2] and count[2] and returns 2, or a negative status. */
int qs_sancho_profile(qs_sancho *s, int32_t enable);
int qs_sancho_profile_read(const qs_sancho *s, double *ms, int64_t *count);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): qs_sancho_self_energies on device arrays with no handle and
 * no allocation.  The work arrays (completed G copies, bubbles, split partial
 * sums, U) live in a fixed buffer inside the library, qs_sancho_static_bytes()
 * = 64 MiB, one copy per device, held until the process ends.  A size whose work
 * arrays need more (qs_sancho_static_need) returns QS_ERR_UNSUPPORTED before any
 * GPU work.  Runs on the calling thread's current device, where GR, GL and the
 * Sig arrays must live, on `stream` (a hipStream_t; NULL for the default stream),
 * and returns after that stream has finished.  Every call on a device shares
 * that device's buffer: calls that overlap on one device race.  Results are
 * those of qs_sancho_self_energies with on_device = 1. */
int64_t qs_sancho_static_bytes(void);
/* bytes of that buffer a call of these sizes uses on the current device; -1 if invalid */
int64_t qs_sancho_static_need(int32_t na, int32_t n_omega, int32_t order);
int qs_sancho_self_energies_static(int32_t na, int32_t n_omega, int32_t order, double delta, const double *U_site,
                                   const double *GR, const double *GL, double *Sig2R, double *Sig2L, double *Sig2G,
                                   double *Sig3R, double *Sig3L, double *Sig3G, void *stream);

#ifdef __cplusplus
}
#endif
#endif
