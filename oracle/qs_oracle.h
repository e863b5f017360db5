/*
 * qs_oracle.h -- CPU FP64 restatement of QUICHOTTE's order-0/1/2 SCF path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker and the host
 * CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product (the HIP library and its host
 * code) never links, loads or calls anything in oracle/.
 *
 * Each function cites the reference routine it restates (paths relative to the
 * reference repository).  Complex numbers are interleaved (re, im) doubles;
 * matrices are column-major, indices are 0-based here where the reference is
 * 1-based.
 *
 * Pinning (see tests/test_oracle_*.py):
 *   - end to end against the reference's committed fixtures
 *     of the order 0, 1 and 2 test cases (1e-6 abs + 1e-5 rel);
 *   - orc_fft against a brute-force DFT and numpy.fft;
 *   - orc_invert against numpy.linalg.inv and the closed form of a 1x1 level;
 *   - orc_sigma2 against a brute-force direct double sum (no FFT) on tiny grids;
 *   - orc_pack_gf/orc_unpack_gf against the index map and round trip;
 *   - orc_broyden on an affine contraction with a known fixed point.
 */
#ifndef QS_ORACLE_H
#define QS_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Problem description: everything solve_scf reads from module state
 * (src/green_functions.f90:38-67, src/parameters.f90, src/scf.f90:38-43). */
typedef struct orc_problem {
    int32_t norb;            /* spin-orbitals, 2*N_sites */
    int32_t n_omega;         /* energy points, omega_i = w_init + i*delta, i = 1..n_omega */
    double w_init, delta;
    const double *H;         /* (norb,norb) complex, column-major */
    const double *gammaL;    /* (norb,norb) complex */
    const double *gammaR;    /* (norb,norb) complex */
    const double *U_site;    /* (norb) real; U on the spin-up index 2i-1 */
    int32_t order;           /* 0, 1, 2 */
    double beta, mu;         /* 1/(kB T), chemical potential */
    int32_t n_atoms;         /* atoms; 0 => one orbital per atom */
    const int32_t *atom_norb;/* (n_atoms) spatial orbitals per atom, or NULL */
    double alpha_mix;
    int32_t nhist, maxscf;
    double tolf, tolf_order1;
    int32_t single_history;  /* broyden_single_precision */
    int32_t nthreads;        /* OpenMP threads (<=0: runtime default) */
} orc_problem;

typedef struct orc_state orc_state;

orc_state *orc_create(const orc_problem *p);
void orc_destroy(orc_state *s);

/* sizes of the active subspace (src/green_functions.f90:90-190) */
void orc_sizes(const orc_state *s, int32_t *nA, int32_t *nS, int32_t *nCL, int32_t *nAs,
               int64_t *n_vec, int32_t *nf);
void orc_maps(const orc_state *s, int32_t *act, int32_t *sup, int32_t *cl);

/* solve_scf (src/scf.f90:78-151).  Returns 0 or a nonzero error status
 * (singular matrix).  Per-iteration max|F| and rms|F| go to log_fmax/log_rms
 * (length >= maxscf, may be NULL). */
int orc_solve_scf(orc_state *s, double V, int32_t fresh, int32_t *converged, int32_t *n_iter,
                  double *fmax, double *log_fmax, double *log_rms);

/* observables of the last final sweep */
void orc_current(const orc_state *s, double *J_up, double *J_down);   /* src/observables.f90:22-39 */
void orc_g_less_obs(const orc_state *s, double *G);                   /* (norb,norb) complex */
void orc_dosd(const orc_state *s, double *dosd);                      /* (norb,n_omega) */
void orc_trUD(const orc_state *s, double *trU, double *trD);          /* (n_omega) each */
void orc_final_flags(const orc_state *s, int32_t *runaway, int32_t *unphysical);
void orc_final_worst(const orc_state *s, int32_t *i_worst, double *n_worst, double *w_worst);

/* kernel-level seams */
int orc_init2(orc_state *s, double V, double *x);                     /* sweep(M_INIT2) + pack_gf */
int orc_residual2(orc_state *s, double V, int64_t n, const double *x, double *F, int32_t *abort_flag);
int orc_final_sweep(orc_state *s, double V);                          /* sweep(M_FINAL) + gathers */
/* the order-2 input of the last residual: G_in%r, G_in%L (nA,nA,n_omega) complex */
void orc_get_gin(const orc_state *s, double *Gr, double *GL);
void orc_get_sigma2(const orc_state *s, double *SigR, double *SigL);  /* (nA,nA,n_omega) */
void orc_get_sig1(const orc_state *s, double *Sig1);                  /* (norb,norb) */
void orc_get_diag(const orc_state *s, double *gRd, double *gLd);      /* (nA,n_omega) complex */

/* standalone pieces */
int orc_good_fft_size(int nmin);                                      /* src/self_energies.f90:75-89 */
void orc_fft(int n, int sign, double *data);                          /* FFTW convention, unnormalised */
int orc_invert(int n, double *A);                                     /* zgetrf + zgetri semantics */
/* Sigma2 from given (nA,nA,n_omega) G^r, G^< on an explicit active structure
 * (src/self_energies.f90:135-219).  U_act: U on each active state (nA);
 * pairs of states (2k, 2k+1) are the active sites. */
void orc_sigma2(int nA, int n_omega, double delta, const double *U_act, const double *Gr,
                const double *GL, double *SigR, double *SigL, int do_lesser);
void orc_pack_gf(const double *Gr, const double *GL, double *x, int nb, int Nw);     /* linear_algebra.f90:155-191 */
void orc_unpack_gf(const double *x, double *Gr, double *GL, int nb, int Nw);         /* linear_algebra.f90:193-233 */
/* check_diag_sanity (common/scf_checks.f90:59-91); state_index may be NULL */
void orc_check_diag_sanity(const double *Grd, const double *GLd, int nb, int nw, double delta,
                           int32_t *runaway, int32_t *unphysical, int32_t *i_worst,
                           double *n_worst, double *w_worst, const int32_t *state_index);
double orc_fermi(double w, double V, double mu, double beta);         /* green_functions.f90:74-84 */

/* broyden_solve (common/broyden_mixing.f90:51-184) on a user residual */
typedef void (*orc_residual_fn)(void *ctx, int64_t n, const double *x, double *F, int32_t *abort_flag);
double orc_bench_dyson(const orc_problem *p);
double orc_bench_sigma2(int n_omega, int n_cols, int n_pairs, int nthreads);
double orc_bench_broyden(int64_t n, int nhist, int nthreads);

void orc_broyden(orc_residual_fn fn, void *ctx, double *x, int64_t n, double alpha_mix, int nhist,
                 int maxscf, double tolf, int single_history, int32_t *converged, int32_t *n_iter,
                 double *fmax, double *log_fmax, double *log_rms);

/* SANCHO (sancho_oracle.c): complete_input_gf + build_self_energies
 * (SANCHO/src/green_functions.f90:163-173, SANCHO/src/self_energies.f90:38-207).
 * GR, GL (na,na,nw) complex; U (na); outputs (na,na,nw) complex, Sig3* only for
 * order 3.  hbar = 1.  Returns 0, or -1 on bad arguments or allocation failure. */
int orc_sancho_sigma(int32_t na, int32_t nw, double delta, const double *U, int32_t order, const double *GR,
                     const double *GL, double *S2R, double *S2L, double *S2G, double *S3R, double *S3L,
                     double *S3G);
/* the bubbles alone: chi (4, na, na, 2nw-1) in the order G, L, R, A with m = -(nw-1).. at m + nw - 1
 * (innermost), then a, then b; phi likewise with n = 2.. at n - 2 (order 3 only) */
int orc_sancho_bubbles(int32_t na, int32_t nw, double delta, int32_t order, const double *GR, const double *GL,
                       double *chi, double *phi);

#ifdef __cplusplus
}
#endif
#endif
