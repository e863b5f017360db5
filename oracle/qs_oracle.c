/*
 * qs_oracle.c -- CPU FP64 restatement of QUICHOTTE's SCF path (orders 0, 1, 2).
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker and the host CPU baseline.
 * The shipped HIP library never links or calls this code.
 *
 * Restated routines (reference paths):
 *   solve_scf, scf_residual1/2, pack_nil/unpack_nil,
 *   final_sanity, sweep, G_at_energy ........... src/scf.f90:78-415
 *   sigma_first_order, good_fft_size,
 *   sigma2_compute, adj2 ........................ src/self_energies.f90:45-228
 *   fermi, setup_storage, equal_time_lesser,
 *   gather_onsite, gather_atoms ................. src/green_functions.f90:74-257
 *   Current ..................................... src/observables.f90:22-39
 *   broyden_solve, dot_ss, dot_sd ............... common/broyden_mixing.f90:51-215
 *   invert_complex, pack_gf, unpack_gf .......... common/linear_algebra.f90:58-88,155-233
 *   check_diag_sanity ........................... common/scf_checks.f90:59-91
 * Third-party arithmetic restated from its published algorithm:
 *   LAPACK zgetf2/ztrti2/zgetri (unblocked LU with partial pivoting on
 *   |Re|+|Im|, inverse of U, back substitution, reverse column swaps),
 *   LAPACK dpotrf/dpotrs for dposv('U'), and a mixed-radix (2,3,4,5,7)
 *   DFT with FFTW's sign convention (forward e^{-2 pi i kn/N}, backward
 *   e^{+...}, both unnormalised).
 */
#include "qs_oracle.h"
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex zc;
#define PI_ (4.0 * atan(1.0))   /* common/constants.f90:20 */
static const double HBAR = 1.0;  /* common/constants.f90:18 */

#define IDX(i, j, ld) ((size_t)(j) * (size_t)(ld) + (size_t)(i))

/* ------------------------------------------------------------------------ */
/* small dense helpers (column-major)                                        */
/* ------------------------------------------------------------------------ */

/* C(m,n) = A(m,k) * B(k,n) */
static void zmm(int m, int n, int k, const zc *A, int lda, const zc *B, int ldb, zc *C, int ldc) {
    for (int j = 0; j < n; j++) {
        for (int i = 0; i < m; i++) C[IDX(i, j, ldc)] = 0;
        for (int l = 0; l < k; l++) {
            zc b = B[IDX(l, j, ldb)];
            if (b == 0) continue;
            const zc *a = A + (size_t)l * lda;
            zc *c = C + (size_t)j * ldc;
            for (int i = 0; i < m; i++) c[i] += a[i] * b;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* LAPACK-semantics inverse: zgetf2 + ztrti2 + zgetri (unblocked)            */
/* ------------------------------------------------------------------------ */
static double cabs1(zc z) { return fabs(creal(z)) + fabs(cimag(z)); }

int orc_invert(int n, double *Ad) {
    zc *A = (zc *)Ad;
    int *ipiv = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
    zc *work = (zc *)malloc(sizeof(zc) * (n > 0 ? n : 1));
    int info = 0;
    /* zgetf2 */
    for (int j = 0; j < n; j++) {
        int jp = j;
        double best = cabs1(A[IDX(j, j, n)]);
        for (int i = j + 1; i < n; i++) {
            double v = cabs1(A[IDX(i, j, n)]);
            if (v > best) { best = v; jp = i; }
        }
        ipiv[j] = jp;
        if (A[IDX(jp, j, n)] != 0) {
            if (jp != j)
                for (int c = 0; c < n; c++) {
                    zc t = A[IDX(j, c, n)]; A[IDX(j, c, n)] = A[IDX(jp, c, n)]; A[IDX(jp, c, n)] = t;
                }
            zc piv = A[IDX(j, j, n)];
            if (cabs(piv) >= DBL_MIN) {
                zc r = 1.0 / piv;
                for (int i = j + 1; i < n; i++) A[IDX(i, j, n)] *= r;
            } else {
                for (int i = j + 1; i < n; i++) A[IDX(i, j, n)] /= piv;
            }
        } else if (info == 0) {
            info = j + 1;
        }
        for (int c = j + 1; c < n; c++) {
            zc u = A[IDX(j, c, n)];
            if (u == 0) continue;
            for (int i = j + 1; i < n; i++) A[IDX(i, c, n)] -= A[IDX(i, j, n)] * u;
        }
    }
    if (info != 0) { free(ipiv); free(work); return info; }
    /* ztrti2: inverse of the upper triangle, non-unit */
    for (int j = 0; j < n; j++) {
        A[IDX(j, j, n)] = 1.0 / A[IDX(j, j, n)];
        zc ajj = -A[IDX(j, j, n)];
        /* x = A(0:j-1, j); x = U^{-1}(0:j-1,0:j-1) * x (upper, non-unit, ztrmv) */
        for (int i = 0; i < j; i++) work[i] = A[IDX(i, j, n)];
        for (int i = 0; i < j; i++) {
            zc s = 0;
            for (int l = i; l < j; l++) s += A[IDX(i, l, n)] * work[l];
            A[IDX(i, j, n)] = s * ajj;
        }
    }
    /* zgetri: solve inv(A)*L = inv(U) */
    for (int j = n - 1; j >= 0; j--) {
        for (int i = j + 1; i < n; i++) { work[i] = A[IDX(i, j, n)]; A[IDX(i, j, n)] = 0; }
        if (j < n - 1)
            for (int l = j + 1; l < n; l++) {
                zc w = work[l];
                if (w == 0) continue;
                for (int i = 0; i < n; i++) A[IDX(i, j, n)] -= A[IDX(i, l, n)] * w;
            }
    }
    for (int j = n - 2; j >= 0; j--) {
        int jp = ipiv[j];
        if (jp != j)
            for (int i = 0; i < n; i++) {
                zc t = A[IDX(i, j, n)]; A[IDX(i, j, n)] = A[IDX(i, jp, n)]; A[IDX(i, jp, n)] = t;
            }
    }
    free(ipiv); free(work);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* mixed-radix DFT, FFTW sign convention                                     */
/* ------------------------------------------------------------------------ */
typedef struct { int n, nfac, fac[40]; zc *tw; } fftplan;

static void plan_make(fftplan *pl, int n) {
    pl->n = n; pl->nfac = 0;
    int m = n;
    static const int rad[] = {4, 2, 3, 5, 7};
    for (int r = 0; r < 5; r++)
        while (m % rad[r] == 0 && m > 1) { pl->fac[pl->nfac++] = rad[r]; m /= rad[r]; }
    if (m > 1) pl->fac[pl->nfac++] = m; /* generic odd factor: direct DFT */
    pl->tw = (zc *)malloc(sizeof(zc) * n);
    for (int k = 0; k < n; k++) {
        double a = -2.0 * PI_ * (double)k / (double)n;
        pl->tw[k] = cos(a) + I * sin(a);
    }
}

/* out[0..n) = DFT of in[0], in[s], in[2s], ... (n points) */
static void fft_rec(const zc *in, zc *out, int n, int s, const int *fac, const fftplan *pl, int sign) {
    if (n == 1) { out[0] = in[0]; return; }
    int p = fac[0], m = n / p;
    for (int r = 0; r < p; r++) fft_rec(in + (size_t)r * s, out + (size_t)r * m, m, s * p, fac + 1, pl, sign);
    int tws = pl->n / n;
    zc t[64], u[64];
    for (int k = 0; k < m; k++) {
        for (int r = 0; r < p; r++) {
            zc w = pl->tw[((size_t)r * k * tws) % pl->n];
            if (sign > 0) w = conj(w);
            t[r] = out[(size_t)r * m + k] * w;
        }
        for (int q = 0; q < p; q++) {
            zc acc = 0;
            for (int r = 0; r < p; r++) {
                zc w = pl->tw[((size_t)((r * q) % p) * m * tws) % pl->n];
                if (sign > 0) w = conj(w);
                acc += t[r] * w;
            }
            u[q] = acc;
        }
        for (int q = 0; q < p; q++) out[(size_t)q * m + k] = u[q];
    }
}

/* Power-of-two lengths: iterative radix-2 Stockham (autosorting, no bit
 * reversal), the same DFT as fft_rec.  Stage with half-length l and stride m
 * (n = 2 l m): b[k + 2jm] = a[k + jm] + a[k + jm + lm],
 * b[k + 2jm + m] = w^j (a[k + jm] - a[k + jm + lm]), w = e^{-+2 pi i m / n}. */
static void fft_pow2(const fftplan *pl, int sign, zc *data, zc *scratch) {
    const int n = pl->n;
    zc *a = data, *b = scratch;
    for (int l = n / 2, m = 1; l >= 1; l /= 2, m *= 2) {
        for (int j = 0; j < l; j++) {
            const zc w = pl->tw[(size_t)j * m];
            const double wr = creal(w), wi = sign > 0 ? -cimag(w) : cimag(w);
            const zc *a0 = a + (size_t)j * m, *a1 = a + (size_t)j * m + (size_t)l * m;
            zc *b0 = b + (size_t)2 * j * m, *b1 = b0 + m;
            for (int k = 0; k < m; k++) {
                const double xr = creal(a0[k]), xi = cimag(a0[k]), yr = creal(a1[k]), yi = cimag(a1[k]);
                const double dr = xr - yr, di = xi - yi;
                b0[k] = (xr + yr) + I * (xi + yi);
                b1[k] = (wr * dr - wi * di) + I * (wr * di + wi * dr);
            }
        }
        zc *t = a; a = b; b = t;
    }
    if (a != data) memcpy(data, a, sizeof(zc) * n);
}

static void fft_exec(const fftplan *pl, int sign, zc *data, zc *scratch) {
    if ((pl->n & (pl->n - 1)) == 0) { fft_pow2(pl, sign, data, scratch); return; }
    fft_rec(data, scratch, pl->n, 1, pl->fac, pl, sign);
    memcpy(data, scratch, sizeof(zc) * pl->n);
}

void orc_fft(int n, int sign, double *data) {
    fftplan pl; plan_make(&pl, n);
    zc *scr = (zc *)malloc(sizeof(zc) * n);
    fft_exec(&pl, sign, (zc *)data, scr);
    free(scr); free(pl.tw);
}

int orc_good_fft_size(int nmin) {
    int n = nmin;
    for (;;) {
        int m = n;
        while (m % 2 == 0) m /= 2;
        while (m % 3 == 0) m /= 3;
        while (m % 5 == 0) m /= 5;
        while (m % 7 == 0) m /= 7;
        if (m == 1) break;
        n++;
    }
    return n;
}

/* ------------------------------------------------------------------------ */
/* state                                                                     */
/* ------------------------------------------------------------------------ */
enum { M_INIT2 = 1, M_SCF2 = 2, M_SCF1 = 3, M_FINAL = 4 };

struct orc_state {
    orc_problem p;
    int norb, nw, nsites, order;
    zc *H, *GamL, *GamR, *Hemb, *Sig1, *GamL_s, *GamR_s, *GamL_cl, *G_less_t0, *G_less_obs;
    double *U, *omega;
    int nA, nS, nCL, nAs, n_int;
    int *act, *pos, *sup, *cl, *asite;
    char *site_int, *site_act;
    int n_atoms, *atom_norb, *atom_first, *atom_off, has_atomL;
    zc *Gin_r, *Gin_L, *SigR2, *SigL2, *gRd, *gLd, *onsL, *atomL;
    double *dosd, *trU, *trD;
    int Nf, n_full, n_su, nblk;
    int64_t n_vec;
    int use_sigma;
    double V_scf;
    int runaway_final, unphys_final, iworst_final;
    double nworst_final, wworst_final;
    volatile int err;
};

double orc_fermi(double w, double V, double mu, double beta) {
    double arg = (w - mu + V / HBAR) * beta;
    double e = HBAR * arg;
    if (e > 700.0) e = 700.0;
    return 1.0 / (exp(e) + 1.0);
}

static zc *zalloc(size_t n) { return (zc *)calloc(n ? n : 1, sizeof(zc)); }

orc_state *orc_create(const orc_problem *p) {
    orc_state *s = (orc_state *)calloc(1, sizeof(orc_state));
    s->p = *p;
    int n = p->norb;
    s->norb = n; s->nw = p->n_omega; s->nsites = n / 2; s->order = p->order;
    s->H = zalloc((size_t)n * n); s->GamL = zalloc((size_t)n * n); s->GamR = zalloc((size_t)n * n);
    memcpy(s->H, p->H, sizeof(zc) * n * n);
    memcpy(s->GamL, p->gammaL, sizeof(zc) * n * n);
    memcpy(s->GamR, p->gammaR, sizeof(zc) * n * n);
    s->U = (double *)malloc(sizeof(double) * n);
    memcpy(s->U, p->U_site, sizeof(double) * n);
    s->omega = (double *)malloc(sizeof(double) * s->nw);
    for (int i = 1; i <= s->nw; i++) s->omega[i - 1] = p->w_init + i * p->delta;  /* quichotte.f90:91-93 */

    /* atoms (src/input_output.f90 set_atom_map) */
    if (p->n_atoms > 0 && p->atom_norb) {
        s->n_atoms = p->n_atoms;
        s->atom_norb = (int *)malloc(sizeof(int) * s->n_atoms);
        for (int a = 0; a < s->n_atoms; a++) s->atom_norb[a] = p->atom_norb[a];
    } else {
        s->n_atoms = s->nsites;
        s->atom_norb = (int *)malloc(sizeof(int) * s->n_atoms);
        for (int a = 0; a < s->n_atoms; a++) s->atom_norb[a] = 1;
    }
    s->atom_first = (int *)malloc(sizeof(int) * s->n_atoms);
    int acc = 0, maxn = 0;
    for (int a = 0; a < s->n_atoms; a++) {
        s->atom_first[a] = acc; acc += s->atom_norb[a];
        if (s->atom_norb[a] > maxn) maxn = s->atom_norb[a];
    }

    /* setup_storage (src/green_functions.f90:90-190) */
    int ns = s->nsites;
    s->site_int = (char *)calloc(ns, 1); s->site_act = (char *)calloc(ns, 1);
    s->pos = (int *)malloc(sizeof(int) * n);
    char *conL = (char *)calloc(n, 1), *conR = (char *)calloc(n, 1);
    s->n_int = 0;
    for (int si = 0; si < ns; si++) {
        s->site_int[si] = (s->order >= 1 && s->U[2 * si] != 0.0);
        s->n_int += s->site_int[si];
        s->site_act[si] = s->site_int[si];
    }
    for (int i = 0; i < n; i++) {
        for (int k = 0; k < n; k++) {
            if (s->GamL[IDX(i, k, n)] != 0 || s->GamL[IDX(k, i, n)] != 0) conL[i] = 1;
            if (s->GamR[IDX(i, k, n)] != 0 || s->GamR[IDX(k, i, n)] != 0) conR[i] = 1;
        }
        if (conL[i] || conR[i]) s->site_act[i / 2] = 1;
    }
    s->nAs = 0;
    for (int si = 0; si < ns; si++) s->nAs += s->site_act[si];
    s->nA = 2 * s->nAs;
    s->act = (int *)malloc(sizeof(int) * (s->nA ? s->nA : 1));
    int k = 0;
    for (int i = 0; i < n; i++) {
        s->pos[i] = -1;
        if (s->site_act[i / 2]) { s->act[k] = i; s->pos[i] = k; k++; }
    }
    s->asite = (int *)malloc(sizeof(int) * (s->nAs ? s->nAs : 1));
    k = 0;
    for (int si = 0; si < ns; si++) if (s->site_act[si]) s->asite[k++] = si;
    s->nCL = 0;
    for (int i = 0; i < n; i++) s->nCL += conL[i];
    s->cl = (int *)malloc(sizeof(int) * (s->nCL ? s->nCL : 1));
    k = 0;
    for (int i = 0; i < n; i++) if (conL[i]) s->cl[k++] = i;
    if (s->order == 2) {
        s->nS = s->nA;
        s->sup = (int *)malloc(sizeof(int) * (s->nS ? s->nS : 1));
        for (int i = 0; i < s->nS; i++) s->sup[i] = s->act[i];
    } else {
        s->nS = 0;
        for (int i = 0; i < n; i++) s->nS += (conL[i] || conR[i]);
        s->sup = (int *)malloc(sizeof(int) * (s->nS ? s->nS : 1));
        k = 0;
        for (int i = 0; i < n; i++) if (conL[i] || conR[i]) s->sup[k++] = i;
    }
    free(conL); free(conR);
    int nS = s->nS, nCL = s->nCL;
    s->GamL_s = zalloc((size_t)nS * nS); s->GamR_s = zalloc((size_t)nS * nS); s->GamL_cl = zalloc((size_t)nCL * nCL);
    for (int b = 0; b < nS; b++)
        for (int a = 0; a < nS; a++) {
            s->GamL_s[IDX(a, b, nS)] = s->GamL[IDX(s->sup[a], s->sup[b], n)];
            s->GamR_s[IDX(a, b, nS)] = s->GamR[IDX(s->sup[a], s->sup[b], n)];
        }
    for (int b = 0; b < nCL; b++)
        for (int a = 0; a < nCL; a++) s->GamL_cl[IDX(a, b, nCL)] = s->GamL[IDX(s->cl[a], s->cl[b], n)];
    s->Hemb = zalloc((size_t)n * n); s->Sig1 = zalloc((size_t)n * n);
    s->G_less_t0 = zalloc((size_t)n * n); s->G_less_obs = zalloc((size_t)n * n);
    s->onsL = zalloc((size_t)4 * ns * s->nw);
    s->dosd = (double *)calloc((size_t)n * s->nw, sizeof(double));
    s->trU = (double *)calloc(s->nw, sizeof(double));
    s->trD = (double *)calloc(s->nw, sizeof(double));
    if (maxn > 1) {
        s->has_atomL = 1;
        s->atom_off = (int *)malloc(sizeof(int) * (s->n_atoms + 1));
        s->atom_off[0] = 0;
        for (int a = 0; a < s->n_atoms; a++) s->atom_off[a + 1] = s->atom_off[a] + 4 * s->atom_norb[a] * s->atom_norb[a];
        s->atomL = zalloc((size_t)s->atom_off[s->n_atoms] * s->nw);
    }
    if (s->order >= 1) { s->gRd = zalloc((size_t)s->nA * s->nw); s->gLd = zalloc((size_t)s->nA * s->nw); }
    int nA = s->nA;
    s->n_full = nA * nA; s->n_su = (s->n_full - nA) / 2;
    s->nblk = 2 * s->n_full + s->n_su + s->n_su + nA;
    if (s->order == 1) s->n_vec = 8 * s->n_int;
    else if (s->order == 2) {
        s->n_vec = (int64_t)s->nblk * s->nw;
        s->Gin_r = zalloc((size_t)nA * nA * s->nw);
        s->Gin_L = zalloc((size_t)nA * nA * s->nw);
    } else s->n_vec = 0;
    s->Nf = orc_good_fft_size(2 * s->nw - 1);
    /* Hemb = -H + (i/2hbar)(GammaL + GammaR)  (src/scf.f90:92) */
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++)
            s->Hemb[IDX(i, j, n)] = -s->H[IDX(i, j, n)] + 0.5 * (I / HBAR) * (s->GamL[IDX(i, j, n)] + s->GamR[IDX(i, j, n)]);
    return s;
}

void orc_destroy(orc_state *s) {
    if (!s) return;
    free(s->H); free(s->GamL); free(s->GamR); free(s->Hemb); free(s->Sig1); free(s->GamL_s);
    free(s->GamR_s); free(s->GamL_cl); free(s->G_less_t0); free(s->G_less_obs); free(s->U);
    free(s->omega); free(s->act); free(s->pos); free(s->sup); free(s->cl); free(s->asite);
    free(s->site_int); free(s->site_act); free(s->atom_norb); free(s->atom_first); free(s->atom_off);
    free(s->Gin_r); free(s->Gin_L); free(s->SigR2); free(s->SigL2); free(s->gRd); free(s->gLd);
    free(s->onsL); free(s->atomL); free(s->dosd); free(s->trU); free(s->trD);
    free(s);
}

void orc_sizes(const orc_state *s, int32_t *nA, int32_t *nS, int32_t *nCL, int32_t *nAs, int64_t *n_vec,
               int32_t *nf) {
    if (nA) *nA = s->nA;
    if (nS) *nS = s->nS;
    if (nCL) *nCL = s->nCL;
    if (nAs) *nAs = s->nAs;
    if (n_vec) *n_vec = s->n_vec;
    if (nf) *nf = s->Nf;
}

void orc_maps(const orc_state *s, int32_t *act, int32_t *sup, int32_t *cl) {
    if (act) for (int i = 0; i < s->nA; i++) act[i] = s->act[i];
    if (sup) for (int i = 0; i < s->nS; i++) sup[i] = s->sup[i];
    if (cl) for (int i = 0; i < s->nCL; i++) cl[i] = s->cl[i];
}

/* ------------------------------------------------------------------------ */
/* pack / unpack (common/linear_algebra.f90:155-233)                         */
/* ------------------------------------------------------------------------ */
void orc_pack_gf(const double *Grd, const double *GLd, double *out, int nb, int Nw) {
    const zc *Gr = (const zc *)Grd, *GL = (const zc *)GLd;
    int n_full = nb * nb, n_su = (n_full - nb) / 2;
    size_t nblock = 2 * (size_t)n_full + n_su + (n_su + nb);
    for (int iw = 0; iw < Nw; iw++) {
        size_t off = iw * nblock;
        const zc *r = Gr + (size_t)iw * n_full, *l = GL + (size_t)iw * n_full;
        size_t m = 0;
        for (int i = 0; i < nb; i++)
            for (int j = 0; j < nb; j++) {
                out[off + m] = creal(r[IDX(i, j, nb)]);
                out[off + m + n_full] = cimag(r[IDX(i, j, nb)]);
                m++;
            }
        size_t mi = 2 * (size_t)n_full + n_su, mr = 2 * (size_t)n_full;
        for (int i = 0; i < nb; i++)
            for (int j = i; j < nb; j++) {
                if (i != j) out[off + mr++] = creal(l[IDX(i, j, nb)]);
                out[off + mi++] = cimag(l[IDX(i, j, nb)]);
            }
    }
}

void orc_unpack_gf(const double *in, double *Grd, double *GLd, int nb, int Nw) {
    zc *Gr = (zc *)Grd, *GL = (zc *)GLd;
    int n_full = nb * nb, n_su = (n_full - nb) / 2;
    size_t nblock = 2 * (size_t)n_full + n_su + (n_su + nb);
#pragma omp parallel for schedule(static)
    for (int iw = 0; iw < Nw; iw++) {
        size_t off = iw * nblock;
        zc *r = Gr + (size_t)iw * n_full, *l = GL + (size_t)iw * n_full;
        size_t m = 0;
        for (int i = 0; i < nb; i++)
            for (int j = 0; j < nb; j++) {
                r[IDX(i, j, nb)] = in[off + m] + I * in[off + m + n_full];
                m++;
            }
        size_t mi = 2 * (size_t)n_full + n_su, mr = 2 * (size_t)n_full;
        for (int i = 0; i < nb; i++)
            for (int j = i; j < nb; j++) {
                double re = 0.0;
                if (i != j) re = in[off + mr++];
                l[IDX(i, j, nb)] = re + I * in[off + mi++];
                l[IDX(j, i, nb)] = -conj(l[IDX(i, j, nb)]);
            }
    }
}

/* ------------------------------------------------------------------------ */
/* sanity (common/scf_checks.f90:59-91)                                      */
/* ------------------------------------------------------------------------ */
void orc_check_diag_sanity(const double *Grdd, const double *GLdd, int nb, int nw, double delta,
                           int32_t *runaway, int32_t *unphysical, int32_t *i_worst, double *n_worst,
                           double *w_worst, const int32_t *state_index) {
    const zc *Grd = (const zc *)Grdd, *GLd = (const zc *)GLdd;
    int ra = 0, un = 0, iw_ = 1;
    double nw_ = 0.5, ww_ = 1.0;
    for (int i = 0; i < nb; i++) {
        int lab = state_index ? state_index[i] : i + 1;
        zc sL = 0;
        double sI = 0;
        for (int k = 0; k < nw; k++) { sL += GLd[IDX(i, k, nb)]; sI += cimag(Grd[IDX(i, k, nb)]); }
        double nocc = creal(-I * HBAR * (delta / (2.0 * PI_)) * sL);
        double wsum = -HBAR * (delta / PI_) * sI;
        if (!isfinite(nocc) || !isfinite(wsum)) {
            *runaway = 1; *unphysical = un; *i_worst = lab; *n_worst = nocc; *w_worst = wsum;
            return;
        }
        if (nocc < -1.0 || nocc > 2.0) ra = 1;
        if (nocc < -0.05 || nocc > 1.05 || wsum < 0.5) un = 1;
        if (fabs(nocc - 0.5) > fabs(nw_ - 0.5) || (wsum < ww_ && fabs(nw_ - 0.5) <= 0.55)) {
            iw_ = lab; nw_ = nocc; ww_ = wsum;
        }
    }
    if (ra) un = 1;
    *runaway = ra; *unphysical = un; *i_worst = iw_; *n_worst = nw_; *w_worst = ww_;
}

/* ------------------------------------------------------------------------ */
/* G_at_energy (src/scf.f90:305-415) and sweep (src/scf.f90:280-290)         */
/* ------------------------------------------------------------------------ */
typedef struct { zc *w1, *SL, *T, *Tfull, *tmp; } ws_t;

static int G_at_energy(orc_state *s, ws_t *ws, int iw, double Volt, int mode, double *F) {
    int n = s->norb, nA = s->nA, nS = s->nS;
    double w = s->omega[iw];
    zc *w1 = ws->w1;
    memcpy(w1, s->Hemb, sizeof(zc) * n * n);
    if (s->use_sigma) {
        if (s->order == 2) {
            const zc *sr = s->SigR2 + (size_t)iw * nA * nA;
            for (int b = 0; b < nA; b++)
                for (int a = 0; a < nA; a++)
                    w1[IDX(s->act[a], s->act[b], n)] -= (s->Sig1[IDX(s->act[a], s->act[b], n)] + sr[IDX(a, b, nA)]);
        } else {
            for (int b = 0; b < nA; b++)
                for (int a = 0; a < nA; a++)
                    w1[IDX(s->act[a], s->act[b], n)] -= s->Sig1[IDX(s->act[a], s->act[b], n)];
        }
    }
    for (int i = 0; i < n; i++) w1[IDX(i, i, n)] += HBAR * w;
    int info = orc_invert(n, (double *)w1);
    if (info != 0) return info;

    double fL = orc_fermi(w, Volt, s->p.mu, s->p.beta), fR = orc_fermi(w, 0.0, s->p.mu, s->p.beta);
    zc *SL = ws->SL;
    for (size_t q = 0; q < (size_t)nS * nS; q++) SL[q] = I * (fL * s->GamL_s[q] + fR * s->GamR_s[q]) / HBAR;
    if (s->use_sigma && s->order == 2) {
        const zc *sl = s->SigL2 + (size_t)iw * nA * nA;
        for (size_t q = 0; q < (size_t)nS * nS; q++) SL[q] += sl[q];
    }

    if (mode == M_INIT2 || mode == M_SCF2) {
        zc *grA = ws->tmp, *T = ws->T, *glA = ws->Tfull;
        for (int b = 0; b < nA; b++)
            for (int a = 0; a < nA; a++) grA[IDX(a, b, nA)] = w1[IDX(s->act[a], s->act[b], n)];
        zmm(nA, nA, nA, grA, nA, SL, nA, T, nA);
        /* glA = T * grA^H */
        for (int j = 0; j < nA; j++)
            for (int i = 0; i < nA; i++) {
                zc acc = 0;
                for (int l = 0; l < nA; l++) acc += T[IDX(i, l, nA)] * conj(grA[IDX(j, l, nA)]);
                glA[IDX(i, j, nA)] = acc;
            }
        if (mode == M_INIT2) {
            memcpy(s->Gin_r + (size_t)iw * nA * nA, grA, sizeof(zc) * nA * nA);
            memcpy(s->Gin_L + (size_t)iw * nA * nA, glA, sizeof(zc) * nA * nA);
        } else {
            orc_pack_gf((double *)grA, (double *)glA, F + (size_t)iw * s->nblk, nA, 1);
            for (int k = 0; k < nA; k++) {
                s->gRd[IDX(k, iw, nA)] = grA[IDX(k, k, nA)];
                s->gLd[IDX(k, iw, nA)] = glA[IDX(k, k, nA)];
            }
        }
    } else if (mode == M_SCF1) {
        zc *T = ws->T, *Wsup = ws->tmp;
        /* T(nA,nS) = w1(act,sup) * SL */
        for (int b = 0; b < nS; b++)
            for (int a = 0; a < nA; a++) Wsup[IDX(a, b, nA)] = w1[IDX(s->act[a], s->sup[b], n)];
        zmm(nA, nS, nS, Wsup, nA, SL, nS, T, nA);
        for (int k = 0; k < nA; k++) {
            s->gRd[IDX(k, iw, nA)] = w1[IDX(s->act[k], s->act[k], n)];
            zc acc = 0;
            for (int l = 0; l < nS; l++) acc += T[IDX(k, l, nA)] * conj(w1[IDX(s->act[k], s->sup[l], n)]);
            s->gLd[IDX(k, iw, nA)] = acc;
        }
        for (int si = 0; si < s->nsites; si++) {
            if (!s->site_int[si]) continue;
            int lo = 2 * si, p1 = s->pos[lo];
            for (int b = 0; b < 2; b++)
                for (int a = 0; a < 2; a++) {
                    zc acc = 0;
                    for (int l = 0; l < nS; l++) acc += T[IDX(p1 + a, l, nA)] * conj(w1[IDX(lo + b, s->sup[l], n)]);
                    s->onsL[IDX(4 * si + 2 * b + a, iw, 4 * s->nsites)] = acc;
                }
        }
    } else { /* M_FINAL */
        zc *T = ws->Tfull, *Wsup = ws->tmp;
        for (int b = 0; b < nS; b++)
            for (int a = 0; a < n; a++) Wsup[IDX(a, b, n)] = w1[IDX(a, s->sup[b], n)];
        zmm(n, nS, nS, Wsup, n, SL, nS, T, n);
        for (int i = 0; i < n; i++) s->dosd[IDX(i, iw, n)] = cimag(w1[IDX(i, i, n)]);
        for (int si = 0; si < s->nsites; si++) {
            int lo = 2 * si;
            for (int b = 0; b < 2; b++)
                for (int a = 0; a < 2; a++) {
                    zc acc = 0;
                    for (int l = 0; l < nS; l++) acc += T[IDX(lo + a, l, n)] * conj(w1[IDX(lo + b, s->sup[l], n)]);
                    s->onsL[IDX(4 * si + 2 * b + a, iw, 4 * s->nsites)] = acc;
                }
        }
        if (s->has_atomL) {
            int ld = s->atom_off[s->n_atoms];
            for (int at = 0; at < s->n_atoms; at++) {
                int lo = 2 * s->atom_first[at], m = 2 * s->atom_norb[at];
                for (int b = 0; b < m; b++)
                    for (int a = 0; a < m; a++) {
                        zc acc = 0;
                        for (int l = 0; l < nS; l++) acc += T[IDX(lo + a, l, n)] * conj(w1[IDX(lo + b, s->sup[l], n)]);
                        s->atomL[IDX(s->atom_off[at] + b * m + a, iw, ld)] = acc;
                    }
            }
        }
        int nCL = s->nCL;
        zc trc_u = 0, trc_d = 0;
        zc *X = (zc *)malloc(sizeof(zc) * (nCL ? nCL * nCL : 1));
        for (int a = 0; a < nCL; a++)
            for (int k = 0; k < nCL; k++) {
                zc glc = 0;
                for (int l = 0; l < nS; l++) glc += T[IDX(s->cl[k], l, n)] * conj(w1[IDX(s->cl[a], s->sup[l], n)]);
                X[IDX(k, a, nCL)] = (fL - 1.0) * glc -
                                    fL * (glc + w1[IDX(s->cl[k], s->cl[a], n)] - conj(w1[IDX(s->cl[a], s->cl[k], n)]));
            }
        for (int a = 0; a < nCL; a++) {
            zc mc = 0;
            for (int l = 0; l < nCL; l++) mc += s->GamL_cl[IDX(a, l, nCL)] * X[IDX(l, a, nCL)];
            mc *= I;
            if (s->cl[a] % 2 == 0) trc_u += mc; else trc_d += mc;   /* 0-based even = 1-based odd = spin up */
        }
        free(X);
        s->trU[iw] = creal(trc_u) / (2.0 * PI_);
        s->trD[iw] = creal(trc_d) / (2.0 * PI_);
    }
    return 0;
}

static void set_threads(const orc_state *s) {
#ifdef _OPENMP
    if (s->p.nthreads > 0) omp_set_num_threads(s->p.nthreads);
#else
    (void)s;
#endif
}

static int sweep_range(orc_state *s, int mode, double *F, int iw0, int iw1) {
    int n = s->norb;
    int err = 0;
    set_threads(s);
#pragma omp parallel
    {
        ws_t ws;
        size_t big = (size_t)n * n + 16;
        ws.w1 = zalloc(big); ws.SL = zalloc(big); ws.T = zalloc(big); ws.Tfull = zalloc(big); ws.tmp = zalloc(big);
#pragma omp for schedule(static)
        for (int iw = iw0; iw < iw1; iw++) {
            int e = G_at_energy(s, &ws, iw, s->V_scf, mode, F);
            if (e) {
#pragma omp atomic write
                err = e;
            }
        }
        free(ws.w1); free(ws.SL); free(ws.T); free(ws.Tfull); free(ws.tmp);
    }
    return err;
}

static int sweep(orc_state *s, int mode, double *F) { return sweep_range(s, mode, F, 0, s->nw); }

/* ------------------------------------------------------------------------ */
/* equal-time, Sigma1, gathers                                               */
/* ------------------------------------------------------------------------ */
static void equal_time_lesser(orc_state *s) {   /* green_functions.f90:192-213 */
    int n = s->norb, nA = s->nA;
    double pp = s->p.delta / (2.0 * PI_);
    memset(s->G_less_t0, 0, sizeof(zc) * n * n);
    for (int is = 0; is < s->nAs; is++)
        for (int a = 2 * s->asite[is]; a <= 2 * s->asite[is] + 1; a++)
            for (int b = 2 * s->asite[is]; b <= 2 * s->asite[is] + 1; b++) {
                zc acc = 0;
                for (int k = 0; k < s->nw; k++) acc += s->Gin_L[(size_t)k * nA * nA + IDX(s->pos[a], s->pos[b], nA)];
                s->G_less_t0[IDX(a, b, n)] = acc * pp;
            }
}

static void sigma_first_order(orc_state *s) {   /* self_energies.f90:45-72 */
    int n = s->norb;
    memset(s->Sig1, 0, sizeof(zc) * n * n);
    for (int i = 0; i < n; i += 2) {
        zc hart = s->G_less_t0[IDX(i, i, n)] + s->G_less_t0[IDX(i + 1, i + 1, n)];
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++) {
                if (a == b) s->Sig1[IDX(i + a, i + b, n)] = I * HBAR * s->U[i] * (s->G_less_t0[IDX(i + a, i + b, n)] - hart);
                else s->Sig1[IDX(i + a, i + b, n)] = I * HBAR * s->U[i] * s->G_less_t0[IDX(i + a, i + b, n)];
            }
    }
}

static void gather_onsite(orc_state *s, const char *mask, zc *G) {   /* green_functions.f90:215-238 */
    int n = s->norb, ld = 4 * s->nsites;
    double pp = s->p.delta / (2.0 * PI_);
    for (int si = 0; si < s->nsites; si++) {
        if (mask && !mask[si]) continue;
        for (int b = 0; b < 2; b++)
            for (int a = 0; a < 2; a++) {
                int k = 4 * si + 2 * b + a;
                zc acc = 0;
                for (int iw = 0; iw < s->nw; iw++) acc += s->onsL[IDX(k, iw, ld)];
                G[IDX(2 * si + a, 2 * si + b, n)] = acc * pp;
            }
    }
}

static void gather_atoms(orc_state *s, zc *G) {   /* green_functions.f90:240-257 */
    if (!s->has_atomL) return;
    int n = s->norb, ld = s->atom_off[s->n_atoms];
    double pp = s->p.delta / (2.0 * PI_);
    for (int at = 0; at < s->n_atoms; at++) {
        int m = 2 * s->atom_norb[at], lo = 2 * s->atom_first[at];
        for (int b = 0; b < m; b++)
            for (int a = 0; a < m; a++) {
                zc acc = 0;
                for (int iw = 0; iw < s->nw; iw++) acc += s->atomL[IDX(s->atom_off[at] + b * m + a, iw, ld)];
                G[IDX(lo + a, lo + b, n)] = acc * pp;
            }
    }
}

/* ------------------------------------------------------------------------ */
/* second-order self-energy (src/self_energies.f90:135-228)                  */
/* ------------------------------------------------------------------------ */
static void adj2(const zc M[4], zc A[4]) {  /* column-major 2x2: [0]=11 [1]=21 [2]=12 [3]=22 */
    A[0] = M[3]; A[2] = -M[2]; A[1] = -M[1]; A[3] = M[0];
}
static void mm2(const zc A[4], const zc B[4], zc C[4]) {
    C[0] = A[0] * B[0] + A[2] * B[1];
    C[1] = A[1] * B[0] + A[3] * B[1];
    C[2] = A[0] * B[2] + A[2] * B[3];
    C[3] = A[1] * B[2] + A[3] * B[3];
}
static void dag2(const zc A[4], zc B[4]) { B[0] = conj(A[0]); B[1] = conj(A[2]); B[2] = conj(A[1]); B[3] = conj(A[3]); }

/* Core of sigma2_compute on an explicit active structure: states (2k, 2k+1)
 * form active site k; U_act[state] carries U on the even (spin-up) state. */
static void sigma2_core(int nA, int nw, double delta, const double *U_act, const zc *Gr, const zc *GL, zc *SigR,
                        zc *SigL, int do_lesser, int nthreads) {
    int Nf = orc_good_fft_size(2 * nw - 1), nAs = nA / 2;
    double pp = delta / (2.0 * PI_);
    zc *Fr = zalloc((size_t)Nf * nA * nA), *FL = zalloc((size_t)Nf * nA * nA);
    fftplan pl; plan_make(&pl, Nf);
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    /* 1. forward FFT of every active element of G^r and G^< */
#pragma omp parallel
    {
        zc *scr = zalloc(Nf);
#pragma omp for collapse(2) schedule(static)
        for (int jj = 0; jj < nA; jj++)
            for (int ii = 0; ii < nA; ii++) {
                zc *dr = Fr + ((size_t)jj * nA + ii) * Nf, *dl = FL + ((size_t)jj * nA + ii) * Nf;
                for (int iw = 0; iw < nw; iw++) {
                    dr[iw] = Gr[(size_t)iw * nA * nA + IDX(ii, jj, nA)];
                    dl[iw] = GL[(size_t)iw * nA * nA + IDX(ii, jj, nA)];
                }
                for (int iw = nw; iw < Nf; iw++) { dr[iw] = 0; dl[iw] = 0; }
                fft_exec(&pl, -1, dr, scr);
                fft_exec(&pl, -1, dl, scr);
            }
        free(scr);
    }
    /* 2. per pair of active sites: 2x2 algebra at every l, inverse FFT */
#pragma omp parallel
    {
        zc *bufR = zalloc((size_t)Nf * 4), *bufL = zalloc((size_t)Nf * 4), *scr = zalloc(Nf);
#pragma omp for collapse(2) schedule(dynamic)
        for (int ja = 0; ja < nAs; ja++)
            for (int ia = 0; ia < nAs; ia++) {
                int io = 2 * ia, jo = 2 * ja;
                if (U_act[io] * U_act[jo] == 0.0) {
                    for (int iw = 0; iw < nw; iw++)
                        for (int b = 0; b < 2; b++)
                            for (int a = 0; a < 2; a++) {
                                SigR[(size_t)iw * nA * nA + IDX(io + a, jo + b, nA)] = 0;
                                SigL[(size_t)iw * nA * nA + IDX(io + a, jo + b, nA)] = 0;
                            }
                    continue;
                }
                for (int l = 0; l < Nf; l++) {
                    int lm = (Nf - l) % Nf;
                    zc r_ij[4], L_ij[4], r_ji[4], L_ji[4], a_ij[4], a_ji[4], G_ij[4], G_ji[4];
                    for (int b = 0; b < 2; b++)
                        for (int a = 0; a < 2; a++) {
                            r_ij[2 * b + a] = Fr[((size_t)(jo + b) * nA + io + a) * Nf + l];
                            L_ij[2 * b + a] = FL[((size_t)(jo + b) * nA + io + a) * Nf + l];
                            r_ji[2 * b + a] = Fr[((size_t)(io + b) * nA + jo + a) * Nf + lm];
                            L_ji[2 * b + a] = FL[((size_t)(io + b) * nA + jo + a) * Nf + lm];
                        }
                    dag2(r_ji, a_ij);
                    dag2(r_ij, a_ji);
                    for (int q = 0; q < 4; q++) { G_ij[q] = L_ij[q] + r_ij[q] - a_ij[q]; G_ji[q] = L_ji[q] + r_ji[q] - a_ji[q]; }
                    zc chi1[4], chi2[4], t1[4], t2[4], ad1[4], ad2[4], sR[4];
                    mm2(L_ji, G_ij, chi1);
                    mm2(L_ji, r_ij, t1); mm2(a_ji, L_ij, t2);
                    for (int q = 0; q < 4; q++) chi2[q] = t1[q] + t2[q];
                    adj2(chi1, ad1); adj2(chi2, ad2);
                    mm2(r_ij, ad1, t1); mm2(L_ij, ad2, t2);
                    for (int q = 0; q < 4; q++) { sR[q] = t1[q] + t2[q]; bufR[(size_t)q * Nf + l] = sR[q]; }
                    if (do_lesser) {
                        zc chiL[4], adL[4], sL[4];
                        mm2(G_ji, L_ij, chiL);
                        adj2(chiL, adL);
                        mm2(L_ij, adL, sL);
                        for (int q = 0; q < 4; q++) bufL[(size_t)q * Nf + l] = sL[q];
                    }
                }
                zc c = U_act[io] * U_act[jo] * HBAR * HBAR * pp * pp / (double)Nf;
                for (int b = 0; b < 2; b++)
                    for (int a = 0; a < 2; a++) {
                        int q = 2 * b + a;
                        fft_exec(&pl, +1, bufR + (size_t)q * Nf, scr);
                        for (int iw = 0; iw < nw; iw++) SigR[(size_t)iw * nA * nA + IDX(io + a, jo + b, nA)] = c * bufR[(size_t)q * Nf + iw];
                        if (do_lesser) {
                            fft_exec(&pl, +1, bufL + (size_t)q * Nf, scr);
                            for (int iw = 0; iw < nw; iw++) SigL[(size_t)iw * nA * nA + IDX(io + a, jo + b, nA)] = c * bufL[(size_t)q * Nf + iw];
                        }
                    }
            }
        free(bufR); free(bufL); free(scr);
    }
    free(Fr); free(FL); free(pl.tw);
}

void orc_sigma2(int nA, int n_omega, double delta, const double *U_act, const double *Gr, const double *GL,
                double *SigR, double *SigL, int do_lesser) {
    sigma2_core(nA, n_omega, delta, U_act, (const zc *)Gr, (const zc *)GL, (zc *)SigR, (zc *)SigL, do_lesser, 0);
}

static void sigma2_compute(orc_state *s) {
    int nA = s->nA;
    if (!s->SigR2) {
        s->SigR2 = zalloc((size_t)nA * nA * s->nw);
        s->SigL2 = zalloc((size_t)nA * nA * s->nw);
    }
    double *Uact = (double *)malloc(sizeof(double) * (nA ? nA : 1));
    for (int k = 0; k < nA; k++) Uact[k] = s->U[s->act[k]];
    sigma2_core(nA, s->nw, s->p.delta, Uact, s->Gin_r, s->Gin_L, s->SigR2, s->SigL2, 1, s->p.nthreads);
    free(Uact);
}

/* ------------------------------------------------------------------------ */
/* Broyden (common/broyden_mixing.f90:51-215)                                */
/* ------------------------------------------------------------------------ */
static double nrm2(int64_t n, const double *v) {
    double acc = 0;
#pragma omp parallel for reduction(+ : acc) schedule(static)
    for (int64_t i = 0; i < n; i++) acc += v[i] * v[i];
    return sqrt(acc);
}

/* dposv('U') on an nh x nh SPD matrix a (leading dim lda), rhs b; 0 on success */
static int dposv_u(int nh, double *a, int lda, double *b) {
    for (int j = 0; j < nh; j++) {
        double ajj = a[IDX(j, j, lda)];
        for (int k = 0; k < j; k++) ajj -= a[IDX(k, j, lda)] * a[IDX(k, j, lda)];
        if (ajj <= 0.0 || isnan(ajj)) return j + 1;
        ajj = sqrt(ajj);
        a[IDX(j, j, lda)] = ajj;
        for (int i = j + 1; i < nh; i++) {
            double t = a[IDX(j, i, lda)];
            for (int k = 0; k < j; k++) t -= a[IDX(k, j, lda)] * a[IDX(k, i, lda)];
            a[IDX(j, i, lda)] = t / ajj;
        }
    }
    for (int i = 0; i < nh; i++) {  /* U^T y = b */
        double t = b[i];
        for (int k = 0; k < i; k++) t -= a[IDX(k, i, lda)] * b[k];
        b[i] = t / a[IDX(i, i, lda)];
    }
    for (int i = nh - 1; i >= 0; i--) {  /* U x = y */
        double t = b[i];
        for (int k = i + 1; k < nh; k++) t -= a[IDX(i, k, lda)] * b[k];
        b[i] = t / a[IDX(i, i, lda)];
    }
    return 0;
}

void orc_broyden(orc_residual_fn fn, void *ctx, double *x, int64_t n, double alpha_mix, int nhist, int maxscf,
                 double tolf, int single_history, int32_t *converged, int32_t *n_iter, double *fmax_out,
                 double *log_fmax, double *log_rms) {
    const double w0 = 0.01;
    int sp = single_history != 0;
    int lda = nhist > 1 ? nhist : 1;
    double *F = (double *)malloc(sizeof(double) * n), *Fold = (double *)malloc(sizeof(double) * n);
    double *xold = (double *)malloc(sizeof(double) * n);
    double *a = (double *)calloc((size_t)lda * lda, sizeof(double)), *cvec = (double *)calloc(lda, sizeof(double));
    double *gram = (double *)calloc((size_t)lda * lda, sizeof(double));
    float *dFs = NULL, *Us = NULL;
    double *dF = NULL, *U = NULL;
    if (sp) { dFs = (float *)calloc((size_t)n * lda, sizeof(float)); Us = (float *)calloc((size_t)n * lda, sizeof(float)); }
    else { dF = (double *)calloc((size_t)n * (nhist ? nhist : 1), sizeof(double)); U = (double *)calloc((size_t)n * (nhist ? nhist : 1), sizeof(double)); }
    *converged = 0;
    int nstored = 0, nh = 0;
    double fmax = DBL_MAX;
    for (int its = 1; its <= maxscf; its++) {
        *n_iter = its;
        int32_t abort_flag = 0;
        fn(ctx, n, x, F, &abort_flag);
        if (abort_flag) break;
        fmax = 0;
#pragma omp parallel for reduction(max : fmax) schedule(static)
        for (int64_t i = 0; i < n; i++) if (fabs(F[i]) > fmax) fmax = fabs(F[i]);
        double rms = nrm2(n, F) / sqrt((double)n);
        if (log_fmax) log_fmax[its - 1] = fmax;
        if (log_rms) log_rms[its - 1] = rms;
        if (fmax < tolf) { *converged = 1; break; }
        if (its > 1 && nhist > 0) {
            int slot = nstored % nhist;
            if (sp) {
#pragma omp parallel for schedule(static)
                for (int64_t i = 0; i < n; i++) Fold[i] = F[i] - Fold[i];
                double fnorm = nrm2(n, Fold);
                if (fnorm > 0.0) {
                    double rf = 1.0 / fnorm;
#pragma omp parallel for
                    for (int64_t i = 0; i < n; i++) {
                        dFs[(size_t)slot * n + i] = (float)(Fold[i] * rf);
                        Us[(size_t)slot * n + i] = (float)(alpha_mix * Fold[i] * rf + (x[i] - xold[i]) * rf);
                    }
                    nstored++;
                    nh = nstored < nhist ? nstored : nhist;
                    for (int k = 0; k < nh; k++) {
                        double acc = 0;
#pragma omp parallel for reduction(+ : acc)
                        for (int64_t i = 0; i < n; i++) acc += (double)dFs[(size_t)k * n + i] * (double)dFs[(size_t)slot * n + i];
                        cvec[k] = acc;
                    }
                    for (int k = 0; k < nh; k++) { gram[IDX(k, slot, lda)] = cvec[k]; gram[IDX(slot, k, lda)] = cvec[k]; }
                }
            } else {
                double *d = dF + (size_t)slot * n;
                for (int64_t i = 0; i < n; i++) d[i] = F[i] - Fold[i];
                double fnorm = nrm2(n, d);
                if (fnorm > 0.0) {
                    double *u = U + (size_t)slot * n;
                    for (int64_t i = 0; i < n; i++) d[i] = d[i] / fnorm;
                    for (int64_t i = 0; i < n; i++) u[i] = alpha_mix * d[i] + (x[i] - xold[i]) / fnorm;
                    nstored++;
                    nh = nstored < nhist ? nstored : nhist;
                    for (int k = 0; k < nh; k++) {
                        double acc = 0;
                        for (int64_t i = 0; i < n; i++) acc += dF[(size_t)k * n + i] * d[i];
                        cvec[k] = acc;
                    }
                    for (int k = 0; k < nh; k++) { gram[IDX(k, slot, lda)] = cvec[k]; gram[IDX(slot, k, lda)] = cvec[k]; }
                }
            }
        }
        nh = nstored < nhist ? nstored : nhist;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; i++) { xold[i] = x[i]; Fold[i] = F[i]; }
        if (nh > 0) {
            for (int j = 0; j < nh; j++)
                for (int i = 0; i < nh; i++) a[IDX(i, j, lda)] = gram[IDX(i, j, lda)];
            for (int i = 0; i < nh; i++) a[IDX(i, i, lda)] += w0 * w0;
            for (int k = 0; k < nh; k++) {
                double acc = 0;
                if (sp) {
#pragma omp parallel for reduction(+ : acc)
                    for (int64_t i = 0; i < n; i++) acc += (double)dFs[(size_t)k * n + i] * F[i];
                } else {
                    for (int64_t i = 0; i < n; i++) acc += dF[(size_t)k * n + i] * F[i];
                }
                cvec[k] = acc;
            }
            if (dposv_u(nh, a, lda, cvec) != 0) { nstored = 0; nh = 0; }
        }
        if (sp) {
#pragma omp parallel for
            for (int64_t i = 0; i < n; i++) {
                double xi = x[i] + alpha_mix * F[i];
                for (int k = 0; k < nh; k++) xi = xi - (double)Us[(size_t)k * n + i] * cvec[k];
                x[i] = xi;
            }
        } else {
            for (int64_t i = 0; i < n; i++) x[i] = x[i] + alpha_mix * F[i];
            for (int k = 0; k < nh; k++)
                for (int64_t i = 0; i < n; i++) x[i] -= U[(size_t)k * n + i] * cvec[k];
        }
    }
    *fmax_out = fmax;
    free(F); free(Fold); free(xold); free(a); free(cvec); free(gram);
    free(dFs); free(Us); free(dF); free(U);
}

/* ------------------------------------------------------------------------ */
/* residuals and the SCF driver                                              */
/* ------------------------------------------------------------------------ */
static void pack_nil(const orc_state *s, const zc *G, double *x) {   /* scf.f90:226-240 */
    int n = s->norb, k = 0;
    for (int si = 0; si < s->nsites; si++) {
        if (!s->site_int[si]) continue;
        for (int b = 2 * si; b <= 2 * si + 1; b++)
            for (int a = 2 * si; a <= 2 * si + 1; a++) {
                x[k] = creal(G[IDX(a, b, n)]); x[k + 1] = cimag(G[IDX(a, b, n)]); k += 2;
            }
    }
}
static void unpack_nil(const orc_state *s, const double *x, zc *G) {   /* scf.f90:242-255 */
    int n = s->norb, k = 0;
    memset(G, 0, sizeof(zc) * n * n);
    for (int si = 0; si < s->nsites; si++) {
        if (!s->site_int[si]) continue;
        for (int b = 2 * si; b <= 2 * si + 1; b++)
            for (int a = 2 * si; a <= 2 * si + 1; a++) { G[IDX(a, b, n)] = x[k] + I * x[k + 1]; k += 2; }
    }
}

int orc_residual2(orc_state *s, double V, int64_t n, const double *x, double *F, int32_t *abort_flag) {
    s->V_scf = V;
    orc_unpack_gf(x, (double *)s->Gin_r, (double *)s->Gin_L, s->nA, s->nw);
    equal_time_lesser(s);
    sigma_first_order(s);
    sigma2_compute(s);
    s->use_sigma = 1;
    int e = sweep(s, M_SCF2, F);
    if (e) { s->err = e; *abort_flag = 1; return e; }
    int32_t ra, un, iwst; double nws, wws;
    orc_check_diag_sanity((double *)s->gRd, (double *)s->gLd, s->nA, s->nw, s->p.delta, &ra, &un, &iwst, &nws, &wws, NULL);
    *abort_flag = ra;
    if (ra) { memset(F, 0, sizeof(double) * n); return 0; }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) F[i] -= x[i];
    return 0;
}

static void residual2_cb(void *ctx, int64_t n, const double *x, double *F, int32_t *abort_flag) {
    orc_state *s = (orc_state *)ctx;
    int e = orc_residual2(s, s->V_scf, n, x, F, abort_flag);
    if (e) *abort_flag = 1;
}

static void residual1_cb(void *ctx, int64_t n, const double *x, double *F, int32_t *abort_flag) {
    orc_state *s = (orc_state *)ctx;
    unpack_nil(s, x, s->G_less_t0);
    sigma_first_order(s);
    s->use_sigma = 1;
    int e = sweep(s, M_SCF1, F);
    if (e) { s->err = e; *abort_flag = 1; return; }
    int32_t ra, un, iwst; double nws, wws;
    orc_check_diag_sanity((double *)s->gRd, (double *)s->gLd, s->nA, s->nw, s->p.delta, &ra, &un, &iwst, &nws, &wws, NULL);
    *abort_flag = ra;
    if (ra) { memset(F, 0, sizeof(double) * n); return; }
    int nn = s->norb;
    zc *Gout = zalloc((size_t)nn * nn);
    gather_onsite(s, s->site_int, Gout);
    pack_nil(s, Gout, F);
    for (int64_t i = 0; i < n; i++) F[i] -= x[i];
    free(Gout);
}

int orc_init2(orc_state *s, double V, double *x) {
    s->V_scf = V;
    s->use_sigma = 0;
    int e = sweep(s, M_INIT2, NULL);
    if (e) return e;
    if (x) orc_pack_gf((double *)s->Gin_r, (double *)s->Gin_L, x, s->nA, s->nw);
    return 0;
}

static void final_sanity(orc_state *s) {   /* scf.f90:260-275 */
    int n = s->norb;
    zc *grd = zalloc((size_t)n * s->nw), *gld = zalloc((size_t)n * s->nw);
    for (int iw = 0; iw < s->nw; iw++)
        for (int i = 0; i < n; i++) {
            grd[IDX(i, iw, n)] = I * s->dosd[IDX(i, iw, n)];
            int si = i / 2, a = i % 2;
            gld[IDX(i, iw, n)] = s->onsL[IDX(4 * si + 3 * a, iw, 4 * s->nsites)];
        }
    int32_t ra, un, iwst; double nws, wws;
    orc_check_diag_sanity((double *)grd, (double *)gld, n, s->nw, s->p.delta, &ra, &un, &iwst, &nws, &wws, NULL);
    s->runaway_final = ra; s->unphys_final = un;
    s->iworst_final = iwst; s->nworst_final = nws; s->wworst_final = wws;
    free(grd); free(gld);
}

int orc_final_sweep(orc_state *s, double V) {
    s->V_scf = V;
    s->use_sigma = (s->order >= 1);
    int e = sweep(s, M_FINAL, NULL);
    if (e) return e;
    int n = s->norb;
    memset(s->G_less_obs, 0, sizeof(zc) * n * n);
    gather_onsite(s, NULL, s->G_less_obs);
    gather_atoms(s, s->G_less_obs);
    final_sanity(s);
    return 0;
}

int orc_solve_scf(orc_state *s, double V, int32_t fresh, int32_t *converged, int32_t *n_iter, double *fmax,
                  double *log_fmax, double *log_rms) {
    int n = s->norb;
    s->V_scf = V;
    s->err = 0;
    *converged = 1; *n_iter = 0; *fmax = 0;
    if (s->order == 1) {
        if (fresh) {
            memset(s->G_less_t0, 0, sizeof(zc) * n * n);
            s->use_sigma = 0;
            int e = sweep(s, M_SCF1, NULL);
            if (e) return e;
            gather_onsite(s, s->site_int, s->G_less_t0);
        }
        double *x = (double *)malloc(sizeof(double) * (s->n_vec ? s->n_vec : 1));
        pack_nil(s, s->G_less_t0, x);
        orc_broyden(residual1_cb, s, x, s->n_vec, s->p.alpha_mix, s->p.nhist, s->p.maxscf, s->p.tolf_order1,
                    s->p.single_history, converged, n_iter, fmax, log_fmax, log_rms);
        free(x);
    } else if (s->order == 2) {
        if (fresh) {
            int e = orc_init2(s, V, NULL);
            if (e) return e;
        }
        double *x = (double *)malloc(sizeof(double) * s->n_vec);
        orc_pack_gf((double *)s->Gin_r, (double *)s->Gin_L, x, s->nA, s->nw);
        orc_broyden(residual2_cb, s, x, s->n_vec, s->p.alpha_mix, s->p.nhist, s->p.maxscf, s->p.tolf,
                    s->p.single_history, converged, n_iter, fmax, log_fmax, log_rms);
        free(x);
    } else {
        *n_iter = 1;
        if (log_fmax) log_fmax[0] = 0;
        if (log_rms) log_rms[0] = 0;
    }
    if (s->err) return s->err;
    int e = orc_final_sweep(s, V);
    if (e) return e;
    if (*converged && s->unphys_final) *converged = 0;
    return 0;
}

void orc_current(const orc_state *s, double *J_up, double *J_down) {
    double u = 0, d = 0;
    for (int iw = 0; iw < s->nw; iw++) { u += s->trU[iw]; d += s->trD[iw]; }
    *J_up = u * (s->p.delta / HBAR);
    *J_down = d * (s->p.delta / HBAR);
}
void orc_g_less_obs(const orc_state *s, double *G) { memcpy(G, s->G_less_obs, sizeof(zc) * s->norb * s->norb); }
void orc_dosd(const orc_state *s, double *d) { memcpy(d, s->dosd, sizeof(double) * s->norb * s->nw); }
void orc_trUD(const orc_state *s, double *u, double *d) {
    memcpy(u, s->trU, sizeof(double) * s->nw);
    memcpy(d, s->trD, sizeof(double) * s->nw);
}
void orc_final_flags(const orc_state *s, int32_t *runaway, int32_t *unphysical) {
    *runaway = s->runaway_final; *unphysical = s->unphys_final;
}
/* worst state of final_sanity (1-based label), its occupation and spectral weight */
void orc_final_worst(const orc_state *s, int32_t *i_worst, double *n_worst, double *w_worst) {
    *i_worst = s->iworst_final; *n_worst = s->nworst_final; *w_worst = s->wworst_final;
}
void orc_get_gin(const orc_state *s, double *Gr, double *GL) {
    size_t m = (size_t)s->nA * s->nA * s->nw;
    if (s->Gin_r) { memcpy(Gr, s->Gin_r, sizeof(zc) * m); memcpy(GL, s->Gin_L, sizeof(zc) * m); }
}
void orc_get_sigma2(const orc_state *s, double *SigR, double *SigL) {
    size_t m = (size_t)s->nA * s->nA * s->nw;
    if (s->SigR2) { memcpy(SigR, s->SigR2, sizeof(zc) * m); memcpy(SigL, s->SigL2, sizeof(zc) * m); }
}
void orc_get_sig1(const orc_state *s, double *Sig1) { memcpy(Sig1, s->Sig1, sizeof(zc) * s->norb * s->norb); }
void orc_get_diag(const orc_state *s, double *gRd, double *gLd) {
    if (!s->gRd) return;
    memcpy(gRd, s->gRd, sizeof(zc) * s->nA * s->nw);
    memcpy(gLd, s->gLd, sizeof(zc) * s->nA * s->nw);
}

/* ------------------------------------------------------------------------ */
/* bounded samples of one order-2 SCF iteration, timed on the host cores      */
/* (CPU baseline of the benchmark; each restates the reference's loop)        */
/* ------------------------------------------------------------------------ */
static double wall(void) {
#ifdef _OPENMP
    return omp_get_wtime();
#else
    return 0.0;
#endif
}

/* sweep(M_SCF2) with Sigma present (src/scf.f90:280-355) over n_omega = p->n_omega
 * energies; returns seconds */
double orc_bench_dyson(const orc_problem *p) {
    orc_state *s = orc_create(p);
    if (!s || s->order != 2) { orc_destroy(s); return -1.0; }
    size_t m = (size_t)s->nA * s->nA * s->nw;
    s->SigR2 = zalloc(m);
    s->SigL2 = zalloc(m);
    double *F = (double *)calloc((size_t)s->n_vec, sizeof(double));
    s->use_sigma = 1;
    double t0 = wall();
    sweep(s, M_SCF2, F);
    double t = wall() - t0;
    free(F);
    orc_destroy(s);
    return t;
}

/* sigma2_compute pieces (src/self_energies.f90:148-216) at FFT length
 * good_fft_size(2 n_omega - 1): n_cols zero-padded forward transforms and
 * n_pairs site pairs (2x2 algebra at every l + 8 backward transforms); seconds */
double orc_bench_sigma2(int n_omega, int n_cols, int n_pairs, int nthreads) {
    int Nf = orc_good_fft_size(2 * n_omega - 1);
    fftplan pl; plan_make(&pl, Nf);
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    zc *cols = zalloc((size_t)Nf * 16);
    for (int q = 0; q < 16; q++)
        for (int l = 0; l < Nf; l++) cols[(size_t)q * Nf + l] = cos(0.001 * l + q) + I * sin(0.002 * l * q);
    double t0 = wall();
#pragma omp parallel
    {
        zc *buf = zalloc(Nf), *scr = zalloc(Nf);
#pragma omp for schedule(static)
        for (int c = 0; c < n_cols; c++) {
            for (int iw = 0; iw < n_omega; iw++) buf[iw] = cols[(size_t)(c % 16) * Nf + iw];
            for (int iw = n_omega; iw < Nf; iw++) buf[iw] = 0;
            fft_exec(&pl, -1, buf, scr);
        }
        zc *bufR = zalloc((size_t)Nf * 4), *bufL = zalloc((size_t)Nf * 4);
#pragma omp for schedule(dynamic)
        for (int pr = 0; pr < n_pairs; pr++) {
            for (int l = 0; l < Nf; l++) {
                int lm = (Nf - l) % Nf;
                zc r_ij[4], L_ij[4], r_ji[4], L_ji[4], a_ij[4], a_ji[4], G_ij[4], G_ji[4];
                for (int q = 0; q < 4; q++) {
                    r_ij[q] = cols[(size_t)q * Nf + l]; L_ij[q] = cols[(size_t)(4 + q) * Nf + l];
                    r_ji[q] = cols[(size_t)(8 + q) * Nf + lm]; L_ji[q] = cols[(size_t)(12 + q) * Nf + lm];
                }
                dag2(r_ji, a_ij); dag2(r_ij, a_ji);
                for (int q = 0; q < 4; q++) { G_ij[q] = L_ij[q] + r_ij[q] - a_ij[q]; G_ji[q] = L_ji[q] + r_ji[q] - a_ji[q]; }
                zc chi1[4], chi2[4], t1[4], t2[4], ad1[4], ad2[4], chiL[4], adL[4], sL[4];
                mm2(L_ji, G_ij, chi1); mm2(L_ji, r_ij, t1); mm2(a_ji, L_ij, t2);
                for (int q = 0; q < 4; q++) chi2[q] = t1[q] + t2[q];
                adj2(chi1, ad1); adj2(chi2, ad2);
                mm2(r_ij, ad1, t1); mm2(L_ij, ad2, t2);
                for (int q = 0; q < 4; q++) bufR[(size_t)q * Nf + l] = t1[q] + t2[q];
                mm2(G_ji, L_ij, chiL); adj2(chiL, adL); mm2(L_ij, adL, sL);
                for (int q = 0; q < 4; q++) bufL[(size_t)q * Nf + l] = sL[q];
            }
            for (int q = 0; q < 4; q++) {
                fft_exec(&pl, +1, bufR + (size_t)q * Nf, scr);
                fft_exec(&pl, +1, bufL + (size_t)q * Nf, scr);
            }
        }
        free(buf); free(scr); free(bufR); free(bufL);
    }
    double t = wall() - t0;
    free(cols); free(pl.tw);
    return t;
}

/* one steady-state step of broyden_solve with an FP32 history of nhist columns
 * (common/broyden_mixing.f90:102-178) plus unpack_gf of the vector, on n
 * elements; seconds */
double orc_bench_broyden(int64_t n, int nhist, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    double *x = (double *)malloc(sizeof(double) * n), *F = (double *)malloc(sizeof(double) * n);
    double *Fold = (double *)malloc(sizeof(double) * n), *xold = (double *)malloc(sizeof(double) * n);
    float *dFs = (float *)malloc(sizeof(float) * n * nhist), *Us = (float *)malloc(sizeof(float) * n * nhist);
#pragma omp parallel for
    for (int64_t i = 0; i < n; i++) {
        x[i] = 1e-3 * (double)(i % 1000); F[i] = 1e-6 * (double)(i % 77); Fold[i] = 0.5 * F[i]; xold[i] = 0.9 * x[i];
        for (int k = 0; k < nhist; k++) { dFs[(size_t)k * n + i] = 1e-3f * (float)k; Us[(size_t)k * n + i] = 1e-4f * (float)k; }
    }
    double cvec[64];
    double t0 = wall();
    /* unpack (a pure copy pass over x into two complex arrays of the same size) */
    double *Gc = (double *)malloc(sizeof(double) * n);
#pragma omp parallel for
    for (int64_t i = 0; i < n; i++) Gc[i] = x[i];
    double fmax = 0, s2 = 0;
#pragma omp parallel for reduction(max : fmax) reduction(+ : s2)
    for (int64_t i = 0; i < n; i++) { double a = fabs(F[i]); if (a > fmax) fmax = a; s2 += F[i] * F[i]; }
    double fn = 0;
#pragma omp parallel for reduction(+ : fn)
    for (int64_t i = 0; i < n; i++) { Fold[i] = F[i] - Fold[i]; fn += Fold[i] * Fold[i]; }
    double rf = 1.0 / sqrt(fn + 1e-300);
    int slot = 0;
#pragma omp parallel for
    for (int64_t i = 0; i < n; i++) {
        dFs[(size_t)slot * n + i] = (float)(Fold[i] * rf);
        Us[(size_t)slot * n + i] = (float)(0.5 * Fold[i] * rf + (x[i] - xold[i]) * rf);
    }
    for (int k = 0; k < nhist; k++) {
        double acc = 0;
#pragma omp parallel for reduction(+ : acc)
        for (int64_t i = 0; i < n; i++) acc += (double)dFs[(size_t)k * n + i] * (double)dFs[(size_t)slot * n + i];
        cvec[k] = acc;
    }
    memcpy(xold, x, sizeof(double) * n);
    memcpy(Fold, F, sizeof(double) * n);
    for (int k = 0; k < nhist; k++) {
        double acc = 0;
#pragma omp parallel for reduction(+ : acc)
        for (int64_t i = 0; i < n; i++) acc += (double)dFs[(size_t)k * n + i] * F[i];
        cvec[k] = acc * 1e-9;
    }
#pragma omp parallel for
    for (int64_t i = 0; i < n; i++) {
        double xi = x[i] + 0.5 * F[i];
        for (int k = 0; k < nhist; k++) xi = xi - (double)Us[(size_t)k * n + i] * cvec[k];
        x[i] = xi;
    }
    double t = wall() - t0;
    free(x); free(F); free(Fold); free(xold); free(dFs); free(Us); free(Gc);
    (void)fmax; (void)s2;
    return t;
}
