/*
 * sancho_oracle.c -- CPU FP64 restatement of SANCHO's interaction self-energies
 * (orders 2 and 3, the factorised bubble form).
 *
 * TEST INFRASTRUCTURE ONLY, like the rest of oracle/: only tests/ and the
 * bench's CPU baseline load it; the product never does.
 *
 * Follows, loop for loop:
 *   complete_input_gf   SANCHO/src/green_functions.f90:163-173
 *   compute_bubbles     SANCHO/src/self_energies.f90:49-113
 *   second_order_sigma  SANCHO/src/self_energies.f90:115-147
 *   third_order_sigma   SANCHO/src/self_energies.f90:149-207
 * Arrays keep the reference's index order, 0-based: G (na,na,nw) column-major;
 * the frequency-first copies g(k,a,b); chi(m,a,b) with m = -(nw-1)..nw-1 stored
 * at m + nw - 1; phi(n,a,b) with n = 2..2nw stored at n - 2.  Both bubble axes
 * have length L = 2 nw - 1.
 */
#include <complex.h>
#include <stdlib.h>
#include <string.h>

#include "qs_oracle.h"

typedef double _Complex cplx;

#define G3(A, k, a, b) (A)[(k) + (size_t)nw * ((a) + (size_t)na * (b))]      /* g(k,a,b) */
#define B3(A, m, a, b) (A)[(m) + (size_t)L * ((a) + (size_t)na * (b))]       /* chi/phi/S/P(m,a,b) */
#define M3(A, a, b, k) (A)[(a) + (size_t)na * ((b) + (size_t)na * (k))]      /* G_in(a,b,k) */

static int imax(int a, int b) { return a > b ? a : b; }
static int imin(int a, int b) { return a < b ? a : b; }

/* frequency-first copies gR, gA, gL, gG (self_energies.f90:66-74) of the completed
 * input: G^a = (G^r)^dagger, G^> = G^< + G^r - G^a (green_functions.f90:163-173) */
static void sn_copies(int na, int nw, const cplx *GR, const cplx *GL, cplx *gR, cplx *gA, cplx *gL, cplx *gG) {
#pragma omp parallel for collapse(2)
    for (int b = 0; b < na; b++)
        for (int a = 0; a < na; a++)
            for (int k = 0; k < nw; k++) {
                const cplx r = M3(GR, a, b, k), l = M3(GL, a, b, k), adv = conj(M3(GR, b, a, k));
                G3(gR, k, a, b) = r;
                G3(gA, k, a, b) = adv;
                G3(gL, k, a, b) = l;
                G3(gG, k, a, b) = l + r - adv;
            }
}

/* particle-hole bubbles chi(m,a,b) (self_energies.f90:77-93) and, if phi is not
 * NULL, particle-particle bubbles phi(n,a,b) (:97-111); 1-based k1, k2 there */
static void sn_bubbles(int na, int nw, double pp, const cplx *gR, const cplx *gA, const cplx *gL, const cplx *gG,
                       cplx *chi[4], cplx *phi[4]) {
    const int L = 2 * nw - 1;
#pragma omp parallel for collapse(3)
    for (int b = 0; b < na; b++)
        for (int a = 0; a < na; a++)
            for (int mi = 0; mi < L; mi++) {
                const int m = mi - (nw - 1);
                cplx s1 = 0, s2 = 0, s3 = 0, s4 = 0;
                for (int k1 = imax(1, 1 + m); k1 <= imin(nw, nw + m); k1++) {
                    const int x = k1 - 1, y = k1 - m - 1;
                    s1 += G3(gG, x, a, b) * G3(gL, y, b, a);
                    s2 += G3(gL, x, a, b) * G3(gG, y, b, a);
                    s3 += G3(gR, x, a, b) * G3(gL, y, b, a) + G3(gL, x, a, b) * G3(gA, y, b, a);
                    s4 += G3(gA, x, a, b) * G3(gL, y, b, a) + G3(gL, x, a, b) * G3(gR, y, b, a);
                }
                /* chi order: G, L, R, A */
                B3(chi[0], mi, a, b) = s1 * pp;
                B3(chi[1], mi, a, b) = s2 * pp;
                B3(chi[2], mi, a, b) = s3 * pp;
                B3(chi[3], mi, a, b) = s4 * pp;
            }
    if (!phi) return;
#pragma omp parallel for collapse(3)
    for (int b = 0; b < na; b++)
        for (int a = 0; a < na; a++)
            for (int ni = 0; ni < L; ni++) {
                const int n = ni + 2;
                cplx s1 = 0, s2 = 0, s3 = 0, s4 = 0;
                for (int k1 = imax(1, n - nw); k1 <= imin(nw, n - 1); k1++) {
                    const int x = k1 - 1, y = n - k1 - 1;
                    s1 += G3(gG, x, a, b) * G3(gG, y, a, b);
                    s2 += G3(gL, x, a, b) * G3(gL, y, a, b);
                    s3 += G3(gR, x, a, b) * G3(gG, y, a, b) + G3(gL, x, a, b) * G3(gR, y, a, b);
                    s4 += G3(gA, x, a, b) * G3(gG, y, a, b) + G3(gL, x, a, b) * G3(gA, y, a, b);
                }
                B3(phi[0], ni, a, b) = s1 * pp;
                B3(phi[1], ni, a, b) = s2 * pp;
                B3(phi[2], ni, a, b) = s3 * pp;
                B3(phi[3], ni, a, b) = s4 * pp;
            }
}

int orc_sancho_bubbles(int32_t na, int32_t nw, double delta, int32_t order, const double *GRd, const double *GLd,
                       double *chi_out, double *phi_out) {
    if (na < 1 || nw < 1 || (order != 2 && order != 3) || !GRd || !GLd || !chi_out) return -1;
    if (order == 3 && !phi_out) return -1;
    const size_t nG = (size_t)na * na * nw, nB = (size_t)na * na * (2 * nw - 1);
    cplx *g = malloc(4 * nG * sizeof(cplx));
    if (!g) return -1;
    cplx *gR = g, *gA = g + nG, *gL = g + 2 * nG, *gG = g + 3 * nG;
    sn_copies(na, nw, (const cplx *)GRd, (const cplx *)GLd, gR, gA, gL, gG);
    cplx *chi[4], *phi[4];
    for (int q = 0; q < 4; q++) {
        chi[q] = (cplx *)chi_out + q * nB;
        phi[q] = order == 3 ? (cplx *)phi_out + q * nB : NULL;
    }
    sn_bubbles(na, nw, delta / (2.0 * 3.14159265358979323846), gR, gA, gL, gG, chi, order == 3 ? phi : NULL);
    free(g);
    return 0;
}

int orc_sancho_sigma(int32_t na, int32_t nw, double delta, const double *U, int32_t order, const double *GRd,
                     const double *GLd, double *S2R, double *S2L, double *S2G, double *S3R, double *S3L,
                     double *S3G) {
    if (na < 1 || nw < 1 || (order != 2 && order != 3) || !GRd || !GLd || !U || !S2R || !S2L || !S2G) return -1;
    if (order == 3 && (!S3R || !S3L || !S3G)) return -1;
    const int L = 2 * nw - 1;
    const size_t nG = (size_t)na * na * nw, nB = (size_t)na * na * L;
    const double pp = delta / (2.0 * 3.14159265358979323846);
    const size_t nbub = order == 3 ? 14 : 4;  /* chi x4 (+ phi x4, S x3, P x3) */
    cplx *g = malloc(4 * nG * sizeof(cplx)), *w = malloc(nbub * nB * sizeof(cplx));
    if (!g || !w) { free(g); free(w); return -1; }
    cplx *gR = g, *gA = g + nG, *gL = g + 2 * nG, *gG = g + 3 * nG;
    cplx *chi[4], *phi[4];
    for (int q = 0; q < 4; q++) chi[q] = w + q * nB, phi[q] = order == 3 ? w + (4 + q) * nB : NULL;
    sn_copies(na, nw, (const cplx *)GRd, (const cplx *)GLd, gR, gA, gL, gG);
    sn_bubbles(na, nw, pp, gR, gA, gL, gG, chi, order == 3 ? phi : NULL);
    cplx *chiG = chi[0], *chiL = chi[1], *chiR = chi[2], *chiA = chi[3];
    const double hbar = 1.0;  /* natural units, SANCHO/common/constants.f90:18 */

    /* second order (self_energies.f90:128-145); c2 = -(i hbar)^2 */
    const cplx c2 = -((I * hbar) * (I * hbar));
    cplx *o2R = (cplx *)S2R, *o2L = (cplx *)S2L, *o2G = (cplx *)S2G;
#pragma omp parallel for collapse(3)
    for (int j = 0; j < na; j++)
        for (int i = 0; i < na; i++)
            for (int iw = 1; iw <= nw; iw++) {
                cplx aR = 0, aL = 0, aG = 0;
                for (int m = imax(iw - nw, -(nw - 1)); m <= imin(iw - 1, nw - 1); m++) {
                    const int kk = iw - m - 1, mi = m + nw - 1;
                    aL += G3(gL, kk, i, j) * B3(chiL, mi, i, j);
                    aG += G3(gG, kk, i, j) * B3(chiG, mi, i, j);
                    aR += G3(gR, kk, i, j) * B3(chiG, mi, i, j) + G3(gL, kk, i, j) * B3(chiR, mi, i, j);
                }
                const cplx s = c2 * U[i] * U[j] * pp;
                M3(o2R, i, j, iw - 1) = s * aR;
                M3(o2L, i, j, iw - 1) = s * aL;
                M3(o2G, i, j, iw - 1) = s * aG;
            }
    if (order == 3) {
        cplx *phiG = phi[0], *phiL = phi[1], *phiR = phi[2], *phiA = phi[3];
        cplx *SG3 = w + 8 * nB, *SL3 = w + 9 * nB, *SR3 = w + 10 * nB;
        cplx *PG3 = w + 11 * nB, *PL3 = w + 12 * nB, *PR3 = w + 13 * nB;
        /* 2. sums over the internal site k (self_energies.f90:160-175) */
#pragma omp parallel for collapse(2)
        for (int j = 0; j < na; j++)
            for (int i = 0; i < na; i++)
                for (int mi = 0; mi < L; mi++) {
                    cplx sg = 0, sl = 0, sr = 0, pg = 0, pl = 0, pr = 0;
                    for (int k = 0; k < na; k++) {
                        sg += U[k] * (B3(chiR, mi, i, k) * B3(chiG, mi, k, j) + B3(chiG, mi, i, k) * B3(chiA, mi, k, j));
                        sl += U[k] * (B3(chiR, mi, i, k) * B3(chiL, mi, k, j) + B3(chiL, mi, i, k) * B3(chiA, mi, k, j));
                        sr += U[k] * B3(chiR, mi, i, k) * B3(chiR, mi, k, j);
                        pg += U[k] * (B3(phiR, mi, i, k) * B3(phiG, mi, k, j) + B3(phiG, mi, i, k) * B3(phiA, mi, k, j));
                        pl += U[k] * (B3(phiR, mi, i, k) * B3(phiL, mi, k, j) + B3(phiL, mi, i, k) * B3(phiA, mi, k, j));
                        pr += U[k] * B3(phiR, mi, i, k) * B3(phiR, mi, k, j);
                    }
                    B3(SG3, mi, i, j) = sg; B3(SL3, mi, i, j) = sl; B3(SR3, mi, i, j) = sr;
                    B3(PG3, mi, i, j) = pg; B3(PL3, mi, i, j) = pl; B3(PR3, mi, i, j) = pr;
                }
        /* 3. convolution with the remaining G (self_energies.f90:178-204); c3 = -(i hbar)^3 */
        const cplx c3 = -((I * hbar) * (I * hbar) * (I * hbar));
        cplx *o3R = (cplx *)S3R, *o3L = (cplx *)S3L, *o3G = (cplx *)S3G;
#pragma omp parallel for collapse(3)
        for (int j = 0; j < na; j++)
            for (int i = 0; i < na; i++)
                for (int iw = 1; iw <= nw; iw++) {
                    cplx aR = 0, aL = 0, aG = 0;
                    for (int m = imax(iw - nw, -(nw - 1)); m <= imin(iw - 1, nw - 1); m++) {
                        const int kk = iw - m - 1, mi = m + nw - 1;
                        aL += G3(gL, kk, i, j) * B3(SL3, mi, i, j);
                        aG += G3(gG, kk, i, j) * B3(SG3, mi, i, j);
                        aR += G3(gR, kk, i, j) * B3(SG3, mi, i, j) + G3(gL, kk, i, j) * B3(SR3, mi, i, j);
                    }
                    for (int n = imax(2, iw + 1); n <= imin(2 * nw, iw + nw); n++) {
                        const int kk = n - iw - 1, ni = n - 2;
                        aL += G3(gG, kk, j, i) * B3(PL3, ni, i, j);
                        aG += G3(gL, kk, j, i) * B3(PG3, ni, i, j);
                        aR += G3(gA, kk, j, i) * B3(PG3, ni, i, j) + G3(gG, kk, j, i) * B3(PR3, ni, i, j);
                    }
                    const cplx s = c3 * U[i] * U[j] * pp;
                    M3(o3R, i, j, iw - 1) = s * aR;
                    M3(o3L, i, j, iw - 1) = s * aL;
                    M3(o3G, i, j, iw - 1) = s * aG;
                }
    }
    free(g);
    free(w);
    return 0;
}
