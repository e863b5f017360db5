// qs_dyson_wave.hip -- per-energy Dyson and Keldysh solve with a small LDS footprint.
//
// Same contract as k_dyson (qs_dyson.hip; src/scf.f90:305-415,
// common/linear_algebra.f90:58-88):
//
//   G^r = [ w + Hemb - Sigma1 - Sigma2^r(w) ]^-1
//   G^< = G^r(:,sup) [ i(f_L GamL + f_R GamR) + Sigma2^<(w) ] G^r(:,sup)^dagger
//
// A workgroup of NT = TI TJ threads (8 x 8: one wavefront, 16 x 16: four)
// solves one energy.  Thread (li, lj) = (tid / TJ, tid % TJ) keeps the RA x RB
// elements (li + TI a, lj + TJ b) of the matrix in registers (norb <= TI RA,
// norb <= TJ RB).  The workgroup keeps only TI-row stages in LDS (about 60 KB at
// norb = 80 against 125 KB for k_dyson), so two energies share a CU and each
// hides the other's barrier and memory latency; M, the operand of both Keldysh
// products, goes through a per-energy scratch slab in HBM.
//
// Steps:
//  1. A = w + Hemb - Sigma1 - Sigma2^r in registers.
//  2. Gauss-Jordan with zgetf2's pivot rule, rows never moved (as k_dyson):
//     the registers end with M, G^r[q_r][p_c] = M[r][c].  One barrier per
//     pivot: after the update of step k each wavefront posts its best row for
//     column k+1 together with that whole row and column k+1 (double-buffered
//     LDS); after the barrier every thread picks the pivot and has its row and
//     column.
//  3. M goes to the slab column-major (Ms): the operand of the Keldysh products,
//     and G^r for the packing through the pivot maps, G^r(i, j) = M(piv[i], qinv[j]).
//  4. T = M S'' with S''[c][c'] = SL[p_c][p_c'] as outer products over c, TI
//     values of c at a time: one round of loads brings TI columns of M (from
//     Ms) and builds TI rows of S'' into two LDS chunks, then TI rank-1
//     updates read them from LDS.  T stays in registers.
//  5. P = T M^dagger the same way: per round, TI columns of T go from registers
//     (they are one register column of every thread) and TI columns of M from
//     Ms into the chunks; G^<[q_r][q_r'] = P[r][r'].
//  6. Epilogues by mode from TI-row LDS stages of G^< and from Ms, with the
//     loads of x and Fold batched ahead of the stores.
#include "qs_device.h"
#include "qs_internal.h"

namespace {

constexpr double TWO_PI = 6.283185307179586476925286766559;
// packing elements per thread in flight: G^r, and G^< (P is live then).  The
// larger batches only where occupancy is two waves per SIMD anyway (25 elements
// per thread); smaller tiles keep their registers for occupancy.
template <int RA, int RB> constexpr int eb_gr() { return RA * RB >= 25 ? 13 : 8; }
template <int RA, int RB> constexpr int eb_gl() { return RA * RB >= 25 ? 5 : 4; }

// packed G^< positions (common/linear_algebra.f90:142-191): Re part over a < b
// row-wise, Im part over a <= b row-wise, after the 2 nA^2 G^r values
__device__ __forceinline__ int su_index(int a, int b, int nA) { return a * (nA - 1) - (a * (a - 1)) / 2 + (b - a - 1); }
__device__ __forceinline__ int u_index(int a, int b, int nA) { return a * nA - (a * (a - 1)) / 2 + (b - a); }

struct WLayout {
    double2 *stage, *crow, *pcol, *srow, *grcl, *glcl, *chunk;
    double2 *cand;  // [2][NW] posted pivot candidates: (value, row as a double)
    double *red;
    int *piv, *qinv, *spc, *actl, *posl;
};

// LDS: stage [SC] | crow [2][NW][TJ RB] | pcol [2][TI RA] | srow [2][TJ RB] |
//      grcl, glcl [nCL^2] | chunk [SC] | cand [2 NW] | red [3 NW] | 5 int maps [TI RA]
// SC = TI TJ RA elements: one store per thread and load batch q < RA, which
// covers TI rows of n (RA = ceil(n / TI), TJ >= TI)
__host__ __device__ inline int wlayout_sc(int n, int RA, int TI, int TJ) {
    return TI * TJ * RA > TI * n ? TI * TJ * RA : TI * n;
}
__host__ __device__ inline size_t wlayout_bytes(int n, int nCL, int RA, int RB, int TI, int TJ) {
    const int NW = TI * TJ / 64, SC = wlayout_sc(n, RA, TI, TJ);
    return (size_t)16 * (2 * SC + 2 * NW * TJ * RB + 2 * TI * RA + 2 * TJ * RB + 2 * nCL * nCL + 2 * NW) +
           (size_t)8 * 3 * NW + (size_t)4 * 5 * TI * RA;
}
__device__ __forceinline__ WLayout wlayout(char *base, int n, int nCL, int RA, int RB, int TI, int TJ) {
    const int NW = TI * TJ / 64, SC = wlayout_sc(n, RA, TI, TJ);
    WLayout L;
    L.stage = (double2 *)base;
    L.crow = L.stage + SC;
    L.pcol = L.crow + 2 * NW * TJ * RB;
    L.srow = L.pcol + 2 * TI * RA;
    L.grcl = L.srow + 2 * TJ * RB;
    L.glcl = L.grcl + nCL * nCL;
    L.chunk = L.glcl + nCL * nCL;
    L.cand = L.chunk + SC;
    L.red = (double *)(L.cand + 2 * NW);
    L.piv = (int *)(L.red + 3 * NW);
    L.qinv = L.piv + TI * RA;
    L.spc = L.qinv + TI * RA;
    L.actl = L.spc + TI * RA;
    L.posl = L.actl + TI * RA;
    return L;
}

// The Gauss-Jordan part alone (PART 1 below) needs only the pivot-search buffers
// and the maps: crow [2][NW][TJ RB] | pcol [2][TI RA] | cand [2 NW] | red [3 NW] |
// 5 int maps [TI RA] (14.6 KB at 80 orbitals, against 58 KB for the whole kernel).
__host__ __device__ inline size_t wlayout_gj_bytes(int RA, int RB, int TI, int TJ) {
    const int NW = TI * TJ / 64;
    return (size_t)16 * (2 * NW * TJ * RB + 2 * TI * RA + 2 * NW) + (size_t)8 * 3 * NW + (size_t)4 * 5 * TI * RA;
}
__device__ __forceinline__ WLayout wlayout_gj(char *base, int RA, int RB, int TI, int TJ) {
    const int NW = TI * TJ / 64;
    WLayout L;
    L.stage = L.srow = L.grcl = L.glcl = L.chunk = nullptr;
    L.crow = (double2 *)base;
    L.pcol = L.crow + 2 * NW * TJ * RB;
    L.cand = L.pcol + 2 * TI * RA;
    L.red = (double *)(L.cand + 2 * NW);
    L.piv = (int *)(L.red + 3 * NW);
    L.qinv = L.piv + TI * RA;
    L.spc = L.qinv + TI * RA;
    L.actl = L.spc + TI * RA;
    L.posl = L.actl + TI * RA;
    return L;
}
// per energy, the slab holds M (n^2 complex, column-major) and then the pivot
// maps piv, qinv (2 n ints) that pass from PART 1 to PART 2
__host__ __device__ inline size_t wave_slab_stride(int n) { return (size_t)n * n + (2 * (size_t)n * 4 + 15) / 16; }
// waves per SIMD asked of the Gauss-Jordan part: three energies per CU where the
// register tile allows it without spilling inside the pivot loop
template <int RA> constexpr int gj_waves_per_simd() { return RA <= 5 ? 3 : 1; }

// One energy's packing: its slices of the packed vectors (offsets below are 32-bit,
// 3 nA^2 < 2^31) and the SCF2 sums of this thread.
struct Epi {
    const double *xe, *oe;
    double *fe, *ye;
    int nsu;
    bool scf2, init2;
    double st_max, st_s2, st_d2;
};

__device__ __forceinline__ Epi epi_begin(const DysonArgs &A, int iw) {
    const int nA = A.nA, nA2 = nA * nA;
    const size_t off = (size_t)iw * 3 * (size_t)nA2;
    Epi E;
    E.xe = A.x + off;
    E.oe = A.Fold + off;
    E.fe = A.F + off;
    E.ye = A.xout + off;
    E.nsu = (nA2 - nA) / 2;
    E.scf2 = A.mode == QS_M_SCF2;
    E.init2 = A.mode == QS_M_INIT2;
    E.st_max = E.st_s2 = E.st_d2 = 0.0;
    return E;
}

// Rows r0 .. r0 + rows - 1 (pivoted order) of G^<, staged in L.stage (row s at
// s n + natural column): the packed G^< entries (INIT2, SCF2), or the diagonal,
// on-site and atom blocks and the contact block (SCF1, FINAL).
template <int NT, int EBL>
__device__ __forceinline__ void epi_rows(const DysonArgs &A, const WLayout &L, Epi &E, int r0, int rows, int iw,
                                         int tid) {
    const int n = A.norb, nA = A.nA, nA2 = nA * nA, nCL = A.nCL;
    const double *xe = E.xe, *oe = E.oe;
    double *fe = E.fe, *ye = E.ye;
    const int nsu = E.nsu;
    const bool scf2 = E.scf2, init2 = E.init2;
    double st_max = E.st_max, st_s2 = E.st_s2, st_d2 = E.st_d2;
    if (init2 || scf2) {
        // element e of the stage: row s = e / nA, active column index bb = e % nA
        for (int e0 = tid; e0 < rows * nA; e0 += NT * EBL) {
            double xv[EBL][2], ov[EBL][2];
            int dst[EBL][2];
#pragma unroll
            for (int u = 0; u < EBL; u++) {
                const int e = e0 + NT * u, s = e / nA, bb = e - s * nA;
                dst[u][0] = dst[u][1] = -1;
                if (e < rows * nA) {
                    const int ar = L.posl[L.qinv[r0 + s]];
                    if (ar >= 0) {
                        if (ar < bb) dst[u][0] = su_index(ar, bb, nA);
                        if (ar <= bb) dst[u][1] = nsu + u_index(ar, bb, nA);
                    }
                }
                if (scf2) {
#pragma unroll
                    for (int h = 0; h < 2; h++) {
                        const int d = 2 * nA2 + (dst[u][h] >= 0 ? dst[u][h] : 0);
                        xv[u][h] = xe[d];
                        ov[u][h] = oe[d];
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < EBL; u++) {
                const int e = e0 + NT * u, s = e / nA, bb = e - s * nA;
                if (e < rows * nA && dst[u][1] >= 0) {
                    const double2 g = L.stage[s * n + L.actl[bb]];
#pragma unroll
                    for (int h = 0; h < 2; h++) {
                        if (dst[u][h] < 0) continue;
                        const double val = h == 0 ? g.x : g.y;
                        const int d = 2 * nA2 + dst[u][h];
                        if (init2) {
                            ye[d] = val;
                        } else {
                            const double f = val - xv[u][h];
                            fe[d] = f;
                            st_max = fmax(st_max, fabs(f));
                            st_s2 += f * f;
                            const double e1 = f - ov[u][h];
                            st_d2 += e1 * e1;
                        }
                    }
                    const int ar = L.posl[L.qinv[r0 + s]];
                    if (scf2 && ar == bb) A.diagL[(size_t)iw * nA + ar] = g.y;
                }
            }
        }
    } else {
        for (int s = 0; s < rows; s++) {
            const int i = L.qinv[r0 + s];
            if (A.mode == QS_M_SCF1 && tid == 0 && L.posl[i] >= 0)
                A.diagL[(size_t)iw * nA + L.posl[i]] = L.stage[s * n + i].y;
            // on-site 2x2 blocks: element (2 si + a2, 2 si + b2) is stored at 4 si + a2 + 2 b2
            if (tid < 2 && (i >> 1) < A.nsites)
                A.onsL[(size_t)iw * 4 * A.nsites + 4 * (i >> 1) + (i & 1) + 2 * tid] = L.stage[s * n + 2 * (i >> 1) + tid];
            if (A.mode == QS_M_FINAL) {
                if (A.atomL) {
                    // atom block of row i: element (lo + a2, lo + b2), stored at b2 * m + a2
                    int at = 0;
                    while (at + 1 < A.n_atoms && 2 * A.atom_first[at + 1] <= i) at++;
                    const int lo = 2 * A.atom_first[at], m = 2 * A.atom_norb[at], a2 = i - lo;
                    if (a2 >= 0 && a2 < m)
                        for (int b2 = tid; b2 < m; b2 += NT)
                            A.atomL[(size_t)iw * A.atom_tot + A.atom_off[at] + b2 * m + a2] = L.stage[s * n + lo + b2];
                }
                for (int e = tid; e < nCL * nCL; e += NT) {
                    const int l = e / nCL, a2 = e - l * nCL;
                    if (A.cl[l] == i) L.glcl[e] = L.stage[s * n + A.cl[a2]];
                }
            }
        }
    }
    E.st_max = st_max;
    E.st_s2 = st_s2;
    E.st_d2 = st_d2;
}

// After the G^< rows: G^r from Ms through the pivot maps (packed for INIT2 and
// SCF2, its diagonal for SCF1, dosd and the current integrand for FINAL), and the
// workgroup's SCF2 sums.
template <int NT, int EB>
__device__ __forceinline__ void epi_gr(const DysonArgs &A, const WLayout &L, Epi &E, const double2 *Ms, int iw,
                                       double fL, int tid) {
    constexpr int NW = NT / 64;
    const int n = A.norb, nA = A.nA, nA2 = nA * nA, nCL = A.nCL;
    const int lane = tid & 63, wv = tid >> 6;
    const double *xe = E.xe, *oe = E.oe;
    double *fe = E.fe, *ye = E.ye;
    const bool scf2 = E.scf2, init2 = E.init2;
    double st_max = E.st_max, st_s2 = E.st_s2, st_d2 = E.st_d2;
    auto GR = [&](int i, int j) { return Ms[L.qinv[j] * n + L.piv[i]]; };
    // G^r from Ms through the pivot maps
    if (init2 || scf2) {
        for (int e0 = tid; e0 < nA2; e0 += NT * EB) {
            double x0[EB], x1[EB], o0[EB], o1[EB];
            double2 g[EB];
#pragma unroll
            for (int u = 0; u < EB; u++) {
                const int e = min(e0 + NT * u, nA2 - 1), ar = e / nA, bb = e - ar * nA;
                g[u] = GR(L.actl[ar], L.actl[bb]);
                if (scf2) {
                    x0[u] = xe[e]; x1[u] = xe[e + nA2];
                    o0[u] = oe[e]; o1[u] = oe[e + nA2];
                }
            }
#pragma unroll
            for (int u = 0; u < EB; u++) {
                const int e = e0 + NT * u;
                if (e >= nA2) continue;
                if (init2) {
                    ye[e] = g[u].x;
                    ye[e + nA2] = g[u].y;
                } else {
                    const double f0 = g[u].x - x0[u], f1 = g[u].y - x1[u];
                    fe[e] = f0;
                    fe[e + nA2] = f1;
                    st_max = fmax(st_max, fmax(fabs(f0), fabs(f1)));
                    st_s2 += f0 * f0 + f1 * f1;
                    const double e0v = f0 - o0[u], e1v = f1 - o1[u];
                    st_d2 += e0v * e0v + e1v * e1v;
                    const int ar = e / nA;
                    if (ar * nA + ar == e) A.diagR[(size_t)iw * nA + ar] = g[u].y;
                }
            }
        }
        if (scf2) {
            st_max = wave_max(st_max);
            st_s2 = wave_sum(st_s2);
            st_d2 = wave_sum(st_d2);
            if (lane == 0) { L.red[3 * wv] = st_max; L.red[3 * wv + 1] = st_s2; L.red[3 * wv + 2] = st_d2; }
            __syncthreads();
            if (tid == 0) {
                double m = L.red[0], s2 = L.red[1], d2 = L.red[2];
                for (int w2 = 1; w2 < NW; w2++) { m = fmax(m, L.red[3 * w2]); s2 += L.red[3 * w2 + 1]; d2 += L.red[3 * w2 + 2]; }
                double *pp = A.partials + (size_t)iw * 4;
                pp[0] = m; pp[1] = s2; pp[2] = d2; pp[3] = 0.0;
            }
        }
    } else if (A.mode == QS_M_SCF1) {
        for (int k2 = tid; k2 < nA; k2 += NT) A.diagR[(size_t)iw * nA + k2] = GR(L.actl[k2], L.actl[k2]).y;
    } else {
        for (int i = tid; i < n; i += NT) A.dosd[(size_t)iw * n + i] = GR(i, i).y;
        for (int e = tid; e < nCL * nCL; e += NT) {
            const int k2 = e / nCL, a2 = e - k2 * nCL;
            L.grcl[e] = GR(A.cl[k2], A.cl[a2]);
        }
        __syncthreads();
        if (tid == 0) {
            // Tr_s{ i GamL [ (f_L - 1) G^< - f_L G^> ] } / 2pi on the left contact  (src/scf.f90:390-409)
            double tu = 0.0, td = 0.0;
            for (int a2 = 0; a2 < nCL; a2++) {
                double2 s = make_double2(0.0, 0.0);
                for (int l = 0; l < nCL; l++) {
                    const double2 glc = L.glcl[l * nCL + a2];
                    const double2 gr_la = L.grcl[l * nCL + a2], gr_al = L.grcl[a2 * nCL + l];
                    const double2 gg = cadd(glc, csub(gr_la, cconj(gr_al)));
                    const double2 X = make_double2((fL - 1.0) * glc.x - fL * gg.x, (fL - 1.0) * glc.y - fL * gg.y);
                    cfma(s, A.gamL_cl[a2 * nCL + l], X);
                }
                if ((A.cl[a2] & 1) == 0) tu += -s.y; else td += -s.y;
            }
            A.trUD[(size_t)iw * 2 + 0] = tu / TWO_PI;
            A.trUD[(size_t)iw * 2 + 1] = td / TWO_PI;
        }
    }
}

// PART 0: the whole per-energy solve in one kernel.  PART 1: the build and the
// Gauss-Jordan inverse, which write M and the pivot maps to the slab; PART 2: the
// Keldysh products and the epilogues, from the slab.  The split lets the
// Gauss-Jordan part, whose steps are bound by the latency of their dependency
// chain, run three energies per CU (its registers and LDS allow it) instead of two.
//
// blk is the workgroup's energy within A's range: energy iw0 + blk, slab entry blk.
template <int RA, int RB, int TI, int TJ, int PART>
__device__ __forceinline__ void dyson_w_body(const DysonArgs &A, const int blk) {
    constexpr int NT = TI * TJ, NW = NT / 64, EB = eb_gr<RA, RB>(), EBL = eb_gl<RA, RB>();
    extern __shared__ __align__(16) char lds_raw[];
    const int n = A.norb, nA = A.nA, nA2 = nA * nA, nCL = A.nCL;
    const int tid = threadIdx.x, li = tid / TJ, lj = tid % TJ, lane = tid & 63, wv = tid >> 6;
    const int iw = A.iw0 + blk;
    const WLayout L = PART == 1 ? wlayout_gj(lds_raw, RA, RB, TI, TJ) : wlayout(lds_raw, n, nCL, RA, RB, TI, TJ);
    const double w = A.w_init + (double)(iw + 1) * A.delta;
    const double fL = qs_fermi(w, A.V, A.mu, A.beta), fR = qs_fermi(w, 0.0, A.mu, A.beta);
    const double2 *sigw = (A.order == 2 && A.use_sigma && A.sig) ? A.sig + (size_t)iw * A.sig_ld : nullptr;
    // 32-bit offsets from uniform bases throughout (3 n^2 < 2^31)
    double2 *Ms = A.mscr + (size_t)blk * wave_slab_stride(n);  // [c][r]: M column-major
    int *maps = (int *)(Ms + (size_t)n * n);                             // piv [n], qinv [n]
    // G^r[q_r][p_c] = M[r][c]: G^r(i, j) = M(piv[i], qinv[j]), read from Ms (epi_gr)
    const int colbase = lj * n + li;                         // (r, c) = (li + TI a, lj + TJ b) at colbase + TJ b n + TI a

    if (PART != 1) {
        for (int i = tid; i < n; i += NT) L.posl[i] = A.pos[i];
        for (int i = tid; i < nA; i += NT) L.actl[i] = A.act[i];
    }

    if constexpr (PART != 2) {
    // ---- 1. A = w + Hemb - Sigma1 - Sigma2^r ------------------------------------
    double2 M[RA][RB];
    // row-clamped base per a, columns past n read the allocation's zero pad
#pragma unroll
    for (int a = 0; a < RA; a++) {
        const double2 *hrow = A.Hemb + min(li + TI * a, n - 1) * n + lj;
#pragma unroll
        for (int b = 0; b < RB; b++) M[a][b] = hrow[TJ * b];
    }
#pragma unroll
    for (int a = 0; a < RA; a++)
#pragma unroll
        for (int b = 0; b < RB; b++)
            if (li + TI * a == lj + TJ * b) M[a][b].x += w;
    if (A.use_sigma && A.sig1) {
#pragma unroll
        for (int a = 0; a < RA; a++)
#pragma unroll
            for (int b = 0; b < RB; b++) {
                const int i = li + TI * a, j = lj + TJ * b;
                if ((i >> 1) == (j >> 1) && i < n) M[a][b] = csub(M[a][b], A.sig1[(i >> 1) * 4 + (i & 1) * 2 + (j & 1)]);
            }
    }
    if (A.use_sigma && sigw) {
        int pc[RB];
#pragma unroll
        for (int b = 0; b < RB; b++) pc[b] = A.pos[min(lj + TJ * b, n - 1)];
#pragma unroll
        for (int a = 0; a < RA; a++) {
            const int pa = A.pos[min(li + TI * a, n - 1)];
            double2 s2v[RB];
#pragma unroll
            for (int b = 0; b < RB; b++) s2v[b] = sigw[(pa >= 0 && pc[b] >= 0) ? sig_r_index(pa, pc[b], A.nAs) : 0];
#pragma unroll
            for (int b = 0; b < RB; b++)
                if (pa >= 0 && pc[b] >= 0) M[a][b] = csub(M[a][b], s2v[b]);
        }
    }
#pragma unroll
    for (int a = 0; a < RA; a++)
#pragma unroll
        for (int b = 0; b < RB; b++) {
            const int i = li + TI * a, j = lj + TJ * b;
            if (i >= n || j >= n) M[a][b] = make_double2(i == j ? 1.0 : 0.0, 0.0);
        }
    // rows already pivoted, as a wavefront-uniform mask (n <= 128: two words)
    uint64_t used_lo = 0, used_hi = 0;

    // ---- 2. Gauss-Jordan with partial pivoting ----------------------------------
    // step k = -1 only posts the candidates of column 0.
    // A SIMD holds waves of two energies.  A step's dependency chain (pivot pick,
    // reciprocal, next candidates and their posting) runs at wave priority 3 and
    // the rank-1 update at 0, so the other energy's bulk work fills this one's
    // stalls instead of delaying its chain (in-box A/B: 37.2/37.8 -> 36.2/36.2 ms
    // per sweep at 80 orbitals and 2^16 energies).
    for (int k = -1; k < n; k++) {
        __builtin_amdgcn_s_setprio(3);
        if (k >= 0) {
            const int buf = k & 1, kb = k / TJ, kl = k % TJ;
            const double2 *pcl = L.pcol + buf * TI * RA;
            // every LDS operand that does not depend on the pivot is read first, in
            // one round: column k for my rows and the NW posted candidates
            double2 fa[RA];
#pragma unroll
            for (int a = 0; a < RA; a++) fa[a] = pcl[li + TI * a];
            double2 cnd[NW];
#pragma unroll
            for (int w2 = 0; w2 < NW; w2++) cnd[w2] = L.cand[buf * NW + w2];
            // largest value, then smallest row; selects, no branches
            double bv = -1.0;
            int p = -1, pw = 0;
#pragma unroll
            for (int w2 = 0; w2 < NW; w2++) {
                const int i = (int)cnd[w2].y;
                const double v = cnd[w2].x;
                const bool take = i >= 0 && (p < 0 || v > bv || (v == bv && i < p));
                bv = take ? v : bv;
                p = take ? i : p;
                pw = take ? w2 : pw;
            }
            p = max(p, 0);  // p < 0 is unreachable: n - k unused rows remain at step k
            p = __builtin_amdgcn_readfirstlane(p);
            pw = __builtin_amdgcn_readfirstlane(pw);
            if (p < 64) used_lo |= 1ull << p;
            else used_hi |= 1ull << (p - 64);
            if (tid == 0) { L.piv[k] = p; L.qinv[p] = k; }
            const double2 *prow = L.crow + (buf * NW + pw) * TJ * RB;
            const int tp = p % TI, ap = p / TI;
            const double2 pv = pcl[p];
            if (tid == 0 && pv.x == 0.0 && pv.y == 0.0) atomicOr(A.status, 1);
            const double2 rinv = cinv(pv);
            double2 sp[RB];
#pragma unroll
            for (int b = 0; b < RB; b++) sp[b] = cmul(prow[lj + TJ * b], rinv);
            if (lj == kl) {
#pragma unroll
                for (int b = 0; b < RB; b++)
                    if (b == kb) {
                        sp[b] = rinv;
#pragma unroll
                        for (int a = 0; a < RA; a++) M[a][b] = make_double2(0.0, 0.0);
                    }
            }
            __builtin_amdgcn_s_setprio(0);
#pragma unroll
            for (int a = 0; a < RA; a++)
#pragma unroll
                for (int b = 0; b < RB; b++) cfms(M[a][b], fa[a], sp[b]);
            __builtin_amdgcn_s_setprio(3);
            if (li == tp) {
#pragma unroll
                for (int a = 0; a < RA; a++)
                    if (a == ap) {
#pragma unroll
                        for (int b = 0; b < RB; b++) M[a][b] = sp[b];
                    }
            }
        }
        if (k + 1 < n) {
            const int kk = k + 1, buf = kk & 1, bkk = kk / TJ, tkk = kk % TJ;
            double2 *pcn = L.pcol + buf * TI * RA;
            // the owners of column kk post it: the next step's multipliers, and the
            // values this wavefront's candidate search reads back
            if (lj == tkk) {
#pragma unroll
                for (int b = 0; b < RB; b++)
                    if (b == bkk)
#pragma unroll
                        for (int a = 0; a < RA; a++) pcn[li + TI * a] = M[a][b];
            }
            // a wavefront's LDS operations complete in order: its own posts are read
            // back without a barrier (the fence only keeps the compiler from moving them)
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            // this wavefront's rows li in [LW wv, LW wv + LW), all a: lane i takes row
            // LW wv + i % LW + TI (i / LW), so lane order is row order.  Key: |Re| + |Im|
            // as ordered bits + 2; a NaN ranks 1, below every number; used rows and
            // padding 0.  The largest key, on the first lane that holds it, is the
            // zgetf2 candidate: largest value, then smallest row.
            constexpr int LW = 64 / TJ;
            uint64_t key = 0;
            if (lane < LW * RA) {
                const int r = LW * wv + lane % LW + TI * (lane / LW);
                const uint64_t um = r < 64 ? used_lo >> (r & 63) : used_hi >> (r & 63);
                if (r < n && !(um & 1)) {
                    const double2 c = pcn[r];
                    const double v = fabs(c.x) + fabs(c.y);
                    key = v != v ? 1 : (uint64_t)__double_as_longlong(v) + 2;
                }
            }
            static_assert(LW * RA <= 32, "the candidate keys sit in lanes 0..31");
            const uint64_t wmax = wave_max_u64_lo32(key);
            int rw = -1;
            double best = -1.0;
            if (wmax != 0) {
                const uint64_t hit = __ballot(key == wmax);
                const int wl = __ffsll((unsigned long long)hit) - 1;
                rw = LW * wv + wl % LW + TI * (wl / LW);
                best = wmax == 1 ? -0.5 : __longlong_as_double((long long)(wmax - 2));
            }
            rw = __builtin_amdgcn_readfirstlane(rw);
            if (lane == 0) L.cand[buf * NW + wv] = make_double2(best, (double)rw);
            if (rw >= 0 && li == rw % TI) {
                const int aw = rw / TI;
                double2 *dst = L.crow + (buf * NW + wv) * TJ * RB;
#pragma unroll
                for (int a = 0; a < RA; a++)
                    if (a == aw)
#pragma unroll
                        for (int b = 0; b < RB; b++) dst[lj + TJ * b] = M[a][b];
            }
        }
        __syncthreads();
    }

    __builtin_amdgcn_s_setprio(0);
    // ---- 3. scratch copy of M: Ms (column-major) ------------------------------------
#pragma unroll
    for (int a = 0; a < RA; a++)
#pragma unroll
        for (int b = 0; b < RB; b++) {
            const int r = li + TI * a, c = lj + TJ * b;
            if (r < n && c < n) {
                Ms[colbase + TJ * b * n + TI * a] = M[a][b];
            }
        }
    if constexpr (PART == 1) {
        // the maps were complete at the last step's barrier
        for (int c = tid; c < n; c += NT) {
            maps[c] = L.piv[c];
            maps[n + c] = L.qinv[c];
        }
        return;
    }
    } else {
        // the first part's M and maps, from the slab (ordered by the kernel boundary)
        for (int c = tid; c < n; c += NT) {
            L.piv[c] = maps[c];
            L.qinv[c] = maps[n + c];
        }
        __syncthreads();
    }
    for (int c = tid; c < n; c += NT) L.spc[c] = A.suppos[L.piv[c]];
    __threadfence_block();
    __syncthreads();

    // ---- 4. T = M S'' -----------------------------------------------------------
    // T = sum_c M(:, c) S''(c, :), S''(c, c') = SL(spc[c], spc[c']) on the
    // support: i (f_L GamL + f_R GamR) + Sigma2^<  (src/scf.f90:333-336).
    // TI steps at a time: columns c of M (from Ms) go to stage[kk][*] and rows
    // c of S'' to chunk[kk][*], all loads of a chunk in one round, then TI
    // rank-1 updates from LDS.
    double2 T[RA][RB];
#pragma unroll
    for (int a = 0; a < RA; a++)
#pragma unroll
        for (int b = 0; b < RB; b++) T[a][b] = make_double2(0.0, 0.0);
    for (int k0 = 0; k0 < n; k0 += TI) {
        const int kc = min(TI, n - k0);
        double2 mv[RA], gl[RA], gr[RA], sg[RA];
        unsigned okm = 0;
#pragma unroll
        for (int q = 0; q < RA; q++) {  // TI n <= NT RA elements per chunk
            const int e = tid + NT * q, kk = e / n, r = e - kk * n;
            const bool in = e < kc * n;
            mv[q] = Ms[k0 * n + (in ? e : 0)];
            const int b1 = in ? L.spc[k0 + kk] : -1, b2 = in ? L.spc[r] : -1;
            const bool ok = b1 >= 0 && b2 >= 0;
            if (ok) okm |= 1u << q;
            const int gi = ok ? b1 * A.nS + b2 : 0;
            gl[q] = A.gamL_s[gi];
            gr[q] = A.gamR_s[gi];
            const SigL sl = sig_l_at(b1, b2, A.nAs);
            const int si = ok ? sl.idx : 0;
            sg[q] = sigw ? sig_l_value(sigw[si], ok && sl.flip) : make_double2(0.0, 0.0);
        }
        // unconditional stores (e < SC); entries past kc n are never read
#pragma unroll
        for (int q = 0; q < RA; q++) {
            const int e = tid + NT * q;
            double2 v = make_double2(0.0, 0.0);
            if (okm & (1u << q)) {
                const double re = fL * gl[q].x + fR * gr[q].x, im = fL * gl[q].y + fR * gr[q].y;
                v = make_double2(-im, re);
                if (sigw) v = cadd(v, sg[q]);
            }
            L.stage[e].x = mv[q].x; L.stage[e].y = mv[q].y;
            L.chunk[e].x = v.x; L.chunk[e].y = v.y;
        }
        __syncthreads();
        for (int kk = 0; kk < kc; kk++) {
            double2 ma[RA];
#pragma unroll
            for (int a = 0; a < RA; a++) ma[a] = L.stage[kk * n + min(li + TI * a, n - 1)];
#pragma unroll
            for (int b = 0; b < RB; b++) {
                const double2 sv = L.chunk[kk * n + min(lj + TJ * b, n - 1)];
#pragma unroll
                for (int a = 0; a < RA; a++) cfma(T[a][b], ma[a], sv);
            }
        }
        __syncthreads();
    }

    // ---- 5. P = T M^dagger --------------------------------------------------------
    // P = sum_c T(:, c) conj(M(:, c))^T, TJ steps at a time.  Columns k0 .. k0 + TJ - 1
    // of T are this thread's register column bb (TI = TJ): they go from registers to
    // stage[kk][*], and the columns of M from Ms to chunk[kk][*].
    // P = M S'' M^dagger is anti-Hermitian (S'' is), so only the register tiles on
    // and above the diagonal are formed (b >= a; a tile below it, (a, b) with a > b,
    // is -conj of tile (b, a) transposed, which the thread (lj, li) holds): 15 of 25
    // tiles at 80 orbitals.  The epilogue stages the missing ones from their mirrors.
    static_assert(RA == RB && TI == TJ, "square register tiles");
    double2 P[RA][RB];
#pragma unroll
    for (int a = 0; a < RA; a++)
#pragma unroll
        for (int b = 0; b < RB; b++) P[a][b] = make_double2(0.0, 0.0);
#pragma unroll
    for (int bb = 0; bb < RB; bb++) {
        const int k0 = TJ * bb;
        if (k0 >= n) break;
        const int kc = min(TJ, n - k0);
        double2 mv[RA];
#pragma unroll
        for (int q = 0; q < RA; q++) {
            const int e = tid + NT * q, ec = e < kc * n ? e : 0;
            mv[q] = Ms[k0 * n + ec];
        }
#pragma unroll
        for (int a = 0; a < RA; a++)
            if (lj < kc && li + TI * a < n) {
                L.stage[lj * n + li + TI * a].x = T[a][bb].x;
                L.stage[lj * n + li + TI * a].y = T[a][bb].y;
            }
#pragma unroll
        for (int q = 0; q < RA; q++) {
            L.chunk[tid + NT * q].x = mv[q].x;
            L.chunk[tid + NT * q].y = mv[q].y;
        }
        __syncthreads();
        for (int kk = 0; kk < kc; kk++) {
            double2 ta[RA];
#pragma unroll
            for (int a = 0; a < RA; a++) ta[a] = L.stage[kk * n + min(li + TI * a, n - 1)];
#pragma unroll
            for (int b = 0; b < RB; b++) {
                const double2 mb = L.chunk[kk * n + min(lj + TJ * b, n - 1)];
#pragma unroll
                for (int a = 0; a <= b; a++) cfmac(P[a][b], ta[a], mb);
            }
        }
        __syncthreads();
    }

    // ---- 6. epilogues ---------------------------------------------------------------
    Epi E = epi_begin(A, iw);
    // G^< in TI-row stages: pivoted row r = TI a + s is natural row qinv[r].  Stage a
    // takes this thread's tiles (a, b >= a) as they are, and for b < a its tile
    // (b, a) transposed and negated-conjugated: element (li + TI b, lj + TI a) of P
    // gives (lj + TI a, li + TI b) as -conj.
#pragma unroll
    for (int a = 0; a < RA; a++) {
        if (TI * a >= n) break;
#pragma unroll
        for (int b = 0; b < RB; b++) {
            if (b >= a) {
                const int r = li + TI * a, c = lj + TJ * b;
                if (r < n && c < n) L.stage[li * n + L.qinv[c]] = P[a][b];
            } else {
                const int r = lj + TI * a, c = li + TI * b;
                if (r < n && c < n) L.stage[lj * n + L.qinv[c]] = make_double2(-P[b][a].x, P[b][a].y);
            }
        }
        __syncthreads();
        epi_rows<NT, EBL>(A, L, E, TI * a, min(TI, n - TI * a), iw, tid);
        __syncthreads();
    }
    epi_gr<NT, EB>(A, L, E, Ms, iw, fL, tid);
}

// ---- the Keldysh products on FP64 MFMA (after the Gauss-Jordan part) ----------------
// G^< = M S'' M^dagger as U = S'' M^dagger, then P = M U (the same product as
// T M^dagger with T = M S''), on v_mfma_f64_16x16x4f64, four real products per
// complex one.  NS column strips of 16 (n <= 16 NS), one wavefront per strip.
//  * U: wavefront J keeps strip J of U in MFMA accumulators: register q of tile I
//    holds U[16 I + lane / 16 + 4 q][16 J + lane % 16].  S'' comes through LDS in
//    chunks of 16 columns, M^dagger (B operand) straight from the slab.
//  * P: for each row chunk I, wavefront J forms the tile P[I][J] with A = rows
//    16 I .. 16 I + 15 of M from the slab and B = its own accumulators of U: an
//    FP64 accumulator register q holds rows 4q .. 4q + 3 of a tile in exactly the
//    lanes in which the B operand of a k-step takes them.  The tile goes to the
//    LDS stage through the pivot map, and the chunk's packing runs (epi_rows).
//  * then G^r and the sums (epi_gr), as in the register kernel.
// A workgroup takes EPW energies (EPW NS wavefronts, each energy its own LDS):
// the waves of a workgroup are given registers on every SIMD in rounds of four, so
// two 5-wavefront workgroups do not share a CU at 168 VGPRs while one of 10 does.
typedef double mf4 __attribute__((vector_size(4 * sizeof(double))));
// packing batches of the MFMA kernel's epilogue (G^<, G^r) and its waves per SIMD:
// three, so that two 5-wavefront workgroups share a CU at 80 orbitals (168 VGPRs)
constexpr int KMF_EBL = 2, KMF_EB = 4, KMF_WPS = 3;

__device__ __forceinline__ mf4 mfma4(double a, double b, mf4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// LDS: stage [16][n] | chunk [16][n] | grcl, glcl [nCL^2] | red [3 NS] | 5 int maps [16 NS]
__host__ __device__ inline size_t klayout_bytes(int n, int nCL, int NS) {
    return (size_t)16 * (32 * (size_t)n + 2 * (size_t)nCL * nCL) + (size_t)8 * 3 * NS + (size_t)4 * 5 * 16 * NS;
}
__device__ __forceinline__ WLayout klayout(char *base, int n, int nCL, int NS) {
    WLayout L;
    L.crow = L.pcol = L.srow = L.cand = nullptr;
    L.stage = (double2 *)base;
    L.chunk = L.stage + 16 * n;
    L.grcl = L.chunk + 16 * n;
    L.glcl = L.grcl + nCL * nCL;
    L.red = (double *)(L.glcl + nCL * nCL);
    L.piv = (int *)(L.red + 3 * NS);
    L.qinv = L.piv + 16 * NS;
    L.spc = L.qinv + 16 * NS;
    L.actl = L.spc + 16 * NS;
    L.posl = L.actl + 16 * NS;
    return L;
}

// an energy's LDS, rounded to 16 bytes
__host__ __device__ inline size_t klayout_stride(int n, int nCL, int NS) { return (klayout_bytes(n, nCL, NS) + 15) & ~(size_t)15; }

// energies EPW blk0 .. EPW blk0 + EPW - 1 of A's range, one per NS wavefronts
template <int NS, int EPW>
__device__ __forceinline__ void keldysh_mf_body(const DysonArgs &A, const int blk0) {
    constexpr int NT = 64 * NS;
    extern __shared__ __align__(16) char lds_raw[];
    const int n = A.norb, nA = A.nA, nCL = A.nCL;
    const int h = EPW > 1 ? (int)threadIdx.x / NT : 0;
    const int tid = (int)threadIdx.x - h * NT, lane = tid & 63, wv = tid >> 6, l16 = lane & 15, l4 = lane >> 4;
    const int blk = EPW * blk0 + h;
    const int iw = A.iw0 + blk;
    const WLayout L = klayout(lds_raw + (size_t)h * klayout_stride(n, nCL, NS), n, nCL, NS);
    const double w = A.w_init + (double)(iw + 1) * A.delta;
    const double fL = qs_fermi(w, A.V, A.mu, A.beta), fR = qs_fermi(w, 0.0, A.mu, A.beta);
    const double2 *sigw = (A.order == 2 && A.use_sigma && A.sig) ? A.sig + (size_t)iw * A.sig_ld : nullptr;
    const double2 *Ms = A.mscr + (size_t)blk * wave_slab_stride(n);  // [c][r]: M column-major
    const int *maps = (const int *)(Ms + (size_t)n * n);                // piv [n], qinv [n]
    for (int i = tid; i < n; i += NT) L.posl[i] = A.pos[i];
    for (int i = tid; i < nA; i += NT) L.actl[i] = A.act[i];
    for (int c = tid; c < n; c += NT) {
        L.piv[c] = maps[c];
        L.qinv[c] = maps[n + c];
    }
    __syncthreads();
    for (int c = tid; c < n; c += NT) L.spc[c] = A.suppos[L.piv[c]];
    __syncthreads();

    // ---- U = S'' M^dagger, strip J = wv ------------------------------------------------
    // S''[i][k] = SL[spc[i]][spc[k]] on the support: i (f_L GamL + f_R GamR) + Sigma2^<
    // (src/scf.f90:333-336)
    mf4 uR[NS], uI[NS];
#pragma unroll
    for (int I = 0; I < NS; I++) {
        uR[I] = mf4{0.0, 0.0, 0.0, 0.0};
        uI[I] = mf4{0.0, 0.0, 0.0, 0.0};
    }
    const int jB = 16 * wv + l16;  // this lane's column of U
    for (int k0 = 0; k0 < n; k0 += 16) {
        const int kc = min(16, n - k0);
        // this lane's B operands of the chunk's four k-steps, M[jB][k0 + 4 s + lane / 16]
        // (B = conj(M)), loaded before the S'' build so that the two overlap
        double2 mb[4];
#pragma unroll
        for (int s = 0; s < 4; s++) {
            const int k = k0 + 4 * s + l4;
            mb[s] = (jB < n && k < n) ? Ms[k * n + jB] : make_double2(0.0, 0.0);
        }
        // chunk[kk][i] = S''[i][k0 + kk]; zero past the chunk's kc columns
        for (int e = tid; e < 16 * n; e += NT) {
            const int kk = e / n, i = e - kk * n;
            double2 v = make_double2(0.0, 0.0);
            if (kk < kc) {
                const int b1 = L.spc[i], b2 = L.spc[k0 + kk];
                if (b1 >= 0 && b2 >= 0) {
                    const double2 gl = A.gamL_s[b1 * A.nS + b2], gr = A.gamR_s[b1 * A.nS + b2];
                    v = make_double2(-(fL * gl.y + fR * gr.y), fL * gl.x + fR * gr.x);
                    if (sigw) {
                        const SigL sl = sig_l_at(b1, b2, A.nAs);
                        v = cadd(v, sig_l_value(sigw[sl.idx], sl.flip));
                    }
                }
            }
            L.chunk[e] = v;
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < 4; s++) {
            const double bR = mb[s].x, bI = -mb[s].y;
#pragma unroll
            for (int I = 0; I < NS; I++) {
                const int i = 16 * I + l16;
                double2 a = make_double2(0.0, 0.0);
                if (i < n) a = L.chunk[(4 * s + l4) * n + i];
                uR[I] = mfma4(a.x, bR, uR[I]);
                uR[I] = mfma4(-a.y, bI, uR[I]);
                uI[I] = mfma4(a.x, bI, uI[I]);
                uI[I] = mfma4(a.y, bR, uI[I]);
            }
        }
        __syncthreads();
    }

    // ---- P = M U by row chunks, each packed as it is formed -----------------------------
    Epi E = epi_begin(A, iw);
    const int cB = 16 * wv + l16;  // this lane's column of the P tile
    const int qc = cB < n ? L.qinv[cB] : 0;
    for (int I = 0; I * 16 < n; I++) {
        // rows 16 I .. 16 I + 15 of M to LDS in one coalesced batch: chunk[k][r - 16 I]
        // (zero past n); the chunk buffer is free since the last barrier
        for (int e = tid; e < 16 * n; e += NT) {
            const int k = e >> 4, r = 16 * I + (e & 15);
            L.chunk[e] = r < n ? Ms[k * n + r] : make_double2(0.0, 0.0);
        }
        __syncthreads();
        mf4 pR = {0.0, 0.0, 0.0, 0.0}, pI = pR;
#pragma unroll
        for (int t = 0; t < 4 * NS; t++) {
            const int k = 4 * t + l4;
            const double2 a = k < n ? L.chunk[k * 16 + l16] : make_double2(0.0, 0.0);  // M[16 I + lane % 16][k]
            const double bR = uR[t / 4][t % 4], bI = uI[t / 4][t % 4];
            pR = mfma4(a.x, bR, pR);
            pI = mfma4(a.x, bI, pI);
            pR = mfma4(-a.y, bI, pR);
            pI = mfma4(a.y, bR, pI);
        }
        // pivoted row 16 I + s of G^<, s = lane / 16 + 4 q, column cB: natural column qinv[cB]
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int srow = l4 + 4 * q;
            if (16 * I + srow < n && cB < n) L.stage[srow * n + qc] = make_double2(pR[q], pI[q]);
        }
        __syncthreads();
        epi_rows<NT, KMF_EBL>(A, L, E, 16 * I, min(16, n - 16 * I), iw, tid);
        __syncthreads();
    }
    epi_gr<NT, KMF_EB>(A, L, E, Ms, iw, fL, tid);
}

// the Keldysh part on MFMA: two energies per workgroup (k_keldysh_mf2), one (the
// last energy of an odd count), or (group) two energies of the members' ranges
template <int NS>
__global__ __launch_bounds__(128 * NS, KMF_WPS) void k_keldysh_mf2(DysonArgs A) {
    keldysh_mf_body<NS, 2>(A, blockIdx.x);
}
template <int NS>
__global__ __launch_bounds__(64 * NS, KMF_WPS) void k_keldysh_mf(DysonArgs A) {
    keldysh_mf_body<NS, 1>(A, blockIdx.x);
}
template <int NS>
__global__ __launch_bounds__(128 * NS, KMF_WPS) void k_keldysh_mfg2(const DysonArgs *__restrict__ tab,
                                                                   DysonGroupIdx g) {
    const int b = blockIdx.x;
    int m = 0;
    for (int k = 1; k < g.nmem; k++) m += b >= g.first[k] ? 1 : 0;
    const DysonArgs A = tab[m];
    keldysh_mf_body<NS, 2>(A, b - g.first[m]);
}

// the MFMA Keldysh part over cnt energies of a (iw0 .. iw0 + cnt - 1, slab entries 0 .. cnt - 1)
template <int NS>
void launch_kmf(const DysonArgs &a, int cnt, hipStream_t s) {
    const size_t lds1 = klayout_stride(a.norb, a.nCL, NS);
    if (cnt >= 2) hipLaunchKernelGGL((k_keldysh_mf2<NS>), dim3(cnt / 2), dim3(128 * NS), 2 * lds1, s, a);
    if (cnt & 1) {
        DysonArgs t = a;
        t.iw0 = a.iw0 + cnt - 1;
        t.mscr = a.mscr + (size_t)(cnt - 1) * wave_slab_stride(a.norb);
        hipLaunchKernelGGL((k_keldysh_mf<NS>), dim3(1), dim3(64 * NS), lds1, s, t);
    }
}

template <int RA, int RB, int TI, int TJ, int PART>
__global__ __launch_bounds__(TI * TJ, PART == 1 ? gj_waves_per_simd<RA>() : 1) void k_dyson_w(DysonArgs A) {
    dyson_w_body<RA, RB, TI, TJ, PART>(A, blockIdx.x);
}

// One launch over the sweeps of several solves (a group step): member m owns
// workgroups [first[m], first[m + 1]) and its arguments are tab[m].  The member
// index is workgroup-uniform, so the arguments come in through scalar loads as
// a single solve's kernel arguments do.
template <int RA, int RB, int TI, int TJ, int PART>
__global__ __launch_bounds__(TI * TJ, PART == 1 ? gj_waves_per_simd<RA>() : 1) void k_dyson_wg(
    const DysonArgs *__restrict__ tab, DysonGroupIdx g) {
    const int b = blockIdx.x;
    int m = 0;
    for (int k = 1; k < g.nmem; k++) m += b >= g.first[k] ? 1 : 0;
    const DysonArgs A = tab[m];
    dyson_w_body<RA, RB, TI, TJ, PART>(A, b - g.first[m]);
}

// form 0: one fused kernel per chunk; 1: the Gauss-Jordan part, then the Keldysh
// part, per chunk; 2: as 1 with the Keldysh part on MFMA (16 x 16 tiles only)
template <int RA, int RB, int TI, int TJ>
hipError_t launch_w(const DysonArgs &a0, int count, int form, hipStream_t s) {
    const size_t lds = wlayout_bytes(a0.norb, a0.nCL, RA, RB, TI, TJ);
    const size_t lds1 = wlayout_gj_bytes(RA, RB, TI, TJ);
    if (hipError_t e = qs_lds_attr((const void *)k_dyson_w<RA, RB, TI, TJ, 0>)) return e;
    if (hipError_t e = qs_lds_attr((const void *)k_dyson_w<RA, RB, TI, TJ, 2>)) return e;
    if constexpr (TI == 16) {
        if (hipError_t e = qs_lds_attr((const void *)k_keldysh_mf<RA>)) return e;
        if (hipError_t e = qs_lds_attr((const void *)k_keldysh_mf2<RA>)) return e;
    }
    if (form == 2 && TI != 16) return hipErrorInvalidValue;
    // the scratch slab holds mscr_energies energies: launch in chunks of that many
    for (int c0 = 0; c0 < count; c0 += a0.mscr_energies) {
        DysonArgs a = a0;
        a.iw0 = a0.iw0 + c0;
        const int cnt = count - c0 < a0.mscr_energies ? count - c0 : a0.mscr_energies;
        if (form == 0) {
            hipLaunchKernelGGL((k_dyson_w<RA, RB, TI, TJ, 0>), dim3(cnt), dim3(TI * TJ), lds, s, a);
        } else {
            hipLaunchKernelGGL((k_dyson_w<RA, RB, TI, TJ, 1>), dim3(cnt), dim3(TI * TJ), lds1, s, a);
            if (form == 1) {
                hipLaunchKernelGGL((k_dyson_w<RA, RB, TI, TJ, 2>), dim3(cnt), dim3(TI * TJ), lds, s, a);
            } else if constexpr (TI == 16) {
                launch_kmf<RA>(a, cnt, s);
            }
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// the members' whole sweeps (iw0 = 0, nw energies each) in one launch per part
template <int RA, int RB, int TI, int TJ>
hipError_t launch_wg(const DysonArgs *ht, const DysonArgs *dt, int nmem, int form, hipStream_t s) {
    DysonGroupIdx g{};
    g.nmem = nmem;
    DysonGroupIdx g2{};   // the MFMA Keldysh part: two energies per workgroup
    g2.nmem = nmem;
    size_t lds = 0, ldsk = 0;
    int tot = 0, tot2 = 0;
    for (int m = 0; m < nmem; m++) {
        g.first[m] = tot;
        g2.first[m] = tot2;
        tot += ht[m].nw;
        tot2 += ht[m].nw / 2;
        const size_t lm = wlayout_bytes(ht[m].norb, ht[m].nCL, RA, RB, TI, TJ);
        lds = lm > lds ? lm : lds;
        const size_t lk = 2 * klayout_stride(ht[m].norb, ht[m].nCL, RA);
        ldsk = lk > ldsk ? lk : ldsk;
        if (form == 2 && (ht[m].nw & 1)) return hipErrorInvalidValue;
    }
    g.first[nmem] = tot;
    g2.first[nmem] = tot2;
    if (lds > 160 * 1024 || (form == 2 && ldsk > 160 * 1024)) return hipErrorInvalidValue;
    if (form == 2 && TI != 16) return hipErrorInvalidValue;
    if (hipError_t e = qs_lds_attr((const void *)k_dyson_wg<RA, RB, TI, TJ, 0>)) return e;
    if (hipError_t e = qs_lds_attr((const void *)k_dyson_wg<RA, RB, TI, TJ, 2>)) return e;
    if (form == 0) {
        hipLaunchKernelGGL((k_dyson_wg<RA, RB, TI, TJ, 0>), dim3(tot), dim3(TI * TJ), lds, s, dt, g);
    } else {
        hipLaunchKernelGGL((k_dyson_wg<RA, RB, TI, TJ, 1>), dim3(tot), dim3(TI * TJ),
                           wlayout_gj_bytes(RA, RB, TI, TJ), s, dt, g);
        if (form == 1) {
            hipLaunchKernelGGL((k_dyson_wg<RA, RB, TI, TJ, 2>), dim3(tot), dim3(TI * TJ), lds, s, dt, g);
        } else if constexpr (TI == 16) {
            if (hipError_t e = qs_lds_attr((const void *)k_keldysh_mfg2<RA>)) return e;
            hipLaunchKernelGGL((k_keldysh_mfg2<RA>), dim3(tot2), dim3(128 * RA), ldsk, s, dt, g2);
        }
    }
    return hipGetLastError();
}

}  // namespace

// the register tile of a size: 8 x 8 threads (one wavefront) up to 16 orbitals, 16 x 16 above
int qs_dyson_group_key(int n, int variant) {
    if (variant != QS_DYSON_WAVE && variant != QS_DYSON_SPLIT && variant != QS_DYSON_SPLIT_MFMA) return -1;
    if (n < 1 || n > qs_dyson_wave_max_norb()) return -1;
    if (variant != QS_DYSON_WAVE && n <= 16) return -1;
    const int T = n <= 16 ? 8 : 16, R = (n + T - 1) / T;
    return 1000 * variant + 10 * T + R;
}

hipError_t qs_launch_dyson_group(const DysonArgs *ht, const DysonArgs *dt, int nmem, int variant, hipStream_t s) {
    if (nmem < 1 || nmem > QS_GROUP_MAX || !ht || !dt) return hipErrorInvalidValue;
    const int key = qs_dyson_group_key(ht[0].norb, variant);
    if (key < 0) return hipErrorInvalidValue;
    for (int m = 0; m < nmem; m++) {
        const DysonArgs &a = ht[m];
        if (qs_dyson_group_key(a.norb, variant) != key || !a.mscr || a.iw0 != 0 || a.nw < 1 ||
            a.mscr_energies < a.nw)
            return hipErrorInvalidValue;
    }
    const int form = variant == QS_DYSON_WAVE ? 0 : variant == QS_DYSON_SPLIT ? 1 : 2;
    const int n = ht[0].norb;
    if (n <= 8) return launch_wg<1, 1, 8, 8>(ht, dt, nmem, form, s);
    if (n <= 16) return launch_wg<2, 2, 8, 8>(ht, dt, nmem, form, s);
    switch ((n + 15) / 16) {
        case 2: return launch_wg<2, 2, 16, 16>(ht, dt, nmem, form, s);
        case 3: return launch_wg<3, 3, 16, 16>(ht, dt, nmem, form, s);
        case 4: return launch_wg<4, 4, 16, 16>(ht, dt, nmem, form, s);
        case 5: return launch_wg<5, 5, 16, 16>(ht, dt, nmem, form, s);
        case 6: return launch_wg<6, 6, 16, 16>(ht, dt, nmem, form, s);
        default: return hipErrorInvalidValue;
    }
}

// 8 x 8 threads (one wavefront) up to 16 orbitals, 16 x 16 above
size_t qs_dyson_wave_lds_bytes(int n, int nCL) {
    const int T = n <= 16 ? 8 : 16, R = (n + T - 1) / T;
    return wlayout_bytes(n, nCL, R, R, T, T);
}

size_t qs_dyson_wave_scratch_bytes(int n) { return wave_slab_stride(n) * sizeof(double2); }

int qs_dyson_wave_max_norb() { return 96; }

hipError_t qs_launch_dyson_wave(const DysonArgs &a, int count, int form, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    if (!a.mscr || a.mscr_energies <= 0 || a.norb > qs_dyson_wave_max_norb()) return hipErrorInvalidValue;
    if (qs_dyson_wave_lds_bytes(a.norb, a.nCL) > 160 * 1024) return hipErrorInvalidValue;
    const int n = a.norb;
    if (form < 0 || form > 2) return hipErrorInvalidValue;
    if (form && n <= 16) return hipErrorInvalidValue;  // one wavefront per energy: nothing to gain
    if (form == 2 && 2 * klayout_stride(n, a.nCL, (n + 15) / 16) > 160 * 1024) return hipErrorInvalidValue;
    if (n <= 8) return launch_w<1, 1, 8, 8>(a, count, form, s);
    if (n <= 16) return launch_w<2, 2, 8, 8>(a, count, form, s);
    switch ((n + 15) / 16) {
        case 2: return launch_w<2, 2, 16, 16>(a, count, form, s);
        case 3: return launch_w<3, 3, 16, 16>(a, count, form, s);
        case 4: return launch_w<4, 4, 16, 16>(a, count, form, s);
        case 5: return launch_w<5, 5, 16, 16>(a, count, form, s);
        case 6: return launch_w<6, 6, 16, 16>(a, count, form, s);
        default: return hipErrorInvalidValue;
    }
}
