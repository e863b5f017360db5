// qs_dyson.hip -- per-energy Dyson and Keldysh solve on gfx950 (FP64).
//
// Replaces G_at_energy + invert_complex (src/scf.f90:305-415,
// common/linear_algebra.f90:58-88) for every energy of the grid at once:
//
//   G^r = [ w + Hemb - Sigma1 - Sigma2^r(w) ]^-1                  (Dyson)
//   G^< = G^r(:,sup) [ i(f_L GamL + f_R GamR) + Sigma2^<(w) ] G^r(:,sup)^dagger   (Keldysh)
//
// One workgroup per energy.  Its TI x 16 threads (TI = 16 or 32: 4 or 8
// wavefronts) form a grid; thread (ti, tj) keeps the RA x RB elements
// (ti + TI a, tj + 16 b) of the matrix in registers ("cyclic" distribution,
// n <= TI RA and n <= 16 RB).  The variant is chosen per size by timing
// (qs_dyson_tune).
//
// Inverse: in-place Gauss-Jordan with partial row pivoting on |Re| + |Im|, the
// pivot rule of LAPACK's zgetf2.  Rows are never moved: step k picks the unused
// physical row p_k with the largest candidate, so the registers end with M where
//   G^r[q_r][p_c] = M[r][c],  q = p^-1.
// The candidates at step k equal LU's column k, so the pivot sequence and the
// stability are those of zgetrf.
//
// Keldysh in the pivoted basis: with S''[c][c'] = SL[p_c][p_c'] (rows/columns
// outside sup are zero), P = M S'' M^dagger gives G^<[q_r][q_r'] = P[r][r'].
//   phase 1: T = M S''   M staged in LDS, S'' streamed in 16-row chunks,
//   phase 2: P = T M^dag T staged in LDS, M^dag chunks written from registers.
// Both products are register-tiled complex FMA loops (FP64 vector FMA and FP64
// MFMA have the same peak on MI355X, 78.6 TFLOP/s).
//
// The epilogue writes, by mode:
//   INIT2  packed non-interacting G^r, G^< of the active block into x
//   SCF2   F = pack(G^r, G^<) - x, Im diagonals, and partial max|F|, |F|^2, |F-Fold|^2
//   SCF1   Im diagonals of the active states, on-site 2x2 blocks of G^<
//   FINAL  Im G^r_ii, on-site and intra-atom blocks of G^<, current integrand
// Packed layout per energy (common/linear_algebra.f90:142-191):
//   [Re G^r row-major | Im G^r row-major | Re G^<(i<j) row-wise | Im G^<(i<=j) row-wise]
#include "qs_device.h"
#include "qs_internal.h"

namespace {

constexpr double TWO_PI = 6.283185307179586476925286766559;

struct LdsLayout {
    double2 *big;   // n x (n+1)
    double2 *chunk; // 16 x 16R
    int *piv, *qinv, *spc, *actl;
    double *redv;
    int *redi;
    double *scratch;
    double2 *grcl;
};

// LDS: big (n x (n+1)), chunk (16 x 16 RB >= 16 x n), 4 int maps of 16 RB,
// 32 pivot candidates, 16 doubles of reduction scratch, the nCL x nCL block.
// The maps (64 RB ints) keep 8-byte alignment for every RB; grcl stays 16-byte aligned.
__device__ __forceinline__ LdsLayout lds_layout(char *base, int n, int RB) {
    LdsLayout L;
    L.big = (double2 *)base;
    L.chunk = L.big + (size_t)n * (n + 1);
    L.piv = (int *)(L.chunk + 16 * 16 * RB);
    L.qinv = L.piv + 16 * RB;
    L.spc = L.qinv + 16 * RB;
    L.actl = L.spc + 16 * RB;
    L.redv = (double *)(L.actl + 16 * RB);
    L.redi = (int *)(L.redv + 32);
    L.scratch = (double *)(L.redi + 32);
    L.grcl = (double2 *)(L.scratch + 16);
    return L;
}

template <int TI, int RA, int RB>
__global__ __launch_bounds__(TI * 16) void k_dyson(DysonArgs A) {
    constexpr int NT = TI * 16;
    extern __shared__ __align__(16) char lds_raw[];
    const int n = A.norb, pitch = n + 1;
    const int tid = threadIdx.x, ti = tid >> 4, tj = tid & 15;
    const int iw = A.iw0 + blockIdx.x;
    LdsLayout L = lds_layout(lds_raw, n, RB);
    const double w = A.w_init + (double)(iw + 1) * A.delta;
    const double fL = qs_fermi(w, A.V, A.mu, A.beta), fR = qs_fermi(w, 0.0, A.mu, A.beta);
    const int nA = A.nA, nA2 = nA * nA;
    const double2 *sigw = (A.order == 2 && A.use_sigma && A.sig) ? A.sig + (size_t)iw * A.sig_ld : nullptr;

    // ---- 1. A = w + Hemb - Sigma1 - Sigma2^r, in registers -------------------------
    // Every load of the tile is issued before any is used (clamped in-bounds
    // addresses and selects, no branches), so the tile costs two memory round
    // trips (pos, then Sigma2) rather than one per element.
    int pr[RA], pc[RB];
#pragma unroll
    for (int a = 0; a < RA; a++) pr[a] = A.pos[min(ti + TI * a, n - 1)];
#pragma unroll
    for (int b = 0; b < RB; b++) pc[b] = A.pos[min(tj + 16 * b, n - 1)];
    double2 M[RA][RB];
#pragma unroll
    for (int a = 0; a < RA; a++)
#pragma unroll
        for (int b = 0; b < RB; b++) M[a][b] = A.Hemb[min(ti + TI * a, n - 1) * n + min(tj + 16 * b, n - 1)];
    const bool use_s1 = A.use_sigma && A.sig1, use_s2 = A.use_sigma && sigw;
    double2 s1v[RA][RB];
    if (use_s1) {
#pragma unroll
        for (int a = 0; a < RA; a++)
#pragma unroll
            for (int b = 0; b < RB; b++) {
                const int i = min(ti + TI * a, n - 1), j = min(tj + 16 * b, n - 1);
                s1v[a][b] = A.sig1[((i >> 1) == (j >> 1)) ? (i >> 1) * 4 + (i & 1) * 2 + (j & 1) : 0];
            }
    }
    double2 s2v[RA][RB];
    if (use_s2) {
#pragma unroll
        for (int a = 0; a < RA; a++)
#pragma unroll
            for (int b = 0; b < RB; b++) {
                const bool ok = pr[a] >= 0 && pc[b] >= 0;
                s2v[a][b] = sigw[ok ? sig_r_index(pr[a], pc[b], A.nAs) : 0];
            }
    }
#pragma unroll
    for (int a = 0; a < RA; a++)
#pragma unroll
        for (int b = 0; b < RB; b++) {
            const int i = ti + TI * a, j = tj + 16 * b;
            double2 v = M[a][b];
            if (i == j) v.x += w;
            if (use_s1 && (i >> 1) == (j >> 1)) v = csub(v, s1v[a][b]);
            if (use_s2 && pr[a] >= 0 && pc[b] >= 0) v = csub(v, s2v[a][b]);
            if (i >= n || j >= n) v = make_double2(i == j ? 1.0 : 0.0, 0.0);
            M[a][b] = v;
        }
    unsigned used = 0;  // bit a: my row ti + TI a already pivoted
#pragma unroll
    for (int a = 0; a < RA; a++)
        if (ti + TI * a >= n) used |= 1u << a;

    double2 *prow = L.chunk, *pcol = L.chunk + 16 * RB;

    // ---- 2. Gauss-Jordan with partial pivoting ----------------------------------
    // Step k: the owners of column k post their best unused row, every wave
    // reduces the TI candidates, the pivot row and column go through LDS, and
    // each thread applies M -= M(:,k) (M(p,:) / M(p,k)) to its tile with
    // column k zeroed first, so that it becomes -M(:,k) / M(p,k); row p is then
    // replaced by M(p,:) / M(p,k) with 1 / M(p,k) at column k.
    const int lane = tid & 63;
    for (int k = 0; k < n; k++) {
        const int bk = k >> 4, tk = k & 15;
        if (tj == tk) {
            double best = -1.0;
            int bi = -1;
#pragma unroll
            for (int b = 0; b < RB; b++)
                if (b == bk)
#pragma unroll
                    for (int a = 0; a < RA; a++)
                        if (!(used & (1u << a))) {
                            double v = fabs(M[a][b].x) + fabs(M[a][b].y);
                            if (v != v) v = -0.5;  // NaN never wins over a number
                            if (bi < 0 || v > best) { best = v; bi = ti + TI * a; }
                        }
            L.redv[ti] = best;
            L.redi[ti] = bi;
        }
        __syncthreads();
        // every wavefront reduces the TI candidates (largest value, then smallest row)
        int p;
        {
            const int t = lane & (TI - 1);
            int ii = L.redi[t];
            double v = L.redv[t];
            if (ii < 0) { v = -3.0; ii = 1 << 30; }
#pragma unroll
            for (int o = TI / 2; o > 0; o >>= 1) {
                const double vo = __shfl_xor(v, o, 64);
                const int io = __shfl_xor(ii, o, 64);
                if (vo > v || (vo == v && io < ii)) { v = vo; ii = io; }
            }
            p = __builtin_amdgcn_readfirstlane(ii);
        }
        const int tp = p % TI, ap = p / TI;
        if (ti == tp) {
#pragma unroll
            for (int a = 0; a < RA; a++)
                if (a == ap)
#pragma unroll
                    for (int b = 0; b < RB; b++) prow[tj + 16 * b] = M[a][b];
        }
        if (tj == tk) {
#pragma unroll
            for (int b = 0; b < RB; b++)
                if (b == bk)
#pragma unroll
                    for (int a = 0; a < RA; a++) pcol[ti + TI * a] = M[a][b];
        }
        if (tid == 0) { L.piv[k] = p; L.qinv[p] = k; }
        __syncthreads();
        const double2 pv = pcol[p];
        if (tid == 0 && pv.x == 0.0 && pv.y == 0.0) atomicOr(A.status, 1);
        const double2 rinv = cinv(pv);
        double2 sp[RB], fr[RA];
#pragma unroll
        for (int b = 0; b < RB; b++) sp[b] = cmul(prow[tj + 16 * b], rinv);
#pragma unroll
        for (int a = 0; a < RA; a++) fr[a] = pcol[ti + TI * a];
        if (tj == tk) {
#pragma unroll
            for (int b = 0; b < RB; b++)
                if (b == bk) {
                    sp[b] = rinv;
#pragma unroll
                    for (int a = 0; a < RA; a++) M[a][b] = make_double2(0.0, 0.0);
                }
        }
#pragma unroll
        for (int a = 0; a < RA; a++)
#pragma unroll
            for (int b = 0; b < RB; b++) cfms(M[a][b], fr[a], sp[b]);
        if (ti == tp) {
#pragma unroll
            for (int a = 0; a < RA; a++)
                if (a == ap) {
                    used |= 1u << a;
#pragma unroll
                    for (int b = 0; b < RB; b++) M[a][b] = sp[b];
                }
        }
    }
    __syncthreads();

    // ---- 3. G^r = permuted M staged in LDS; mode extraction -----------------------
#pragma unroll
    for (int a = 0; a < RA; a++)
#pragma unroll
        for (int b = 0; b < RB; b++) {
            const int r = ti + TI * a, c = tj + 16 * b;
            if (r < n && c < n) L.big[L.qinv[r] * pitch + L.piv[c]] = M[a][b];
        }
    for (int c = tid; c < n; c += NT) L.spc[c] = A.suppos[L.piv[c]];
    for (int c = tid; c < nA; c += NT) L.actl[c] = A.act[c];
    __syncthreads();

    const size_t nblk = 3 * (size_t)nA2, off = (size_t)iw * nblk;
    double st_max = 0.0, st_s2 = 0.0, st_d2 = 0.0;
    // Packing epilogues read x and Fold in batches of EU elements per thread,
    // all loads of a batch ahead of its stores.
    constexpr int EU = 8;
    if (A.mode == QS_M_INIT2) {
        for (int e = tid; e < nA2; e += NT) {
            const int a = e / nA, b = e - a * nA;
            const double2 g = L.big[L.actl[a] * pitch + L.actl[b]];
            A.xout[off + e] = g.x;
            A.xout[off + nA2 + e] = g.y;
        }
    } else if (A.mode == QS_M_SCF2) {
        for (int e0 = tid; e0 < nA2; e0 += EU * NT) {
            double x0[EU], x1[EU], o0[EU], o1[EU];
#pragma unroll
            for (int u = 0; u < EU; u++) {
                const size_t d0 = off + min(e0 + u * NT, nA2 - 1), d1 = d0 + nA2;
                x0[u] = A.x[d0]; x1[u] = A.x[d1];
                o0[u] = A.Fold[d0]; o1[u] = A.Fold[d1];
            }
#pragma unroll
            for (int u = 0; u < EU; u++) {
                const int e = e0 + u * NT;
                if (e < nA2) {
                    const int a = e / nA, b = e - a * nA;
                    const double2 g = L.big[L.actl[a] * pitch + L.actl[b]];
                    const size_t d0 = off + e, d1 = off + nA2 + e;
                    const double f0 = g.x - x0[u], f1 = g.y - x1[u];
                    A.F[d0] = f0;
                    A.F[d1] = f1;
                    st_max = fmax(st_max, fmax(fabs(f0), fabs(f1)));
                    st_s2 += f0 * f0 + f1 * f1;
                    const double e0v = f0 - o0[u], e1v = f1 - o1[u];
                    st_d2 += e0v * e0v + e1v * e1v;
                    if (a == b) A.diagR[(size_t)iw * nA + a] = g.y;
                }
            }
        }
    } else if (A.mode == QS_M_SCF1) {
        for (int k2 = tid; k2 < nA; k2 += NT) A.diagR[(size_t)iw * nA + k2] = L.big[L.actl[k2] * pitch + L.actl[k2]].y;
    } else {
        for (int i = tid; i < n; i += NT) A.dosd[(size_t)iw * n + i] = L.big[i * pitch + i].y;
        for (int e = tid; e < A.nCL * A.nCL; e += NT) {
            const int k2 = e / A.nCL, a = e - k2 * A.nCL;
            L.grcl[e] = L.big[A.cl[k2] * pitch + A.cl[a]];
        }
    }
    __syncthreads();

    // ---- 4. phase 1: T = M S'' --------------------------------------------------
#pragma unroll
    for (int a = 0; a < RA; a++)
#pragma unroll
        for (int b = 0; b < RB; b++) {
            const int r = ti + TI * a, c = tj + 16 * b;
            if (r < n && c < n) L.big[r * pitch + c] = M[a][b];
        }
    double2 acc[RA][RB];
#pragma unroll
    for (int a = 0; a < RA; a++)
#pragma unroll
        for (int b = 0; b < RB; b++) acc[a][b] = make_double2(0.0, 0.0);
    // S'' rows k0 .. k0+15 are fetched into registers one chunk ahead, so their
    // memory latency hides behind the previous chunk's FMA loop.
    constexpr int QF = (16 * 16 * RB + NT - 1) / NT;
    double2 fgl[QF], fgr[QF], fsg[QF];
    unsigned fok = 0;  // bit q: element q lies on the support (else S'' is zero there)
    auto fetch = [&](int k0) {
        fok = 0;
#pragma unroll
        for (int q = 0; q < QF; q++) {
            const int e = tid + q * NT, kk = e / n, c = e - kk * n, k = k0 + kk;
            const bool in = e < 16 * n && k < n;
            const int b1 = in ? L.spc[k] : -1, b2 = in ? L.spc[c] : -1;
            const bool ok = b1 >= 0 && b2 >= 0;
            const int gi = ok ? b1 * A.nS + b2 : 0;
            fgl[q] = A.gamL_s[gi];
            fgr[q] = A.gamR_s[gi];
            const SigL sl = sig_l_at(b1, b2, A.nAs);
            const int si = ok ? sl.idx : 0;
            fsg[q] = sigw ? sig_l_value(sigw[si], ok && sl.flip) : make_double2(0.0, 0.0);
            if (ok) fok |= 1u << q;
        }
    };
    fetch(0);
    for (int k0 = 0; k0 < n; k0 += 16) {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < QF; q++) {
            const int e = tid + q * NT;
            if (e < 16 * n) {
                double2 v = make_double2(0.0, 0.0);
                if (fok & (1u << q)) {
                    // i (f_L GamL + f_R GamR) + Sigma2^<  (src/scf.f90:333-336)
                    const double re = fL * fgl[q].x + fR * fgr[q].x, im = fL * fgl[q].y + fR * fgr[q].y;
                    v = make_double2(-im, re);
                    if (sigw) v = cadd(v, fsg[q]);
                }
                L.chunk[e] = v;
            }
        }
        __syncthreads();
        if (k0 + 16 < n) fetch(k0 + 16);
        const int kmax = min(16, n - k0);
        for (int kk = 0; kk < kmax; kk++) {
            double2 av[RA], bv[RB];
#pragma unroll
            for (int a = 0; a < RA; a++) {
                const int r = ti + TI * a;
                av[a] = r < n ? L.big[r * pitch + k0 + kk] : make_double2(0.0, 0.0);
            }
#pragma unroll
            for (int b = 0; b < RB; b++) {
                const int c = tj + 16 * b;
                bv[b] = c < n ? L.chunk[kk * n + c] : make_double2(0.0, 0.0);
            }
#pragma unroll
            for (int a = 0; a < RA; a++)
#pragma unroll
                for (int b = 0; b < RB; b++) cfma(acc[a][b], av[a], bv[b]);
        }
    }
    __syncthreads();
    // ---- 5. phase 2: P = T M^dagger ---------------------------------------------
#pragma unroll
    for (int a = 0; a < RA; a++)
#pragma unroll
        for (int b = 0; b < RB; b++) {
            const int r = ti + TI * a, c = tj + 16 * b;
            if (r < n && c < n) L.big[r * pitch + c] = acc[a][b];
            acc[a][b] = make_double2(0.0, 0.0);
        }
#pragma unroll
    for (int bb = 0; bb < RB; bb++) {
        const int k0 = 16 * bb;
        if (k0 >= n) break;
        __syncthreads();
        // chunk[kk][j] = conj(M[j][k0 + kk]); thread (ti, tj) owns column k0 + tj
#pragma unroll
        for (int a = 0; a < RA; a++) {
            const int j = ti + TI * a;
            if (j < n) L.chunk[tj * n + j] = cconj(M[a][bb]);
        }
        __syncthreads();
        const int kmax = min(16, n - k0);
        for (int kk = 0; kk < kmax; kk++) {
            double2 av[RA], bv[RB];
#pragma unroll
            for (int a = 0; a < RA; a++) {
                const int r = ti + TI * a;
                av[a] = r < n ? L.big[r * pitch + k0 + kk] : make_double2(0.0, 0.0);
            }
#pragma unroll
            for (int b = 0; b < RB; b++) {
                const int c = tj + 16 * b;
                bv[b] = c < n ? L.chunk[kk * n + c] : make_double2(0.0, 0.0);
            }
#pragma unroll
            for (int a = 0; a < RA; a++)
#pragma unroll
                for (int b = 0; b < RB; b++) cfma(acc[a][b], av[a], bv[b]);
        }
    }
    __syncthreads();
    // ---- 6. G^< = permuted P staged in LDS; mode extraction ------------------------
#pragma unroll
    for (int a = 0; a < RA; a++)
#pragma unroll
        for (int b = 0; b < RB; b++) {
            const int r = ti + TI * a, c = tj + 16 * b;
            if (r < n && c < n) L.big[L.qinv[r] * pitch + L.qinv[c]] = acc[a][b];
        }
    __syncthreads();

    if (A.mode == QS_M_INIT2 || A.mode == QS_M_SCF2) {
        const int nsu = (nA2 - nA) / 2;
        const bool scf2 = A.mode == QS_M_SCF2;
        for (int e0 = tid; e0 < nA2; e0 += EU * NT) {
            int code[EU];
            double xv[EU], ov[EU];
#pragma unroll
            for (int u = 0; u < EU; u++) {
                const int ec = min(e0 + u * NT, nA2 - 1);
                code[u] = A.tri_map[ec];
                if (scf2) {
                    const size_t d = off + 2 * (size_t)nA2 + ec;
                    xv[u] = A.x[d];
                    ov[u] = A.Fold[d];
                }
            }
#pragma unroll
            for (int u = 0; u < EU; u++) {
                const int e = e0 + u * NT;
                if (e < nA2) {
                    const int i = code[u] >> 16, j = code[u] & 0xffff;
                    const double2 g = L.big[L.actl[i] * pitch + L.actl[j]];
                    const double val = e < nsu ? g.x : g.y;
                    const size_t d = off + 2 * (size_t)nA2 + e;
                    if (!scf2) {
                        A.xout[d] = val;
                    } else {
                        const double f = val - xv[u];
                        A.F[d] = f;
                        st_max = fmax(st_max, fabs(f));
                        st_s2 += f * f;
                        const double e0v = f - ov[u];
                        st_d2 += e0v * e0v;
                        if (e >= nsu && i == j) A.diagL[(size_t)iw * nA + i] = g.y;
                    }
                }
            }
        }
        if (A.mode == QS_M_SCF2) {
            const double m = block_max(st_max, L.scratch);
            const double s2 = block_sum(st_s2, L.scratch);
            const double d2 = block_sum(st_d2, L.scratch);
            if (tid == 0) {
                double *pp = A.partials + (size_t)blockIdx.x * 4;
                pp[0] = m; pp[1] = s2; pp[2] = d2; pp[3] = 0.0;
            }
        }
    } else {
        if (A.mode == QS_M_SCF1)
            for (int k2 = tid; k2 < nA; k2 += NT) A.diagL[(size_t)iw * nA + k2] = L.big[L.actl[k2] * pitch + L.actl[k2]].y;
        for (int e = tid; e < 4 * A.nsites; e += NT) {
            const int si = e >> 2, a = e & 1, b = (e >> 1) & 1;
            A.onsL[(size_t)iw * 4 * A.nsites + e] = L.big[(2 * si + a) * pitch + 2 * si + b];
        }
        if (A.mode == QS_M_FINAL) {
            if (A.atomL) {
                for (int e = tid; e < A.atom_tot; e += NT) {
                    int at = 0;
                    while (at + 1 < A.n_atoms && A.atom_off[at + 1] <= e) at++;
                    const int m = 2 * A.atom_norb[at], loc = e - A.atom_off[at];
                    const int b = loc / m, a = loc - b * m, lo = 2 * A.atom_first[at];
                    A.atomL[(size_t)iw * A.atom_tot + e] = L.big[(lo + a) * pitch + lo + b];
                }
            }
            if (tid == 0) {
                // Tr_s{ i GamL [ (f_L - 1) G^< - f_L G^> ] } / 2pi on the left contact  (src/scf.f90:390-409)
                const int nc = A.nCL;
                double tu = 0.0, td = 0.0;
                for (int a = 0; a < nc; a++) {
                    double2 s = make_double2(0.0, 0.0);
                    for (int l = 0; l < nc; l++) {
                        const double2 glc = L.big[A.cl[l] * pitch + A.cl[a]];
                        const double2 gr_la = L.grcl[l * nc + a], gr_al = L.grcl[a * nc + l];
                        const double2 gg = cadd(glc, csub(gr_la, cconj(gr_al)));
                        const double2 X = make_double2((fL - 1.0) * glc.x - fL * gg.x, (fL - 1.0) * glc.y - fL * gg.y);
                        cfma(s, A.gamL_cl[a * nc + l], X);
                    }
                    // Mc(a,a) = i s
                    if ((A.cl[a] & 1) == 0) tu += -s.y; else td += -s.y;
                }
                A.trUD[(size_t)iw * 2 + 0] = tu / TWO_PI;
                A.trUD[(size_t)iw * 2 + 1] = td / TWO_PI;
            }
        }
    }
}

template <int TI, int RA, int RB>
hipError_t launch_v(const DysonArgs &a, int count, hipStream_t s) {
    if (hipError_t e = qs_lds_attr((const void *)k_dyson<TI, RA, RB>)) return e;
    const size_t lds = qs_dyson_lds_bytes(a.norb, a.nCL);
    hipLaunchKernelGGL((k_dyson<TI, RA, RB>), dim3(count), dim3(TI * 16), lds, s, a);
    return hipGetLastError();
}

}  // namespace

size_t qs_dyson_lds_bytes(int n, int nCL) {
    const int RB = (n + 15) / 16;
    size_t b = (size_t)n * (n + 1) * 16 + (size_t)16 * 16 * RB * 16;
    b += (size_t)4 * 16 * RB * 4 + 32 * 8 + 32 * 4 + 16 * 8;
    b = (b + 15) & ~(size_t)15;
    b += (size_t)nCL * nCL * 16;
    return b;
}

int qs_dyson_max_norb() { return 88; }

bool qs_dyson_shape_ok(int norb, int shape) {
    // k_dyson<16, 5, 5> (WG256 at 65..80 orbitals) is switched off: it returned
    // wrong init2 values at 80 orbitals and a GPU memory fault followed.  Its
    // source runs clean and correct in CPU builds under AddressSanitizer and
    // ThreadSanitizer (tests/test_dyson_emu_cpu.py), so no cause is known; it is
    // also the only shape the compiler gives 256 VGPRs + 94 AGPRs.  The other
    // shapes cover these sizes.
    if (shape == QS_DYSON_WG256) return norb <= qs_dyson_max_norb() && (norb + 15) / 16 != 5;
    if (shape == QS_DYSON_WG512) return norb > 16 && norb <= qs_dyson_max_norb();
    if (shape == QS_DYSON_WAVE) return norb <= qs_dyson_wave_max_norb();
    if (shape == QS_DYSON_SPLIT || shape == QS_DYSON_SPLIT_MFMA) return norb > 16 && norb <= qs_dyson_wave_max_norb();
    if (shape == QS_DYSON_BIG) return norb <= qs_dyson_big_max_norb();
    return false;
}

int qs_dyson_variants(int norb) { (void)norb; return QS_DYSON_NSHAPES; }

const char *qs_dyson_shape_desc(int shape, int *uses_mfma) {
    static const char *const names[QS_DYSON_NSHAPES] = {
        "256-thread workgroup, matrix in registers and LDS (FP64 vector FMA)",
        "512-thread workgroup, matrix in registers and LDS (FP64 vector FMA)",
        "16x16 workgroup, matrix in registers, scratch slab (FP64 vector FMA)",
        "large-Norb chunks through HBM: blocked Gauss-Jordan and products on FP64 MFMA",
        "16x16 workgroups, matrix in registers, scratch slab: Gauss-Jordan kernel (three energies per CU) "
        "then Keldysh kernel (FP64 vector FMA)",
        "16x16 workgroups, matrix in registers, scratch slab: Gauss-Jordan kernel (three energies per CU) "
        "then Keldysh kernel on FP64 MFMA (one wavefront per 16-column strip)"};
    static const int mfma[QS_DYSON_NSHAPES] = {0, 0, 0, 1, 0, 1};
    if (shape < 0 || shape >= QS_DYSON_NSHAPES) {
        if (uses_mfma) *uses_mfma = 0;
        return "";
    }
    if (uses_mfma) *uses_mfma = mfma[shape];
    return names[shape];
}

hipError_t qs_launch_dyson_variant(const DysonArgs &a, int count, int variant, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    if (!qs_dyson_shape_ok(a.norb, variant)) return hipErrorInvalidValue;
    if (variant == QS_DYSON_WAVE) return qs_launch_dyson_wave(a, count, 0, s);
    if (variant == QS_DYSON_SPLIT) return qs_launch_dyson_wave(a, count, 1, s);
    if (variant == QS_DYSON_SPLIT_MFMA) return qs_launch_dyson_wave(a, count, 2, s);
    if (variant == QS_DYSON_BIG) return qs_launch_dyson_big(a, count, s);
    if (qs_dyson_lds_bytes(a.norb, a.nCL) > 160 * 1024) return hipErrorInvalidValue;
    const int RB = (a.norb + 15) / 16;
    if (variant == QS_DYSON_WG256) {
        switch (RB) {
            case 1: return launch_v<16, 1, 1>(a, count, s);
            case 2: return launch_v<16, 2, 2>(a, count, s);
            case 3: return launch_v<16, 3, 3>(a, count, s);
            case 4: return launch_v<16, 4, 4>(a, count, s);
            case 6: return launch_v<16, 6, 6>(a, count, s);
            default: return hipErrorInvalidValue;
        }
    }
    const int RA = (a.norb + 31) / 32;
    switch (RB * 4 + RA) {
        case 2 * 4 + 1: return launch_v<32, 1, 2>(a, count, s);
        case 3 * 4 + 2: return launch_v<32, 2, 3>(a, count, s);
        case 4 * 4 + 2: return launch_v<32, 2, 4>(a, count, s);
        case 5 * 4 + 3: return launch_v<32, 3, 5>(a, count, s);
        case 6 * 4 + 3: return launch_v<32, 3, 6>(a, count, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t qs_launch_dyson(const DysonArgs &a, int count, hipStream_t s) {
    return qs_launch_dyson_variant(a, count, 0, s);
}
