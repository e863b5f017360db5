// qs_sigma2.hip -- second-order self-energy on the whole energy grid (FP64).
//
// Replaces sigma2_compute / load_fft / adj2 (src/self_energies.f90:116-228):
// per pair of active sites (i, j) and frequency l of the zero-padded DFT,
//   a_ij(l) = r_ji(-l)^dag,  G_ij = L_ij + r_ij - a_ij  (greater)
//   chi1 = L_ji(-l) G_ij(l),  chi2 = L_ji(-l) r_ij(l) + a_ji(-l) L_ij(l)
//   S^r_ij = c [ r_ij adj(chi1) + L_ij adj(chi2) ],  S^<_ij = c L_ij adj(G_ji(-l) L_ij),
//   adj(M) = Tr(M) 1 - M,  c = U_i U_j hbar^2 (delta/2pi)^2 / Nf,
// then an inverse DFT keeps the first n_omega points.  Zero padding to
// Nf >= 2 n_omega - 1 makes each term an exact linear convolution, so the result
// does not depend on Nf; this file uses the power of two Nf = 2^m (the reference
// uses the smallest 7-smooth length, good_fft_size).
//
// Data layout ("blocked"): per energy, 2 nA^2 complex values ordered as
//   c = ((ia * nAs + ja) * 2 + kind) * 4 + sa * 2 + sb
// (active sites ia, ja; kind 0 = G^r, 1 = G^<; spins sa, sb), so the 8 values a
// site pair needs are one 128-byte segment.  Gt and Sig use [omega][c]; the
// frequency-domain buffer Z uses [k1][n2][c].
//
// The length-Nf DFT is split as Nf = N1 * N2 (n = n2 + N2 n1, l = k1 + N1 k2):
//   F1: DFT over n1 (length N1) for every (c, n2), twiddle e^{-2 pi i k1 n2/Nf}
//   M : per site pair and row pair (k1, -k1): DFT over n2 (length N2), the 2x2
//       spin algebra at l and -l, inverse DFT over k2, twiddle e^{+2 pi i k1 n2/Nf}
//   I2: inverse DFT over k1 (length N1) for every (c, n2), keep n < n_omega.
// Each pass reads and writes HBM once in 128-byte segments; the DFTs run in LDS
// (radix-2, decimation in time, bit-reversed load).

#include "qs_device.h"
#include "qs_internal.h"

namespace {

constexpr int NT = 256;
constexpr int CB = 8;   // columns per block of the passes over n1 / k1
constexpr int BPW = 8;  // column blocks per workgroup in those passes

__device__ __forceinline__ int bitrev(int v, int bits) { return __brev((unsigned)v) >> (32 - bits); }

// In-place decimation-in-time DFT on NSEQ sequences of length L = 2^LG stored
// with stride ld(L) (complex), input in bit-reversed order; SIGN = -1 forward,
// +1 inverse; twl[tsw(m)] = e^{-2 pi i m / L} (m < L / 2), a table in LDS.
// The LG radix-2 levels are done in groups of up to three: one task loads the
// 8 elements j + t h (t < 8) of a block, runs the group's levels in registers
// and stores them back, so a length-512 transform takes 3 LDS sweeps and 3
// barriers.  The arithmetic is exactly that of the radix-2 levels.  Every size
// is a compile-time constant, so the element addresses of a task are one base
// plus immediate offsets.
//
// LDS banks (gfx950, measured with a one-wavefront probe of ds_read_b128 /
// ds_write_b128): a 16-byte read is served in four phases of 16 lanes (lanes
// 0-3, 12-15, 20-23, 24-27; the other 16 of lanes 0-31; the same two sets + 32)
// over 16 slots of 16 bytes, a 16-byte write in eight phases of 8 consecutive
// lanes over 8 slots; each further distinct slot that a phase puts on one bank
// costs a cycle.
// Layout: element x of sequence s sits at slot s * LD + (x ^ sx(x) ^ ss(s)).
// sx XORs a few shifted bit fields of x into its low four bits; every field
// comes from bits >= 3, and bits >= 4 where it changes bit 3, so x -> slot is
// one-to-one inside each aligned run of 16, and sx is linear over the bits of x.
// The fields, the sequence term and the pad of LD were chosen (by a search over
// these kernels' access patterns under the bank model above) so that the
// butterfly sweeps, the bit-reversed loads, the column-interleaved reads and
// the row accesses of the middle pass put each phase on distinct banks (the
// middle pass at N2 = 256 keeps about 64 conflict cycles per wavefront, against
// about 2300 cycles of LDS work).
struct LdsLayout { int a1, m1, a2, m2, a3, m3, k, pad; };
__host__ __device__ constexpr LdsLayout lds_layout(int LG, int NSEQ) {
    if (NSEQ == 12)  // the middle pass: 12 columns of one site pair
        return LG >= 8 ? LdsLayout{3, 7, 4, 11, 5, 7, 5, 0} : LdsLayout{3, 7, 6, 11, 0, 0, 1, 8};
    return LG == 7 ? LdsLayout{3, 7, 6, 3, 0, 0, 9, 0} : LdsLayout{3, 7, 4, 11, 0, 0, 1, 8};
}
template <int LG, int NSEQ>
struct Lay {
    static constexpr LdsLayout P = lds_layout(LG, NSEQ);
    static constexpr int L = 1 << LG, LD = L + P.pad, M = L < 16 ? L - 1 : 15;
    __host__ __device__ static constexpr int sx(int x) {
        return (((x >> P.a1) & P.m1) ^ ((x >> P.a2) & P.m2) ^ ((x >> P.a3) & P.m3)) & M;
    }
    __host__ __device__ static constexpr int ss(int s) { return (s * P.k) & M; }
    __host__ __device__ static constexpr int kx(int x) { return x ^ sx(x); }
    __host__ __device__ static constexpr int at(int s, int x) { return s * LD + (kx(x) ^ ss(s)); }
    static constexpr size_t bytes() { return ((size_t)L / 2 + (size_t)NSEQ * LD) * sizeof(double2); }
};
// x -> s LD + (kx(x) ^ ss(s)) maps [0, L) one-to-one onto a run of L slots
template <int LG, int NSEQ>
constexpr bool lay_one_to_one() {
    using Y = Lay<LG, NSEQ>;
    for (int s = 0; s < NSEQ; s++) {
        bool seen[Y::L] = {};
        for (int x = 0; x < Y::L; x++) {
            const int v = Y::kx(x) ^ Y::ss(s);
            if (v < 0 || v >= Y::L || seen[v]) return false;
            seen[v] = true;
        }
    }
    return true;
}
// The twiddle table is read at strides 2^k by lanes whose tasks differ in j;
// entry m sits at tsw(m), which keeps those reads on distinct banks.
__device__ __forceinline__ int tsw(int m) { return m ^ ((m >> 4) & 15); }

template <int G, int LG, int S0, int NSEQ, int SIGN>
__device__ __forceinline__ void fft_group(double2 *buf, const double2 *twl) {
    using Y = Lay<LG, NSEQ>;
    constexpr int T = 1 << G, LT = LG - G, LH = S0 - 1, H = 1 << LH;  // tasks per sequence 2^LT, element stride H
    constexpr int TOTAL = NSEQ << LT;
#pragma unroll 1
    for (int task = threadIdx.x; task < TOTAL; task += NT) {
        const int seq = task >> LT, r = task & ((1 << LT) - 1);
        const int blk = r >> LH, j = r & (H - 1);
        // x0 and t H have no bit in common and sx is linear, so element x0 + t H
        // sits at seq LD + (o ^ kx(t H))
        const int x0 = (blk << (LH + G)) + j;
        // byte offsets, so that each element address is one XOR (and add)
        const int o16 = (Y::kx(x0) ^ Y::ss(seq)) << 4;
        char *b = (char *)(buf + seq * Y::LD);
        auto el = [&](int t) -> double2 & { return *(double2 *)(b + (o16 ^ (Y::kx(t << LH) << 4))); };
        double2 v[T];
#pragma unroll
        for (int t = 0; t < T; t++) v[t] = el(t);
#pragma unroll
        for (int e = 0; e < G; e++) {
            // radix-2 level S0 + e, half-length H 2^e: w = e^{-2 pi i m / (2 H 2^e)} = twl[m L / (2 H 2^e)]
            const int tws = LG - 1 - LH - e;
#pragma unroll
            for (int t = 0; t < T; t++)
                if (!(t & (1 << e))) {
                    const int m = j + ((t & ((1 << e) - 1)) << LH);  // position inside the level's group
                    const double2 u = v[t], y = v[t + (1 << e)];
                    // first sweep (H = 1, j = 0): m is a compile-time constant and the
                    // angle 2 pi m / 2^(e+1) a multiple of pi / 4; w = 1 and w = -/+ i
                    // need only additions
                    const int q8 = S0 == 1 ? (m << (2 - e)) & 7 : -1;  // angle in units of pi / 4
                    if (q8 == 0) {
                        v[t] = make_double2(u.x + y.x, u.y + y.y);
                        v[t + (1 << e)] = make_double2(u.x - y.x, u.y - y.y);
                        continue;
                    }
                    if (q8 == 2) {  // w = -i (forward), +i (inverse)
                        const double2 wy = SIGN < 0 ? make_double2(y.y, -y.x) : make_double2(-y.y, y.x);
                        v[t] = make_double2(u.x + wy.x, u.y + wy.y);
                        v[t + (1 << e)] = make_double2(u.x - wy.x, u.y - wy.y);
                        continue;
                    }
                    double2 w;
                    if (q8 > 0) {  // e^{-i pi q8 / 4}, q8 = 1 or 3
                        constexpr double R = 0.70710678118654752440;
                        w = make_double2(q8 == 1 ? R : -R, -R);
                    } else {
                        w = twl[tsw(m << tws)];
                    }
                    if (SIGN > 0) w.y = -w.y;
                    // (u + w y, u - w y): four FMAs for the sum, the difference as 2u - sum
                    const double sr = fma(w.x, y.x, fma(-w.y, y.y, u.x));
                    const double si = fma(w.x, y.y, fma(w.y, y.x, u.y));
                    v[t] = make_double2(sr, si);
                    v[t + (1 << e)] = make_double2(fma(2.0, u.x, -sr), fma(2.0, u.y, -si));
                }
        }
#pragma unroll
        for (int t = 0; t < T; t++) el(t) = v[t];
    }
    __syncthreads();
}

template <int LG, int NSEQ, int SIGN, int S0 = 1>
__device__ __forceinline__ void lds_fft(double2 *buf, const double2 *twl) {
    constexpr int REM = LG - S0 + 1;
    constexpr int G = REM >= 3 ? 3 : REM;
    if constexpr (REM >= 1) fft_group<G, LG, S0, NSEQ, SIGN>(buf, twl);
    if constexpr (REM > 3) lds_fft<LG, NSEQ, SIGN, S0 + 3>(buf, twl);
}

// twl[tsw(m)] = e^{-2 pi i m / L} for m < L / 2, from the length-Nf table
template <int L>
__device__ __forceinline__ void load_twl(double2 *twl, const double2 *tw, int Nf) {
    for (int m = threadIdx.x; m < L / 2; m += NT) twl[tsw(m)] = tw[m * (Nf / L)];
}

// ---- x (reference packed layout) -> Gt[omega][c] (blocked) ----------------------
// One workgroup per energy: the gather of an energy's packed slice (3 nA^2
// doubles) stays inside one XCD's L2, so each line of x is fetched once.
constexpr int PU = 8;   // columns per thread per batch
constexpr int EPW = 4;  // energies per workgroup: one read of the column map serves all of them
__global__ __launch_bounds__(NT) void k_permute_in(const double *__restrict__ x, double2 *__restrict__ Gt,
                                                   const int2 *__restrict__ src, int nA, int ncol, int nw) {
    const size_t nblk = 3 * (size_t)nA * nA;
    for (int iw0 = blockIdx.x * EPW; iw0 < nw; iw0 += gridDim.x * EPW) {
        for (int c0 = threadIdx.x; c0 < ncol; c0 += PU * NT) {
            // column c: real part from +/- x[|sc.x| - 1] (0: zero), imaginary part from x[sc.y] (< 0: padding)
            int ir[PU], ii[PU], kr[PU], ki[PU];  // kr: 1 copy, -1 negate, 0 zero; ki: 1 copy, 0 zero
#pragma unroll
            for (int u = 0; u < PU; u++) {
                const int2 sc = src[min(c0 + u * NT, ncol - 1)];
                ir[u] = sc.x > 0 ? sc.x - 1 : (sc.x < 0 ? -sc.x - 1 : 0);
                kr[u] = sc.x > 0 ? 1 : (sc.x < 0 ? -1 : 0);
                ii[u] = sc.y >= 0 ? sc.y : 0;
                ki[u] = sc.y >= 0;
            }
            double re[EPW][PU], im[EPW][PU];
#pragma unroll
            for (int e = 0; e < EPW; e++) {
                const double *xb = x + (size_t)min(iw0 + e, nw - 1) * nblk;
#pragma unroll
                for (int u = 0; u < PU; u++) {
                    re[e][u] = xb[ir[u]];
                    im[e][u] = xb[ii[u]];
                }
            }
#pragma unroll
            for (int e = 0; e < EPW; e++) {
                if (iw0 + e >= nw) break;
                double2 *g = Gt + (size_t)(iw0 + e) * ncol;
#pragma unroll
                for (int u = 0; u < PU; u++) {
                    const int c = c0 + u * NT;
                    const double r = kr[u] == 0 ? 0.0 : (kr[u] < 0 ? -re[e][u] : re[e][u]);
                    if (c < ncol) g[c] = make_double2(r, ki[u] ? im[e][u] : 0.0);
                }
            }
        }
    }
}

// Same gather with the energy's slice staged in LDS (3 nA^2 doubles; nA <= 82
// fits one CU's 160 KB): one 1024-thread workgroup per CU, persistent over the
// energies.  The slice comes in with 16-byte loads that read each line of x
// once, the gather runs from LDS, and Gt rows go out as contiguous 16-byte
// stores.  The next energy's loads are in flight during the current gather;
// each thread keeps its columns' source pairs in registers for every energy.
constexpr int PL_NT = 1024, PL_QL = 10, PL_QC = 11;  // slice <= PL_QL NT double2, ncol <= PL_QC NT
// this thread's columns c = tid + PL_NT u of the site-pair layout: their source pairs
__device__ __forceinline__ void load_colmap(const int2 *__restrict__ src, int2 (&sc)[PL_QC], int ncol, int tid) {
#pragma unroll
    for (int u = 0; u < PL_QC; u++) {
        const int c = tid + PL_NT * u;
        sc[u] = c < ncol ? src[c] : make_int2(0, -1);
    }
}
// one energy's Gt row from its packed slice xs (in LDS): column c takes its real
// part from +/- xs[|sc.x| - 1] (sc.x = 0: zero) and its imaginary part from xs[sc.y]
// (sc.y < 0: zero); the row goes out as contiguous 16-byte stores
__device__ __forceinline__ void gather_gt(const double *xs, const int2 (&sc)[PL_QC], double2 *__restrict__ g, int ncol,
                                          int tid) {
#pragma unroll
    for (int u = 0; u < PL_QC; u++) {
        const int c = tid + PL_NT * u;
        const int ir = sc[u].x > 0 ? sc[u].x - 1 : (sc[u].x < 0 ? -sc[u].x - 1 : 0);
        const int ii = sc[u].y >= 0 ? sc[u].y : 0;
        const double xr = xs[ir], xi = xs[ii];
        const double re = sc[u].x > 0 ? xr : (sc[u].x < 0 ? -xr : 0.0);
        if (c < ncol) g[c] = make_double2(re, sc[u].y >= 0 ? xi : 0.0);
    }
}
__global__ __launch_bounds__(PL_NT) void k_permute_lds(const double *__restrict__ x, double2 *__restrict__ Gt,
                                                       const int2 *__restrict__ src, int nA, int ncol, int nw) {
    extern __shared__ __align__(16) char lds_raw[];
    double *xs = (double *)lds_raw;
    const int tid = threadIdx.x, nh = 3 * nA * nA / 2;  // double2 per energy (nA is even)
    int2 sc[PL_QC];
    load_colmap(src, sc, ncol, tid);
    double2 v[PL_QL];
    auto fetch = [&](int iw) {
        const double2 *xb = (const double2 *)x + (size_t)min(iw, nw - 1) * nh;
#pragma unroll
        for (int q = 0; q < PL_QL; q++) {
            const int i = tid + PL_NT * q;
            v[q] = i < nh ? xb[i] : make_double2(0.0, 0.0);
        }
    };
    fetch(blockIdx.x);
    for (int iw = blockIdx.x; iw < nw; iw += gridDim.x) {
#pragma unroll
        for (int q = 0; q < PL_QL; q++) {
            const int i = tid + PL_NT * q;
            if (i < nh) ((double2 *)xs)[i] = v[q];
        }
        __syncthreads();
        if (iw + (int)gridDim.x < nw) fetch(iw + gridDim.x);
        gather_gt(xs, sc, Gt + (size_t)iw * ncol, ncol, tid);
        __syncthreads();  // the next slice overwrites xs
    }
}

// Broyden update fused with the gather of the next residual's FFT input:
// x_new = x + alpha F - sum_k gamma_k U_k over the FP32 history (the arithmetic
// of the plain update kernel, so x_new is the same to the bit), computed on whole
// energy slices; each slice also goes to LDS and out to Gt as k_permute_lds
// writes it.  With dx, x_new - x also goes out: the next history pass reads it
// in place of x and xold (the same FP64 difference).  One workgroup per CU,
// persistent over the energies.
// ctl (not null): gamma and nh from the control block; nothing when the step
// stopped, and no Gt once its >= maxscf
struct GamH { double g[16]; };
__global__ __launch_bounds__(PL_NT) void k_update_f32_gt(int64_t n, const double *__restrict__ x,
                                                         const double *__restrict__ F, const float *__restrict__ Us,
                                                         GamH gm, int nh, double alpha, double *__restrict__ xnew,
                                                         double *__restrict__ dx, double2 *__restrict__ Gt,
                                                         const int2 *__restrict__ src, int nA, int ncol, int nw,
                                                         const QsCtrl *__restrict__ ctl, int maxscf) {
    if (ctl) {
        if (ctl->stop) return;
        nh = ctl->nh_use;
#pragma unroll
        for (int k = 0; k < 16; k++) gm.g[k] = k < nh ? ctl->gamma[k] : 0.0;
        if (ctl->its >= maxscf) Gt = nullptr;
    }
    extern __shared__ __align__(16) char lds_raw[];
    double *xs = (double *)lds_raw;
    const int tid = threadIdx.x, nb = 3 * nA * nA, nq = nb / 4;  // 4 | nb (nA is even)
    for (int iw = blockIdx.x; iw < nw; iw += gridDim.x) {
        // the energy's slices from uniform bases, lane offsets 32-bit
        const int64_t base = (int64_t)iw * nb;
        const double4 *xb = (const double4 *)(x + base), *fb = (const double4 *)(F + base);
        double4 *xnb = (double4 *)(xnew + base);
        for (int q = tid; q < nq; q += PL_NT) {
            const double4 xv = xb[q];
            const double4 f = fb[q];
            float4 hv[16];
#pragma unroll
            for (int k = 0; k < 16; k++)
                if (k < nh) hv[k] = ((const float4 *)(Us + (size_t)k * n + base))[q];
            double v0 = xv.x + alpha * f.x, v1 = xv.y + alpha * f.y, v2 = xv.z + alpha * f.z, v3 = xv.w + alpha * f.w;
#pragma unroll
            for (int k = 0; k < 16; k++)
                if (k < nh) {
                    v0 = v0 - (double)hv[k].x * gm.g[k];
                    v1 = v1 - (double)hv[k].y * gm.g[k];
                    v2 = v2 - (double)hv[k].z * gm.g[k];
                    v3 = v3 - (double)hv[k].w * gm.g[k];
                }
            xnb[q] = make_double4(v0, v1, v2, v3);
            if (dx) ((double4 *)(dx + base))[q] = make_double4(v0 - xv.x, v1 - xv.y, v2 - xv.z, v3 - xv.w);
            ((double2 *)xs)[2 * q] = make_double2(v0, v1);
            ((double2 *)xs)[2 * q + 1] = make_double2(v2, v3);
        }
        __syncthreads();
        if (Gt) {
            // the column map comes from L2 each time: registers go to the history loads
            int2 sc[PL_QC];
            load_colmap(src, sc, ncol, tid);
            gather_gt(xs, sc, Gt + (size_t)iw * ncol, ncol, tid);
        }
        __syncthreads();  // the next slice overwrites xs
    }
}

// ---- F1 -----------------------------------------------------------------------
// grid: (ncol / (CB BPW), N2).  LDS: twiddles (N1 / 2) | CB padded sequences.
// Thread work is unrolled over constant bounds (QS = CB N1 / NT elements per
// thread, QI of them nonzero inputs) so that every global load of a round is in
// flight at once.
template <int LG1>
__global__ __launch_bounds__(NT) void k_fft_f1(const double2 *__restrict__ Gt, double2 *__restrict__ Z,
                                               Sigma2Plan p) {
    extern __shared__ __align__(16) char lds_raw[];
    using Y = Lay<LG1, CB>;
    constexpr int N1 = 1 << LG1, HALF = CB * N1 / 2;
    constexpr int QS = (CB * N1 + NT - 1) / NT, QI = (HALF + NT - 1) / NT;
    double2 *twl = (double2 *)lds_raw, *buf = twl + N1 / 2;
    const int ncol = p.ncol, n2 = blockIdx.y, nblk = ncol / CB;
    const int cb0 = blockIdx.x * BPW, cb1 = min(cb0 + BPW, nblk);
    load_twl<N1>(twl, p.tw, p.Nf);
    // rows n1 >= N1 / 2 are the zero padding: the inputs of a block are QI rounds
    auto fetch = [&](int cb, double2 *v) {
#pragma unroll
        for (int q = 0; q < QI; q++) {
            const int t = threadIdx.x + q * NT, n = n2 + p.N2 * (t >> 3);
            v[q] = make_double2(0.0, 0.0);
            if (t < HALF && n < p.nw) v[q] = Gt[(size_t)n * ncol + cb * CB + (t & 7)];
        }
    };
    double2 v[QI];
    fetch(cb0, v);
    // BPW column blocks of row n2; the next block's inputs load during this one's transform
    for (int cb = cb0; cb < cb1; cb++) {
#pragma unroll
        for (int q = 0; q < QS; q++) {
            const int t = threadIdx.x + q * NT;
            if (t < CB * N1) buf[Y::at(t & 7, bitrev(t >> 3, LG1))] = q < QI ? v[q] : make_double2(0.0, 0.0);
        }
        __syncthreads();
        // the next block's loads go out at raised wave priority, ahead of the other
        // workgroup's transform work (in-box A/B: F1 and I2 each about 0.25 ms faster)
        __builtin_amdgcn_s_setprio(2);
        if (cb + 1 < cb1) fetch(cb + 1, v);
        __builtin_amdgcn_s_setprio(0);
        lds_fft<LG1, CB, -1>(buf, twl);
        // output twiddles e^{-2 pi i k1 n2 / Nf}, k1 = (tid >> 3) + (NT / 8) i (cached: the same for every block)
        double2 w[QS];
#pragma unroll
        for (int i = 0; i < QS; i++) {
            const int k1 = (threadIdx.x >> 3) + (NT / 8) * i;
            w[i] = k1 < N1 ? p.tw[k1 * n2] : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int i = 0; i < QS; i++) {
            const int t = threadIdx.x + i * NT;
            if (t < CB * N1) {
                const int col = t & 7, k1 = t >> 3;
                Z[((size_t)k1 * p.N2 + n2) * ncol + cb * CB + col] = cmul(buf[Y::at(col, k1)], w[i]);
            }
        }
        __syncthreads();  // the next block overwrites buf
    }
}

// ---- M ------------------------------------------------------------------------
// a0 b0 + a1 b1 (complex) as one multiply and three FMAs per real component
__device__ __forceinline__ double2 cdot2(double2 a0, double2 b0, double2 a1, double2 b1) {
    return make_double2(fma(a0.x, b0.x, fma(-a0.y, b0.y, fma(a1.x, b1.x, -a1.y * b1.y))),
                        fma(a0.x, b0.y, fma(a0.y, b0.x, fma(a1.x, b1.y, a1.y * b1.x))));
}
__device__ __forceinline__ void mm2(const double2 *A, const double2 *B, double2 *C) {
    // 2x2 row-major: [0]=11 [1]=12 [2]=21 [3]=22
    C[0] = cdot2(A[0], B[0], A[1], B[2]);
    C[1] = cdot2(A[0], B[1], A[1], B[3]);
    C[2] = cdot2(A[2], B[0], A[3], B[2]);
    C[3] = cdot2(A[2], B[1], A[3], B[3]);
}
__device__ __forceinline__ void adj2(const double2 *M, double2 *A) {
    A[0] = M[3];
    A[1] = make_double2(-M[1].x, -M[1].y);
    A[2] = make_double2(-M[2].x, -M[2].y);
    A[3] = M[0];
}
__device__ __forceinline__ void dag2(const double2 *M, double2 *A) {
    A[0] = cconj(M[0]);
    A[1] = cconj(M[2]);
    A[2] = cconj(M[1]);
    A[3] = cconj(M[3]);
}

// Sigma^r, Sigma^< of the ordered pair (i, j) at +l from blocks at +l (r_ij, L_ij)
// and at -l (r_ji, L_ji); src/self_energies.f90:175-198.
__device__ __forceinline__ void pair_sigma(const double2 *r_ij, const double2 *L_ij, const double2 *r_ji,
                                           const double2 *L_ji, double c, double2 *sR, double2 *sL) {
    double2 a_ij[4], a_ji[4], G_ij[4], G_ji[4], chi1[4], chi2[4], t1[4], t2[4], ad1[4], ad2[4];
    dag2(r_ji, a_ij);
    dag2(r_ij, a_ji);
#pragma unroll
    for (int q = 0; q < 4; q++) {
        G_ij[q] = csub(cadd(L_ij[q], r_ij[q]), a_ij[q]);
        G_ji[q] = csub(cadd(L_ji[q], r_ji[q]), a_ji[q]);
    }
    mm2(L_ji, G_ij, chi1);
    mm2(L_ji, r_ij, t1);
    mm2(a_ji, L_ij, t2);
#pragma unroll
    for (int q = 0; q < 4; q++) chi2[q] = cadd(t1[q], t2[q]);
    adj2(chi1, ad1);
    adj2(chi2, ad2);
    mm2(r_ij, ad1, t1);
    mm2(L_ij, ad2, t2);
#pragma unroll
    for (int q = 0; q < 4; q++) sR[q] = cscale(cadd(t1[q], t2[q]), c);
    mm2(G_ji, L_ij, chi1);
    adj2(chi1, ad1);
    mm2(L_ij, ad1, t1);
#pragma unroll
    for (int q = 0; q < 4; q++) sL[q] = cscale(t1[q], c);
}

// Middle pass with one row of the (k1, -k1) pair held in registers: LDS holds
// a single row of the pair's 12 columns (12 padded sequences, 55 KB at
// N2 = 256) and the row-DFT twiddles, so two workgroups share a CU.  Thread t
// owns position t of row k1 and, for the spin algebra, the partner position
// (-k1, N2-1-t); both outputs are computed by the same thread from the same
// two inputs.
//
// The lesser block (J, I) is not among the columns: G^< is anti-Hermitian at
// every energy, G^<_JI(w) = -G^<_IJ(w)^dagger, so after the forward transform
// L_JI(l) = -L_IJ(-l)^dagger, which the thread has at the partner position.
// Likewise only Sigma^<_IJ is stored.
constexpr int PS = 12;  // columns per site pair

__device__ __forceinline__ void neg_dag2(const double2 *M, double2 *A) {  // A = -M^dagger (2x2 row-major)
    A[0] = make_double2(-M[0].x, M[0].y);
    A[1] = make_double2(-M[2].x, M[2].y);
    A[2] = make_double2(-M[1].x, M[1].y);
    A[3] = make_double2(-M[3].x, M[3].y);
}

template <int LG2>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(2, 8))) void k_fft_mid(double2 *__restrict__ Z,
                                                                                          Sigma2Plan p) {
    extern __shared__ __align__(16) char lds_raw[];
    using Y = Lay<LG2, PS>;
    constexpr int N2 = 1 << LG2;  // N2 <= NT
    const int N1 = p.N1, nAs = p.nAs, ncol = p.ncol;
    double2 *twl = (double2 *)lds_raw, *buf = twl + N2 / 2;  // buf: 12 sequences of N2, slots Y::at(s, x)
    // Work item w = (k1, pair), pair fastest.  Workgroup L of the dispatch runs on
    // XCD L % 8, so the items are dealt out in 8 contiguous ranges: neighbouring
    // pairs, whose 192-byte column groups share 128-byte lines, then meet in one
    // L2.  The tail that does not fill 8 equal ranges keeps w = L.
    const int npairs = gridDim.x, total = npairs * gridDim.y, L = blockIdx.y * npairs + blockIdx.x;
    const int per = total / 8;
    const int w = L < 8 * per ? (L % 8) * per + L / 8 : L;
    const int pair = w % npairs;
    int pid = pair, ia = 0;
    while (pid >= nAs - ia) { pid -= nAs - ia; ia++; }
    const int ja = ia + pid;
    const int k1 = w / npairs, k1p = (N1 - k1) & (N1 - 1);
    const int cb = PS * pair;  // the pair's columns (pair = sig_pair(ia, ja))
    const double c = p.Upair[ia * nAs + ja];
    if (c == 0.0) {
        const int nrow = (k1 == k1p) ? 1 : 2;
        for (int t = threadIdx.x; t < PS * nrow * N2; t += NT) {
            const int s = t % PS, rest = t / PS, n2 = rest % N2, rr = rest / N2;
            const int row = rr ? k1p : k1;
            Z[((size_t)row * N2 + n2) * ncol + cb + s] = make_double2(0.0, 0.0);
        }
        return;
    }
    const int t = threadIdx.x;
    load_twl<N2>(twl, p.tw, p.Nf);
    // A row is PS N2 <= PS NT values: one round of loads into registers, then
    // into LDS at bit-reversed positions.  Row k1p is fetched while row k1 is
    // transformed.  The row goes in two passes of power-of-two shape: columns
    // 0..7 with 8 lanes per n2, then columns 8..11 with 4 lanes per n2.  A lane
    // keeps its column, and its n2 = n2a + (NT / 8) u (or n2b + (NT / 4) u) has
    // the lane part and the step in disjoint bits; bit reversal and the LDS
    // swizzle are linear over those bits, so every slot is a per-lane base XOR a
    // compile-time constant, and every global offset a per-lane base plus a
    // uniform step.
    constexpr int NA = (8 * N2 + NT - 1) / NT, NB = (4 * N2 + NT - 1) / NT, NE = NA + NB;  // NE <= PS
    const int sa = t & 7, n2a = t >> 3, sb = 8 + (t & 3), n2b = t >> 2;
    const int oga = n2a * ncol + sa, ogb = n2b * ncol + sb;  // offsets in a row of Z from its column base
    const int lra = sa * Y::LD * 16, lrb = sb * Y::LD * 16;  // byte offsets of the lane's sequence in buf
    const int pba = (Y::kx(bitrev(n2a, LG2)) ^ Y::ss(sa)) << 4, pbb = (Y::kx(bitrev(n2b, LG2)) ^ Y::ss(sb)) << 4;
    const int nba = (Y::kx(n2a) ^ Y::ss(sa)) << 4, nbb = (Y::kx(n2b) ^ Y::ss(sb)) << 4;
    char *bufc = (char *)buf;
    auto e_n2 = [&](int u) { return u < NA ? n2a + (NT / 8) * u : n2b + (NT / 4) * (u - NA); };
    auto e_og = [&](int u) { return u < NA ? oga + (NT / 8) * u * ncol : ogb + (NT / 4) * (u - NA) * ncol; };
    // slot of element u at bit-reversed (put) or natural (store) position
    auto e_put = [&](int u) -> double2 & {
        return u < NA ? *(double2 *)(bufc + lra + (pba ^ (Y::kx(bitrev((NT / 8) * u, LG2)) << 4)))
                      : *(double2 *)(bufc + lrb + (pbb ^ (Y::kx(bitrev((NT / 4) * (u - NA), LG2)) << 4)));
    };
    auto e_nat = [&](int u) -> double2 & {
        return u < NA ? *(double2 *)(bufc + lra + (nba ^ (Y::kx((NT / 8) * u) << 4)))
                      : *(double2 *)(bufc + lrb + (nbb ^ (Y::kx((NT / 4) * (u - NA)) << 4)));
    };
    double2 pre[NE];
    auto fetch_row = [&](int row) {
        const double2 *Zr = Z + (size_t)row * N2 * ncol + cb;
#pragma unroll
        for (int u = 0; u < NE; u++) {
            pre[u] = make_double2(0.0, 0.0);
            if (e_n2(u) < N2) pre[u] = Zr[e_og(u)];
        }
    };
    auto put_row = [&]() {
#pragma unroll
        for (int u = 0; u < NE; u++)
            if (e_n2(u) < N2) e_put(u) = pre[u];
    };
    // twiddle e^{+2 pi i row n2 / Nf} between the inverse row DFT and the
    // inverse pass over k1
    auto store_row = [&](int row) {
        double2 w[NE];
#pragma unroll
        for (int u = 0; u < NE; u++) w[u] = e_n2(u) < N2 ? cconj(p.tw[row * e_n2(u)]) : make_double2(0.0, 0.0);
        double2 *Zr = Z + (size_t)row * N2 * ncol + cb;
#pragma unroll
        for (int u = 0; u < NE; u++)
            if (e_n2(u) < N2) Zr[e_og(u)] = cmul(e_nat(u), w[u]);
    };
    // outputs at position x from its values X (at +l) and the partner's Y (at -l):
    // [Sigma^r_IJ, Sigma^<_IJ, Sigma^r_JI], src/self_energies.f90:175-198
    auto pair_out = [&](const double2 *X, const double2 *Y, double2 *O) {
        double2 LjiY[4], LjiX[4], scratch[4];
        neg_dag2(X + 4, LjiY);  // L_JI at -l
        neg_dag2(Y + 4, LjiX);  // L_JI at +l
        pair_sigma(X + 0, X + 4, Y + 8, LjiY, c, O + 0, O + 4);
        pair_sigma(X + 8, LjiX, Y + 0, Y + 4, c, O + 8, scratch);
    };
    fetch_row(k1);
    put_row();
    __syncthreads();
    if (k1 != k1p) fetch_row(k1p);
    lds_fft<LG2, PS, -1>(buf, twl);
    if (k1 != k1p) {
        double2 A[PS];
        if (t < N2)
#pragma unroll
            for (int s = 0; s < PS; s++) A[s] = buf[Y::at(s, t)];
        __syncthreads();
        put_row();
        __syncthreads();
        lds_fft<LG2, PS, -1>(buf, twl);
        const int m2 = N2 - 1 - t;  // partner of (k1, t) is (k1p, N2 - 1 - t)
        double2 Q[PS];
        if (t < N2)
#pragma unroll
            for (int s = 0; s < PS; s++) Q[s] = buf[Y::at(s, m2)];
        __syncthreads();
        if (t < N2) {
            double2 O[PS];
            pair_out(Q, A, O);
            const int bm = bitrev(m2, LG2);
#pragma unroll
            for (int s = 0; s < PS; s++) buf[Y::at(s, bm)] = O[s];
            pair_out(A, Q, O);
#pragma unroll
            for (int s = 0; s < PS; s++) A[s] = O[s];
        }
        __syncthreads();
        lds_fft<LG2, PS, +1>(buf, twl);
        store_row(k1p);
        __syncthreads();
        if (t < N2)
#pragma unroll
            for (int s = 0; s < PS; s++) buf[Y::at(s, bitrev(t, LG2))] = A[s];
        __syncthreads();
    } else {
        // k1 = 0 or N1/2: the partner position is in the same row
        const int m2 = (k1 == 0) ? ((N2 - t) & (N2 - 1)) : (N2 - 1 - t);
        double2 O[PS];
        if (t < N2) {
            double2 P[PS], Q[PS];
#pragma unroll
            for (int s = 0; s < PS; s++) { P[s] = buf[Y::at(s, t)]; Q[s] = buf[Y::at(s, m2)]; }
            pair_out(P, Q, O);
        }
        __syncthreads();
        if (t < N2)
#pragma unroll
            for (int s = 0; s < PS; s++) buf[Y::at(s, bitrev(t, LG2))] = O[s];
        __syncthreads();
    }
    lds_fft<LG2, PS, +1>(buf, twl);
    store_row(k1);
}

// ---- I2 -----------------------------------------------------------------------
// grid: (ncol / (CB BPW), N2).  LDS as in F1.  Only the outputs n < nw are kept.
template <int LG1>
__global__ __launch_bounds__(NT) void k_fft_i2(const double2 *__restrict__ Z, double2 *__restrict__ Sig,
                                               Sigma2Plan p) {
    extern __shared__ __align__(16) char lds_raw[];
    using Y = Lay<LG1, CB>;
    constexpr int N1 = 1 << LG1, QS = (CB * N1 + NT - 1) / NT;
    double2 *twl = (double2 *)lds_raw, *buf = twl + N1 / 2;
    const int ncol = p.ncol, n2 = blockIdx.y, nblk = ncol / CB;
    const int cb0 = blockIdx.x * BPW, cb1 = min(cb0 + BPW, nblk);
    load_twl<N1>(twl, p.tw, p.Nf);
    auto fetch = [&](int cb, double2 *v) {
#pragma unroll
        for (int q = 0; q < QS; q++) {
            const int t = threadIdx.x + q * NT;
            v[q] = make_double2(0.0, 0.0);
            if (t < CB * N1) v[q] = Z[((size_t)(t >> 3) * p.N2 + n2) * ncol + cb * CB + (t & 7)];
        }
    };
    double2 v[QS];
    fetch(cb0, v);
    // BPW column blocks of row n2; the next block's inputs load during this one's transform
    for (int cb = cb0; cb < cb1; cb++) {
#pragma unroll
        for (int q = 0; q < QS; q++) {
            const int t = threadIdx.x + q * NT;
            if (t < CB * N1) buf[Y::at(t & 7, bitrev(t >> 3, LG1))] = v[q];
        }
        __syncthreads();
        // the next block's loads go out at raised wave priority, ahead of the other
        // workgroup's transform work (in-box A/B: F1 and I2 each about 0.25 ms faster)
        __builtin_amdgcn_s_setprio(2);
        if (cb + 1 < cb1) fetch(cb + 1, v);
        __builtin_amdgcn_s_setprio(0);
        lds_fft<LG1, CB, +1>(buf, twl);
#pragma unroll
        for (int i = 0; i < QS; i++) {
            const int t = threadIdx.x + i * NT;
            const int col = t & 7, n1 = t >> 3, n = n2 + p.N2 * n1;
            if (t < CB * N1 && n < p.nw) Sig[(size_t)n * ncol + cb * CB + col] = buf[Y::at(col, n1)];
        }
        __syncthreads();  // the next block overwrites buf
    }
}

}  // namespace

namespace {
// an energy's packed slice fits the LDS-staged gather
bool slice_fits_lds(int nA, int ncol) {
    const size_t slice = (size_t)3 * nA * nA * sizeof(double);
    return nA % 2 == 0 && slice <= 160 * 1024 && 3 * nA * nA / 2 <= PL_QL * PL_NT && ncol <= PL_QC * PL_NT;
}
// compute units of the current device (0 on failure); the persistent kernels take one workgroup per CU
int cu_count() {
    static int ncu = 0;
    if (!ncu) {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
            ncu = v;
    }
    return ncu;
}
}  // namespace

bool qs_update_gt_fits(int nA, int ncol, int64_t n, int nh) { return slice_fits_lds(nA, ncol) && n % 4 == 0 && nh <= 16; }

hipError_t qs_launch_broyden_update_f32_gt(int64_t n, const double *x, const double *F, const float *Us,
                                           const double *gamma, int nh, double alpha, double *xnew, double *dx,
                                           double2 *Gt, const int *blk_src, int nA, int ncol, int nw, hipStream_t s) {
    if (!qs_update_gt_fits(nA, ncol, n, nh) || (int64_t)3 * nA * nA * nw != n) return hipErrorInvalidValue;
    const int ncu = cu_count();
    if (ncu <= 0) return hipErrorInvalidValue;
    if (hipError_t e = qs_lds_attr((const void *)k_update_f32_gt)) return e;
    GamH gm;
    for (int k = 0; k < 16; k++) gm.g[k] = k < nh ? gamma[k] : 0.0;
    const int blocks = nw < ncu ? nw : ncu;
    hipLaunchKernelGGL(k_update_f32_gt, dim3(blocks), dim3(PL_NT), (size_t)3 * nA * nA * sizeof(double), s, n, x, F,
                       Us, gm, nh, alpha, xnew, dx, Gt, (const int2 *)blk_src, nA, ncol, nw, (const QsCtrl *)nullptr, 0);
    return hipGetLastError();
}

hipError_t qs_launch_broyden_update_f32_gt_ctl(int64_t n, const double *x, const double *F, const float *Us,
                                               const QsCtrl *c, int maxscf, double alpha, double *xnew, double *dx,
                                               double2 *Gt, const int *blk_src, int nA, int ncol, int nw,
                                               hipStream_t s) {
    if (!c || !qs_update_gt_fits(nA, ncol, n, 16) || (int64_t)3 * nA * nA * nw != n) return hipErrorInvalidValue;
    const int ncu = cu_count();
    if (ncu <= 0) return hipErrorInvalidValue;
    if (hipError_t e = qs_lds_attr((const void *)k_update_f32_gt)) return e;
    GamH gm{};
    const int blocks = nw < ncu ? nw : ncu;
    hipLaunchKernelGGL(k_update_f32_gt, dim3(blocks), dim3(PL_NT), (size_t)3 * nA * nA * sizeof(double), s, n, x, F,
                       Us, gm, 0, alpha, xnew, dx, Gt, (const int2 *)blk_src, nA, ncol, nw, c, maxscf);
    return hipGetLastError();
}

hipError_t qs_launch_permute_in(const double *x, double2 *Gt, const int *blk_src, int nA, int ncol, int nw,
                                hipStream_t s) {
    const size_t slice = (size_t)3 * nA * nA * sizeof(double);
    if (slice_fits_lds(nA, ncol)) {
        const int ncu = cu_count();
        if (ncu <= 0) return hipErrorInvalidValue;
        if (hipError_t e = qs_lds_attr((const void *)k_permute_lds)) return e;
        const int blocks = nw < ncu ? nw : ncu;  // one per CU (the slice takes most of its LDS)
        hipLaunchKernelGGL(k_permute_lds, dim3(blocks), dim3(PL_NT), slice, s, x, Gt, (const int2 *)blk_src, nA, ncol,
                           nw);
        return hipGetLastError();
    }
    const int groups = (nw + EPW - 1) / EPW, blocks = groups < 65536 ? groups : 65536;
    hipLaunchKernelGGL(k_permute_in, dim3(blocks), dim3(NT), 0, s, x, Gt, (const int2 *)blk_src, nA, ncol, nw);
    return hipGetLastError();
}

namespace {

template <int LG1>
hipError_t launch_f1(const Sigma2Plan &p, const double2 *Gt, double2 *Z, hipStream_t s) {
    if (hipError_t e = qs_lds_attr((const void *)k_fft_f1<LG1>)) return e;
    static_assert(lay_one_to_one<LG1, CB>(), "LDS layout");
    const size_t lds = Lay<LG1, CB>::bytes();
    hipLaunchKernelGGL(k_fft_f1<LG1>, dim3((p.ncol / CB + BPW - 1) / BPW, p.N2), dim3(NT), lds, s, Gt, Z, p);
    return hipGetLastError();
}
template <int LG1>
hipError_t launch_i2(const Sigma2Plan &p, const double2 *Z, double2 *Sig, hipStream_t s) {
    if (hipError_t e = qs_lds_attr((const void *)k_fft_i2<LG1>)) return e;
    static_assert(lay_one_to_one<LG1, CB>(), "LDS layout");
    const size_t lds = Lay<LG1, CB>::bytes();
    hipLaunchKernelGGL(k_fft_i2<LG1>, dim3((p.ncol / CB + BPW - 1) / BPW, p.N2), dim3(NT), lds, s, Z, Sig, p);
    return hipGetLastError();
}
template <int LG2>
hipError_t launch_mid(const Sigma2Plan &p, double2 *Z, hipStream_t s) {
    if (hipError_t e = qs_lds_attr((const void *)k_fft_mid<LG2>)) return e;
    static_assert(lay_one_to_one<LG2, PS>(), "LDS layout");
    const size_t lds = Lay<LG2, PS>::bytes();
    const int npairs = p.nAs * (p.nAs + 1) / 2;
    hipLaunchKernelGGL(k_fft_mid<LG2>, dim3(npairs, p.N1 / 2 + 1), dim3(NT), lds, s, Z, p);
    return hipGetLastError();
}


}  // namespace

hipError_t qs_launch_sigma2(const Sigma2Plan &p, const double2 *Gt, double2 *Z, double2 *Sig, int parts,
                            hipStream_t s) {
    if (p.log2N1 < 1 || p.log2N1 > 10 || p.log2N2 < 1 || p.log2N2 > 8 || p.ncol % CB) return hipErrorInvalidValue;
    hipError_t e = hipSuccess;
#define QS_LG1_CASES(F, ...)                                                                            \
    switch (p.log2N1) {                                                                                 \
        case 1: e = F<1>(__VA_ARGS__); break;                                                           \
        case 2: e = F<2>(__VA_ARGS__); break;                                                           \
        case 3: e = F<3>(__VA_ARGS__); break;                                                           \
        case 4: e = F<4>(__VA_ARGS__); break;                                                           \
        case 5: e = F<5>(__VA_ARGS__); break;                                                           \
        case 6: e = F<6>(__VA_ARGS__); break;                                                           \
        case 7: e = F<7>(__VA_ARGS__); break;                                                           \
        case 8: e = F<8>(__VA_ARGS__); break;                                                           \
        case 9: e = F<9>(__VA_ARGS__); break;                                                           \
        default: e = F<10>(__VA_ARGS__); break;                                                         \
    }
    if (parts & 1) {
        QS_LG1_CASES(launch_f1, p, Gt, Z, s)
        if (e != hipSuccess) return e;
    }
    if (parts & 2) {
        switch (p.log2N2) {
            case 1: e = launch_mid<1>(p, Z, s); break;
            case 2: e = launch_mid<2>(p, Z, s); break;
            case 3: e = launch_mid<3>(p, Z, s); break;
            case 4: e = launch_mid<4>(p, Z, s); break;
            case 5: e = launch_mid<5>(p, Z, s); break;
            case 6: e = launch_mid<6>(p, Z, s); break;
            case 7: e = launch_mid<7>(p, Z, s); break;
            default: e = launch_mid<8>(p, Z, s); break;
        }
        if (e != hipSuccess) return e;
    }
    if (parts & 4) {
        QS_LG1_CASES(launch_i2, p, Z, Sig, s)
    }
#undef QS_LG1_CASES
    return e;
}
