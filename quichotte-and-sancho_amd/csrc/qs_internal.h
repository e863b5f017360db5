// qs_internal.h -- launch interfaces between the engine (qs_engine.cpp) and the
// gfx950 kernels.  Not part of the public C ABI (include/qs_hip.h).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <set>

// Raise a kernel's dynamic-LDS limit to the CU's 160 KiB, once per kernel and
// process; safe to call from several host threads (handles of a serving loop
// are created and launched concurrently).
inline hipError_t qs_lds_attr(const void *kernel) {
    static std::mutex mu;
    static std::set<const void *> done;
    std::lock_guard<std::mutex> lk(mu);
    if (done.count(kernel)) return hipSuccess;
    const hipError_t e = hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    if (e == hipSuccess) done.insert(kernel);
    return e;
}

// modes of the per-energy solve, as G_at_energy's (src/scf.f90:49)
enum { QS_M_INIT2 = 1, QS_M_SCF2 = 2, QS_M_SCF1 = 3, QS_M_FINAL = 4 };

// Arguments of the fused per-energy Dyson + Keldysh kernel.  Matrices are
// row-major complex (double2).  The second-order self-energy uses the
// site-pair blocked layout described in qs_sigma2.hip.
struct DysonArgs {
    int norb, nA, nS, nCL, nsites, nAs, n_atoms, atom_tot;
    int nw, iw0;             // energies handled: iw0 .. iw0 + gridDim.x - 1
    double w_init, delta;    // omega(iw) = w_init + (iw + 1) * delta
    double mu, beta, V;
    int mode, use_sigma, order;
    const double2 *Hemb;     // (norb, norb): -H + (i/2)(GammaL + GammaR)
    const double2 *sig1;     // (nsites, 2, 2) on-site first-order self-energy
    const double2 *sig;      // [nw][sig_ld] second-order self-energy, site-pair layout (qs_device.h)
    int sig_ld;              // complex values per energy in sig: 12 per unordered active site pair
    const double2 *gamL_s, *gamR_s;  // (nS, nS)
    const double2 *gamL_cl;  // (nCL, nCL)
    const int *act, *cl, *pos, *suppos;  // suppos: orbital -> index in sup or -1
    const int *tri_map;      // (n_su + n_u): (i << 16) | j of the packed G^< entries
    const int *atom_first, *atom_norb, *atom_off;
    const double *x;         // packed input vector (SCF2)
    double *xout;            // INIT2: packed non-interacting G written here
    double *F;               // SCF2: pack(G_out) - x
    const double *Fold;      // SCF2: previous residual (for |F - Fold|^2)
    double *partials;        // SCF2: [block][4] max|F|, sum F^2, sum (F-Fold)^2
    double *diagR, *diagL;   // SCF1/SCF2: [iw][nA] Im G^r_kk, Im G^<_kk
    double2 *onsL;           // SCF1/FINAL: [iw][4 nsites]
    double *dosd;            // FINAL: [iw][norb] Im G^r_ii
    double2 *atomL;          // FINAL: [iw][atom_tot]
    double *trUD;            // FINAL: [iw][2] current integrand up / down
    int *status;             // bit 0: singular pivot met
    double2 *mscr;           // small-LDS kernel: 2 norb^2 complex of scratch per energy (M, T)
    int mscr_energies;       //   for this many energies (launches are chunked to it)
    void *big;               // large-Norb path: its QsBigWork (chunk buffers)
    int fp32;                // large-Norb path: complex FP32 matrices, inverse and products
};

// Dyson + Keldysh at every energy of [iw0, iw0 + count).  Returns the hip error.
hipError_t qs_launch_dyson(const DysonArgs &a, int count, hipStream_t s);
// launch shapes of the Dyson kernel for a size: qs_dyson_variants(norb) choices
int qs_dyson_variants(int norb);
hipError_t qs_launch_dyson_variant(const DysonArgs &a, int count, int variant, hipStream_t s);
size_t qs_dyson_lds_bytes(int norb, int nCL);
int qs_dyson_max_norb();
// launch shapes: 0 = 256-thread workgroup, 1 = 512-thread workgroup (norb > 16),
// 2 = one wavefront per energy (norb <= qs_dyson_wave_max_norb(), needs mscr)
// 4 = shape 2 split into a Gauss-Jordan kernel (three energies per CU) and a Keldysh kernel
// 5 = shape 4 with the Keldysh part on FP64 MFMA
enum { QS_DYSON_WG256 = 0, QS_DYSON_WG512 = 1, QS_DYSON_WAVE = 2, QS_DYSON_BIG = 3, QS_DYSON_SPLIT = 4,
       QS_DYSON_SPLIT_MFMA = 5, QS_DYSON_NSHAPES = 6 };
// large-Norb path (qs_dyson_big.hip): energies in chunks through HBM, batched
// library factorisation and products, hand-written build and packing kernels
struct QsBigWork;
int qs_dyson_big_max_norb();
size_t qs_dyson_big_bytes(int norb, int nw, size_t budget);
QsBigWork *qs_big_create(int norb, int nw, size_t budget, size_t *bytes);
void qs_big_destroy(QsBigWork *bw);
hipError_t qs_launch_dyson_big(const DysonArgs &a, int count, hipStream_t s);
bool qs_dyson_shape_ok(int norb, int shape);
// what a launch shape is, and whether its products issue MFMA instructions
const char *qs_dyson_shape_desc(int shape, int *uses_mfma);
// batched complex inverse and products on MFMA (qs_gj_mfma.hip); row-major np x np
// matrices, np a multiple of 16 and at most qs_gj_max_np(); fp32 selects complex FP32
int qs_gj_max_np();
size_t qs_gj_lds_bytes(int np, bool fp32);
// G[b] = M[b]^-1 for count matrices (M is overwritten); bit 0 of *status on an exact zero pivot
hipError_t qs_launch_gj_inv(int np, int count, void *M, void *G, int *status, bool fp32, hipStream_t s);
// C[b] = A[b] op(B[b]), op = identity or conjugate transpose.  upper: only the
// QS_CGEMM_TILE x QS_CGEMM_TILE tiles of C on and above the diagonal are formed (for
// a product known to be Hermitian or anti-Hermitian; the tiles below keep what they held)
constexpr int QS_CGEMM_TILE = 64;
hipError_t qs_launch_cgemm(int np, int count, const void *A, const void *B, void *C, bool conjB, bool fp32,
                           hipStream_t s, bool upper = false);
// small-LDS kernels (qs_dyson_wave.hip): form 0, one fused kernel; 1, the Gauss-Jordan
// part and the Keldysh part as two kernels through the slab (17..96 orbitals); 2, as 1
// with the Keldysh part on FP64 MFMA
hipError_t qs_launch_dyson_wave(const DysonArgs &a, int count, int form, hipStream_t s);
size_t qs_dyson_wave_lds_bytes(int norb, int nCL);
// One launch over the whole sweeps of several solves (a group step).  Member m
// owns workgroups [first[m], first[m + 1]) of each part's launch; its arguments
// are tab[m] in device memory.
constexpr int QS_GROUP_MAX = 64;
struct DysonGroupIdx { int nmem; int first[QS_GROUP_MAX + 1]; };
// the kernel instance a size runs under launch shape `variant` (2 or 4): members
// with equal keys can share a launch, whatever their orbital counts and grids; -1
// when the shape has no grouped form
int qs_dyson_group_key(int norb, int variant);
// ht: the members' arguments on the host (sizes, checks), dt: the same in device
// memory.  Each member sweeps its nw energies (iw0 = 0) through its own slab,
// which must hold them all (mscr_energies >= nw); shape 5 also needs nw even.
hipError_t qs_launch_dyson_group(const DysonArgs *ht, const DysonArgs *dt, int nmem, int variant, hipStream_t s);
size_t qs_dyson_wave_scratch_bytes(int norb);
int qs_dyson_wave_max_norb();

// Device-side control of order-2 SCF steps (qs_control.hip): the scalar state of
// broyden_solve (common/broyden_mixing.f90:51-184) and each step's outcome, kept
// in device memory so that a whole step runs without the host (and can be a
// captured graph).  fmax, s2, d2 are contiguous: k_reduce_partials writes them.
struct QsCtrl {
    int its, nstored, nh, slot;          // residuals so far; history ring bookkeeping
    int nh_use, stop, converged, abort;  // nh_use: history columns in the update (0 after a reset)
    int singular, unphysical, i_worst, pad;
    double fmax, s2, d2, rf;             // max|F|, |F|^2, |F - Fold|^2 of the last residual; 1/|F - Fold|
    double n_worst, w_worst;
    double gamma[32];
    double gram[32 * 32];
};
hipError_t qs_launch_ctl_reset(QsCtrl *c, int *status, hipStream_t s);
hipError_t qs_launch_ctl_status(int *status, hipStream_t s);
// after the residual: sanity from the column sums of Im G^r_kk, Im G^<_kk (label: act, 0-based),
// its += 1, the stop decision, and the history slot / 1 / |F - Fold| of this step
hipError_t qs_launch_ctl_check(QsCtrl *c, const double *sumR, const double *sumL, int nb, double delta,
                               const int *label, const int *status, int nhist, double tolf, hipStream_t s);
// after the history pass: the Gram row of the new slot, the projections and the
// (w0^2 + Gram) gamma = cvec solve (dposv; a failed factorisation resets the history)
hipError_t qs_launch_ctl_gram(QsCtrl *c, const double *sums, int nhist, hipStream_t s);

// Second-order self-energy stage (qs_sigma2.hip)
struct Sigma2Plan {
    int nA, nAs, nw, Nf, N1, N2, log2N1, log2N2;
    int ncol;                // complex columns per energy: 12 per unordered active site pair
    double delta;
    const double2 *tw;       // (Nf) e^{-2 pi i k / Nf}
    const double *Upair;     // (nAs, nAs) U_i * U_j * hbar^2 (delta/2pi)^2 / Nf
};
// x (packed, reference layout) -> Gt [nw][ncol] site-pair layout
hipError_t qs_launch_permute_in(const double *x, double2 *Gt, const int *blk_src, int nA, int ncol, int nw,
                                hipStream_t s);
// Broyden update (FP32 history) that also writes the next residual's Gt from x_new,
// and x_new - x into dx when dx is not null; qs_update_gt_fits tells whether the
// shape allows it (energy slice in LDS, nh <= 16)
bool qs_update_gt_fits(int nA, int ncol, int64_t n, int nh);
hipError_t qs_launch_broyden_update_f32_gt(int64_t n, const double *x, const double *F, const float *Us,
                                           const double *gamma, int nh, double alpha, double *xnew, double *dx,
                                           double2 *Gt, const int *blk_src, int nA, int ncol, int nw, hipStream_t s);
// the same with gamma and the history count from the control block; nothing when the
// step stopped, and no Gt once its >= maxscf (the final sweep reads Sigma2 there)
hipError_t qs_launch_broyden_update_f32_gt_ctl(int64_t n, const double *x, const double *F, const float *Us,
                                               const QsCtrl *c, int maxscf, double alpha, double *xnew, double *dx,
                                               double2 *Gt, const int *blk_src, int nA, int ncol, int nw,
                                               hipStream_t s);
// parts: bit 0 forward pass over n1, bit 1 middle pass (row DFTs + 2x2 algebra),
// bit 2 inverse pass over k1; the three in order make one Sigma2 evaluation.
hipError_t qs_launch_sigma2(const Sigma2Plan &p, const double2 *Gt, double2 *Z, double2 *Sig, int parts,
                            hipStream_t s);

// small vector kernels (qs_vector.hip)
// part: >= 256 * 4 nAs doubles of scratch
hipError_t qs_launch_equal_time_sigma1(const double *x, const int *srcidx, const int *act_site,
                                       const double *U_site_act, double2 *gt0, double2 *sig1, int nA, int nAs, int nw,
                                       double pp, double *part, hipStream_t s);
// out[c] = sum_r in[r][c] (deterministic order); scratch >= qs_colsum_chunks(ncols, nrows) * ncols
// doubles, at most QS_COLSUM_CHUNKS * ncols
constexpr int QS_COLSUM_CHUNKS = 1024;
int qs_colsum_chunks(int ncols, int nrows, int *rows_per);
hipError_t qs_launch_reduce_cols(const double *in, int ncols, int nrows, double *out, double *scratch,
                                 hipStream_t s);
hipError_t qs_launch_reduce_partials(const double *partials, int nblocks, double *out, hipStream_t s);
// dx (may be null): x - xold as the last update wrote it; read in place of x and xold
hipError_t qs_launch_broyden_hist_f32(int64_t n, const double *F, const double *Fold, const double *x,
                                      const double *xold, const double *dx, double rf, double alpha, float *dFs,
                                      float *Us, int slot, int nh, double *partials, int nblocks, hipStream_t s);
// the FP32-history pass with slot, 1/|F - Fold| and the history count from the control
// block (dx as above, or x - xold formed from x and xold when dx is null: the same FP64
// difference, the same bits; nothing is stored when the step stopped or slot < 0)
hipError_t qs_launch_broyden_hist_f32_ctl(int64_t n, const double *F, const double *Fold, const double *x,
                                          const double *xold, const double *dx, double alpha, float *dFs, float *Us,
                                          const QsCtrl *c, double *partials, int nblocks, hipStream_t s);
hipError_t qs_launch_broyden_hist_f64(int64_t n, const double *F, const double *Fold, const double *x,
                                      const double *xold, double fnorm, double alpha, double *dF, double *U,
                                      int slot, int nh, double *partials, int nblocks, hipStream_t s);
hipError_t qs_launch_broyden_update_f32(int64_t n, const double *x, const double *F, const float *Us,
                                        const double *gamma, int nh, double alpha, double *xnew, hipStream_t s);
hipError_t qs_launch_broyden_update_f64(int64_t n, const double *x, const double *F, const double *U,
                                        const double *gamma, int nh, double alpha, double *xnew, hipStream_t s);
hipError_t qs_launch_vec_stats(int64_t n, const double *F, const double *Fold, double *partials, int nblocks,
                               hipStream_t s);
hipError_t qs_launch_fill(double *p, int64_t n, double v, hipStream_t s);
int qs_vec_blocks(int64_t n);
