// qs_sancho.hip -- SANCHO's interaction self-energies of orders 2 and 3 on gfx950.
//
// Replaces complete_input_gf + build_self_energies of SANCHO
// (SANCHO/src/green_functions.f90:163-173, SANCHO/src/self_energies.f90:38-207):
// from the input G^r, G^< of one SCF step, for all energies at once,
//   chi(m,a,b)  particle-hole bubbles      (self_energies.f90:77-93)
//   phi(n,a,b)  particle-particle bubbles  (:97-111, order 3)
//   S, P        sums over the internal site k with U_k   (:160-175, order 3)
//   Sigma2, Sigma3 (retarded, lesser, greater)           (:128-145, :178-204)
//
// Every sum over the energy axis is a correlation or a convolution
//   out(p) = sum_k sum_t x_t(k) y_t(sig (k - p) + c),   sig = +-1,
// evaluated by one LDS-tiled kernel (k_sn_sum): a workgroup owns 256
// consecutive outputs p of one site pair, stages a 256-long tile of the x
// arrays and the 511-long window of the y arrays that the tile meets, and each
// thread sums over the tile.  Entries outside the grid are staged as zeros,
// which is exactly the reference's loop bounds (k1 = max(1, 1+m) .. min(nw, nw+m)
// and the like).  Tiles whose window misses the grid are skipped.  x reads are
// LDS broadcasts; y reads by consecutive lanes hit consecutive 16-byte slots.
// A launch too small to fill the GPU splits each output block's run of tiles
// over several workgroups, which write partial sums; k_sn_reduce adds them in
// a fixed order, so the results do not depend on scheduling.
//
// Layouts in HBM (na sites, nw energies, L = 2 nw - 1):
//   g   [X][a][b][k]   X = R, A, L, G (frequency-first copies of the completed input)
//   bub [Q][a][b][m]   Q = chi G, L, R, A | phi G, L, R, A | S G, L, R | P G, L, R
//                      chi at m + nw - 1 (m = -(nw-1)..nw-1), phi at n - 2 (n = 2..2nw)
//   Sigma outputs (na, na, nw) complex, column-major, as the reference's Sig2R etc.
#include <hip/hip_runtime.h>

#include <string>
#include <cstdarg>
#include <cstdio>
#include <mutex>

#include "qs_device.h"
#include "qs_hip.h"

namespace {

constexpr int SN_NT = 256;               // threads and outputs per workgroup
constexpr int SN_TK = 256;               // summed-axis tile
constexpr int SN_W = SN_TK + SN_NT;      // y window (511 used)

// G^a = (G^r)^dagger, G^> = G^< + G^r - G^a (green_functions.f90:163-173), and
// the frequency-first copies (self_energies.f90:66-74).  Thread per (a, b, k), k fastest.
__global__ __launch_bounds__(256) void k_sn_complete(int na, int nw, const double2 *__restrict__ GR,
                                                     const double2 *__restrict__ GL, double2 *__restrict__ g) {
    const size_t n = (size_t)na * na * nw;
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    const int k = (int)(idx % nw);
    const int pair = (int)(idx / nw), a = pair / na, b = pair % na;
    const size_t ab = a + (size_t)na * b + (size_t)na * na * k, ba = b + (size_t)na * a + (size_t)na * na * k;
    const double2 r = GR[ab], l = GL[ab], adv = cconj(GR[ba]);
    g[idx] = r;
    g[n + idx] = adv;
    g[2 * n + idx] = l;
    g[3 * n + idx] = csub(cadd(l, r), adv);
}

struct SnSeg {
    const double2 *x[4];  // x arrays: pair (a, b) at (a na + b) nx
    const double2 *y[4];  // y arrays: pair at (a na + b) ny, or (b na + a) ny when ytr
    int nx, ny, sig, c, ytr;
};
struct SnJob {
    SnSeg seg[2];
    int nseg;
    int np;            // outputs per pair
    double2 *out[4];
    int fortran;       // 0: out[q][pair np + p] scaled by pp; 1: out[q](a, b, p) scaled by cs U_a U_b pp
    double2 cs;
    double pp;
};
struct SnLaunch {
    SnJob job[2];
    const double *U;
    int na;
    int nsplit;        // workgroups sharing one output block, each over its own run of tiles
    double2 *part;     // nsplit > 1: raw partial sums [split][job][q][pair][p]
};

// acc of output p of pair (a, b), scaled and stored as job J asks
template <int NX>
__device__ __forceinline__ void sn_store(const SnLaunch &P, const SnJob &J, int a, int b, int p, const double2 *acc) {
    const int na = P.na;
    if (!J.fortran) {
        const size_t o = ((size_t)a * na + b) * J.np + p;
#pragma unroll
        for (int q = 0; q < NX; q++) J.out[q][o] = cscale(acc[q], J.pp);
    } else {
        // (c U_a U_b pp) acc, as second_order_sigma / third_order_sigma scale it
        const double2 s = cscale(J.cs, P.U[a] * P.U[b] * J.pp);
        const size_t o = a + (size_t)na * b + (size_t)na * na * p;
#pragma unroll
        for (int q = 0; q < NX; q++) J.out[q][o] = cmul(s, acc[q]);
    }
}

// NX = 4: the bubble sums  a0 += x0 y0, a1 += x1 y1, a2 += x2 y0 + x1 y2, a3 += x3 y0 + x1 y3
//   (chi: x = G, L, R, A at (a,b), y = L, G, A, R at (b,a); phi: y = G, L, R, A at (a,b)).
// NX = 3: the self-energy sums  a0 += x0 y0, a1 += x1 y1, a2 += x1 y2 + x2 y0
//   (particle-hole: x = (chi|S) L, G, R, y = g L, G, R at (i,j);
//    particle-particle: x = P L, G, R, y = g G, L, A at (j,i)); a0, a1, a2 = lesser, greater, retarded.
template <int NX>
__global__ __launch_bounds__(SN_NT) void k_sn_sum(SnLaunch P) {
    __shared__ double2 xs[NX][SN_TK];
    __shared__ double2 ys[NX][SN_W];
    const SnJob &J = P.job[blockIdx.z];
    const int na = P.na, pair = blockIdx.y, a = pair / na, b = pair % na;
    const int split = blockIdx.x % P.nsplit, p0 = (blockIdx.x / P.nsplit) * SN_NT;
    if (p0 >= J.np) return;  // uniform over the workgroup, before any barrier
    // the job's tiles, numbered over its segments in order; this workgroup sums [tlo, thi)
    int ntile = 0;
    for (int s = 0; s < J.nseg; s++) ntile += (J.seg[s].nx + SN_TK - 1) / SN_TK;
    const int tlo = (int)((long)ntile * split / P.nsplit), thi = (int)((long)ntile * (split + 1) / P.nsplit);
    int tix = 0;
    const int t = threadIdx.x, p = p0 + t;
    double2 acc[NX];
#pragma unroll
    for (int q = 0; q < NX; q++) acc[q] = make_double2(0.0, 0.0);
    for (int s = 0; s < J.nseg; s++) {
        const SnSeg &S = J.seg[s];
        const size_t xo = (size_t)pair * S.nx, yo = (size_t)(S.ytr ? b * na + a : pair) * S.ny;
        for (int k0 = 0; k0 < S.nx; k0 += SN_TK) {
            const int mine = tix++;
            if (mine < tlo || mine >= thi) continue;  // another split's tile
            // y index of (k, p) is yb + local, local = kk - t + NT - 1 (sig > 0) or t - kk + TK - 1
            const int yb = S.sig > 0 ? k0 - p0 - (SN_NT - 1) + S.c : p0 - k0 - (SN_TK - 1) + S.c;
            if (yb + (SN_W - 2) < 0 || yb >= S.ny) continue;  // the window misses the grid
            __syncthreads();  // the previous tile's reads are done
            for (int i = t; i < SN_TK; i += SN_NT) {
                const int k = k0 + i;
#pragma unroll
                for (int q = 0; q < NX; q++) xs[q][i] = k < S.nx ? S.x[q][xo + k] : make_double2(0.0, 0.0);
            }
            for (int i = t; i < SN_W; i += SN_NT) {
                const int yi = yb + i;
                const bool in = yi >= 0 && yi < S.ny;
#pragma unroll
                for (int q = 0; q < NX; q++) ys[q][i] = in ? S.y[q][yo + yi] : make_double2(0.0, 0.0);
            }
            __syncthreads();
            // The wavefront's outputs pw .. pw + 63 meet the grid (0 <= y index < ny) only for
            // k in the union of their ranges; outside it every product is with a staged zero.
            const int pw = __builtin_amdgcn_readfirstlane(p0 + (t & ~(QS_WAVE - 1)));
            int lo, hi;
            if (S.sig > 0) lo = pw - S.c - k0, hi = pw + (QS_WAVE - 1) - S.c + S.ny - k0;
            else lo = pw + S.c - S.ny + 1 - k0, hi = pw + (QS_WAVE - 1) + S.c + 1 - k0;
            lo = max(lo, 0);
            hi = min(hi, min(SN_TK, S.nx - k0));
            const int yl0 = S.sig > 0 ? SN_NT - 1 - t : t + SN_TK - 1, ys1 = S.sig > 0 ? 1 : -1;
#pragma unroll 4
            for (int kk = lo; kk < hi; kk++) {
                const int yl = yl0 + ys1 * kk;
                double2 xv[NX], yv[NX];
#pragma unroll
                for (int q = 0; q < NX; q++) { xv[q] = xs[q][kk]; yv[q] = ys[q][yl]; }
                cfma(acc[0], xv[0], yv[0]);
                cfma(acc[1], xv[1], yv[1]);
                if constexpr (NX == 4) {
                    cfma(acc[2], xv[2], yv[0]);
                    cfma(acc[2], xv[1], yv[2]);
                    cfma(acc[3], xv[3], yv[0]);
                    cfma(acc[3], xv[1], yv[3]);
                } else {
                    cfma(acc[2], xv[1], yv[2]);
                    cfma(acc[2], xv[2], yv[0]);
                }
            }
        }
    }
    if (p >= J.np) return;
    if (P.nsplit == 1) {
        sn_store<NX>(P, J, a, b, p, acc);
    } else {
        const size_t base = ((size_t)(split * gridDim.z + blockIdx.z) * NX * na * na + pair) * J.np + p;
        const size_t qs = (size_t)na * na * J.np;
#pragma unroll
        for (int q = 0; q < NX; q++) P.part[base + q * qs] = acc[q];
    }
}

// sum of the nsplit partials of each output, in split order, then stored as k_sn_sum would
template <int NX>
__global__ __launch_bounds__(256) void k_sn_reduce(SnLaunch P) {
    const SnJob &J = P.job[blockIdx.z];
    const int na = P.na, pair = blockIdx.y, p = blockIdx.x * 256 + threadIdx.x;
    if (p >= J.np) return;
    const size_t qs = (size_t)na * na * J.np;
    double2 acc[NX];
#pragma unroll
    for (int q = 0; q < NX; q++) acc[q] = make_double2(0.0, 0.0);
    for (int sp = 0; sp < P.nsplit; sp++) {
        const size_t base = ((size_t)(sp * gridDim.z + blockIdx.z) * NX * na * na + pair) * J.np + p;
#pragma unroll
        for (int q = 0; q < NX; q++) acc[q] = cadd(acc[q], P.part[base + q * qs]);
    }
    sn_store<NX>(P, J, pair / na, pair % na, p, acc);
}

// S and P: sums over the internal site k with U_k (self_energies.f90:160-175), pointwise
// in the bubble axis.  Thread per (m, i, j), m fastest; blockIdx.z = 0: S from chi, 1: P from phi.
__global__ __launch_bounds__(256) void k_sn_sp(int na, int L, const double *__restrict__ U,
                                               double2 *__restrict__ bub) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= L) return;
    const int pair = blockIdx.y, i = pair / na, j = pair % na;
    const size_t nB = (size_t)na * na * L;
    const double2 *B = bub + (size_t)blockIdx.z * 4 * nB;  // chi or phi: G, L, R, A
    double2 *O = bub + (8 + 3 * (size_t)blockIdx.z) * nB;  // S or P: G, L, R
    auto at = [&](int q, int r, int c) { return B[q * nB + ((size_t)r * na + c) * L + m]; };
    double2 sg = make_double2(0.0, 0.0), sl = sg, sr = sg;
    for (int k = 0; k < na; k++) {
        const double2 Rik = at(2, i, k), Gik = at(0, i, k), Lik = at(1, i, k);
        const double2 Gkj = at(0, k, j), Lkj = at(1, k, j), Rkj = at(2, k, j), Akj = at(3, k, j);
        double2 tg = cmul(Rik, Gkj), tl = cmul(Rik, Lkj);
        cfma(tg, Gik, Akj);
        cfma(tl, Lik, Akj);
        const double u = U[k];
        sg = cadd(sg, cscale(tg, u));
        sl = cadd(sl, cscale(tl, u));
        sr = cadd(sr, cscale(cmul(Rik, Rkj), u));
    }
    const size_t o = ((size_t)i * na + j) * L + m;
    O[o] = sg;
    O[nB + o] = sl;
    O[2 * nB + o] = sr;
}

// U_site into device memory from a kernel argument (no host-memory copy)
struct SnU {
    double u[256];
};
__global__ void k_sn_set_u(int na, SnU U, double *__restrict__ dst) {
    const int i = threadIdx.x;
    if (i < na) dst[i] = U.u[i];
}

}  // namespace

struct qs_sancho {
    int na = 0, nw = 0, order = 0, device = 0, cus = 256;
    hipStream_t stream = nullptr;
    double2 *part = nullptr;  // partial sums of split launches
    size_t part_elems = 0;
    // optional HIP-event timing of the bubble and self-energy stages on the handle's stream
    bool prof = false;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    double prof_ms[2] = {0.0, 0.0};
    int64_t prof_n[2] = {0, 0};
    double2 *g = nullptr;     // 4 na^2 nw
    double2 *bub = nullptr;   // (4 | 14) na^2 L
    double2 *gin = nullptr;   // 2 na^2 nw: host inputs staged
    double2 *out = nullptr;   // 6 na^2 nw: outputs staged for host callers
    double *U = nullptr;      // na
    std::string err;
};

namespace {
int sn_fail(qs_sancho *s, int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (s) s->err = buf;
    return code;
}
#define SNCHK(s, call)                                                                                   \
    do {                                                                                                 \
        hipError_t e_ = (call);                                                                          \
        if (e_ != hipSuccess) return sn_fail((s), QS_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_));   \
    } while (0)

size_t sn_nbub(int order) { return order == 3 ? 14 : 4; }

// A launch of fewer than three 256-output workgroups per CU is latency-bound on
// each workgroup's serial run of tiles: split that run over up to 4 workgroups
// (never more than its tiles), which write partials for k_sn_reduce.
int sn_nsplit(int np, int na, int nz, int ntile, int cus) {
    const long wg = (long)((np + 255) / 256) * na * na * nz;
    if (wg >= 3L * cus) return 1;
    long n = (3L * cus + wg - 1) / wg;
    n = n < 4 ? n : 4;
    n = n < ntile ? n : ntile;
    return n > 1 ? (int)n : 1;
}
size_t sn_part_elems(int nq, int np, int na, int nz, int ntile, int cus) {
    const int n = sn_nsplit(np, na, nz, ntile, cus);
    return n > 1 ? (size_t)n * nz * nq * na * na * np : 0;
}
template <int NX>
hipError_t sn_launch(qs_sancho *s, SnLaunch &P, int np, int nz, int ntile, hipStream_t st) {
    const int na = s->na, nb = (np + 255) / 256;
    const int n = sn_nsplit(np, na, nz, ntile, s->cus);
    P.nsplit = 1, P.part = nullptr;
    if (n > 1 && (size_t)n * nz * NX * na * na * np <= s->part_elems) P.nsplit = n, P.part = s->part;
    hipLaunchKernelGGL(k_sn_sum<NX>, dim3(nb * P.nsplit, na * na, nz), dim3(SN_NT), 0, st, P);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || P.nsplit == 1) return e;
    hipLaunchKernelGGL(k_sn_reduce<NX>, dim3(nb, na * na, nz), dim3(256), 0, st, P);
    return hipGetLastError();
}
}  // namespace

extern "C" {

int qs_sancho_create(int32_t na, int32_t n_omega, int32_t order, int32_t device, qs_sancho **out) {
    if (!out) return QS_ERR_ARG;
    *out = nullptr;
    if (na < 1 || na > 255 || n_omega < 1 || n_omega > (1 << 24) || (order != 2 && order != 3)) return QS_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return QS_ERR_HIP;
    qs_sancho *s = new qs_sancho;
    s->na = na, s->nw = n_omega, s->order = order, s->device = device;
    const size_t nG = (size_t)na * na * n_omega, nB = (size_t)na * na * (2 * (size_t)n_omega - 1);
    int rc = QS_OK;
    const int L = 2 * n_omega - 1, nz = order == 3 ? 2 : 1;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
        hipDeviceGetAttribute(&s->cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        rc = QS_ERR_HIP;
    if (rc == QS_OK && (hipMalloc(&s->g, 4 * nG * sizeof(double2)) != hipSuccess ||
                        hipMalloc(&s->bub, sn_nbub(order) * nB * sizeof(double2)) != hipSuccess ||
                        hipMalloc(&s->gin, 2 * nG * sizeof(double2)) != hipSuccess ||
                        hipMalloc(&s->out, 6 * nG * sizeof(double2)) != hipSuccess ||
                        hipMalloc(&s->U, na * sizeof(double)) != hipSuccess))
        rc = QS_ERR_NOMEM;
    if (rc == QS_OK) {  // partials of the bubble and self-energy launches, when they split
        const size_t e4 = sn_part_elems(4, L, na, nz, (n_omega + 255) / 256, s->cus);
        const size_t e3 = sn_part_elems(3, n_omega, na, nz, (L + 255) / 256, s->cus);
        s->part_elems = e4 > e3 ? e4 : e3;
        if (s->part_elems && hipMalloc(&s->part, s->part_elems * sizeof(double2)) != hipSuccess) rc = QS_ERR_NOMEM;
    }
    if (rc != QS_OK) {
        qs_sancho_destroy(s);
        return rc;
    }
    *out = s;
    return QS_OK;
}

void qs_sancho_destroy(qs_sancho *s) {
    if (!s) return;
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    (void)hipFree(s->g);
    (void)hipFree(s->bub);
    (void)hipFree(s->gin);
    (void)hipFree(s->out);
    (void)hipFree(s->U);
    (void)hipFree(s->part);
    for (hipEvent_t e : s->ev)
        if (e) (void)hipEventDestroy(e);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

const char *qs_sancho_last_error(const qs_sancho *s) { return s ? s->err.c_str() : "null handle"; }

int qs_sancho_profile(qs_sancho *s, int32_t enable) {
    if (!s) return QS_ERR_ARG;
    if (enable) {
        SNCHK(s, hipSetDevice(s->device));
        for (hipEvent_t &e : s->ev)
            if (!e) SNCHK(s, hipEventCreate(&e));
    }
    s->prof = enable != 0;
    s->prof_ms[0] = s->prof_ms[1] = 0.0;
    s->prof_n[0] = s->prof_n[1] = 0;
    return QS_OK;
}

int qs_sancho_profile_read(const qs_sancho *s, double *ms, int64_t *count) {
    if (!s || !ms || !count) return -QS_ERR_ARG;
    for (int k = 0; k < 2; k++) ms[k] = s->prof_ms[k], count[k] = s->prof_n[k];
    return 2;
}

int64_t qs_sancho_device_bytes(const qs_sancho *s) {
    if (!s) return -1;
    const int64_t nG = (int64_t)s->na * s->na * s->nw, nB = (int64_t)s->na * s->na * (2 * (int64_t)s->nw - 1);
    return (int64_t)sizeof(double2) * (12 * nG + (int64_t)sn_nbub(s->order) * nB + (int64_t)s->part_elems) +
           (int64_t)sizeof(double) * s->na;
}

}  // extern "C"

namespace {
int sn_self_energies(qs_sancho *s, double delta, const double *U_site, const double *GR, const double *GL,
                     double *Sig2R, double *Sig2L, double *Sig2G, double *Sig3R, double *Sig3L, double *Sig3G,
                     int32_t on_device) {
    if (!U_site || !GR || !GL || !Sig2R || !Sig2L || !Sig2G)
        return sn_fail(s, QS_ERR_ARG, "null input or second-order output");
    if (s->order == 3 && (!Sig3R || !Sig3L || !Sig3G)) return sn_fail(s, QS_ERR_ARG, "order 3 needs the Sig3 outputs");
    if (!(delta > 0.0)) return sn_fail(s, QS_ERR_ARG, "delta must be positive");
    const int na = s->na, nw = s->nw, L = 2 * nw - 1;
    const size_t nG = (size_t)na * na * nw, nB = (size_t)na * na * L;
    SNCHK(s, hipSetDevice(s->device));
    hipStream_t st = s->stream;
    {
        SnU u{};
        for (int i = 0; i < na; i++) u.u[i] = U_site[i];
        hipLaunchKernelGGL(k_sn_set_u, dim3(1), dim3(256), 0, st, na, u, s->U);
        SNCHK(s, hipGetLastError());
    }
    const double2 *gr = (const double2 *)GR, *gl = (const double2 *)GL;
    if (!on_device) {
        SNCHK(s, hipMemcpyAsync(s->gin, GR, nG * sizeof(double2), hipMemcpyHostToDevice, st));
        SNCHK(s, hipMemcpyAsync(s->gin + nG, GL, nG * sizeof(double2), hipMemcpyHostToDevice, st));
        gr = s->gin, gl = s->gin + nG;
    }
    double2 *o[6];
    double *user[6] = {Sig2R, Sig2L, Sig2G, Sig3R, Sig3L, Sig3G};
    for (int q = 0; q < 6; q++) o[q] = on_device ? (double2 *)user[q] : s->out + q * nG;
    const double pp = delta / (2.0 * 3.14159265358979323846);
    hipLaunchKernelGGL(k_sn_complete, dim3((unsigned)((nG + 255) / 256)), dim3(256), 0, st, na, nw, gr, gl, s->g);
    SNCHK(s, hipGetLastError());
    double2 *gR = s->g, *gA = s->g + nG, *gL = s->g + 2 * nG, *gG = s->g + 3 * nG;
    auto B = [&](int q) { return s->bub + q * nB; };
    const bool o3 = s->order == 3;
    {   // bubbles: chi (and phi)
        SnLaunch P{};
        P.na = na, P.U = s->U;
        SnJob &c = P.job[0];
        c.nseg = 1, c.np = L, c.fortran = 0, c.pp = pp;
        c.seg[0] = SnSeg{{gG, gL, gR, gA}, {gL, gG, gA, gR}, nw, nw, +1, nw - 1, 1};
        for (int q = 0; q < 4; q++) c.out[q] = B(q);
        SnJob &f = P.job[1];
        f = c;
        f.seg[0] = SnSeg{{gG, gL, gR, gA}, {gG, gL, gR, gA}, nw, nw, -1, 0, 0};
        for (int q = 0; q < 4; q++) f.out[q] = B(4 + q);
        if (s->prof) SNCHK(s, hipEventRecord(s->ev[0], st));
        SNCHK(s, sn_launch<4>(s, P, L, o3 ? 2 : 1, (nw + SN_TK - 1) / SN_TK, st));
        if (s->prof) SNCHK(s, hipEventRecord(s->ev[1], st));
    }
    if (o3) {
        hipLaunchKernelGGL(k_sn_sp, dim3((L + 255) / 256, na * na, 2), dim3(256), 0, st, na, L, (const double *)s->U,
                           s->bub);
        SNCHK(s, hipGetLastError());
    }
    {   // Sigma2 from chi; Sigma3 from S (particle-hole) and P (particle-particle)
        SnLaunch P{};
        P.na = na, P.U = s->U;
        SnJob &j2 = P.job[0];
        j2.nseg = 1, j2.np = nw, j2.fortran = 1, j2.pp = pp;
        j2.cs = make_double2(1.0, 0.0);  // -(i hbar)^2, hbar = 1
        j2.seg[0] = SnSeg{{B(1), B(0), B(2), nullptr}, {gL, gG, gR, nullptr}, L, nw, -1, nw - 1, 0};
        j2.out[0] = o[1], j2.out[1] = o[2], j2.out[2] = o[0], j2.out[3] = nullptr;  // lesser, greater, retarded
        SnJob &j3 = P.job[1];
        j3 = j2;
        j3.nseg = 2;
        j3.cs = make_double2(0.0, 1.0);  // -(i hbar)^3
        j3.seg[0] = SnSeg{{B(9), B(8), B(10), nullptr}, {gL, gG, gR, nullptr}, L, nw, -1, nw - 1, 0};
        j3.seg[1] = SnSeg{{B(12), B(11), B(13), nullptr}, {gG, gL, gA, nullptr}, L, nw, +1, 0, 1};
        j3.out[0] = o[4], j3.out[1] = o[5], j3.out[2] = o[3];
        if (s->prof) SNCHK(s, hipEventRecord(s->ev[2], st));
        SNCHK(s, sn_launch<3>(s, P, nw, o3 ? 2 : 1, (L + SN_TK - 1) / SN_TK, st));
        if (s->prof) SNCHK(s, hipEventRecord(s->ev[3], st));
    }
    if (!on_device)
        for (int q = 0; q < (o3 ? 6 : 3); q++)
            SNCHK(s, hipMemcpyAsync(user[q], o[q], nG * sizeof(double2), hipMemcpyDeviceToHost, st));
    SNCHK(s, hipStreamSynchronize(st));
    if (s->prof)
        for (int k = 0; k < 2; k++) {
            float ms = 0.0f;
            SNCHK(s, hipEventElapsedTime(&ms, s->ev[2 * k], s->ev[2 * k + 1]));
            s->prof_ms[k] += ms;
            s->prof_n[k]++;
        }
    return QS_OK;
}
}  // namespace

// ---------------------------------------------------------------------------
// Entry point that allocates nothing: the work arrays live in a fixed buffer of
// the library, one copy per device (held until the process ends).
// ---------------------------------------------------------------------------
constexpr size_t kStaticBytes = (size_t)64 << 20;
static __device__ __align__(256) double2 scratch[kStaticBytes / sizeof(double2)];

namespace {
std::mutex g_scratch_mu;
double2 *g_scratch_ptr[64] = {nullptr};

// the device address of this device's copy of scratch (the array's name is a host
// stand-in: only this address may be passed to kernels or copies)
int scratch_address(int dev, double2 **out) {
    if (dev < 0 || dev >= 64) return QS_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
            (void)hipGetLastError();
            return QS_ERR_HIP;
        }
        g_scratch_ptr[dev] = (double2 *)p;
    }
    *out = g_scratch_ptr[dev];
    return QS_OK;
}
// the work arrays of one call, in double2 slots: g, bubbles, split partials, U
struct StaticLayout {
    size_t nG, nB, part_elems, total;
};
StaticLayout static_layout(int na, int n_omega, int order, int cus) {
    StaticLayout l;
    const size_t L = 2 * (size_t)n_omega - 1;
    l.nG = (size_t)na * na * n_omega;
    l.nB = (size_t)na * na * L;
    const int nz = order == 3 ? 2 : 1;
    const size_t e4 = sn_part_elems(4, (int)L, na, nz, (n_omega + 255) / 256, cus);
    const size_t e3 = sn_part_elems(3, n_omega, na, nz, ((int)L + 255) / 256, cus);
    l.part_elems = e4 > e3 ? e4 : e3;
    l.total = 4 * l.nG + sn_nbub(order) * l.nB + l.part_elems + (na + 1) / 2;
    return l;
}
bool static_args_ok(int na, int n_omega, int order) {
    return na >= 1 && na <= 255 && n_omega >= 1 && n_omega <= (1 << 24) && (order == 2 || order == 3);
}
int current_cus(int *dev, int *cus) {
    if (hipGetDevice(dev) != hipSuccess) return QS_ERR_HIP;
    if (hipDeviceGetAttribute(cus, hipDeviceAttributeMultiprocessorCount, *dev) != hipSuccess) return QS_ERR_HIP;
    return QS_OK;
}
}  // namespace

extern "C" {

int64_t qs_sancho_static_bytes(void) { return (int64_t)kStaticBytes; }

int64_t qs_sancho_static_need(int32_t na, int32_t n_omega, int32_t order) {
    if (!static_args_ok(na, n_omega, order)) return -1;
    int dev = 0, cus = 256;
    if (current_cus(&dev, &cus) != QS_OK) return -1;
    return (int64_t)(sizeof(double2) * static_layout(na, n_omega, order, cus).total);
}

int qs_sancho_self_energies_static(int32_t na, int32_t n_omega, int32_t order, double delta, const double *U_site,
                                   const double *GR, const double *GL, double *Sig2R, double *Sig2L, double *Sig2G,
                                   double *Sig3R, double *Sig3L, double *Sig3G, void *stream) {
    if (!static_args_ok(na, n_omega, order)) return QS_ERR_ARG;
    int dev = 0, cus = 256;
    if (int rc = current_cus(&dev, &cus)) return rc;
    const StaticLayout l = static_layout(na, n_omega, order, cus);
    if (l.total * sizeof(double2) > kStaticBytes) return QS_ERR_UNSUPPORTED;
    double2 *base = nullptr;
    if (int rc = scratch_address(dev, &base)) return rc;
    // a descriptor of the work arrays inside scratch, laid out as qs_sancho_create's
    qs_sancho w;
    w.na = na, w.nw = n_omega, w.order = order, w.device = dev, w.cus = cus;
    w.stream = (hipStream_t)stream;
    w.part_elems = l.part_elems;
    w.g = base;
    w.bub = w.g + 4 * l.nG;
    w.part = l.part_elems ? w.bub + sn_nbub(order) * l.nB : nullptr;
    w.U = (double *)(w.bub + sn_nbub(order) * l.nB + l.part_elems);
    const int rc = sn_self_energies(&w, delta, U_site, GR, GL, Sig2R, Sig2L, Sig2G, Sig3R, Sig3L, Sig3G, 1);
    if (rc != QS_OK) (void)hipStreamSynchronize(w.stream);
    w.stream = nullptr;   // the caller's stream: not destroyed here
    return rc;
}

int qs_sancho_self_energies(qs_sancho *s, double delta, const double *U_site, const double *GR, const double *GL,
                            double *Sig2R, double *Sig2L, double *Sig2G, double *Sig3R, double *Sig3L,
                            double *Sig3G, int32_t on_device) {
    if (!s) return QS_ERR_ARG;
    const int rc = sn_self_energies(s, delta, U_site, GR, GL, Sig2R, Sig2L, Sig2G, Sig3R, Sig3L, Sig3G, on_device);
    // after a failure part-way, work already queued may still read or write the
    // caller's arrays: let it finish before handing them back
    if (rc != QS_OK && s->stream) (void)hipStreamSynchronize(s->stream);
    return rc;
}

}  // extern "C"
