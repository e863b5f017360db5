// qs_engine.cpp -- the SCF engine behind include/qs_hip.h.
//
// Host side of one bias-point solve.  State that the reference keeps in module
// variables (src/green_functions.f90:38-67, src/self_energies.f90:23-26,
// src/scf.f90:46-54) lives in a handle; every array that grows with the energy
// grid lives in HBM.  Per SCF step only scalars cross PCIe: max|F|, |F|^2,
// |F - Fold|^2, the 2 nA sanity sums and 2 nh Broyden dot products.
//
//   solve_scf          src/scf.f90:78-151
//   scf_residual2      src/scf.f90:158-188   (device: Sigma1, Sigma2, Dyson/Keldysh)
//   scf_residual1      src/scf.f90:194-224   (device sweep, host vector of 8/site)
//   broyden_solve      common/broyden_mixing.f90:51-184 (host scalars, device vectors)
//   final_sanity       src/scf.f90:260-275
//   check_diag_sanity  common/scf_checks.f90:59-91
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstddef>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "qs_hip.h"
#include "qs_device.h"
#include "qs_internal.h"

namespace {

using zd = std::complex<double>;
constexpr double PI = 3.14159265358979323846;
constexpr int MAXH_DEV = 32;

constexpr size_t kWaveSlab = (size_t)8 << 30;  // small-LDS Dyson scratch: energies per launch up to 8 GB of M

// energies per launch of the small-LDS Dyson kernel (its slab holds one M per energy)
int wave_chunk(int n, int nw) {
    const size_t per = qs_dyson_wave_scratch_bytes(n);
    const size_t c = per ? kWaveSlab / per : (size_t)nw;
    return (int)std::max<size_t>(1, std::min<size_t>((size_t)nw, c));
}
constexpr size_t kBigBudget = (size_t)2 << 30;  // chunk buffers of the large-Norb Dyson path

// stages timed with HIP events on the handle's stream (qs_profile)
enum Stage {
    ST_EQT = 0, ST_PERMUTE, ST_FFT_F1, ST_FFT_MID, ST_FFT_I2, ST_DYSON, ST_REDUCE, ST_HIST, ST_UPDATE, ST_FINAL,
    ST_DYSON32, ST_COUNT
};
const char *const kStageNames[ST_COUNT] = {"equal_time_sigma1", "permute_in",      "sigma2_fft_f1",  "sigma2_fft_mid",
                                           "sigma2_fft_i2",     "dyson_keldysh",   "reductions",     "broyden_history",
                                           "broyden_update",    "final_sweep",     "dyson_keldysh_fp32"};

struct Prof {
    bool on = false;
    std::vector<hipEvent_t> pool;
    struct Rec { int stage; hipEvent_t a, b; };
    std::vector<Rec> pending;
    // while a graph is captured, its intervals go to capture_to; each replay
    // records the same events again and is read with prof_collect_graph
    std::vector<Rec> *capture_to = nullptr;
    std::vector<Rec> graph_recs[2];
    double ms[ST_COUNT] = {0};
    int64_t count[ST_COUNT] = {0};
};

}  // namespace

struct qs_group;

struct qs_handle {
    qs_problem p{};
    int norb = 0, nw = 0, nsites = 0, order = 0;
    int nA = 0, nS = 0, nCL = 0, nAs = 0, n_int = 0;
    int Nf = 0, N1 = 0, N2 = 0, log2N1 = 0, log2N2 = 0;
    int64_t n_vec = 0;
    int n_atoms = 0, atom_tot = 0;
    bool has_atomL = false;
    std::vector<int> act, pos, sup, suppos, cl, asite, atom_norb, atom_first, atom_off;
    std::vector<char> site_int;
    std::vector<double> U;
    std::vector<zd> H, GamL, GamR, Hemb, GamL_s, GamR_s, GamL_cl;
    hipStream_t stream = nullptr;
    std::string err;
    int64_t dev_bytes = 0;
    std::vector<void *> allocs;

    // device constants
    double2 *d_Hemb = nullptr, *d_gamL_s = nullptr, *d_gamR_s = nullptr, *d_gamL_cl = nullptr;
    int *d_act = nullptr, *d_cl = nullptr, *d_pos = nullptr, *d_suppos = nullptr, *d_tri = nullptr;
    int *d_blk_src = nullptr, *d_et_src = nullptr, *d_act_site = nullptr;
    int *d_atom_first = nullptr, *d_atom_norb = nullptr, *d_atom_off = nullptr;
    double *d_U_act = nullptr, *d_Upair = nullptr;
    double2 *d_tw = nullptr;
    // device state
    double *d_x = nullptr, *d_F = nullptr, *d_Fold = nullptr, *d_xold = nullptr;
    double *d_dx = nullptr;  // optional: x - xold from the fused Broyden update (FP32 history)
    void *d_hist_dF = nullptr, *d_hist_U = nullptr;
    double2 *d_Gt = nullptr, *d_Z = nullptr;  // Gt doubles as the Sigma2 output
    int ncolS = 0;                            // site-pair layout columns: 12 per unordered active site pair
    double2 *d_sig1 = nullptr, *d_gt0 = nullptr;
    double *d_partials = nullptr, *d_red = nullptr, *d_scratch = nullptr;
    double2 *d_mscr = nullptr;   // one-wavefront Dyson kernel scratch (mscr_energies energies)
    int mscr_energies = 0;
    QsBigWork *big = nullptr;    // large-Norb Dyson path: library handle and chunk buffers
    double *d_diagR = nullptr, *d_diagL = nullptr, *d_dosd = nullptr, *d_trUD = nullptr;
    double2 *d_onsL = nullptr, *d_atomL = nullptr;
    int *d_status = nullptr;
    double *d_sums = nullptr;  // column sums
    size_t scratch_doubles = 0;
    double *h_pin = nullptr;   // pinned scalars

    // host state
    bool x_valid = false;      // d_x holds the input of the last residual / a fresh start
    bool sigma_valid = false;
    double V_scf = 0.0;
    std::vector<zd> G_less_t0, Sig1, G_less_obs;
    std::vector<double> trU_sum;  // [2]
    double J_up = 0, J_down = 0;
    int runaway_final = 0, unphys_final = 0, iworst_final = 1;
    double nworst_final = 0.5, wworst_final = 1.0;
    bool final_valid = false;
    Prof prof;
    int dyson_variant = 0;
    double dyson_tune_ms[QS_DYSON_NSHAPES] = {};
    // precision of the order-2 SCF sweeps: 0 = FP64; 1 = complex-FP32 Dyson/Keldysh
    // (MFMA) while max|F| >= mixed_switch, then FP64 to the end (and the final sweep)
    int precision = 0;
    double mixed_switch = 1e-5;
    int steps32 = 0, steps64 = 0;   // residuals of the last solve in each precision
    // order-2 SCF steps: 0 = host loop (scalars through the host each step), 1 = the
    // step's scalar logic on the device (QsCtrl; the default where the configuration
    // allows it), 2 = as 1 with steps replayed from captured graphs
    int step_mode = 1;
    QsCtrl *d_ctl = nullptr;
    QsCtrl *h_ctl = nullptr;        // pinned copy of the control header
    double *d_csum = nullptr;       // column sums: 2 nA sanity + 2 MAXH history
    int graph_steps = 0;            // steps of the last solve replayed from a graph
    qs_group *group = nullptr;      // the group stepping this handle's solve (qs_group_add)
};

// Several order-2 solves stepped together (qs_group_*): each step runs every
// member's Sigma1/Sigma2 stage on the member's own stream, then, on the group's
// stream once those have finished, one Dyson sweep launch per kernel instance for
// all members that share it, then every member's sanity check and Broyden passes
// on its own stream again.  Members' stages overlap on the device; the shared
// sweep is one launch.
struct qs_group {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    struct Member {
        qs_handle *h = nullptr;
        double V = 0.0;
        int its = 0;          // residuals of this solve so far
        double fmax = 0.0;    // max|F| of the last residual that was not aborted
        hipEvent_t ev = nullptr;  // the end of the member's stage before the shared sweep
    };
    std::vector<Member> mem;
    DysonArgs *h_tab = nullptr;  // pinned: the members' Dyson arguments of a step
    DysonArgs *d_tab = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};  // around the step's shared Dyson sweeps
    int last_launches = 0, last_fused = 0;  // Dyson launches of the last step; members in grouped launches
    double dyson_ms = 0.0;                  // the shared Dyson sweeps of every step so far
    int64_t steps = 0;
    bool in_step = false;                   // a step was begun and not yet ended
    std::vector<int> rc_step;               // each member's launch status in that step
};

namespace {

int fail(qs_handle *h, int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (h) h->err = buf;
    return code;
}

#define HIPCHK(h, call)                                                                                  \
    do {                                                                                                 \
        hipError_t e_ = (call);                                                                          \
        if (e_ != hipSuccess) return fail((h), QS_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_));      \
    } while (0)

hipEvent_t prof_event(qs_handle *h) {
    if (!h->prof.pool.empty()) {
        hipEvent_t e = h->prof.pool.back();
        h->prof.pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}
// open a timed region on the handle's stream; returns the start event (or null)
// an event record on the handle's stream; while a graph is captured it becomes an
// external event-record node, so each replay records the event again
bool prof_record(qs_handle *h, hipEvent_t e) {
    const hipError_t r = h->prof.capture_to ? hipEventRecordWithFlags(e, h->stream, hipEventRecordExternal)
                                            : hipEventRecord(e, h->stream);
    if (r != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return true;
}
hipEvent_t prof_begin(qs_handle *h) {
    if (!h->prof.on) return nullptr;
    hipEvent_t a = prof_event(h);
    if (a && !prof_record(h, a)) {
        h->prof.pool.push_back(a);
        return nullptr;
    }
    return a;
}
void prof_end(qs_handle *h, int stage, hipEvent_t a) {
    if (!a) return;
    hipEvent_t b = prof_event(h);
    if (!b || !prof_record(h, b)) {
        h->prof.pool.push_back(a);
        if (b) h->prof.pool.push_back(b);
        return;
    }
    (h->prof.capture_to ? *h->prof.capture_to : h->prof.pending).push_back({stage, a, b});
}
// after a replay of graph `par` has finished: add its intervals
void prof_collect_graph(qs_handle *h, int par) {
    for (auto &r : h->prof.graph_recs[par]) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            h->prof.ms[r.stage] += ms;
            h->prof.count[r.stage] += 1;
        }
    }
}
void prof_release_graph(qs_handle *h, int par) {
    for (auto &r : h->prof.graph_recs[par]) {
        h->prof.pool.push_back(r.a);
        h->prof.pool.push_back(r.b);
    }
    h->prof.graph_recs[par].clear();
}
// after a stream synchronisation: fold the pending intervals into the totals
void prof_collect(qs_handle *h) {
    for (auto &r : h->prof.pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            h->prof.ms[r.stage] += ms;
            h->prof.count[r.stage] += 1;
        }
        h->prof.pool.push_back(r.a);
        h->prof.pool.push_back(r.b);
    }
    h->prof.pending.clear();
}

// the order-2 FP32-history path keeps x - xold from the fused Broyden update
bool uses_dx(const qs_problem &p, int nA, int ncolS, int64_t n_vec) {
    return p.order == 2 && p.single_history && p.nhist > 0 && qs_update_gt_fits(nA, ncolS, n_vec, std::min(p.nhist, 16));
}

// QS_FILL=nan|zero in the environment fills every device buffer of a new handle
// (NaN bytes or zeros) before use: no result may depend on what the memory held
int fill_mode() {
    const char *e = getenv("QS_FILL");
    if (!e) return 0;
    if (!strcmp(e, "nan")) return 1;
    if (!strcmp(e, "zero")) return 2;
    return 0;
}

int fill_new(qs_handle *h, void *p, size_t bytes) {
    const int m = fill_mode();
    if (m) HIPCHK(h, hipMemsetAsync(p, m == 1 ? 0xff : 0, bytes, h->stream));
    return QS_OK;
}

template <class T>
int dalloc(qs_handle *h, T **ptr, size_t count) {
    size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(h, QS_ERR_NOMEM, "hipMalloc of %zu bytes failed: %s", bytes, hipGetErrorString(e));
    }
    if (int rc = fill_new(h, p, bytes)) {
        (void)hipFree(p);
        return rc;
    }
    h->allocs.push_back(p);
    h->dev_bytes += (int64_t)bytes;
    *ptr = (T *)p;
    return QS_OK;
}

// Every write into a handle's buffers goes through the handle's stream (created
// non-blocking), so each is ordered before the kernels that read it; a copy from a
// host vector is waited for before the vector can go away.
int copy_in(qs_handle *h, void *dst, const void *src, size_t bytes) {
    if (!bytes) return QS_OK;
    HIPCHK(h, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return QS_OK;
}

template <class T>
int upload(qs_handle *h, T **ptr, const std::vector<T> &v) {
    int rc = dalloc(h, ptr, v.size());
    if (rc) return rc;
    return copy_in(h, *ptr, v.data(), v.size() * sizeof(T));
}

int upload_z(qs_handle *h, double2 **ptr, const std::vector<zd> &v) {
    int rc = dalloc(h, ptr, v.size());
    if (rc) return rc;
    return copy_in(h, *ptr, v.data(), v.size() * sizeof(double2));
}

inline int su_idx(int i, int j, int n) { return i * n - i * (i + 1) / 2 + (j - i - 1); }
inline int u_idx(int i, int j, int n) { return i * n - i * (i - 1) / 2 + (j - i); }

// active subspace exactly as setup_storage (src/green_functions.f90:90-190)
void setup_storage(qs_handle *h) {
    const int n = h->norb, ns = h->nsites;
    h->site_int.assign(ns, 0);
    std::vector<char> site_act(ns, 0), conL(n, 0), conR(n, 0);
    h->n_int = 0;
    for (int s = 0; s < ns; s++) {
        h->site_int[s] = (h->order >= 1 && h->U[2 * s] != 0.0);
        h->n_int += h->site_int[s];
        site_act[s] = h->site_int[s];
    }
    for (int i = 0; i < n; i++) {
        for (int k = 0; k < n; k++) {
            if (h->GamL[(size_t)k * n + i] != 0.0 || h->GamL[(size_t)i * n + k] != 0.0) conL[i] = 1;
            if (h->GamR[(size_t)k * n + i] != 0.0 || h->GamR[(size_t)i * n + k] != 0.0) conR[i] = 1;
        }
        if (conL[i] || conR[i]) site_act[i / 2] = 1;
    }
    h->act.clear();
    h->pos.assign(n, -1);
    for (int i = 0; i < n; i++)
        if (site_act[i / 2]) { h->pos[i] = (int)h->act.size(); h->act.push_back(i); }
    h->nA = (int)h->act.size();
    h->asite.clear();
    for (int s = 0; s < ns; s++)
        if (site_act[s]) h->asite.push_back(s);
    h->nAs = (int)h->asite.size();
    h->cl.clear();
    for (int i = 0; i < n; i++)
        if (conL[i]) h->cl.push_back(i);
    h->nCL = (int)h->cl.size();
    h->sup.clear();
    if (h->order == 2) h->sup = h->act;
    else
        for (int i = 0; i < n; i++)
            if (conL[i] || conR[i]) h->sup.push_back(i);
    h->nS = (int)h->sup.size();
    h->suppos.assign(n, -1);
    for (int b = 0; b < h->nS; b++) h->suppos[h->sup[b]] = b;
    const int nS = h->nS, nCL = h->nCL;
    h->GamL_s.assign((size_t)nS * nS, 0.0);
    h->GamR_s.assign((size_t)nS * nS, 0.0);
    h->GamL_cl.assign((size_t)nCL * nCL, 0.0);
    // row-major on the device
    for (int a = 0; a < nS; a++)
        for (int b = 0; b < nS; b++) {
            h->GamL_s[(size_t)a * nS + b] = h->GamL[(size_t)h->sup[b] * n + h->sup[a]];
            h->GamR_s[(size_t)a * nS + b] = h->GamR[(size_t)h->sup[b] * n + h->sup[a]];
        }
    for (int a = 0; a < nCL; a++)
        for (int b = 0; b < nCL; b++) h->GamL_cl[(size_t)a * nCL + b] = h->GamL[(size_t)h->cl[b] * n + h->cl[a]];
    // Hemb = -H + (i/2hbar)(GammaL + GammaR), row-major (src/scf.f90:92)
    h->Hemb.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            const size_t cm = (size_t)j * n + i;
            h->Hemb[(size_t)i * n + j] = -h->H[cm] + zd(0.0, 0.5) * (h->GamL[cm] + h->GamR[cm]);
        }
    if (h->order == 1) h->n_vec = 8 * (int64_t)h->n_int;
    else if (h->order == 2) h->n_vec = 3 * (int64_t)h->nA * h->nA * h->nw;
    else h->n_vec = 0;
}

int setup_device(qs_handle *h) {
    int rc;
    const int n = h->norb, nA = h->nA, nA2 = nA * nA, nw = h->nw;
    HIPCHK(h, hipSetDevice(h->p.device));
    // non-blocking: work on the null stream (another handle's synchronous copies)
    // neither waits for this stream nor disturbs a graph being captured on it
    HIPCHK(h, hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    HIPCHK(h, hipHostMalloc((void **)&h->h_pin, 4096 * sizeof(double), hipHostMallocDefault));
    {
        // zero pad past n^2: the small-footprint Dyson kernel reads whole
        // 16-column blocks of the last row
        std::vector<zd> hp(h->Hemb);
        hp.resize(h->Hemb.size() + 64, zd(0.0, 0.0));
        if ((rc = upload_z(h, &h->d_Hemb, hp))) return rc;
    }
    if ((rc = upload_z(h, &h->d_gamL_s, h->GamL_s))) return rc;
    if ((rc = upload_z(h, &h->d_gamR_s, h->GamR_s))) return rc;
    if ((rc = upload_z(h, &h->d_gamL_cl, h->GamL_cl))) return rc;
    if ((rc = upload(h, &h->d_act, h->act))) return rc;
    if ((rc = upload(h, &h->d_cl, h->cl))) return rc;
    if ((rc = upload(h, &h->d_pos, h->pos))) return rc;
    if ((rc = upload(h, &h->d_suppos, h->suppos))) return rc;
    if ((rc = upload(h, &h->d_atom_first, h->atom_first))) return rc;
    if ((rc = upload(h, &h->d_atom_norb, h->atom_norb))) return rc;
    if ((rc = upload(h, &h->d_atom_off, h->atom_off))) return rc;
    if ((rc = dalloc(h, &h->d_status, 1))) return rc;
    HIPCHK(h, hipMemsetAsync(h->d_status, 0, sizeof(int), h->stream));
    if ((rc = dalloc(h, &h->d_sig1, (size_t)4 * h->nsites))) return rc;
    HIPCHK(h, hipMemsetAsync(h->d_sig1, 0, sizeof(double2) * 4 * h->nsites, h->stream));
    if ((rc = dalloc(h, &h->d_onsL, (size_t)nw * 4 * h->nsites))) return rc;
    if ((rc = dalloc(h, &h->d_dosd, (size_t)nw * n))) return rc;
    if ((rc = dalloc(h, &h->d_trUD, (size_t)nw * 2))) return rc;
    if (h->has_atomL && (rc = dalloc(h, &h->d_atomL, (size_t)nw * h->atom_tot))) return rc;
    if ((rc = dalloc(h, &h->d_diagR, (size_t)nw * std::max(nA, 1)))) return rc;
    if ((rc = dalloc(h, &h->d_diagL, (size_t)nw * std::max(nA, 1)))) return rc;
    if ((rc = dalloc(h, &h->d_partials, (size_t)nw * 4))) return rc;
    if (n <= qs_dyson_wave_max_norb()) {
        h->mscr_energies = wave_chunk(n, nw);
        if ((rc = dalloc(h, &h->d_mscr, h->mscr_energies * qs_dyson_wave_scratch_bytes(n) / sizeof(double2)))) return rc;
    }
    // the large-Norb Dyson path, needed above the small-LDS kernel's size or when forced
    const char *fv = getenv("QS_DYSON_VARIANT");
    if (n > qs_dyson_wave_max_norb() || (fv && atoi(fv) == QS_DYSON_BIG)) {
        size_t bytes = 0;
        h->big = qs_big_create(n, nw, kBigBudget, &bytes);
        if (!h->big) return fail(h, QS_ERR_HIP, "large-Norb Dyson workspace (chunk buffers)");
        h->dev_bytes += (int64_t)bytes;
    }
    if ((rc = dalloc(h, &h->d_red, 64))) return rc;
    // column-sum scratch: the row chunks of the widest column set over the grid, and
    // behind the Broyden partials the chunks of their own column sums
    int widest = std::max({8 * h->nsites, n, 2 * std::max(nA, 1), 2 * std::max(h->atom_tot, 1), 2});
    const int hist_blocks = qs_vec_blocks(std::max<int64_t>(h->n_vec, 1));
    size_t need = (size_t)qs_colsum_chunks(widest, nw, nullptr) * widest;
    for (int c : {8 * h->nsites, n, std::max(nA, 1), 2 * std::max(h->atom_tot, 1), 2})
        need = std::max(need, (size_t)qs_colsum_chunks(c, nw, nullptr) * c);
    size_t hist_part = (size_t)hist_blocks * 2 * MAXH_DEV;
    size_t hist_sums = (size_t)qs_colsum_chunks(2 * MAXH_DEV, hist_blocks, nullptr) * 2 * MAXH_DEV;
    size_t et_part = (size_t)256 * 4 * std::max(h->nAs, 1);
    h->scratch_doubles = std::max({need, hist_part + hist_sums, et_part, (size_t)1024});
    if ((rc = dalloc(h, &h->d_scratch, h->scratch_doubles))) return rc;
    if ((rc = dalloc(h, &h->d_sums, (size_t)std::max(widest, 64)))) return rc;

    if (h->order == 2) {
        // packed-vector index tables
        const int nsu = (nA2 - nA) / 2;
        std::vector<int> tri(nA2);
        int e = 0;
        for (int i = 0; i < nA; i++)
            for (int j = i + 1; j < nA; j++) tri[e++] = (i << 16) | j;
        for (int i = 0; i < nA; i++)
            for (int j = i; j < nA; j++) tri[e++] = (i << 16) | j;
        if ((rc = upload(h, &h->d_tri, tri))) return rc;
        // source of every column of the site-pair layout in the packed vector x:
        // re = 1 + index (negated for -Re), 0 for a zero real part; im = index, -1 for zero
        const int nAs = h->nAs;
        // 12 columns per unordered pair, padded to the FFT column block of 8 (zero columns)
        h->ncolS = (12 * (nAs * (nAs + 1) / 2) + 7) / 8 * 8;
        std::vector<int> blk(2 * (size_t)h->ncolS, 0);
        for (size_t c = 12 * (size_t)(nAs * (nAs + 1) / 2); c < (size_t)h->ncolS; c++) blk[2 * c + 1] = -1;
        auto source = [&](int i, int j, bool lesser, int &re, int &im) {
            if (!lesser) {
                re = i * nA + j + 1;
                im = nA2 + i * nA + j;
            } else if (i < j) {
                re = 2 * nA2 + su_idx(i, j, nA) + 1;
                im = 2 * nA2 + nsu + u_idx(i, j, nA);
            } else if (i == j) {
                re = 0;
                im = 2 * nA2 + nsu + u_idx(i, i, nA);
            } else {
                re = -(2 * nA2 + su_idx(j, i, nA) + 1);
                im = 2 * nA2 + nsu + u_idx(j, i, nA);
            }
        };
        for (int I = 0; I < nAs; I++)
            for (int J = I; J < nAs; J++)
                for (int slot = 0; slot < 12; slot++) {
                    const int c = 12 * sig_pair(I, J, nAs) + slot, s1 = (slot >> 1) & 1, s2 = slot & 1;
                    int re, im;
                    if (slot < 8) source(2 * I + s1, 2 * J + s2, slot >= 4, re, im);
                    else source(2 * J + s1, 2 * I + s2, false, re, im);
                    blk[2 * (size_t)c] = re;
                    blk[2 * (size_t)c + 1] = im;
                }
        if ((rc = upload(h, &h->d_blk_src, blk))) return rc;
        std::vector<int> et(4 * (size_t)h->nAs), asite_v(h->nAs);
        std::vector<double> uact(h->nAs);
        for (int ia = 0; ia < h->nAs; ia++) {
            const int i = 2 * ia;
            et[4 * ia + 0] = 2 * nA2 + nsu + u_idx(i, i, nA);
            et[4 * ia + 1] = 2 * nA2 + nsu + u_idx(i + 1, i + 1, nA);
            et[4 * ia + 2] = 2 * nA2 + su_idx(i, i + 1, nA);
            et[4 * ia + 3] = 2 * nA2 + nsu + u_idx(i, i + 1, nA);
            asite_v[ia] = h->asite[ia];
            uact[ia] = h->U[2 * h->asite[ia]];
        }
        if ((rc = upload(h, &h->d_et_src, et))) return rc;
        if ((rc = upload(h, &h->d_act_site, asite_v))) return rc;
        if ((rc = upload(h, &h->d_U_act, uact))) return rc;
        const double pp = h->p.delta / (2.0 * PI);
        std::vector<double> upair((size_t)h->nAs * h->nAs);
        for (int ia = 0; ia < h->nAs; ia++)
            for (int ja = 0; ja < h->nAs; ja++) upair[(size_t)ia * h->nAs + ja] = uact[ia] * uact[ja] * pp * pp / h->Nf;
        if ((rc = upload(h, &h->d_Upair, upair))) return rc;
        std::vector<zd> tw(h->Nf);
        for (int k = 0; k < h->Nf; k++) {
            long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)h->Nf;
            tw[k] = zd((double)cosl(a), (double)-sinl(a));
        }
        if ((rc = upload_z(h, &h->d_tw, tw))) return rc;
        const size_t nv = (size_t)h->n_vec;
        if ((rc = dalloc(h, &h->d_x, nv))) return rc;
        if ((rc = dalloc(h, &h->d_F, nv))) return rc;
        if ((rc = dalloc(h, &h->d_Fold, nv))) return rc;
        if ((rc = dalloc(h, &h->d_xold, nv))) return rc;
        HIPCHK(h, hipMemsetAsync(h->d_Fold, 0, nv * sizeof(double), h->stream));
        HIPCHK(h, hipMemsetAsync(h->d_xold, 0, nv * sizeof(double), h->stream));
        const int nhist = h->p.nhist;
        if (nhist > 0) {
            if (h->p.single_history) {
                float *a, *b;
                if ((rc = dalloc(h, &a, nv * nhist))) return rc;
                if ((rc = dalloc(h, &b, nv * nhist))) return rc;
                h->d_hist_dF = a;
                h->d_hist_U = b;
            } else {
                double *a, *b;
                if ((rc = dalloc(h, &a, nv * nhist))) return rc;
                if ((rc = dalloc(h, &b, nv * nhist))) return rc;
                h->d_hist_dF = a;
                h->d_hist_U = b;
            }
        }
        if ((rc = dalloc(h, &h->d_Gt, (size_t)nw * h->ncolS))) return rc;
        if ((rc = dalloc(h, &h->d_Z, (size_t)h->Nf * h->ncolS))) return rc;
        // x - xold for the history pass (saves it reading x and xold), where the fused
        // update runs; counted in qs_footprint_bytes.  Without it (allocation
        // failure, or QS_DX_BUFFER=0 in the environment) the history pass reads x
        // and xold, with the same results.
        const char *dxe = getenv("QS_DX_BUFFER");
        if (uses_dx(h->p, h->nA, h->ncolS, h->n_vec) && !(dxe && atoi(dxe) == 0)) {
            void *p = nullptr;
            if (hipMalloc(&p, nv * sizeof(double)) == hipSuccess) {
                h->allocs.push_back(p);
                if (int rc = fill_new(h, p, nv * sizeof(double))) return rc;
                h->dev_bytes += (int64_t)(nv * sizeof(double));
                h->d_dx = (double *)p;
            } else {
                (void)hipGetLastError();
            }
        }
        if ((rc = dalloc(h, &h->d_gt0, (size_t)4 * h->nAs))) return rc;
        if ((rc = dalloc(h, &h->d_ctl, 1))) return rc;
        if ((rc = dalloc(h, &h->d_csum, (size_t)2 * std::max(nA, 1) + 2 * MAXH_DEV))) return rc;
        HIPCHK(h, hipHostMalloc((void **)&h->h_ctl, sizeof(QsCtrl), hipHostMallocDefault));
    }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return QS_OK;
}

DysonArgs dyson_args(qs_handle *h, int mode, bool use_sigma, double V) {
    DysonArgs a{};
    a.norb = h->norb; a.nA = h->nA; a.nS = h->nS; a.nCL = h->nCL; a.nsites = h->nsites; a.nAs = h->nAs;
    a.n_atoms = h->n_atoms; a.atom_tot = h->atom_tot;
    a.nw = h->nw; a.iw0 = 0;
    a.w_init = h->p.w_init; a.delta = h->p.delta;
    a.mu = h->p.mu; a.beta = h->p.beta; a.V = V;
    a.mode = mode; a.use_sigma = use_sigma ? 1 : 0; a.order = h->order;
    a.Hemb = h->d_Hemb;
    a.sig1 = h->d_sig1;
    a.sig = (h->order == 2 && use_sigma) ? h->d_Gt : nullptr;
    a.sig_ld = h->ncolS;
    a.gamL_s = h->d_gamL_s; a.gamR_s = h->d_gamR_s; a.gamL_cl = h->d_gamL_cl;
    a.act = h->d_act; a.cl = h->d_cl; a.pos = h->d_pos; a.suppos = h->d_suppos;
    a.tri_map = h->d_tri;
    a.atom_first = h->d_atom_first; a.atom_norb = h->d_atom_norb; a.atom_off = h->d_atom_off;
    a.x = h->d_x; a.xout = h->d_x; a.F = h->d_F; a.Fold = h->d_Fold;
    a.partials = h->d_partials;
    a.diagR = h->d_diagR; a.diagL = h->d_diagL;
    a.onsL = h->d_onsL; a.dosd = h->d_dosd; a.atomL = h->d_atomL; a.trUD = h->d_trUD;
    a.status = h->d_status;
    a.mscr = h->d_mscr;
    a.mscr_energies = h->mscr_energies;
    a.big = h->big;
    return a;
}

// fp32: the complex-FP32 large-Norb path (mixed precision), whatever the tuned shape
int run_dyson(qs_handle *h, int mode, bool use_sigma, double V, bool fp32 = false) {
    DysonArgs a = dyson_args(h, mode, use_sigma, V);
    HIPCHK(h, hipMemsetAsync(h->d_status, 0, sizeof(int), h->stream));
    hipEvent_t t0 = prof_begin(h);
    if (fp32) {
        a.fp32 = 1;
        HIPCHK(h, qs_launch_dyson_variant(a, h->nw, QS_DYSON_BIG, h->stream));
    } else {
        HIPCHK(h, qs_launch_dyson_variant(a, h->nw, h->dyson_variant, h->stream));
    }
    prof_end(h, mode == QS_M_FINAL ? ST_FINAL : (fp32 ? ST_DYSON32 : ST_DYSON), t0);
    return QS_OK;
}

int check_status(qs_handle *h) {
    int st = 0;
    HIPCHK(h, hipMemcpyAsync(&st, h->d_status, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    prof_collect(h);
    if (st & 1) return fail(h, QS_ERR_SINGULAR, "singular matrix in the Dyson inverse (zgetrf INFO != 0)");
    return QS_OK;
}

// column sums over the energy grid of a [nw][ncols] array -> host vector
int colsum(qs_handle *h, const double *in, int ncols, std::vector<double> &out) {
    HIPCHK(h, qs_launch_reduce_cols(in, ncols, h->nw, h->d_sums, h->d_scratch, h->stream));
    out.resize(ncols);
    HIPCHK(h, hipMemcpyAsync(out.data(), h->d_sums, ncols * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return QS_OK;
}

struct Sanity {
    int runaway = 0, unphysical = 0, i_worst = 1;
    double n_worst = 0.5, w_worst = 1.0;
};

// check_diag_sanity (common/scf_checks.f90:59-91) from the sums over omega
Sanity diag_sanity(const double *sumImGL, const double *sumImGR, int nb, double delta, const int *label) {
    Sanity s;
    bool ra = false, un = false;
    for (int i = 0; i < nb; i++) {
        const int lab = label ? label[i] + 1 : i + 1;
        const double nocc = (delta / (2.0 * PI)) * sumImGL[i];
        const double wsum = -(delta / PI) * sumImGR[i];
        if (!std::isfinite(nocc) || !std::isfinite(wsum)) {
            s.runaway = 1; s.unphysical = un ? 1 : 0; s.i_worst = lab; s.n_worst = nocc; s.w_worst = wsum;
            return s;
        }
        if (nocc < -1.0 || nocc > 2.0) ra = true;
        if (nocc < -0.05 || nocc > 1.05 || wsum < 0.5) un = true;
        if (std::fabs(nocc - 0.5) > std::fabs(s.n_worst - 0.5) || (wsum < s.w_worst && std::fabs(s.n_worst - 0.5) <= 0.55)) {
            s.i_worst = lab; s.n_worst = nocc; s.w_worst = wsum;
        }
    }
    s.runaway = ra;
    s.unphysical = (ra || un);
    return s;
}

// Pick the Dyson launch shape for this size: time each variant on a slice of
// the grid (mode FINAL, which every order runs) and keep the fastest (the
// lowest-numbered shape within 3 % of it).
// QS_DYSON_VARIANT=<k> in the environment forces variant k.
// The choice is kept for the process, per (orbitals, contact orbitals, slice
// length): a handle created while other work runs (a serving loop) then takes
// the shape an earlier handle of that size measured, and gives the same
// results bit for bit.
struct TuneRecord { int variant; double ms[QS_DYSON_NSHAPES]; };
std::mutex g_tune_mu;
std::mutex g_capture_mu;   // graph captures of SCF steps, one at a time in the process
std::map<std::tuple<int, int, int>, TuneRecord> g_tune;

int tune_dyson(qs_handle *h) {
    const int nv = qs_dyson_variants(h->norb);
    h->dyson_variant = 0;
    if (const char *e = getenv("QS_DYSON_VARIANT")) {
        const int v = atoi(e);
        if (v >= 0 && v < nv && qs_dyson_shape_ok(h->norb, v)) {
            h->dyson_variant = v;
            return QS_OK;
        }
        // a shape this size cannot use: time the ones it can
    }
    const int count = std::min(h->nw, 2048);
    const auto key = std::make_tuple(h->norb, h->nCL, count);
    // one tuning at a time in the process: a second creator of the same size
    // waits here and then takes the first one's record
    std::lock_guard<std::mutex> lk(g_tune_mu);
    {
        const auto it = g_tune.find(key);
        if (it != g_tune.end()) {
            h->dyson_variant = it->second.variant;
            for (int i = 0; i < QS_DYSON_NSHAPES; i++) h->dyson_tune_ms[i] = it->second.ms[i];
            return QS_OK;
        }
    }
    DysonArgs a = dyson_args(h, QS_M_FINAL, false, 0.0);
    hipEvent_t e0, e1;
    HIPCHK(h, hipEventCreate(&e0));
    HIPCHK(h, hipEventCreate(&e1));
    double best = 1e300;
    for (int v = 0; v < nv; v++) {
        if (!qs_dyson_shape_ok(h->norb, v)) continue;
        // the large-Norb path only where no register-resident shape applies
        if (v == QS_DYSON_BIG && h->norb <= qs_dyson_wave_max_norb()) continue;
        HIPCHK(h, qs_launch_dyson_variant(a, count, v, h->stream));  // warm-up (code load)
        HIPCHK(h, hipEventRecord(e0, h->stream));
        for (int r = 0; r < 2; r++) HIPCHK(h, qs_launch_dyson_variant(a, count, v, h->stream));
        HIPCHK(h, hipEventRecord(e1, h->stream));
        HIPCHK(h, hipEventSynchronize(e1));
        float ms = 0.f;
        HIPCHK(h, hipEventElapsedTime(&ms, e0, e1));
        if (v < QS_DYSON_NSHAPES) h->dyson_tune_ms[v] = ms / 2;
        if (ms < best) best = ms;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    // the lowest-numbered shape within 3 % of the fastest: near-ties, whose order
    // timing noise decides, resolve the same way in every process, so that another
    // run picks the same shape (and the same rounding) for a size more often
    for (int v = 0; v < nv && v < QS_DYSON_NSHAPES; v++)
        if (h->dyson_tune_ms[v] > 0.0 && 2 * h->dyson_tune_ms[v] <= 1.03 * best) {
            h->dyson_variant = v;
            break;
        }
    TuneRecord r{h->dyson_variant, {}};
    for (int i = 0; i < QS_DYSON_NSHAPES; i++) r.ms[i] = h->dyson_tune_ms[i];
    g_tune.emplace(key, r);
    return QS_OK;
}

// scf_residual2 on the device: d_x -> d_F; stats[3] = max|F|, |F|^2, |F-Fold|^2
// gt_ready: Gt already holds x in the FFT input layout (written by the fused Broyden update)
int residual2_device(qs_handle *h, double V, double stats[3], int *abort_flag, Sanity *san, bool gt_ready = false,
                     bool fp32 = false) {
    const int nA = h->nA;
    h->V_scf = V;
    const double pp = h->p.delta / (2.0 * PI);
    hipEvent_t t0 = prof_begin(h);
    HIPCHK(h, qs_launch_equal_time_sigma1(h->d_x, h->d_et_src, h->d_act_site, h->d_U_act, h->d_gt0, h->d_sig1, nA,
                                          h->nAs, h->nw, pp, h->d_scratch, h->stream));
    prof_end(h, ST_EQT, t0);
    if (!gt_ready) {
        t0 = prof_begin(h);
        HIPCHK(h, qs_launch_permute_in(h->d_x, h->d_Gt, h->d_blk_src, nA, h->ncolS, h->nw, h->stream));
        prof_end(h, ST_PERMUTE, t0);
    }
    Sigma2Plan sp{};
    sp.nA = nA; sp.nAs = h->nAs; sp.nw = h->nw; sp.Nf = h->Nf; sp.N1 = h->N1; sp.N2 = h->N2;
    sp.log2N1 = h->log2N1; sp.log2N2 = h->log2N2; sp.delta = h->p.delta; sp.tw = h->d_tw; sp.Upair = h->d_Upair;
    sp.ncol = h->ncolS;
    for (int part = 0; part < 3; part++) {
        t0 = prof_begin(h);
        HIPCHK(h, qs_launch_sigma2(sp, h->d_Gt, h->d_Z, h->d_Gt, 1 << part, h->stream));
        prof_end(h, ST_FFT_F1 + part, t0);
    }
    h->sigma_valid = true;
    int rc = run_dyson(h, QS_M_SCF2, true, V, fp32);
    if (rc) return rc;
    t0 = prof_begin(h);
    HIPCHK(h, qs_launch_reduce_partials(h->d_partials, h->nw, h->d_red, h->stream));
    prof_end(h, ST_REDUCE, t0);
    HIPCHK(h, hipMemcpyAsync(h->h_pin, h->d_red, 3 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    if ((rc = check_status(h))) return rc;
    stats[0] = h->h_pin[0]; stats[1] = h->h_pin[1]; stats[2] = h->h_pin[2];
    std::vector<double> sR, sL;
    if ((rc = colsum(h, h->d_diagR, nA, sR))) return rc;
    if ((rc = colsum(h, h->d_diagL, nA, sL))) return rc;
    *san = diag_sanity(sL.data(), sR.data(), nA, h->p.delta, h->act.data());
    *abort_flag = san->runaway;
    return QS_OK;
}

// host Broyden on small vectors (order 1), common/broyden_mixing.f90:51-184
template <class Fn>
void broyden_host(Fn residual, std::vector<double> &x, double alpha, int nhist, int maxscf, double tolf, bool sp,
                  int *converged, int *n_iter, double *fmax_out, double *lf, double *lr) {
    const int64_t n = (int64_t)x.size();
    const double w0 = 0.01;
    const int lda = std::max(nhist, 1);
    std::vector<double> F(n), Fold(n, 0.0), xold(n, 0.0), a((size_t)lda * lda), cvec(lda), gram((size_t)lda * lda, 0.0);
    std::vector<float> dFs, Us;
    std::vector<double> dF, U;
    if (sp) { dFs.assign((size_t)n * lda, 0.f); Us.assign((size_t)n * lda, 0.f); }
    else { dF.assign((size_t)n * lda, 0.0); U.assign((size_t)n * lda, 0.0); }
    *converged = 0;
    int nstored = 0, nh = 0;
    double fmax = 1.79769313486231570e308;
    for (int its = 1; its <= maxscf; its++) {
        *n_iter = its;
        int ab = 0;
        residual(x, F, &ab);
        if (ab) break;
        fmax = 0.0;
        double ss = 0.0;
        for (int64_t i = 0; i < n; i++) { fmax = std::max(fmax, std::fabs(F[i])); ss += F[i] * F[i]; }
        if (lf) lf[its - 1] = fmax;
        if (lr) lr[its - 1] = std::sqrt(ss) / std::sqrt((double)n);
        if (fmax < tolf) { *converged = 1; break; }
        if (its > 1 && nhist > 0) {
            const int slot = nstored % nhist;
            if (sp) {
                for (int64_t i = 0; i < n; i++) Fold[i] = F[i] - Fold[i];
                double fn = 0.0;
                for (int64_t i = 0; i < n; i++) fn += Fold[i] * Fold[i];
                fn = std::sqrt(fn);
                if (fn > 0.0) {
                    const double rf = 1.0 / fn;
                    for (int64_t i = 0; i < n; i++) {
                        dFs[(size_t)slot * n + i] = (float)(Fold[i] * rf);
                        Us[(size_t)slot * n + i] = (float)(alpha * Fold[i] * rf + (x[i] - xold[i]) * rf);
                    }
                    nstored++;
                    nh = std::min(nstored, nhist);
                    for (int k = 0; k < nh; k++) {
                        double s = 0.0;
                        for (int64_t i = 0; i < n; i++) s += (double)dFs[(size_t)k * n + i] * (double)dFs[(size_t)slot * n + i];
                        gram[(size_t)slot * lda + k] = s;
                        gram[(size_t)k * lda + slot] = s;
                    }
                }
            } else {
                double fn = 0.0;
                for (int64_t i = 0; i < n; i++) {
                    dF[(size_t)slot * n + i] = F[i] - Fold[i];
                    fn += dF[(size_t)slot * n + i] * dF[(size_t)slot * n + i];
                }
                fn = std::sqrt(fn);
                if (fn > 0.0) {
                    for (int64_t i = 0; i < n; i++) {
                        dF[(size_t)slot * n + i] /= fn;
                        U[(size_t)slot * n + i] = alpha * dF[(size_t)slot * n + i] + (x[i] - xold[i]) / fn;
                    }
                    nstored++;
                    nh = std::min(nstored, nhist);
                    for (int k = 0; k < nh; k++) {
                        double s = 0.0;
                        for (int64_t i = 0; i < n; i++) s += dF[(size_t)k * n + i] * dF[(size_t)slot * n + i];
                        gram[(size_t)slot * lda + k] = s;
                        gram[(size_t)k * lda + slot] = s;
                    }
                }
            }
        }
        nh = std::min(nstored, nhist);
        xold = x;
        Fold = F;
        if (nh > 0) {
            for (int j = 0; j < nh; j++)
                for (int i = 0; i < nh; i++) a[(size_t)j * lda + i] = gram[(size_t)j * lda + i];
            for (int i = 0; i < nh; i++) a[(size_t)i * lda + i] += w0 * w0;
            for (int k = 0; k < nh; k++) {
                double s = 0.0;
                for (int64_t i = 0; i < n; i++) s += (sp ? (double)dFs[(size_t)k * n + i] : dF[(size_t)k * n + i]) * F[i];
                cvec[k] = s;
            }
            // dposv('U')
            bool ok = true;
            for (int j = 0; j < nh && ok; j++) {
                double ajj = a[(size_t)j * lda + j];
                for (int k = 0; k < j; k++) ajj -= a[(size_t)j * lda + k] * a[(size_t)j * lda + k];
                if (!(ajj > 0.0)) { ok = false; break; }
                ajj = std::sqrt(ajj);
                a[(size_t)j * lda + j] = ajj;
                for (int i = j + 1; i < nh; i++) {
                    double t = a[(size_t)i * lda + j];
                    for (int k = 0; k < j; k++) t -= a[(size_t)j * lda + k] * a[(size_t)i * lda + k];
                    a[(size_t)i * lda + j] = t / ajj;
                }
            }
            if (ok) {
                for (int i = 0; i < nh; i++) {
                    double t = cvec[i];
                    for (int k = 0; k < i; k++) t -= a[(size_t)i * lda + k] * cvec[k];
                    cvec[i] = t / a[(size_t)i * lda + i];
                }
                for (int i = nh - 1; i >= 0; i--) {
                    double t = cvec[i];
                    for (int k = i + 1; k < nh; k++) t -= a[(size_t)k * lda + i] * cvec[k];
                    cvec[i] = t / a[(size_t)i * lda + i];
                }
            } else {
                nstored = 0;
                nh = 0;
            }
        }
        for (int64_t i = 0; i < n; i++) {
            double xi = x[i] + alpha * F[i];
            for (int k = 0; k < nh; k++) xi -= (sp ? (double)Us[(size_t)k * n + i] : U[(size_t)k * n + i]) * cvec[k];
            x[i] = xi;
        }
    }
    *fmax_out = fmax;
}

// dposv('U') on the Gram system; returns false when not positive definite
bool solve_spd(std::vector<double> &a, int nh, int lda, std::vector<double> &b) {
    for (int j = 0; j < nh; j++) {
        double ajj = a[(size_t)j * lda + j];
        for (int k = 0; k < j; k++) ajj -= a[(size_t)j * lda + k] * a[(size_t)j * lda + k];
        if (!(ajj > 0.0)) return false;
        ajj = std::sqrt(ajj);
        a[(size_t)j * lda + j] = ajj;
        for (int i = j + 1; i < nh; i++) {
            double t = a[(size_t)i * lda + j];
            for (int k = 0; k < j; k++) t -= a[(size_t)j * lda + k] * a[(size_t)i * lda + k];
            a[(size_t)i * lda + j] = t / ajj;
        }
    }
    for (int i = 0; i < nh; i++) {
        double t = b[i];
        for (int k = 0; k < i; k++) t -= a[(size_t)i * lda + k] * b[k];
        b[i] = t / a[(size_t)i * lda + i];
    }
    for (int i = nh - 1; i >= 0; i--) {
        double t = b[i];
        for (int k = i + 1; k < nh; k++) t -= a[(size_t)k * lda + i] * b[k];
        b[i] = t / a[(size_t)i * lda + i];
    }
    return true;
}

// broyden_solve with x, F and the history in HBM (order 2)
int broyden_device(qs_handle *h, double V, int *converged, int *n_iter, double *fmax_out, Sanity *last, double *lf,
                   double *lr) {
    const int64_t n = h->n_vec;
    const int nhist = h->p.nhist, lda = std::max(nhist, 1);
    const double alpha = h->p.alpha_mix, w0 = 0.01;
    const bool sp = h->p.single_history != 0;
    std::vector<double> gram((size_t)lda * lda, 0.0), a((size_t)lda * lda), cvec(lda), gam(lda);
    const int nblocks = qs_vec_blocks(n);
    int nstored = 0, nh = 0, rc;
    double fmax = 1.79769313486231570e308;
    *converged = 0;
    bool gt_ready = false;  // the last update also wrote the FFT input of the next residual
    bool dx_ready = false;  // ... and x - xold into d_dx
    // mixed precision: FP32 sweeps while max|F| stays above the switch level and keeps
    // falling (it stalls at the FP32 rounding floor), FP64 from then on
    bool use32 = h->precision == 1 && h->order == 2;
    double best32 = 1.79769313486231570e308;
    int stall32 = 0;
    h->steps32 = h->steps64 = 0;
    for (int its = 1; its <= h->p.maxscf; its++) {
        *n_iter = its;
        double st[3];
        int ab = 0;
        if ((rc = residual2_device(h, V, st, &ab, last, gt_ready, use32))) return rc;
        (use32 ? h->steps32 : h->steps64)++;
        gt_ready = false;
        if (ab) break;
        fmax = st[0];
        const double rms = std::sqrt(st[1]) / std::sqrt((double)n);
        if (lf) lf[its - 1] = fmax;
        if (lr) lr[its - 1] = rms;
        if (use32) {
            if (fmax < 0.5 * best32) {
                best32 = fmax;
                stall32 = 0;
            } else {
                stall32++;
            }
            // an FP32 residual never ends the loop: the next one is FP64
            if (fmax < h->mixed_switch || fmax < h->p.tolf || stall32 >= 6) use32 = false;
        } else if (fmax < (h->precision == 1 ? 0.1 * h->p.tolf : h->p.tolf)) {
            // mixed precision ends a decade below tolf, so that its solution sits well
            // inside the tolf-wide set of FP64 solutions and not just anywhere in it
            *converged = 1;
            break;
        }
        bool need_proj = false;
        int slot = -1;
        double fnorm = 0.0;
        if (its > 1 && nhist > 0) {
            fnorm = std::sqrt(st[2]);
            if (fnorm > 0.0) {
                slot = nstored % nhist;
                nstored++;
            }
        }
        nh = std::min(nstored, nhist);
        if (nh > 0) need_proj = true;
        if (need_proj) {
            hipEvent_t t0 = prof_begin(h);
            if (sp)
                HIPCHK(h, qs_launch_broyden_hist_f32(n, h->d_F, h->d_Fold, h->d_x, h->d_xold,
                                                     dx_ready ? h->d_dx : nullptr, slot >= 0 ? 1.0 / fnorm : 0.0,
                                                     alpha, (float *)h->d_hist_dF, (float *)h->d_hist_U, slot, nh,
                                                     h->d_scratch, nblocks, h->stream));
            else
                HIPCHK(h, qs_launch_broyden_hist_f64(n, h->d_F, h->d_Fold, h->d_x, h->d_xold, fnorm, alpha,
                                                     (double *)h->d_hist_dF, (double *)h->d_hist_U, slot, nh,
                                                     h->d_scratch, nblocks, h->stream));
            prof_end(h, ST_HIST, t0);
            // sum the per-block partials [nblocks][2 MAXH]
            HIPCHK(h, qs_launch_reduce_cols(h->d_scratch, 2 * MAXH_DEV, nblocks, h->d_sums,
                                            h->d_scratch + (size_t)nblocks * 2 * MAXH_DEV, h->stream));
            HIPCHK(h, hipMemcpyAsync(h->h_pin, h->d_sums, 2 * MAXH_DEV * sizeof(double), hipMemcpyDeviceToHost,
                                     h->stream));
            HIPCHK(h, hipStreamSynchronize(h->stream));
            if (slot >= 0)
                for (int k = 0; k < nh; k++) {
                    gram[(size_t)slot * lda + k] = h->h_pin[k];
                    gram[(size_t)k * lda + slot] = h->h_pin[k];
                }
            for (int k = 0; k < nh; k++) cvec[k] = h->h_pin[MAXH_DEV + k];
        }
        // xold = x, Fold = F happen as buffer swaps after the update
        int nh_use = nh;
        if (nh > 0) {
            for (int j = 0; j < nh; j++)
                for (int i = 0; i < nh; i++) a[(size_t)j * lda + i] = gram[(size_t)j * lda + i];
            for (int i = 0; i < nh; i++) a[(size_t)i * lda + i] += w0 * w0;
            for (int k = 0; k < nh; k++) gam[k] = cvec[k];
            if (!solve_spd(a, nh, lda, gam)) {
                nstored = 0;
                nh_use = 0;
            }
        }
        hipEvent_t t1 = prof_begin(h);
        // Gt shares memory with Sigma2, which the final sweep reads after the last
        // step: the step that can end the loop at maxscf keeps the plain update
        const bool fuse = sp && its < h->p.maxscf && qs_update_gt_fits(h->nA, h->ncolS, n, nh_use);
        if (fuse)
            HIPCHK(h, qs_launch_broyden_update_f32_gt(n, h->d_x, h->d_F, (const float *)h->d_hist_U, gam.data(), nh_use,
                                                      alpha, h->d_xold, h->d_dx, h->d_Gt, h->d_blk_src, h->nA,
                                                      h->ncolS, h->nw, h->stream));
        else if (sp)
            HIPCHK(h, qs_launch_broyden_update_f32(n, h->d_x, h->d_F, (const float *)h->d_hist_U, gam.data(), nh_use,
                                                   alpha, h->d_xold, h->stream));
        else
            HIPCHK(h, qs_launch_broyden_update_f64(n, h->d_x, h->d_F, (const double *)h->d_hist_U, gam.data(), nh_use,
                                                   alpha, h->d_xold, h->stream));
        prof_end(h, ST_UPDATE, t1);
        gt_ready = fuse;
        dx_ready = fuse && h->d_dx;
        std::swap(h->d_x, h->d_xold);  // x <- x_new, xold <- x
        std::swap(h->d_F, h->d_Fold);  // Fold <- F
    }
    *fmax_out = fmax;
    return QS_OK;
}

// whether the device-controlled steps apply: order 2, FP64 sweeps, FP32 history of at
// most 16 columns with the fused update.  The optional x - xold buffer only saves the
// history pass reading x and xold (the same difference, the same bits), so whether its
// allocation succeeded changes neither the path nor the results.
bool ctl_steps_ok(const qs_handle *h) {
    return h->order == 2 && h->precision == 0 && h->p.single_history && h->p.nhist > 0 && h->p.nhist <= 16 &&
           h->d_ctl && qs_update_gt_fits(h->nA, h->ncolS, h->n_vec, 16);
}

// One order-2 SCF step with its scalar logic on the device: the residual
// (scf_residual2, src/scf.f90:158-188), the sanity check, the history pass, the
// Gram solve and the update (broyden_solve, common/broyden_mixing.f90:51-184),
// then a copy of the control header to pinned host memory.  No host decision
// inside, so the same sequence can be captured as a graph.  A group step
// (qs_group_step) runs the part before the Dyson sweep and the part after it for
// each member around one sweep launch for all of them.
int enqueue_ctl_pre(qs_handle *h) {
    const int nA = h->nA;
    const double pp = h->p.delta / (2.0 * PI);
    hipEvent_t t0 = prof_begin(h);
    HIPCHK(h, qs_launch_equal_time_sigma1(h->d_x, h->d_et_src, h->d_act_site, h->d_U_act, h->d_gt0, h->d_sig1, nA,
                                          h->nAs, h->nw, pp, h->d_scratch, h->stream));
    prof_end(h, ST_EQT, t0);
    Sigma2Plan sp{};
    sp.nA = nA; sp.nAs = h->nAs; sp.nw = h->nw; sp.Nf = h->Nf; sp.N1 = h->N1; sp.N2 = h->N2;
    sp.log2N1 = h->log2N1; sp.log2N2 = h->log2N2; sp.delta = h->p.delta; sp.tw = h->d_tw; sp.Upair = h->d_Upair;
    sp.ncol = h->ncolS;
    for (int part = 0; part < 3; part++) {
        t0 = prof_begin(h);
        HIPCHK(h, qs_launch_sigma2(sp, h->d_Gt, h->d_Z, h->d_Gt, 1 << part, h->stream));
        prof_end(h, ST_FFT_F1 + part, t0);
    }
    h->sigma_valid = true;
    HIPCHK(h, qs_launch_ctl_status(h->d_status, h->stream));
    return QS_OK;
}

int enqueue_ctl_post(qs_handle *h) {
    const int nA = h->nA;
    const int64_t n = h->n_vec;
    hipEvent_t t0 = prof_begin(h);
    HIPCHK(h, qs_launch_reduce_partials(h->d_partials, h->nw, &h->d_ctl->fmax, h->stream));
    HIPCHK(h, qs_launch_reduce_cols(h->d_diagR, nA, h->nw, h->d_csum, h->d_scratch, h->stream));
    HIPCHK(h, qs_launch_reduce_cols(h->d_diagL, nA, h->nw, h->d_csum + nA, h->d_scratch, h->stream));
    HIPCHK(h, qs_launch_ctl_check(h->d_ctl, h->d_csum, h->d_csum + nA, nA, h->p.delta, h->d_act, h->d_status,
                                  h->p.nhist, h->p.tolf, h->stream));
    prof_end(h, ST_REDUCE, t0);
    const int nblocks = qs_vec_blocks(n);
    t0 = prof_begin(h);
    HIPCHK(h, qs_launch_broyden_hist_f32_ctl(n, h->d_F, h->d_Fold, h->d_x, h->d_xold, h->d_dx, h->p.alpha_mix,
                                             (float *)h->d_hist_dF, (float *)h->d_hist_U, h->d_ctl, h->d_scratch,
                                             nblocks, h->stream));
    prof_end(h, ST_HIST, t0);
    HIPCHK(h, qs_launch_reduce_cols(h->d_scratch, 2 * MAXH_DEV, nblocks, h->d_csum + 2 * nA,
                                    h->d_scratch + (size_t)nblocks * 2 * MAXH_DEV, h->stream));
    HIPCHK(h, qs_launch_ctl_gram(h->d_ctl, h->d_csum + 2 * nA, h->p.nhist, h->stream));
    t0 = prof_begin(h);
    HIPCHK(h, qs_launch_broyden_update_f32_gt_ctl(n, h->d_x, h->d_F, (const float *)h->d_hist_U, h->d_ctl,
                                                  h->p.maxscf, h->p.alpha_mix, h->d_xold, h->d_dx, h->d_Gt,
                                                  h->d_blk_src, nA, h->ncolS, h->nw, h->stream));
    prof_end(h, ST_UPDATE, t0);
    HIPCHK(h, hipMemcpyAsync(h->h_ctl, h->d_ctl, offsetof(QsCtrl, gamma), hipMemcpyDeviceToHost, h->stream));
    return QS_OK;
}

int enqueue_ctl_step(qs_handle *h, double V) {
    int rc;
    if ((rc = enqueue_ctl_pre(h))) return rc;
    DysonArgs a = dyson_args(h, QS_M_SCF2, true, V);
    hipEvent_t t0 = prof_begin(h);
    HIPCHK(h, qs_launch_dyson_variant(a, h->nw, h->dyson_variant, h->stream));
    prof_end(h, ST_DYSON, t0);
    return enqueue_ctl_post(h);
}

// broyden_solve with the scalar logic on the device.  Steps 1 and 2 run as plain
// launches; later steps replay one of two graphs captured from the same launch
// sequence (x / xold and F / Fold trade places every step, so the buffers of
// step k+2 are those of step k).
int broyden_device_ctl(qs_handle *h, double V, int *converged, int *n_iter, double *fmax_out, Sanity *last,
                       double *lf, double *lr) {
    const int64_t n = h->n_vec;
    int rc = QS_OK;
    *converged = 0;
    h->graph_steps = 0;
    HIPCHK(h, qs_launch_ctl_reset(h->d_ctl, h->d_status, h->stream));
    // the first residual's FFT input from x (later ones come from the fused update)
    hipEvent_t t0 = prof_begin(h);
    HIPCHK(h, qs_launch_permute_in(h->d_x, h->d_Gt, h->d_blk_src, h->nA, h->ncolS, h->nw, h->stream));
    prof_end(h, ST_PERMUTE, t0);
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    // destroys what exists and forgets it, so that a second call is harmless
    auto release = [&]() {
        for (int p = 0; p < 2; p++) {
            if (exec[p]) (void)hipGraphExecDestroy(exec[p]);
            if (graph[p]) (void)hipGraphDestroy(graph[p]);
            exec[p] = nullptr;
            graph[p] = nullptr;
            prof_release_graph(h, p);
        }
    };
    double fmax = 1.79769313486231570e308;
    bool graphs = h->step_mode == 2;
    for (int its = 1; its <= h->p.maxscf; its++) {
        *n_iter = its;
        const int par = its & 1;
        if (graphs && its > 2 && !exec[par]) {
            // one capture at a time in the process; if the capture fails anyway (work
            // of other threads can invalidate it), the solve goes on with plain
            // launches of the same sequence, which give the same results
            std::lock_guard<std::mutex> lk(g_capture_mu);
            h->prof.capture_to = &h->prof.graph_recs[par];
            hipError_t e = hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal);
            int crc = QS_OK;
            if (e == hipSuccess) {
                crc = enqueue_ctl_step(h, V);
                e = hipStreamEndCapture(h->stream, &graph[par]);
                if (crc == QS_OK && e == hipSuccess)
                    e = hipGraphInstantiate(&exec[par], graph[par], nullptr, nullptr, 0);
            }
            h->prof.capture_to = nullptr;
            if (crc != QS_OK || e != hipSuccess) {
                (void)hipGetLastError();
                h->err.clear();
                release();
                graphs = false;
            }
        }
        if (graphs && its > 2) {
            hipError_t e = hipGraphLaunch(exec[par], h->stream);
            if (e != hipSuccess) {
                release();
                return fail(h, QS_ERR_HIP, "hipGraphLaunch: %s", hipGetErrorString(e));
            }
            h->graph_steps++;
        } else if ((rc = enqueue_ctl_step(h, V))) {
            release();
            return rc;
        }
        {
            hipError_t e = hipStreamSynchronize(h->stream);
            if (e != hipSuccess) {
                release();
                return fail(h, QS_ERR_HIP, "SCF step: %s", hipGetErrorString(e));
            }
        }
        if (h->prof.on) {
            if (graphs && its > 2) prof_collect_graph(h, par);
            else prof_collect(h);
        }
        const QsCtrl &c = *h->h_ctl;
        if (c.singular) {
            release();
            return fail(h, QS_ERR_SINGULAR, "singular matrix in the Dyson inverse (zgetrf INFO != 0)");
        }
        last->runaway = c.abort;
        last->unphysical = c.unphysical;
        last->i_worst = c.i_worst;
        last->n_worst = c.n_worst;
        last->w_worst = c.w_worst;
        if (c.abort) break;
        fmax = c.fmax;
        if (lf) lf[its - 1] = fmax;
        if (lr) lr[its - 1] = std::sqrt(c.s2) / std::sqrt((double)n);
        if (c.converged) {
            *converged = 1;
            break;
        }
        std::swap(h->d_x, h->d_xold);  // x <- x_new, xold <- x
        std::swap(h->d_F, h->d_Fold);  // Fold <- F
    }
    release();
    *fmax_out = fmax;
    return QS_OK;
}

int gfail(qs_group *g, int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g->err = buf;
    return code;
}

// the member leaves the group (its stream and the group's are idle)
void group_release(qs_group *g, size_t i) {
    qs_group::Member &m = g->mem[i];
    if (m.ev) (void)hipEventDestroy(m.ev);
    m.h->group = nullptr;
    g->mem.erase(g->mem.begin() + (long)i);
    if (i < g->rc_step.size()) g->rc_step.erase(g->rc_step.begin() + (long)i);
}

}  // namespace

namespace {

// order-1 vector: Re, Im of the 2x2 block of G^<(t=0) of each interacting site (src/scf.f90:226-255)
void pack_nil(const qs_handle *h, const std::vector<zd> &G, std::vector<double> &x) {
    const int n = h->norb;
    x.assign(8 * (size_t)h->n_int, 0.0);
    size_t k = 0;
    for (int s = 0; s < h->nsites; s++) {
        if (!h->site_int[s]) continue;
        for (int b = 2 * s; b <= 2 * s + 1; b++)
            for (int a = 2 * s; a <= 2 * s + 1; a++) {
                x[k] = G[(size_t)b * n + a].real();
                x[k + 1] = G[(size_t)b * n + a].imag();
                k += 2;
            }
    }
}
void unpack_nil(const qs_handle *h, const std::vector<double> &x, std::vector<zd> &G) {
    const int n = h->norb;
    G.assign((size_t)n * n, 0.0);
    size_t k = 0;
    for (int s = 0; s < h->nsites; s++) {
        if (!h->site_int[s]) continue;
        for (int b = 2 * s; b <= 2 * s + 1; b++)
            for (int a = 2 * s; a <= 2 * s + 1; a++) {
                G[(size_t)b * n + a] = zd(x[k], x[k + 1]);
                k += 2;
            }
    }
}

// sigma_first_order (src/self_energies.f90:45-72) on the host (order 1), uploaded as (nsites, 2, 2) row-major
int upload_sigma1(qs_handle *h, const std::vector<zd> &Gt0) {
    const int n = h->norb;
    std::vector<zd> s1((size_t)4 * h->nsites, 0.0);
    h->Sig1.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; i += 2) {
        const zd hart = Gt0[(size_t)i * n + i] + Gt0[(size_t)(i + 1) * n + i + 1];
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++) {
                const zd g = Gt0[(size_t)(i + b) * n + i + a];
                const zd v = (a == b) ? zd(0, 1) * h->U[i] * (g - hart) : zd(0, 1) * h->U[i] * g;
                s1[(size_t)(i / 2) * 4 + a * 2 + b] = v;
                h->Sig1[(size_t)(i + b) * n + i + a] = v;
            }
    }
    // the host vector dies at return: finish the copy before that
    HIPCHK(h, hipMemcpyAsync(h->d_sig1, s1.data(), s1.size() * sizeof(double2), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return QS_OK;
}

// gather_onsite (src/green_functions.f90:215-238) from the summed onsL
void gather_onsite(const qs_handle *h, const std::vector<double> &sum_ons, const char *mask, std::vector<zd> &G) {
    const int n = h->norb;
    const double pp = h->p.delta / (2.0 * PI);
    for (int s = 0; s < h->nsites; s++) {
        if (mask && !mask[s]) continue;
        for (int b = 0; b < 2; b++)
            for (int a = 0; a < 2; a++) {
                const int k = 4 * s + 2 * b + a;
                G[(size_t)(2 * s + b) * n + 2 * s + a] = zd(sum_ons[2 * k], sum_ons[2 * k + 1]) * pp;
            }
    }
}

// M_SCF1 sweep + sums: gRd/gLd sanity and the on-site blocks
int sweep_scf1(qs_handle *h, bool use_sigma, std::vector<double> &sum_ons, Sanity *san) {
    int rc = run_dyson(h, QS_M_SCF1, use_sigma, h->V_scf);
    if (rc) return rc;
    if ((rc = check_status(h))) return rc;
    if ((rc = colsum(h, (const double *)h->d_onsL, 8 * h->nsites, sum_ons))) return rc;
    if (san) {
        std::vector<double> sR, sL;
        if ((rc = colsum(h, h->d_diagR, h->nA, sR))) return rc;
        if ((rc = colsum(h, h->d_diagL, h->nA, sL))) return rc;
        *san = diag_sanity(sL.data(), sR.data(), h->nA, h->p.delta, h->act.data());
    }
    return QS_OK;
}

int final_sweep(qs_handle *h, double V) {
    h->V_scf = V;
    int rc = run_dyson(h, QS_M_FINAL, h->order >= 1, V);
    if (rc) return rc;
    if ((rc = check_status(h))) return rc;
    const int n = h->norb;
    std::vector<double> sum_ons, sum_dos, sum_tr, sum_atom;
    if ((rc = colsum(h, (const double *)h->d_onsL, 8 * h->nsites, sum_ons))) return rc;
    if ((rc = colsum(h, h->d_dosd, n, sum_dos))) return rc;
    if ((rc = colsum(h, h->d_trUD, 2, sum_tr))) return rc;
    h->G_less_obs.assign((size_t)n * n, 0.0);
    gather_onsite(h, sum_ons, nullptr, h->G_less_obs);
    if (h->has_atomL) {
        if ((rc = colsum(h, (const double *)h->d_atomL, 2 * h->atom_tot, sum_atom))) return rc;
        const double pp = h->p.delta / (2.0 * PI);
        for (int at = 0; at < h->n_atoms; at++) {
            const int m = 2 * h->atom_norb[at], lo = 2 * h->atom_first[at];
            for (int b = 0; b < m; b++)
                for (int a = 0; a < m; a++) {
                    const int e = h->atom_off[at] + b * m + a;
                    h->G_less_obs[(size_t)(lo + b) * n + lo + a] = zd(sum_atom[2 * e], sum_atom[2 * e + 1]) * pp;
                }
        }
    }
    h->J_up = sum_tr[0] * h->p.delta;
    h->J_down = sum_tr[1] * h->p.delta;
    // final_sanity: G^r_ii = i dosd, G^<_ii from the on-site blocks
    std::vector<double> sImGL(n), sImGR(n);
    for (int i = 0; i < n; i++) {
        const int s = i / 2, a = i % 2, k = 4 * s + 3 * a;
        sImGL[i] = sum_ons[2 * k + 1];
        sImGR[i] = sum_dos[i];
    }
    Sanity fs = diag_sanity(sImGL.data(), sImGR.data(), n, h->p.delta, nullptr);
    h->runaway_final = fs.runaway;
    h->unphys_final = fs.unphysical;
    h->iworst_final = fs.i_worst;
    h->nworst_final = fs.n_worst;
    h->wworst_final = fs.w_worst;
    h->final_valid = true;
    return QS_OK;
}

// atoms and their intra-atom blocks of G^< (gather_atoms, src/green_functions.f90:240-257):
// first orbital, orbital count and offset of each atom's (2m)^2 block; false if
// the atoms' orbitals do not add up to the sites
bool setup_atoms(qs_handle *h, const qs_problem *p) {
    if (p->n_atoms > 0 && p->atom_norb) {
        h->n_atoms = p->n_atoms;
        h->atom_norb.assign(p->atom_norb, p->atom_norb + p->n_atoms);
    } else {
        h->n_atoms = h->nsites;
        h->atom_norb.assign(h->nsites, 1);
    }
    int acc = 0, maxn = 0;
    h->atom_first.resize(h->n_atoms);
    for (int a = 0; a < h->n_atoms; a++) {
        h->atom_first[a] = acc;
        acc += h->atom_norb[a];
        maxn = std::max(maxn, h->atom_norb[a]);
    }
    if (acc != h->nsites) return false;
    h->atom_off.assign(h->n_atoms + 1, 0);
    for (int a = 0; a < h->n_atoms; a++) h->atom_off[a + 1] = h->atom_off[a] + 4 * h->atom_norb[a] * h->atom_norb[a];
    h->has_atomL = maxn > 1;
    h->atom_tot = h->has_atomL ? h->atom_off[h->n_atoms] : 0;
    return true;
}

int check_problem(const qs_problem *p, std::string &msg) {
    if (!p) { msg = "null problem"; return QS_ERR_ARG; }
    if (p->norb < 2 || (p->norb & 1)) { msg = "norb must be even and >= 2"; return QS_ERR_ARG; }
    if (p->n_omega < 1) { msg = "n_omega must be >= 1"; return QS_ERR_ARG; }
    if (!(p->delta > 0)) { msg = "delta must be > 0"; return QS_ERR_ARG; }
    if (!p->H || !p->gammaL || !p->gammaR || !p->U_site) { msg = "H, gammaL, gammaR and U_site are required"; return QS_ERR_ARG; }
    if (p->order < 0 || p->order > 2) { msg = "order must be 0, 1 or 2"; return QS_ERR_ARG; }
    if (p->nhist < 0 || p->nhist > MAXH_DEV) { msg = "nhist must be in [0, 32]"; return QS_ERR_ARG; }
    if (p->maxscf < 1) { msg = "maxscf must be >= 1"; return QS_ERR_ARG; }
    if (p->norb > qs_dyson_big_max_norb()) {
        char b[128];
        snprintf(b, sizeof b, "norb = %d exceeds the Dyson limit of %d", p->norb, qs_dyson_big_max_norb());
        msg = b;
        return QS_ERR_UNSUPPORTED;
    }
    if (p->n_omega > (1 << 17)) { msg = "n_omega must be <= 131072"; return QS_ERR_UNSUPPORTED; }
    return QS_OK;
}

}  // namespace

extern "C" {

int qs_create(const qs_problem *p, qs_handle **out) {
    if (!out) return QS_ERR_ARG;
    *out = nullptr;
    std::string msg;
    int rc = check_problem(p, msg);
    if (rc) {
        fprintf(stderr, "qs_create: %s\n", msg.c_str());
        return rc;
    }
    qs_handle *h = new qs_handle();
    h->p = *p;
    h->norb = p->norb; h->nw = p->n_omega; h->nsites = p->norb / 2; h->order = p->order;
    const size_t nn = (size_t)h->norb * h->norb;
    h->H.assign((const zd *)p->H, (const zd *)p->H + nn);
    h->GamL.assign((const zd *)p->gammaL, (const zd *)p->gammaL + nn);
    h->GamR.assign((const zd *)p->gammaR, (const zd *)p->gammaR + nn);
    h->U.assign(p->U_site, p->U_site + h->norb);
    if (!setup_atoms(h, p)) { fprintf(stderr, "qs_create: atom_norb does not sum to norb/2\n"); delete h; return QS_ERR_ARG; }
    int m = 1;
    while ((1 << m) < 2 * h->nw - 1) m++;
    h->Nf = 1 << m;
    h->log2N2 = std::min(m / 2, 8);
    h->log2N1 = m - h->log2N2;
    h->N1 = 1 << h->log2N1;
    h->N2 = 1 << h->log2N2;
    setup_storage(h);
    if (h->nCL > 64) { fprintf(stderr, "qs_create: more than 64 left-contact states\n"); delete h; return QS_ERR_UNSUPPORTED; }
    if (h->order == 2 && (h->log2N1 > 10 || h->log2N2 < 1)) { delete h; return QS_ERR_UNSUPPORTED; }
    rc = setup_device(h);
    if (!rc) rc = tune_dyson(h);
    if (rc) {
        fprintf(stderr, "qs_create: %s\n", h->err.c_str());
        qs_destroy(h);
        return rc;
    }
    h->G_less_t0.assign(nn, 0.0);
    h->Sig1.assign(nn, 0.0);
    h->G_less_obs.assign(nn, 0.0);
    *out = h;
    return QS_OK;
}

void qs_destroy(qs_handle *h) {
    if (!h) return;
    if (h->group) {
        qs_group *g = h->group;
        (void)hipStreamSynchronize(g->stream);
        (void)hipStreamSynchronize(h->stream);
        for (size_t i = 0; i < g->mem.size(); i++)
            if (g->mem[i].h == h) {
                group_release(g, i);
                break;
            }
    }
    if (h->p.device >= 0) (void)hipSetDevice(h->p.device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (void *p : h->allocs) (void)hipFree(p);
    qs_big_destroy(h->big);
    for (auto &r : h->prof.pending) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
    for (hipEvent_t e : h->prof.pool) (void)hipEventDestroy(e);
    if (h->h_pin) (void)hipHostFree(h->h_pin);
    if (h->h_ctl) (void)hipHostFree(h->h_ctl);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

const char *qs_last_error(const qs_handle *h) { return h ? h->err.c_str() : "null handle"; }

int qs_sizes(const qs_handle *h, int32_t *nA, int32_t *nS, int32_t *nCL, int32_t *nAs, int64_t *n_vec, int32_t *nf) {
    if (!h) return QS_ERR_ARG;
    if (nA) *nA = h->nA;
    if (nS) *nS = h->nS;
    if (nCL) *nCL = h->nCL;
    if (nAs) *nAs = h->nAs;
    if (n_vec) *n_vec = h->n_vec;
    if (nf) *nf = h->Nf;
    return QS_OK;
}

int64_t qs_device_bytes(const qs_handle *h) { return h ? h->dev_bytes : 0; }

int qs_init2(qs_handle *h, double V, double *x_out) {
    if (!h) return QS_ERR_ARG;
    if (h->order != 2) return fail(h, QS_ERR_STATE, "qs_init2 needs order 2");
    HIPCHK(h, hipSetDevice(h->p.device));
    h->V_scf = V;
    int rc = run_dyson(h, QS_M_INIT2, false, V);
    if (rc) return rc;
    if ((rc = check_status(h))) return rc;
    h->x_valid = true;
    if (x_out) {
        HIPCHK(h, hipMemcpyAsync(x_out, h->d_x, h->n_vec * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
    }
    return QS_OK;
}

int qs_residual2(qs_handle *h, double V, int64_t n, const double *x, double *F, int32_t *abort_flag) {
    if (!h || !x || !F || !abort_flag) return QS_ERR_ARG;
    if (h->order != 2) return fail(h, QS_ERR_STATE, "qs_residual2 needs order 2");
    if (n != h->n_vec) return fail(h, QS_ERR_ARG, "n = %lld but n_vec = %lld", (long long)n, (long long)h->n_vec);
    HIPCHK(h, hipSetDevice(h->p.device));
    HIPCHK(h, hipMemcpyAsync(h->d_x, x, n * sizeof(double), hipMemcpyHostToDevice, h->stream));
    double st[3];
    int ab = 0;
    Sanity san;
    int rc = residual2_device(h, V, st, &ab, &san);
    if (rc) return rc;
    h->x_valid = true;
    *abort_flag = ab;
    if (ab) {
        std::memset(F, 0, n * sizeof(double));
        return QS_OK;
    }
    HIPCHK(h, hipMemcpyAsync(F, h->d_F, n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return QS_OK;
}

int qs_get_sigma(const qs_handle *hc, double *SigR2, double *SigL2, double *Sig1) {
    qs_handle *h = const_cast<qs_handle *>(hc);
    if (!h) return QS_ERR_ARG;
    HIPCHK(h, hipSetDevice(h->p.device));
    const int n = h->norb, nA = h->nA, nA2 = nA * nA;
    if (Sig1) {
        if (h->order == 2) {
            std::vector<zd> s1((size_t)4 * h->nsites);
            HIPCHK(h, hipMemcpyAsync(s1.data(), h->d_sig1, s1.size() * sizeof(double2), hipMemcpyDeviceToHost, h->stream));
            HIPCHK(h, hipStreamSynchronize(h->stream));
            zd *o = (zd *)Sig1;
            std::fill(o, o + (size_t)n * n, zd(0.0));
            for (int s = 0; s < h->nsites; s++)
                for (int a = 0; a < 2; a++)
                    for (int b = 0; b < 2; b++) o[(size_t)(2 * s + b) * n + 2 * s + a] = s1[(size_t)s * 4 + a * 2 + b];
        } else {
            std::memcpy(Sig1, h->Sig1.data(), sizeof(zd) * n * n);
        }
    }
    if ((SigR2 || SigL2) && h->order == 2) {
        if (!h->sigma_valid) return fail(h, QS_ERR_STATE, "no second-order self-energy computed yet");
        std::vector<zd> blk((size_t)h->nw * h->ncolS);
        HIPCHK(h, hipMemcpyAsync(blk.data(), h->d_Gt, blk.size() * sizeof(double2), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        zd *r = (zd *)SigR2, *l = (zd *)SigL2;
        for (int iw = 0; iw < h->nw; iw++)
            for (int j = 0; j < nA; j++)
                for (int i = 0; i < nA; i++) {
                    const size_t d = (size_t)iw * nA2 + (size_t)j * nA + i, b = (size_t)iw * h->ncolS;
                    if (r) r[d] = blk[b + sig_r_index(i, j, h->nAs)];
                    if (l) {
                        const SigL sl = sig_l_at(i, j, h->nAs);
                        const zd v = blk[b + sl.idx];
                        l[d] = sl.flip ? -std::conj(v) : v;
                    }
                }
    }
    return QS_OK;
}

int qs_final_sweep(qs_handle *h, double V, qs_scf_result *res) {
    if (!h) return QS_ERR_ARG;
    HIPCHK(h, hipSetDevice(h->p.device));
    int rc = final_sweep(h, V);
    if (rc) return rc;
    if (res) {
        res->runaway = h->runaway_final;
        res->unphysical = h->unphys_final;
        res->i_worst = h->iworst_final;
        res->n_worst = h->nworst_final;
        res->w_worst = h->wworst_final;
    }
    return QS_OK;
}

int qs_solve_scf(qs_handle *h, double V, int32_t fresh, qs_scf_result *res, double *log_fmax, double *log_rms) {
    if (!h) return QS_ERR_ARG;
    HIPCHK(h, hipSetDevice(h->p.device));
    qs_scf_result r{};
    r.converged = 1;
    h->V_scf = V;
    h->final_valid = false;
    int rc;
    if (h->order == 1) {
        const int n = h->norb;
        if (fresh) {
            // G^<(t=0) of the non-interacting system (src/scf.f90:103-108)
            h->G_less_t0.assign((size_t)n * n, 0.0);
            std::vector<double> sum_ons;
            if ((rc = sweep_scf1(h, false, sum_ons, nullptr))) return rc;
            gather_onsite(h, sum_ons, h->site_int.data(), h->G_less_t0);
        }
        std::vector<double> x;
        pack_nil(h, h->G_less_t0, x);
        int err = QS_OK;
        auto resid = [&](const std::vector<double> &xin, std::vector<double> &F, int *ab) {
            *ab = 0;
            if (err) { *ab = 1; return; }
            unpack_nil(h, xin, h->G_less_t0);
            if ((err = upload_sigma1(h, h->G_less_t0))) { *ab = 1; return; }
            std::vector<double> sum_ons;
            Sanity san;
            if ((err = sweep_scf1(h, true, sum_ons, &san))) { *ab = 1; return; }
            if (san.runaway) {
                *ab = 1;
                r.runaway = 1;
                std::fill(F.begin(), F.end(), 0.0);
                return;
            }
            std::vector<zd> Gout((size_t)n * n, 0.0);
            gather_onsite(h, sum_ons, h->site_int.data(), Gout);
            std::vector<double> xo;
            pack_nil(h, Gout, xo);
            for (size_t i = 0; i < F.size(); i++) F[i] = xo[i] - xin[i];
        };
        int conv = 0, nit = 0;
        double fmax = 0.0;
        broyden_host(resid, x, h->p.alpha_mix, h->p.nhist, h->p.maxscf, h->p.tolf_order1, h->p.single_history != 0,
                     &conv, &nit, &fmax, log_fmax, log_rms);
        if (err) return err;
        r.converged = conv;
        r.n_iter = nit;
        r.fmax = fmax;
    } else if (h->order == 2) {
        if (fresh || !h->x_valid) {
            if ((rc = run_dyson(h, QS_M_INIT2, false, V))) return rc;
            if ((rc = check_status(h))) return rc;
            h->x_valid = true;
        }
        int conv = 0, nit = 0;
        double fmax = 0.0;
        Sanity last;
        if (h->step_mode > 0 && ctl_steps_ok(h)) {
            if ((rc = broyden_device_ctl(h, V, &conv, &nit, &fmax, &last, log_fmax, log_rms))) return rc;
        } else {
            h->graph_steps = 0;
            if ((rc = broyden_device(h, V, &conv, &nit, &fmax, &last, log_fmax, log_rms))) return rc;
        }
        r.converged = conv;
        r.n_iter = nit;
        r.fmax = fmax;
        r.runaway = last.runaway;
        // the last residual's input is x (converged) or the last x before abort/maxscf
    } else {
        r.n_iter = 1;
        if (log_fmax) log_fmax[0] = 0.0;
        if (log_rms) log_rms[0] = 0.0;
    }
    if ((rc = final_sweep(h, V))) return rc;
    if (r.converged && h->unphys_final) r.converged = 0;
    r.unphysical = h->unphys_final;
    r.i_worst = h->iworst_final;
    r.n_worst = h->nworst_final;
    r.w_worst = h->wworst_final;
    if (res) *res = r;
    return QS_OK;
}

int qs_current(const qs_handle *h, double *J_up, double *J_down) {
    if (!h || !h->final_valid) return QS_ERR_STATE;
    if (J_up) *J_up = h->J_up;
    if (J_down) *J_down = h->J_down;
    return QS_OK;
}

int qs_g_less_obs(const qs_handle *h, double *G) {
    if (!h || !h->final_valid || !G) return QS_ERR_STATE;
    std::memcpy(G, h->G_less_obs.data(), sizeof(zd) * h->G_less_obs.size());
    return QS_OK;
}

int qs_dosd(const qs_handle *hc, double *dosd) {
    qs_handle *h = const_cast<qs_handle *>(hc);
    if (!h || !h->final_valid || !dosd) return QS_ERR_STATE;
    HIPCHK(h, hipSetDevice(h->p.device));
    std::vector<double> tmp((size_t)h->nw * h->norb);
    HIPCHK(h, hipMemcpyAsync(tmp.data(), h->d_dosd, tmp.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    for (int iw = 0; iw < h->nw; iw++)
        for (int i = 0; i < h->norb; i++) dosd[(size_t)iw * h->norb + i] = tmp[(size_t)iw * h->norb + i];
    return QS_OK;
}

int qs_trUD(const qs_handle *hc, double *trU, double *trD) {
    qs_handle *h = const_cast<qs_handle *>(hc);
    if (!h || !h->final_valid) return QS_ERR_STATE;
    HIPCHK(h, hipSetDevice(h->p.device));
    std::vector<double> tmp((size_t)h->nw * 2);
    HIPCHK(h, hipMemcpyAsync(tmp.data(), h->d_trUD, tmp.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    for (int iw = 0; iw < h->nw; iw++) {
        if (trU) trU[iw] = tmp[2 * (size_t)iw];
        if (trD) trD[iw] = tmp[2 * (size_t)iw + 1];
    }
    return QS_OK;
}

int qs_profile(qs_handle *h, int32_t enable) {
    if (!h) return QS_ERR_ARG;
    h->prof.on = enable != 0;
    for (int i = 0; i < ST_COUNT; i++) { h->prof.ms[i] = 0.0; h->prof.count[i] = 0; }
    return QS_OK;
}

int qs_profile_read(qs_handle *h, double *ms, int64_t *count, int32_t nstage) {
    if (!h) return QS_ERR_ARG;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    prof_collect(h);
    for (int i = 0; i < nstage && i < ST_COUNT; i++) {
        if (ms) ms[i] = h->prof.ms[i];
        if (count) count[i] = h->prof.count[i];
    }
    return ST_COUNT;
}

const char *qs_stage_name(int32_t stage) { return (stage >= 0 && stage < ST_COUNT) ? kStageNames[stage] : ""; }

int64_t qs_footprint_bytes(const qs_problem *p) {
    std::string msg;
    if (check_problem(p, msg) != QS_OK) return -1;
    qs_handle tmp;
    tmp.p = *p;
    tmp.norb = p->norb; tmp.nw = p->n_omega; tmp.nsites = p->norb / 2; tmp.order = p->order;
    const size_t nn = (size_t)tmp.norb * tmp.norb;
    tmp.H.assign((const zd *)p->H, (const zd *)p->H + nn);
    tmp.GamL.assign((const zd *)p->gammaL, (const zd *)p->gammaL + nn);
    tmp.GamR.assign((const zd *)p->gammaR, (const zd *)p->gammaR + nn);
    tmp.U.assign(p->U_site, p->U_site + tmp.norb);
    if (!setup_atoms(&tmp, p)) return -1;
    setup_storage(&tmp);
    const int64_t nw = tmp.nw, n = tmp.norb, nA = tmp.nA;
    int m = 1;
    while ((1 << m) < 2 * tmp.nw - 1) m++;
    const int64_t Nf = (int64_t)1 << m;
    // per energy: onsL, dosd, trUD, the two sanity diagonals, the Dyson partials, atom blocks
    int64_t b = nw * (4 * (n / 2) * 16 + n * 8 + 16 + 2 * std::max<int64_t>(nA, 1) * 8 + 32 +
                      (tmp.has_atomL ? (int64_t)tmp.atom_tot * 16 : 0));
    if (tmp.order == 2) {
        const int64_t nv = tmp.n_vec;
        b += 4 * nv * 8;
        if (p->nhist > 0) b += 2 * nv * (int64_t)p->nhist * (p->single_history ? 4 : 8);
        const int64_t nAs = nA / 2, ncolS = (12 * (nAs * (nAs + 1) / 2) + 7) / 8 * 8;
        b += (nw + Nf) * ncolS * 16;
        if (uses_dx(*p, (int)nA, (int)ncolS, nv)) b += nv * 8;
    }
    if (n <= qs_dyson_wave_max_norb()) b += (int64_t)wave_chunk((int)n, (int)nw) * (int64_t)qs_dyson_wave_scratch_bytes((int)n);
    else b += (int64_t)qs_dyson_big_bytes((int)n, (int)nw, kBigBudget);
    return b + ((int64_t)64 << 20);
}

int qs_device_memory(int32_t device, int64_t *free_bytes, int64_t *total_bytes) {
    if (hipSetDevice(device) != hipSuccess) return QS_ERR_HIP;
    size_t f = 0, t = 0;
    if (hipMemGetInfo(&f, &t) != hipSuccess) return QS_ERR_HIP;
    if (free_bytes) *free_bytes = (int64_t)f;
    if (total_bytes) *total_bytes = (int64_t)t;
    return QS_OK;
}

int qs_set_precision(qs_handle *h, int32_t mode, double switch_fmax) {
    if (!h) return QS_ERR_ARG;
    if (mode != 0 && mode != 1) return fail(h, QS_ERR_ARG, "precision mode must be 0 (FP64) or 1 (mixed)");
    if (mode == 1 && !(switch_fmax > 0.0)) return fail(h, QS_ERR_ARG, "the switch level must be positive");
    if (mode == 1 && h->order != 2) return fail(h, QS_ERR_STATE, "mixed precision applies to order 2");
    if (mode == 1 && h->norb > qs_dyson_big_max_norb())
        return fail(h, QS_ERR_UNSUPPORTED, "no complex-FP32 path above %d orbitals", qs_dyson_big_max_norb());
    HIPCHK(h, hipSetDevice(h->p.device));
    if (mode == 1 && !h->big) {
        size_t bytes = 0;
        h->big = qs_big_create(h->norb, h->nw, kBigBudget, &bytes);
        if (!h->big) return fail(h, QS_ERR_NOMEM, "complex-FP32 Dyson workspace");
        h->dev_bytes += (int64_t)bytes;
    }
    h->precision = mode;
    if (mode == 1) h->mixed_switch = switch_fmax;
    return QS_OK;
}

int qs_reset_problem(qs_handle *h, const qs_problem *p) {
    if (!h || !p) return QS_ERR_ARG;
    std::string msg;
    if (int rc = check_problem(p, msg)) return fail(h, rc, "%s", msg.c_str());
    if (p->norb != h->norb || p->n_omega != h->nw || p->order != h->order || p->nhist != h->p.nhist ||
        (p->single_history != 0) != (h->p.single_history != 0) || p->device != h->p.device)
        return fail(h, QS_ERR_STATE, "the new problem needs other sizes than this handle's");
    // the new system's host setup, compared with the handle's
    qs_handle t;
    t.p = *p;
    t.norb = p->norb; t.nw = p->n_omega; t.nsites = p->norb / 2; t.order = p->order;
    const size_t nn = (size_t)t.norb * t.norb;
    t.H.assign((const zd *)p->H, (const zd *)p->H + nn);
    t.GamL.assign((const zd *)p->gammaL, (const zd *)p->gammaL + nn);
    t.GamR.assign((const zd *)p->gammaR, (const zd *)p->gammaR + nn);
    t.U.assign(p->U_site, p->U_site + t.norb);
    if (!setup_atoms(&t, p)) return fail(h, QS_ERR_ARG, "atom_norb does not sum to norb/2");
    t.Nf = h->Nf;
    setup_storage(&t);
    if (t.act != h->act || t.cl != h->cl || t.sup != h->sup || t.asite != h->asite || t.site_int != h->site_int ||
        t.atom_norb != h->atom_norb)
        return fail(h, QS_ERR_STATE, "the new problem has another active subspace, contact set or atom structure");
    HIPCHK(h, hipSetDevice(h->p.device));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->p = *p;
    h->H.swap(t.H); h->GamL.swap(t.GamL); h->GamR.swap(t.GamR); h->U.swap(t.U);
    h->Hemb.swap(t.Hemb); h->GamL_s.swap(t.GamL_s); h->GamR_s.swap(t.GamR_s); h->GamL_cl.swap(t.GamL_cl);
    int rc;
    if ((rc = copy_in(h, h->d_Hemb, h->Hemb.data(), h->Hemb.size() * sizeof(double2)))) return rc;
    if ((rc = copy_in(h, h->d_gamL_s, h->GamL_s.data(), h->GamL_s.size() * sizeof(double2)))) return rc;
    if ((rc = copy_in(h, h->d_gamR_s, h->GamR_s.data(), h->GamR_s.size() * sizeof(double2)))) return rc;
    if ((rc = copy_in(h, h->d_gamL_cl, h->GamL_cl.data(), h->GamL_cl.size() * sizeof(double2)))) return rc;
    if (h->order == 2) {
        std::vector<double> uact(h->nAs);
        for (int ia = 0; ia < h->nAs; ia++) uact[ia] = h->U[2 * h->asite[ia]];
        const double pp = h->p.delta / (2.0 * PI);
        std::vector<double> upair((size_t)h->nAs * h->nAs);
        for (int ia = 0; ia < h->nAs; ia++)
            for (int ja = 0; ja < h->nAs; ja++) upair[(size_t)ia * h->nAs + ja] = uact[ia] * uact[ja] * pp * pp / h->Nf;
        if ((rc = copy_in(h, h->d_U_act, uact.data(), uact.size() * sizeof(double)))) return rc;
        if ((rc = copy_in(h, h->d_Upair, upair.data(), upair.size() * sizeof(double)))) return rc;
    }
    h->x_valid = false;
    h->sigma_valid = false;
    h->final_valid = false;
    h->G_less_t0.assign(nn, 0.0);
    h->Sig1.assign(nn, 0.0);
    h->G_less_obs.assign(nn, 0.0);
    return QS_OK;
}

int qs_set_step_mode(qs_handle *h, int32_t mode) {
    if (!h) return QS_ERR_ARG;
    if (mode < 0 || mode > 2) return fail(h, QS_ERR_ARG, "step mode must be 0, 1 or 2");
    h->step_mode = mode;
    return QS_OK;
}

int qs_step_info(const qs_handle *h, int32_t *device_steps_active, int32_t *graph_steps) {
    if (!h) return QS_ERR_ARG;
    if (device_steps_active) *device_steps_active = (h->step_mode > 0 && ctl_steps_ok(h)) ? 1 : 0;
    if (graph_steps) *graph_steps = h->graph_steps;
    return QS_OK;
}

int qs_precision_steps(const qs_handle *h, int32_t *fp32_steps, int32_t *fp64_steps) {
    if (!h) return QS_ERR_ARG;
    if (fp32_steps) *fp32_steps = h->steps32;
    if (fp64_steps) *fp64_steps = h->steps64;
    return QS_OK;
}

const char *qs_dyson_shape(const qs_handle *h, int32_t *uses_mfma) {
    if (!h) {
        if (uses_mfma) *uses_mfma = 0;
        return "";
    }
    int m = 0;
    const char *d = qs_dyson_shape_desc(h->dyson_variant, &m);
    if (uses_mfma) *uses_mfma = m;
    return d;
}

int qs_dyson_choice(const qs_handle *h, int32_t *variant, double *ms_per_variant, int32_t nmax) {
    if (!h) return QS_ERR_ARG;
    if (variant) *variant = h->dyson_variant;
    for (int i = 0; i < nmax && i < QS_DYSON_NSHAPES; i++)
        if (ms_per_variant) ms_per_variant[i] = h->dyson_tune_ms[i];
    return qs_dyson_variants(h->norb);
}

}  // extern "C"

namespace {

// the end of a member's solve as qs_solve_scf ends it: the final sweep with the
// converged (or last) Sigma, then final_sanity (src/scf.f90:138-150,260-275)
int group_finish(qs_group::Member &m, int converged, int runaway, qs_scf_result *res) {
    qs_handle *h = m.h;
    qs_scf_result r{};
    r.converged = converged;
    r.n_iter = m.its;
    r.fmax = m.fmax;
    r.runaway = runaway;
    int rc = final_sweep(h, m.V);
    if (rc == QS_OK) {
        if (r.converged && h->unphys_final) r.converged = 0;
        r.unphysical = h->unphys_final;
        r.i_worst = h->iworst_final;
        r.n_worst = h->nworst_final;
        r.w_worst = h->wworst_final;
    }
    if (res) *res = r;
    return rc;
}

}  // namespace

extern "C" {

int qs_group_create(int32_t device, qs_group **out) {
    if (!out) return QS_ERR_ARG;
    *out = nullptr;
    if (hipSetDevice(device) != hipSuccess) return QS_ERR_HIP;
    qs_group *g = new qs_group;
    g->device = device;
    bool ok = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) == hipSuccess &&
              hipHostMalloc((void **)&g->h_tab, sizeof(DysonArgs) * QS_GROUP_MAX, 0) == hipSuccess &&
              hipMalloc((void **)&g->d_tab, sizeof(DysonArgs) * QS_GROUP_MAX) == hipSuccess &&
              hipEventCreate(&g->ev[0]) == hipSuccess && hipEventCreate(&g->ev[1]) == hipSuccess;
    if (!ok) {
        qs_group_destroy(g);
        return QS_ERR_HIP;
    }
    *out = g;
    return QS_OK;
}

void qs_group_destroy(qs_group *g) {
    if (!g) return;
    (void)hipSetDevice(g->device);
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    for (auto &m : g->mem) (void)hipStreamSynchronize(m.h->stream);
    while (!g->mem.empty()) group_release(g, g->mem.size() - 1);
    for (hipEvent_t e : g->ev)
        if (e) (void)hipEventDestroy(e);
    if (g->d_tab) (void)hipFree(g->d_tab);
    if (g->h_tab) (void)hipHostFree(g->h_tab);
    if (g->stream) (void)hipStreamDestroy(g->stream);
    delete g;
}

const char *qs_group_last_error(const qs_group *g) { return g ? g->err.c_str() : "null group"; }

int qs_group_members(const qs_group *g) { return g ? (int)g->mem.size() : QS_ERR_ARG; }

int qs_group_add(qs_group *g, qs_handle *h, double V, int32_t fresh) {
    if (!g || !h) return QS_ERR_ARG;
    if (h->group) return gfail(g, QS_ERR_STATE, "the handle is already a member of a group");
    if (g->in_step) return gfail(g, QS_ERR_STATE, "solves join between steps (a step is in flight)");
    if ((int)g->mem.size() >= QS_GROUP_MAX) return gfail(g, QS_ERR_UNSUPPORTED, "a group holds at most %d solves", QS_GROUP_MAX);
    if (h->p.device != g->device) return gfail(g, QS_ERR_ARG, "the handle is on device %d, the group on %d", h->p.device, g->device);
    if (!(h->step_mode > 0 && ctl_steps_ok(h)))
        return gfail(g, QS_ERR_UNSUPPORTED, "group steps need an order-2 FP64 solve with device-controlled steps "
                                            "(FP32 history of at most 16 columns)");
    HIPCHK(h, hipSetDevice(h->p.device));
    qs_group::Member m;
    m.h = h;
    m.V = V;
    HIPCHK(h, hipEventCreateWithFlags(&m.ev, hipEventDisableTiming));
    h->group = g;
    h->V_scf = V;
    h->final_valid = false;
    h->graph_steps = 0;
    h->steps32 = 0;
    int rc = QS_OK;
    // the fresh start (src/scf.f90:114-120), as qs_solve_scf
    if (fresh || !h->x_valid) {
        if ((rc = run_dyson(h, QS_M_INIT2, false, V)) == QS_OK && (rc = check_status(h)) == QS_OK) h->x_valid = true;
    }
    hipError_t e = hipSuccess;
    if (rc == QS_OK && (e = qs_launch_ctl_reset(h->d_ctl, h->d_status, h->stream)) == hipSuccess) {
        hipEvent_t t0 = prof_begin(h);
        e = qs_launch_permute_in(h->d_x, h->d_Gt, h->d_blk_src, h->nA, h->ncolS, h->nw, h->stream);
        prof_end(h, ST_PERMUTE, t0);
    }
    if (rc == QS_OK && e != hipSuccess) rc = fail(h, QS_ERR_HIP, "group start: %s", hipGetErrorString(e));
    if (rc != QS_OK) {
        (void)hipStreamSynchronize(h->stream);
        (void)hipEventDestroy(m.ev);
        h->group = nullptr;
        return gfail(g, rc, "%s", h->err.c_str());
    }
    g->mem.push_back(m);
    return QS_OK;
}

int qs_group_step_begin(qs_group *g) {
    if (!g) return QS_ERR_ARG;
    if (g->in_step) return gfail(g, QS_ERR_STATE, "a step of this group is already in flight");
    const int nm = (int)g->mem.size();
    if (nm && hipSetDevice(g->device) != hipSuccess) return gfail(g, QS_ERR_HIP, "hipSetDevice");
    g->rc_step.assign(nm, QS_OK);
    g->in_step = true;
    if (nm == 0) return QS_OK;
    // a member whose launches fail leaves with the error; the others go on
    std::vector<int> &rcm = g->rc_step;
    // members that share a Dyson kernel instance sweep in one launch; the others
    // (the shapes without a grouped form, a slab shorter than the grid) launch on
    // their own streams
    std::vector<int> key(nm, -1);
    for (int i = 0; i < nm; i++) {
        const qs_handle *h = g->mem[i].h;
        const bool fits = h->mscr_energies >= h->nw && !(h->dyson_variant == QS_DYSON_SPLIT_MFMA && (h->nw & 1));
        key[i] = fits ? qs_dyson_group_key(h->norb, h->dyson_variant) : -1;
    }
    for (int i = 0; i < nm; i++) {
        qs_group::Member &m = g->mem[i];
        qs_handle *h = m.h;
        if ((rcm[i] = enqueue_ctl_pre(h))) continue;
        if (key[i] >= 0) {
            hipError_t e = hipEventRecord(m.ev, h->stream);
            if (e != hipSuccess) rcm[i] = fail(h, QS_ERR_HIP, "group step: %s", hipGetErrorString(e));
        } else {
            DysonArgs a = dyson_args(h, QS_M_SCF2, true, m.V);
            hipEvent_t t0 = prof_begin(h);
            hipError_t e = qs_launch_dyson_variant(a, h->nw, h->dyson_variant, h->stream);
            prof_end(h, ST_DYSON, t0);
            if (e != hipSuccess) rcm[i] = fail(h, QS_ERR_HIP, "Dyson sweep: %s", hipGetErrorString(e));
        }
    }
    std::vector<std::pair<int, int>> order;  // (key, member) of the shared sweeps
    for (int i = 0; i < nm; i++)
        if (!rcm[i] && key[i] >= 0) order.push_back({key[i], i});
    std::stable_sort(order.begin(), order.end(), [](const std::pair<int, int> &a, const std::pair<int, int> &b) {
        return a.first < b.first;
    });
    const int ntab = (int)order.size();
    for (int t = 0; t < ntab; t++)
        g->h_tab[t] = dyson_args(g->mem[order[t].second].h, QS_M_SCF2, true, g->mem[order[t].second].V);
    hipError_t e = hipSuccess;
    if (ntab) e = hipMemcpyAsync(g->d_tab, g->h_tab, sizeof(DysonArgs) * ntab, hipMemcpyHostToDevice, g->stream);
    for (int t = 0; t < ntab && e == hipSuccess; t++) e = hipStreamWaitEvent(g->stream, g->mem[order[t].second].ev, 0);
    if (e == hipSuccess) e = hipEventRecord(g->ev[0], g->stream);
    g->last_launches = nm - ntab;
    g->last_fused = ntab;
    for (int t = 0; t < ntab && e == hipSuccess;) {
        int t2 = t;
        while (t2 < ntab && order[t2].first == order[t].first) t2++;
        e = qs_launch_dyson_group(g->h_tab + t, g->d_tab + t, t2 - t, g->mem[order[t].second].h->dyson_variant,
                                  g->stream);
        g->last_launches++;
        t = t2;
    }
    if (e == hipSuccess) e = hipEventRecord(g->ev[1], g->stream);
    for (int t = 0; t < ntab && e == hipSuccess; t++) e = hipStreamWaitEvent(g->mem[order[t].second].h->stream, g->ev[1], 0);
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(g->stream);
        for (auto &m : g->mem) (void)hipStreamSynchronize(m.h->stream);
        g->in_step = false;
        return gfail(g, QS_ERR_HIP, "group Dyson sweep: %s", hipGetErrorString(e));
    }
    for (int i = 0; i < nm; i++)
        if (!rcm[i]) rcm[i] = enqueue_ctl_post(g->mem[i].h);
    return QS_OK;
}

int qs_group_step_end(qs_group *g, int32_t max_done, qs_handle **done, qs_scf_result *res, int32_t *status,
                      int32_t *n_done) {
    if (!g || !n_done) return QS_ERR_ARG;
    *n_done = 0;
    if (!g->in_step) return gfail(g, QS_ERR_STATE, "no step of this group is in flight");
    const int nm = (int)g->mem.size();
    if (nm && (max_done < nm || !done || !res || !status))
        return gfail(g, QS_ERR_ARG, "room for %d finished solves is needed", nm);
    g->in_step = false;
    if (nm == 0) return QS_OK;
    if (hipSetDevice(g->device) != hipSuccess) return gfail(g, QS_ERR_HIP, "hipSetDevice");
    std::vector<int> &rcm = g->rc_step;
    hipError_t e = hipStreamSynchronize(g->stream);
    for (int i = 0; i < nm; i++) {
        const hipError_t e2 = hipStreamSynchronize(g->mem[i].h->stream);
        if (e == hipSuccess) e = e2;
    }
    if (e != hipSuccess) return gfail(g, QS_ERR_HIP, "group step: %s", hipGetErrorString(e));
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, g->ev[0], g->ev[1]) == hipSuccess) g->dyson_ms += ms;
    g->steps++;
    // each member's step outcome, as in broyden_device_ctl
    std::vector<char> leave(nm, 0);
    for (int i = 0; i < nm; i++) {
        qs_group::Member &m = g->mem[i];
        qs_handle *h = m.h;
        if (h->prof.on) prof_collect(h);
        int fin = 0, conv = 0, runaway = 0, rc = rcm[i];
        if (rc == QS_OK) {
            m.its++;
            const QsCtrl &c = *h->h_ctl;
            if (c.singular) {
                rc = fail(h, QS_ERR_SINGULAR, "singular matrix in the Dyson inverse (zgetrf INFO != 0)");
            } else if (c.abort) {
                fin = 1;
                runaway = 1;
            } else {
                m.fmax = c.fmax;
                if (c.converged) fin = conv = 1;
                else if (m.its >= h->p.maxscf) fin = 1;
                else {
                    std::swap(h->d_x, h->d_xold);  // x <- x_new, xold <- x
                    std::swap(h->d_F, h->d_Fold);  // Fold <- F
                }
            }
        }
        if (rc == QS_OK && !fin) continue;
        const int k = *n_done;
        done[k] = h;
        res[k] = qs_scf_result{};
        if (rc == QS_OK) rc = group_finish(m, conv, runaway, &res[k]);
        status[k] = rc;
        (*n_done)++;
        leave[i] = 1;
    }
    for (int i = nm - 1; i >= 0; i--)
        if (leave[i]) group_release(g, (size_t)i);
    return QS_OK;
}

int qs_group_step(qs_group *g, int32_t max_done, qs_handle **done, qs_scf_result *res, int32_t *status,
                  int32_t *n_done) {
    if (!g || !n_done) return QS_ERR_ARG;
    *n_done = 0;
    const int nm = (int)g->mem.size();
    if (nm && (max_done < nm || !done || !res || !status))
        return gfail(g, QS_ERR_ARG, "room for %d finished solves is needed", nm);
    int rc = qs_group_step_begin(g);
    if (rc) return rc;
    return qs_group_step_end(g, max_done, done, res, status, n_done);
}

int qs_group_info(const qs_group *g, int32_t *launches, int32_t *fused_members, double *dyson_ms, int64_t *steps) {
    if (!g) return QS_ERR_ARG;
    if (launches) *launches = g->last_launches;
    if (fused_members) *fused_members = g->last_fused;
    if (dyson_ms) *dyson_ms = g->dyson_ms;
    if (steps) *steps = g->steps;
    return QS_OK;
}

}  // extern "C"
