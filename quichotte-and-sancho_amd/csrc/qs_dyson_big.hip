// qs_dyson_big.hip -- per-energy Dyson and Keldysh solve for systems whose matrix
// does not fit one workgroup's registers (Norb above the small-LDS kernel's 96).
//
// Same contract as k_dyson / k_dyson_w (src/scf.f90:305-415):
//   G^r = [ w + Hemb - Sigma1 - Sigma2^r(w) ]^-1
//   G^< = G^r(:,sup) [ i(f_L GamL + f_R GamR) + Sigma2^<(w) ] G^r(:,sup)^dagger
// with every epilogue mode.  The energies go in chunks through HBM, all in
// hand-written kernels:
//   k_big_build   A and S'' = SL on the support (zero elsewhere), row-major,
//                 padded to np = a multiple of 16 with an identity block
//   k_gj_inv      blocked Gauss-Jordan inverse with the zgetf2 pivot rule and
//                 MFMA trailing updates (qs_gj_mfma.hip), in place of
//                 zgetrf + zgetri (common/linear_algebra.f90:58-88)
//   k_cgemm x 2   T = G^r S'', G^< = T G^r^dagger on MFMA (src/scf.f90:343-344)
//   k_big_pack    the packing epilogues of the mode (as the other two kernels)
// With DysonArgs::fp32 the matrices, the inverse and the products are complex FP32
// (FP32 MFMA, FP32 accumulation); A and S'' are built in FP64 and rounded, and the
// packing epilogues widen the results to FP64.
// The padding block is the identity in A and zero in S'', so it stays out of
// G^r's top-left n x n block and of G^<.
#include <cstdlib>
#include <cstring>

#include "qs_device.h"
#include "qs_internal.h"

struct QsBigWork {
    double2 *A = nullptr, *S = nullptr, *T = nullptr;  // [chunk][np][np] each
    int n = 0, np = 0, chunk = 0;
};

namespace {

constexpr int NT = 256;
constexpr double TWO_PI = 6.283185307179586476925286766559;

// A(i, j) = w + Hemb - Sigma1 - Sigma2^r and S(i, j) = i (f_L GamL + f_R GamR) + Sigma2^<
// on the support (zero elsewhere), for energy iwc + blockIdx.x; row-major np x np,
// identity (A) and zero (S) past n
// the matrices are stored as CT: double2 (FP64 path) or float2 (complex FP32 path)
__device__ __forceinline__ void put(double2 *p, double2 v) { *p = v; }
__device__ __forceinline__ void put(float2 *p, double2 v) { *p = make_float2((float)v.x, (float)v.y); }
__device__ __forceinline__ double2 get(const double2 *p) { return *p; }
__device__ __forceinline__ double2 get(const float2 *p) {
    const float2 v = *p;
    return make_double2((double)v.x, (double)v.y);
}

template <typename CT>
__global__ __launch_bounds__(NT) void k_big_build(DysonArgs A, int iwc, int np, CT *__restrict__ Ab,
                                                  CT *__restrict__ Sb) {
    const int n = A.norb, iw = iwc + blockIdx.x;
    const double w = A.w_init + (double)(iw + 1) * A.delta;
    const double fL = qs_fermi(w, A.V, A.mu, A.beta), fR = qs_fermi(w, 0.0, A.mu, A.beta);
    const double2 *sigw = (A.order == 2 && A.use_sigma && A.sig) ? A.sig + (size_t)iw * A.sig_ld : nullptr;
    const bool use_s1 = A.use_sigma && A.sig1;
    CT *a = Ab + (size_t)blockIdx.x * np * np, *s = Sb + (size_t)blockIdx.x * np * np;
    for (int idx = threadIdx.x; idx < np * np; idx += NT) {
        const int i = idx / np, j = idx - i * np;  // row-major (i, j)
        if (i >= n || j >= n) {
            put(a + idx, make_double2(i == j ? 1.0 : 0.0, 0.0));
            put(s + idx, make_double2(0.0, 0.0));
            continue;
        }
        double2 v = A.Hemb[(size_t)i * n + j];
        if (i == j) v.x += w;
        if (use_s1 && (i >> 1) == (j >> 1)) v = csub(v, A.sig1[(i >> 1) * 4 + (i & 1) * 2 + (j & 1)]);
        const int pi = A.pos[i], pj = A.pos[j];
        if (sigw && pi >= 0 && pj >= 0) v = csub(v, sigw[sig_r_index(pi, pj, A.nAs)]);
        put(a + idx, v);
        const int si = A.suppos[i], sj = A.suppos[j];
        double2 sv = make_double2(0.0, 0.0);
        if (si >= 0 && sj >= 0) {
            const double2 gl = A.gamL_s[si * A.nS + sj], gr = A.gamR_s[si * A.nS + sj];
            const double re = fL * gl.x + fR * gr.x, im = fL * gl.y + fR * gr.y;
            sv = make_double2(-im, re);  // (src/scf.f90:333-336)
            if (sigw) {
                const SigL sl = sig_l_at(si, sj, A.nAs);
                sv = cadd(sv, sig_l_value(sigw[sl.idx], sl.flip));
            }
        }
        put(s + idx, sv);
    }
}

__device__ __forceinline__ double block_max_d(double v, double *red) {
    v = wave_max(v);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double m = red[0];
    for (int k = 1; k < NT / 64; k++) m = fmax(m, red[k]);
    return m;
}
__device__ __forceinline__ double block_sum_d(double v, double *red) {
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double t = red[0];
    for (int k = 1; k < NT / 64; k++) t += red[k];
    return t;
}

// the epilogue of the mode for energy iwc + blockIdx.x from G^r and G^< (row-major np x np)
template <typename CT>
__global__ __launch_bounds__(NT) void k_big_pack(DysonArgs A, int iwc, int np, const CT *__restrict__ Gr,
                                                 const CT *__restrict__ Gl) {
    __shared__ double red[NT / 64];
    const int n = A.norb, nA = A.nA, nA2 = nA * nA, nCL = A.nCL, iw = iwc + blockIdx.x, tid = threadIdx.x;
    const CT *gr = Gr + (size_t)blockIdx.x * np * np, *gl = Gl + (size_t)blockIdx.x * np * np;
    auto GR = [&](int i, int j) { return get(gr + (size_t)i * np + j); };
    // G^< is anti-Hermitian and its product forms only the tiles on and above the
    // diagonal: an element of a tile below it is -conj of its mirror
    auto GL = [&](int i, int j) {
        if (i / QS_CGEMM_TILE > j / QS_CGEMM_TILE) {
            const double2 v = get(gl + (size_t)j * np + i);
            return make_double2(-v.x, v.y);
        }
        return get(gl + (size_t)i * np + j);
    };
    const double w = A.w_init + (double)(iw + 1) * A.delta;
    const double fL = qs_fermi(w, A.V, A.mu, A.beta);
    if (A.mode == QS_M_INIT2 || A.mode == QS_M_SCF2) {
        const bool scf2 = A.mode == QS_M_SCF2;
        const size_t off = (size_t)iw * 3 * nA2;
        const int nsu = (nA2 - nA) / 2;
        double st_max = 0.0, st_s2 = 0.0, st_d2 = 0.0;
        // packed layout (common/linear_algebra.f90:142-191): Re G^r, Im G^r row-major,
        // then Re G^<(i < j), Im G^<(i <= j) row-wise (tri_map)
        for (int e = tid; e < nA2; e += NT) {
            const int ar = e / nA, bb = e - ar * nA;
            const double2 g = GR(A.act[ar], A.act[bb]);
            if (!scf2) {
                A.xout[off + e] = g.x;
                A.xout[off + nA2 + e] = g.y;
            } else {
                const double f0 = g.x - A.x[off + e], f1 = g.y - A.x[off + nA2 + e];
                A.F[off + e] = f0;
                A.F[off + nA2 + e] = f1;
                st_max = fmax(st_max, fmax(fabs(f0), fabs(f1)));
                st_s2 += f0 * f0 + f1 * f1;
                const double e0 = f0 - A.Fold[off + e], e1 = f1 - A.Fold[off + nA2 + e];
                st_d2 += e0 * e0 + e1 * e1;
                if (ar == bb) A.diagR[(size_t)iw * nA + ar] = g.y;
            }
        }
        for (int e = tid; e < nA2; e += NT) {
            const int code = A.tri_map[e], i = code >> 16, j = code & 0xffff;
            const double2 g = GL(A.act[i], A.act[j]);
            const double val = e < nsu ? g.x : g.y;
            const size_t d = off + 2 * (size_t)nA2 + e;
            if (!scf2) {
                A.xout[d] = val;
            } else {
                const double f = val - A.x[d];
                A.F[d] = f;
                st_max = fmax(st_max, fabs(f));
                st_s2 += f * f;
                const double e1 = f - A.Fold[d];
                st_d2 += e1 * e1;
                if (e >= nsu && i == j) A.diagL[(size_t)iw * nA + i] = g.y;
            }
        }
        if (scf2) {
            const double m = block_max_d(st_max, red);
            const double s2 = block_sum_d(st_s2, red);
            const double d2 = block_sum_d(st_d2, red);
            if (tid == 0) {
                double *pp = A.partials + (size_t)iw * 4;
                pp[0] = m; pp[1] = s2; pp[2] = d2; pp[3] = 0.0;
            }
        }
        return;
    }
    // SCF1 / FINAL: on-site blocks, element (2 si + a2, 2 si + b2) at 4 si + a2 + 2 b2
    for (int e = tid; e < 4 * A.nsites; e += NT) {
        const int si = e >> 2, a2 = e & 1, b2 = (e >> 1) & 1;
        A.onsL[(size_t)iw * 4 * A.nsites + e] = GL(2 * si + a2, 2 * si + b2);
    }
    if (A.mode == QS_M_SCF1) {
        for (int k = tid; k < nA; k += NT) {
            const int i = A.act[k];
            A.diagR[(size_t)iw * nA + k] = GR(i, i).y;
            A.diagL[(size_t)iw * nA + k] = GL(i, i).y;
        }
        return;
    }
    // FINAL
    for (int i = tid; i < n; i += NT) A.dosd[(size_t)iw * n + i] = GR(i, i).y;
    if (A.atomL) {
        for (int e = tid; e < A.atom_tot; e += NT) {
            int at = 0;
            while (at + 1 < A.n_atoms && A.atom_off[at + 1] <= e) at++;
            const int m = 2 * A.atom_norb[at], loc = e - A.atom_off[at];
            const int b2 = loc / m, a2 = loc - b2 * m, lo = 2 * A.atom_first[at];
            A.atomL[(size_t)iw * A.atom_tot + e] = GL(lo + a2, lo + b2);
        }
    }
    if (tid == 0) {
        // Tr_s{ i GamL [ (f_L - 1) G^< - f_L G^> ] } / 2pi on the left contact  (src/scf.f90:390-409)
        double tu = 0.0, td = 0.0;
        for (int a2 = 0; a2 < nCL; a2++) {
            double2 sacc = make_double2(0.0, 0.0);
            for (int l = 0; l < nCL; l++) {
                const double2 glc = GL(A.cl[l], A.cl[a2]);
                const double2 gr_la = GR(A.cl[l], A.cl[a2]), gr_al = GR(A.cl[a2], A.cl[l]);
                const double2 gg = cadd(glc, csub(gr_la, cconj(gr_al)));
                const double2 X = make_double2((fL - 1.0) * glc.x - fL * gg.x, (fL - 1.0) * glc.y - fL * gg.y);
                cfma(sacc, A.gamL_cl[a2 * nCL + l], X);
            }
            if ((A.cl[a2] & 1) == 0) tu += -sacc.y; else td += -sacc.y;
        }
        A.trUD[(size_t)iw * 2 + 0] = tu / TWO_PI;
        A.trUD[(size_t)iw * 2 + 1] = td / TWO_PI;
    }
}

}  // namespace

int qs_dyson_big_max_norb() { return qs_gj_max_np(); }

static size_t big_per_energy(int n) {
    const size_t np = (size_t)((n + 15) / 16 * 16);
    return 3 * np * np * sizeof(double2);
}

// energies per chunk: what the budget holds, rounded down to whole rounds of the
// inverse kernel (one matrix per CU at a time), at most the grid
static size_t big_chunk(int n, int nw, size_t budget) {
    size_t chunk = budget / big_per_energy(n);
    const size_t ncu = 256;   // MI355X compute units
    if (chunk >= ncu) chunk = chunk / ncu * ncu;
    if (chunk < 1) chunk = 1;
    if (chunk > (size_t)nw) chunk = nw;
    return chunk;
}

size_t qs_dyson_big_bytes(int n, int nw, size_t budget) { return big_chunk(n, nw, budget) * big_per_energy(n); }

QsBigWork *qs_big_create(int n, int nw, size_t budget, size_t *bytes) {
    QsBigWork *bw = new QsBigWork();
    bw->n = n;
    bw->np = (n + 15) / 16 * 16;
    const size_t chunk = big_chunk(n, nw, budget);
    bw->chunk = (int)chunk;
    const size_t mat = chunk * (size_t)bw->np * bw->np * sizeof(double2);
    bool ok = hipMalloc((void **)&bw->A, mat) == hipSuccess;
    ok = ok && hipMalloc((void **)&bw->S, mat) == hipSuccess;
    ok = ok && hipMalloc((void **)&bw->T, mat) == hipSuccess;
    if (!ok) {
        (void)hipGetLastError();
        qs_big_destroy(bw);
        return nullptr;
    }
    // QS_FILL=nan|zero: the new buffers filled before use (as the engine's own)
    if (const char *f = getenv("QS_FILL")) {
        const int v = !strcmp(f, "nan") ? 0xff : !strcmp(f, "zero") ? 0 : -1;
        if (v >= 0) {
            for (void *p : {(void *)bw->A, (void *)bw->S, (void *)bw->T}) (void)hipMemset(p, v, mat);
            (void)hipDeviceSynchronize();
        }
    }
    if (bytes) *bytes = 3 * mat;
    return bw;
}

void qs_big_destroy(QsBigWork *bw) {
    if (!bw) return;
    (void)hipFree(bw->A);
    (void)hipFree(bw->S);
    (void)hipFree(bw->T);
    delete bw;
}

hipError_t qs_launch_dyson_big(const DysonArgs &a, int count, hipStream_t s) {
    QsBigWork *bw = (QsBigWork *)a.big;
    if (count <= 0) return hipSuccess;
    if (!bw || bw->n != a.norb || a.norb > qs_dyson_big_max_norb()) return hipErrorInvalidValue;
    const int np = bw->np;
    const bool f32 = a.fp32 != 0;
    hipError_t e;
    for (int c0 = 0; c0 < count; c0 += bw->chunk) {
        const int cnt = count - c0 < bw->chunk ? count - c0 : bw->chunk, iwc = a.iw0 + c0;
        if (f32)
            hipLaunchKernelGGL(k_big_build<float2>, dim3(cnt), dim3(NT), 0, s, a, iwc, np, (float2 *)bw->A,
                               (float2 *)bw->S);
        else
            hipLaunchKernelGGL(k_big_build<double2>, dim3(cnt), dim3(NT), 0, s, a, iwc, np, bw->A, bw->S);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        // G^r into T (A is left as scratch)
        if ((e = qs_launch_gj_inv(np, cnt, bw->A, bw->T, a.status, f32, s)) != hipSuccess) return e;
        // T1 = G^r S'' into A;  G^< = T1 G^r^dagger into S, the tiles on and above the
        // diagonal only (G^< = G^r S'' G^r^dagger is anti-Hermitian, as S'' is)
        if ((e = qs_launch_cgemm(np, cnt, bw->T, bw->S, bw->A, false, f32, s)) != hipSuccess) return e;
        if ((e = qs_launch_cgemm(np, cnt, bw->A, bw->T, bw->S, true, f32, s, true)) != hipSuccess) return e;
        if (f32)
            hipLaunchKernelGGL(k_big_pack<float2>, dim3(cnt), dim3(NT), 0, s, a, iwc, np, (const float2 *)bw->T,
                               (const float2 *)bw->S);
        else
            hipLaunchKernelGGL(k_big_pack<double2>, dim3(cnt), dim3(NT), 0, s, a, iwc, np, bw->T, bw->S);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    return hipSuccess;
}
