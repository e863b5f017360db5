// qs_gj_mfma.hip -- batched complex inverse and products on the matrix cores
// (gfx950 MFMA), for matrices too large to stay in one workgroup's registers.
//
// Replaces the LAPACK pair zgetrf + zgetri of invert_complex
// (common/linear_algebra.f90:58-88) and the two products of G_at_energy
// (src/scf.f90:343-344, T = G^r SL, G^< = T G^r^dagger) for a batch of
// matrices, one per energy.  Matrices are row-major np x np (np a multiple of
// 16) with leading dimension np; R = double for the FP64 path, float for the
// mixed-precision one.
//
// k_gj_inv<R, NB>: blocked Gauss-Jordan with partial pivoting, one workgroup
// (4 wavefronts) per matrix.  Rows are never swapped: the pivot of column c is
// the row p_c with the largest |Re| + |Im| among the rows not yet used (the
// zgetf2 rule, on the same Schur-complement values LU sees).  A panel of NB
// columns is factored with one matrix row per thread (registers); then every
// column outside the panel takes the panel's transform at once,
//     A[i][c] <- [i not a pivot of the panel] A[i][c] + sum_k P[i][k] X[k][c],
// with P the factored panel and X the panel's pivot rows before the step: a
// rank-NB update on MFMA 16x16x4 (4 real products per complex one).  At the end
// the inverse is read out through the pivot maps, G[a][b] = M[p_a][q_b], q the
// inverse map of p.
//
// k_cgemm<R, CONJB>: C = A op(B), op(B) = B or B^dagger, 64 x 64 output tile
// per workgroup, operands staged through LDS in chunks of 16.
#include <type_traits>

#include "qs_device.h"
#include "qs_internal.h"

namespace {

template <typename R> struct Cx;
template <> struct Cx<double> { using t = double2; };
template <> struct Cx<float> { using t = float2; };

// D (16 x 16) += A (16 x 4) B (4 x 16).  Operands: lane l holds A[l % 16][l / 16]
// and B[l / 16][l % 16].  Result register q of lane l holds D[row(l / 16, q)][l % 16],
// with row = l/16 + 4 q for FP64 and 4 (l/16) + q for FP32 (measured on gfx950 with
// a one-wavefront probe of both builtins).
template <typename R> struct Mfma;
template <> struct Mfma<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 run(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int q) { return lk + 4 * q; }
};
template <> struct Mfma<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 run(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int q) { return 4 * lk + q; }
};

template <typename C> __device__ __forceinline__ C cz() { C r; r.x = 0; r.y = 0; return r; }
template <typename C> __device__ __forceinline__ C cm(C a, C b) {
    C r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
}
// a - f b
template <typename C> __device__ __forceinline__ C cms(C a, C f, C b) {
    C r;
    r.x = a.x - f.x * b.x + f.y * b.y;
    r.y = a.y - f.x * b.y - f.y * b.x;
    return r;
}
template <typename C> __device__ __forceinline__ C crecip(C a) {
    const auto d = a.x * a.x + a.y * a.y;
    C r;
    r.x = a.x / d;
    r.y = -a.y / d;
    return r;
}

constexpr int GJ_NT = 256;

// f(integral_constant<int, I>) for I = I0 .. N-1, unrolled at compile time (register
// arrays indexed by I stay in registers)
template <int I, int N, class F> __device__ __forceinline__ void static_for(F &&f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, N>(f);
    }
}

// LDS of k_gj_inv: X [NB][np] complex; the candidate rows of the four
// wavefronts, double-buffered by step parity [2][4][NB] complex; their keys
// [2][4] and row indices [2][4]; then pivrow / isR / colof [np] ints
template <typename R, int NB> size_t gj_lds_bytes(int np) {
    using C = typename Cx<R>::t;
    return (size_t)NB * np * sizeof(C) + 8 * NB * sizeof(C) + 8 * sizeof(uint64_t) + 8 * sizeof(int) +
           3 * (size_t)np * sizeof(int);
}

template <typename R, int NB>
__global__ __launch_bounds__(GJ_NT) void k_gj_inv(int np, typename Cx<R>::t *__restrict__ Mb,
                                                  typename Cx<R>::t *__restrict__ Gb, int *__restrict__ status) {
    using C = typename Cx<R>::t;
    using MF = Mfma<R>;
    using v4 = typename MF::v4;
    extern __shared__ __align__(16) char smem[];
    C *X = (C *)smem;                   // [NB][np]
    C *cand = X + (size_t)NB * np;      // [2][4][NB]: each wavefront's candidate row
    uint64_t *ckey = (uint64_t *)(cand + 8 * NB);   // [2][4]
    int *cidx = (int *)(ckey + 8);      // [2][4]
    int *pivrow = cidx + 8;             // [np]: pivot row of each column
    int *isR = pivrow + np;           // [np]: the row is a pivot of the current panel
    int *colof = isR + np;            // [np]: column whose pivot the row is
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    C *M = Mb + (size_t)blockIdx.x * np * np;
    bool used = false, singular = false;

    for (int j0 = 0; j0 < np; j0 += NB) {
        const int pw = np - j0 < NB ? np - j0 : NB;   // 16 or NB
        // --- panel rows into registers: thread t holds row t (its loads all in flight at once)
        if (t < np) isR[t] = 0;
        C row[NB];
        {
            const C *src = M + (size_t)(t < np ? t : 0) * np + j0;
#pragma unroll
            for (int c = 0; c < NB; c++) row[c] = (t < np && c < pw) ? src[c] : cz<C>();
        }
        // --- factor the panel: one Gauss-Jordan step per column, one barrier each.
        // Each wavefront posts its best candidate (key, row index, the whole row); after
        // the barrier every thread picks the same winner p.  Rows other than p take
        // row - (f r) prow with r = 1 / pivot, f = row[jj], and row[jj] = -f r.  Row p
        // would become r row with row[jj] = r; as row operations are linear, its scaling
        // is deferred to the end of the panel (it keeps row[jj] = 1 meanwhile and takes
        // the later steps like any other row), so that no wavefront runs two paths.
        // The row is rotated by one element after each step, so the step's column is
        // always element 0 and one loop body serves every step (a small code footprint).
        C rp = cz<C>();   // this thread's pending scale, if its row is a pivot of the panel
        bool pend = false;
#pragma unroll 1
        for (int jj = 0; jj < pw; jj++) {
            // key: |Re| + |Im| as ordered bits, + 1 so that any candidate beats a
            // non-candidate (0); a NaN magnitude ranks as 0
            uint64_t key = 0;
            if (t < np && !used) {
                double v = (double)(fabs(row[0].x) + fabs(row[0].y));
                if (!(v >= 0.0)) v = 0.0;
                key = (uint64_t)__double_as_longlong(v) + 1;
            }
            const uint64_t wmax = wave_max_u64(key);
            const uint64_t hit = __ballot(key == wmax);
            const int wl = __ffsll((unsigned long long)hit) - 1;   // first such lane: the first row
            const int par = (jj & 1) * 4;
            if (lane == wl) {
                C *slot = cand + (par + wv) * NB;
#pragma unroll
                for (int c = 0; c < NB; c++) slot[c] = row[c];
                ckey[par + wv] = wmax;
                cidx[par + wv] = t;
            }
            __syncthreads();
            uint64_t bk = ckey[par];
            int bw = 0;
#pragma unroll
            for (int q = 1; q < 4; q++) {
                const uint64_t kq = ckey[par + q];
                if (kq > bk) { bk = kq; bw = q; }   // wavefront order is row order: ties keep the first
            }
            const int bi = cidx[par + bw];
            const C *prow = cand + (par + bw) * NB;
            const C pv = prow[0];
            const R id = (R)1 / (pv.x * pv.x + pv.y * pv.y);
            C r;
            r.x = pv.x * id;
            r.y = -pv.y * id;
            const bool me = (t == bi);
            // the pivot row: f = 0 (no change) and element 0 = 1; the others: f = row[0] r
            C f = cm(row[0], r);
            if (me) f = cz<C>();
            C h;   // the new element 0, which moves to the end
            h.x = me ? (R)1 : -f.x;
            h.y = me ? (R)0 : -f.y;
#pragma unroll
            for (int c = 1; c < NB; c++) row[c - 1] = cms(row[c], f, prow[c]);
            row[NB - 1] = h;
            if (me) {
                if (pv.x == 0 && pv.y == 0) singular = true;   // zgetrf INFO > 0
                rp = r;
                pend = true;
                used = true;
                pivrow[j0 + jj] = t;
                isR[t] = 1;
            }
        }
        // after pw rotations element c holds column (c + pw) mod NB: undo a partial turn
        if (pw < NB) {
            static_assert(NB % 2 == 0, "panel width");
            if (pw == NB / 2) {
#pragma unroll
                for (int c = 0; c < NB / 2; c++) {
                    const C tmp = row[c];
                    row[c] = row[c + NB / 2];
                    row[c + NB / 2] = tmp;
                }
            }
        }
        if (pend) {
#pragma unroll
            for (int c = 0; c < NB; c++) row[c] = cm(row[c], rp);
        }
        // --- panel back to M; X = the panel's pivot rows before the step (coalesced rows)
        if (t < np) {
            C *dst = M + (size_t)t * np + j0;
#pragma unroll
            for (int c = 0; c < NB; c++)
                if (c < pw) dst[c] = row[c];
        }
        __syncthreads();   // pivrow of the whole panel is visible
        if (t < np) {
            const bool inpanel = t >= j0 && t < j0 + pw;
            C xv[NB];
#pragma unroll
            for (int k = 0; k < NB; k++) xv[k] = (k < pw && !inpanel) ? M[(size_t)pivrow[j0 + k] * np + t] : cz<C>();
#pragma unroll
            for (int k = 0; k < NB; k++)
                if (k < pw) X[k * np + t] = xv[k];
        }
        __syncthreads();
        // --- rank-pw update of every column block outside the panel, on MFMA:
        // D[i][c] (16 x 16) += P[i][k] X[k][c]; wavefront wv takes row blocks wv, wv + 4, ...
        const int nb16 = np >> 4, pb0 = j0 >> 4, pb1 = (j0 + pw) >> 4, ks = pw >> 2;
        const int li = lane & 15, lk = lane >> 4;
        // GEMM-PHASE-BEGIN
        // Tiles of this wavefront, row block major: (wv + 4 (q / ncb), the q % ncb-th column
        // block outside the panel).  Each tile's C and P fragments are loaded two tiles
        // ahead of its products, so two tiles' loads are in flight behind every tile's MFMAs.
        const int ncb = nb16 - (pb1 - pb0), nrbw = nb16 > wv ? (nb16 - wv + 3) / 4 : 0;
        const int ntile = ncb * nrbw;
        struct Tile {
            v4 dr, di;
            R pr[NB / 4], pi[NB / 4];
        };
        auto tile_load = [&](int q, Tile &T) {
            const int rb = wv + 4 * (q / ncb), cq = q % ncb, cb = cq < pb0 ? cq : cq + (pb1 - pb0);
            const C *pg = M + (size_t)(rb * 16 + li) * np + j0 + lk;   // P[rb*16 + li][4 s + lk]
#pragma unroll
            for (int s = 0; s < NB / 4; s++) {
                const C p = s < ks ? pg[4 * s] : cz<C>();
                T.pr[s] = p.x;
                T.pi[s] = p.y;
            }
            const C *src = M + (size_t)(rb * 16) * np + cb * 16 + li;
#pragma unroll
            for (int q2 = 0; q2 < 4; q2++) {
                const int r = MF::row(lk, q2);
                const C old = isR[rb * 16 + r] ? cz<C>() : src[(size_t)r * np];   // pivot rows start at zero
                T.dr[q2] = old.x;
                T.di[q2] = old.y;
            }
        };
        Tile t0, t1, t2;
        if (ntile > 0) tile_load(0, t0);
        if (ntile > 1) tile_load(1, t1);
        for (int q = 0; q < ntile; q++) {
            if (q + 2 < ntile) tile_load(q + 2, t2);
            const int rb = wv + 4 * (q / ncb), cq = q % ncb, cb = cq < pb0 ? cq : cq + (pb1 - pb0);
#pragma unroll
            for (int s = 0; s < NB / 4; s++) {
                if (s < ks) {
                    const C x = X[(4 * s + lk) * np + cb * 16 + li];   // B operand: X[k][c]
                    t0.dr = MF::run(t0.pr[s], x.x, t0.dr);
                    t0.di = MF::run(t0.pr[s], x.y, t0.di);
                    t0.dr = MF::run(-t0.pi[s], x.y, t0.dr);
                    t0.di = MF::run(t0.pi[s], x.x, t0.di);
                }
            }
            C *dst = M + (size_t)(rb * 16) * np + cb * 16 + li;
#pragma unroll
            for (int q2 = 0; q2 < 4; q2++) {
                C v;
                v.x = t0.dr[q2];
                v.y = t0.di[q2];
                dst[(size_t)MF::row(lk, q2) * np] = v;
            }
            t0 = t1;
            t1 = t2;
        }
        // GEMM-PHASE-END
        __syncthreads();
    }
    if (singular) atomicOr(status, 1);
    // --- read the inverse out through the pivot maps: G[a][b] = M[p_a][q_b]
    for (int c = t; c < np; c += GJ_NT) colof[pivrow[c]] = c;
    __syncthreads();
    C *G = Gb + (size_t)blockIdx.x * np * np;
    for (int e = t; e < np * np; e += GJ_NT) {
        const int a = e / np, b = e - a * np;
        G[e] = M[(size_t)pivrow[a] * np + colof[b]];
    }
}

// ---------------------------------------------------------------------------
// C = A op(B), row-major np x np, one 64 x 64 tile of C per workgroup
// ---------------------------------------------------------------------------
constexpr int GM_T = QS_CGEMM_TILE, GM_K = 16, GM_LD = GM_K + 1;

// upper: blockIdx.x numbers the nt (nt + 1) / 2 tiles (ti, tj >= ti) row by row
template <typename R, bool CONJB>
__global__ __launch_bounds__(256) void k_cgemm(int np, const typename Cx<R>::t *__restrict__ Ab,
                                               const typename Cx<R>::t *__restrict__ Bb,
                                               typename Cx<R>::t *__restrict__ Cb, int upper) {
    using C = typename Cx<R>::t;
    using MF = Mfma<R>;
    using v4 = typename MF::v4;
    __shared__ C As[GM_T * GM_LD];      // As[i][k]
    __shared__ C Bs[GM_K * GM_T];       // Bs[k][j] = op(B)[k][j]
    const int nt = (np + GM_T - 1) / GM_T;
    int ti = blockIdx.x / nt, tj = blockIdx.x - ti * nt;
    if (upper) {
        int r = 0, rest = blockIdx.x;
        while (rest >= nt - r) { rest -= nt - r; r++; }
        ti = r;
        tj = r + rest;
    }
    const int i0 = ti * GM_T, j0 = tj * GM_T;
    const size_t mo = (size_t)blockIdx.y * np * np;
    const C *A = Ab + mo, *B = Bb + mo;
    C *Cm = Cb + mo;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int wi = i0 + 32 * (wv >> 1), wj = j0 + 32 * (wv & 1);   // this wavefront's 32 x 32
    v4 accr[2][2], acci[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int q = 0; q < 4; q++) { accr[a][b][q] = 0; acci[a][b][q] = 0; }
    for (int k0 = 0; k0 < np; k0 += GM_K) {
        // A chunk: rows i0..i0+63, columns k0..k0+15 (zero past np)
        for (int e = t; e < GM_T * GM_K; e += 256) {
            const int r = e >> 4, c = e & 15;
            As[r * GM_LD + c] = (i0 + r < np) ? A[(size_t)(i0 + r) * np + k0 + c] : cz<C>();
        }
        for (int e = t; e < GM_T * GM_K; e += 256) {
            if (!CONJB) {
                const int k = e >> 6, c = e & 63;   // B[k0 + k][j0 + c]
                Bs[k * GM_T + c] = (j0 + c < np) ? B[(size_t)(k0 + k) * np + j0 + c] : cz<C>();
            } else {
                const int c = e >> 4, k = e & 15;   // conj(B[j0 + c][k0 + k])
                C v = (j0 + c < np) ? B[(size_t)(j0 + c) * np + k0 + k] : cz<C>();
                v.y = -v.y;
                Bs[k * GM_T + c] = v;
            }
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < GM_K / 4; s++) {
            C a[2], b[2];
#pragma unroll
            for (int q = 0; q < 2; q++) a[q] = As[(wi - i0 + 16 * q + li) * GM_LD + 4 * s + lk];
#pragma unroll
            for (int q = 0; q < 2; q++) b[q] = Bs[(4 * s + lk) * GM_T + wj - j0 + 16 * q + li];
#pragma unroll
            for (int x = 0; x < 2; x++)
#pragma unroll
                for (int y = 0; y < 2; y++) {
                    accr[x][y] = MF::run(a[x].x, b[y].x, accr[x][y]);
                    accr[x][y] = MF::run(-a[x].y, b[y].y, accr[x][y]);
                    acci[x][y] = MF::run(a[x].x, b[y].y, acci[x][y]);
                    acci[x][y] = MF::run(a[x].y, b[y].x, acci[x][y]);
                }
        }
        __syncthreads();
    }
    // D[i][j]: lane holds rows MF::row(lk, q), column li of each 16 x 16 tile
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++) {
            const int col = wj + 16 * y + li;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int r = wi + 16 * x + MF::row(lk, q);
                if (r < np && col < np) {
                    C v;
                    v.x = accr[x][y][q];
                    v.y = acci[x][y][q];
                    Cm[(size_t)r * np + col] = v;
                }
            }
        }
}

template <typename R, int NB> hipError_t launch_gj(int np, int count, void *M, void *G, int *status, hipStream_t s) {
    using C = typename Cx<R>::t;
    if (hipError_t e = qs_lds_attr((const void *)k_gj_inv<R, NB>)) return e;
    const size_t lds = gj_lds_bytes<R, NB>(np);
    hipLaunchKernelGGL((k_gj_inv<R, NB>), dim3(count), dim3(GJ_NT), lds, s, np, (C *)M, (C *)G, status);
    return hipGetLastError();
}

}  // namespace

int qs_gj_max_np() { return 256; }

size_t qs_gj_lds_bytes(int np, bool fp32) {
    return fp32 ? gj_lds_bytes<float, 32>(np) : gj_lds_bytes<double, 32>(np);
}

hipError_t qs_launch_gj_inv(int np, int count, void *M, void *G, int *status, bool fp32, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    if (np <= 0 || np > qs_gj_max_np() || (np & 15)) return hipErrorInvalidValue;
    return fp32 ? launch_gj<float, 32>(np, count, M, G, status, s) : launch_gj<double, 32>(np, count, M, G, status, s);
}

hipError_t qs_launch_cgemm(int np, int count, const void *A, const void *B, void *Cm, bool conjB, bool fp32,
                           hipStream_t s, bool upper) {
    if (count <= 0) return hipSuccess;
    if (np <= 0 || (np & 15)) return hipErrorInvalidValue;
    const int nt = (np + GM_T - 1) / GM_T, up = upper ? 1 : 0;
    const dim3 grid(upper ? nt * (nt + 1) / 2 : nt * nt, count);
    if (fp32) {
        if (conjB)
            hipLaunchKernelGGL((k_cgemm<float, true>), grid, dim3(256), 0, s, np, (const float2 *)A, (const float2 *)B,
                               (float2 *)Cm, up);
        else
            hipLaunchKernelGGL((k_cgemm<float, false>), grid, dim3(256), 0, s, np, (const float2 *)A, (const float2 *)B,
                               (float2 *)Cm, up);
    } else {
        if (conjB)
            hipLaunchKernelGGL((k_cgemm<double, true>), grid, dim3(256), 0, s, np, (const double2 *)A,
                               (const double2 *)B, (double2 *)Cm, up);
        else
            hipLaunchKernelGGL((k_cgemm<double, false>), grid, dim3(256), 0, s, np, (const double2 *)A,
                               (const double2 *)B, (double2 *)Cm, up);
    }
    return hipGetLastError();
}
