// qs_vector.hip -- HBM-streaming kernels of the SCF iteration (FP64, gfx950).
//
//   equal-time G^<(t=0) on the active sites + first-order Sigma
//        src/green_functions.f90:192-213 (equal_time_lesser),
//        src/self_energies.f90:45-72 (sigma_first_order)
//   column sums over the energy grid (sanity sums, gathers, current)
//        common/scf_checks.f90:59-91, src/green_functions.f90:215-257,
//        src/observables.f90:22-39
//   Broyden history + Gram/projection dots in one pass, and the update pass
//        common/broyden_mixing.f90:102-178 (dF, u columns; dot_ss / dot_sd;
//        x <- x + alpha F - sum_k gamma_k u_k)
// Every reduction is two-stage with a fixed order, so results do not depend on
// scheduling.
#include "qs_device.h"
#include "qs_internal.h"

namespace {

constexpr int NT = 256;
constexpr int MAXH = 32;     // largest Broyden history handled on the device

// ---- equal-time lesser + Sigma1 ------------------------------------------------
// srcidx[4 ia + q]: offsets in an energy block of Im G<(i,i), Im G<(i+1,i+1),
// Re G<(i,i+1), Im G<(i,i+1) for the active site ia (i = 2 ia).
__global__ __launch_bounds__(NT) void k_et_partial(const double *__restrict__ x, const int *__restrict__ srcidx,
                                                   int nq, int nw, size_t nblk, double *__restrict__ part) {
    const int q = blockIdx.x * NT + threadIdx.x;
    const int chunk = blockIdx.y, nch = gridDim.y;
    if (q >= nq) return;
    const int off = srcidx[q];
    const int per = (nw + nch - 1) / nch, w0 = chunk * per, w1 = min(nw, w0 + per);
    // four interleaved sums keep four loads in flight; combined in a fixed order
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int iw = w0;
    for (; iw + 3 < w1; iw += 4) {
        const double *xp = x + (size_t)iw * nblk + off;
        s0 += xp[0];
        s1 += xp[nblk];
        s2 += xp[2 * nblk];
        s3 += xp[3 * nblk];
    }
    for (; iw < w1; iw++) s0 += x[(size_t)iw * nblk + off];
    part[(size_t)chunk * nq + q] = (s0 + s1) + (s2 + s3);
}

// one block per active site: its NT threads add the chunks' partials of the
// site's four sums, then a block reduction (fixed order)
__global__ __launch_bounds__(NT) void k_et_final(const double *__restrict__ part, int nq, int nch,
                                                 const int *__restrict__ act_site,
                                                 const double *__restrict__ U_site_act, double pp,
                                                 double2 *__restrict__ gt0, double2 *__restrict__ sig1) {
    __shared__ double scratch[16];
    const int ia = blockIdx.x;
    double S[4];
    for (int q = 0; q < 4; q++) {
        double v = 0.0;
        for (int c = threadIdx.x; c < nch; c += NT) v += part[(size_t)c * nq + 4 * ia + q];
        S[q] = block_sum(v, scratch);
    }
    if (threadIdx.x != 0) return;
    // G<(t=0) block: g00 = i pp S0, g11 = i pp S1, g01 = pp (S2 + i S3), g10 = -conj(g01)
    double2 g[4];
    g[0] = make_double2(0.0, pp * S[0]);
    g[3] = make_double2(0.0, pp * S[1]);
    g[1] = make_double2(pp * S[2], pp * S[3]);
    g[2] = make_double2(-g[1].x, g[1].y);
    const int site = act_site[ia];
    for (int q = 0; q < 4; q++) gt0[4 * ia + q] = g[q];
    const double U = U_site_act[ia];
    const double2 H = cadd(g[0], g[3]);
    // Sigma1(a,a) = i U (g_aa - H);  Sigma1(a,b) = i U g_ab   (hbar = 1)
    double2 s1[4];
    s1[0] = make_double2(-U * (g[0].y - H.y), U * (g[0].x - H.x));
    s1[3] = make_double2(-U * (g[3].y - H.y), U * (g[3].x - H.x));
    s1[1] = make_double2(-U * g[1].y, U * g[1].x);
    s1[2] = make_double2(-U * g[2].y, U * g[2].x);
    for (int q = 0; q < 4; q++) sig1[4 * site + q] = s1[q];
}

// ---- column sums ---------------------------------------------------------------
// Column sums of a [nrows][ncols] array in two launches.  Partial: block `chunk`
// sums rows [chunk * rows_per, ...) of every column; with ncols <= NT its
// threads are Q = NT / ncols row lanes of one column each (t = q * ncols + c),
// so a wavefront's loads are contiguous rows, and four interleaved sums keep
// four loads in flight.  Final: one block per column adds the chunks' partials.
// The chunking depends only on the sizes, so the sums are deterministic.
__global__ __launch_bounds__(NT) void k_colsum_partial(const double *__restrict__ in, int ncols, int nrows,
                                                       int rows_per, double *__restrict__ part) {
    __shared__ double red[NT];
    const int chunk = blockIdx.x, t = threadIdx.x;
    const int r0 = chunk * rows_per, r1 = min(nrows, r0 + rows_per);
    if (ncols <= NT) {
        const int Q = NT / ncols, c = t % ncols, q = t / ncols;
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
        if (q < Q) {
            int r = r0 + q;
            for (; r + 3 * Q < r1; r += 4 * Q) {
                const double *p = in + (size_t)r * ncols + c;
                s0 += p[0];
                s1 += p[(size_t)Q * ncols];
                s2 += p[(size_t)2 * Q * ncols];
                s3 += p[(size_t)3 * Q * ncols];
            }
            for (; r < r1; r += Q) s0 += in[(size_t)r * ncols + c];
        }
        red[t] = (s0 + s1) + (s2 + s3);
        __syncthreads();
        if (t < ncols) {
            double acc = 0.0;
            for (int k = 0; k < Q; k++) acc += red[k * ncols + t];
            part[(size_t)chunk * ncols + t] = acc;
        }
    } else {
        for (int c = t; c < ncols; c += NT) {
            double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
            int r = r0;
            for (; r + 3 < r1; r += 4) {
                const double *p = in + (size_t)r * ncols + c;
                s0 += p[0];
                s1 += p[ncols];
                s2 += p[2 * (size_t)ncols];
                s3 += p[3 * (size_t)ncols];
            }
            for (; r < r1; r++) s0 += in[(size_t)r * ncols + c];
            part[(size_t)chunk * ncols + c] = (s0 + s1) + (s2 + s3);
        }
    }
}
__global__ __launch_bounds__(NT) void k_colsum_final(const double *__restrict__ part, int ncols, int nchunks,
                                                     double *__restrict__ out) {
    __shared__ double scratch[16];
    const int c = blockIdx.x;
    double s = 0.0;
    for (int k = threadIdx.x; k < nchunks; k += NT) s += part[(size_t)k * ncols + c];
    s = block_sum(s, scratch);
    if (threadIdx.x == 0) out[c] = s;
}

// partials [nblocks][4] = (max, sum, sum, -) -> out[3]
__global__ __launch_bounds__(1024) void k_reduce_partials(const double *__restrict__ p, int nb, double *out) {
    __shared__ double scratch[16];
    double m = 0.0, s = 0.0, d = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
        m = fmax(m, p[(size_t)b * 4 + 0]);
        s += p[(size_t)b * 4 + 1];
        d += p[(size_t)b * 4 + 2];
    }
    m = block_max(m, scratch);
    s = block_sum(s, scratch);
    d = block_sum(d, scratch);
    if (threadIdx.x == 0) { out[0] = m; out[1] = s; out[2] = d; }
}

// ---- Broyden -------------------------------------------------------------------
struct Gam { double g[MAXH]; };

// One pass: store the new (dF, u) pair (slot >= 0) and accumulate, for k < nh,
//   gram_k = <dF_k, dF_slot>,  proj_k = <dF_k, F>   (double sums of FP32 history)
// partials[block][2 MAXH].
template <int MH>
__global__ __launch_bounds__(NT) void k_hist_f32(int64_t n, const double *__restrict__ F,
                                                 const double *__restrict__ Fold, const double *__restrict__ x,
                                                 const double *__restrict__ xold, double rf, double alpha,
                                                 float *__restrict__ dFs, float *__restrict__ Us, int slot, int nh,
                                                 double *__restrict__ part) {
    __shared__ double red[NT / QS_WAVE][2 * MAXH];
    double g[MH], pr[MH];
#pragma unroll
    for (int k = 0; k < MH; k++) { g[k] = 0.0; pr[k] = 0.0; }
    for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT) {
        const double f = F[i];
        float df = 0.0f;
        if (slot >= 0) {
            const double d = f - Fold[i];
            df = (float)(d * rf);
            const float u = (float)(alpha * d * rf + (x[i] - xold[i]) * rf);
            dFs[(size_t)slot * n + i] = df;
            Us[(size_t)slot * n + i] = u;
        }
#pragma unroll
        for (int k = 0; k < MH; k++)
            if (k < nh) {
                const double hk = (k == slot) ? (double)df : (double)dFs[(size_t)k * n + i];
                g[k] = fma(hk, (double)df, g[k]);
                pr[k] = fma(hk, f, pr[k]);
            }
    }
    const int w = threadIdx.x / QS_WAVE, lane = threadIdx.x & (QS_WAVE - 1);
#pragma unroll
    for (int k = 0; k < MH; k++)
        if (k < nh) {
            const double a = wave_sum(g[k]), b = wave_sum(pr[k]);
            if (lane == 0) { red[w][k] = a; red[w][MAXH + k] = b; }
        }
    __syncthreads();
    if (threadIdx.x < 2 * MAXH) {
        double s = 0.0;
        for (int ww = 0; ww < NT / QS_WAVE; ww++) s += ((threadIdx.x & (MAXH - 1)) < nh) ? red[ww][threadIdx.x] : 0.0;
        part[(size_t)blockIdx.x * 2 * MAXH + threadIdx.x] = s;
    }
}

template <int MH>
__global__ __launch_bounds__(NT) void k_hist_f64(int64_t n, const double *__restrict__ F,
                                                 const double *__restrict__ Fold, const double *__restrict__ x,
                                                 const double *__restrict__ xold, double fnorm, double alpha,
                                                 double *__restrict__ dF, double *__restrict__ U, int slot, int nh,
                                                 double *__restrict__ part) {
    __shared__ double red[NT / QS_WAVE][2 * MAXH];
    double g[MH], pr[MH];
#pragma unroll
    for (int k = 0; k < MH; k++) { g[k] = 0.0; pr[k] = 0.0; }
    for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT) {
        const double f = F[i];
        double dfn = 0.0;
        if (slot >= 0) {
            dfn = (f - Fold[i]) / fnorm;
            dF[(size_t)slot * n + i] = dfn;
            U[(size_t)slot * n + i] = alpha * dfn + (x[i] - xold[i]) / fnorm;
        }
#pragma unroll
        for (int k = 0; k < MH; k++)
            if (k < nh) {
                const double hk = (k == slot) ? dfn : dF[(size_t)k * n + i];
                g[k] = fma(hk, dfn, g[k]);
                pr[k] = fma(hk, f, pr[k]);
            }
    }
    const int w = threadIdx.x / QS_WAVE, lane = threadIdx.x & (QS_WAVE - 1);
#pragma unroll
    for (int k = 0; k < MH; k++)
        if (k < nh) {
            const double a = wave_sum(g[k]), b = wave_sum(pr[k]);
            if (lane == 0) { red[w][k] = a; red[w][MAXH + k] = b; }
        }
    __syncthreads();
    if (threadIdx.x < 2 * MAXH) {
        double s = 0.0;
        for (int ww = 0; ww < NT / QS_WAVE; ww++) s += ((threadIdx.x & (MAXH - 1)) < nh) ? red[ww][threadIdx.x] : 0.0;
        part[(size_t)blockIdx.x * 2 * MAXH + threadIdx.x] = s;
    }
}

// x_new = x + alpha F - sum_{k < nh} gamma_k u_k   (k in storage order, as the reference)
__global__ __launch_bounds__(NT) void k_update_f32(int64_t n, const double *__restrict__ x,
                                                   const double *__restrict__ F, const float *__restrict__ Us,
                                                   Gam gm, int nh, double alpha, double *__restrict__ xnew) {
    for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT) {
        double v = x[i] + alpha * F[i];
#pragma unroll
        for (int k = 0; k < MAXH; k++)
            if (k < nh) v = v - (double)Us[(size_t)k * n + i] * gm.g[k];
        xnew[i] = v;
    }
}
__global__ __launch_bounds__(NT) void k_update_f64(int64_t n, const double *__restrict__ x,
                                                   const double *__restrict__ F, const double *__restrict__ U,
                                                   Gam gm, int nh, double alpha, double *__restrict__ xnew) {
    for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT) {
        double v = x[i] + alpha * F[i];
#pragma unroll
        for (int k = 0; k < MAXH; k++)
            if (k < nh) v = v - U[(size_t)k * n + i] * gm.g[k];
        xnew[i] = v;
    }
}


// FP32 history, 4 consecutive elements per thread: every column is read with
// 16-byte loads, so each thread keeps ~30 independent loads in flight.
// n4 = n / 4 quads; the n % 4 tail is handled by the scalar kernel.
// DX: x - xold comes from dx (written by the fused update) instead of x and xold
// ctl (not null): slot, rf and nh come from the control block, and a stopped step stores nothing
template <int MH, bool DX>
__global__ __launch_bounds__(NT) void k_hist_f32x4(int64_t n, int64_t n4, const double *__restrict__ F,
                                                   const double *__restrict__ Fold, const double *__restrict__ x,
                                                   const double *__restrict__ xold, const double *__restrict__ dx,
                                                   double rf, double alpha,
                                                   float *__restrict__ dFs, float *__restrict__ Us, int slot, int nh,
                                                   double *__restrict__ part, const QsCtrl *__restrict__ ctl) {
    if (ctl) {
        if (ctl->stop) return;
        slot = ctl->slot;
        rf = ctl->rf;
        nh = ctl->nh;
    }
    __shared__ double red[NT / QS_WAVE][2 * MAXH];
    double g[MH], pr[MH];
#pragma unroll
    for (int k = 0; k < MH; k++) { g[k] = 0.0; pr[k] = 0.0; }
    for (int64_t q = (int64_t)blockIdx.x * NT + threadIdx.x; q < n4; q += (int64_t)gridDim.x * NT) {
        const int64_t i = 4 * q;
        const double4 f = *(const double4 *)(F + i);
        float4 df = make_float4(0.f, 0.f, 0.f, 0.f);
        if (slot >= 0) {
            const double4 fo = *(const double4 *)(Fold + i);
            double4 dd;
            if (DX) {
                dd = *(const double4 *)(dx + i);
            } else {
                const double4 xv = *(const double4 *)(x + i);
                const double4 xo = *(const double4 *)(xold + i);
                dd = make_double4(xv.x - xo.x, xv.y - xo.y, xv.z - xo.z, xv.w - xo.w);
            }
            const double d0 = f.x - fo.x, d1 = f.y - fo.y, d2 = f.z - fo.z, d3 = f.w - fo.w;
            df = make_float4((float)(d0 * rf), (float)(d1 * rf), (float)(d2 * rf), (float)(d3 * rf));
            const float4 u = make_float4((float)(alpha * d0 * rf + dd.x * rf), (float)(alpha * d1 * rf + dd.y * rf),
                                         (float)(alpha * d2 * rf + dd.z * rf), (float)(alpha * d3 * rf + dd.w * rf));
            *(float4 *)(dFs + (size_t)slot * n + i) = df;
            *(float4 *)(Us + (size_t)slot * n + i) = u;
        }
        float4 hv[MH];
#pragma unroll
        for (int k = 0; k < MH; k++)
            if (k < nh && k != slot) hv[k] = *(const float4 *)(dFs + (size_t)k * n + i);
#pragma unroll
        for (int k = 0; k < MH; k++)
            if (k < nh) {
                const float4 h = (k == slot) ? df : hv[k];
                g[k] = fma((double)h.x, (double)df.x, g[k]);
                g[k] = fma((double)h.y, (double)df.y, g[k]);
                g[k] = fma((double)h.z, (double)df.z, g[k]);
                g[k] = fma((double)h.w, (double)df.w, g[k]);
                pr[k] = fma((double)h.x, f.x, pr[k]);
                pr[k] = fma((double)h.y, f.y, pr[k]);
                pr[k] = fma((double)h.z, f.z, pr[k]);
                pr[k] = fma((double)h.w, f.w, pr[k]);
            }
    }
    const int w = threadIdx.x / QS_WAVE, lane = threadIdx.x & (QS_WAVE - 1);
#pragma unroll
    for (int k = 0; k < MH; k++)
        if (k < nh) {
            const double a = wave_sum(g[k]), b = wave_sum(pr[k]);
            if (lane == 0) { red[w][k] = a; red[w][MAXH + k] = b; }
        }
    __syncthreads();
    if (threadIdx.x < 2 * MAXH) {
        double s = 0.0;
        for (int ww = 0; ww < NT / QS_WAVE; ww++) s += ((threadIdx.x & (MAXH - 1)) < nh) ? red[ww][threadIdx.x] : 0.0;
        part[(size_t)blockIdx.x * 2 * MAXH + threadIdx.x] = s;
    }
}

template <int MH>
__global__ __launch_bounds__(NT) void k_update_f32x4(int64_t n, int64_t n4, const double *__restrict__ x,
                                                     const double *__restrict__ F, const float *__restrict__ Us,
                                                     Gam gm, int nh, double alpha, double *__restrict__ xnew) {
    for (int64_t q = (int64_t)blockIdx.x * NT + threadIdx.x; q < n4; q += (int64_t)gridDim.x * NT) {
        const int64_t i = 4 * q;
        const double4 xv = *(const double4 *)(x + i);
        const double4 f = *(const double4 *)(F + i);
        float4 hv[MH];
#pragma unroll
        for (int k = 0; k < MH; k++)
            if (k < nh) hv[k] = *(const float4 *)(Us + (size_t)k * n + i);
        double v0 = xv.x + alpha * f.x, v1 = xv.y + alpha * f.y, v2 = xv.z + alpha * f.z, v3 = xv.w + alpha * f.w;
#pragma unroll
        for (int k = 0; k < MH; k++)
            if (k < nh) {
                v0 = v0 - (double)hv[k].x * gm.g[k];
                v1 = v1 - (double)hv[k].y * gm.g[k];
                v2 = v2 - (double)hv[k].z * gm.g[k];
                v3 = v3 - (double)hv[k].w * gm.g[k];
            }
        *(double4 *)(xnew + i) = make_double4(v0, v1, v2, v3);
    }
}

__global__ __launch_bounds__(NT) void k_fill(double *p, int64_t n, double v) {
    for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT) p[i] = v;
}

}  // namespace

int qs_vec_blocks(int64_t n) {
    int64_t b = (n + NT - 1) / NT;
    if (b > 8192) b = 8192;
    if (b < 1) b = 1;
    return (int)b;
}

hipError_t qs_launch_equal_time_sigma1(const double *x, const int *srcidx, const int *act_site,
                                       const double *U_site_act, double2 *gt0, double2 *sig1, int nA, int nAs, int nw,
                                       double pp, double *part, hipStream_t s) {
    if (nAs <= 0) return hipSuccess;
    const int nq = 4 * nAs, nch = 256;
    const size_t nblk = 3 * (size_t)nA * nA;
    hipLaunchKernelGGL(k_et_partial, dim3((nq + NT - 1) / NT, nch), dim3(NT), 0, s, x, srcidx, nq, nw, nblk, part);
    hipLaunchKernelGGL(k_et_final, dim3(nAs), dim3(NT), 0, s, part, nq, nch, act_site, U_site_act, pp, gt0, sig1);
    return hipGetLastError();
}

int qs_colsum_chunks(int ncols, int nrows, int *rows_per) {
    // at least four rows per thread; at most QS_COLSUM_CHUNKS chunks
    const int Q = ncols <= NT ? NT / ncols : 1;
    int rp = std::max(4 * Q, (nrows + QS_COLSUM_CHUNKS - 1) / QS_COLSUM_CHUNKS);
    if (rows_per) *rows_per = rp;
    return std::max(1, (nrows + rp - 1) / rp);
}

hipError_t qs_launch_reduce_cols(const double *in, int ncols, int nrows, double *out, double *scratch,
                                 hipStream_t s) {
    if (ncols <= 0) return hipSuccess;
    int rows_per = 0;
    const int nch = qs_colsum_chunks(ncols, nrows, &rows_per);
    hipLaunchKernelGGL(k_colsum_partial, dim3(nch), dim3(NT), 0, s, in, ncols, nrows, rows_per, scratch);
    hipLaunchKernelGGL(k_colsum_final, dim3(ncols), dim3(NT), 0, s, scratch, ncols, nch, out);
    return hipGetLastError();
}

hipError_t qs_launch_reduce_partials(const double *partials, int nblocks, double *out, hipStream_t s) {
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(1024), 0, s, partials, nblocks, out);
    return hipGetLastError();
}

hipError_t qs_launch_broyden_hist_f32(int64_t n, const double *F, const double *Fold, const double *x,
                                      const double *xold, const double *dx, double rf, double alpha, float *dFs,
                                      float *Us, int slot, int nh, double *partials, int nblocks, hipStream_t s) {
    if (nh > MAXH) return hipErrorInvalidValue;
    if (n % 4 == 0 && nh <= 16 && dx)
        hipLaunchKernelGGL((k_hist_f32x4<16, true>), dim3(nblocks), dim3(NT), 0, s, n, n / 4, F, Fold, x, xold, dx, rf,
                           alpha, dFs, Us, slot, nh, partials, (const QsCtrl *)nullptr);
    else if (n % 4 == 0 && nh <= 16)
        hipLaunchKernelGGL((k_hist_f32x4<16, false>), dim3(nblocks), dim3(NT), 0, s, n, n / 4, F, Fold, x, xold,
                           nullptr, rf, alpha, dFs, Us, slot, nh, partials, (const QsCtrl *)nullptr);
    else if (nh <= 16)
        hipLaunchKernelGGL(k_hist_f32<16>, dim3(nblocks), dim3(NT), 0, s, n, F, Fold, x, xold, rf, alpha, dFs, Us, slot,
                           nh, partials);
    else
        hipLaunchKernelGGL(k_hist_f32<MAXH>, dim3(nblocks), dim3(NT), 0, s, n, F, Fold, x, xold, rf, alpha, dFs, Us,
                           slot, nh, partials);
    return hipGetLastError();
}

hipError_t qs_launch_broyden_hist_f32_ctl(int64_t n, const double *F, const double *Fold, const double *x,
                                          const double *xold, const double *dx, double alpha, float *dFs, float *Us,
                                          const QsCtrl *c, double *partials, int nblocks, hipStream_t s) {
    if (n % 4 != 0 || !c || (!dx && (!x || !xold))) return hipErrorInvalidValue;
    if (dx)
        hipLaunchKernelGGL((k_hist_f32x4<16, true>), dim3(nblocks), dim3(NT), 0, s, n, n / 4, F, Fold, nullptr, nullptr,
                           dx, 0.0, alpha, dFs, Us, -1, 0, partials, c);
    else
        hipLaunchKernelGGL((k_hist_f32x4<16, false>), dim3(nblocks), dim3(NT), 0, s, n, n / 4, F, Fold, x, xold,
                           nullptr, 0.0, alpha, dFs, Us, -1, 0, partials, c);
    return hipGetLastError();
}

hipError_t qs_launch_broyden_hist_f64(int64_t n, const double *F, const double *Fold, const double *x,
                                      const double *xold, double fnorm, double alpha, double *dF, double *U, int slot,
                                      int nh, double *partials, int nblocks, hipStream_t s) {
    if (nh > MAXH) return hipErrorInvalidValue;
    if (nh <= 16)
        hipLaunchKernelGGL(k_hist_f64<16>, dim3(nblocks), dim3(NT), 0, s, n, F, Fold, x, xold, fnorm, alpha, dF, U, slot,
                           nh, partials);
    else
        hipLaunchKernelGGL(k_hist_f64<MAXH>, dim3(nblocks), dim3(NT), 0, s, n, F, Fold, x, xold, fnorm, alpha, dF, U,
                           slot, nh, partials);
    return hipGetLastError();
}

hipError_t qs_launch_broyden_update_f32(int64_t n, const double *x, const double *F, const float *Us,
                                        const double *gamma, int nh, double alpha, double *xnew, hipStream_t s) {
    if (nh > MAXH) return hipErrorInvalidValue;
    Gam gm;
    for (int k = 0; k < MAXH; k++) gm.g[k] = k < nh ? gamma[k] : 0.0;
    if (n % 4 == 0 && nh <= 16)
        hipLaunchKernelGGL(k_update_f32x4<16>, dim3(qs_vec_blocks(n / 4)), dim3(NT), 0, s, n, n / 4, x, F, Us, gm, nh,
                           alpha, xnew);
    else
        hipLaunchKernelGGL(k_update_f32, dim3(qs_vec_blocks(n)), dim3(NT), 0, s, n, x, F, Us, gm, nh, alpha, xnew);
    return hipGetLastError();
}

hipError_t qs_launch_broyden_update_f64(int64_t n, const double *x, const double *F, const double *U,
                                        const double *gamma, int nh, double alpha, double *xnew, hipStream_t s) {
    if (nh > MAXH) return hipErrorInvalidValue;
    Gam gm;
    for (int k = 0; k < MAXH; k++) gm.g[k] = k < nh ? gamma[k] : 0.0;
    hipLaunchKernelGGL(k_update_f64, dim3(qs_vec_blocks(n)), dim3(NT), 0, s, n, x, F, U, gm, nh, alpha, xnew);
    return hipGetLastError();
}

hipError_t qs_launch_vec_stats(int64_t, const double *, const double *, double *, int, hipStream_t) {
    return hipErrorNotSupported;
}

hipError_t qs_launch_fill(double *p, int64_t n, double v, hipStream_t s) {
    hipLaunchKernelGGL(k_fill, dim3(qs_vec_blocks(n)), dim3(NT), 0, s, p, n, v);
    return hipGetLastError();
}
