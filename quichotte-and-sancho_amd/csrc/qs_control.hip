// qs_control.hip -- the scalar logic of an order-2 SCF step on the device.
//
// The host loop of broyden_solve (common/broyden_mixing.f90:51-184) and the
// sanity check of scf_residual2 (common/scf_checks.f90:59-91,
// src/scf.f90:180-186) take a few scalars per step: max|F|, |F - Fold|, the
// diagonal sums, nh Gram entries and projections, and a tiny SPD solve.  These
// single-workgroup kernels do that work in device memory (QsCtrl), so a whole
// step -- residual, history pass, solve, update -- is a fixed launch sequence
// that needs no host decision and can be captured as a graph; the host reads
// one small copy of QsCtrl per step.  The arithmetic and the order of the
// operations are those of the host versions in qs_engine.cpp, so the two give
// the same results.
#include "qs_device.h"
#include "qs_internal.h"

namespace {

constexpr int MAXH_C = 32;
constexpr double PI_C = 3.14159265358979323846;

__global__ void k_ctl_reset(QsCtrl *c, int *status) {
    if (threadIdx.x == 0) {
        c->its = 0;
        c->nstored = 0;
        c->nh = 0;
        c->slot = -1;
        c->nh_use = 0;
        c->stop = 0;
        c->converged = 0;
        c->abort = 0;
        c->singular = 0;
        c->unphysical = 0;
        c->i_worst = 1;
        c->fmax = 1.79769313486231570e308;
        c->s2 = 0.0;
        c->d2 = 0.0;
        c->rf = 0.0;
        c->n_worst = 0.5;
        c->w_worst = 1.0;
        *status = 0;
    }
    for (int k = threadIdx.x; k < MAXH_C * MAXH_C; k += blockDim.x) c->gram[k] = 0.0;
    for (int k = threadIdx.x; k < MAXH_C; k += blockDim.x) c->gamma[k] = 0.0;
}

__global__ void k_ctl_status(int *status) {
    if (threadIdx.x == 0) *status = 0;
}

// check_diag_sanity on the sums over omega, then the step's bookkeeping
__global__ void k_ctl_check(QsCtrl *c, const double *__restrict__ sumR, const double *__restrict__ sumL, int nb,
                            double delta, const int *__restrict__ label, const int *__restrict__ status, int nhist,
                            double tolf) {
    if (threadIdx.x != 0) return;
    const int its = c->its + 1;
    c->its = its;
    c->singular = (*status & 1) ? 1 : 0;
    // check_diag_sanity (common/scf_checks.f90:59-91)
    int ra = 0, un = 0, iw = 1, fin = 1;
    double nw_ = 0.5, ww = 1.0;
    for (int i = 0; i < nb; i++) {
        const int lab = label ? label[i] + 1 : i + 1;
        const double nocc = (delta / (2.0 * PI_C)) * sumL[i];
        const double wsum = -(delta / PI_C) * sumR[i];
        if (!isfinite(nocc) || !isfinite(wsum)) {
            ra = 1;
            iw = lab;
            nw_ = nocc;
            ww = wsum;
            fin = 0;
            break;
        }
        if (nocc < -1.0 || nocc > 2.0) ra = 1;
        if (nocc < -0.05 || nocc > 1.05 || wsum < 0.5) un = 1;
        if (fabs(nocc - 0.5) > fabs(nw_ - 0.5) || (wsum < ww && fabs(nw_ - 0.5) <= 0.55)) {
            iw = lab;
            nw_ = nocc;
            ww = wsum;
        }
    }
    c->abort = ra;
    c->unphysical = fin ? (ra || un) : un;
    c->i_worst = iw;
    c->n_worst = nw_;
    c->w_worst = ww;
    c->stop = (c->singular || ra) ? 1 : 0;
    c->converged = (!c->stop && c->fmax < tolf) ? 1 : 0;
    if (c->converged) c->stop = 1;
    c->slot = -1;
    if (!c->stop && its > 1 && nhist > 0) {
        const double fnorm = sqrt(c->d2);
        if (fnorm > 0.0) {
            c->slot = c->nstored % nhist;
            c->nstored += 1;
            c->rf = 1.0 / fnorm;
        }
    }
    c->nh = c->nstored < nhist ? c->nstored : nhist;
    if (c->stop) c->nh = 0;
}

// the Gram row of the new slot, the projections and dposv('U') on w0^2 1 + Gram
__global__ void k_ctl_gram(QsCtrl *c, const double *__restrict__ sums, int nhist) {
    if (threadIdx.x != 0) return;
    c->nh_use = 0;
    if (c->stop) return;
    const int nh = c->nh, slot = c->slot, lda = MAXH_C;
    if (nh <= 0) return;
    if (slot >= 0)
        for (int k = 0; k < nh; k++) {
            c->gram[slot * lda + k] = sums[k];
            c->gram[k * lda + slot] = sums[k];
        }
    const double w0 = 0.01;
    __shared__ double a[MAXH_C * MAXH_C];
    __shared__ double b[MAXH_C];
    for (int j = 0; j < nh; j++)
        for (int i = 0; i < nh; i++) a[j * lda + i] = c->gram[j * lda + i];
    for (int i = 0; i < nh; i++) a[i * lda + i] += w0 * w0;
    for (int k = 0; k < nh; k++) b[k] = sums[MAXH_C + k];
    bool ok = true;
    for (int j = 0; j < nh && ok; j++) {
        double ajj = a[j * lda + j];
        for (int k = 0; k < j; k++) ajj -= a[j * lda + k] * a[j * lda + k];
        if (!(ajj > 0.0)) {
            ok = false;
            break;
        }
        ajj = sqrt(ajj);
        a[j * lda + j] = ajj;
        for (int i = j + 1; i < nh; i++) {
            double t = a[i * lda + j];
            for (int k = 0; k < j; k++) t -= a[j * lda + k] * a[i * lda + k];
            a[i * lda + j] = t / ajj;
        }
    }
    if (!ok) {
        c->nstored = 0;   // broyden_solve resets the history when dposv fails
        return;
    }
    for (int i = 0; i < nh; i++) {
        double t = b[i];
        for (int k = 0; k < i; k++) t -= a[i * lda + k] * b[k];
        b[i] = t / a[i * lda + i];
    }
    for (int i = nh - 1; i >= 0; i--) {
        double t = b[i];
        for (int k = i + 1; k < nh; k++) t -= a[k * lda + i] * b[k];
        b[i] = t / a[i * lda + i];
    }
    for (int k = 0; k < nh; k++) c->gamma[k] = b[k];
    c->nh_use = nh;
}

}  // namespace

hipError_t qs_launch_ctl_reset(QsCtrl *c, int *status, hipStream_t s) {
    hipLaunchKernelGGL(k_ctl_reset, dim3(1), dim3(256), 0, s, c, status);
    return hipGetLastError();
}
hipError_t qs_launch_ctl_status(int *status, hipStream_t s) {
    hipLaunchKernelGGL(k_ctl_status, dim3(1), dim3(64), 0, s, status);
    return hipGetLastError();
}
hipError_t qs_launch_ctl_check(QsCtrl *c, const double *sumR, const double *sumL, int nb, double delta,
                               const int *label, const int *status, int nhist, double tolf, hipStream_t s) {
    hipLaunchKernelGGL(k_ctl_check, dim3(1), dim3(64), 0, s, c, sumR, sumL, nb, delta, label, status, nhist, tolf);
    return hipGetLastError();
}
hipError_t qs_launch_ctl_gram(QsCtrl *c, const double *sums, int nhist, hipStream_t s) {
    if (nhist > MAXH_C) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_ctl_gram, dim3(1), dim3(64), 0, s, c, sums, nhist);
    return hipGetLastError();
}
