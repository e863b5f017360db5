// CPU build of the streaming and control kernels of an SCF step (quichotte-and-
// sancho_amd/csrc/qs_vector.hip and qs_control.hip) for the host sanitizers: test
// infrastructure, see hip/hip_runtime.h.  The sources are included as they are.
//
// usage: vector_emu <nA> <nw> <nhist> <seed>
// Random inputs of one step's sizes go through the launches that follow the Dyson
// sweep in a device-controlled step, in the engine's order: the partials'
// reduction into the control block, the column sums of the diagonals, the sanity
// check, the history pass (from dx, and from x and xold: the same bits), the
// column sums of its partials, the Gram solve and the update; before them the
// equal-time G^< and Sigma1 pass.  Each result is checked against a plain host
// loop (sums to 1e-12 relative, element-wise results to the bit); a mismatch
// prints and exits 4.  Several steps run, so the history ring wraps.
#include QS_VECTOR_SOURCE
#include QS_CONTROL_SOURCE

#include <random>

static int bad = 0;
static void check_close(const char *what, double got, double want, double rel) {
    if (!(std::fabs(got - want) <= rel * std::max(1.0, std::fabs(want)))) {
        fprintf(stderr, "MISMATCH %s: %.17g vs %.17g\n", what, got, want);
        bad = 1;
    }
}

int main(int argc, char **argv) {
    if (argc != 5) return 2;
    const int nA = atoi(argv[1]), nw = atoi(argv[2]), nhist = atoi(argv[3]), seed = atoi(argv[4]);
    if (nA < 2 || nA % 2 || nw < 1 || nhist < 1 || nhist > 16) return 2;
    const int nAs = nA / 2, nblk = 3 * nA * nA;
    const int64_t n = (int64_t)nblk * nw;
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> uni(-1.0, 1.0);

    // equal-time G^< and Sigma1
    {
        std::vector<double> x(n);
        for (auto &v : x) v = uni(rng);
        std::vector<int> srcidx(4 * nAs), act(nAs);
        std::vector<double> Uact(nAs);
        std::uniform_int_distribution<int> pick(0, nblk - 1);
        for (auto &v : srcidx) v = pick(rng);
        for (int i = 0; i < nAs; i++) { act[i] = nAs - 1 - i; Uact[i] = 1.0 + i; }
        std::vector<double2> gt0(4 * nAs), sig1(4 * nAs);
        std::vector<double> part(256 * 4 * (size_t)nAs);
        const double pp = 0.01;
        if (qs_launch_equal_time_sigma1(x.data(), srcidx.data(), act.data(), Uact.data(), gt0.data(), sig1.data(), nA,
                                        nAs, nw, pp, part.data(), nullptr) != hipSuccess)
            return 3;
        for (int ia = 0; ia < nAs; ia++) {
            double S[4] = {0, 0, 0, 0};
            for (int q = 0; q < 4; q++)
                for (int iw = 0; iw < nw; iw++) S[q] += x[(size_t)iw * nblk + srcidx[4 * ia + q]];
            check_close("gt0 00", gt0[4 * ia].y, pp * S[0], 1e-12);
            check_close("gt0 11", gt0[4 * ia + 3].y, pp * S[1], 1e-12);
            check_close("gt0 01 re", gt0[4 * ia + 1].x, pp * S[2], 1e-12);
            check_close("gt0 01 im", gt0[4 * ia + 1].y, pp * S[3], 1e-12);
            const double U = Uact[ia];
            check_close("sig1 01 re", sig1[4 * act[ia] + 1].x, -U * pp * S[3], 1e-12);
            check_close("sig1 00 re", sig1[4 * act[ia]].x, U * pp * S[1], 1e-12);
        }
    }
    // column sums (narrow and wide column sets)
    for (int ncols : {nA, 2 * 64, 300}) {
        std::vector<double> in((size_t)nw * ncols), out(ncols);
        for (auto &v : in) v = uni(rng);
        std::vector<double> scr((size_t)QS_COLSUM_CHUNKS * ncols);
        if (qs_launch_reduce_cols(in.data(), ncols, nw, out.data(), scr.data(), nullptr) != hipSuccess) return 3;
        for (int c = 0; c < ncols; c++) {
            double s = 0.0;
            for (int r = 0; r < nw; r++) s += in[(size_t)r * ncols + c];
            check_close("colsum", out[c], s, 1e-12);
        }
    }

    // device-controlled steps
    const int nblocks = qs_vec_blocks(n);
    std::vector<double> x(n), xold(n), F(n), Fold(n), dx(n), xnew(n), xnew_h(n);
    for (auto &v : x) v = uni(rng);
    for (auto &v : xold) v = uni(rng);
    for (auto &v : Fold) v = 1e-2 * uni(rng);
    std::vector<float> dFs((size_t)16 * n, 0.f), Us((size_t)16 * n, 0.f);
    std::vector<float> dFs2((size_t)16 * n, 0.f), Us2((size_t)16 * n, 0.f);
    std::vector<double> partials((size_t)nw * 4), diagR((size_t)nw * nA), diagL((size_t)nw * nA);
    std::vector<double> csum(2 * nA + 64), scratch((size_t)nblocks * 64 + QS_COLSUM_CHUNKS * 64 + nw * 300);
    std::vector<double> hpart((size_t)nblocks * 64), hpart2((size_t)nblocks * 64);
    std::vector<int> label(nA), status(1, 0);
    for (int i = 0; i < nA; i++) label[i] = i;
    QsCtrl ctl;
    memset(&ctl, 0xff, sizeof ctl);
    if (qs_launch_ctl_reset(&ctl, status.data(), nullptr) != hipSuccess) return 3;
    const double alpha = 0.3, delta = 0.05;
    for (int step = 1; step <= nhist + 3; step++) {
        for (auto &v : F) v = (1e-2 / step) * uni(rng);
        // the Dyson sweep's per-energy partials and diagonal sums: occupations near 1/2, spectral weight near 1
        for (int iw = 0; iw < nw; iw++) {
            partials[4 * (size_t)iw + 0] = std::fabs(uni(rng)) / step;
            partials[4 * (size_t)iw + 1] = std::fabs(uni(rng));
            partials[4 * (size_t)iw + 2] = std::fabs(uni(rng));
            partials[4 * (size_t)iw + 3] = 0.0;
        }
        for (int64_t k = 0; k < (int64_t)nw * nA; k++) {
            diagL[k] = (0.5 + 0.01 * uni(rng)) * 2.0 * 3.14159265358979323846 / (delta * nw);
            diagR[k] = -(1.0 + 0.01 * uni(rng)) * 3.14159265358979323846 / (delta * nw);
        }
        const int its0 = ctl.its, nstored0 = ctl.nstored;
        if (qs_launch_ctl_status(status.data(), nullptr) != hipSuccess) return 3;
        if (qs_launch_reduce_partials(partials.data(), nw, &ctl.fmax, nullptr) != hipSuccess) return 3;
        double m = 0.0, s = 0.0, d = 0.0;
        for (int iw = 0; iw < nw; iw++) {
            m = std::fmax(m, partials[4 * (size_t)iw]);
            s += partials[4 * (size_t)iw + 1];
            d += partials[4 * (size_t)iw + 2];
        }
        if (ctl.fmax != m) { fprintf(stderr, "MISMATCH fmax\n"); bad = 1; }
        check_close("s2", ctl.s2, s, 1e-12);
        check_close("d2", ctl.d2, d, 1e-12);
        if (qs_launch_reduce_cols(diagR.data(), nA, nw, csum.data(), scratch.data(), nullptr) != hipSuccess) return 3;
        if (qs_launch_reduce_cols(diagL.data(), nA, nw, csum.data() + nA, scratch.data(), nullptr) != hipSuccess)
            return 3;
        if (qs_launch_ctl_check(&ctl, csum.data(), csum.data() + nA, nA, delta, label.data(), status.data(), nhist,
                                1e-30, nullptr) != hipSuccess)
            return 3;
        if (ctl.its != its0 + 1 || ctl.stop || ctl.abort) {
            fprintf(stderr, "MISMATCH control: its %d stop %d abort %d unphysical %d\n", ctl.its, ctl.stop, ctl.abort,
                    ctl.unphysical);
            return 4;
        }
        const int want_slot = step > 1 ? nstored0 % nhist : -1;
        if (ctl.slot != want_slot) { fprintf(stderr, "MISMATCH slot %d vs %d\n", ctl.slot, want_slot); bad = 1; }
        // history pass from dx and from x, xold: the same bits
        for (int64_t i = 0; i < n; i++) dx[i] = x[i] - xold[i];
        dFs2 = dFs;
        Us2 = Us;
        if (qs_launch_broyden_hist_f32_ctl(n, F.data(), Fold.data(), x.data(), xold.data(), dx.data(), alpha,
                                           dFs.data(), Us.data(), &ctl, hpart.data(), nblocks, nullptr) != hipSuccess)
            return 3;
        if (qs_launch_broyden_hist_f32_ctl(n, F.data(), Fold.data(), x.data(), xold.data(), nullptr, alpha,
                                           dFs2.data(), Us2.data(), &ctl, hpart2.data(), nblocks, nullptr) != hipSuccess)
            return 3;
        if (memcmp(dFs.data(), dFs2.data(), dFs.size() * 4) || memcmp(Us.data(), Us2.data(), Us.size() * 4) ||
            memcmp(hpart.data(), hpart2.data(), hpart.size() * 8)) {
            fprintf(stderr, "MISMATCH history pass with and without dx\n");
            return 4;
        }
        // the stored column and the Gram / projection sums against the host
        const int slot = ctl.slot, nh = ctl.nh;
        if (slot >= 0)
            for (int64_t i = 0; i < n; i++) {
                const double dd = F[i] - Fold[i];
                const float df = (float)(dd * ctl.rf);
                const float u = (float)(alpha * dd * ctl.rf + (x[i] - xold[i]) * ctl.rf);
                if (dFs[(size_t)slot * n + i] != df || Us[(size_t)slot * n + i] != u) {
                    fprintf(stderr, "MISMATCH history column at %lld\n", (long long)i);
                    return 4;
                }
            }
        if (qs_launch_reduce_cols(hpart.data(), 64, nblocks, csum.data() + 2 * nA, scratch.data() + (size_t)nblocks * 64,
                                  nullptr) != hipSuccess)
            return 3;
        for (int k = 0; k < nh; k++) {
            double g = 0.0, pr = 0.0;
            for (int64_t i = 0; i < n; i++) {
                const double hk = dFs[(size_t)k * n + i];
                g += hk * (double)(slot >= 0 ? dFs[(size_t)slot * n + i] : 0.f);
                pr += hk * F[i];
            }
            if (slot >= 0) check_close("gram", csum[2 * nA + k], g, 1e-9);
            check_close("proj", csum[2 * nA + 32 + k], pr, 1e-9);
        }
        if (qs_launch_ctl_gram(&ctl, csum.data() + 2 * nA, nhist, nullptr) != hipSuccess) return 3;
        // the update (plain kernel, gamma from the control block) against the host
        double gam[32];
        for (int k = 0; k < 32; k++) gam[k] = k < ctl.nh_use ? ctl.gamma[k] : 0.0;
        if (qs_launch_broyden_update_f32(n, x.data(), F.data(), Us.data(), gam, ctl.nh_use, alpha, xnew.data(),
                                         nullptr) != hipSuccess)
            return 3;
        for (int64_t i = 0; i < n; i++) {
            double v = x[i] + alpha * F[i];
            for (int k = 0; k < ctl.nh_use; k++) v = v - (double)Us[(size_t)k * n + i] * gam[k];
            xnew_h[i] = v;
        }
        if (memcmp(xnew.data(), xnew_h.data(), n * 8)) {
            fprintf(stderr, "MISMATCH update\n");
            return 4;
        }
        if (step > 2 && ctl.nh_use == 0) {
            fprintf(stderr, "MISMATCH: no history in use at step %d\n", step);
            bad = 1;
        }
        xold.swap(x);
        x.swap(xnew);
        Fold.swap(F);
    }
    return bad ? 4 : 0;
}
