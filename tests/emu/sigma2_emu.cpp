// CPU build of the second-order self-energy kernels (quichotte-and-sancho_amd/csrc/
// qs_sigma2.hip) for the host sanitizers: test infrastructure, see hip/hip_runtime.h.
// The kernel source is included as it is, after one textual change made by the test
// that builds this file: its dynamic LDS declaration becomes the emulated
// workgroup's LDS block.
//
// usage: sigma2_emu <nA> <nw> <seed> <out.bin> [fft]
// Random packed vectors x (3 nA^2 per energy) and a random column map run through
// the launches of one SCF step: the gather of x into the site-pair layout (the
// per-energy form and the LDS-staged persistent form, both equal to a host gather),
// the three passes of the self-energy (F1, middle, I2, in place over Gt as the
// engine runs them), and the Broyden update fused with the next gather (without and
// with a control block; its x_new must be the plain update's to the bit, its Gt the
// gather of x_new, its dx x_new - x).  A mismatch prints and exits 4.
// With "fft" only the three passes run, on the host's gather: the gather kernels
// take a workgroup per four energies or two barriers per energy, which makes long
// grids slow to emulate.
// out.bin: complex128 Sig (nw * ncol), float64 x_new (3 nA^2 nw).
#include QS_SIGMA2_SOURCE

#include <random>

template <class T>
static bool same_bits(const std::vector<T> &a, const std::vector<T> &b) {
    return a.size() == b.size() && !memcmp(a.data(), b.data(), a.size() * sizeof(T));
}

int main(int argc, char **argv) {
    if (argc != 5 && argc != 6) return 2;
    const bool fft_only = argc == 6 && !strcmp(argv[5], "fft");
    const int nA = atoi(argv[1]), nw = atoi(argv[2]), seed = atoi(argv[3]);
    if (nA < 2 || nA % 2 || nw < 2) return 2;
    const int nAs = nA / 2, npairs = nAs * (nAs + 1) / 2, ncol = (12 * npairs + 7) / 8 * 8;
    int m = 1;
    while ((1 << m) < 2 * nw - 1) m++;
    Sigma2Plan p{};
    p.nA = nA; p.nAs = nAs; p.nw = nw; p.Nf = 1 << m;
    p.log2N2 = m / 2 < 8 ? m / 2 : 8;
    p.log2N1 = m - p.log2N2;
    p.N1 = 1 << p.log2N1; p.N2 = 1 << p.log2N2; p.ncol = ncol; p.delta = 0.01;
    const int nblk = 3 * nA * nA;
    const int64_t n = (int64_t)nblk * nw;
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> uni(-1.0, 1.0);
    std::vector<double> x(n), F(n);
    for (auto &v : x) v = uni(rng);
    for (auto &v : F) v = 1e-3 * uni(rng);
    // column map: real part from +/- x[|re| - 1] or zero (re = 0), imaginary part
    // from x[im] or zero (im = -1); the padding columns are zero
    std::vector<int> src(2 * (size_t)ncol);
    std::uniform_int_distribution<int> pick(0, nblk - 1), kind(0, 5);
    for (int c = 0; c < ncol; c++) {
        if (c >= 12 * npairs) {
            src[2 * c] = 0;
            src[2 * c + 1] = -1;
            continue;
        }
        const int k = kind(rng);
        src[2 * c] = k == 0 ? 0 : (k == 1 ? -(pick(rng) + 1) : pick(rng) + 1);
        src[2 * c + 1] = k == 2 ? -1 : pick(rng);
    }
    std::vector<double2> tw(p.Nf);
    for (int k = 0; k < p.Nf; k++) {
        const long double a = 2.0L * 3.14159265358979323846264338327950288L * k / p.Nf;
        tw[k] = make_double2((double)cosl(a), (double)-sinl(a));
    }
    // pair couplings; site 0 carries no U, so its pairs take the kernel's zero path
    std::vector<double> Upair((size_t)nAs * nAs);
    for (int i = 0; i < nAs; i++)
        for (int j = 0; j < nAs; j++) Upair[(size_t)i * nAs + j] = (i == 0 || j == 0) ? 0.0 : 0.5 + 0.25 * (i + j);
    p.tw = tw.data();
    p.Upair = Upair.data();

    // gather on the host, with the per-energy kernel and with the LDS-staged persistent kernel
    std::vector<double2> Gh((size_t)nw * ncol), Gt1((size_t)nw * ncol), Gt((size_t)nw * ncol);
    for (int iw = 0; iw < nw; iw++)
        for (int c = 0; c < ncol; c++) {
            const double *xb = x.data() + (size_t)iw * nblk;
            const int re = src[2 * c], im = src[2 * c + 1];
            Gh[(size_t)iw * ncol + c] =
                make_double2(re > 0 ? xb[re - 1] : (re < 0 ? -xb[-re - 1] : 0.0), im >= 0 ? xb[im] : 0.0);
        }
    if (!slice_fits_lds(nA, ncol)) {
        fprintf(stderr, "slice does not fit the LDS-staged gather\n");
        return 2;
    }
    if (fft_only) {
        Gt = Gh;
    } else {
        const int groups = (nw + EPW - 1) / EPW;
        hipLaunchKernelGGL(k_permute_in, dim3(groups), dim3(NT), 0, nullptr, x.data(), Gt1.data(),
                           (const int2 *)src.data(), nA, ncol, nw);
        if (qs_launch_permute_in(x.data(), Gt.data(), src.data(), nA, ncol, nw, nullptr) != hipSuccess) return 3;
        if (!same_bits(Gh, Gt1) || !same_bits(Gh, Gt)) {
            fprintf(stderr, "MISMATCH: the gathers differ (per-energy %d, persistent %d)\n", (int)same_bits(Gh, Gt1),
                    (int)same_bits(Gh, Gt));
            return 4;
        }
    }
    // the self-energy, in place over Gt, one launch per pass
    std::vector<double2> Z((size_t)p.Nf * ncol);
    for (int part = 0; part < 3; part++)
        if (qs_launch_sigma2(p, Gt.data(), Z.data(), Gt.data(), 1 << part, nullptr) != hipSuccess) return 3;
    for (size_t i = 0; i < Gt.size(); i++)
        if (!std::isfinite(Gt[i].x) || !std::isfinite(Gt[i].y)) {
            fprintf(stderr, "MISMATCH: Sig[%zu] not finite\n", i);
            return 4;
        }
    std::vector<double> xnew(n);
    if (fft_only) {
        FILE *f = fopen(argv[4], "wb");
        if (!f) return 2;
        fwrite(Gt.data(), sizeof(double2), Gt.size(), f);
        fwrite(xnew.data(), sizeof(double), xnew.size(), f);
        fclose(f);
        return 0;
    }

    // fused update + gather, plain and with a control block
    const int nh = 5;
    std::vector<float> Us((size_t)16 * n);
    for (auto &v : Us) v = (float)uni(rng);
    double gamma[16] = {0};
    for (int k = 0; k < nh; k++) gamma[k] = 0.1 * uni(rng);
    std::vector<double> ref(n);
    for (int64_t i = 0; i < n; i++) {
        double v = x[i] + 0.2 * F[i];
        for (int k = 0; k < nh; k++) v = v - (double)Us[(size_t)k * n + i] * gamma[k];
        ref[i] = v;
    }
    std::vector<double> dx(n), refdx(n);
    for (int64_t i = 0; i < n; i++) refdx[i] = ref[i] - x[i];
    std::vector<double2> Gu((size_t)nw * ncol), Gref((size_t)nw * ncol);
    if (qs_launch_permute_in(ref.data(), Gref.data(), src.data(), nA, ncol, nw, nullptr) != hipSuccess) return 3;
    if (qs_launch_broyden_update_f32_gt(n, x.data(), F.data(), Us.data(), gamma, nh, 0.2, xnew.data(), dx.data(),
                                        Gu.data(), src.data(), nA, ncol, nw, nullptr) != hipSuccess)
        return 3;
    if (!same_bits(xnew, ref) || !same_bits(dx, refdx) || !same_bits(Gu, Gref)) {
        fprintf(stderr, "MISMATCH: fused update (x_new %d, dx %d, Gt %d)\n", (int)same_bits(xnew, ref),
                (int)same_bits(dx, refdx), (int)same_bits(Gu, Gref));
        return 4;
    }
    QsCtrl ctl{};
    ctl.nh_use = nh;
    ctl.its = 1;
    for (int k = 0; k < nh; k++) ctl.gamma[k] = gamma[k];
    std::vector<double> xnew2(n);
    std::vector<double2> Gu2((size_t)nw * ncol);
    if (qs_launch_broyden_update_f32_gt_ctl(n, x.data(), F.data(), Us.data(), &ctl, 10, 0.2, xnew2.data(), nullptr,
                                            Gu2.data(), src.data(), nA, ncol, nw, nullptr) != hipSuccess)
        return 3;
    if (!same_bits(xnew2, ref) || !same_bits(Gu2, Gref)) {
        fprintf(stderr, "MISMATCH: fused update with the control block\n");
        return 4;
    }
    // at maxscf the update still runs, and Gt is left alone
    ctl.its = 10;
    std::vector<double2> Gu3((size_t)nw * ncol, make_double2(7.0, 7.0));
    if (qs_launch_broyden_update_f32_gt_ctl(n, x.data(), F.data(), Us.data(), &ctl, 10, 0.2, xnew2.data(), nullptr,
                                            Gu3.data(), src.data(), nA, ncol, nw, nullptr) != hipSuccess)
        return 3;
    for (const auto &v : Gu3)
        if (v.x != 7.0 || v.y != 7.0) {
            fprintf(stderr, "MISMATCH: Gt written at maxscf\n");
            return 4;
        }
    FILE *f = fopen(argv[4], "wb");
    if (!f) return 2;
    fwrite(Gt.data(), sizeof(double2), Gt.size(), f);
    fwrite(xnew.data(), sizeof(double), xnew.size(), f);
    fclose(f);
    return 0;
}
