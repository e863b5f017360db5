// CPU build of the register-resident Dyson/Keldysh kernel (k_dyson in
// quichotte-and-sancho_amd/csrc/qs_dyson.hip) for the host sanitizers: test
// infrastructure, see hip/hip_runtime.h.  The kernel source is included as it
// is, after one textual change made by the test that builds this file: its
// dynamic LDS declaration becomes the emulated workgroup's LDS block.
//
// usage: dyson_emu <problem.bin> <out.bin> <TI> <RA> <RB>   (k_dyson<TI, RA, RB>)
//        dyson_emu <problem.bin> <out.bin> wave 0 0      (built with QS_EMU_WAVE and
//        qs_dyson_wave.hip as the source: the small-LDS kernel's own launch)
//        dyson_emu <problem.bin> <out.bin> split 0 0     (the same kernel as its
//        Gauss-Jordan and Keldysh parts, two launches through the slab)
//        dyson_emu <problem.bin> <out.bin> mfsplit 0 0   (the Gauss-Jordan kernel, then
//        the Keldysh kernel on MFMA, emulated lane by lane)
//        dyson_emu <p1.bin> <o1.bin> group|gsplit|gmfsplit <p2.bin> <o2.bin>  (two
//        problems as the members of one grouped launch: fused, split, split on MFMA)
//        dyson_emu <problem.bin> <out.bin> big|big32 0 0   (built with QS_EMU_BIG,
//        qs_dyson_big.hip as the source and qs_gj_mfma.hip linked: the large-Norb
//        chunked path in FP64, or in complex FP32, its MFMA emulated lane by lane)
// problem.bin: int32 n, count, mode, use_sigma, nCL; float64 w_init, delta,
// mu, beta, V; complex128 Hemb (n*n), GamL (n*n), GamR (n*n), GamL_cl
// (nCL*nCL); int32 cl (nCL); complex128 sig1 (n/2*4); complex128 sig
// (count*sig_ld) if use_sigma; float64 x, Fold (count*3*n*n) for mode SCF2.
// Every orbital is active and on the support (order 2 with U on every site).
// out.bin: xout or F (count*3n^2), partials (count*4), diagR, diagL
// (count*n each), onsL (count*4*(n/2) complex), dosd (count*n), trUD (count*2),
// status (int32).
#include QS_DYSON_SOURCE

#ifndef QS_EMU_WAVE
hipError_t qs_launch_dyson_wave(const DysonArgs &, int, int, hipStream_t) { return hipErrorInvalidValue; }
int qs_dyson_wave_max_norb() { return 0; }
#endif
#ifndef QS_EMU_BIG
hipError_t qs_launch_dyson_big(const DysonArgs &, int, hipStream_t) { return hipErrorInvalidValue; }
int qs_dyson_big_max_norb() { return 0; }
#endif

#include <string>

template <class T>
static std::vector<T> rd(FILE *f, size_t n) {
    std::vector<T> v(n);
    if (n && fread(v.data(), sizeof(T), n, f) != n) {
        fprintf(stderr, "short read\n");
        exit(2);
    }
    return v;
}
template <class T>
static void wr(FILE *f, const std::vector<T> &v) {
    fwrite(v.data(), sizeof(T), v.size(), f);
}

// one problem file: the arrays it holds, and the kernel arguments over them
struct Problem {
    int n = 0, count = 0, mode = 0;
    std::vector<double2> Hemb, gl, gr, glcl, sig1, sig, onsL, mscr;
    std::vector<int> cl, act, pos, suppos, tri, status;
    std::vector<double> x, Fold, xout, partials, diagR, diagL, dosd, trUD;
    DysonArgs a{};
};

static bool load(const char *path, Problem &P) {
    FILE *f = fopen(path, "rb");
    if (!f) return false;
    auto hi = rd<int>(f, 5);
    const int n = hi[0], count = hi[1], mode = hi[2], use_sigma = hi[3], nCL = hi[4];
    P.n = n; P.count = count; P.mode = mode;
    auto hd = rd<double>(f, 5);
    // Hemb keeps the library's zero pad of 64 entries past n^2
    P.Hemb = rd<double2>(f, (size_t)n * n);
    P.Hemb.resize((size_t)n * n + 64, make_double2(0.0, 0.0));
    P.gl = rd<double2>(f, (size_t)n * n);
    P.gr = rd<double2>(f, (size_t)n * n);
    P.glcl = rd<double2>(f, (size_t)nCL * nCL);
    P.cl = rd<int>(f, nCL);
    const int ns = n / 2, nAs = ns, nA = n, nA2 = n * n;
    P.sig1 = rd<double2>(f, (size_t)ns * 4);
    const int sig_ld = (12 * (nAs * (nAs + 1) / 2) + 7) / 8 * 8;
    if (use_sigma) P.sig = rd<double2>(f, (size_t)count * sig_ld);
    if (mode == QS_M_SCF2) {
        P.x = rd<double>(f, (size_t)count * 3 * nA2);
        P.Fold = rd<double>(f, (size_t)count * 3 * nA2);
    }
    fclose(f);
    P.act.resize(n); P.pos.resize(n); P.suppos.resize(n); P.tri.resize(nA2);
    for (int i = 0; i < n; i++) P.act[i] = P.pos[i] = P.suppos[i] = i;
    int e = 0;
    for (int i = 0; i < nA; i++)
        for (int j = i + 1; j < nA; j++) P.tri[e++] = (i << 16) | j;
    for (int i = 0; i < nA; i++)
        for (int j = i; j < nA; j++) P.tri[e++] = (i << 16) | j;
    const double qnan = nan("");
    P.xout.assign((size_t)count * 3 * nA2, qnan);
    P.partials.assign((size_t)count * 4, qnan);
    P.diagR.assign((size_t)count * nA, qnan);
    P.diagL.assign((size_t)count * nA, qnan);
    P.onsL.assign((size_t)count * 4 * ns, make_double2(qnan, qnan));
    P.dosd.assign((size_t)count * n, qnan);
    P.trUD.assign((size_t)count * 2, qnan);
    P.status.assign(1, 0);

    DysonArgs &a = P.a;
    a.norb = n; a.nA = nA; a.nS = n; a.nCL = nCL; a.nsites = ns; a.nAs = nAs; a.n_atoms = 0; a.atom_tot = 0;
    a.nw = count; a.iw0 = 0;
    a.w_init = hd[0]; a.delta = hd[1]; a.mu = hd[2]; a.beta = hd[3]; a.V = hd[4];
    a.mode = mode; a.use_sigma = use_sigma; a.order = 2;
    a.Hemb = P.Hemb.data(); a.sig1 = P.sig1.data(); a.sig = use_sigma ? P.sig.data() : nullptr; a.sig_ld = sig_ld;
    a.gamL_s = P.gl.data(); a.gamR_s = P.gr.data(); a.gamL_cl = P.glcl.data();
    a.act = P.act.data(); a.cl = P.cl.data(); a.pos = P.pos.data(); a.suppos = P.suppos.data();
    a.tri_map = P.tri.data();
    a.atom_first = a.atom_norb = a.atom_off = nullptr;
    a.x = P.x.empty() ? nullptr : P.x.data();
    a.xout = mode == QS_M_INIT2 ? P.xout.data() : nullptr;
    a.F = mode == QS_M_SCF2 ? P.xout.data() : nullptr;
    a.Fold = P.Fold.empty() ? nullptr : P.Fold.data();
    a.partials = P.partials.data();
    a.diagR = P.diagR.data(); a.diagL = P.diagL.data(); a.onsL = P.onsL.data(); a.dosd = P.dosd.data();
    a.atomL = nullptr; a.trUD = P.trUD.data(); a.status = P.status.data();
#ifdef QS_EMU_WAVE
    // the small-LDS kernel's slab holds M and the pivot maps per energy
    P.mscr.assign(qs_dyson_wave_scratch_bytes(n) / sizeof(double2) * count, make_double2(qnan, qnan));
    a.mscr = P.mscr.data();
    a.mscr_energies = count;
#endif
    return true;
}

static bool save(const char *path, const Problem &P) {
    FILE *o = fopen(path, "wb");
    if (!o) return false;
    wr(o, P.xout); wr(o, P.partials); wr(o, P.diagR); wr(o, P.diagL); wr(o, P.onsL); wr(o, P.dosd); wr(o, P.trUD);
    wr(o, P.status);
    fclose(o);
    return true;
}

int main(int argc, char **argv) {
    if (argc != 6) return 2;
#ifdef QS_EMU_WAVE
    // group / gsplit: the problems of argv[1] and argv[4] as the two members of
    // one fused launch (launch shape 2 / 4); outputs to argv[2] and argv[5]
    const std::string first = argv[3];
    if (first == "group" || first == "gsplit" || first == "gmfsplit") {
        Problem P[2];
        if (!load(argv[1], P[0]) || !load(argv[4], P[1])) return 2;
        DysonArgs tab[2] = {P[0].a, P[1].a};
        const int variant = first == "group" ? QS_DYSON_WAVE : first == "gsplit" ? QS_DYSON_SPLIT : QS_DYSON_SPLIT_MFMA;
        const hipError_t rc = qs_launch_dyson_group(tab, tab, 2, variant, nullptr);
        if (rc != hipSuccess) {
            fprintf(stderr, "grouped launch failed: %d\n", rc);
            return 3;
        }
        return save(argv[2], P[0]) && save(argv[5], P[1]) ? 0 : 2;
    }
#endif
    Problem P;
    if (!load(argv[1], P)) return 2;
    const int n = P.n, count = P.count;
    DysonArgs &a = P.a;
    const std::string shape = std::string(argv[3]) + "," + argv[4] + "," + argv[5];
    hipError_t rc = hipErrorInvalidValue;
#if defined(QS_EMU_BIG)
    (void)n;
    if (shape == "big,0,0" || shape == "big32,0,0") {
        size_t bytes = 0;
        QsBigWork *bw = qs_big_create(n, count, (size_t)1 << 30, &bytes);
        if (!bw) return 3;
        a.big = bw;
        a.fp32 = shape == "big32,0,0";
        rc = qs_launch_dyson_big(a, count, nullptr);
        qs_big_destroy(bw);
    }
#elif defined(QS_EMU_WAVE)
    (void)n;
    if (shape == "wave,0,0") rc = qs_launch_dyson_wave(a, count, 0, nullptr);
    else if (shape == "split,0,0") rc = qs_launch_dyson_wave(a, count, 1, nullptr);
    else if (shape == "mfsplit,0,0") rc = qs_launch_dyson_wave(a, count, 2, nullptr);
#else
    if (shape == "16,5,5") rc = launch_v<16, 5, 5>(a, count, nullptr);
    else if (shape == "32,3,5") rc = launch_v<32, 3, 5>(a, count, nullptr);
    else if (shape == "16,6,6") rc = launch_v<16, 6, 6>(a, count, nullptr);
    else if (shape == "32,3,6") rc = launch_v<32, 3, 6>(a, count, nullptr);
    else if (shape == "16,2,2") rc = launch_v<16, 2, 2>(a, count, nullptr);
#endif
    if (rc != hipSuccess) {
        fprintf(stderr, "launch of shape %s failed: %d\n", shape.c_str(), rc);
        return 3;
    }
    return save(argv[2], P) ? 0 : 2;
}
