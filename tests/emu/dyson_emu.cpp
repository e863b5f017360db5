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
hipError_t qs_launch_dyson_wave(const DysonArgs &, int, bool, hipStream_t) { return hipErrorInvalidValue; }
int qs_dyson_wave_max_norb() { return 0; }
#endif
hipError_t qs_launch_dyson_big(const DysonArgs &, int, hipStream_t) { return hipErrorInvalidValue; }
int qs_dyson_big_max_norb() { return 0; }

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

int main(int argc, char **argv) {
    if (argc != 6) return 2;
    FILE *f = fopen(argv[1], "rb");
    if (!f) return 2;
    auto hi = rd<int>(f, 5);
    const int n = hi[0], count = hi[1], mode = hi[2], use_sigma = hi[3], nCL = hi[4];
    auto hd = rd<double>(f, 5);
    // Hemb keeps the library's zero pad of 64 entries past n^2
    auto Hemb = rd<double2>(f, (size_t)n * n);
    Hemb.resize((size_t)n * n + 64, make_double2(0.0, 0.0));
    auto gl = rd<double2>(f, (size_t)n * n), gr = rd<double2>(f, (size_t)n * n);
    auto glcl = rd<double2>(f, (size_t)nCL * nCL);
    auto cl = rd<int>(f, nCL);
    const int ns = n / 2, nAs = ns, nA = n, nA2 = n * n;
    auto sig1 = rd<double2>(f, (size_t)ns * 4);
    const int sig_ld = (12 * (nAs * (nAs + 1) / 2) + 7) / 8 * 8;
    std::vector<double2> sig;
    if (use_sigma) sig = rd<double2>(f, (size_t)count * sig_ld);
    std::vector<double> x, Fold;
    if (mode == QS_M_SCF2) {
        x = rd<double>(f, (size_t)count * 3 * nA2);
        Fold = rd<double>(f, (size_t)count * 3 * nA2);
    }
    fclose(f);
    std::vector<int> act(n), pos(n), suppos(n), tri(nA2);
    for (int i = 0; i < n; i++) act[i] = pos[i] = suppos[i] = i;
    int e = 0;
    for (int i = 0; i < nA; i++)
        for (int j = i + 1; j < nA; j++) tri[e++] = (i << 16) | j;
    for (int i = 0; i < nA; i++)
        for (int j = i; j < nA; j++) tri[e++] = (i << 16) | j;
    const double qnan = nan("");
    std::vector<double> xout((size_t)count * 3 * nA2, qnan), partials((size_t)count * 4, qnan);
    std::vector<double> diagR((size_t)count * nA, qnan), diagL((size_t)count * nA, qnan);
    std::vector<double2> onsL((size_t)count * 4 * ns, make_double2(qnan, qnan));
    std::vector<double> dosd((size_t)count * n, qnan), trUD((size_t)count * 2, qnan);
    std::vector<int> status(1, 0);

    DysonArgs a{};
    a.norb = n; a.nA = nA; a.nS = n; a.nCL = nCL; a.nsites = ns; a.nAs = nAs; a.n_atoms = 0; a.atom_tot = 0;
    a.nw = count; a.iw0 = 0;
    a.w_init = hd[0]; a.delta = hd[1]; a.mu = hd[2]; a.beta = hd[3]; a.V = hd[4];
    a.mode = mode; a.use_sigma = use_sigma; a.order = 2;
    a.Hemb = Hemb.data(); a.sig1 = sig1.data(); a.sig = use_sigma ? sig.data() : nullptr; a.sig_ld = sig_ld;
    a.gamL_s = gl.data(); a.gamR_s = gr.data(); a.gamL_cl = glcl.data();
    a.act = act.data(); a.cl = cl.data(); a.pos = pos.data(); a.suppos = suppos.data(); a.tri_map = tri.data();
    a.atom_first = a.atom_norb = a.atom_off = nullptr;
    a.x = x.empty() ? nullptr : x.data();
    a.xout = mode == QS_M_INIT2 ? xout.data() : nullptr;
    a.F = mode == QS_M_SCF2 ? xout.data() : nullptr;
    a.Fold = Fold.empty() ? nullptr : Fold.data();
    a.partials = partials.data();
    a.diagR = diagR.data(); a.diagL = diagL.data(); a.onsL = onsL.data(); a.dosd = dosd.data();
    a.atomL = nullptr; a.trUD = trUD.data(); a.status = status.data();

    const std::string shape = std::string(argv[3]) + "," + argv[4] + "," + argv[5];
    hipError_t rc = hipErrorInvalidValue;
#ifdef QS_EMU_WAVE
    // the small-LDS kernel picks its tile from n; its slab holds M per energy
    std::vector<double2> mscr(qs_dyson_wave_scratch_bytes(n) / sizeof(double2) * count);
    a.mscr = mscr.data();
    a.mscr_energies = count;
    if (shape == "wave,0,0") rc = qs_launch_dyson_wave(a, count, false, nullptr);
    else if (shape == "split,0,0") rc = qs_launch_dyson_wave(a, count, true, nullptr);
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
    FILE *o = fopen(argv[2], "wb");
    if (!o) return 2;
    wr(o, xout); wr(o, partials); wr(o, diagR); wr(o, diagL); wr(o, onsL); wr(o, dosd); wr(o, trUD); wr(o, status);
    fclose(o);
    return 0;
}
