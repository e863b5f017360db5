// CPU build of SANCHO's self-energy kernels (quichotte-and-sancho_amd/csrc/qs_sancho.hip)
// for the host sanitizers: test infrastructure, see hip/hip_runtime.h.  The source is
// included as it is; device memory is host memory, so on_device = 0 and 1 mean the same.
//
// usage: sancho_emu <in.bin> <out.bin>
// in.bin: int32 na, n_omega, order; float64 delta; float64 U (na); complex128 GR, GL
// (na, na, n_omega) column-major.  out.bin: complex128 Sig2R, Sig2L, Sig2G and, at
// order 3, Sig3R, Sig3L, Sig3G, from qs_sancho_self_energies on a handle.  The entry
// point that allocates nothing (qs_sancho_self_energies_static) runs twice as well,
// before and after the handle's call: its outputs must equal the handle's to the bit,
// and a size whose work arrays exceed its buffer must be refused.
#include QS_SANCHO_SOURCE

#include <vector>

template <class T>
static std::vector<T> rd(FILE *f, size_t n) {
    std::vector<T> v(n);
    if (n && fread(v.data(), sizeof(T), n, f) != n) {
        fprintf(stderr, "short read\n");
        exit(2);
    }
    return v;
}

int main(int argc, char **argv) {
    if (argc != 3) return 2;
    FILE *f = fopen(argv[1], "rb");
    if (!f) return 2;
    const auto hi = rd<int32_t>(f, 3);
    const int na = hi[0], nw = hi[1], order = hi[2];
    const double delta = rd<double>(f, 1)[0];
    const auto U = rd<double>(f, na);
    const size_t nm = (size_t)na * na * nw;
    const auto GR = rd<double>(f, 2 * nm), GL = rd<double>(f, 2 * nm);
    fclose(f);
    const int nout = order == 3 ? 6 : 3;
    std::vector<std::vector<double>> out(6), st(6);
    for (int i = 0; i < 6; i++) {
        out[i].assign(i < nout ? 2 * nm : 0, std::nan(""));
        st[i].assign(i < nout ? 2 * nm : 0, std::nan(""));
    }
    auto p = [&](std::vector<std::vector<double>> &v, int i) { return i < nout ? v[i].data() : nullptr; };
    // the entry point without allocation first: its buffer starts as the process left it
    if (qs_sancho_static_need(na, nw, order) > qs_sancho_static_bytes()) {
        fprintf(stderr, "sizes too large for the static buffer\n");
        return 2;
    }
    int rc = qs_sancho_self_energies_static(na, nw, order, delta, U.data(), GR.data(), GL.data(), p(st, 0), p(st, 1),
                                            p(st, 2), p(st, 3), p(st, 4), p(st, 5), nullptr);
    if (rc) {
        fprintf(stderr, "static call failed: %d\n", rc);
        return 3;
    }
    qs_sancho *s = nullptr;
    if ((rc = qs_sancho_create(na, nw, order, 0, &s))) {
        fprintf(stderr, "create failed: %d\n", rc);
        return 3;
    }
    rc = qs_sancho_self_energies(s, delta, U.data(), GR.data(), GL.data(), p(out, 0), p(out, 1), p(out, 2), p(out, 3),
                                 p(out, 4), p(out, 5), 0);
    if (rc) {
        fprintf(stderr, "handle call failed: %d (%s)\n", rc, qs_sancho_last_error(s));
        return 3;
    }
    qs_sancho_destroy(s);
    for (int i = 0; i < nout; i++)
        if (memcmp(out[i].data(), st[i].data(), out[i].size() * sizeof(double))) {
            fprintf(stderr, "MISMATCH: output %d of the static entry point\n", i);
            return 4;
        }
    // a second static call on a smaller problem (the first energies of this one),
    // then one that does not fit: refused, with nothing written
    const int nw2 = nw / 2;
    std::vector<std::vector<double>> st2(6);
    for (int i = 0; i < 6; i++) st2[i].assign(i < nout ? 2 * (size_t)na * na * nw2 : 0, 0.0);
    std::vector<double> GR2(GR.begin(), GR.begin() + 2 * (size_t)na * na * nw2);
    std::vector<double> GL2(GL.begin(), GL.begin() + 2 * (size_t)na * na * nw2);
    rc = qs_sancho_self_energies_static(na, nw2, order, delta, U.data(), GR2.data(), GL2.data(), p(st2, 0),
                                        p(st2, 1), p(st2, 2), p(st2, 3), p(st2, 4), p(st2, 5), nullptr);
    if (rc) {
        fprintf(stderr, "second static call failed: %d\n", rc);
        return 3;
    }
    int big = 1;
    while (qs_sancho_static_need(na, big, order) >= 0 && qs_sancho_static_need(na, big, order) <= qs_sancho_static_bytes())
        big *= 2;
    if (qs_sancho_self_energies_static(na, big, order, delta, U.data(), nullptr, nullptr, nullptr, nullptr, nullptr,
                                       nullptr, nullptr, nullptr, nullptr) != QS_ERR_UNSUPPORTED) {
        fprintf(stderr, "MISMATCH: an oversized static call was not refused\n");
        return 4;
    }
    FILE *o = fopen(argv[2], "wb");
    if (!o) return 2;
    for (int i = 0; i < nout; i++) fwrite(out[i].data(), sizeof(double), out[i].size(), o);
    for (int i = 0; i < nout; i++) fwrite(st2[i].data(), sizeof(double), st2[i].size(), o);
    fclose(o);
    return 0;
}
