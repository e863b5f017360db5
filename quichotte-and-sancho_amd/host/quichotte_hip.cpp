// quichotte_hip.cpp -- drop-in driver for QUICHOTTE runs on one MI355X.
//
// Reads input.dat (or the legacy inputSOC.dat) from the current directory with
// the reference's keys, aliases and defaults (src/input_output.f90:31-139,
// common/input_parser.f90), builds the Hamiltonian and lead couplings, runs the
// voltage loop through the C ABI of libqs_hip.so (include/qs_hip.h) and writes
// the reference's output files:
//   Volt_Current_<order>.dat  populations.dat  populations_atoms.dat
//   spectral_function.dat     scf_V_<V>.dat    info_Hamiltonian.dat
//   runtime_datasheet.dat     hamiltonian_matrix.dat (write_hamiltonian = T)
// Exit codes as the reference (src/quichotte.f90:161-165): 0 all voltages
// converged, 1 bad input, 3 some voltages did not converge.
//
// Models: helix, peptide, xyz, custom (src/hamiltonian.f90:83-326,741-755).
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "qs_hip.h"

using zd = std::complex<double>;
static const double PI = 3.14159265358979323846;
static const double KB = 8.6173303e-5;  // common/constants.f90:19

[[noreturn]] static void die(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    fprintf(stderr, " ERROR: ");
    vfprintf(stderr, fmt, ap);
    fprintf(stderr, "\n");
    va_end(ap);
    exit(1);
}

// ---------------------------------------------------------------------------
// key = value input (common/input_parser.f90 semantics)
// ---------------------------------------------------------------------------
struct Input {
    struct Entry { std::string key, val; int line; bool used; };
    std::vector<Entry> e;
    std::string file;
    int n_missing = 0;
    bool dup = false;

    static std::string trim(const std::string &s) {
        size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
        return a == std::string::npos ? "" : s.substr(a, b - a + 1);
    }
    void load(const std::vector<std::string> &cands) {
        for (auto &c : cands) {
            std::ifstream f(c);
            if (f.good()) { file = c; break; }
        }
        if (file.empty()) {
            fprintf(stderr, " ERROR: no input file found; looked for:\n");
            for (auto &c : cands) fprintf(stderr, "    %s\n", c.c_str());
            exit(1);
        }
        std::ifstream f(file);
        std::string line;
        int n = 0;
        while (std::getline(f, line)) {
            n++;
            line = trim(line);
            if (line.empty() || line[0] == '#') continue;
            size_t h = line.find('#');
            if (h != std::string::npos) line = line.substr(0, h);
            size_t eq = line.find('=');
            if (eq == std::string::npos) {
                printf(" WARNING: line %d of %s has no \"=\" (ignored): %s\n", n, file.c_str(), line.c_str());
                continue;
            }
            std::string k = trim(line.substr(0, eq)), v = trim(line.substr(eq + 1));
            for (auto &x : e)
                if (x.key == k) {
                    fprintf(stderr, " ERROR: key \"%s\" given twice, lines %d and %d\n", k.c_str(), x.line, n);
                    dup = true;
                }
            e.push_back({k, v, n, false});
        }
    }
    Entry *find(const std::string &key, const std::string &alias) {
        for (auto &x : e)
            if (x.key == key) { x.used = true; return &x; }
        std::stringstream ss(alias);
        std::string a;
        while (std::getline(ss, a, ',')) {
            a = trim(a);
            if (a.empty()) continue;
            for (auto &x : e)
                if (x.key == a) { x.used = true; return &x; }
        }
        return nullptr;
    }
    [[noreturn]] void bad(const Entry *x) {
        fprintf(stderr, " ERROR: cannot read the value of \"%s\" on line %d of %s\n", x->key.c_str(), x->line,
                file.c_str());
        fprintf(stderr, "        value was: \"%s\"\n", x->val.c_str());
        exit(1);
    }
    static bool parse_real(std::string s, double &out) {
        s = trim(s);
        if (s.empty()) return false;
        size_t sp = s.find_first_of(" \t,");
        if (sp != std::string::npos) s = s.substr(0, sp);
        for (auto &c : s)
            if (c == 'd' || c == 'D') c = 'e';
        char *end = nullptr;
        out = strtod(s.c_str(), &end);
        return end && *end == '\0';
    }
    void get(const std::string &key, double &var, const double *def = nullptr, const std::string &alias = "") {
        Entry *x = find(key, alias);
        if (!x) {
            if (def) var = *def;
            else { fprintf(stderr, " ERROR: required key \"%s\" missing from %s\n", key.c_str(), file.c_str()); n_missing++; }
            return;
        }
        if (!parse_real(x->val, var)) bad(x);
    }
    void get(const std::string &key, int &var, const int *def = nullptr, const std::string &alias = "") {
        Entry *x = find(key, alias);
        if (!x) {
            if (def) var = *def;
            else { fprintf(stderr, " ERROR: required key \"%s\" missing from %s\n", key.c_str(), file.c_str()); n_missing++; }
            return;
        }
        double r;
        if (!parse_real(x->val, r) || std::fabs(r - std::nearbyint(r)) > 1e-9) bad(x);
        var = (int)std::nearbyint(r);
    }
    void get(const std::string &key, bool &var, const bool *def = nullptr, const std::string &alias = "") {
        Entry *x = find(key, alias);
        if (!x) {
            if (def) var = *def;
            else { fprintf(stderr, " ERROR: required key \"%s\" missing from %s\n", key.c_str(), file.c_str()); n_missing++; }
            return;
        }
        std::string v = trim(x->val);
        std::string l;
        for (char c : v) l += (char)tolower(c);
        if (!l.empty() && l[0] == '.') l = l.substr(1);
        if (!l.empty() && (l[0] == 't')) var = true;
        else if (!l.empty() && (l[0] == 'f')) var = false;
        else bad(x);
    }
    void get(const std::string &key, std::string &var, const char *def = nullptr, const std::string &alias = "") {
        Entry *x = find(key, alias);
        if (!x) {
            if (def) var = def;
            else { fprintf(stderr, " ERROR: required key \"%s\" missing from %s\n", key.c_str(), file.c_str()); n_missing++; }
            return;
        }
        var = x->val;
    }
    void finish() {
        for (auto &x : e)
            if (!x.used) printf(" WARNING: unknown input key \"%s\" on line %d (ignored -- typo?)\n", x.key.c_str(), x.line);
        if (n_missing > 0 || dup) {
            fprintf(stderr, " Stopping: fix the input file (see the examples/ folder).\n");
            exit(1);
        }
    }
};

// ---------------------------------------------------------------------------
// run parameters (src/parameters.f90, src/input_output.f90:37-116)
// ---------------------------------------------------------------------------
struct Params {
    double V = 0, Vf = 0, delv = 0, T = 0, mu = 0;
    std::string model = "helix", xyz_file, ham_file, overlap_file, max_l, soc_xi, soc_form, hubbard_file, lead_left,
                lead_right, soc_model;
    bool write_ham = false;
    int N_sites = 0, N_ions = 0, N_turns = 0, hop_range = 1, order = 0, nhist = 16, maxscf = 300;
    double hel_length = 0, hel_radius = 0, hand = 1, pep_twist = 100, pep_rise = 1.5, hop_decay = 1, dipole_ramp = 0;
    double E_CC = 0, t_hop = 0, lamb = 0, U_int = 0, GammaL_up = 0, GammaL_down = 0, GammaR = 0;
    double dw = 0, up = 0, delta = 0, alpha_mix = 0.5, tolf = 1e-6, tolf_order1 = 1e-8;
    bool single_history = true, continuation = true;
    int Norb = 0, n_volt_steps = 0;
    double beta = 0;
    int N_atoms = 0;
    std::vector<int> atom_norb, atom_first;
    std::vector<std::string> atom_elem, elem_name;
    std::vector<int> elem_lmax;
    std::vector<double> elem_xi;
};

// 'BLOCK n | Atom_I: a Atom_J: b | Dist: d A | Dim: 8x8' -> a, b, 8, 8
static bool parse_block_header(const std::string &line, int &ai, int &aj, int &di, int &dj) {
    size_t k1 = line.find("Atom_I:"), k2 = line.find("Atom_J:"), k3 = line.find("Dim:");
    if (k1 == std::string::npos || k2 == std::string::npos || k3 == std::string::npos) return false;
    if (sscanf(line.c_str() + k1 + 7, "%d", &ai) != 1) return false;
    if (sscanf(line.c_str() + k2 + 7, "%d", &aj) != 1) return false;
    size_t kx = line.find('x', k3);
    if (kx == std::string::npos) return false;
    if (sscanf(line.substr(k3 + 4, kx - k3 - 4).c_str(), "%d", &di) != 1) return false;
    if (sscanf(line.c_str() + kx + 1, "%d", &dj) != 1) return false;
    return ai >= 1 && aj >= 1 && di >= 1 && dj >= 1;
}

// atoms and orbitals per atom from the BLOCK headers (src/hamiltonian.f90:349-396)
static void dftb_blocks_scan(const std::string &path, Params &P) {
    std::ifstream f(path);
    if (!f.good()) die("cannot open ham_file %s", path.c_str());
    std::string line;
    std::vector<std::array<int, 4>> hdr;
    int nat = 0;
    while (std::getline(f, line)) {
        std::string t = line.substr(std::min(line.find_first_not_of(" \t"), line.size()));
        if (t.rfind("BLOCK", 0) != 0) continue;
        int ai, aj, di, dj;
        if (!parse_block_header(line, ai, aj, di, dj)) die("cannot read block header in %s: %s", path.c_str(), line.c_str());
        if (di % 2 || dj % 2) die("block dimensions must be even (orbitals x 2 spins): %s", line.c_str());
        nat = std::max({nat, ai, aj});
        hdr.push_back({ai, aj, di, dj});
    }
    std::vector<int> norb(nat, 0);
    auto setn = [&](int a, int n) {
        if (norb[a - 1] != 0 && norb[a - 1] != n) die("atom %d has inconsistent block dimensions in %s", a, path.c_str());
        norb[a - 1] = n;
    };
    for (auto &h : hdr) { setn(h[0], h[2] / 2); setn(h[1], h[3] / 2); }
    if (nat < 1 || std::count(norb.begin(), norb.end(), 0) > 0)
        die("%s has no BLOCK for some atom (every atom needs its on-site block)", path.c_str());
    P.N_atoms = nat;
    P.atom_norb = norb;
}

static std::string element_case(std::string e) {
    e = Input::trim(e);
    if (e.size() > 2) e = e.substr(0, 2);
    if (!e.empty() && e[0] >= 'a' && e[0] <= 'z') e[0] = (char)(e[0] - 32);
    if (e.size() > 1 && e[1] >= 'A' && e[1] <= 'Z') e[1] = (char)(e[1] + 32);
    return e;
}

static std::vector<std::pair<std::string, std::string>> split_pairs(const std::string &str, const char *key) {
    std::string s = str;
    for (auto &c : s)
        if (c == ',') c = ' ';
    std::stringstream ss(s);
    std::string tok;
    std::vector<std::pair<std::string, std::string>> out;
    while (ss >> tok) {
        size_t c = tok.find(':');
        if (c == std::string::npos || c < 1)
            die("%s: entries must look like Element:value, got \"%s\"", key, tok.c_str());
        out.push_back({tok.substr(0, c), tok.substr(c + 1)});
    }
    if (out.empty()) die("%s is empty", key);
    return out;
}

// elements of xyz_file, shells from max_angular_momentum, xi from soc_xi (src/input_output.f90:185-238)
static void setup_dftb_elements(Params &P) {
    std::ifstream f(P.xyz_file);
    if (!f.good()) die("cannot open xyz_file %s", P.xyz_file.c_str());
    int nat = 0;
    if (!(f >> nat) || nat < 1) die("first line of %s must be the number of atoms", P.xyz_file.c_str());
    std::string line;
    std::getline(f, line);
    std::getline(f, line);
    P.atom_elem.clear();
    for (int a = 0; a < nat; a++) {
        if (!std::getline(f, line)) die("%s has fewer than %d atom lines", P.xyz_file.c_str(), nat);
        std::stringstream ls(line);
        std::string el;
        ls >> el;
        P.atom_elem.push_back(element_case(el));
    }
    for (auto &pr : split_pairs(P.max_l, "max_angular_momentum")) {
        P.elem_name.push_back(element_case(pr.first));
        std::string v = Input::trim(pr.second);
        if (v == "s" || v == "S") P.elem_lmax.push_back(0);
        else if (v == "p" || v == "P") P.elem_lmax.push_back(1);
        else die("max_angular_momentum: only s and p shells are supported, not \"%s\" for %s", v.c_str(), pr.first.c_str());
        P.elem_xi.push_back(0.0);
    }
    if (!Input::trim(P.soc_xi).empty())
        for (auto &pr : split_pairs(P.soc_xi, "soc_xi")) {
            auto it = std::find(P.elem_name.begin(), P.elem_name.end(), element_case(pr.first));
            double xv;
            if (it == P.elem_name.end() || !Input::parse_real(pr.second, xv))
                die("soc_xi: bad entry %s:%s (element must also appear in max_angular_momentum)", pr.first.c_str(),
                    pr.second.c_str());
            P.elem_xi[it - P.elem_name.begin()] = xv;
        }
    P.N_atoms = nat;
    P.atom_norb.assign(nat, 1);
    for (int a = 0; a < nat; a++) {
        auto it = std::find(P.elem_name.begin(), P.elem_name.end(), P.atom_elem[a]);
        if (it == P.elem_name.end())
            die("element %s of %s is missing from max_angular_momentum", P.atom_elem[a].c_str(), P.xyz_file.c_str());
        P.atom_norb[a] = P.elem_lmax[it - P.elem_name.begin()] == 1 ? 4 : 1;
    }
}

static int xyz_site_count(const std::string &fn) {
    std::ifstream f(fn);
    if (!f.good()) die("cannot open xyz_file %s", fn.c_str());
    int n = 0;
    if (!(f >> n) || n < 1) die("first line of %s must be the number of atoms", fn.c_str());
    return n;
}

static Params read_input() {
    Input in;
    Params P;
    in.load({"input.dat", "inputSOC.dat"});
    in.get("V", P.V);
    in.get("Vf", P.Vf);
    in.get("delv", P.delv);
    in.get("T", P.T);
    in.get("mu", P.mu);
    in.get("hamiltonian", P.model, "helix");
    in.get("xyz_file", P.xyz_file, "molecule.xyz");
    if (P.model == "dftb_hs") {
        in.get("ham_file", P.ham_file, "hamsqr1.dat");
        in.get("overlap_file", P.overlap_file, "oversqr.dat");
        in.get("max_angular_momentum", P.max_l);
    } else {
        in.get("ham_file", P.ham_file, "hamiltonian_blocks_fortran.txt");
        in.get("overlap_file", P.overlap_file, "oversqr.dat");
        in.get("max_angular_momentum", P.max_l, "");
    }
    in.get("soc_xi", P.soc_xi, "");
    in.get("soc_form", P.soc_form, "dual");
    const bool f = false;
    in.get("write_hamiltonian", P.write_ham, &f);
    const int zero = 0, one = 1;
    in.get("N_sites", P.N_sites, &zero);
    const double d0 = 0, d1 = 1;
    if (P.model == "helix") {
        in.get("N_ions", P.N_ions);
        in.get("N_turns", P.N_turns);
        in.get("hel_length", P.hel_length);
        in.get("hel_radius", P.hel_radius);
        in.get("hand", P.hand);
    } else {
        in.get("N_ions", P.N_ions, &zero);
        in.get("N_turns", P.N_turns, &zero);
        in.get("hel_length", P.hel_length, &d0);
        in.get("hel_radius", P.hel_radius, &d0);
        in.get("hand", P.hand, &d1);
    }
    const double tw = 100, rise = 1.5;
    in.get("pep_twist", P.pep_twist, &tw);
    in.get("pep_rise", P.pep_rise, &rise);
    in.get("hop_range", P.hop_range, &one);
    in.get("hop_decay", P.hop_decay, &d1);
    in.get("soc_model", P.soc_model, "radial");
    in.get("dipole_ramp", P.dipole_ramp, &d0);
    if (P.model == "dftb" || P.model == "dftb_hs") {
        in.get("E_CC", P.E_CC, &d0);
        in.get("t_hop", P.t_hop, &d0);
    } else {
        in.get("E_CC", P.E_CC);
        in.get("t_hop", P.t_hop);
    }
    if (P.model == "helix" || P.model == "xyz" || P.model == "peptide") in.get("lamb", P.lamb);
    else in.get("lamb", P.lamb, &d0);
    in.get("U_int", P.U_int, nullptr, "Hubbard");
    in.get("hubbard_file", P.hubbard_file, "");
    in.get("order", P.order);
    in.get("GammaL_up", P.GammaL_up);
    in.get("GammaL_down", P.GammaL_down);
    in.get("GammaR", P.GammaR, nullptr, "GammaR_both");
    in.get("lead_left_atoms", P.lead_left, "");
    in.get("lead_right_atoms", P.lead_right, "");
    in.get("dw", P.dw);
    in.get("up", P.up);
    in.get("delta", P.delta);
    const double a5 = 0.5, t6 = 1e-6, t8 = 1e-8;
    const int h16 = 16, m300 = 300;
    const bool tr = true;
    in.get("alpha_mix", P.alpha_mix, &a5, "pulay");
    in.get("nhist", P.nhist, &h16);
    in.get("maxscf", P.maxscf, &m300);
    in.get("tolf", P.tolf, &t6);
    in.get("tolf_order1", P.tolf_order1, &t8);
    in.get("broyden_single_precision", P.single_history, &tr);
    in.get("continuation", P.continuation, &tr);
    bool restart = false;
    in.get("restart", restart, &f);
    in.finish();

    // set_number_of_sites (src/input_output.f90:141-183)
    int n_model = -1;
    if (P.model == "helix") n_model = P.N_ions * P.N_turns;
    else if (P.model == "xyz") n_model = xyz_site_count(P.xyz_file);
    else if (P.model == "dftb") {
        dftb_blocks_scan(P.ham_file, P);
        n_model = 0;
        for (int a : P.atom_norb) n_model += a;
    } else if (P.model == "dftb_hs") {
        setup_dftb_elements(P);
        n_model = 0;
        for (int a : P.atom_norb) n_model += a;
    } else if (P.N_sites < 1) {
        fprintf(stderr, " ERROR: hamiltonian = %s needs the key N_sites\n", P.model.c_str());
        fprintf(stderr, "        (built-in models: helix, xyz, custom)\n");
        exit(1);
    }
    if (n_model >= 0) {
        if (P.N_sites > 0 && P.N_sites != n_model) {
            fprintf(stderr, " ERROR: N_sites = %d but the %s model has %d\n", P.N_sites, P.model.c_str(), n_model);
            exit(1);
        }
        P.N_sites = n_model;
    }
    P.beta = 1.0 / (KB * P.T);
    P.Norb = 2 * P.N_sites;
    if (P.atom_norb.empty()) {  // one orbital per atom (set_atom_map, src/input_output.f90:288-308)
        P.N_atoms = P.N_sites;
        P.atom_norb.assign(P.N_sites, 1);
    }
    P.atom_first.assign(P.N_atoms, 0);
    for (int a = 1; a < P.N_atoms; a++) P.atom_first[a] = P.atom_first[a - 1] + P.atom_norb[a - 1];
    P.n_volt_steps = (int)std::nearbyint((P.Vf - P.V) / P.delv);

    // check_input (src/input_output.f90:396-430)
    bool badin = false;
    auto complain = [&](const char *m) { fprintf(stderr, " ERROR in input: %s\n", m); badin = true; };
    if (P.model == "helix" && (P.N_ions < 2 || P.N_turns < 1)) complain("need N_ions >= 2 and N_turns >= 1");
    if (P.N_sites < 1) complain("need at least one site");
    if (P.model == "peptide") {
        if (P.N_sites < 2) complain("peptide: need N_sites >= 2");
        if (P.hel_radius <= 0) complain("peptide: hel_radius must be > 0 (alpha-helix: 2.5)");
        if (std::fabs(std::fabs(P.hand) - 1.0) > 1e-12) complain("peptide: hand must be +1 or -1");
        if (P.hop_range < 1) complain("peptide: hop_range must be >= 1");
        if (P.hop_decay <= 0) complain("peptide: hop_decay must be > 0");
        if (P.soc_model != "radial" && P.soc_model != "torsion") complain("peptide: soc_model must be radial or torsion");
    }
    if (P.order < 0 || P.order > 2) complain("order must be 0, 1 or 2");
    if (P.delta <= 0) complain("delta (energy grid step) must be > 0");
    if (P.delv <= 0) complain("delv (voltage step) must be > 0");
    if (P.Vf < P.V) complain("Vf must be >= V");
    if (P.T <= 0) complain("T must be > 0");
    if (P.nhist < 0) complain("nhist must be >= 0");
    if (P.maxscf < 1) complain("maxscf must be >= 1");
    if (P.tolf <= 0) complain("tolf must be > 0");
    if (P.tolf_order1 <= 0) complain("tolf_order1 must be > 0");
    if (badin) exit(1);
    printf(" Input read from %s\n", in.file.c_str());
    printf(" Hamiltonian: %s   N_sites (orbitals): %d   N_atoms: %d\n", P.model.c_str(), P.N_sites, P.N_atoms);
    return P;
}

// ---------------------------------------------------------------------------
// Hamiltonian builders (column-major H(i,j) = H[j*n + i])
// ---------------------------------------------------------------------------
struct Mat {
    int n;
    std::vector<zd> a;
    explicit Mat(int n_) : n(n_), a((size_t)n_ * n_, 0.0) {}
    zd &operator()(int i, int j) { return a[(size_t)j * n + i]; }  // 1-based like the reference
    zd operator()(int i, int j) const { return a[(size_t)j * n + i]; }
};
static zd &H1(Mat &H, int i, int j) { return H.a[(size_t)(j - 1) * H.n + (i - 1)]; }

static void make_hermitian(Mat &H) {
    for (int i = 1; i <= H.n; i++)
        for (int j = i; j <= H.n; j++) H1(H, j, i) = std::conj(H1(H, i, j));
}

static void cross(const double a[3], const double b[3], double c[3]) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - b[0] * a[1];
}
static double norm3(const double r[3]) { return std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]); }

// H_ij = i lamb (d1 x d2) . sigma on the block of sites i, i+2 (ii, jj: spin-down indices, 1-based)
static void soc_block(Mat &H, double lamb, const double *r, const double *r1, const double *r2, int ii, int jj, bool add) {
    double d1[3], d2[3], v[3], t1[3], t2[3];
    for (int k = 0; k < 3; k++) { t1[k] = r[k] - r1[k]; t2[k] = r[k] - r2[k]; }
    double n1 = norm3(t1), n2 = norm3(t2);
    for (int k = 0; k < 3; k++) { d1[k] = t1[k] / n1; d2[k] = t2[k] / n2; }
    cross(d1, d2, v);
    const zd I(0, 1);
    zd b[4] = {I * lamb * v[2], I * lamb * (v[0] - I * v[1]), I * lamb * (v[0] + I * v[1]), I * lamb * (-v[2])};
    if (add) {
        H1(H, ii - 1, jj - 1) += b[0]; H1(H, ii - 1, jj) += b[1]; H1(H, ii, jj - 1) += b[2]; H1(H, ii, jj) += b[3];
    } else {
        H1(H, ii - 1, jj - 1) = b[0]; H1(H, ii - 1, jj) = b[1]; H1(H, ii, jj - 1) = b[2]; H1(H, ii, jj) = b[3];
    }
}

// A non-magnetic H is time-reversal invariant: every 2x2 spin block h0*1 + h.sigma
// has h0 real and h purely imaginary.  Im h0 and Re h act like random magnetic
// fields, so the report puts them beside the spin-orbit part Im h
// (src/hamiltonian.f90:477-504).
static void report_time_reversal(Mat &H, int N_sites) {
    const zd I(0, 1);
    double odd0 = 0, oddh = 0, soc = 0;
    for (int j = 1; j <= N_sites; j++)
        for (int i = 1; i <= N_sites; i++) {
            const zd b11 = H1(H, 2 * i - 1, 2 * j - 1), b12 = H1(H, 2 * i - 1, 2 * j);
            const zd b21 = H1(H, 2 * i, 2 * j - 1), b22 = H1(H, 2 * i, 2 * j);
            const zd h0 = 0.5 * (b11 + b22), hz = 0.5 * (b11 - b22);
            const zd hx = 0.5 * (b12 + b21), hy = 0.5 * I * (b12 - b21);
            odd0 = std::max(odd0, std::abs(h0.imag()));
            oddh = std::max({oddh, std::abs(hx.real()), std::abs(hy.real()), std::abs(hz.real())});
            soc = std::max({soc, std::abs(hx.imag()), std::abs(hy.imag()), std::abs(hz.imag())});
        }
    printf(" Time reversal (all spin blocks h0*1 + h.sigma):\n");
    printf("   max |Im h0| =%10.2E eV   (should be 0)\n", odd0);
    printf("   max |Re h|  =%10.2E eV   (should be 0: Zeeman-like)\n", oddh);
    printf("   max |Im h|  =%10.2E eV   (spin-orbit coupling)\n", soc);
    if (std::max(odd0, oddh) > 1e-8)
        printf(" WARNING: H breaks time reversal by more than 1e-8 eV; spin polarisation of that size is not physical\n");
}

static void chain_from_positions(Mat &H, const Params &P, const std::vector<std::array<double, 3>> &rs) {
    const int nat = (int)rs.size();
    for (int i = 1; i <= nat; i++) {
        const int ii = 2 * i;
        H1(H, ii - 1, ii - 1) = P.E_CC;
        H1(H, ii, ii) = P.E_CC;
        if (i <= nat - 1) {
            const int jj = 2 * (i + 1);
            H1(H, ii - 1, jj - 1) = P.t_hop;
            H1(H, ii, jj) = P.t_hop;
        }
        if (i <= nat - 2) soc_block(H, P.lamb, rs[i - 1].data(), rs[i].data(), rs[i + 1].data(), ii, 2 * (i + 2), false);
    }
    make_hermitian(H);
}

static Mat build_helix(const Params &P) {
    Mat H(P.Norb);
    std::vector<std::array<double, 3>> rs(P.N_sites);
    for (int i = 1; i <= P.N_sites; i++) {
        const double a = P.hand * (i - 1) * 2 * PI / P.N_ions;
        rs[i - 1] = {P.hel_radius * std::cos(a), P.hel_radius * std::sin(a),
                     P.hel_length * (i - 1) / (P.N_turns * (P.N_ions - 1))};
    }
    chain_from_positions(H, P, rs);
    return H;
}

static Mat build_xyz(const Params &P) {
    std::ifstream f(P.xyz_file);
    if (!f.good()) die("cannot open XYZ file: %s", P.xyz_file.c_str());
    int nat = 0;
    f >> nat;
    if (2 * nat != P.Norb) die("%s has %d atoms but the input gives %d sites", P.xyz_file.c_str(), nat, P.Norb / 2);
    std::string line;
    std::getline(f, line);
    std::getline(f, line);
    std::vector<std::array<double, 3>> rs(nat);
    for (int i = 0; i < nat; i++) {
        std::string el;
        f >> el >> rs[i][0] >> rs[i][1] >> rs[i][2];
    }
    Mat H(P.Norb);
    chain_from_positions(H, P, rs);
    return H;
}

static Mat build_custom(const Params &P) {
    Mat H(P.Norb);
    for (int i = 1; i <= P.N_sites; i++) {
        const int ii = 2 * i;
        H1(H, ii - 1, ii - 1) = P.E_CC;
        H1(H, ii, ii) = P.E_CC;
        if (i < P.N_sites) {
            H1(H, ii - 1, ii + 1) = P.t_hop;
            H1(H, ii, ii + 2) = P.t_hop;
        }
    }
    make_hermitian(H);
    return H;
}

// peptide alpha-helix (src/hamiltonian.f90:195-273)
static Mat build_peptide(const Params &P) {
    const int nat = P.N_sites, jmax = std::min(P.hop_range, nat - 1);
    const double dphi = P.hand * P.pep_twist * PI / 180.0;
    std::vector<std::array<double, 3>> rs(nat);
    std::vector<double> phi(nat);
    for (int i = 1; i <= nat; i++) {
        phi[i - 1] = i * dphi;
        rs[i - 1] = {P.hel_radius * std::cos(phi[i - 1]), P.hel_radius * std::sin(phi[i - 1]), (i - 1) * P.pep_rise};
    }
    double t0[3] = {rs[1][0] - rs[0][0], rs[1][1] - rs[0][1], rs[1][2] - rs[0][2]};
    const double l1 = norm3(t0);
    Mat H(P.Norb);
    for (int i = 1; i <= nat; i++) {
        const int ii = 2 * i;
        H1(H, ii - 1, ii - 1) = P.E_CC + P.dipole_ramp * ((double)(i - 1) / (double)(nat - 1) - 0.5);
        H1(H, ii, ii) = H1(H, ii - 1, ii - 1);
    }
    if (P.soc_model == "torsion")
        for (int i = 1; i <= nat - 2; i++) soc_block(H, P.lamb, rs[i - 1].data(), rs[i].data(), rs[i + 1].data(), 2 * i, 2 * (i + 2), false);
    const zd I(0, 1);
    for (int j = 1; j <= jmax; j++)
        for (int i = 1; i <= nat - j; i++) {
            const int m = i + j, ii = 2 * i, jj = 2 * m;
            double d[3] = {rs[m - 1][0] - rs[i - 1][0], rs[m - 1][1] - rs[i - 1][1], rs[m - 1][2] - rs[i - 1][2]};
            const double l = norm3(d);
            double dhat[3] = {d[0] / l, d[1] / l, d[2] / l};
            const double fac = std::exp(-(l - l1) / P.hop_decay);
            H1(H, ii - 1, jj - 1) += P.t_hop * fac;
            H1(H, ii, jj) += P.t_hop * fac;
            if (P.soc_model == "radial") {
                const double phibar = 0.5 * (phi[i - 1] + phi[m - 1]);
                double rhat[3] = {std::cos(phibar), std::sin(phibar), 0.0}, nu[3];
                cross(rhat, dhat, nu);
                const double lj = 2.0 * P.lamb * fac * std::cos(0.5 * (phi[m - 1] - phi[i - 1]));
                H1(H, ii - 1, jj - 1) += I * lj * nu[2];
                H1(H, ii - 1, jj) += I * lj * (nu[0] - I * nu[1]);
                H1(H, ii, jj - 1) += I * lj * (nu[0] + I * nu[1]);
                H1(H, ii, jj) -= I * lj * nu[2];
            }
        }
    make_hermitian(H);
    FILE *g = fopen("peptide_geometry.xyz", "w");
    if (g) {
        fprintf(g, "%d\nQUICHOTTE peptide sites [Angstrom]\n", nat);
        for (auto &r : rs) fprintf(g, "C %14.8f%14.8f%14.8f\n", r[0], r[1], r[2]);
        fclose(g);
    }
    // geometry summary and hopping table (src/hamiltonian.f90:260-270)
    printf(" Peptide helix: twist%7.2f deg, rise%6.3f A, radius%6.3f A, hand%5.1f, SOC %s\n", P.pep_twist, P.pep_rise,
           P.hel_radius, P.hand, P.soc_model.c_str());
    printf(" Dipole ramp:%9.5f eV  (eps_1 =%9.5f, eps_N =%9.5f eV)\n", P.dipole_ramp, H1(H, 1, 1).real(),
           H1(H, P.Norb, P.Norb).real());
    printf("    j    l_j [A]    t_j/t_hop     (hop_decay =%6.3f A)\n", P.hop_decay);
    for (int j = 1; j <= jmax; j++) {
        double d[3] = {rs[j][0] - rs[0][0], rs[j][1] - rs[0][1], rs[j][2] - rs[0][2]};
        const double l = norm3(d);
        printf("%5d%11.4f%14.4E\n", j, l, std::exp(-(l - l1) / P.hop_decay));
    }
    report_time_reversal(H, nat);
    return H;
}

// dftb: orthogonal spin-orbit Hamiltonian from soc_dftb blocks (src/hamiltonian.f90:398-475)
static Mat build_dftb(const Params &P) {
    Mat H(P.Norb);
    std::ifstream f(P.ham_file);
    if (!f.good()) die("cannot open ham_file %s", P.ham_file.c_str());
    std::vector<char> onsite(P.N_atoms, 0);
    std::string line;
    int nblk = 0;
    while (std::getline(f, line)) {
        std::string t = line.substr(std::min(line.find_first_not_of(" \t"), line.size()));
        if (t.rfind("BLOCK", 0) != 0) continue;
        int ai, aj, di, dj;
        if (!parse_block_header(line, ai, aj, di, dj) || ai > P.N_atoms || aj > P.N_atoms ||
            di != 2 * P.atom_norb[ai - 1] || dj != 2 * P.atom_norb[aj - 1])
            die("bad block header in %s: %s", P.ham_file.c_str(), line.c_str());
        const int r0 = 2 * P.atom_first[ai - 1], c0 = 2 * P.atom_first[aj - 1];
        for (int r = 1; r <= di; r++)
            for (int c = 1; c <= dj; c++) {
                std::string dl;
                double re, imv;
                if (!std::getline(f, dl) || sscanf(dl.c_str(), "%lf %lf", &re, &imv) != 2)
                    die("block too short in %s: %s", P.ham_file.c_str(), line.c_str());
                H1(H, r0 + r, c0 + c) = zd(re, imv);
            }
        if (ai == aj) onsite[ai - 1] = 1;
        nblk++;
    }
    for (char o : onsite)
        if (!o) die("%s is missing the on-site block of some atom", P.ham_file.c_str());
    double herm = 0;
    Mat Hs(P.Norb);
    for (int i = 1; i <= P.Norb; i++)
        for (int j = 1; j <= P.Norb; j++) {
            herm = std::max(herm, std::abs(H1(H, i, j) - std::conj(H1(H, j, i))));
            H1(Hs, i, j) = 0.5 * (H1(H, i, j) + std::conj(H1(H, j, i)));
        }
    printf(" DFTB Hamiltonian: %d blocks, %d atoms, %d orbitals\n", nblk, P.N_atoms, P.N_sites);
    printf(" max |H - H^dagger| in file = %10.2E eV (H symmetrised)\n", herm);
    report_time_reversal(Hs, P.N_sites);
    return Hs;
}

static void hermitian_eig(const Mat &H, std::vector<double> &ev, Mat &V);

// a DFTB+ square matrix file: '#' lines, 'T NALLORB NKPOINT', k-point index, n rows of n numbers
static std::vector<double> read_dftb_square(const std::string &path, int n) {
    std::ifstream f(path);
    if (!f.good()) die("cannot open %s", path.c_str());
    auto next_data = [&](std::string &l) {
        while (std::getline(f, l)) {
            std::string t = Input::trim(l);
            if (!t.empty() && t[0] != '#') return;
        }
        die("unexpected end of %s", path.c_str());
    };
    std::string l;
    next_data(l);
    std::stringstream hs(l);
    std::string flag;
    int nall = -1, nk = -1;
    hs >> flag >> nall >> nk;
    if (nall != n) die("%s has %d orbitals; the geometry and shells give %d", path.c_str(), nall, n);
    if (nk != 1) die("%s: only one k-point (a molecule) is supported", path.c_str());
    next_data(l);  // k-point index
    std::vector<double> A((size_t)n * n);
    for (int i = 0; i < n; i++) {
        next_data(l);
        std::stringstream rs(l);
        for (int j = 0; j < n; j++) {
            std::string tok;
            if (!(rs >> tok) || !Input::parse_real(tok, A[(size_t)j * n + i])) die("%s: cannot read row %d", path.c_str(), i + 1);
        }
    }
    return A;
}

// dftb_hs: H = (S^-1/2 x 1)[H0 x 1 + H_SOC](S^-1/2 x 1) (src/hamiltonian.f90:611-708)
static Mat build_dftb_hs(const Params &P) {
    const double hartree = 27.211386245988;
    const int N = P.N_sites, n = P.Norb;
    std::vector<double> S = read_dftb_square(P.overlap_file, N), H0 = read_dftb_square(P.ham_file, N);
    Mat Sm(N);
    for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++) {
            Sm.a[(size_t)j * N + i] = 0.5 * (S[(size_t)j * N + i] + S[(size_t)i * N + j]);
        }
    std::vector<double> Ssym((size_t)N * N), H0s((size_t)N * N);
    for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++) {
            Ssym[(size_t)j * N + i] = Sm.a[(size_t)j * N + i].real();
            H0s[(size_t)j * N + i] = 0.5 * (H0[(size_t)j * N + i] + H0[(size_t)i * N + j]) * hartree;
        }
    std::vector<double> ev;
    Mat U(N);
    hermitian_eig(Sm, ev, U);
    if (ev.empty() || *std::min_element(ev.begin(), ev.end()) <= 1e-8)
        die("overlap matrix is not positive definite (min eigenvalue %g)", ev.empty() ? 0.0 : ev[0]);
    std::vector<double> Smh((size_t)N * N, 0.0);  // U diag(s^-1/2) U^T (U real up to phases)
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) {
            zd acc = 0;
            for (int k = 0; k < N; k++) acc += U.a[(size_t)k * N + i] * std::conj(U.a[(size_t)k * N + j]) / std::sqrt(ev[k]);
            Smh[(size_t)j * N + i] = acc.real();
        }
    printf(" DFTB H0/S: smallest overlap eigenvalue = %10.2E\n", ev[0]);
    // Lambda = sum_atoms xi L.S on the p shell (L_k)_ab = -i eps_kab, DFTB+ order s, p_y, p_z, p_x
    const zd I(0, 1);
    auto Lk = [&](int k, int a, int b) -> zd {  // 1-based cartesian indices
        if (k == 1) return (a == 2 && b == 3) ? -I : (a == 3 && b == 2) ? I : zd(0);
        if (k == 2) return (a == 3 && b == 1) ? -I : (a == 1 && b == 3) ? I : zd(0);
        return (a == 1 && b == 2) ? -I : (a == 2 && b == 1) ? I : zd(0);
    };
    auto sig = [&](int s1, int s2, int k) -> zd {
        if (k == 1) return (s1 != s2) ? zd(1) : zd(0);
        if (k == 2) return (s1 == 1 && s2 == 2) ? -I : (s1 == 2 && s2 == 1) ? I : zd(0);
        return (s1 == s2) ? zd(s1 == 1 ? 1.0 : -1.0) : zd(0);
    };
    const int pidx[3] = {2, 3, 1};
    Mat Lam(n);
    for (int a = 0; a < P.N_atoms; a++) {
        if (P.atom_norb[a] < 4) continue;
        auto it = std::find(P.elem_name.begin(), P.elem_name.end(), P.atom_elem[a]);
        const double xi = P.elem_xi[it - P.elem_name.begin()];
        if (xi == 0.0) continue;
        const int i0 = P.atom_first[a] + 1;  // 1-based site of the s orbital
        for (int p = 1; p <= 3; p++)
            for (int q = 1; q <= 3; q++)
                for (int s1 = 1; s1 <= 2; s1++)
                    for (int s2 = 1; s2 <= 2; s2++) {
                        const int i = 2 * (i0 + p - 1) + s1, j = 2 * (i0 + q - 1) + s2;
                        for (int k = 1; k <= 3; k++) H1(Lam, i, j) += xi * Lk(k, pidx[p - 1], pidx[q - 1]) * sig(s1, s2, k) / 2.0;
                    }
    }
    Mat Sx2(n), Sm2(n), Hao(n);
    for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++)
            for (int s1 = 1; s1 <= 2; s1++) {
                H1(Sx2, 2 * i + s1, 2 * j + s1) = Ssym[(size_t)j * N + i];
                H1(Sm2, 2 * i + s1, 2 * j + s1) = Smh[(size_t)j * N + i];
                H1(Hao, 2 * i + s1, 2 * j + s1) = H0s[(size_t)j * N + i];
            }
    auto mm = [&](const Mat &A, const Mat &B) {
        Mat C(n);
        for (int j = 0; j < n; j++)
            for (int k = 0; k < n; k++) {
                const zd b = B.a[(size_t)j * n + k];
                if (b == 0.0) continue;
                for (int i = 0; i < n; i++) C.a[(size_t)j * n + i] += A.a[(size_t)k * n + i] * b;
            }
        return C;
    };
    if (P.soc_form == "dual") {
        Mat A = mm(Sx2, Lam), B = mm(Lam, Sx2);
        for (size_t q = 0; q < Hao.a.size(); q++) Hao.a[q] += 0.5 * (A.a[q] + B.a[q]);
    } else if (P.soc_form == "onsite") {
        for (size_t q = 0; q < Hao.a.size(); q++) Hao.a[q] += Lam.a[q];
    } else {
        die("soc_form must be dual or onsite, not %s", P.soc_form.c_str());
    }
    Mat Ht = mm(Sm2, mm(Hao, Sm2)), H(n);
    for (int i = 1; i <= n; i++)
        for (int j = 1; j <= n; j++) H1(H, i, j) = 0.5 * (H1(Ht, i, j) + std::conj(H1(Ht, j, i)));
    printf(" DFTB H0/S Hamiltonian: %d atoms, %d orbitals, SOC form %s\n", P.N_atoms, N, P.soc_form.c_str());
    for (size_t e = 0; e < P.elem_name.size(); e++)
        printf("   element %-2s: l_max = %d, xi(p) = %12.4E eV\n", P.elem_name[e].c_str(), P.elem_lmax[e], P.elem_xi[e]);
    report_time_reversal(H, P.N_sites);
    return H;
}

// ---------------------------------------------------------------------------
// Hermitian eigen-decomposition by cyclic complex Jacobi (setup only: the grid
// bounds of src/quichotte.f90:81-86 and info_Hamiltonian.dat)
// ---------------------------------------------------------------------------
static void hermitian_eig(const Mat &H, std::vector<double> &ev, Mat &V) {
    const int n = H.n;
    std::vector<zd> A = H.a;  // column-major
    auto a = [&](int i, int j) -> zd & { return A[(size_t)j * n + i]; };
    V = Mat(n);
    for (int i = 0; i < n; i++) V.a[(size_t)i * n + i] = 1.0;
    auto v = [&](int i, int j) -> zd & { return V.a[(size_t)j * n + i]; };
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0, tot = 0;
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) {
                const double s = std::norm(a(i, j));
                tot += s;
                if (i != j) off += s;
            }
        if (off <= 1e-32 * tot || off == 0.0) break;
        for (int p = 0; p < n - 1; p++)
            for (int q = p + 1; q < n; q++) {
                const zd b = a(p, q);
                const double ab = std::abs(b);
                if (ab == 0.0) continue;
                const zd ph = b / ab;  // e^{i phi}
                const double app = a(p, p).real(), aqq = a(q, q).real();
                const double tau = (aqq - app) / (2.0 * ab);
                const double t = (tau >= 0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
                const double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
                // U = [[c, s ph], [-s conj(ph), c]] on columns (p, q): A <- U^dag A U
                for (int k = 0; k < n; k++) {
                    const zd akp = a(k, p), akq = a(k, q);
                    a(k, p) = c * akp - s * std::conj(ph) * akq;
                    a(k, q) = s * ph * akp + c * akq;
                }
                for (int k = 0; k < n; k++) {
                    const zd apk = a(p, k), aqk = a(q, k);
                    a(p, k) = c * apk - s * ph * aqk;
                    a(q, k) = s * std::conj(ph) * apk + c * aqk;
                }
                a(p, q) = 0.0;
                a(q, p) = 0.0;
                for (int k = 0; k < n; k++) {
                    const zd vkp = v(k, p), vkq = v(k, q);
                    v(k, p) = c * vkp - s * std::conj(ph) * vkq;
                    v(k, q) = s * ph * vkp + c * vkq;
                }
            }
    }
    std::vector<int> idx(n);
    for (int i = 0; i < n; i++) idx[i] = i;
    std::sort(idx.begin(), idx.end(), [&](int x, int y) { return a(x, x).real() < a(y, y).real(); });
    ev.resize(n);
    Mat Vs(n);
    for (int k = 0; k < n; k++) {
        ev[k] = a(idx[k], idx[k]).real();
        for (int i = 0; i < n; i++) Vs.a[(size_t)k * n + i] = v(i, idx[k]);
    }
    V = Vs;
}

// ---------------------------------------------------------------------------
// list-directed output close to gfortran's (numbers only are compared)
// ---------------------------------------------------------------------------
static std::string fr(double x) {
    char b[64];
    const double ax = std::fabs(x);
    if (x == 0.0 || (ax >= 0.1 && ax < 1e16)) {
        int digits = x == 0.0 ? 1 : (int)std::floor(std::log10(ax)) + 1;
        int dec = std::max(17 - digits, 0);
        snprintf(b, sizeof b, "%*.*f", 25, dec, x);
    } else {
        snprintf(b, sizeof b, "%25.16E", x);
        // gfortran writes three exponent digits
        std::string s(b);
        size_t e = s.find('E');
        if (e != std::string::npos && s.size() - e == 4) s.insert(e + 2, "0");
        return s;
    }
    return b;
}
static std::string fi(long v) {
    char b[32];
    snprintf(b, sizeof b, "%12ld", v);
    return b;
}

static std::vector<int> parse_atom_list(const std::string &str, int default_atom, const char *key, int n_atoms) {
    std::string s = str;
    for (auto &c : s)
        if (c == ',') c = ' ';
    std::stringstream ss(s);
    std::string tok;
    std::vector<int> out;
    while (ss >> tok) {
        size_t k = tok.find(':');
        int a = 0, b = 0;
        bool ok = true;
        try {
            if (k != std::string::npos) { a = std::stoi(tok.substr(0, k)); b = std::stoi(tok.substr(k + 1)); }
            else { a = b = std::stoi(tok); }
        } catch (...) { ok = false; }
        if (!ok || a < 1 || b > n_atoms || b < a)
            die("%s: bad atom or range \"%s\" (atoms are 1..%d)", key, tok.c_str(), n_atoms);
        for (int x = a; x <= b; x++) out.push_back(x);
    }
    if (out.empty()) out.push_back(default_atom);
    return out;
}

int main() {
    auto t_start = std::chrono::steady_clock::now();
    FILE *frt = fopen("runtime_datasheet.dat", "w");
    fprintf(frt, " >>>>>>>>>>>>>>>>>>>>>>>>>>>>Runtime Data Sheet>>>>>>>>>>>>>>>>>>>>>>>>>>>>\n");
    fprintf(frt, " QUICHOTTE version: quichotte-mi355x (libqs_hip, gfx950)\n");
    Params P = read_input();
    Mat H(P.Norb);
    if (P.model == "helix") H = build_helix(P);
    else if (P.model == "xyz") H = build_xyz(P);
    else if (P.model == "custom") H = build_custom(P);
    else if (P.model == "peptide") H = build_peptide(P);
    else if (P.model == "dftb") H = build_dftb(P);
    else if (P.model == "dftb_hs") H = build_dftb_hs(P);
    else {
        printf(" ERROR: unknown hamiltonian = %s\n", P.model.c_str());
        printf("        built in: helix, peptide, xyz, dftb, dftb_hs, custom\n");
        exit(1);
    }
    const int n = P.Norb;
    if (P.write_ham) {
        FILE *g = fopen("hamiltonian_matrix.dat", "w");
        fprintf(g, "# QUICHOTTE Hamiltonian [eV]; state index 2*site-1 = up, 2*site = down\n# Norb = %d\n", n);
        fprintf(g, "#   i   j   Re H_ij   Im H_ij\n");
        // elements that vanish by symmetry (e.g. the on-site spin flip of an s orbital after the
        // Lowdin transform) come out as rounding noise of ~1e-17 eV here; they are written as zeros
        for (int j = 1; j <= n; j++)
            for (int i = 1; i <= n; i++)
                if (std::abs(H1(H, i, j)) > 1e-12)
                    fprintf(g, "%6d%6d%25.16E%25.16E\n", i, j, H1(H, i, j).real(), H1(H, i, j).imag());
        fclose(g);
    }
    // leads on every orbital of the contact atoms (src/input_output.f90:310-346)
    const int n_atoms = P.N_atoms;
    std::vector<int> left = parse_atom_list(P.lead_left, 1, "lead_left_atoms", n_atoms);
    std::vector<int> right = parse_atom_list(P.lead_right, n_atoms, "lead_right_atoms", n_atoms);
    for (int a : left)
        for (int b : right)
            if (a == b) die("atom %d is in both lead_left_atoms and lead_right_atoms", a);
    Mat GL(n), GR(n);
    for (int a : left)
        for (int i = P.atom_first[a - 1] + 1; i <= P.atom_first[a - 1] + P.atom_norb[a - 1]; i++) {
            H1(GL, 2 * i - 1, 2 * i - 1) = P.GammaL_up;
            H1(GL, 2 * i, 2 * i) = P.GammaL_down;
        }
    for (int a : right)
        for (int i = P.atom_first[a - 1] + 1; i <= P.atom_first[a - 1] + P.atom_norb[a - 1]; i++) {
            H1(GR, 2 * i - 1, 2 * i - 1) = P.GammaR;
            H1(GR, 2 * i, 2 * i) = P.GammaR;
        }
    // Hubbard U (src/input_output.f90:432-462)
    std::vector<double> U(n, 0.0);
    if (P.hubbard_file.empty()) {
        for (int i = 0; i < n; i += 2) U[i] = P.U_int;
    } else {
        std::ifstream f(P.hubbard_file);
        if (!f.good()) die("cannot open hubbard_file %s", P.hubbard_file.c_str());
        for (int i = 0; i < n; i += 2) {
            std::string line;
            double val;
            if (!std::getline(f, line) || !Input::parse_real(line, val))
                die("%s needs one U per site/orbital (%d lines)", P.hubbard_file.c_str(), n / 2);
            U[i] = val;
        }
        printf(" Site-dependent U read from %s\n", P.hubbard_file.c_str());
    }
    // energy grid from the spectrum (src/quichotte.f90:81-93)
    std::vector<double> ev;
    Mat Evec(n);
    hermitian_eig(H, ev, Evec);
    const double w_init = ev[0] - P.dw, w_fin = ev[n - 1] + P.up;
    const int n_omega = (int)((w_fin - w_init) / P.delta);
    printf(" n_omega: %d w_fin: %.10g w_init: %.10g\n", n_omega, w_fin, w_init);
    {
        FILE *g = fopen("info_Hamiltonian.dat", "w");
        fprintf(g, "\n... Hamiltonian:\n");
        for (int i = 1; i <= n; i++) {
            for (int j = 1; j <= n; j++) fprintf(g, " [%10.5f%10.5f] ", H1(H, i, j).real(), H1(H, i, j).imag());
            fprintf(g, "\n");
        }
        fprintf(g, "\n... Checking the Hamiltonian is Hermitian:\n");
        for (int i = 1; i <= n; i++)
            for (int j = i; j <= n; j++) {
                zd d = H1(H, i, j) - std::conj(H1(H, j, i));
                fprintf(g, "%12d%12d %s %s\n", i, j, fr(d.real()).c_str(), fr(d.imag()).c_str());
            }
        fprintf(g, "\n... Eigenvalues:\n");
        for (int i = 0; i < n; i++) fprintf(g, "%12d . %s\n", i + 1, fr(ev[i]).c_str());
        fprintf(g, "\n... Eigenvectors:\n");
        for (int j = 1; j <= n; j++) {
            fprintf(g, "%3d", j);
            for (int i = 1; i <= n; i++) fprintf(g, " [%10.5f%10.5f] ", H1(Evec, i, j).real(), H1(Evec, i, j).imag());
            fprintf(g, "\n");
        }
        fprintf(g, " Norb: %12d\n", n);
        fclose(g);
    }
    qs_problem qp{};
    qp.norb = n;
    qp.n_omega = n_omega;
    qp.w_init = w_init;
    qp.delta = P.delta;
    qp.H = (const double *)H.a.data();
    qp.gammaL = (const double *)GL.a.data();
    qp.gammaR = (const double *)GR.a.data();
    qp.U_site = U.data();
    qp.order = P.order;
    qp.beta = P.beta;
    qp.mu = P.mu;
    qp.n_atoms = P.N_atoms;
    qp.atom_norb = P.atom_norb.data();
    qp.alpha_mix = P.alpha_mix;
    qp.nhist = P.nhist;
    qp.maxscf = P.maxscf;
    qp.tolf = P.tolf;
    qp.tolf_order1 = P.tolf_order1;
    qp.single_history = P.single_history ? 1 : 0;
    qp.device = 0;
    qs_handle *h = nullptr;
    int rc = qs_create(&qp, &h);
    if (rc != QS_OK) die("the device engine could not be set up (status %d)", rc);
    char vfn[32];
    snprintf(vfn, sizeof vfn, "Volt_Current_%d.dat", P.order);
    FILE *fiv = fopen(vfn, "w"), *fpop = fopen("populations.dat", "w"), *fspec = fopen("spectral_function.dat", "w");
    const bool per_atom = *std::max_element(P.atom_norb.begin(), P.atom_norb.end()) > 1;
    FILE *fpa = per_atom ? fopen("populations_atoms.dat", "w") : nullptr;
    int n_unconv = 0;
    bool fresh = true;
    std::vector<double> lf(std::max(P.maxscf, 1)), lr(std::max(P.maxscf, 1)), dosd((size_t)n * n_omega);
    std::vector<zd> G((size_t)n * n);
    for (int k = 0; k <= P.n_volt_steps; k++) {
        const double V1 = P.V + k * P.delv;
        fprintf(frt, " ........SCF Calculations at Voltage: %s ..........\n", fr(V1).c_str());
        printf(" >>>>>>>>>>VOLTAGE: %s%s\n", fr(V1).c_str(), fresh ? " (fresh start)" : " (from previous voltage)");
        qs_scf_result res{};
        rc = qs_solve_scf(h, V1, fresh ? 1 : 0, &res, lf.data(), lr.data());
        if (rc != QS_OK) {
            fprintf(stderr, " ERROR: %s\n", qs_last_error(h));
            qs_destroy(h);
            return 1;
        }
        char fn1[32];
        snprintf(fn1, sizeof fn1, "%8.3f", V1);
        std::string vs = Input::trim(fn1);
        FILE *flog = fopen(("scf_V_" + vs + ".dat").c_str(), "w");
        fprintf(flog, "#  step        max|F|          rms|F|   nhist\n");
        for (int it = 0; it < res.n_iter; it++)
            fprintf(flog, "%6d%16.6E%16.6E%6d\n", it + 1, lf[it], lr[it], P.order == 0 ? 0 : std::min(std::max(it - 1, 0), P.nhist));
        fclose(flog);
        if (!res.converged && !res.unphysical)
            printf(" NOT CONVERGED after%5d steps, max|F| =%11.3E\n", res.n_iter, res.fmax);
        else if (!res.converged)
            printf(" CONVERGED TO AN UNPHYSICAL SOLUTION: state%4d n =%11.3E, spectral weight =%11.3E\n", res.i_worst,
                   res.n_worst, res.w_worst);
        else
            printf(" ... CONVERGED in%5d steps ...\n", res.n_iter);
        if (!res.converged) n_unconv++;
        fresh = (!P.continuation) || (!res.converged);
        double Ju = 0, Jd = 0;
        qs_current(h, &Ju, &Jd);
        fprintf(fiv, " %s %s %s %s %s\n", fr(V1).c_str(), fr(Ju).c_str(), fr(Jd).c_str(), fr(Ju + Jd).c_str(),
                fi(res.converged ? 1 : 0).c_str());
        fflush(fiv);
        // spin texture (src/observables.f90:41-86)
        qs_g_less_obs(h, (double *)G.data());
        auto g = [&](int i, int j) { return G[(size_t)j * n + i]; };
        const zd I(0, 1), pref = -I * 0.5;
        for (int i = 0; i < n; i += 2) {
            zd tx = 0, ty = 0, tz = 0;
            // sum over s1, s2 of sigma(s1,s2) G(i+s2, i+s1)
            tx = g(i + 1, i) + g(i, i + 1);
            ty = -I * g(i + 1, i) + I * g(i, i + 1);
            tz = g(i, i) - g(i + 1, i + 1);
            const zd up = -I * g(i, i), dn = -I * g(i + 1, i + 1);
            fprintf(fpop, " %s %s %s %s %s %s %s %s\n", fr(V1).c_str(), fi(i / 2 + 1).c_str(), fr((tx * pref).real()).c_str(),
                    fr((ty * pref).real()).c_str(), fr((tz * pref).real()).c_str(), fr(dn.real()).c_str(),
                    fr(up.real()).c_str(), fr((up + dn).real()).c_str());
        }
        fflush(fpop);
        // per-atom S, populations and L over the p shell (src/observables.f90:104-158)
        if (fpa) {
            const int cart[3] = {2, 3, 1};
            auto Lk = [&](int k, int a, int b) -> zd {  // (L_k)_ab = -i eps_kab, 1-based
                const int kp1 = k % 3 + 1, kp2 = (k + 1) % 3 + 1;
                if (a == kp1 && b == kp2) return -I;
                if (a == kp2 && b == kp1) return I;
                return 0.0;
            };
            for (int at = 0; at < P.N_atoms; at++) {
                double Sv[3] = {0, 0, 0}, Lv[3] = {0, 0, 0}, nu = 0, nd = 0;
                for (int o = P.atom_first[at]; o < P.atom_first[at] + P.atom_norb[at]; o++) {
                    zd rho[2][2];
                    for (int s1 = 0; s1 < 2; s1++)
                        for (int s2 = 0; s2 < 2; s2++) rho[s1][s2] = -I * g(2 * o + s1, 2 * o + s2);
                    nu += rho[0][0].real();
                    nd += rho[1][1].real();
                    // 0.5 Re sum(sigma * transpose(rho))
                    Sv[0] += 0.5 * (rho[1][0] + rho[0][1]).real();
                    Sv[1] += 0.5 * (-I * rho[1][0] + I * rho[0][1]).real();
                    Sv[2] += 0.5 * (rho[0][0] - rho[1][1]).real();
                }
                if (P.atom_norb[at] == 4)
                    for (int p = 1; p <= 3; p++)
                        for (int q = 1; q <= 3; q++) {
                            const int ip = P.atom_first[at] + p, iq = P.atom_first[at] + q;  // 0-based sites
                            const zd rhop = -I * (g(2 * iq, 2 * ip) + g(2 * iq + 1, 2 * ip + 1));
                            for (int k = 1; k <= 3; k++) Lv[k - 1] += (Lk(k, cart[p - 1], cart[q - 1]) * rhop).real();
                        }
                fprintf(fpa, " %s %s %s %s %s %s %s %s %s %s %s\n", fr(V1).c_str(), fi(at + 1).c_str(), fr(Sv[0]).c_str(),
                        fr(Sv[1]).c_str(), fr(Sv[2]).c_str(), fr(nd).c_str(), fr(nu).c_str(), fr(nd + nu).c_str(),
                        fr(Lv[0]).c_str(), fr(Lv[1]).c_str(), fr(Lv[2]).c_str());
            }
            fflush(fpa);
        }
        // spectral function (src/observables.f90:89-102)
        qs_dosd(h, dosd.data());
        fprintf(fspec, "# V = %10.5f   columns: V, omega, A_ii for i = 1..Norb\n", V1);
        for (int iw = 0; iw < n_omega; iw++) {
            fprintf(fspec, " %s %s", fr(V1).c_str(), fr(w_init + (iw + 1) * P.delta).c_str());
            for (int i = 0; i < n; i++) fprintf(fspec, " %s", fr(-2.0 * dosd[(size_t)iw * n + i]).c_str());
            fprintf(fspec, "\n");
        }
        fprintf(fspec, "\n\n");
        fflush(fspec);
        printf(" Progress: %s %% %s %s\n", fr(100.0 * k / std::max(P.n_volt_steps, 1)).c_str(), fr(Ju).c_str(), fr(Jd).c_str());
    }
    fclose(fiv);
    fclose(fpop);
    fclose(fspec);
    if (fpa) fclose(fpa);
    qs_destroy(h);
    const double mins = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count() / 60.0;
    fprintf(frt, " <<<<<<<<<<<<<<<<<<<<<<<<<<<<<<<<<<<<<<<<<<<<<\n Total Runtime: %s mins\n", fr(mins).c_str());
    fprintf(frt, " <<<<<<<<<<<<<<<<<<<<<<<<<<<<<<<<<<<<<<<<<<<<<\n");
    fclose(frt);
    printf(" Total Runtime: %s mins\n", fr(mins).c_str());
    if (n_unconv > 0) {
        printf(" WARNING: %d voltage(s) did not converge (column 5 = 0 in Volt_Current)\n", n_unconv);
        return 3;
    }
    return 0;
}
