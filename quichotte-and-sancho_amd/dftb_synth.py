"""Seeded synthetic DFTB-like carbon helix: the inputs of a `hamiltonian = dftb_hs`
run (src/hamiltonian.f90:506-708) without DFTB+.

n_atoms carbon atoms sit on a helix of radius 3 A and pitch 3.5 A, consecutive
atoms 1.42 A apart; the seed draws the start angle and a small random jitter of
the positions.  Each atom carries s, p_y, p_z, p_x (DFTB+ order).  The
two-centre integrals are Slater-Koster forms (direction cosines of the bond
times the bond integrals) with every bond integral scaled by
f(r) = exp(-(r - 1.42)/0.3):

    H:  on-site eps_s = -13.6 eV, eps_p = -5.4 eV;
        V_ss_sigma = -2.7 f, V_sp_sigma = +2.7 f, V_pp_sigma = +2.7 f, V_pp_pi = -2.7 f  [eV]
    S:  1 on the diagonal (orthonormal orbitals within an atom), and
        S_X = -(0.25 / 2.7) V_X between atoms: 0.25 f in magnitude, the sign opposite
        to H's (H ~ -K S, as in extended Hueckel)

H0 goes to hamsqr1.dat in Hartree and S to oversqr.dat, in DFTB+'s square-matrix
format (WriteHS = Yes), with a geometry file and an input.dat for the drop-in
driver.  `dftb_hs_hamiltonian` restates the driver's orthogonal spin-orbit
Hamiltonian H = (S^-1/2 x 1)[H0 x 1 + H_SOC](S^-1/2 x 1) with the dual SOC form
(S Lambda + Lambda S)/2, Lambda = xi L.S on each p shell, so a Python caller
can build the same system without the driver.
"""
import os

import numpy as np

HARTREE = 27.211386245988  # eV
EPS_S, EPS_P = -13.6, -5.4
V_SS, V_SP, V_PP_SIGMA, V_PP_PI = -2.7, 2.7, 2.7, -2.7
S_OVER_V = -0.25 / 2.7
R0, DECAY = 1.42, 0.3
# DFTB+ orbital order s, p_y, p_z, p_x: Cartesian axis of each p orbital
P_AXIS = (1, 2, 0)


def helix_positions(n_atoms=30, radius=3.0, pitch=3.5, d_cc=R0, seed=0, jitter=0.01):
    """Atom positions [A]: equal angular steps chosen so that consecutive atoms are
    d_cc apart, a seeded start angle, and a seeded Gaussian jitter (sigma `jitter`)."""
    rng = np.random.default_rng(seed)
    # chord^2 = (2 R sin(dphi/2))^2 + (pitch dphi / 2 pi)^2 = d_cc^2, solved by bisection
    lo, hi = 0.0, np.pi
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        d = np.hypot(2 * radius * np.sin(mid / 2), pitch * mid / (2 * np.pi))
        lo, hi = (mid, hi) if d < d_cc else (lo, mid)
    dphi = 0.5 * (lo + hi)
    phi = rng.uniform(0, 2 * np.pi) + dphi * np.arange(n_atoms)
    pos = np.stack([radius * np.cos(phi), radius * np.sin(phi), pitch * (phi - phi[0]) / (2 * np.pi)], 1)
    return pos + jitter * rng.standard_normal(pos.shape)


def sk_block(d, v_ss, v_sp, v_pps, v_ppp):
    """4x4 two-centre block <a_i | . | b_j> for the bond vector d = r_j - r_i."""
    r = np.linalg.norm(d)
    c = d / r
    B = np.zeros((4, 4))
    B[0, 0] = v_ss
    for a in range(3):
        B[0, 1 + a] = c[P_AXIS[a]] * v_sp
        B[1 + a, 0] = -c[P_AXIS[a]] * v_sp
        for b in range(3):
            ca, cb = c[P_AXIS[a]], c[P_AXIS[b]]
            B[1 + a, 1 + b] = ca * cb * (v_pps - v_ppp) + (v_ppp if a == b else 0.0)
    return B


def synth_h0_s(pos, cutoff=1e-8):
    """Real symmetric H0 [eV] and S (4 orbitals per atom, DFTB+ order)."""
    na = len(pos)
    N = 4 * na
    H0 = np.zeros((N, N))
    S = np.eye(N)
    for i in range(na):
        H0[4 * i, 4 * i] = EPS_S
        for a in range(1, 4):
            H0[4 * i + a, 4 * i + a] = EPS_P
        for j in range(na):
            if i == j:
                continue
            d = pos[j] - pos[i]
            f = np.exp(-(np.linalg.norm(d) - R0) / DECAY)
            if f < cutoff:
                continue
            Vb = sk_block(d, V_SS * f, V_SP * f, V_PP_SIGMA * f, V_PP_PI * f)
            H0[4 * i:4 * i + 4, 4 * j:4 * j + 4] = Vb
            S[4 * i:4 * i + 4, 4 * j:4 * j + 4] = S_OVER_V * Vb
    return H0, S


def write_square(path, A):
    """DFTB+ square-matrix file (hamsqr1.dat / oversqr.dat of a WriteHS run)."""
    n = A.shape[0]
    with open(path, "w") as f:
        f.write("#      REAL   NALLORB   NKPOINT\n          T%10d         1\n#   IKPOINT\n          1\n# MATRIX\n" % n)
        for i in range(n):
            f.write(" ".join("%.17e" % v for v in A[i]) + "\n")


def write_inputs(dirpath, n_atoms=30, seed=0, xi_p=0.00849, U=1.5, gamma=0.2, delta=0.05, volts=(-0.5, 0.5, 0.5),
                 order=2, T=300.0, mu=None):
    """hamsqr1.dat, oversqr.dat, helix.xyz and input.dat of a dftb_hs run of the
    drop-in driver.  Its leads couple to every orbital of the first and the last
    atom (lead_left_atoms / lead_right_atoms: the driver's, and the reference's,
    only form).  Returns the positions, H0 [eV] and S."""
    pos = helix_positions(n_atoms, seed=seed)
    H0, S = synth_h0_s(pos)
    os.makedirs(dirpath, exist_ok=True)
    write_square(os.path.join(dirpath, "hamsqr1.dat"), H0 / HARTREE)
    write_square(os.path.join(dirpath, "oversqr.dat"), S)
    with open(os.path.join(dirpath, "helix.xyz"), "w") as f:
        f.write("%d\nsynthetic carbon helix, radius 3 A, pitch 3.5 A, seed %d\n" % (n_atoms, seed))
        for r in pos:
            f.write("    C %15.8f %15.8f %15.8f\n" % tuple(r))
    if mu is None:
        mu = half_filling_mu(dftb_hs_hamiltonian(H0, S, xi_p))
    V0, Vf, dv = volts
    with open(os.path.join(dirpath, "input.dat"), "w") as f:
        f.write("# synthetic DFTB-like carbon helix (%d atoms, s/p, dual SOC), seed %d\n" % (n_atoms, seed))
        f.write("hamiltonian = dftb_hs\nham_file = hamsqr1.dat\noverlap_file = oversqr.dat\nxyz_file = helix.xyz\n")
        f.write("max_angular_momentum = C:p\nsoc_xi = C:%.8g\nsoc_form = dual\n" % xi_p)
        f.write("lead_left_atoms = 1\nlead_right_atoms = %d\n" % n_atoms)
        f.write("T = %.8g\nV = %.8g\nVf = %.8g\ndelv = %.8g\nmu = %.10g\norder = %d\nU_int = %.8g\n"
                % (T, V0, Vf, dv, mu, order, U))
        f.write("GammaL_up = %.8g\nGammaL_down = %.8g\nGammaR = %.8g\ndw = 1\nup = 1\ndelta = %.8g\n"
                % (gamma, gamma, gamma, delta))
    return pos, H0, S


def dftb_hs_hamiltonian(H0, S, xi_p, dual=True):
    """Orthogonal spin-orbit Hamiltonian of a dftb_hs run (src/hamiltonian.f90:611-708):
    spin fastest, index 2 * orbital + spin, every atom carrying s, p_y, p_z, p_x."""
    N = H0.shape[0]
    ev, U = np.linalg.eigh(0.5 * (S + S.T))
    if ev.min() <= 1e-8:
        raise ValueError("overlap matrix is not positive definite (min eigenvalue %g)" % ev.min())
    Smh = (U / np.sqrt(ev)) @ U.T
    eps = np.zeros((3, 3, 3))
    eps[0, 1, 2] = eps[1, 2, 0] = eps[2, 0, 1] = 1.0
    eps[0, 2, 1] = eps[2, 1, 0] = eps[1, 0, 2] = -1.0
    Lk = -1j * eps                                   # (L_k)_ab = -i eps_kab, a, b in (x, y, z)
    sig = np.array([[[0, 1], [1, 0]], [[0, -1j], [1j, 0]], [[1, 0], [0, -1]]], complex)
    Lam = np.zeros((2 * N, 2 * N), complex)
    for a0 in range(0, N, 4):
        for p in range(3):
            for q in range(3):
                blk = xi_p * sum(Lk[k, P_AXIS[p], P_AXIS[q]] * sig[k] for k in range(3)) / 2
                i, j = 2 * (a0 + 1 + p), 2 * (a0 + 1 + q)
                Lam[i:i + 2, j:j + 2] += blk
    I2 = np.eye(2)
    Sx2, Sm2, Hao = np.kron(S, I2), np.kron(Smh, I2), np.kron(0.5 * (H0 + H0.T), I2).astype(complex)
    Hao += 0.5 * (Sx2 @ Lam + Lam @ Sx2) if dual else Lam
    H = Sm2 @ Hao @ Sm2
    return 0.5 * (H + H.conj().T)


def half_filling_mu(H):
    """Midpoint of the gap at half filling (one electron per spin-orbital pair)."""
    e = np.linalg.eigvalsh(H)
    n = len(e) // 2
    return 0.5 * (e[n - 1] + e[n])


def problem(n_atoms=30, seed=0, n_omega=1 << 12, xi_p=0.00849, U=1.5, gamma=0.2, T=300.0, pz_leads=True,
            nhist=16, tolf=1e-8, maxscf=300):
    """ScfProblem keyword arguments of the synthetic helix: U on every spatial
    orbital, Gamma on the first and the last atom's p_z (both spins) or, with
    pz_leads False, on all their orbitals as the driver's leads; the grid spans
    the spectrum with 1 eV margins (src/quichotte.f90:81-86) in n_omega points;
    mu at the half-filling gap."""
    pos = helix_positions(n_atoms, seed=seed)
    H0, S = synth_h0_s(pos)
    H = dftb_hs_hamiltonian(H0, S, xi_p)
    n = H.shape[0]
    GL = np.zeros((n, n), complex)
    GR = np.zeros((n, n), complex)
    last = 4 * (n_atoms - 1)
    orbs_l = [2] if pz_leads else [0, 1, 2, 3]
    for o in orbs_l:
        for s in (0, 1):
            GL[2 * o + s, 2 * o + s] = gamma
            GR[2 * (last + o) + s, 2 * (last + o) + s] = gamma
    e = np.linalg.eigvalsh(H)
    w_init, w_fin = e[0] - 1.0, e[-1] + 1.0
    delta = (w_fin - w_init) / n_omega
    Us = np.zeros(n)
    Us[0::2] = U
    KB = 8.617333262e-5
    return dict(H=H, gammaL=GL, gammaR=GR, U_site=Us, order=2, n_omega=n_omega, w_init=w_init, delta=delta,
                beta=1.0 / (KB * T), mu=half_filling_mu(H), alpha_mix=0.5, nhist=nhist, maxscf=maxscf, tolf=tolf,
                single_history=True)
