"""Pins of the oracle that do not restate its formulas.

Second-order self-energy (sigma2_compute, src/self_energies.f90:158-212,222-228):
  * SU(2) covariance: one spin rotation R on every site, G -> (1 x R) G (1 x R)^+,
    gives Sigma2 -> (1 x R) Sigma2 (1 x R)^+ (U n_up n_dn is rotation invariant,
    and adj(M) = Tr(M) 1 - M is covariant), for any input data;
  * the spin-degenerate limit: for G = g (x) 1_spin with a spinless g, each spin
    block of Sigma2 is SANCHO's second-order self-energy of g
    (SANCHO/src/self_energies.f90:115-147), a different code (direct bubble sums,
    pinned by SANCHO's own regression cases).
  A numpy restatement with planted mistakes (a dropped a_ji L_ij term, a
  conjugated L_ji, a transposed adjugate, a transposed instead of adjoint a_ij)
  is first shown equal to the oracle without them; every plant then fails one of
  the two pins.

Broyden (broyden_solve, common/broyden_mixing.f90:51-184): on a one-dimensional
affine residual the second iterate is the secant root shifted by the w0^2 = 1e-4
regulariser, by a closed form; on a two-dimensional affine residual the third
iterate is the root up to O(w0^2) (two stored secant pairs span the space).

End to end: QUICHOTTE's custom chain without SOC and spin-independent leads
reproduces SANCHO's order-2 regression case (J = J_up + J_dn, n_up = n_dn =
n per spin).
"""
import ctypes
import os

import numpy as np
import pytest

from qs_testlib import GOLDEN, KB, Oracle, Problem, custom_hamiltonian, leads, oracle_lib, parse_input, ptr, freal, \
    read_table, sancho_oracle


# ---------------------------------------------------------------------------
# Sigma2
# ---------------------------------------------------------------------------
def oracle_sigma2(Gr, GL, U_act, delta):
    nA, _, nw = Gr.shape
    Gr = np.asfortranarray(Gr, dtype=complex)
    GL = np.asfortranarray(GL, dtype=complex)
    SR = np.zeros((nA, nA, nw), complex, order="F")
    SL = np.zeros((nA, nA, nw), complex, order="F")
    L = oracle_lib()
    L.orc_sigma2(nA, nw, delta, ptr(np.ascontiguousarray(U_act, dtype=float)), ptr(Gr), ptr(GL), ptr(SR), ptr(SL), 1)
    return SR, SL


def numpy_sigma2(Gr, GL, U_act, delta, plant=None):
    """sigma2_compute restated with numpy FFTs (same zero padding); `plant`
    names a deliberate mistake."""
    nA, _, nw = Gr.shape
    Nf = oracle_lib().orc_good_fft_size(2 * nw - 1)
    Fr = np.fft.fft(Gr, n=Nf, axis=2)
    FL = np.fft.fft(GL, n=Nf, axis=2)
    lm = (Nf - np.arange(Nf)) % Nf
    dag = (lambda M: np.swapaxes(M, 1, 2)) if plant == "transposed_adjoint" else \
        (lambda M: np.conj(np.swapaxes(M, 1, 2)))
    eye = np.eye(2)

    def adj(M):
        tr = (M[:, 0, 0] + M[:, 1, 1])[:, None, None]
        return tr * eye - (np.swapaxes(M, 1, 2) if plant == "transposed_adjugate" else M)

    SR = np.zeros((nA, nA, nw), complex)
    SL = np.zeros((nA, nA, nw), complex)
    pp = delta / (2 * np.pi)
    for ia in range(nA // 2):
        for ja in range(nA // 2):
            io, jo = 2 * ia, 2 * ja
            c = U_act[io] * U_act[jo] * pp * pp / Nf
            if c == 0.0:
                continue
            blk = lambda F, i, j, idx: np.moveaxis(F[i:i + 2, j:j + 2, :], 2, 0)[idx]
            ar = np.arange(Nf)
            r_ij, L_ij = blk(Fr, io, jo, ar), blk(FL, io, jo, ar)
            r_ji, L_ji = blk(Fr, jo, io, lm), blk(FL, jo, io, lm)
            if plant == "conjugated_Lji":
                L_ji = np.conj(L_ji)
            a_ij, a_ji = dag(r_ji), dag(r_ij)
            G_ij = L_ij + r_ij - a_ij
            G_ji = L_ji + r_ji - a_ji
            chi1 = L_ji @ G_ij
            chi2 = L_ji @ r_ij + (0 if plant == "dropped_aL" else a_ji @ L_ij)
            sr = r_ij @ adj(chi1) + L_ij @ adj(chi2)
            sl = L_ij @ adj(G_ji @ L_ij)
            SR[io:io + 2, jo:jo + 2, :] = np.moveaxis(np.fft.ifft(sr, axis=0)[:nw] * Nf * c, 0, 2)
            SL[io:io + 2, jo:jo + 2, :] = np.moveaxis(np.fft.ifft(sl, axis=0)[:nw] * Nf * c, 0, 2)
    return SR, SL


def random_gf(nA, nw, seed):
    rng = np.random.default_rng(seed)
    Gr = rng.standard_normal((nA, nA, nw)) + 1j * rng.standard_normal((nA, nA, nw))
    A = rng.standard_normal((nA, nA, nw)) + 1j * rng.standard_normal((nA, nA, nw))
    GL = A - np.conj(np.swapaxes(A, 0, 1))    # anti-Hermitian at every energy, as G^< is
    U = np.zeros(nA)
    U[0::2] = 0.5 + rng.random(nA // 2)
    return Gr, GL, U


def spin_rotation(theta, axis):
    axis = np.asarray(axis, float) / np.linalg.norm(axis)
    sx = np.array([[0, 1], [1, 0]], complex)
    sy = np.array([[0, -1j], [1j, 0]], complex)
    sz = np.array([[1, 0], [0, -1]], complex)
    n_sigma = axis[0] * sx + axis[1] * sy + axis[2] * sz
    return np.cos(theta / 2) * np.eye(2) - 1j * np.sin(theta / 2) * n_sigma


def rotate(G, R):
    nA = G.shape[0]
    Rb = np.kron(np.eye(nA // 2), R)
    return np.einsum("ab,bcw,dc->adw", Rb, G, np.conj(Rb))


def covariance_error(fn, Gr, GL, U, delta, R):
    SR, SL = fn(Gr, GL, U, delta)
    SR2, SL2 = fn(rotate(Gr, R), rotate(GL, R), U, delta)
    scale = max(np.abs(SR).max(), np.abs(SL).max())
    return max(np.abs(SR2 - rotate(SR, R)).max(), np.abs(SL2 - rotate(SL, R)).max()) / scale


def sancho_degenerate_error(fn, seed=5, na=3, nw=40):
    """Each spin block of Sigma2[g (x) 1] against SANCHO's Sigma2[g]."""
    from qs_testlib import sancho_inputs
    gR, gL, Us, delta = sancho_inputs(na, nw, seed=seed)
    nA = 2 * na
    Gr = np.zeros((nA, nA, nw), complex)
    GL = np.zeros((nA, nA, nw), complex)
    for s in range(2):
        Gr[s::2, s::2, :] = gR
        GL[s::2, s::2, :] = gL
    U = np.zeros(nA)
    U[0::2] = Us
    SR, SL = fn(Gr, GL, U, delta)
    s2r, s2l, _ = sancho_oracle(gR, gL, Us, delta, 2)
    err = 0.0
    for s in range(2):
        err = max(err, np.abs(SR[s::2, s::2, :] - s2r).max() / np.abs(s2r).max(),
                  np.abs(SL[s::2, s::2, :] - s2l).max() / np.abs(s2l).max())
    err = max(err, np.abs(SR[0::2, 1::2, :]).max() / np.abs(s2r).max())   # no spin mixing
    return err


PLANTS = ["dropped_aL", "conjugated_Lji", "transposed_adjugate", "transposed_adjoint"]


def test_numpy_restatement_equals_oracle():
    Gr, GL, U = random_gf(6, 37, seed=1)
    SRo, SLo = oracle_sigma2(Gr, GL, U, 0.05)
    SRn, SLn = numpy_sigma2(Gr, GL, U, 0.05)
    assert np.abs(SRn - SRo).max() <= 1e-12 * np.abs(SRo).max()
    assert np.abs(SLn - SLo).max() <= 1e-12 * np.abs(SLo).max()


@pytest.mark.parametrize("theta,axis", [(0.7, (1, 0, 0)), (2.1, (0.3, -0.5, 0.8)), (np.pi, (0, 0, 1))])
def test_oracle_sigma2_is_spin_rotation_covariant(theta, axis):
    Gr, GL, U = random_gf(6, 41, seed=2)
    assert covariance_error(oracle_sigma2, Gr, GL, U, 0.03, spin_rotation(theta, axis)) < 1e-12


def test_oracle_sigma2_spin_degenerate_limit_is_sanchos():
    assert sancho_degenerate_error(oracle_sigma2) < 1e-12


@pytest.mark.parametrize("plant", PLANTS)
def test_planted_mistakes_fail_a_pin(plant):
    Gr, GL, U = random_gf(6, 41, seed=2)
    fn = lambda a, b, c, d: numpy_sigma2(a, b, c, d, plant=plant)
    cov = covariance_error(fn, Gr, GL, U, 0.03, spin_rotation(2.1, (0.3, -0.5, 0.8)))
    deg = sancho_degenerate_error(fn)
    assert cov > 1e-3 or deg > 1e-3, (plant, cov, deg)


# ---------------------------------------------------------------------------
# Broyden
# ---------------------------------------------------------------------------
CB = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_double),
                      ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int32))


def broyden_iterates(A, b, x0, nhist, maxscf, single):
    """x of every residual call of orc_broyden on F(x) = A x + b"""
    L = oracle_lib()
    n = len(x0)
    xs = []

    def res(ctx, nn, xp, Fp, ab):
        xv = np.ctypeslib.as_array(xp, (nn,)).copy()
        xs.append(xv)
        np.ctypeslib.as_array(Fp, (nn,))[:] = A @ xv + b
        ab[0] = 0

    cb = CB(res)
    L.orc_broyden.argtypes = [CB, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_double, ctypes.c_int,
                              ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                              ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    x = np.array(x0, float)
    conv, nit, fmax = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_double()
    L.orc_broyden(cb, None, ptr(x), n, 0.5, nhist, maxscf, 1e-300, int(single), ctypes.byref(conv),
                  ctypes.byref(nit), ctypes.byref(fmax), None, None)
    return xs


@pytest.mark.parametrize("single,tol", [(0, 1e-14), (1, 1e-6)])
def test_broyden_second_iterate_is_regularised_secant_root(single, tol):
    """n = 1, F(x) = a x + b: x1 = x0 + alpha F0, and with dF = +-1, gram = 1,
    x2 = x1 + alpha F1 - (alpha + dx/dF) F1 / (1 + w0^2); as w0 -> 0 this is
    the secant root x1 - F1 dx/dF.  A wrong w0 or a dropped (x - xold) term in
    the history vector u moves x2 off this value."""
    a, b, alpha, w0 = -0.7, 0.9, 0.5, 0.01
    xs = broyden_iterates(np.array([[a]]), np.array([b]), [0.0], 4, 3, single)
    x0, x1, x2 = xs[0][0], xs[1][0], xs[2][0]
    F0, F1 = a * x0 + b, a * x1 + b
    assert x1 == x0 + alpha * F0
    dx, dF = x1 - x0, F1 - F0
    expect = x1 + alpha * F1 - (alpha + dx / dF) * F1 / (1 + w0 ** 2)
    assert abs(x2 - expect) <= tol * abs(expect)
    root = -b / a
    assert abs(x1 - dx / dF * F1 - root) <= 1e-15     # the secant root of an affine F is exact
    assert 1e-6 < abs(x2 - root) / abs(x1 - root) < 1e-3   # off it only by the O(w0^2) regulariser


def test_broyden_two_secant_pairs_give_the_root_in_2d():
    """n = 2, affine F: after two stored pairs the multi-secant update is exact
    on the plane, so x3 is the root up to O(w0^2); x2 is not."""
    A = np.array([[-0.8, 0.3], [0.1, -0.5]])
    b = np.array([0.4, -0.7])
    root = np.linalg.solve(A, -b)
    xs = broyden_iterates(A, b, [0.0, 0.0], 4, 4, 0)
    e2 = np.linalg.norm(xs[2] - root)
    e3 = np.linalg.norm(xs[3] - root)
    assert e2 > 1e-2 * np.linalg.norm(root)
    # with w0 = 0 the Gram solve removes F's whole component in the span of the
    # normalised dF_k, which is the plane; the w0^2 regulariser leaves a part of
    # relative size w0^2 / lambda_min(Gram)
    F = [A @ x + b for x in xs]
    dF = [(F[k + 1] - F[k]) / np.linalg.norm(F[k + 1] - F[k]) for k in range(2)]
    lam = np.linalg.eigvalsh(np.array([[u @ v for v in dF] for u in dF])).min()
    assert 1e-4 / lam < 0.05
    assert e3 < 5 * (1e-4 / lam) * e2, (e3, e2, lam)


# ---------------------------------------------------------------------------
# QUICHOTTE's chain against SANCHO's regression case
# ---------------------------------------------------------------------------
def test_custom_chain_reproduces_sanchos_order2_case():
    cd = os.path.join(GOLDEN, "sancho", "order2_second_order")
    d = parse_input(os.path.join(cd, "input.dat"))
    N = int(d["n_sites"]) if "n_sites" in d else int(d["n"])
    E_CC, t_hop = freal(d["e_cc"]), freal(d["t_hop"])
    H = custom_hamiltonian(N, E_CC, t_hop)
    gl = freal(d.get("gammal", d.get("gamma_l", "1.0")))
    gr = freal(d.get("gammar", d.get("gamma_r", "1.0")))
    GL, GR = leads(2 * N, gl, gl, gr)
    U = np.zeros(2 * N)
    U[0::2] = freal(d.get("u_int", d.get("u", "0")))
    V, Vf, dv = freal(d["v"]), freal(d["vf"]), freal(d["delv"])
    volts = [V + k * dv for k in range(int(round((Vf - V) / dv)) + 1)]
    prob = Problem(H, GL, GR, U, 2, freal(d["t"]), freal(d["mu"]), freal(d["dw"]), freal(d["up"]),
                   freal(d["delta"]), volts, alpha_mix=freal(d.get("alpha_mix", "0.5")),
                   nhist=int(d.get("nhist", 16)), maxscf=int(d.get("maxscf", 300)),
                   tolf=freal(d.get("tolf", "1e-6")))
    o = Oracle(prob)
    ref = read_table(os.path.join(cd, "reference", [f for f in os.listdir(os.path.join(cd, "reference"))
                                                    if f.startswith("Volt_Current")][0]))
    pops = read_table(os.path.join(cd, "reference", "populations.dat"))
    fresh = True
    worst = 0.0
    for V, row in zip(volts, ref):
        r = o.solve(V, fresh)
        fresh = not r["converged"]
        ju, jd = o.current()
        assert abs(row[0] - V) < 1e-9
        worst = max(worst, abs(ju + jd - row[1]) / (1e-6 + 1e-5 * abs(row[1])))
        G = o.g_less_obs()
        for p in [p for p in pops if abs(p[0] - V) < 1e-9]:
            i = int(p[1]) - 1
            n_up, n_dn = (-1j * G[2 * i, 2 * i]).real, (-1j * G[2 * i + 1, 2 * i + 1]).real
            for n in (n_up, n_dn):
                worst = max(worst, abs(n - p[2]) / (1e-6 + 1e-5 * abs(p[2])))
    assert worst <= 1.0, worst


@pytest.mark.parametrize("V", [0.0, 0.3, -0.45])
def test_order0_current_is_landauer(V):
    """The non-interacting current (order 0) equals the Landauer form per energy and
    per spin, with spin-orbit coupling mixing the spins (src/observables.f90:22-39,
    src/scf.f90:390-409); J(V = 0) = 0."""
    from qs_testlib import landauer_integrand, polarized_helix
    prob = polarized_helix(V)
    o = Oracle(prob)
    o.solve(V, True)
    tu, td = o.trUD()
    lu, ld = landauer_integrand(prob, V)
    scale = max(np.abs(lu).max(), np.abs(ld).max(), 1e-300)
    if V == 0.0:
        assert np.abs(tu).max() < 1e-15 and np.abs(td).max() < 1e-15
        assert np.abs(np.array(o.current())).max() < 1e-15
    else:
        assert np.abs(tu - lu).max() <= 1e-12 * scale and np.abs(td - ld).max() <= 1e-12 * scale
        # the spin-polarized lead makes the spin channels differ
        assert np.abs(lu - ld).max() > 1e-3 * scale
        J = np.array(o.current())
        assert np.allclose(J, [prob.delta * lu.sum(), prob.delta * ld.sum()], rtol=1e-11, atol=0)
