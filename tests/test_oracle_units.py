"""Pins of the oracle's pieces against mathematics and library routines.

Each check is chosen so that a plausible mistake (wrong sign, dropped term,
transposed operand, off-by-one index) fails it.
"""
import ctypes

import numpy as np
import pytest

from qs_testlib import Oracle, helix_problem, oracle_lib, ptr


def test_good_fft_size():
    L = oracle_lib()
    assert L.orc_good_fft_size(719) == 720
    assert L.orc_good_fft_size(2 * 2 ** 14 - 1) == 2 ** 15
    assert L.orc_good_fft_size(11) == 12
    assert L.orc_good_fft_size(13) == 14
    assert L.orc_good_fft_size(1) == 1


@pytest.mark.parametrize("n", [1, 2, 7, 12, 35, 49, 720, 1024, 2 ** 15])
def test_fft_matches_numpy(n):
    L = oracle_lib()
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    f = x.copy()
    L.orc_fft(n, -1, ptr(f))
    np.testing.assert_allclose(f, np.fft.fft(x), rtol=0, atol=1e-12 * max(1, np.sqrt(n)) * 10)
    b = x.copy()
    L.orc_fft(n, +1, ptr(b))
    np.testing.assert_allclose(b, np.fft.ifft(x) * n, rtol=0, atol=1e-12 * max(1, np.sqrt(n)) * 10)


def test_fft_bruteforce_small():
    L = oracle_lib()
    n = 30
    x = np.arange(n) * (0.3 + 0.1j) + np.cos(np.arange(n))
    k = np.arange(n)
    ref = np.array([np.sum(x * np.exp(-2j * np.pi * l * k / n)) for l in range(n)])
    f = x.copy()
    L.orc_fft(n, -1, ptr(f))
    np.testing.assert_allclose(f, ref, atol=1e-11)


@pytest.mark.parametrize("n", [1, 2, 3, 8, 16, 33])
def test_invert_matches_lapack(n):
    L = oracle_lib()
    rng = np.random.default_rng(100 + n)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    B = np.asfortranarray(A.copy())
    assert L.orc_invert(n, ptr(B)) == 0
    np.testing.assert_allclose(B, np.linalg.inv(A), rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(A @ B, np.eye(n), atol=1e-11)


def test_invert_needs_pivoting_and_reports_singular():
    L = oracle_lib()
    A = np.asfortranarray(np.array([[0, 1], [1, 0]], complex))
    assert L.orc_invert(2, ptr(A)) == 0
    np.testing.assert_allclose(A, [[0, 1], [1, 0]], atol=1e-15)
    S = np.asfortranarray(np.array([[1, 2], [2, 4]], complex))
    assert L.orc_invert(2, ptr(S)) != 0


def test_single_level_closed_form():
    """One site, no U: G^r = 1/(w - E + i Gamma/2) per spin (Dyson, scf.f90:318-330)."""
    E, g = 0.3, 0.2
    from qs_testlib import Problem
    H = np.diag([E, E]).astype(complex)
    GL = np.diag([g, g]).astype(complex)
    GR = np.zeros((2, 2), complex)
    prob = Problem(H, GL, GR, np.zeros(2), 0, 300.0, 0.0, 0, 0, 0.01, [0.0], n_omega=200, w_init=-1.0)
    o = Oracle(prob)
    o.solve(0.0, True)
    w = prob.omega
    dos = o.dosd()
    np.testing.assert_allclose(dos[0], (1.0 / (w - E + 0.5j * g)).imag, rtol=1e-13, atol=1e-15)


def brute_sigma2(Gr, GL, U_act, delta):
    """Direct O(N^3) double sum: every Sigma2 term is sum_{n1-n2+n3=m} X(n1) Y(n2) Z(n3)
    (src/self_energies.f90:28-40); no FFT, no padding."""
    nA, _, N = Gr.shape
    pp = delta / (2 * np.pi)
    SR = np.zeros_like(Gr)
    SL = np.zeros_like(Gr)

    def adj(M):
        return np.array([[M[1, 1], -M[0, 1]], [-M[1, 0], M[0, 0]]])

    for ia in range(nA // 2):
        for ja in range(nA // 2):
            io, jo = 2 * ia, 2 * ja
            c = U_act[io] * U_act[jo] * pp * pp
            if c == 0:
                continue
            r_ij = Gr[io:io + 2, jo:jo + 2, :]
            L_ij = GL[io:io + 2, jo:jo + 2, :]
            r_ji = Gr[jo:jo + 2, io:io + 2, :]
            L_ji = GL[jo:jo + 2, io:io + 2, :]
            a_ji = np.conj(np.transpose(r_ij, (1, 0, 2)))
            a_ij = np.conj(np.transpose(r_ji, (1, 0, 2)))
            g_ij = L_ij + r_ij - a_ij
            g_ji = L_ji + r_ji - a_ji
            for m in range(N):
                sr = np.zeros((2, 2), complex)
                sl = np.zeros((2, 2), complex)
                for n1 in range(N):
                    for n2 in range(N):
                        n3 = m - n1 + n2
                        if n3 < 0 or n3 >= N:
                            continue
                        sr += r_ij[:, :, n1] @ adj(L_ji[:, :, n2] @ g_ij[:, :, n3])
                        sr += L_ij[:, :, n1] @ adj(L_ji[:, :, n2] @ r_ij[:, :, n3] + a_ji[:, :, n2] @ L_ij[:, :, n3])
                        sl += L_ij[:, :, n1] @ adj(g_ji[:, :, n2] @ L_ij[:, :, n3])
                SR[io:io + 2, jo:jo + 2, m] = c * sr
                SL[io:io + 2, jo:jo + 2, m] = c * sl
    return SR, SL


@pytest.mark.parametrize("N", [5, 8, 13])
def test_sigma2_fft_equals_direct_sum(N):
    L = oracle_lib()
    rng = np.random.default_rng(N)
    nA = 6
    Gr = rng.standard_normal((nA, nA, N)) + 1j * rng.standard_normal((nA, nA, N))
    A = rng.standard_normal((nA, nA, N)) + 1j * rng.standard_normal((nA, nA, N))
    GL = A - np.conj(np.transpose(A, (1, 0, 2)))          # anti-Hermitian at each energy
    U = np.array([0.8, 0.0, 0.0, 0.0, 1.3, 0.0])           # site 1 non-interacting
    delta = 0.05
    # oracle layout: (nA, nA, N) column-major == numpy order='F' on (nA,nA,N)
    GrF = np.asfortranarray(Gr)
    GLF = np.asfortranarray(GL)
    SR = np.zeros((nA, nA, N), complex, order="F")
    SLo = np.zeros((nA, nA, N), complex, order="F")
    L.orc_sigma2(nA, N, delta, ptr(U), ptr(GrF), ptr(GLF), ptr(SR), ptr(SLo), 1)
    bR, bL = brute_sigma2(Gr, GL, U, delta)
    np.testing.assert_allclose(SR, bR, atol=1e-12 * np.abs(bR).max())
    np.testing.assert_allclose(SLo, bL, atol=1e-12 * np.abs(bL).max())
    assert np.abs(SR[0:2, 2:4]).max() == 0 and np.abs(SR[2:4, :]).max() == 0


def test_pack_index_map_and_round_trip():
    L = oracle_lib()
    nb, Nw = 3, 2
    rng = np.random.default_rng(7)
    Gr = np.asfortranarray(rng.standard_normal((nb, nb, Nw)) + 1j * rng.standard_normal((nb, nb, Nw)))
    A = rng.standard_normal((nb, nb, Nw)) + 1j * rng.standard_normal((nb, nb, Nw))
    GL = np.asfortranarray(A - np.conj(np.transpose(A, (1, 0, 2))))
    nblk = 3 * nb * nb
    x = np.zeros(nblk * Nw)
    L.orc_pack_gf(ptr(Gr), ptr(GL), ptr(x), nb, Nw)
    # explicit index map (common/linear_algebra.f90:142-191): row-major, j fastest
    for w in range(Nw):
        off = w * nblk
        assert x[off + 0 * nb + 1] == Gr[0, 1, w].real
        assert x[off + nb * nb + 2 * nb + 0] == Gr[2, 0, w].imag
        assert x[off + 2 * nb * nb + 0] == GL[0, 1, w].real        # Re G<(1,2)
        assert x[off + 2 * nb * nb + 2] == GL[1, 2, w].real        # Re G<(2,3)
        nsu = nb * (nb - 1) // 2
        assert x[off + 2 * nb * nb + nsu + 0] == GL[0, 0, w].imag  # Im G<(1,1)
        assert x[off + 2 * nb * nb + nsu + 3] == GL[1, 1, w].imag  # Im G<(2,2)
    Gr2 = np.zeros_like(Gr)
    GL2 = np.zeros_like(GL)
    L.orc_unpack_gf(ptr(x), ptr(Gr2), ptr(GL2), nb, Nw)
    np.testing.assert_array_equal(Gr2, Gr)
    GLexp = GL.copy()
    for w in range(Nw):
        for i in range(nb):
            GLexp[i, i, w] = 1j * GL[i, i, w].imag
    np.testing.assert_allclose(GL2, GLexp, atol=0)


def test_keldysh_fluctuation_dissipation_order0():
    """Order 0, V = 0: G^< = -f (G^r - G^a) in equilibrium (both leads at mu)."""
    prob = helix_problem(4, 1, 0.0, order=0, n_omega=128, voltages=[0.0])
    o = Oracle(prob)
    o.solve(0.0, True)
    G = o.g_less_obs()
    # occupations from the final sweep equal the integral of f(w) A(w)/2pi
    dos = o.dosd()
    f = np.array([o.L.orc_fermi(w, 0.0, prob.mu, prob.beta) for w in prob.omega])
    nocc = -(prob.delta / np.pi) * (dos * f[None, :]).sum(axis=1)
    np.testing.assert_allclose((-1j * np.diag(G)).real, nocc, rtol=1e-10, atol=1e-13)
    J = o.current()
    assert abs(J[0]) < 1e-14 and abs(J[1]) < 1e-14


def test_sanity_flags():
    L = oracle_lib()
    nb, nw = 2, 4
    Grd = np.zeros((nb, nw), complex, order="F")
    GLd = np.zeros((nb, nw), complex, order="F")
    Grd[:, :] = -1j * 10.0
    GLd[:, :] = 1j * 0.5 * 2 * np.pi / (0.1 * nw)   # n = 0.5
    out = [ctypes.c_int32() for _ in range(3)] + [ctypes.c_double() for _ in range(2)]

    def run():
        L.orc_check_diag_sanity(ptr(Grd), ptr(GLd), nb, nw, 0.1, *[ctypes.byref(v) for v in out], None)
        return [v.value for v in out]

    ra, un, iw, nw_, ww = run()
    assert (ra, un) == (0, 0) and abs(nw_ - 0.5) < 1e-12
    GLd[1, :] *= 3.0                                 # n = 1.5: unphysical, not runaway
    ra, un, iw, nw_, ww = run()
    assert (ra, un, iw) == (0, 1, 2) and abs(nw_ - 1.5) < 1e-12
    GLd[1, :] *= 2.0                                 # n = 3: runaway
    assert run()[0:2] == [1, 1]
    GLd[0, 1] = np.nan
    ra, un, iw, _, _ = run()
    assert ra == 1 and iw == 1


def test_broyden_affine_fixed_point():
    """F(x) = A x + b - x with a contraction A: Broyden reaches x* = (I-A)^-1 b."""
    L = oracle_lib()
    n = 50
    rng = np.random.default_rng(3)
    Q = rng.standard_normal((n, n))
    A = 0.9 * Q / np.linalg.norm(Q, 2)
    b = rng.standard_normal(n)
    xstar = np.linalg.solve(np.eye(n) - A, b)
    CB = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_double),
                          ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int32))
    calls = []

    def res(ctx, nn, xp, Fp, ab):
        xv = np.ctypeslib.as_array(xp, (nn,))
        Fv = np.ctypeslib.as_array(Fp, (nn,))
        Fv[:] = A @ xv + b - xv
        ab[0] = 0
        calls.append(1)

    cb = CB(res)
    L.orc_broyden.argtypes = [CB, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_double, ctypes.c_int,
                              ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                              ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    for sp in (0, 1):
        x = np.zeros(n)
        conv, nit, fmax = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_double()
        L.orc_broyden(cb, None, ptr(x), n, 0.5, 16, 300, 1e-12, sp, ctypes.byref(conv), ctypes.byref(nit),
                      ctypes.byref(fmax), None, None)
        assert conv.value == 1 and fmax.value < 1e-12
        np.testing.assert_allclose(x, xstar, atol=1e-10)
        assert nit.value < 100   # Broyden beats linear mixing (which needs ~300 here)
    # nhist = 0 is plain linear mixing: x_{k+1} = x_k + alpha F_k, checked step by step
    x = np.zeros(n)
    conv, nit, fmax = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_double()
    L.orc_broyden(cb, None, ptr(x), n, 0.5, 0, 5, 1e-30, 0, ctypes.byref(conv), ctypes.byref(nit),
                  ctypes.byref(fmax), None, None)
    xr = np.zeros(n)
    for _ in range(5):
        xr = xr + 0.5 * (A @ xr + b - xr)
    np.testing.assert_allclose(x, xr, rtol=1e-13, atol=1e-14)
