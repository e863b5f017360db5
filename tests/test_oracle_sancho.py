"""Pins of the SANCHO self-energy oracle (oracle/sancho_oracle.c).

The reference holds no fixture at the level of the bubbles, so the oracle is
pinned by routes that do not share its loops:
  * every sum rebuilt with numpy's full convolution, whose index offsets are
    derived here from the definitions (a correlation is a convolution with one
    operand reversed);
  * identities that hold term by term on the truncated grid for physical inputs
    (G^a = G^r^dagger, G^< and G^> anti-Hermitian, G^> - G^< = G^r - G^a);
  * the closed form of one site on one energy.
A swapped (a,b)/(b,a) operand, a sign of m, an off-by-one between the chi and
phi axes or a dropped term each fail one of them.
"""
import numpy as np
import pytest

from qs_testlib import oracle_lib, sancho_inputs, sancho_oracle, sancho_oracle_bubbles


def completed(GR, GL):
    GA = np.conj(GR.transpose(1, 0, 2))
    return GR, GA, GL, GL + GR - GA


def conv_bubbles(GR, GL, delta):
    """chi(m) = pp sum_k x_ab(k) y_ba(k - m)  -> convolve(x, y[::-1])[m + nw - 1]
    phi(n) = pp sum_k x_ab(k) y_ab(n - 2 - k) -> convolve(x, y)[n - 2]"""
    na, nw = GR.shape[0], GR.shape[2]
    pp = delta / (2 * np.pi)
    R, A, L, G = completed(GR, GL)
    chi = np.zeros((4, na, na, 2 * nw - 1), complex)
    phi = np.zeros_like(chi)
    c = lambda x, y: np.convolve(x, y[::-1])
    for a in range(na):
        for b in range(na):
            chi[0, a, b] = c(G[a, b], L[b, a])
            chi[1, a, b] = c(L[a, b], G[b, a])
            chi[2, a, b] = c(R[a, b], L[b, a]) + c(L[a, b], A[b, a])
            chi[3, a, b] = c(A[a, b], L[b, a]) + c(L[a, b], R[b, a])
            phi[0, a, b] = np.convolve(G[a, b], G[a, b])
            phi[1, a, b] = np.convolve(L[a, b], L[a, b])
            phi[2, a, b] = np.convolve(R[a, b], G[a, b]) + np.convolve(L[a, b], R[a, b])
            phi[3, a, b] = np.convolve(A[a, b], G[a, b]) + np.convolve(L[a, b], A[a, b])
    return chi * pp, phi * pp


def conv_sigma(GR, GL, U, delta, order):
    na, nw = GR.shape[0], GR.shape[2]
    pp = delta / (2 * np.pi)
    R, A, L, G = completed(GR, GL)
    chi, phi = conv_bubbles(GR, GL, delta)
    cG, cL, cR, cA = chi
    sl = slice(nw - 1, 2 * nw - 1)  # w + nw - 1 for w = 0..nw-1
    out2 = np.zeros((3, na, na, nw), complex)
    for i in range(na):
        for j in range(na):
            s = -(1j) ** 2 * U[i] * U[j] * pp
            out2[0, i, j] = s * (np.convolve(cG[i, j], R[i, j]) + np.convolve(cR[i, j], L[i, j]))[sl]
            out2[1, i, j] = s * np.convolve(cL[i, j], L[i, j])[sl]
            out2[2, i, j] = s * np.convolve(cG[i, j], G[i, j])[sl]
    if order == 2:
        return out2
    pG, pL, pR, pA = phi
    ein = lambda X, Y: np.einsum("ikm,k,kjm->ijm", X, U, Y)
    SG, SL, SR = ein(cR, cG) + ein(cG, cA), ein(cR, cL) + ein(cL, cA), ein(cR, cR)
    PG, PL, PR = ein(pR, pG) + ein(pG, pA), ein(pR, pL) + ein(pL, pA), ein(pR, pR)
    out3 = np.zeros((3, na, na, nw), complex)
    for i in range(na):
        for j in range(na):
            s = -(1j) ** 3 * U[i] * U[j] * pp
            ph = lambda X, Y: np.convolve(X[i, j], Y[i, j])[sl]
            pq = lambda P, Y: np.convolve(P[i, j], Y[j, i][::-1])[sl]  # sum_t P(w + t) y_ji(t)
            out3[0, i, j] = s * (ph(SG, R) + ph(SR, L) + pq(PG, A) + pq(PR, G))
            out3[1, i, j] = s * (ph(SL, L) + pq(PL, G))
            out3[2, i, j] = s * (ph(SG, G) + pq(PG, L))
    return out2, out3


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("na,nw,seed", [(1, 1, 0), (2, 7, 1), (3, 40, 2), (4, 33, 3)])
def test_bubbles_match_convolution(na, nw, seed):
    GR, GL, U, delta = sancho_inputs(na, nw, seed)
    chi, phi = sancho_oracle_bubbles(GR, GL, delta, 3)
    chi_c, phi_c = conv_bubbles(GR, GL, delta)
    assert rel(chi, chi_c) < 1e-13
    assert rel(phi, phi_c) < 1e-13


@pytest.mark.parametrize("order", [2, 3])
@pytest.mark.parametrize("na,nw,seed", [(1, 1, 4), (2, 9, 5), (3, 31, 6)])
def test_sigma_matches_convolution(order, na, nw, seed):
    GR, GL, U, delta = sancho_inputs(na, nw, seed)
    outs = sancho_oracle(GR, GL, U, delta, order)
    ref = conv_sigma(GR, GL, U, delta, order)
    ref2 = ref if order == 2 else ref[0]
    for q in range(3):
        assert rel(outs[q], ref2[q]) < 1e-12, q
    if order == 3:
        for q in range(3):
            assert rel(outs[3 + q], ref[1][q]) < 1e-12, q


def test_bubble_identities():
    GR, GL, U, delta = sancho_inputs(4, 50, 7)
    chi, phi = sancho_oracle_bubbles(GR, GL, delta, 3)
    cG, cL, cR, cA = chi
    pG, pL, pR, pA = phi
    scale = np.max(np.abs(chi))
    # retarded - advanced = greater - lesser, term by term on the grid
    assert np.max(np.abs((cR - cA) - (cG - cL))) < 1e-14 * scale * 50
    assert np.max(np.abs((pR - pA) - (pG - pL))) < 1e-14 * np.max(np.abs(phi)) * 50
    # chi^r(m,a,b)* = -chi^a(m,b,a); chi^<(m,a,b)* = chi^>(-m,a,b)
    assert np.max(np.abs(np.conj(cR) + cA.transpose(1, 0, 2))) < 1e-14 * scale * 50
    assert np.max(np.abs(np.conj(cL) - cG[:, :, ::-1])) < 1e-14 * scale * 50


@pytest.mark.parametrize("order", [2, 3])
def test_sigma_spectral_identity(order):
    """Sigma^r - Sigma^a = Sigma^> - Sigma^< with Sigma^a(i,j) = Sigma^r(j,i)*,
    the check the reference keeps Sig2G / Sig3G for."""
    GR, GL, U, delta = sancho_inputs(4, 48, 8)
    outs = sancho_oracle(GR, GL, U, delta, order)
    for R, L, G in ([outs[0:3]] if order == 2 else [outs[0:3], outs[3:6]]):
        lhs = R - np.conj(R.transpose(1, 0, 2))
        assert np.max(np.abs(lhs - (G - L))) < 1e-12 * np.max(np.abs(R))
        # the lesser self-energy is anti-Hermitian
        assert np.max(np.abs(L + np.conj(L.transpose(1, 0, 2)))) < 1e-12 * np.max(np.abs(L))


def test_one_site_one_energy_closed_form():
    """na = nw = 1: chi = pp g^> g^<, the only m is 0, and Sigma2^< = U^2 pp^2 g^< g^< g^>"""
    GR = np.array([[[0.3 - 0.8j]]], order="F")
    GL = np.array([[[0.45j]]], order="F")
    U, delta = np.array([1.7]), 0.02
    pp = delta / (2 * np.pi)
    S2R, S2L, S2G = sancho_oracle(GR, GL, U, delta, 2)
    r, l = GR[0, 0, 0], GL[0, 0, 0]
    a = np.conj(r)
    g = l + r - a
    assert abs(S2L[0, 0, 0] - U[0] ** 2 * pp ** 2 * l * l * g) < 1e-15
    assert abs(S2G[0, 0, 0] - U[0] ** 2 * pp ** 2 * g * g * l) < 1e-15
    assert abs(S2R[0, 0, 0] - U[0] ** 2 * pp ** 2 * (r * g * l + l * (r * l + l * a))) < 1e-15


def test_rejects_bad_arguments():
    L = oracle_lib()
    GR, GL, U, delta = sancho_inputs(2, 4, 9)
    from qs_testlib import ptr
    o = np.zeros((2, 2, 4), complex, order="F")
    assert L.orc_sancho_sigma(2, 4, delta, ptr(U), 4, ptr(GR), ptr(GL), ptr(o), ptr(o), ptr(o), None, None, None) == -1
    assert L.orc_sancho_sigma(0, 4, delta, ptr(U), 2, ptr(GR), ptr(GL), ptr(o), ptr(o), ptr(o), None, None, None) == -1
    assert L.orc_sancho_sigma(2, 4, delta, ptr(U), 3, ptr(GR), ptr(GL), ptr(o), ptr(o), ptr(o), None, None, None) == -1
