"""SANCHO's SCF voltage sweep around a pluggable self-energy function.

TEST INFRASTRUCTURE: it lets the reference's own SANCHO regression fixtures
(tests/golden/sancho/, copied from SANCHO/tests/cases/) pin the self-energies
end to end.  Everything except the self-energies follows the reference:
  grid and chain           SANCHO/src/sancho.f90:64-91, src/hamiltonian.f90:48-65,126-137
  fermi                    SANCHO/src/green_functions.f90:45-53 (no overflow clamp)
  non-interacting start    SANCHO/src/green_functions.f90:55-95
  scf_residual             SANCHO/src/scf.f90:102-135 (unpack, complete, equal_time_lesser,
                           self-energies, interacting_gf, check_gf_sanity, pack)
  interacting_gf           SANCHO/src/scf.f90:147-195
  solve_scf / sweep        SANCHO/src/scf.f90:53-95, src/sancho.f90:118-143
  Current, populations     SANCHO/src/observables.f90:10-55
The Broyden loop and pack/unpack are the oracle's (common/ is the same file
in SANCHO and QUICHOTTE); the per-energy Dyson/Keldysh step is numpy's LAPACK.
sigma_fn(GR, GL, U_site, delta, order) returns [Sig2R, Sig2L, Sig2G(, Sig3R, Sig3L, Sig3G)].
"""
import ctypes

import numpy as np

from qs_testlib import flogical, freal, oracle_lib, parse_input, ptr

HBAR, KB = 1.0, 8.6173303e-5


def key(inp, *names, default=None):
    for n in names:
        if n.lower() in inp:
            return inp[n.lower()]
    if default is None:
        raise KeyError(names[0])
    return default


class SanchoSweep:
    def __init__(self, input_path, sigma_fn):
        inp = parse_input(input_path)
        self.sigma_fn = sigma_fn
        N = int(key(inp, "N_sites", "Nsites", "Natoms"))
        E, t = freal(key(inp, "E_CC")), freal(key(inp, "t_hop"))
        H = np.zeros((N, N), complex)
        for i in range(N):
            H[i, i] = E
        for i in range(N - 1):
            H[i, i + 1] = t
            H[i + 1, i] = np.conj(H[i, i + 1])
        self.N, self.H = N, H
        self.U = np.full(N, freal(key(inp, "U_int")))
        self.order = int(key(inp, "order"))
        self.GamL = np.zeros((N, N), complex)
        self.GamR = np.zeros((N, N), complex)
        self.GamL[0, 0] = freal(key(inp, "GammaL", default="1.d0"))
        self.GamR[N - 1, N - 1] = freal(key(inp, "GammaR", default="1.d0"))
        self.mu = freal(key(inp, "mu"))
        self.beta = 1.0 / (KB * freal(key(inp, "T")))
        self.delta = freal(key(inp, "delta"))
        ev = np.linalg.eigvalsh(H)
        self.w_init = ev[0] - freal(key(inp, "dw"))
        w_fin = ev[-1] + freal(key(inp, "up"))
        self.nw = int((w_fin - self.w_init) / self.delta)
        self.omega = self.w_init + self.delta * np.arange(1, self.nw + 1)
        self.alpha = freal(key(inp, "alpha_mix", default="0.3"))
        self.nhist = int(key(inp, "nhist", default="8"))
        self.maxscf = int(key(inp, "maxscf", default="200"))
        self.tolf = freal(key(inp, "tolf", default="1.d-6"))
        self.continuation = flogical(key(inp, "continuation", default="T"))
        V0, Vf, dv = freal(key(inp, "V")), freal(key(inp, "Vf")), freal(key(inp, "delv"))
        self.volts = [V0 + k * dv for k in range(int(np.rint((Vf - V0) / dv)) + 1)]
        nf = N * N
        nsu = (nf - N) // 2
        self.n_vec = (nf + nf + nsu + nsu + N) * self.nw
        self.pp = self.delta / (2 * np.pi)

    def fermi(self, w, V):
        return 1.0 / (np.exp(HBAR * ((w - self.mu + V / HBAR) * self.beta)) + 1.0)

    # (N, N, nw) Fortran-ordered complex <-> packed x
    def pack(self, GR, GL):
        x = np.zeros(self.n_vec)
        R, L = np.asfortranarray(GR), np.asfortranarray(GL)
        oracle_lib().orc_pack_gf(ptr(R), ptr(L), ptr(x), self.N, self.nw)
        return x

    def unpack(self, x):
        GR = np.zeros((self.N, self.N, self.nw), complex, order="F")
        GL = np.zeros_like(GR)
        xc = np.ascontiguousarray(x)
        oracle_lib().orc_unpack_gf(ptr(xc), ptr(GR), ptr(GL), self.N, self.nw)
        return GR, GL

    def embedding(self):
        return -self.H + 0.5 * (1j / HBAR) * (self.GamL + self.GamR)

    def lesser_in(self, V):
        fL, fR = self.fermi(self.omega, V), self.fermi(self.omega, 0.0)
        return 1j * (fL[:, None, None] * self.GamL + fR[:, None, None] * self.GamR) / HBAR  # (nw, N, N)

    def noninteracting(self, V):
        W = self.embedding()[None] + HBAR * self.omega[:, None, None] * np.eye(self.N)
        R = np.linalg.inv(W)
        A = np.conj(np.transpose(R, (0, 2, 1)))
        L = R @ self.lesser_in(V) @ A
        return np.asfortranarray(np.transpose(R, (1, 2, 0))), np.asfortranarray(np.transpose(L, (1, 2, 0)))

    def residual(self, V, x):
        GR, GL = self.unpack(x)
        GA = np.conj(np.transpose(GR, (1, 0, 2)))
        GG = GL + GR - GA
        glt0 = self.pp * np.einsum("iik->i", GL)
        SigR = np.zeros((self.nw, self.N, self.N), complex)
        SigL = np.zeros_like(SigR)
        if self.order >= 2:
            outs = self.sigma_fn(GR, GL, self.U, self.delta, self.order)
            SigR += np.transpose(outs[0], (2, 0, 1))
            SigL += np.transpose(outs[1], (2, 0, 1))
            if self.order == 3:
                SigR += np.transpose(outs[3], (2, 0, 1))
                SigL += np.transpose(outs[4], (2, 0, 1))
        if self.order >= 1:
            SigR[:, np.arange(self.N), np.arange(self.N)] -= 1j * HBAR * self.U * glt0
        W = self.embedding()[None] - SigR + HBAR * self.omega[:, None, None] * np.eye(self.N)
        R = np.linalg.inv(W)
        A = np.conj(np.transpose(R, (0, 2, 1)))
        L = R @ (self.lesser_in(V) + SigL) @ A
        Ro, Lo = np.transpose(R, (1, 2, 0)), np.transpose(L, (1, 2, 0))
        self.G_in = (GR, GL, GG, glt0)
        # check_gf_sanity (common/scf_checks.f90:30-57)
        nocc = np.real(-1j * HBAR * self.pp * np.einsum("iik->i", Lo))
        wsum = -HBAR * (self.delta / np.pi) * np.sum(np.imag(np.einsum("iik->ik", Ro)), axis=1)
        runaway = (not np.all(np.isfinite(nocc))) or (not np.all(np.isfinite(wsum))) or \
            bool(np.any((nocc < -1.0) | (nocc > 2.0)))
        self.unphysical = bool(np.any((nocc < -0.05) | (nocc > 1.05) | (wsum < 0.5)))
        if runaway:
            return np.zeros(self.n_vec), True
        return self.pack(Ro, Lo) - x, False

    def solve(self, V, x):
        CB = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_double),
                              ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int32))

        def res(ctx, n, xp, Fp, ab):
            F, a = self.residual(V, np.ctypeslib.as_array(xp, (n,)).copy())
            np.ctypeslib.as_array(Fp, (n,))[:] = F
            ab[0] = int(a)

        cb = CB(res)
        L = oracle_lib()
        L.orc_broyden.argtypes = [CB, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_double,
                                  ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_void_p,
                                  ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        conv, nit, fmax = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_double()
        L.orc_broyden(cb, None, ptr(x), self.n_vec, self.alpha, self.nhist, self.maxscf, self.tolf, 0,
                      ctypes.byref(conv), ctypes.byref(nit), ctypes.byref(fmax), None, None)
        return bool(conv.value) and not self.unphysical, nit.value

    def current(self, V):
        _, GL, GG, _ = self.G_in
        f = self.fermi(self.omega, V)
        L, G = np.transpose(GL, (2, 0, 1)), np.transpose(GG, (2, 0, 1))
        M = 1j * (self.GamL[None] @ ((f - 1.0)[:, None, None] * L - f[:, None, None] * G))
        trns = np.real(np.einsum("kii->k", M)) / (2 * np.pi)
        return 2.0 * np.sum(trns) * (self.delta / HBAR)

    def run(self):
        """Volt_Current rows (V, J, converged) and populations rows (V, i, n, 2n)"""
        iv, pops = [], []
        fresh, x = True, None
        for V in self.volts:
            if fresh:
                x = self.pack(*self.noninteracting(V))
            else:
                x = self.pack(self.G_in[0], self.G_in[1])
            conv, _ = self.solve(V, x)
            fresh = (not self.continuation) or (not conv)
            iv.append([V, self.current(V), 1.0 if conv else 0.0])
            n = np.real(-1j * HBAR * self.G_in[3])
            pops += [[V, i + 1, n[i], 2 * n[i]] for i in range(self.N)]
        return iv, pops
