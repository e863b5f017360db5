"""The C ABI on the host, without a GPU: the built library loads and exports
every function include/qs_hip.h declares, and the host-only queries answer
(no compute call is made)."""
import ctypes
import os
import re

import pytest

from qs_testlib import REPO, random_problem

HEADER = os.path.join(REPO, "include", "qs_hip.h")
SO = os.path.join(REPO, "quichotte-and-sancho_amd", "lib", "libqs_hip.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)  # comments
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*(qs_[a-z0-9_]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(SO):
        pytest.fail("libqs_hip.so is not built; run __graft_entry__.build()")
    return ctypes.CDLL(SO)


def test_header_declares_the_abi():
    names = declared_functions()
    for must in ("qs_create", "qs_destroy", "qs_solve_scf", "qs_residual2", "qs_init2", "qs_get_sigma",
                 "qs_final_sweep", "qs_footprint_bytes"):
        assert must in names
    assert len(names) >= 15


def test_library_exports_every_declared_function(lib):
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_stage_names_and_invalid_problem(lib):
    lib.qs_stage_name.restype = ctypes.c_char_p
    lib.qs_stage_name.argtypes = [ctypes.c_int32]
    names = [lib.qs_stage_name(i) for i in range(10)]
    assert b"broyden_update" in names and b"sigma2_fft_mid" in names
    lib.qs_footprint_bytes.restype = ctypes.c_int64
    lib.qs_footprint_bytes.argtypes = [ctypes.c_void_p]
    assert lib.qs_footprint_bytes(None) == -1


def test_footprint_counts_the_update_difference_buffer():
    """Order 2 with FP32 history keeps x_new - x (n_vec doubles) for the history
    pass: relative to no history, nh FP32 columns add 2 nh 4 n_vec bytes plus
    that buffer (8 n_vec); FP64 columns add 2 nh 8 n_vec and no buffer."""
    import importlib
    pkg = importlib.import_module("quichotte-and-sancho_amd")
    prob = random_problem(16, 64, seed=4)

    def footprint(nhist, single):
        kw = dict(H=prob.H, gammaL=prob.GL, gammaR=prob.GR, U_site=prob.U_site, order=2, n_omega=prob.n_omega,
                  w_init=prob.w_init, delta=prob.delta, beta=prob.beta, mu=prob.mu, alpha_mix=prob.alpha_mix,
                  nhist=nhist, maxscf=prob.maxscf, tolf=prob.tolf, tolf_order1=prob.tolf_order1,
                  single_history=single)
        return pkg.footprint_bytes(pkg.ScfProblem(**kw))

    base = footprint(0, True)
    d4 = footprint(4, True) - base
    assert d4 > 0 and d4 % 40 == 0
    n_vec = d4 // 40  # 2 * 4 * 4 + 8 bytes per element
    assert n_vec % (3 * prob.n_omega) == 0  # 3 nA^2 per energy
    assert footprint(8, True) - base == (2 * 8 * 4 + 8) * n_vec
    assert footprint(4, False) - base == 2 * 4 * 8 * n_vec
    assert footprint(0, False) == base


def test_footprint_counts_the_atom_blocks():
    """Multi-orbital atoms (dftb) keep their intra-atom blocks of G^< per energy
    (gather_atoms, src/green_functions.f90:240-257): 16 bytes for each of the
    sum over atoms of (2 m)^2 entries, per energy, on top of one-orbital atoms."""
    import importlib

    import numpy as np
    pkg = importlib.import_module("quichotte-and-sancho_amd")
    prob = random_problem(16, 64, seed=4)

    def footprint(atom_norb):
        kw = dict(H=prob.H, gammaL=prob.GL, gammaR=prob.GR, U_site=prob.U_site, order=2, n_omega=prob.n_omega,
                  w_init=prob.w_init, delta=prob.delta, beta=prob.beta, mu=prob.mu, nhist=4,
                  atom_norb=atom_norb)
        return pkg.footprint_bytes(pkg.ScfProblem(**kw))

    base = footprint(None)
    atoms = np.array([4, 4], dtype=np.int32)          # 8 sites as two 4-orbital atoms
    assert footprint(atoms) - base == 16 * prob.n_omega * 2 * (2 * 4) ** 2
    assert footprint(np.array([1] * 8, dtype=np.int32)) == base
    import pytest
    with pytest.raises(pkg.QsError):
        footprint(np.array([4, 3], dtype=np.int32))   # does not add up to the 8 sites


def test_sancho_create_rejects_bad_sizes_before_any_device_call(lib):
    """argument checks come first: no HIP call is made for these"""
    P = ctypes.c_void_p
    lib.qs_sancho_create.argtypes = [ctypes.c_int32] * 4 + [ctypes.POINTER(P)]
    lib.qs_sancho_last_error.restype = ctypes.c_char_p
    lib.qs_sancho_last_error.argtypes = [P]
    lib.qs_sancho_device_bytes.restype = ctypes.c_int64
    lib.qs_sancho_device_bytes.argtypes = [P]
    h = P(1)
    for na, nw, order in [(0, 10, 2), (4, 0, 3), (4, 10, 1), (4, 10, 4), (256, 10, 2)]:
        assert lib.qs_sancho_create(na, nw, order, 0, ctypes.byref(h)) == 2
        assert not h.value
    assert lib.qs_sancho_last_error(None) == b"null handle"
    assert lib.qs_sancho_device_bytes(None) == -1
