"""The entry point that allocates nothing (qs_sancho_self_energies_static)
against the handle-based entry point it varies (qs_sancho_self_energies with
device arrays): same outputs bit for bit, on two calls in a row, the second
with a smaller input, and no change in free device memory across the calls."""
import importlib

import numpy as np
import pytest

from qs_testlib import sancho_inputs

pytestmark = pytest.mark.gpu


def pkg():
    return importlib.import_module("quichotte-and-sancho_amd")


def run_both(na, nw, order, seed):
    P = pkg()
    GR, GL, U, delta = sancho_inputs(na, nw, seed=seed)
    nbytes = 16 * na * na * nw
    gr, gl = P.DeviceBuffer(nbytes), P.DeviceBuffer(nbytes)
    gr.upload(GR.ravel(order="F"))
    gl.upload(GL.ravel(order="F"))
    n_out = 6 if order == 3 else 3
    ref = [P.DeviceBuffer(nbytes) for _ in range(n_out)]
    sta = [P.DeviceBuffer(nbytes) for _ in range(n_out)]
    s = P.SanchoSelfEnergies(na, nw, order)
    s.build_device(gr, gl, ref, U, delta)
    del s
    free0, _ = P.device_memory(0)
    P.sancho_self_energies_static(na, nw, order, gr, gl, sta, U, delta)
    free1, _ = P.device_memory(0)
    out = []
    for a, b in zip(ref, sta):
        x = np.zeros(na * na * nw, complex)
        y = np.zeros(na * na * nw, complex)
        a.download(x)
        b.download(y)
        out.append((x, y))
    return out, free0 - free1


def test_static_entry_matches_handle_entry_twice():
    # the first call loads nothing new; the second is smaller and must not see the first's leftovers
    L = pkg().load_library()
    for k, (na, nw, order, seed) in enumerate([(6, 300, 3, 1), (3, 97, 2, 2)]):
        outs, used = run_both(na, nw, order, seed)
        for x, y in outs:
            assert np.array_equal(x, y), (na, nw, order, np.abs(x - y).max())
            assert np.abs(x).max() > 0
        # the work arrays are not allocated: free device memory does not drop by them
        # (the first call may still see the runtime set up its default stream)
        need = L.qs_sancho_static_need(na, nw, order)
        assert used < need if k == 0 else used == 0, (k, used, need)


def test_static_entry_rejects_sizes_beyond_its_buffer():
    P = pkg()
    L = P.load_library()
    cap = L.qs_sancho_static_bytes()
    assert cap == 64 << 20
    need = L.qs_sancho_static_need(64, 2000, 3)
    assert need > cap
    with pytest.raises(P.QsError) as e:
        P.sancho_self_energies_static(64, 2000, 3, 1, 2, [3, 4, 5, 6, 7, 8], np.ones(64), 0.01)
    assert e.value.code == 5   # QS_ERR_UNSUPPORTED, before any GPU work touches those addresses
    assert 0 < L.qs_sancho_static_need(6, 300, 3) <= cap
