"""SANCHO's order-2/3 self-energies on the GPU (k_sn_* kernels of libqs_hip.so)
against the oracle (oracle/sancho_oracle.c, SANCHO/src/self_energies.f90:38-207).

Floating point: |hip - oracle| <= 1e-12 max|oracle| per output array (the sums
run in another order than the reference's loops).
"""
import numpy as np
import pytest

from qs_testlib import product, sancho_inputs, sancho_oracle

TOL = 1e-12
NAMES = ["Sig2R", "Sig2L", "Sig2G", "Sig3R", "Sig3L", "Sig3G"]


def check(outs, ref, case):
    for q, (o, r) in enumerate(zip(outs, ref)):
        err = np.max(np.abs(o - r)) / max(np.max(np.abs(r)), 1e-300)
        assert err < TOL, "%s %s rel err %.3e" % (case, NAMES[q], err)


@pytest.mark.gpu
@pytest.mark.parametrize("order", [2, 3])
@pytest.mark.parametrize("na,nw,seed", [(1, 1, 0), (2, 7, 1), (3, 256, 2), (4, 333, 3), (5, 513, 4), (8, 600, 5),
                                        (12, 700, 6)])
def test_sancho_sigma_matches_oracle(order, na, nw, seed):
    print("sancho parity order=%d na=%d nw=%d seed=%d" % (order, na, nw, seed), flush=True)
    GR, GL, U, delta = sancho_inputs(na, nw, seed)
    ref = sancho_oracle(GR, GL, U, delta, order)
    s = product().SanchoSelfEnergies(na, nw, order)
    outs = s.build(GR, GL, U, delta)
    assert len(outs) == len(ref)
    check(outs, ref, (order, na, nw))


@pytest.mark.gpu
@pytest.mark.parametrize("order", [2, 3])
def test_sancho_device_arrays_match_host_arrays(order):
    """on_device = 1 (device memory in and out) gives the host call's bits"""
    pkg = product()
    na, nw = 4, 300
    print("sancho device/host order=%d na=%d nw=%d" % (order, na, nw), flush=True)
    GR, GL, U, delta = sancho_inputs(na, nw, 11)
    s = pkg.SanchoSelfEnergies(na, nw, order)
    host = s.build(GR, GL, U, delta)
    nbytes = 16 * na * na * nw
    gr, gl = pkg.DeviceBuffer(nbytes), pkg.DeviceBuffer(nbytes)
    gr.upload(GR.ravel(order="F"))
    gl.upload(GL.ravel(order="F"))
    outs = [pkg.DeviceBuffer(nbytes) for _ in range(s.n_outputs)]
    s.build_device(gr, gl, outs, U, delta)
    for h, d in zip(host, outs):
        back = np.empty(na * na * nw, complex)
        d.download(back)
        np.testing.assert_array_equal(h.ravel(order="F"), back)


@pytest.mark.gpu
def test_sancho_repeat_calls_and_spectral_identity():
    """a second call on the same handle with other inputs gives that input's
    result, and the outputs keep Sigma^r - Sigma^a = Sigma^> - Sigma^<"""
    na, nw = 3, 200
    print("sancho repeat na=%d nw=%d" % (na, nw), flush=True)
    s = product().SanchoSelfEnergies(na, nw, 3)
    for seed in (21, 22):
        GR, GL, U, delta = sancho_inputs(na, nw, seed)
        outs = s.build(GR, GL, U, delta)
        check(outs, sancho_oracle(GR, GL, U, delta, 3), ("repeat", seed))
        for R, L, G in (outs[0:3], outs[3:6]):
            lhs = R - np.conj(R.transpose(1, 0, 2))
            assert np.max(np.abs(lhs - (G - L))) < 1e-12 * np.max(np.abs(R))


@pytest.mark.gpu
def test_sancho_rejects_bad_arguments():
    pkg = product()
    with pytest.raises(pkg.QsError):
        pkg.SanchoSelfEnergies(4, 100, 4)
    s = pkg.SanchoSelfEnergies(2, 10, 2)
    GR, GL, U, delta = sancho_inputs(2, 10, 0)
    with pytest.raises(pkg.QsError):
        s.build(GR, GL, U, -1.0)
    with pytest.raises(ValueError):
        s.build(GR[:, :, :5], GL[:, :, :5], U, delta)


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["order2_second_order", "order3_third_order", "order2_asymmetric_leads",
                                  "legacy_natoms", "order2_custom_template"])
def test_sancho_cases_with_hip_self_energies(case):
    """SANCHO's own regression cases (SANCHO/tests/cases/) with the GPU
    self-energies inside the reference's voltage sweep (tests/sancho_scf.py):
    Volt_Current and populations within the reference's tolerance."""
    import os
    from qs_testlib import assert_close_table, read_table
    from sancho_scf import SanchoSweep
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sancho", case)
    print("sancho case %s with HIP self-energies" % case, flush=True)
    handle = {}

    def hip_sigma(GR, GL, U, delta, order):
        if "s" not in handle:
            handle["s"] = product().SanchoSelfEnergies(GR.shape[0], GR.shape[2], order)
        return handle["s"].build(GR, GL, U, delta)

    sw = SanchoSweep(os.path.join(gold, "input.dat"), hip_sigma)
    iv, pops = sw.run()
    for f in os.listdir(os.path.join(gold, "reference")):
        new = iv if f.startswith("Volt_Current") else pops
        assert_close_table(new, read_table(os.path.join(gold, "reference", f)))


@pytest.mark.gpu
def test_sancho_profile_counts_builds():
    na, nw = 4, 300
    print("sancho profile na=%d nw=%d" % (na, nw), flush=True)
    GR, GL, U, delta = sancho_inputs(na, nw, 41)
    s = product().SanchoSelfEnergies(na, nw, 3)
    s.profile(True)
    for _ in range(3):
        s.build(GR, GL, U, delta)
    prof = s.profile_read()
    assert prof["bubbles"][1] == 3 and prof["self_energies"][1] == 3
    assert prof["bubbles"][0] > 0 and prof["self_energies"][0] > 0
    s.profile(False)
    s.build(GR, GL, U, delta)
    assert s.profile_read()["bubbles"][1] == 0
