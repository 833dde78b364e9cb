"""Full-size parity of the 15-mode heat-bath workload against the oracle's
CPU runs (tests/golden/c1_oracle.json: 20 minutes without PT2, 48 minutes
with it): the basis size after every heat-bath iteration, the final basis
(checksum over the quanta in order), the 100 variational energies, the PT2
external space and the 100 PT2-corrected energies; the mixed form must
select the identical basis and agree within 0.01 cm^-1."""
import json
import os

import numpy as np
import pytest

from helpers import GOLDEN

import vhci_amd
from vhci_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ref():
    with open(os.path.join(GOLDEN, "c1_oracle.json")) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def fp64_run():
    p = vhci_amd.Problem(text=synth.config_input(1, pt2=False))
    ctx = vhci_amd.Context(p, vhci_amd.FP64)
    r = ctx.run()
    r["sizes"] = ctx.iteration_sizes()
    r["C"] = ctx.vectors(p.neig)
    r["assign"] = ctx.freq_assignments(p.neig)
    ctx.close()
    return r


def test_c1_fp64_matches_oracle(ref, fp64_run):
    print("c1 fp64 sizes", fp64_run["sizes"], flush=True)
    assert fp64_run["sizes"] == ref["sizes"]
    assert fp64_run["checksum"] == int(ref["checksum"])
    err = np.max(np.abs(fp64_run["E_var"] - np.array(ref["E_var"])))
    print("max |E - oracle| = %.3e cm^-1" % err, flush=True)
    assert err < 1e-6


def test_c1_mixed_matches_fp64(ref, fp64_run):
    p = vhci_amd.Problem(text=synth.config_input(1, pt2=False))
    ctx = vhci_amd.Context(p, vhci_amd.MIXED)
    r = ctx.run()
    assert ctx.iteration_sizes() == ref["sizes"]
    assert r["checksum"] == fp64_run["checksum"]
    assert np.max(np.abs(r["E_var"] - fp64_run["E_var"])) < 1e-2


@pytest.mark.parametrize("form", [vhci_amd.FP64, vhci_amd.MIXED])
def test_c1_pt2_matches_oracle(ref, form):
    """The whole workload with PT2 (eps2 0.01): the external space of the
    oracle's run (5,381,485 states) and its 100 corrected energies, within
    1e-6 cm^-1 (FP64) or 0.01 cm^-1 (mixed), with the oracle's basis."""
    p = vhci_amd.Problem(text=synth.config_input(1))
    ctx = vhci_amd.Context(p, form)
    r = ctx.run()
    sizes = ctx.iteration_sizes()
    ctx.close()
    err = float(np.max(np.abs(r["E"] - np.array(ref["E_pt2"]))))
    print("c1 PT2 form %d: external %d, max |E - oracle| = %.3e cm^-1" % (form, r["ns2"], err), flush=True)
    assert sizes == ref["sizes"]
    assert r["checksum"] == int(ref["checksum"])
    assert r["ns2"] == ref["pt2_space"]
    assert err < (1e-6 if form == vhci_amd.FP64 else 1e-2)


def test_c1_freq_assignments_match_full_scan(fp64_run):
    """PrintFreqs' assignment scan on the device (rows with |C| >= max|C|/2,
    ascending, Ham.cpp:448-502) equals the same scan over the downloaded
    eigenvectors, row for row and bit for bit."""
    C = fp64_run["C"]
    for c, (rows, vals) in enumerate(fp64_run["assign"]):
        col = np.abs(C[:, c])
        want = np.nonzero(col >= col.max() / 2)[0]
        assert np.array_equal(rows, want), c
        assert np.array_equal(vals, col[want]), c


def test_c1_matrix_free_matches_oracle(ref, fp64_run, monkeypatch):
    """Products that regenerate H from the basis hash instead of reading a
    stored matrix (the form for bases whose CSR does not fit): same
    per-iteration sizes and basis as the oracle, energies within 1e-6."""
    monkeypatch.setenv("VHCI_MATRIX_FREE", "1")
    p = vhci_amd.Problem(text=synth.config_input(1, pt2=False))
    ctx = vhci_amd.Context(p, vhci_amd.FP64)
    r = ctx.run()
    assert ctx.matrix_free
    sizes = ctx.iteration_sizes()
    ctx.close()
    print("c1 matrix-free sizes", sizes, r["timings"], flush=True)
    assert sizes == ref["sizes"]
    assert r["checksum"] == int(ref["checksum"])
    assert np.max(np.abs(r["E_var"] - np.array(ref["E_var"]))) < 1e-6
    assert np.max(np.abs(r["E_var"] - fp64_run["E_var"])) < 1e-6


def test_device_orthonormalisation_repeat_path(monkeypatch):
    """Lanczos steps are orthonormalised on the device and only a step the
    device declines is repeated on the host.  With every step declined
    (VHCI_HOST_ORTHO=repeat) the solve must equal the host-only solve
    (VHCI_HOST_ORTHO=1) bit for bit, and the device form must agree with both
    within the eigensolver's tolerance and select the same basis."""
    out = {}
    for mode in ("1", "repeat", "0"):
        monkeypatch.setenv("VHCI_HOST_ORTHO", mode)
        ctx = vhci_amd.Context(vhci_amd.Problem(text=synth.config_input(1, pt2=False)), vhci_amd.FP64)
        out[mode] = ctx.run()
        ctx.close()
    assert out["repeat"]["checksum"] == out["1"]["checksum"] == out["0"]["checksum"]
    assert np.array_equal(out["repeat"]["E_var"], out["1"]["E_var"])
    assert np.max(np.abs(out["0"]["E_var"] - out["1"]["E_var"])) < 1e-6


@pytest.mark.parametrize("form", [vhci_amd.FP64, vhci_amd.MIXED])
def test_graph_captured_cycles_equal_eager(monkeypatch, form):
    """VHCI_GRAPH=1 captures a cycle's device Lanczos steps once in a HIP
    graph and replays it in the cycles of the same shape (the mixed form's
    FP32 and FP64 cycles each get their own); the solve must equal the eager
    one bit for bit."""
    out = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("VHCI_GRAPH", mode)
        ctx = vhci_amd.Context(vhci_amd.Problem(text=synth.config_input(1, pt2=False)), form)
        out[mode] = ctx.run()
        ctx.close()
    assert out["1"]["checksum"] == out["0"]["checksum"]
    assert np.array_equal(out["1"]["E_var"], out["0"]["E_var"])
