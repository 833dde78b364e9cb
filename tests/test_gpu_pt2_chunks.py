"""PT2 over the external space in chunks (ranges of the states' key hash):
forced small budgets split the space into many chunks; the corrections
match the one-chunk sum and the oracle, and the external space is the same
size."""
import os

import numpy as np
import pytest

from helpers import GOLDEN

import vhci_amd
from vhci_amd import synth

pyoracle = pytest.importorskip("oracle.pyoracle")

pytestmark = pytest.mark.gpu


def _clo2_pt2(form, env, monkeypatch):
    path = os.path.join(GOLDEN, "clo2_pt2.inp")
    o = pyoracle.OProblem(path)
    q = o.initial_basis()
    w, Z = pyoracle.eigh(o.build_h_dense(q))
    nopt = 20
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    ctx = vhci_amd.Context(vhci_amd.Problem(path), form)
    ctx.set_vectors(Z[:, :nopt])
    dE, ns2 = ctx.pt2(w, nopt, 0.0)
    ph = ctx.phase_stats()
    ctx.close()
    for k in env:
        monkeypatch.delenv(k)
    return dE, ns2, ph, (o, q, Z, w, nopt)


@pytest.mark.parametrize("form", [vhci_amd.FP64, vhci_amd.MIXED])
@pytest.mark.parametrize("env", [{"VHCI_PT2_CHUNK_STATES": "100"}, {"VHCI_PT2_CHUNK_PAIRS": "2000"},
                                 {"VHCI_PT2_CHUNK_STATES": "64", "VHCI_PT2_CHUNK_PAIRS": "700", "VHCI_PT2_PULL": "1"}])
def test_clo2_chunked_matches_one_chunk_and_oracle(form, env, monkeypatch):
    dE1, ns1, ph1, (o, q, Z, w, nopt) = _clo2_pt2(form, {}, monkeypatch)
    dEc, nsc, phc, _ = _clo2_pt2(form, env, monkeypatch)
    print("ClO2 chunks", ph1["pt2_chunks"], "->", phc["pt2_chunks"], "external", ns1, nsc, flush=True)
    assert ph1["pt2_chunks"] == 1
    assert phc["pt2_chunks"] > 1
    assert nsc == ns1
    # the pull form sums each numerator in pattern order, the push form in
    # ascending i: in FP32 (mixed) that changes the rounding
    rel = 1e-5 if (form == vhci_amd.MIXED and "VHCI_PT2_PULL" in env) else 1e-12
    assert np.max(np.abs(dEc - dE1)) <= rel * np.max(np.abs(dE1))
    dE_ref, ns_ref = o.pt2(q, Z[:, :nopt], w[:nopt], nopt, 0.0)
    assert nsc == ns_ref
    tol = 1e-9 if form == vhci_amd.FP64 else 1e-4
    assert np.max(np.abs(dEc - dE_ref)) < tol


def test_c1_chunked_matches_one_chunk(monkeypatch):
    """The 15-mode workload's PT2 (5.4 M external states): 2^20 states and
    2^26 pairs per chunk against one chunk; same space, corrections within
    1e-9 cm^-1."""
    p = vhci_amd.Problem(text=synth.config_input(1))
    ctx = vhci_amd.Context(p, vhci_amd.FP64)
    r1 = ctx.run()
    ph1 = ctx.phase_stats()
    E = r1["E_var"]
    nopt = len(E)
    monkeypatch.setenv("VHCI_PT2_CHUNK_STATES", str(1 << 20))
    monkeypatch.setenv("VHCI_PT2_CHUNK_PAIRS", str(1 << 26))
    dEc, nsc = ctx.pt2(E, nopt, p.eps2)
    phc = ctx.phase_stats()
    ctx.close()
    print("c1 PT2 chunks", ph1["pt2_chunks"], "->", phc["pt2_chunks"], "external", r1["ns2"], nsc,
          "pairs", ph1["pairs"], phc["pairs"], flush=True)
    assert ph1["pt2_chunks"] == 1 and phc["pt2_chunks"] >= 4
    assert nsc == r1["ns2"]
    assert phc["pairs"] == ph1["pairs"]
    assert np.max(np.abs((E + dEc) - r1["E"])) < 1e-9


@pytest.mark.parametrize("form", [vhci_amd.FP64, vhci_amd.MIXED])
def test_c1_one_pass_matches_counted(form, monkeypatch):
    """PT2 pairs written in one pass (hits queued per wave, sorted by
    (a, i)) give the counted form's sorted pairs, so the corrections are
    the same bits; also with a pair budget that forces several chunks."""
    p = vhci_amd.Problem(text=synth.config_input(1))
    ctx = vhci_amd.Context(p, form)
    r1 = ctx.run()
    ph1 = ctx.phase_stats()
    E = r1["E_var"]
    nopt = len(E)
    monkeypatch.setenv("VHCI_PT2_COUNTED", "1")
    dEc, nsc = ctx.pt2(E, nopt, p.eps2)
    phc = ctx.phase_stats()
    monkeypatch.setenv("VHCI_PT2_CHUNK_PAIRS", str(1 << 26))
    monkeypatch.delenv("VHCI_PT2_COUNTED")
    dE4, ns4 = ctx.pt2(E, nopt, p.eps2)
    monkeypatch.setenv("VHCI_PT2_COUNTED", "1")
    dE4c, ns4c = ctx.pt2(E, nopt, p.eps2)
    ctx.close()
    print("c1 PT2 one pass vs counted: pairs", ph1["pairs"], phc["pairs"], flush=True)
    assert nsc == r1["ns2"] and ns4 == ns4c == r1["ns2"]
    assert phc["pairs"] == ph1["pairs"]
    assert np.array_equal(E + dEc, r1["E"])
    assert np.array_equal(dE4, dE4c)
