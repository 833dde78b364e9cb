"""GPU parity of the product kernels: H X through the solver's own kernels
(vhci_apply_h) against a float64 product of the stored H copied to the host,
for every block-width class the dispatcher distinguishes, in each precision
form, in basis order and in the solver's row order, in matrix-free mode and
with only the upper triangle stored (half storage).

The error of each output element is measured against sum_j |H_ij| |X_jc|,
the scale its rounding error is proportional to:
- FP64 products (and the mixed form's hi + lo product, whose values get_h
  returns exactly as hi + lo): below 1e-12 of that scale;
- the mixed form's FP32 product (FP32 values, FP32 vectors, FP32 sums):
  below 1e-4 of it, and above 1e-9 (it must really be the FP32 product).
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp

from helpers import GOLDEN

import vhci_amd
from vhci_amd import synth

pytestmark = pytest.mark.gpu

MIX = dict(m_range=(8, 14), eps1_set=(1.0, 2.0), nstate_range=(20, 200), eps2_set=(0.0, 0.01, 0.1))
WIDTHS = (1, 2, 3, 4, 5, 8, 12, 16, 24, 40, 64)


def _say(*a):
    print(*a, flush=True)


def _stored(ctx):
    rp, col, val, diag = ctx.get_h()
    n = len(diag)
    return sp.csr_matrix((val, col, rp), shape=(n, n)), diag


def _errors(ctx, H, diag, fp32, widths=WIDTHS):
    rng = np.random.default_rng(11)
    Habs = abs(H)
    out = {}
    for b in widths:
        X = rng.standard_normal((H.shape[0], b))
        want = H @ X + diag[:, None] * X
        scale = Habs @ np.abs(X) + np.abs(diag)[:, None] * np.abs(X)
        got = ctx.apply_h(X if b > 1 else X[:, 0], fp32=fp32).reshape(-1, b)
        out[b] = float(np.max(np.abs(got - want) / scale))
        _say("  b=%d fp32=%s max scaled error %.2e" % (b, fp32, out[b]))
    return out


@pytest.fixture(scope="module")
def grown_text():
    # a request that grows past 16384 states, so its final H is stored in
    # the solver's row order
    return synth.request_input(5, **MIX)


@pytest.mark.parametrize("form", [vhci_amd.FP64, vhci_amd.MIXED])
def test_products_solver_order(grown_text, form):
    _say("products, serving-mix request 5, form", form)
    ctx = vhci_amd.Context(vhci_amd.Problem(text=grown_text), form)
    ctx.run()
    assert ctx.n >= 16384
    H, diag = _stored(ctx)
    for b, e in _errors(ctx, H, diag, fp32=False).items():
        assert e < 1e-12, (b, e)
    if form == vhci_amd.MIXED:
        for b, e in _errors(ctx, H, diag, fp32=True).items():
            assert 1e-9 < e < 1e-4, (b, e)
    ctx.close()


def test_products_basis_order():
    _say("products, ClO2 VCI (basis order)")
    ctx = vhci_amd.Context(vhci_amd.Problem(os.path.join(GOLDEN, "clo2_vci.inp")))
    ctx.run()
    assert ctx.n < 16384
    H, diag = _stored(ctx)
    for b, e in _errors(ctx, H, diag, fp32=False).items():
        assert e < 1e-12, (b, e)
    ctx.close()


def test_products_matrix_free(grown_text, monkeypatch):
    _say("products, serving-mix request 5, matrix-free against stored")
    p = vhci_amd.Problem(text=grown_text)
    ref = vhci_amd.Context(p)
    r0 = ref.run()
    H, diag = _stored(ref)
    ref.close()
    monkeypatch.setenv("VHCI_MATRIX_FREE", "1")
    ctx = vhci_amd.Context(p)
    r1 = ctx.run()
    assert ctx.matrix_free and r1["checksum"] == r0["checksum"]
    for b, e in _errors(ctx, H, diag, fp32=False, widths=(1, 3, 8, 16, 40)).items():
        assert e < 1e-12, (b, e)
    ctx.close()
    ctx = vhci_amd.Context(p, vhci_amd.MIXED)
    ctx.run()
    assert ctx.matrix_free
    with pytest.raises(RuntimeError, match="FP64 only"):
        ctx.apply_h(np.zeros((ctx.n, 2)), fp32=True)
    ctx.close()


def test_apply_h_rejects_bad_arguments():
    ctx = vhci_amd.Context(vhci_amd.Problem(os.path.join(GOLDEN, "clo2_vci.inp")))
    ctx.run()
    with pytest.raises(RuntimeError, match="block width"):
        ctx.apply_h(np.zeros((ctx.n, 65)))
    with pytest.raises(RuntimeError, match="mixed form only"):
        ctx.apply_h(np.zeros((ctx.n, 2)), fp32=True)
    ctx.close()


@pytest.mark.parametrize("form", [vhci_amd.FP64, vhci_amd.MIXED])
def test_products_half_storage(grown_text, form, monkeypatch):
    """Upper triangle only: the same matrix (get_h and get_h_rows mirror it,
    bit for bit equal to the full CSR's), the same basis, products to the
    bars above, and every product the same bit for bit when repeated (the
    transposed terms are summed in integer fixed point)."""
    _say("products, serving-mix request 5, half storage against full, form", form)
    p = vhci_amd.Problem(text=grown_text)
    monkeypatch.setenv("VHCI_H_STORAGE", "full")
    ref = vhci_amd.Context(p, form)
    r0 = ref.run()
    assert ref.phase_stats()["h_storage"] == 0
    H, diag = _stored(ref)
    rows = np.random.default_rng(5).choice(ref.n, 300, replace=False)
    rows_ref = ref.get_h_rows(rows)
    ref.close()
    monkeypatch.setenv("VHCI_H_STORAGE", "half")
    ctx = vhci_amd.Context(p, form)
    r1 = ctx.run()
    ph = ctx.phase_stats()
    assert ph["h_storage"] == 1 and ph["nnz"] == H.nnz
    assert r1["checksum"] == r0["checksum"] and r1["nvar"] == r0["nvar"]
    assert np.max(np.abs(r1["E"] - r0["E"])) < 1e-7
    H2, diag2 = _stored(ctx)
    assert (H2 != H).nnz == 0 and np.array_equal(diag2, diag)
    for a, b in zip(ctx.get_h_rows(rows), rows_ref):
        assert np.array_equal(a, b)
    for b, e in _errors(ctx, H, diag, fp32=False, widths=(1, 3, 4, 8, 16, 24, 40)).items():
        assert e < 1e-12, (b, e)
    if form == vhci_amd.MIXED:
        for b, e in _errors(ctx, H, diag, fp32=True, widths=(1, 8, 16)).items():
            assert 1e-9 < e < 1e-4, (b, e)
    X = np.random.default_rng(3).standard_normal((ctx.n, 8))
    y1 = ctx.apply_h(X)
    y2 = ctx.apply_h(X)
    assert np.array_equal(y1, y2)
    ctx.close()


def test_layout_knobs_leave_results_unchanged(monkeypatch):
    """The performance choices change where work runs, not what it computes:
    XCD row runs in the products (VHCI_XCD_RUN_ROWS=0 against the default),
    the count pass's hit words (VHCI_HIT_WORDS=0) and the Bloom filter size
    (VHCI_BLOOM_BITS=2) must give the same H, products, basis and energies
    bit for bit on the 15-mode workload."""
    text = synth.config_input(1, pt2=False)
    runs = {}
    for name, env in (("default", {}), ("plain", {"VHCI_XCD_RUN_ROWS": "0", "VHCI_HIT_WORDS": "0",
                                                   "VHCI_BLOOM_BITS": "2"})):
        for k in ("VHCI_XCD_RUN_ROWS", "VHCI_HIT_WORDS", "VHCI_BLOOM_BITS"):
            monkeypatch.delenv(k, raising=False)
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        ctx = vhci_amd.Context(vhci_amd.Problem(text=text), vhci_amd.FP64)
        r = ctx.run()
        rp, col, val, diag = ctx.get_h()
        X = np.random.default_rng(3).standard_normal((ctx.n, 8))
        runs[name] = (r, rp, col, val, diag, ctx.apply_h(X), ctx.apply_h(X[:, 0]))
        ctx.close()
    a, b = runs["default"], runs["plain"]
    assert a[0]["checksum"] == b[0]["checksum"]
    assert np.array_equal(a[0]["E_var"], b[0]["E_var"])
    for x, y in zip(a[1:], b[1:]):
        assert np.array_equal(x, y)
