"""The 60-mode sextic stress field (37,820 cubic, 43,722 quartic, 7,140
quintic and 8,910 sextic constants, up to 6th order) at a reduced selection
threshold (eps1 4, Nstate 20, PT2 eps2 0.5), small enough to check in
seconds:
  - 200 rows of the device H (seeded sample), quintic and sextic terms
    included, against the oracle's hash-based rows over the final basis,
    bit for bit (the hash-based rows equal the all-pairs rows bit for bit,
    tests/test_oracle.py);
  - every Ritz pair's residual from an explicit product within the
    eigensolver tolerance;
  - the mixed form: the same basis and every energy within 0.01 cm^-1;
  - PT2 summed in many hash-tag chunks equals the one-chunk sum."""
import numpy as np
import pytest

import vhci_amd
from vhci_amd import synth

pyoracle = pytest.importorskip("oracle.pyoracle")

pytestmark = pytest.mark.gpu

PARAMS = dict(eps1=4.0, nstate=20, eps2=0.5)


@pytest.fixture(scope="module")
def c4(tmp_path_factory):
    text = synth.config_input(4, **PARAMS)
    ctx = vhci_amd.Context(vhci_amd.Problem(text=text), vhci_amd.FP64)
    r = ctx.run()
    r["sizes"] = ctx.iteration_sizes()
    r["phase"] = ctx.phase_stats()
    print("c4 reduced fp64: sizes", r["sizes"], "PT2 space", r["ns2"], "timings", r["timings"], flush=True)
    path = str(tmp_path_factory.mktemp("c4") / "c4.inp")
    with open(path, "w") as f:
        f.write(text)
    yield ctx, r, path, text
    ctx.close()


def test_c4_basis_grows_and_converges(c4):
    ctx, r, _, _ = c4
    assert r["sizes"][0] == 1891  # C(62, 2): Ntot 2 over 60 modes
    assert r["sizes"][-1] > r["sizes"][0]
    assert r["ns2"] > 0


def test_c4_sampled_rows_match_oracle(c4):
    ctx, r, path, _ = c4
    q = ctx.basis()
    n = q.shape[0]
    rows = np.sort(np.random.default_rng(4).choice(n, size=min(200, n), replace=False))
    rp, col, val, diag = ctx.get_h_rows(rows)
    o = pyoracle.OProblem(path)
    pyoracle.set_threads(16)
    orp, ocol, oval = o.fast_h_rows(q, rows)
    entries = 0
    for k, i in enumerate(rows):
        oc, ov = ocol[orp[k]:orp[k + 1]], oval[orp[k]:orp[k + 1]]
        dc, dv = col[rp[k]:rp[k + 1]], val[rp[k]:rp[k + 1]]
        isd = oc == i
        assert isd.sum() == 1 and ov[isd][0] == diag[k], i
        assert np.array_equal(oc[~isd], dc), i
        assert np.array_equal(ov[~isd], dv), i
        entries += len(dc)
    print("c4: %d sampled rows, %d entries equal to the oracle's" % (len(rows), entries), flush=True)
    assert entries > len(rows)


def test_c4_ritz_residuals(c4):
    ctx, r, _, _ = c4
    E = r["E_var"]
    C = ctx.vectors(len(E))
    Y = ctx.apply_h(np.ascontiguousarray(C))
    worst = float(np.max(np.linalg.norm(Y - C * E, axis=0) / np.abs(E)))
    print("c4: largest relative Ritz residual %.3e" % worst, flush=True)
    assert worst < 2e-10


def test_c4_mixed_and_chunked_pt2(c4, monkeypatch):
    ctx, r, _, text = c4
    m = vhci_amd.Context(vhci_amd.Problem(text=text), vhci_amd.MIXED)
    rm = m.run()
    m.close()
    assert rm["checksum"] == r["checksum"] and rm["ns2"] == r["ns2"]
    assert float(np.max(np.abs(rm["E"] - r["E"]))) <= 0.01
    pairs = int(r["phase"]["pairs"])
    monkeypatch.setenv("VHCI_PT2_CHUNK_PAIRS", str(max(1, pairs // 6)))
    c = vhci_amd.Context(vhci_amd.Problem(text=text), vhci_amd.FP64)
    rc = c.run()
    chunks = c.phase_stats()["pt2_chunks"]
    c.close()
    print("c4: PT2 in %d chunks, max |dE| against one chunk %.3e" % (chunks, np.max(np.abs(rc["E"] - r["E"]))),
          flush=True)
    assert chunks > 1 and rc["ns2"] == r["ns2"] and rc["checksum"] == r["checksum"]
    assert np.max(np.abs(rc["E"] - r["E"])) < 1e-9 * np.max(np.abs(r["E"]))


def test_c4_half_storage(c4, monkeypatch):
    """Half storage (upper triangle, transposed products through the
    fixed-point accumulator) selects the same basis and gives the same
    energies within 1e-6 cm^-1."""
    ctx, r, _, text = c4
    monkeypatch.setenv("VHCI_H_STORAGE", "half")
    c = vhci_amd.Context(vhci_amd.Problem(text=text), vhci_amd.FP64)
    rs = c.run()
    c.close()
    err = float(np.max(np.abs(rs["E"] - r["E"])))
    print("c4 half storage: timings %s, max |dE| %.3e cm^-1" % (rs["timings"], err), flush=True)
    assert rs["checksum"] == r["checksum"] and rs["ns2"] == r["ns2"]
    assert err < 1e-6


def test_c4_matrix_free(monkeypatch):
    """Matrix-free products (every product regenerates the quintic and
    sextic terms from the hash) against the stored form, at a higher
    threshold (eps1 10) that keeps the matrix-free solve short: the same
    basis and energies within 1e-6 cm^-1."""
    text = synth.config_input(4, eps1=10.0, nstate=20, eps2=0.5)
    out = {}
    for storage in ("full", "free"):
        monkeypatch.setenv("VHCI_H_STORAGE", storage)
        c = vhci_amd.Context(vhci_amd.Problem(text=text), vhci_amd.FP64)
        out[storage] = c.run()
        out[storage]["sizes"] = c.iteration_sizes()
        c.close()
    err = float(np.max(np.abs(out["free"]["E"] - out["full"]["E"])))
    print("c4 matrix-free (eps1 10): sizes %s, timings %s, max |dE| %.3e cm^-1"
          % (out["free"]["sizes"], out["free"]["timings"], err), flush=True)
    assert out["free"]["checksum"] == out["full"]["checksum"] and out["free"]["ns2"] == out["full"]["ns2"]
    assert err < 1e-6
