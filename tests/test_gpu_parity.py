"""GPU parity: each stage of the device path against the CPU oracle on the
same inputs, and the drop-in program against the reference's fixtures."""
import os
import re
import subprocess

import numpy as np
import pytest

from helpers import GOLDEN, absolute, manifest, read_freqs

import vhci_amd
from vhci_amd import synth

pyoracle = pytest.importorskip("oracle.pyoracle")

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def c1_iter1(tmp_path_factory):
    """Synthetic workload c1 after one heat-bath iteration (oracle)."""
    d = tmp_path_factory.mktemp("c1")
    path = str(d / "c1.inp")
    synth.config_input(1, path)
    o = pyoracle.OProblem(path)
    q0 = o.initial_basis()
    H = o.build_h_dense(q0)
    w, Z = pyoracle.eigh(H)
    nopt = min(o.neig, q0.shape[0])
    cm = pyoracle.cmax(Z, nopt)
    new = o.hb_select(q0, cm, o.eps1)
    return path, q0, cm, new


def _keys(rp, col):
    n = len(rp) - 1
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(rp))
    return rows * n + col.astype(np.int64), rows


def _compare_h(ctx, oprob, q, mixed=False):
    """Device CSR vs the oracle's all-pairs CSR: diagonal bit for bit; every
    device entry is an oracle entry with the same value (FP64) or within
    2^-46 (hi+lo of the mixed form); every nonzero oracle off-diagonal entry
    is on the device (the oracle also keeps the reference's explicit zeros)."""
    rp, col, val, diag = ctx.get_h()
    orp, ocol, oval = oprob.build_h(q)
    n = q.shape[0]
    ok, orows = _keys(orp, ocol)
    isdiag = orows == ocol
    odiag = np.zeros(n)
    odiag[orows[isdiag]] = oval[isdiag]
    assert isdiag.sum() == n
    assert np.array_equal(diag, odiag)
    ok, oval = ok[~isdiag], oval[~isdiag]
    order = np.argsort(ok)
    ok, oval = ok[order], oval[order]
    dk, drows = _keys(rp, col)
    assert not np.any(drows == col)
    assert len(np.unique(dk)) == len(dk)
    pos = np.searchsorted(ok, dk)
    assert np.all(pos < len(ok)) and np.array_equal(ok[pos], dk)
    if mixed:
        assert np.all(np.abs(val - oval[pos]) <= 2.0 ** -46 * np.abs(oval[pos]))
    else:
        bad = np.nonzero(val != oval[pos])[0]
        assert len(bad) == 0, (dk[bad[:5]], val[bad[:5]], oval[pos][bad[:5]])
    nz = ok[oval != 0.0]
    assert np.all(np.isin(nz, dk))
    return len(dk)


@pytest.mark.parametrize("form", [vhci_amd.FP64, vhci_amd.MIXED])
def test_h_build_clo2(form):
    path = os.path.join(GOLDEN, "clo2_vci.inp")
    _say("H build ClO2 VCI N=3276 form", form)
    p = vhci_amd.Problem(path)
    ctx = vhci_amd.Context(p, form)
    st = ctx.build_h()
    assert st["n"] == 3276
    nnz = _compare_h(ctx, pyoracle.OProblem(path), p.initial_basis(), mixed=form == vhci_amd.MIXED)
    assert nnz == st["nnz_offdiag"]


def test_h_build_c1_grown(c1_iter1):
    path, q0, cm, new = c1_iter1
    q = np.vstack([q0, new])
    _say("H build c1 after one HB iteration N=", q.shape[0])
    p = vhci_amd.Problem(path)
    ctx = vhci_amd.Context(p)
    ctx.set_basis(q)
    ctx.build_h()
    _compare_h(ctx, pyoracle.OProblem(path), q)
    # structural count of the reference's triplet list (explicit zeros)
    orp, _, _ = pyoracle.OProblem(path).build_h(q)
    assert ctx.structural_nnz() == orp[-1]


@pytest.mark.parametrize("form", [vhci_amd.FP64, vhci_amd.MIXED])
def test_h_build_sorted_in_slices(c1_iter1, monkeypatch, form):
    """The row sort of the H build runs in slices of rows (rocPRIM counts
    entries in 32 bits); many small slices give the same matrix as one."""
    path, q0, cm, new = c1_iter1
    q = np.vstack([q0, new])
    p = vhci_amd.Problem(path)
    ctx = vhci_amd.Context(p, form)
    ctx.set_basis(q)
    ctx.build_h()
    ref = ctx.get_h()
    monkeypatch.setenv("VHCI_SORT_SLICE", "20000")  # above any row length (8440 patterns)
    ctx.build_h()
    got = ctx.get_h()
    _say("H build in slices: nnz", len(ref[1]), "slices of 20000 entries")
    assert len(ref[1]) > 50 * 20000
    for a, b in zip(ref, got):
        assert np.array_equal(a, b)


def test_hb_select_matches_oracle_order(c1_iter1):
    """Same Cmax in, same new states out, in the same (first-discovery) order."""
    path, q0, cm, new = c1_iter1
    _say("HB select c1 initial basis")
    p = vhci_amd.Problem(path)
    ctx = vhci_amd.Context(p)
    nn = ctx.hb_select(p.eps1, cmax=cm)
    assert nn == new.shape[0]
    assert np.array_equal(ctx.new_states(nn), new)
    # ClO2 with random Cmax and a small threshold
    pc = os.path.join(GOLDEN, "clo2_vhci.inp")
    o = pyoracle.OProblem(pc)
    q = o.initial_basis()
    rng = np.random.default_rng(11)
    c = rng.standard_normal(q.shape[0]) * np.exp(-rng.uniform(0, 10, q.shape[0]))
    ref = o.hb_select(q, c, 0.003)
    ctx2 = vhci_amd.Context(vhci_amd.Problem(pc))
    n2 = ctx2.hb_select(0.003, cmax=c)
    assert np.array_equal(ctx2.new_states(n2), ref)
    # eps = 0 keeps every connected state
    ref0 = o.hb_select(q, c, 0.0)
    n3 = ctx2.hb_select(0.0, cmax=c)
    assert np.array_equal(ctx2.new_states(n3), ref0)


@pytest.mark.parametrize("form", [vhci_amd.FP64, vhci_amd.MIXED])
def test_eigsolve_c0(tmp_path, form):
    """Workload c0 (N = 1820, 50 states): Davidson vs the oracle's dense
    eigensolver on the oracle's H; residuals checked with the oracle H."""
    path = str(tmp_path / "c0.inp")
    synth.config_input(0, path)
    _say("eigsolve c0 form", form)
    p = vhci_amd.Problem(path)
    o = pyoracle.OProblem(path)
    q = p.initial_basis()
    ctx = vhci_amd.Context(p, form)
    ctx.build_h()
    E, C = ctx.eigsolve(50, tol=1e-12)
    Hd = o.build_h_dense(q)
    w, _ = pyoracle.eigh(Hd)
    assert np.max(np.abs(E - w[:50])) < 1e-8
    R = Hd @ C - C * E
    assert np.max(np.linalg.norm(R, axis=0)) < 1e-6
    assert np.allclose(C.T @ C, np.eye(50), atol=1e-10)


def test_dense_diag_matches_oracle():
    path = os.path.join(GOLDEN, "clo2_vhci.inp")
    p = vhci_amd.Problem(path)
    o = pyoracle.OProblem(path)
    ctx = vhci_amd.Context(p)
    ctx.build_h()
    E, C = ctx.dense_diag()
    w, Z = pyoracle.eigh(o.build_h_dense(p.initial_basis()))
    assert np.max(np.abs(E - w)) < 1e-9
    ov = np.abs(np.sum(C[:, :20] * Z[:, :20], axis=0))
    assert np.allclose(ov, 1.0, atol=1e-9)


@pytest.mark.parametrize("form", [vhci_amd.FP64, vhci_amd.MIXED])
def test_pt2_matches_oracle(form):
    """ClO2 PT2 (eps2 = 0, all connected states): same external space size
    and corrections as the oracle for the same vectors and energies."""
    path = os.path.join(GOLDEN, "clo2_pt2.inp")
    o = pyoracle.OProblem(path)
    q = o.initial_basis()
    H = o.build_h_dense(q)
    w, Z = pyoracle.eigh(H)
    nopt = 20
    dE_ref, ns2_ref = o.pt2(q, Z[:, :nopt], w[:nopt], nopt, 0.0)
    p = vhci_amd.Problem(path)
    ctx = vhci_amd.Context(p, form)
    ctx.set_vectors(Z[:, :nopt])
    dE, ns2 = ctx.pt2(w, nopt, 0.0)
    assert ns2 == ns2_ref
    tol = 1e-9 if form == vhci_amd.FP64 else 1e-4
    assert np.max(np.abs(dE - dE_ref)) < tol
    assert dE[0] < 0  # the ground state lies below every external state


def _run_cli(tmp_path, name, env=None):
    out = str(tmp_path / (name + ".freqs"))
    e = dict(os.environ)
    if env:
        e.update(env)
    r = subprocess.run([vhci_amd.CLI_PATH, "-n", "4", "-i", os.path.join(GOLDEN, name + ".inp"), "-o", out],
                       capture_output=True, text=True, env=e, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return out, r.stdout


@pytest.mark.parametrize("name", ["clo2_vci", "clo2_vhci", "clo2_pt2"])
def test_cli_clo2_fixtures(tmp_path, name):
    """The drop-in `vhci` on the ClO2 examples: energies within 1e-8 cm^-1 of
    the reference's freqs.dat, identical assignment strings, and the same
    basis sizes / PT2 space in the log."""
    _say("CLI", name)
    out, log = _run_cli(tmp_path, name)
    e, a = read_freqs(out)
    er, ar = read_freqs(os.path.join(GOLDEN, name + ".freqs"))
    assert np.max(np.abs(absolute(e) - absolute(er))) < 1e-8
    assert a == ar
    pins = manifest()[name]
    sizes = [int(x) for x in re.findall(r"Basis size after iteration \d+ is: (\d+)", log)]
    assert sizes == pins["iterations"]
    m = re.search(r"Perturbative space contains (\d+) states", log)
    if "pt2_space" in pins:
        assert int(m.group(1)) == pins["pt2_space"]
    # the structural "% sparse" lines (explicit zeros counted)
    pct = [float(x) for x in re.findall(r"The Hamiltonian is ([\d.]+)% sparse", log)]
    assert pct == pins["sparse_pct"]


@pytest.mark.parametrize("name", ["acn_vci", "acn_vhci", "acn_pt2"])
def test_cli_acetonitrile_fixtures(tmp_path, name):
    """The drop-in `vhci` on the acetonitrile examples (12 modes): the exact
    basis sizes 91 -> 9749 -> 22857 -> 28666 -> 29900, the exact PT2 space
    of 1,978,296 states, the "% sparse" lines, and all 70 energies within
    1e-7 cm^-1 of the reference's freqs.dat.  Assignment strings of the
    near-degenerate E pairs depend on the eigenvector rotation and are not
    compared."""
    _say("CLI", name)
    out, log = _run_cli(tmp_path, name)
    e, _ = read_freqs(out)
    er, _ = read_freqs(os.path.join(GOLDEN, name + ".freqs"))
    pins = manifest()[name]
    sizes = [int(x) for x in re.findall(r"Basis size after iteration \d+ is: (\d+)", log)]
    d = np.abs(absolute(e) - absolute(er))
    _say("sizes", sizes, "pins", pins["iterations"], "max |dE|", d.max(), "at state", int(d.argmax()),
         re.findall(r"Perturbative space contains (\d+)", log))
    assert len(e) == len(er)
    assert np.max(np.abs(absolute(e) - absolute(er))) < 1e-7
    assert sizes == pins["iterations"]
    m = re.search(r"Perturbative space contains (\d+) states", log)
    if "pt2_space" in pins:
        assert int(m.group(1)) == pins["pt2_space"]
    pct = [float(x) for x in re.findall(r"The Hamiltonian is ([\d.]+)% sparse", log)]
    assert pct == pins["sparse_pct"]


def test_cli_mixed_matches_fp64(tmp_path):
    """Mixed form on the ClO2 PT2 example: same basis sizes and PT2 space,
    every energy within 0.01 cm^-1 (observed far tighter) of the FP64 run."""
    o64, l64 = _run_cli(tmp_path, "clo2_pt2")
    e64, _ = read_freqs(o64)
    os.rename(o64, o64 + ".fp64")
    o32, l32 = _run_cli(tmp_path, "clo2_pt2", env={"VHCI_FORM": "mixed"})
    e32, _ = read_freqs(o32)
    grab = lambda s: re.findall(r"(Basis size after iteration \d+ is: \d+|Perturbative space contains \d+)", s)
    assert grab(l64) == grab(l32)
    assert np.max(np.abs(absolute(e64) - absolute(e32))) < 1e-2


def test_solve_c0_vs_oracle(tmp_path):
    """Whole solve of workload c0 on the device (both forms) against the
    oracle's energies."""
    path = str(tmp_path / "c0.inp")
    synth.config_input(0, path)
    js = str(tmp_path / "c0.json")
    pyoracle.run(path, str(tmp_path / "c0.freqs"), threads=8, json_out=js)
    import json
    ref = np.array(json.load(open(js))["E"])
    p = vhci_amd.Problem(path)
    for form in (vhci_amd.FP64, vhci_amd.MIXED):
        r = vhci_amd.solve(p, form)
        assert r["nvar"] == 1820
        assert np.max(np.abs(r["E"] - ref[:len(r["E"])])) < 1e-8


def _capped(text, caps, ntot):
    """The input with new per-mode caps and Max_quanta."""
    lines = text.splitlines()
    out, m = [], 0
    for ln in lines:
        f = ln.split()
        if ln.startswith("Max_quanta"):
            ln = "Max_quanta: %d" % ntot
        elif len(f) == 3 and f[0].isdigit() and m < len(caps) and "." in f[1]:
            ln = " %s %s %d" % (f[0], f[1], caps[m])
            m += 1
        out.append(ln)
    return "\n".join(out) + "\n"


def test_initial_basis_on_device(tmp_path):
    """The initial product basis enumerated on the device (unranked in the
    reference's order, Input_Reader.cpp:308-364) equals the host enumeration
    row for row, including per-mode caps below Max_quanta."""
    texts = {}
    for f in ("clo2_vci.inp", "clo2_vhci.inp", "acn_vci.inp", "acn_vhci.inp"):
        texts[f] = open(os.path.join(GOLDEN, f)).read()
    for f in sorted(os.listdir(os.path.join(GOLDEN, "fields"))):
        texts[f] = open(os.path.join(GOLDEN, "fields", f)).read()
    for c in (0, 1, 2):
        texts["c%d" % c] = synth.config_input(c)
    for r in range(4):
        texts["request%d" % r] = synth.request_input(r)
    clo2 = texts["clo2_vci.inp"]
    texts["clo2 caps 2,6,3 ntot 6"] = _capped(clo2, [2, 6, 3], 6)
    texts["clo2 caps 0,9,1 ntot 4"] = _capped(clo2, [0, 9, 1], 4)
    for name, text in texts.items():
        p = vhci_amd.Problem(text=text)
        ctx = vhci_amd.Context(p)
        qd, qh = ctx.basis(), p.initial_basis()
        _say("[initial basis]", name, qh.shape)
        assert qd.shape == qh.shape and np.array_equal(qd, qh), name
        ctx.close()


@pytest.mark.parametrize("name", ["clo2_pt2.inp", "clo2_vhci.inp"])
def test_matrix_free_matches_stored(name, monkeypatch):
    """ClO2 through matrix-free products (forced) against the stored CSR:
    the dense first diagonalization sees the same matrix (H times identity
    columns), so basis and energies agree; the stored run is the default."""
    p = vhci_amd.Problem(os.path.join(GOLDEN, name))
    ctx = vhci_amd.Context(p)
    ref = ctx.run()
    assert not ctx.matrix_free
    ctx.close()
    monkeypatch.setenv("VHCI_MATRIX_FREE", "1")
    ctx = vhci_amd.Context(p)
    r = ctx.run()
    assert ctx.matrix_free
    ctx.close()
    assert r["nvar"] == ref["nvar"] and r["ns2"] == ref["ns2"] and r["checksum"] == ref["checksum"]
    assert np.max(np.abs(r["E"] - ref["E"])) < 1e-8


SERVE_MIX = dict(m_range=(8, 14), eps1_set=(1.0, 2.0), nstate_range=(20, 200), eps2_set=(0.0, 0.01, 0.1))


def test_davidson_matches_krylov_schur(monkeypatch):
    """The opt-in block Davidson (VHCI_EIGSOLVER=davidson) against the
    default Krylov-Schur on a request that grows past 16384 states: each
    diagonalization warm-starts from the previous basis's vectors, placed on
    the solver's row order.  Same basis, energies within 1e-6 cm^-1."""
    text = synth.request_input(5, **SERVE_MIX)
    p = vhci_amd.Problem(text=text)
    ctx = vhci_amd.Context(p)
    ref = ctx.run()
    ctx.close()
    monkeypatch.setenv("VHCI_EIGSOLVER", "davidson")
    ctx = vhci_amd.Context(p)
    r = ctx.run()
    ctx.close()
    assert r["nvar"] >= 16384
    assert r["nvar"] == ref["nvar"] and r["ns2"] == ref["ns2"] and r["checksum"] == ref["checksum"]
    assert np.max(np.abs(r["E"] - ref["E"])) < 1e-6


@pytest.mark.parametrize("form", [vhci_amd.FP64, vhci_amd.MIXED])
def test_incremental_build_matches_full(form, monkeypatch):
    """Each heat-bath iteration rebuilds H from the previous one (new rows
    probed, old rows carried over with the transposes of the new entries);
    the matrix must be the one a full build gives, bit for bit.  The
    request grows past 16384 states, so the solver order appears midway."""
    text = synth.request_input(5, **SERVE_MIX)
    out = {}
    for full in (False, True):
        if full:
            monkeypatch.setenv("VHCI_FULL_BUILD", "1")
        else:
            monkeypatch.delenv("VHCI_FULL_BUILD", raising=False)
        ctx = vhci_amd.Context(vhci_amd.Problem(text=text), form)
        r = ctx.run()
        out[full] = (r, ctx.iteration_sizes(), ctx.get_h())
        ctx.close()
    (ri, si, hi), (rf, sf, hf) = out[False], out[True]
    _say("[incremental build] sizes", si)
    assert si == sf and si[-1] > 16384
    assert ri["checksum"] == rf["checksum"] and ri["ns2"] == rf["ns2"]
    for a, b in zip(hi, hf):
        assert np.array_equal(a, b)
    assert np.array_equal(ri["E"], rf["E"])
