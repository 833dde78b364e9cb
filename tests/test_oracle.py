"""Pins of the CPU oracle: the reference's own regression fixtures, closed
forms of the ladder algebra, brute force on small bases and library
eigensolvers.  CPU only."""
import itertools
import json
import math
import os

import numpy as np
import pytest

from helpers import GOLDEN, absolute, manifest, read_freqs

pyoracle = pytest.importorskip("oracle.pyoracle")


def _run(tmp_path, name, threads=8):
    inp = os.path.join(GOLDEN, name + ".inp")
    out = str(tmp_path / (name + ".freqs"))
    js = str(tmp_path / (name + ".json"))
    pyoracle.run(inp, out, threads=threads, json_out=js)
    with open(js) as f:
        return out, json.load(f)


@pytest.mark.parametrize("name", ["clo2_vci", "clo2_vhci", "clo2_pt2"])
def test_clo2_fixtures(tmp_path, name):
    """ClO2 examples: every energy within 1e-8 cm^-1 of the reference's
    freqs.dat, identical assignment strings, exact basis sizes per heat-bath
    iteration and exact PT2 space size (no degeneracies in ClO2)."""
    out, js = _run(tmp_path, name)
    e, a = read_freqs(out)
    er, ar = read_freqs(os.path.join(GOLDEN, name + ".freqs"))
    assert len(e) == len(er)
    assert np.max(np.abs(absolute(e) - absolute(er))) < 1e-8
    assert a == ar
    pins = manifest()[name]
    assert js["N0"] == pins["basis_functions"]
    sizes = [it["N"] for it in js["iters"]]
    if pins["iterations"]:
        assert sizes == pins["iterations"]
        assert js["N_final"] == pins["converged_size"]
    if "pt2_space" in pins:
        assert js["pt2_space"] == pins["pt2_space"]
    # structural sparsity line (Ham.cpp:341-342) printed with 2 decimals
    for it, pct in zip(js["iters"], pins["sparse_pct"]):
        n = it["N"]
        assert round(100 * (1 - it["nnz"] / (n * n)), 2) == pytest.approx(pct, abs=1e-9)


@pytest.mark.slow
def test_acetonitrile_vhci(tmp_path):
    """Acetonitrile VHCI: exact basis sizes 91 -> 9749 -> 22857 -> 28666 ->
    29900 and all 70 energies within 1e-8 cm^-1 (assignments of the
    degenerate E pairs are rotation dependent and not compared)."""
    out, js = _run(tmp_path, "acn_vhci")
    pins = manifest()["acn_vhci"]
    assert [it["N"] for it in js["iters"]] == pins["iterations"]
    e, _ = read_freqs(out)
    er, _ = read_freqs(os.path.join(GOLDEN, "acn_vhci.freqs"))
    assert np.max(np.abs(absolute(e) - absolute(er))) < 1e-8


@pytest.mark.slow
def test_acetonitrile_vci(tmp_path):
    """Acetonitrile VCI (N = 50388, Lanczos path): 70 energies within 1e-8."""
    out, js = _run(tmp_path, "acn_vci")
    assert js["N0"] == 50388
    e, _ = read_freqs(out)
    er, _ = read_freqs(os.path.join(GOLDEN, "acn_vci.freqs"))
    assert np.max(np.abs(absolute(e) - absolute(er))) < 1e-8


def ladder_matrix(nmax):
    """(a + a^dag) on the first nmax levels."""
    X = np.zeros((nmax, nmax))
    for n in range(nmax - 1):
        X[n + 1, n] = X[n, n + 1] = math.sqrt(n + 1)
    return X


def test_coeffsum_closed_forms():
    """<m|(a+a^dag)^k|n> against textbook values and matrix powers."""
    for n in range(12):
        assert pyoracle.coeffsum(1, n, n + 1) == pytest.approx(math.sqrt(n + 1), rel=1e-15)
        assert pyoracle.coeffsum(2, n, n) == pytest.approx(2 * n + 1, rel=1e-14)
        assert pyoracle.coeffsum(4, n, n) == pytest.approx(6 * n * n + 6 * n + 3, rel=1e-14)
        assert pyoracle.coeffsum(3, n, n + 1) == pytest.approx(3 * (n + 1) ** 1.5, rel=1e-14)
    X = ladder_matrix(40)
    P = np.eye(40)
    for k in range(1, 7):
        P = P @ X
        for n in range(20):
            for m in range(max(0, n - k), n + k + 1):
                assert pyoracle.coeffsum(k, n, m) == pytest.approx(P[m, n], rel=1e-12, abs=1e-12)


def _brute_h(prob, q, fcs, freqs=None):
    """H from the mathematics: harmonic diagonal plus, for every force
    constant and every pair, f * prod_modes ((a+a^dag)^k)[q_j, q_i]; no
    screening or prefilter."""
    nmax = int(q.max()) + 8
    X = ladder_matrix(nmax)
    pw = [np.eye(nmax)]
    for _ in range(6):
        pw.append(pw[-1] @ X)
    n = q.shape[0]
    H = np.zeros((n, n))
    for i in range(n):
        H[i, i] = sum((0.5 + q[i, j]) * freqs[j] for j in range(q.shape[1]))
    for f, order, smp in fcs:
        modes = [m for m, _ in smp]
        others = [j for j in range(q.shape[1]) if j not in modes]
        for i in range(n):
            for j in range(n):
                if any(q[i, o] != q[j, o] for o in others):
                    continue
                v = f
                for m, k in smp:
                    v *= pw[k][q[j, m], q[i, m]]
                H[i, j] += v
    return H


def test_dense_h_brute_force(tmp_path):
    """Oracle H (screening, ladder paths, even highest order so the odd
    prefilter drops nothing) equals the brute-force operator algebra on the
    ClO2 6th-order field over a 4-quanta basis."""
    prob = pyoracle.OProblem(os.path.join(GOLDEN, "fields", "CLO2.inp"))
    q = np.array([s for s in itertools.product(range(5), repeat=3) if sum(s) <= 4], dtype=np.int32)
    H = prob.build_h_dense(q)
    fcs = prob.fc_list(0)
    Hb = _brute_h(prob, q, fcs, [455.617542, 960.149798, 1127.824246])
    assert np.allclose(H, Hb, rtol=1e-12, atol=1e-9)
    # sparse and dense builds agree entry by entry
    rp, col, val = prob.build_h(q)
    Hs = np.zeros_like(H)
    for i in range(len(q)):
        Hs[i, col[rp[i]:rp[i + 1]]] = val[rp[i]:rp[i + 1]]
    assert np.array_equal(Hs, H)


def test_odd_prefilter_quirk(tmp_path):
    """With cubic-only force constants (fcmax = 3) the reference's odd
    prefilter (qdiff <= fcmax-1, Ham.cpp:283) drops Delta = 3 couplings and
    three-mode ijk couplings; the oracle must reproduce that."""
    txt = ("c\nHCI_Eps: 0\nNstate: 2\nPT2: 0\nPT2_Eps: 0\nMax_quanta: 3\nModes: 3\n"
           " 0 1000.0 5\n 1 1500.0 5\n 2 2000.0 5\nForce_constants: 3\n"
           " 3 0 0 0 50.0\n 3 0 1 2 20.0\n 3 0 0 1 10.0\n")
    p = tmp_path / "cub.inp"
    p.write_text(txt)
    prob = pyoracle.OProblem(str(p))
    assert prob.fcmax == 3
    q = np.array([s for s in itertools.product(range(4), repeat=3) if sum(s) <= 3], dtype=np.int32)
    H = prob.build_h_dense(q)
    Hb = _brute_h(prob, q, prob.fc_list(0), [1000.0, 1500.0, 2000.0])
    diff = np.abs(H - Hb) > 1e-9
    for i, j in zip(*np.nonzero(diff)):
        d = np.abs(q[i] - q[j])
        assert d.sum() == 3  # only qdiff = 3 couplings differ
        assert H[i, j] == 0.0
    assert diff.any()


def test_eigh_matches_numpy():
    rng = np.random.default_rng(7)
    for n in (1, 2, 3, 17, 120):
        A = rng.standard_normal((n, n))
        A = A + A.T
        w, Z = pyoracle.eigh(A)
        wr = np.linalg.eigvalsh(A)
        assert np.allclose(w, wr, atol=1e-10 * max(1, np.abs(wr).max()))
        assert np.allclose(A @ Z, Z * w, atol=1e-9 * max(1, np.abs(wr).max()))
        assert np.allclose(Z.T @ Z, np.eye(n), atol=1e-10)


def test_lanczos_matches_numpy():
    """Krylov-Schur Lanczos on the ClO2 VCI Hamiltonian (N = 3276) against
    numpy's dense eigvalsh."""
    prob = pyoracle.OProblem(os.path.join(GOLDEN, "clo2_vci.inp"))
    q = prob.initial_basis()
    assert q.shape[0] == 3276 == math.comb(3 + 25, 25)
    rp, col, val = prob.build_h(q)
    E, C = pyoracle.lanczos(rp, col, val, 20)
    H = np.zeros((q.shape[0],) * 2)
    for i in range(q.shape[0]):
        H[i, col[rp[i]:rp[i + 1]]] = val[rp[i]:rp[i + 1]]
    wr = np.linalg.eigvalsh(H)[:20]
    assert np.allclose(E, wr, rtol=1e-13, atol=1e-9)


@pytest.mark.parametrize("name,m,ntot", [("clo2_vci", 3, 25), ("clo2_vhci", 3, 8), ("acn_vci", 12, 7),
                                         ("acn_vhci", 12, 2)])
def test_initial_basis_count(name, m, ntot):
    """C(M + Ntot, Ntot) when every cap >= Ntot (the logs' 3276, 165, 50388, 91)."""
    prob = pyoracle.OProblem(os.path.join(GOLDEN, name + ".inp"))
    assert prob.n0 == math.comb(m + ntot, ntot) == manifest()[name]["basis_functions"]
    q = prob.initial_basis()
    # mode 0 varies fastest, the last mode is the slowest key
    assert tuple(q[0]) == (0,) * m and q[1][0] == 1
    assert len({tuple(r) for r in q}) == len(q)


def test_scale_fc_known_answers():
    """ScaleFC: f * 1/sqrt(2^p) / prod k_j!  (f_iii -> f/sqrt(8)/6)."""
    prob = pyoracle.OProblem(os.path.join(GOLDEN, "clo2_vci.inp"))
    raw = {}
    with open(os.path.join(GOLDEN, "clo2_vci.inp")) as f:
        lines = f.read().split("\n")
    start = [i for i, l in enumerate(lines) if l.startswith("Force_constants")][0]
    for l in lines[start + 1:]:
        t = l.split()
        if t:
            raw[tuple(int(x) for x in t[1:-1])] = float(t[-1])
    fcs = prob.fc_list(0)
    assert len(fcs) == len(raw)
    for (f, order, smp), (modes, v) in zip(fcs, raw.items()):
        s = 1 / math.sqrt(2.0 ** order)
        for _, k in smp:
            s /= math.factorial(k)
        assert f == v * s
    f, order, smp = fcs[0]
    assert order == 3 and smp == [(0, 3)]
    assert f == raw[(0, 0, 0)] * ((1 / math.sqrt(8.0)) / 6.0)


def test_hb_select_brute_force():
    """Heat-bath selection against a direct statement of the rule on the
    ClO2 initial basis: every candidate q + delta (delta in the FC's
    {-k..k step 2} product, not all zero, q + delta >= 0) of every FC with
    |C_n f| >= eps that is not already in the basis."""
    prob = pyoracle.OProblem(os.path.join(GOLDEN, "clo2_vhci.inp"))
    q = prob.initial_basis()
    rng = np.random.default_rng(3)
    cmax = rng.standard_normal(q.shape[0]) * np.exp(-rng.uniform(0, 8, q.shape[0]))
    eps = 0.01
    new = prob.hb_select(q, cmax, eps)
    basis = {tuple(r) for r in q}
    expect = []
    seen = set()
    for n in range(q.shape[0]):
        for f, order, smp in prob.fc_list(6):
            if not abs(cmax[n] * f) >= eps:
                break
            ranges = [range(-k, k + 1, 2) for _, k in smp]
            for d in itertools.product(*ranges):
                if not any(d):
                    continue
                t = list(q[n])
                for (m, _), dd in zip(smp, d):
                    t[m] += dd
                t = tuple(t)
                if min(t) < 0 or t in basis or t in seen:
                    continue
                seen.add(t)
                expect.append(t)
    assert [tuple(r) for r in new] == expect


def _rows_without_zeros(rp, col, val, rows):
    out = []
    for k, r in enumerate(rows):
        c, v = col[rp[k]:rp[k + 1]], val[rp[k]:rp[k + 1]]
        keep = (v != 0.0) | (c == r)
        out.append((c[keep], v[keep]))
    return out


@pytest.mark.parametrize("field", ["clo2", "cubic", "acn"])
def test_fast_form_h_rows_equal_all_pairs(tmp_path, field):
    """The hash-based host form (the benchmark's stronger CPU baseline)
    gives the all-pairs rows bit for bit, stored zeros aside: the 6th-order
    ClO2 field, the cubic-only field whose odd prefilter drops couplings
    (fcmax = 3), and the acetonitrile field on a heat-bath grown basis."""
    rng = np.random.default_rng(11)
    if field == "clo2":
        prob = pyoracle.OProblem(os.path.join(GOLDEN, "fields", "CLO2.inp"))
        q = np.array([s for s in itertools.product(range(7), repeat=3) if sum(s) <= 6], dtype=np.int32)
        rows = np.arange(len(q))
    elif field == "cubic":
        txt = ("c\nHCI_Eps: 0\nNstate: 2\nPT2: 0\nPT2_Eps: 0\nMax_quanta: 3\nModes: 3\n"
               " 0 1000.0 5\n 1 1500.0 5\n 2 2000.0 5\nForce_constants: 3\n"
               " 3 0 0 0 50.0\n 3 0 1 2 20.0\n 3 0 0 1 10.0\n")
        p = tmp_path / "cub.inp"
        p.write_text(txt)
        prob = pyoracle.OProblem(str(p))
        q = np.array([s for s in itertools.product(range(6), repeat=3) if sum(s) <= 5], dtype=np.int32)
        rows = np.arange(len(q))
    else:
        prob = pyoracle.OProblem(os.path.join(GOLDEN, "acn_vhci.inp"))
        q0 = prob.initial_basis()
        new = prob.hb_select(q0, rng.standard_normal(q0.shape[0]) * 10, 1.0)
        q = np.concatenate([q0, new])
        rows = np.sort(rng.choice(len(q), 200, replace=False))
    a = _rows_without_zeros(*prob.h_rows(q, rows), rows)
    b = _rows_without_zeros(*prob.fast_h_rows(q, rows), rows)
    assert sum(len(x[0]) for x in a) > len(rows)
    for x, y in zip(a, b):
        assert np.array_equal(x[0], y[0])
        assert np.array_equal(x[1], y[1])


def test_fast_form_pt2_equals_all_pairs():
    """The hash-based host form's PT2 corrections over the external space
    that the heat-bath rule selects equal the all-pairs gather's bit for bit
    (same thread count); a corrupted coefficient changes them."""
    pyoracle.set_threads(4)
    prob = pyoracle.OProblem(os.path.join(GOLDEN, "acn_vhci.inp"))
    q0 = prob.initial_basis()
    rng = np.random.default_rng(5)
    q = np.concatenate([q0, prob.hb_select(q0, rng.standard_normal(q0.shape[0]) * 10, 2.0)])
    C = rng.standard_normal((len(q), 5))
    C /= np.linalg.norm(C, axis=0)
    E = np.array([9000.0, 9100.0, 9200.0, 9300.0, 9400.0])
    nopt, eps2 = 4, 0.2
    dE = prob.pt2(q, C, E, nopt, eps2)
    dE = np.asarray(dE[0] if isinstance(dE, tuple) else dE)
    qA = prob.hb_select(q, pyoracle.cmax(C, nopt), eps2)
    fast, pairs = prob.fast_pt2(q, C, E, nopt, qA)
    assert len(qA) > 1000 and pairs > len(qA)
    assert np.array_equal(dE, fast)
    C2 = C.copy()
    C2[len(q) // 2, 0] += 0.1
    bad, _ = prob.fast_pt2(q, C2, E, nopt, qA)
    assert bad[0] != fast[0]
