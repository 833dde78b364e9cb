"""Host-side logic of libvhci (no GPU calls): the C ABI loads and exports
every declared symbol; input parsing, force-constant scaling, the heat-bath
list, the initial basis and the ladder sums are bit-identical to the oracle."""
import math
import os
import re

import numpy as np
import pytest

from helpers import GOLDEN

import vhci_amd
from vhci_amd import synth

pyoracle = pytest.importorskip("oracle.pyoracle")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    with open(os.path.join(ROOT, "include", "vhci.h")) as f:
        txt = f.read()
    return sorted(set(re.findall(r"\b(vhci_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_header_symbols():
    L = vhci_amd.lib()
    names = header_functions()
    assert len(names) >= 30
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(vhci_amd.EXPORTED)


def _inputs(tmp_path):
    files = [os.path.join(GOLDEN, f) for f in ("clo2_vci.inp", "clo2_vhci.inp", "acn_vhci.inp")]
    files += [os.path.join(GOLDEN, "fields", f) for f in sorted(os.listdir(os.path.join(GOLDEN, "fields")))]
    for c in (0, 1, 2):
        p = str(tmp_path / ("c%d.inp" % c))
        synth.config_input(c, p)
        files.append(p)
    return files


def test_parser_matches_oracle(tmp_path):
    """Every FC list (AnharmFC, the four order lists, AnharmHB in
    construction and sorted order) equal value for value, bit for bit."""
    for path in _inputs(tmp_path):
        p = vhci_amd.Problem(path)
        o = pyoracle.OProblem(path)
        assert p.M == o.M and p.n0 == o.n0 and p.fcmax == o.fcmax
        assert p.eps1 == o.eps1 and p.eps2 == o.eps2
        for which in range(7):
            a = p.fc_list(which)
            b = o.fc_list(which)
            assert len(a) == len(b), (path, which)
            for x, y in zip(a, b):
                assert x[0] == y[0] and x[1] == y[1] and x[2] == y[2], (path, which, x, y)
        assert np.array_equal(p.initial_basis(), o.initial_basis())


def test_restart_builds_ground_state_only():
    p = vhci_amd.Problem(os.path.join(GOLDEN, "clo2_vhci.inp"), restart=True)
    assert p.n0 == 1
    assert np.array_equal(p.initial_basis(), np.zeros((1, 3), dtype=np.int32))


def test_ladder_sum_bit_exact():
    for k in range(7):
        for qs in range(0, 40):
            for qe in range(max(0, qs - k), qs + k + 1):
                assert vhci_amd.ladder_sum(k, qs, qe) == pyoracle.coeffsum(k, qs, qe)
    assert vhci_amd.ladder_sum(4, 3, 3) == pytest.approx(6 * 9 + 18 + 3)


def test_parser_errors(tmp_path):
    bad_mode = tmp_path / "bad.inp"
    bad_mode.write_text("c\nHCI_Eps: 0\nNstate: 2\nPT2: 0\nPT2_Eps: 0\nMax_quanta: 2\nModes: 2\n"
                        " 0 100.0 3\n 2 200.0 3\nForce_constants: 0\n")
    with pytest.raises(vhci_amd.VhciError, match="Expected mode 1 but read data for mode 2"):
        vhci_amd.Problem(str(bad_mode))
    bad_order = tmp_path / "bad2.inp"
    bad_order.write_text("c\nHCI_Eps: 0\nNstate: 2\nPT2: 0\nPT2_Eps: 0\nMax_quanta: 2\nModes: 1\n"
                         " 0 100.0 3\nForce_constants: 1\n 2 0 0 5.0\n")
    with pytest.raises(vhci_amd.VhciError, match="only supported up to 6th order"):
        vhci_amd.Problem(str(bad_order))
    with pytest.raises(vhci_amd.VhciError, match="Could not open input file"):
        vhci_amd.Problem(str(tmp_path / "missing.inp"))


def test_pt2_flag_zeroes_eps2():
    txt = open(os.path.join(GOLDEN, "clo2_vhci.inp")).read().replace("PT2_Eps: 0", "PT2_Eps: 0.5")
    p = vhci_amd.Problem(text=txt)
    assert not p.perturb and p.eps2 == 0.0


def test_synthetic_generator_deterministic(tmp_path):
    a = synth.config_input(0)
    b = synth.config_input(0)
    assert a == b
    p = vhci_amd.Problem(text=a)
    assert p.M == 12 and p.n0 == math.comb(16, 4) == 1820
    assert p.nfc == 364 + 210
    p1 = vhci_amd.Problem(text=synth.config_input(1))
    assert p1.M == 15 and p1.n0 == 136 and p1.nfc == 680 + 600
    # first draws of the xoshiro256** stream seeded by splitmix64(0)
    r = synth.Rng(0)
    assert [r.next() for _ in range(3)] == [11091344671253066420, 13793997310169335082, 1900383378846508768]


# ---- checkpoint files (WriteCheckpoint / ReadCheckpoint) ------------------

def _ck_data(P, n, ncol, seed=5):
    r = np.random.default_rng(seed)
    q = r.integers(0, 7, size=(n, P.M)).astype(np.int32)
    E = np.sort(r.uniform(1000, 20000, size=ncol))
    C = r.standard_normal((n, ncol)) * 10.0 ** r.integers(-19, 1, size=(n, ncol))
    C[0, 0] = 0.0
    C[0, 1] = -0.0
    C[1, 0] = 1.0 - 2.0 ** -53
    return q, E, C


def _ck_text(P, q, E, C):
    """The reference's checkpoint text: fixed notation with 17 decimals
    (std::fixed << std::setprecision(17), i.e. printf %.17f), one space after
    every value, one line per state."""
    n = q.shape[0]
    lines = ["%d %d %d\n" % (n, P.M, P.neig)]
    lines.append("".join("%.17f " % x for x in list(P.freqs()) + list(E)) + "\n")
    for i in range(n):
        lines.append("".join("%d " % x for x in q[i]) + "".join("%.17f " % x for x in C[i]) + "\n")
    return "".join(lines).encode()


@pytest.fixture(scope="module")
def ck_problem():
    return vhci_amd.Problem(os.path.join(GOLDEN, "clo2_vhci.inp"))


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_checkpoint_text_is_the_reference_format(ck_problem, tmp_path, threads):
    P = ck_problem
    q, E, C = _ck_data(P, 5000, P.neig)
    path = str(tmp_path / "t.ck")
    P.write_checkpoint(path, q, E, C, threads=threads)
    with open(path, "rb") as f:
        assert f.read() == _ck_text(P, q, E, C)


def test_checkpoint_text_round_trip(ck_problem, tmp_path):
    P = ck_problem
    q, E, C = _ck_data(P, 7000, P.neig, seed=9)
    path = str(tmp_path / "r.ck")
    P.write_checkpoint(path, q, E, C, threads=4)
    for threads in (1, 5):
        q2, E2, C2 = P.read_checkpoint(path, threads=threads)
        assert np.array_equal(q2, q)
        # the values come back as the 17-decimal text rounds them
        assert np.array_equal(E2, np.array([float("%.17f" % x) for x in E]))
        assert np.array_equal(C2, np.vectorize(lambda x: float("%.17f" % x))(C))


def test_checkpoint_binary_round_trip_is_exact(ck_problem, tmp_path):
    P = ck_problem
    q, E, C = _ck_data(P, 3000, P.neig, seed=11)
    path = str(tmp_path / "b.ck")
    P.write_checkpoint(path, q, E, C, binary=True)
    with open(path, "rb") as f:
        assert f.read(8) == b"VHCICKB1"
    q2, E2, C2 = P.read_checkpoint(path)
    assert np.array_equal(q2, q) and np.array_equal(E2, E) and np.array_equal(C2, C)


def test_checkpoint_errors(ck_problem, tmp_path):
    P = ck_problem
    q, E, C = _ck_data(P, 100, P.neig)
    good = str(tmp_path / "g.ck")
    P.write_checkpoint(good, q, E, C)
    text = open(good, "rb").read()
    lines = text.split(b"\n")

    def bad(name, data, msg):
        p = str(tmp_path / name)
        with open(p, "wb") as f:
            f.write(data)
        with pytest.raises(vhci_amd.VhciError, match=re.escape(msg)):
            P.read_checkpoint(p)

    n, M, ne = lines[0].split()
    bad("m.ck", b" ".join([n, b"%d" % (int(M) + 1), ne]) + b"\n" + b"\n".join(lines[1:]),
        "wrong number of fundamental modes")
    bad("e.ck", b" ".join([n, M, b"%d" % (int(ne) + 1)]) + b"\n" + b"\n".join(lines[1:]),
        "wrong number of eigenstates")
    f0 = lines[1].split(b" ")
    f0[0] = b"%.17f" % (float(f0[0]) + 1.0)
    bad("f.ck", b"\n".join([lines[0], b" ".join(f0)] + lines[2:]), "wrong fundamental frequencies")
    bad("t.ck", text[: len(text) // 2], "truncated")
    P.write_checkpoint(str(tmp_path / "bb.ck"), q, E, C, binary=True)
    bb = open(str(tmp_path / "bb.ck"), "rb").read()
    bad("bt.ck", bb[: len(bb) - 8], "truncated")


def test_published_run_inputs():
    """Each published run's input (the committed force field with the header
    its file name states) parses with that eps1, Nstate, PT2 flag and eps2,
    total quanta 2, and an initial basis of C(M + 2, 2) states."""
    from test_gpu_results import CASES, NAME, _input_text

    assert len(CASES) == 14
    for name in CASES:
        mol, eps1, nstate, pt2, eps2 = NAME.match(name).groups()
        p = vhci_amd.Problem(text=_input_text(name))
        assert p.eps1 == float(eps1) and p.neig == int(nstate) and p.perturb == (pt2 == "1")
        assert p.eps2 == (float(eps2) if pt2 == "1" else 0.0)
        assert p.ntot == 2 and p.n0 == math.comb(p.M + 2, 2)


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: a request for more exits with status 2 and
    says so, before it touches a device."""
    import subprocess
    import sys

    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and "--gpus 2" in r.stderr
    assert r.stdout == ""
