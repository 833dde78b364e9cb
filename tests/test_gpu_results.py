"""The drop-in `vhci` against the reference's published runs on real force
fields (tests/golden/results: C2H4 in two fields, acetonitrile, ethylene
oxide, naphthalene; heat-bath VCI with and without PT2, bases up to 3
million and PT2 spaces up to 7.2 million states; about a minute for all).
Each input is the committed force field with its header set to the eps1,
Nstate, PT2 and eps2 that the result's file name states and total quanta 2
(the per-mode cap of 2 that the name also states cannot bind when the total
is 2).

Compared: the number of lines; the ZPE line (energy and assignment); the
excitation energies as sorted lists; and, for every state with no other
state within 0.01 cm^-1, the set of assignment terms of its line.  The
terms' order within a line follows the basis row order, which depends on
the reference's hash-set iteration, and the reference prints a few
near-degenerate pairs of naphthalene out of ascending order, so neither
order is compared.  The energies agree to a few 1e-9 cm^-1; the test allows
1e-7."""
import os
import re
import subprocess

import numpy as np
import pytest

from helpers import GOLDEN

import vhci_amd

pytestmark = pytest.mark.gpu

RESULTS = os.path.join(GOLDEN, "results")
FIELD = {"C2H4": "C2H4.inp", "C2H4_quartic": "C2H4_quartic.inp", "acetonitrile": "acetonitrile.inp",
         "ethylene_oxide": "ethylene_oxide.inp", "naphthalene": "naphthalene.inp"}
NAME = re.compile(r"freq_(.+)_eps1-([\d.]+)_nstate-(\d+)_perturb-(\d)_eps2-([\d.]+)_nmax-2_ntot-2\.dat")
CASES = sorted(f for f in os.listdir(RESULTS) if NAME.match(f))


def _say(*a):
    print(*a, flush=True)


def _input_text(name):
    mol, eps1, nstate, pt2, eps2 = NAME.match(name).groups()
    lines = open(os.path.join(GOLDEN, "fields", FIELD[mol])).read().split("\n")
    # header lines 2-6: HCI_Eps, Nstate, PT2, PT2_Eps, Max_quanta (label, value)
    for i, v in enumerate((eps1, nstate, pt2, eps2, "2")):
        lines[1 + i] = lines[1 + i].split()[0] + " " + v
    return "\n".join(lines)


def _read(path):
    rows = []
    for line in open(path):
        if not line.strip():
            continue
        head, _, rest = line.rstrip("\n").partition("   ")
        terms = sorted(t.strip() for t in re.split(r"  +", rest.strip()) if t.strip())
        rows.append((float(head), terms))
    return rows


def test_results_fixtures_present():
    assert len(CASES) == 14


@pytest.mark.parametrize("name", CASES)
def test_published_run(tmp_path, name):
    _say("published run", name)
    inp = tmp_path / "in.inp"
    inp.write_text(_input_text(name))
    out = str(tmp_path / "freqs.dat")
    r = subprocess.run([vhci_amd.CLI_PATH, "-n", "8", "-i", str(inp), "-o", out], capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0 and os.path.exists(out), r.stdout[-2000:] + r.stderr[-2000:]
    got, exp = _read(out), _read(os.path.join(RESULTS, name))
    assert len(got) == len(exp)
    assert abs(got[0][0] - exp[0][0]) < 1e-7 and got[0][1] == exp[0][1]
    gs = sorted(got[1:], key=lambda x: x[0])
    es = sorted(exp[1:], key=lambda x: x[0])
    de = np.abs(np.array([g[0] for g in gs]) - np.array([e[0] for e in es]))
    _say("  max |dE| %.2e cm^-1 over %d lines" % (max(de.max(), abs(got[0][0] - exp[0][0])), len(exp)))
    assert de.max() < 1e-7
    ev = [e[0] for e in es]
    for i, (g, e) in enumerate(zip(gs, es)):
        near = (i > 0 and ev[i] - ev[i - 1] < 1e-2) or (i + 1 < len(ev) and ev[i + 1] - ev[i] < 1e-2)
        if not near:
            assert g[1] == e[1], (i, g, e)
