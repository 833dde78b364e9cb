"""The serving loop: requests of different sizes, run concurrently on one GPU,
give the same results as alone; admission by predicted memory; bad inputs
are rejected at submission without disturbing the others."""
import os

import numpy as np
import pytest

from helpers import GOLDEN

import vhci_amd
from vhci_amd import synth

pytestmark = pytest.mark.gpu

# small serving requests: 6-9 modes, eps1 1 or 2, PT2 with eps2 0.01 or 0.1
RANGES = dict(m_range=(6, 9), eps1_set=(1.0, 2.0), nstate_range=(10, 60), eps2_set=(0.01, 0.1))
NREQ = 8
TOL_CM = 1e-9  # energies alone vs in the mix


def _say(*a):
    print(*a, flush=True)


def _texts():
    return [synth.request_input(r, **RANGES) for r in range(NREQ)]


@pytest.fixture(scope="module")
def alone():
    """Every request on a one-slot server, one after another."""
    srv = vhci_amd.Server(slots=1)
    out = []
    for r, t in enumerate(_texts()):
        _say("[serving] alone request", r, synth.request_params(r, **RANGES)[:4])
        rid = srv.submit(t)
        assert srv.wait(rid, 300)
        out.append(srv.result(rid))
    assert srv.stats()["max_running"] == 1
    srv.close()
    return out


def test_mix_matches_alone(alone):
    srv = vhci_amd.Server(slots=4)
    ids = [srv.submit(t) for t in _texts()]
    _say("[serving] mix submitted", len(ids))
    res = []
    for rid in ids:
        assert srv.wait(rid, 300)
        res.append(srv.result(rid))
    st = srv.stats()
    srv.close()
    _say("[serving] mix stats", st)
    assert st["finished"] == NREQ and st["max_running"] > 1
    for a, b in zip(alone, res):
        assert a["state"] == b["state"] == "done"
        assert a["nvar"] == b["nvar"] and a["ns2"] == b["ns2"] and a["checksum"] == b["checksum"]
        assert np.max(np.abs(a["E"] - b["E"])) <= TOL_CM
        assert np.max(np.abs(a["E_var"] - b["E_var"])) <= TOL_CM


def test_large_request_in_mix(alone):
    """A request of the serving workload's own distribution with 34 modes
    (request 8: eps1 2, Nstate 60, eps2 0.1) alone, then in a mix with the
    small ones on four slots: every result the same as alone."""
    big = synth.request_input(8)
    assert synth.request_params(8)[0] >= 30
    srv = vhci_amd.Server(slots=1)
    rid = srv.submit(big)
    assert srv.wait(rid, 300)
    big_alone = srv.result(rid)
    srv.close()
    _say("[serving] large request alone: %d states, PT2 space %d, %.2f s" %
         (big_alone["nvar"], big_alone["ns2"], big_alone["t_end"] - big_alone["t_start"]))
    srv = vhci_amd.Server(slots=4)
    ids = [srv.submit(big)] + [srv.submit(t) for t in _texts()]
    res = []
    for rid in ids:
        assert srv.wait(rid, 300)
        res.append(srv.result(rid))
    st = srv.stats()
    srv.close()
    _say("[serving] large + small mix stats", st)
    assert st["max_running"] > 1
    for a, b in zip([big_alone] + alone, res):
        assert a["state"] == b["state"] == "done"
        assert a["nvar"] == b["nvar"] and a["ns2"] == b["ns2"] and a["checksum"] == b["checksum"]
        assert np.max(np.abs(a["E"] - b["E"])) <= TOL_CM
        assert np.max(np.abs(a["E_var"] - b["E_var"])) <= TOL_CM


def test_requests_overlap_in_time(alone):
    """With four slots, later requests start before earlier ones end."""
    srv = vhci_amd.Server(slots=4)
    ids = [srv.submit(t) for t in _texts()]
    for rid in ids:
        assert srv.wait(rid, 300)
    rec = [srv.result(rid) for rid in ids]
    srv.close()
    overlaps = sum(1 for i in range(len(rec)) for j in range(i + 1, len(rec))
                   if rec[j]["t_start"] < rec[i]["t_end"] and rec[i]["t_start"] < rec[j]["t_end"])
    assert overlaps > 0
    assert len({r["slot"] for r in rec}) > 1


def test_memory_limit_serializes(alone):
    """A limit below any prediction admits one request at a time (a request
    is always admitted when none runs), and all of them still finish."""
    srv = vhci_amd.Server(slots=4, mem_limit=1)
    ids = [srv.submit(t) for t in _texts()[:4]]
    for rid in ids:
        assert srv.wait(rid, 300)
    st = srv.stats()
    rec = [srv.result(rid) for rid in ids]
    srv.close()
    assert st["max_running"] == 1
    for a, b in zip(alone, rec):
        assert a["checksum"] == b["checksum"]
        assert np.max(np.abs(a["E"] - b["E"])) <= TOL_CM
    # predictions and measured peaks are reported
    assert all(r["predicted_bytes"] > 0 and r["peak_bytes"] > 0 for r in rec)


def test_bad_input_rejected_others_unaffected(alone):
    srv = vhci_amd.Server(slots=2)
    good = srv.submit(_texts()[0])
    with pytest.raises(vhci_amd.VhciError):
        srv.submit("Comment: broken\nHCI_Eps: 1.0\nNstate: 5\n")
    assert srv.wait(good, 300)
    r = srv.result(good)
    srv.close()
    assert r["checksum"] == alone[0]["checksum"]


def test_server_matches_solve_on_fixture():
    """The ClO2 VHCI+PT2 fixture through the server equals vhci_amd.solve."""
    path = os.path.join(GOLDEN, "clo2_pt2.inp")
    text = open(path).read()
    ref = vhci_amd.solve(vhci_amd.Problem(path))
    srv = vhci_amd.Server(slots=2)
    ids = [srv.submit(text), srv.submit(text, form=vhci_amd.MIXED)]
    for rid in ids:
        assert srv.wait(rid, 300)
    r64, r32 = srv.result(ids[0]), srv.result(ids[1])
    srv.close()
    assert r64["nvar"] == ref["nvar"] and r64["ns2"] == ref["ns2"]
    assert np.max(np.abs(r64["E"] - ref["E"])) <= TOL_CM
    assert r32["checksum"] == r64["checksum"]
    assert np.max(np.abs(r32["E"] - r64["E"])) <= 0.01
