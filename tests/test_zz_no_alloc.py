"""The product entry point that allocates no device memory
(vhci_apply_h_noalloc): through the C ABI, its results equal those of
vhci_apply_h bit for bit on two calls in a row (the second with a smaller
input), it takes no block from the context's device-memory pool, and it
rejects a block beyond its scratch limit."""
import os

import numpy as np
import pytest

from helpers import GOLDEN

import vhci_amd
from vhci_amd import synth

pytestmark = pytest.mark.gpu


def _ctx(text=None, path=None, form=vhci_amd.FP64):
    p = vhci_amd.Problem(path=path, text=text)
    ctx = vhci_amd.Context(p, form)
    ctx.build_h()
    return ctx


@pytest.mark.parametrize("form", [vhci_amd.FP64, vhci_amd.MIXED])
def test_noalloc_matches_apply_h(form):
    # first call: a basis of 50,388 states (solver order on), 16 vectors;
    # second: a 3,276-state basis (basis order), 5 vectors
    big = _ctx(path=os.path.join(GOLDEN, "acn_vci.inp"), form=form)
    small = _ctx(path=os.path.join(GOLDEN, "clo2_vci.inp"), form=form)
    rng = np.random.default_rng(11)
    for ctx, b in ((big, 16), (small, 5)):
        X = rng.standard_normal((ctx.n, b))
        before = ctx.pool_stats()
        Y0 = ctx.apply_h_noalloc(X)
        after = ctx.pool_stats()
        Y1 = ctx.apply_h(X)
        print("n %d b %d pool before %s after %s" % (ctx.n, b, before, after), flush=True)
        assert after["mallocs"] == before["mallocs"] and after["reuses"] == before["reuses"]
        assert np.array_equal(Y0, Y1)
        if form == vhci_amd.MIXED:
            Z0 = ctx.apply_h_noalloc(X, fp32=True)
            Z1 = ctx.apply_h(X, fp32=True)
            assert np.array_equal(Z0, Z1)
            assert np.max(np.abs(Z0 - Y1)) < 1e-4 * np.max(np.abs(Y1))
    big.close()
    small.close()


def test_noalloc_rejects_beyond_limit():
    ctx = _ctx(text=synth.config_input(0))
    n = ctx.n  # 1820 states: 64 vectors fit
    X = np.ones((n, 64))
    assert np.array_equal(ctx.apply_h_noalloc(X), ctx.apply_h(X))
    ctx.close()
    big = _ctx(path=os.path.join(GOLDEN, "acn_vci.inp"))
    X = np.ones((big.n, 64))  # 50,388 x 64 = 3.2 M elements: within the limit
    assert np.array_equal(big.apply_h_noalloc(X), big.apply_h(X))
    big.close()
    # 4,194,304 elements at most: an n x b beyond it is refused cleanly
    from vhci_amd import VhciError
    p = vhci_amd.Problem(text=synth.config_input(1, pt2=False))
    ctx = vhci_amd.Context(p, vhci_amd.FP64)
    ctx.run()  # 161,747 states
    X = np.ones((ctx.n, 32))  # 5.2 M elements
    with pytest.raises(VhciError, match="scratch limit"):
        ctx.apply_h_noalloc(X)
    Y = ctx.apply_h_noalloc(np.ones((ctx.n, 16)))  # 2.6 M: fine after the refusal
    assert np.array_equal(Y, ctx.apply_h(np.ones((ctx.n, 16))))
    ctx.close()
