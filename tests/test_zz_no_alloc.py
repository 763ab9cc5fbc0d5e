"""The allocation-free batched RHS (phi_hip_rhs_noalloc) through the C ABI:
bit-for-bit the output of phi_hip_rhs on the same input, on two calls in a
row (the 7-site hierarchy at depth 8, 315315 complex elements, then the
smaller 2-site dimer at depth 8, 1980 elements, so the second call reuses
the device's staging array with a shorter input), and a clear error for a
state beyond its staging limit (the 24-site ring at depth 4)."""

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import phi_amd as P  # noqa: E402

pytestmark = pytest.mark.gpu

CFG = os.path.join(ROOT, "phi_amd", "configs")


def _random_state(ctx, seed):
    rng = np.random.default_rng(seed)
    shape = (ctx.num_nodes, ctx.N, ctx.N)
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def test_rhs_noalloc_equals_rhs_on_two_calls():
    for name, seed in (("c2", 11), ("c1", 12)):
        print("case: rhs_noalloc vs rhs,", name, flush=True)
        ctx = P.Context(P.PhiParameters.read(os.path.join(CFG, name + ".prm"), "rk4"))
        x = _random_state(ctx, seed)
        want = ctx.rhs(x)
        got = ctx.rhs_noalloc(x)
        print("  %d complex elements, max |d| = %.3e"
              % (ctx.n_complex, np.abs(got - want).max()), flush=True)
        assert np.array_equal(got, want)
        assert np.abs(want).max() > 0
        ctx.close()


def test_rhs_noalloc_refuses_a_state_beyond_its_staging():
    print("case: rhs_noalloc on the 24-site ring (11.8 M elements)", flush=True)
    ctx = P.Context(P.PhiParameters.read(os.path.join(CFG, "c3.prm"), "rk4"))
    x = np.zeros((ctx.num_nodes, ctx.N, ctx.N), dtype=complex)
    with pytest.raises(P.PhiError) as e:
        ctx.rhs_noalloc(x)
    assert e.value.code == -1
    assert "allocation-free staging" in str(e.value)
    ctx.close()
