"""GPU: tdgl_euler_step_noalloc, the Forward-Euler step that allocates nothing
(its staging buffer is the library's per-GPU scratch array), against
tdgl_euler_step through the C ABI: bit for bit on two calls in a row on a film
with a hole, the second on a smaller film (the array keeps the first call's
contents), each from a fresh state and after a step of its own; a state
longer than the array is refused before any launch."""

from __future__ import annotations

import numpy as np
import pytest

import specs as S
from helpers import device, rand_state

pytestmark = pytest.mark.gpu


def _ctx(case, field):
    from nanowire_tdgl_amd.engine import engine_for
    dev = device(case)
    eng = engine_for(dev.params, dev.idx, dev.material)
    eng.set_field(*field)
    return eng.ctx, dev


@pytest.mark.parametrize("dt", [0.01, 0.05])
def test_noalloc_euler_matches_euler_step(dt):
    import torch
    for name in ("hole_bz", "plain2d"):  # the second film is smaller than the first
        case = S.EVALF_CASES[name]
        ctx, dev = _ctx(case, case["field"])
        assert ctx.len <= ctx.noalloc_max_len()
        nphi = 3 if dev.params.Nz > 1 else 2
        X = ctx.to_internal(rand_state(dev.params.n_interior, nphi, 21, 0.2, 0.1))
        A, B = X.clone(), X.clone()
        for _ in range(2):
            ctx.euler_step(A, dt)
            ctx.euler_step_noalloc(B, dt)
            torch.cuda.synchronize()
            assert torch.equal(A, B), name
        assert not torch.equal(A, X)


def test_noalloc_rejects_states_beyond_the_scratch_array():
    import torch
    from nanowire_tdgl_amd import _native
    big = _native.Context(256, 256, 64, 0.5, 0.5, 0.5, 3.0)
    assert big.len > big.noalloc_max_len() == 4 << 20
    X = torch.zeros(big.len, dtype=torch.float64, device=big.device)
    with pytest.raises(ValueError, match="exceeds the scratch array"):
        big.euler_step_noalloc(X, 0.1)
    assert not torch.any(X)
    big.close()
