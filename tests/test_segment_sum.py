"""The deterministic segment sum behind every row sum of the adaptive engine and the wire
format: from a tensor, from a host array and from a prepared plan it gives the same sums bit
for bit, and they are numpy's unbuffered sums to rounding."""

from __future__ import annotations

import numpy as np
import pytest

torch = pytest.importorskip("torch")


@pytest.mark.parametrize("n", [0, 1, 7, 100, 2048])
def test_plans_from_host_and_device_agree(n):
    from scatrs_amd.ensemble import segment_plan, segment_sum

    rng = np.random.default_rng(n)
    idx = rng.integers(0, 10, n)
    src = torch.from_numpy(rng.standard_normal((n, 3, 2)))
    a = segment_sum(torch.from_numpy(idx), src, 10)
    b = segment_sum(idx, src, 10)
    plan = segment_plan(idx, "cpu")
    c = segment_sum(plan, src, 10)
    assert torch.equal(a, b) and torch.equal(a, c)
    assert np.array_equal(plan.index.numpy(), idx)
    ref = np.zeros((10, 3, 2))
    np.add.at(ref, idx, src.numpy())
    assert np.allclose(a.numpy(), ref, rtol=1e-13, atol=1e-13)
