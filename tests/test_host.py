"""Host logic: workloads, cost accounting and the Python API's checks (CPU only)."""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from scatrs_amd import workloads
from scatrs_amd.errors import InvalidInputError


def test_workloads_are_deterministic():
    a = workloads.spheres(1000)
    b = workloads.spheres(1000)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    x, m = workloads.spheres(1 << 20)
    assert x.min() >= 0.1 and x.max() <= 1000.0
    assert m.real.min() >= 1.3 and m.real.max() <= 1.7
    assert m.imag.min() >= 1e-8 and m.imag.max() <= 1e-1
    kind, p0, p1, wl, m = workloads.lognormals(8, 4)
    assert p0.size == 32 and np.all(p1 >= 1.2) and np.all(p1 <= 2.5)


def test_cost_accounting_follows_the_oracle_start_index():
    x, m = workloads.spheres(200, seed=5)
    n = workloads.default_n_max(x)
    for i in range(x.size):
        assert n[i] == oracle.default_n_max(x[i])
        assert workloads.start_index(n[i] + 1, abs(m[i] * x[i])) == oracle.start_index(
            n[i] + 1, abs(m[i] * x[i]))
    terms, flops = workloads.mie_cost(x, m)
    assert terms == n.sum() and flops > 100 * terms


def test_layered_sphere_checks():
    from scatrs_amd.particles import LayeredSphere

    with pytest.raises(InvalidInputError):
        LayeredSphere([0.5, 0.4], [1.5, 1.6, 1.7])
    with pytest.raises(InvalidInputError):
        LayeredSphere([0.5], [1.5, 1.6, 1.7])
    s = LayeredSphere.coated(1.9 + 0.7j, 1.43, core_fraction=0.6)
    assert s.layers == 2
    assert s.parameter_names()[-1] == "layer0.radius_fraction"


def test_gpu_entry_points_fail_loudly_without_a_gpu():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    from scatrs_amd import NativeLibraryError
    from scatrs_amd.mie import efficiencies

    with pytest.raises((NativeLibraryError, RuntimeError, AssertionError)):
        efficiencies(np.array([1.0]), np.array([1.5 + 0j]))
