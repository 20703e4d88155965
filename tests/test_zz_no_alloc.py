"""scatrs_mie_efficiencies_noalloc through the C ABI: the same results as
scatrs_mie_efficiencies on two calls in a row (the second smaller, so it runs on what the first
left in the shared scratch array), with and without the Jacobian, and its size limits."""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _call(name, x, m, derivatives, n_max=0):
    from scatrs_amd import _native

    ctx = _native.context(0)
    n = x.numel()
    values = torch.full((5, n), float("nan"), dtype=torch.float64, device=x.device)
    jac = (torch.full((3, 5, n), float("nan"), dtype=torch.float64, device=x.device)
           if derivatives else None)
    terms = C.c_int64(-1)
    rc = getattr(ctx.lib, name)(ctx.handle, C.c_void_p(x.data_ptr()), C.c_void_p(m.data_ptr()),
                                n, n_max, 0, C.c_void_p(values.data_ptr()),
                                C.c_void_p(jac.data_ptr()) if jac is not None else None,
                                C.byref(terms))
    ctx.check(rc)
    torch.cuda.synchronize()
    return values, jac, terms.value


def _spheres(n, seed):
    rng = np.random.default_rng(seed)
    x = np.exp(rng.uniform(np.log(0.1), np.log(1000.0), n))
    m = rng.uniform(1.3, 1.7, n) + 1j * np.exp(rng.uniform(np.log(1e-8), np.log(1e-1), n))
    return (torch.from_numpy(x).to("cuda"), torch.from_numpy(m).to("cuda"))


@pytest.mark.parametrize("derivatives", [False, True])
def test_noalloc_equals_the_allocating_entry_point(derivatives):
    for n, seed in ((200_000, 1), (3_001, 2)):  # the second call smaller than the first
        print(f"noalloc: {n} spheres, seed {seed}, derivatives {derivatives}", flush=True)
        x, m = _spheres(n, seed)
        v_ref, j_ref, t_ref = _call("scatrs_mie_efficiencies", x, m, derivatives)
        v, j, t = _call("scatrs_mie_efficiencies_noalloc", x, m, derivatives)
        assert t == t_ref > 0
        # the same kernels on the same inputs: bitwise the same results
        assert torch.equal(v, v_ref)
        if derivatives:
            assert torch.equal(j, j_ref)


def test_noalloc_limits_and_errors():
    from scatrs_amd.errors import InvalidInputError

    x, m = _spheres(64, 3)
    # the allocating entry point takes any count; the scratch holds at most 2^20 spheres
    big_x = torch.full(((1 << 20) + 1,), 1.0, dtype=torch.float64, device="cuda")
    big_m = torch.full(((1 << 20) + 1,), 1.5 + 0.01j, dtype=torch.complex128, device="cuda")
    with pytest.raises(InvalidInputError, match="at most 1048576 spheres"):
        _call("scatrs_mie_efficiencies_noalloc", big_x, big_m, False)
    # an invalid element maps to the reference's message, as the allocating entry point's does
    xb = x.clone()
    xb[5] = -1.0
    with pytest.raises(InvalidInputError, match="element 5"):
        _call("scatrs_mie_efficiencies_noalloc", xb, m, False)
    # and the scratch still works after the errors
    v_ref, _, _ = _call("scatrs_mie_efficiencies", x, m, False)
    v, _, _ = _call("scatrs_mie_efficiencies_noalloc", x, m, False)
    assert torch.equal(v, v_ref)
