"""Captured ensemble calls (scatrs_ensemble_fixed_graph) through the C ABI: every replay equals
the eager call bit for bit, with eager work and synchronizations between the replays; inputs
changed in place (same truncations) are recomputed; a change of a node's truncation is flagged;
the graph keeps working after the context allocates afresh."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def _sulfate():
    from scatrs_amd import workloads

    kind, p0, p1, wl, m = workloads.sulfate()
    return [_dev(v) for v in (kind, p0, p1, wl, m)]


def test_replays_equal_the_eager_call():
    """The sulfate lognormal at three wavelengths, 128 radius nodes, Greek to 32: each of
    several replays (outputs poisoned with NaN before each) equals the eager call bitwise."""
    from scatrs_amd.ensemble import EnsembleGraph, ensemble_fixed_device

    t = _sulfate()
    kw = dict(n_radius=128, n_angles=64, order=32)
    sc, gr, ph, terms = ensemble_fixed_device(*t, **kw)
    g = EnsembleGraph(*t, **kw)
    assert g.terms == terms
    assert torch.equal(g.scalars, sc) and torch.equal(g.greek, gr) and torch.equal(g.phase, ph)
    for rep in range(4):
        g.scalars.fill_(float("nan"))
        g.greek.fill_(float("nan"))
        g.phase.fill_(float("nan"))
        # eager work of the program's own between replays, with synchronizations
        sc2, gr2, _, _ = ensemble_fixed_device(*t, **kw)
        torch.cuda.synchronize()
        g.launch()
        g.status()
        torch.cuda.synchronize()
        assert torch.equal(g.scalars, sc), rep
        assert torch.equal(g.greek, gr), rep
        assert torch.equal(g.phase, ph), rep
        assert torch.equal(sc2, sc)
    g.close()


def test_inputs_changed_in_place_are_recomputed():
    """A new refractive index keeps every node's truncation: the replay equals the eager call
    with the new index. A new wavelength changes the truncations: the status says so."""
    from scatrs_amd import InvalidInputError
    from scatrs_amd.ensemble import EnsembleGraph, ensemble_fixed_device

    t = _sulfate()
    kw = dict(n_radius=128, n_angles=64, order=32, want_phase=False)
    g = EnsembleGraph(*t, **kw)
    g.inputs[4].copy_(_dev(np.full(3, 1.50 + 0.01j)))
    for _ in range(2):
        g.launch()
        g.status()
    t2 = list(t)
    t2[4] = _dev(np.full(3, 1.50 + 0.01j))
    sc, gr, _, _ = ensemble_fixed_device(*t2, **kw)
    torch.cuda.synchronize()
    assert torch.equal(g.scalars, sc) and torch.equal(g.greek, gr)
    g.inputs[3].copy_(_dev(np.array([0.2, 0.3, 0.4])))
    g.launch()
    with pytest.raises(InvalidInputError, match="changed the number of terms"):
        g.status()
    g.close()


@pytest.mark.parametrize("kernel", [0, 2])
def test_layered_graph_survives_later_calls(kernel):
    """A 3-layer gamma batch captured with either angular kernel; later calls on the context
    (which then allocates afresh, here a larger batch) leave the graph's replays unchanged."""
    from scatrs_amd import _native, workloads
    from scatrs_amd.ensemble import EnsembleGraph, ensemble_fixed_device

    W = workloads.layered(n_dist=16, n_wl=4, layers=3)
    kind, p0, p1, wl, fr, m = W[3]
    t = [_dev(v) for v in (kind, p0, p1, wl, m)]
    tfr = _dev(fr)
    kw = dict(n_radius=64, n_angles=96, order=24)
    ctx = _native.context(0)
    ctx.set_angular_kernel(kernel)
    try:
        sc, gr, ph, _ = ensemble_fixed_device(*t, tfr, **kw)
        g = EnsembleGraph(*t, tfr, **kw)
        big = workloads.layered(n_dist=64, n_wl=8, layers=3)[3]
        tb = [_dev(v) for v in (big[0], big[1], big[2], big[3], big[5])]
        ensemble_fixed_device(*tb, _dev(big[4]), n_radius=96, n_angles=128, order=40)
        for _ in range(3):
            g.scalars.zero_()
            g.launch()
            g.status()
            torch.cuda.synchronize()
            assert torch.equal(g.scalars, sc) and torch.equal(g.greek, gr)
            assert torch.equal(g.phase, ph)
        g.close()
    finally:
        ctx.set_angular_kernel(0)


def test_graph_errors():
    """Inputs the eager call rejects are rejected the same way, and no graph is made."""
    from scatrs_amd import InvalidInputError
    from scatrs_amd.ensemble import EnsembleGraph

    t = _sulfate()
    t[2] = _dev(np.array([1.6, 0.9, 1.6]))  # sigma_g < 1
    with pytest.raises(InvalidInputError, match="ensemble 1"):
        EnsembleGraph(*t, n_radius=32, n_angles=32, order=8)
