"""The oracle's forward-mode ensemble Jacobian (orc_ensemble_fixed_jacobian) against central
differences of the oracle's forward pass: with the nodes moving with the distribution's
parameters (the derivative of the fixed discrete path) and with the ln r range held fixed (the
reference's definition, docs/theory/ensembles.md:150,339-340), for spheres and layered
spheres. N per node is held at the base point's value so the differences see a smooth path."""

from __future__ import annotations

import numpy as np
import pytest

import oracle

CASES = [
    ("lognormal sphere", 0, 0.3, 1.6, 0.55, [1.43 + 0.01j], None, 0.0),
    ("gamma sphere", 1, 0.5, 0.2, 0.55, [1.5 + 0.001j], None, 0.0),
    ("gamma 3 layers", 1, 0.4, 0.15, 0.55, [1.6 + 0.1j, 1.45 + 0.01j, 1.33 + 1e-4j], [0.4, 0.7],
     0.0),
    ("modified gamma sphere", 2, 2.0, 3.0, 0.55, [1.45 + 0.01j], None, 1.2),
    ("power law sphere", 3, 3.0, 0.05, 0.55, [1.5 + 0.001j], None, 2.0),
    ("monodisperse 2 layers", 4, 0.8, 0.0, 0.55, [1.6 + 0.1j, 1.4 + 0.01j], [0.5], 0.0),
]
# The distribution's differentiable parameters among (p0, p1, p2) per kind: the power law's
# limits are fixed (distribution.rs:40-47,189-198); the others' columns are 0.
FREE = {0: 2, 1: 2, 2: 3, 3: 1, 4: 1}


def _perturbed(p, m, fr, name, sgn, h):
    p = list(p)
    m = np.array(m, np.complex128)
    fr = None if fr is None else np.array(fr, np.float64)
    if name[0] == "p":
        k = int(name[1])
        d = h * abs(p[k])
        p[k] += sgn * d
        return p, m, fr, d
    if name[0] == "m":
        l, c = int(name[1]), name[2]
        d = h * abs(m[l])
        m[l] += sgn * (d if c == "r" else 1j * d)
        return p, m, fr, d
    l = int(name[1])
    d = h * fr[l]
    fr[l] += sgn * d
    return p, m, fr, d


@pytest.mark.parametrize("fixed_range", [False, True])
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_oracle_jacobian_matches_central_differences(case, fixed_range):
    _, kind, p0, p1, wl, m, fr, p2 = case
    nr, na, order = 24, 32, 16
    # gamma enters through r^gamma over 15 units of ln r: a smaller step for its differences
    h = 4e-6 if kind == oracle.MODIFIED_GAMMA else 2e-5
    J = oracle.ensemble_fixed_jacobian(kind, p0, p1, wl, m, fr, nr, na, order, True, fixed_range,
                                       p2=p2)
    L = len(m)
    names = ["p0", "p1", "p2"] + [f"m{l}{c}" for l in range(L) for c in "ri"] + \
        [f"f{l}" for l in range(L - 1)]
    assert J["d_scalars"].shape == (len(names), 11)
    # a monodisperse law's range is its radius: its nodes move under either definition
    hold = fixed_range and kind != oracle.MONODISPERSE
    rng = oracle.radius_range(kind, p0, p1, p2) if hold else None
    y = J["scalars"]
    for p, name in enumerate(names):
        if p < 3 and p >= FREE[kind]:
            assert not J["d_scalars"][p].any() and not J["d_greek"][p].any(), name
            continue
        # Richardson-extrapolated central differences at steps h |p| and h |p| / 2
        fds = []
        for hh in (h, h / 2):
            out = []
            for sgn in (1, -1):
                q, mm, ff, d = _perturbed((p0, p1, p2), m, fr, name, sgn, hh)
                out.append(oracle.ensemble_fixed(kind, q[0], q[1], wl, mm, ff, nr, na, order, True,
                                                 t_range=rng, n_nodes=J["n_nodes"], p2=q[2]))
            fds.append({k: (out[0][k] - out[1][k]) / (2 * d) for k in ("scalars", "greek",
                                                                       "phase")})
        # log-sensitivities: d y / d ln p relative to |y| (the last step is d = (h/2) |p|)
        scale_p = 2 * d / h
        for key, dkey, idx in (("scalars", "d_scalars", [0, 1, 2, 3, 4, 5, 6, 8, 9, 10]),
                               ("greek", "d_greek", None), ("phase", "d_phase", None)):
            fd = (4.0 * fds[1][key] - fds[0][key]) / 3.0
            jac = J[dkey][p]
            if idx is not None:
                err = scale_p * np.abs(fd - jac)[idx] / np.abs(y[idx])
            else:
                ref = J[key]
                err = scale_p * np.abs(fd - jac).max() / np.abs(ref).max()
            assert np.max(err) < 1e-8, (name, key, err)


def test_fixed_range_and_moving_nodes_agree_for_the_index():
    """The refractive index does not move the nodes: both definitions give the same m
    derivatives; they differ for the distribution's parameters by node motion."""
    a = oracle.ensemble_fixed_jacobian(0, 0.3, 1.6, 0.55, [1.43 + 0.01j], None, 24, 32, 16,
                                       True, False)
    b = oracle.ensemble_fixed_jacobian(0, 0.3, 1.6, 0.55, [1.43 + 0.01j], None, 24, 32, 16,
                                       True, True)
    assert np.array_equal(a["d_scalars"][3:], b["d_scalars"][3:])
    assert np.array_equal(a["d_greek"][3:], b["d_greek"][3:])
    assert not np.allclose(a["d_scalars"][:2], b["d_scalars"][:2], rtol=1e-3)
