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
    ("lognormal sphere", 0, 0.3, 1.6, 0.55, [1.43 + 0.01j], None),
    ("gamma sphere", 1, 0.5, 0.2, 0.55, [1.5 + 0.001j], None),
    ("gamma 3 layers", 1, 0.4, 0.15, 0.55, [1.6 + 0.1j, 1.45 + 0.01j, 1.33 + 1e-4j], [0.4, 0.7]),
]


def _perturbed(kind, p0, p1, m, fr, name, sgn, h):
    m = np.array(m, np.complex128)
    fr = None if fr is None else np.array(fr, np.float64)
    if name == "p0":
        d = h * p0
        return p0 + sgn * d, p1, m, fr, d
    if name == "p1":
        d = h * p1
        return p0, p1 + sgn * d, m, fr, d
    if name[0] == "m":
        l, c = int(name[1]), name[2]
        d = h * abs(m[l])
        m[l] += sgn * (d if c == "r" else 1j * d)
        return p0, p1, m, fr, d
    l = int(name[1])
    d = h * fr[l]
    fr[l] += sgn * d
    return p0, p1, m, fr, d


@pytest.mark.parametrize("fixed_range", [False, True])
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_oracle_jacobian_matches_central_differences(case, fixed_range):
    _, kind, p0, p1, wl, m, fr = case
    nr, na, order, h = 24, 32, 16, 2e-5
    J = oracle.ensemble_fixed_jacobian(kind, p0, p1, wl, m, fr, nr, na, order, True, fixed_range)
    L = len(m)
    names = ["p0", "p1"] + [f"m{l}{c}" for l in range(L) for c in "ri"] + \
        [f"f{l}" for l in range(L - 1)]
    assert J["d_scalars"].shape == (len(names), 11)
    rng = oracle.radius_range(kind, p0, p1) if fixed_range else None
    y = J["scalars"]
    for p, name in enumerate(names):
        # Richardson-extrapolated central differences at steps h |p| and h |p| / 2
        fds = []
        for hh in (h, h / 2):
            out = []
            for sgn in (1, -1):
                q0, q1, mm, ff, d = _perturbed(kind, p0, p1, m, fr, name, sgn, hh)
                out.append(oracle.ensemble_fixed(kind, q0, q1, wl, mm, ff, nr, na, order, True,
                                                 t_range=rng, n_nodes=J["n_nodes"]))
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
    assert np.array_equal(a["d_scalars"][2:], b["d_scalars"][2:])
    assert np.array_equal(a["d_greek"][2:], b["d_greek"][2:])
    assert not np.allclose(a["d_scalars"][:2], b["d_scalars"][:2], rtol=1e-3)
