"""The adaptive engine (scatrs_amd.adaptive) and its node evaluator scatrs_ensemble_nodes on
the GPU: the caller's nodes reproduce the fixed path; adaptive results against the oracle's
converged integral over the same range (2048-8192 Gauss-Legendre nodes); refinement where
the fixed rule fails, extension of the range, the panel limit, and the wire format's
adaptive option."""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from scatrs_amd.errors import InvalidInputError

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def _fixed_nodes(kind, p, n):
    """The fixed path's nodes (r, W) from the host helpers: Gauss-Legendre in ln r over the
    fixed range, W = w half u r."""
    from scatrs_amd.adaptive import log_density, size_range

    lo, hi = size_range(kind, p)
    x, w = np.polynomial.legendre.leggauss(n)
    half, mid = 0.5 * (hi - lo), 0.5 * (hi + lo)
    r = np.exp(mid + half * x)
    lu = log_density(kind, p, r)
    return np.stack([r, w * half * np.exp(lu - lu.max()) * r], -1)


def test_caller_nodes_reproduce_the_fixed_path():
    from scatrs_amd.ensemble import ensemble_fixed_device, ensemble_nodes_device

    cases = [(0, (0.3, 1.6, 0.0), 0.55, [1.45 + 0.01j], None),
             (1, (0.5, 0.2, 0.0), 0.9, [1.52 + 1e-3j], None),
             (2, (2.0, 3.0, 1.2), 0.7, [1.6 + 0.1j, 1.35 + 1e-5j], [0.6]),
             (3, (3.5, 0.1, 1.5), 1.1, [1.33 + 1e-4j], None)]
    n, na, L = 48, 40, 12
    for kind, p, wl, m, fr in cases:
        nodes = _fixed_nodes(kind, p, n)[None]
        mm = np.array(m)[None] if len(m) > 1 else np.array(m)
        f = None if fr is None else _dev(np.array([fr]))
        a = ensemble_nodes_device(_dev(nodes), _dev(np.array([wl])), _dev(mm), f, n_angles=na,
                                  order=L)
        b = ensemble_fixed_device(_dev(np.array([kind], np.int32)), _dev(np.array([p[0]])),
                                  _dev(np.array([p[1]])), _dev(np.array([wl])), _dev(mm), f,
                                  p2=_dev(np.array([p[2]])), n_radius=n, n_angles=na, order=L)
        for x, y in zip(a[:3], b[:3]):
            x, y = x.cpu().numpy(), y.cpu().numpy()
            assert np.max(np.abs(x - y)) <= 1e-12 * np.max(np.abs(y)), kind
        assert a[3] == b[3]
    with pytest.raises(InvalidInputError, match="node 3: the radius must be finite"):
        bad = _fixed_nodes(0, (0.3, 1.6, 0.0), 8)[None].copy()
        bad[0, 3, 0] = -1.0
        ensemble_nodes_device(_dev(bad), _dev(np.array([0.5])), _dev(np.array([1.5 + 0j])))


def _adaptive(kind, p, wl, m, rtol, **kw):
    from scatrs_amd.adaptive import parts_adaptive
    from scatrs_amd.particles import Sphere

    K, G, terms, conv, panels, ranges, na = parts_adaptive(
        [(kind, *p, Sphere(m))], np.asarray(wl, np.float64), order=kw.pop("order", 16),
        n_angles=kw.pop("n_angles", None), rtol=rtol, max_panels=kw.pop("max_panels", 4096),
        allow_unvalidated=False, want_phase=False, **kw)
    return K[0][0].cpu().numpy(), K[1][0].cpu().numpy(), conv[0], panels[0], ranges[0], na


def _law(kind, p):
    """The number fraction in [r_lo, r_hi] and <r^2>, <r^3> of a lognormal or gamma law in
    closed form (distribution.rs:221-268)."""
    from scipy.special import gammainc
    from scipy.stats import norm

    if kind == 0:
        s = np.log(p[1])
        frac = lambda lo, hi: norm.cdf((np.log(hi / p[0])) / s) - norm.cdf(  # noqa: E731
            np.log(lo / p[0]) / s)
        return frac, p[0] ** 2 * np.exp(2 * s * s), p[0] ** 3 * np.exp(4.5 * s * s)
    k, th = (1 - 2 * p[1]) / p[1], p[0] * p[1]
    frac = lambda lo, hi: gammainc(k, hi / th) - gammainc(k, lo / th)  # noqa: E731
    return frac, th ** 2 * k * (k + 1), th ** 3 * k * (k + 1) * (k + 2)


@pytest.mark.parametrize("case", [
    (0, (0.3, 1.6, 0.0), [0.55, 1.2], 1.5 + 0.05j, 1e-8),   # absorbing lognormal
    (0, (0.08, 1.6, 0.0), [0.55], 1.43 + 1e-8j, 1e-6),      # resonant: the fixed rule fails
    (0, (0.002, 2.0, 0.0), [1.0], 1.5 + 1e-3j, 1e-7),       # Rayleigh: C_sca ~ r^6 tail
    (1, (0.6, 0.25, 0.0), [0.8], 1.38 + 2e-3j, 1e-7),       # gamma
])
def test_adaptive_matches_the_converged_integral_over_its_range(case):
    """normalized=False averages against the oracle's fixed path with t_range set to the
    engine's range and 4096 nodes (8192 for the resonant case): the oracle divides by the
    number inside that range, the engine by the whole law's (its moment integration), so the
    oracle's averages times the law's number fraction inside the range; <r^2> and <r^3>
    against the closed forms. Within a few rtol."""
    from scatrs_amd.adaptive import integration_range

    kind, p, wl, m, rtol = case
    sc, gr, conv, panels, ranges, na = _adaptive(kind, p, wl, m, rtol)
    assert conv.all(), panels
    n_ref = 8192 if m.imag < 1e-6 else 4096
    frac, r2, r3 = _law(kind, p)
    for i, w in enumerate(wl):
        t = np.log(ranges[i])
        # an angle rule exact for every node (the engine's rules are, band by band)
        nq = oracle.default_n_max(2 * np.pi * ranges[i][1] / w) + 9
        ref = oracle.ensemble_fixed(kind, p[0], p[1], w, m, n_radius=n_ref, n_angles=nq,
                                    order=16, normalized=False, t_range=t, p2=p[2])
        f = frac(*ranges[i])
        rs = ref["scalars"] * f
        for k in (0, 1):  # C_ext, C_sca
            assert abs(sc[i, k] - rs[k]) <= 5 * rtol * abs(rs[k]), (k, sc[i, k], rs[k])
        assert abs(sc[i, 5] * sc[i, 1] - rs[5] / f * rs[1]) <= 5 * rtol * abs(rs[1])
        assert abs(sc[i, 2] - rs[2]) <= 5 * rtol * abs(rs[0])
        assert abs(sc[i, 9] - np.pi * r2) <= 5 * rtol * np.pi * r2
        assert abs(sc[i, 10] - 4 * np.pi / 3 * r3) <= 5 * rtol * 4 * np.pi / 3 * r3
        for c in range(6):  # every Greek coefficient, exact projections on both sides
            ref_c = ref["greek"][c] * f
            assert np.max(np.abs(gr[i, c] - ref_c)) <= 20 * rtol * np.max(np.abs(ref["greek"][0]))
    if m.imag < 1e-6:
        assert panels.max() > 4  # refined beyond the initial panels
    if p[0] == 0.002:
        # C_sca ~ r^6 outweighs the area: the range grew past the area rule's upper end
        _, t_hi, _ = integration_range(kind, p, rtol)
        assert np.log(ranges[0][1]) > t_hi + 1e-9


def test_panel_limit_leaves_the_task_unconverged():
    sc, _, conv, panels, _, _ = _adaptive(0, (0.08, 1.6, 0.0), [0.55], 1.43 + 1e-8j, 1e-10,
                                          max_panels=6)
    assert not conv.any() and panels.max() == 6
    assert np.all(np.isfinite(sc))


def test_unresolvable_derivatives_are_flagged():
    """The reference's case (scatrs/tests/ensemble.rs:664-680): without absorption the
    derivative with respect to Im m has a spike at every resonance, which pass 1 cannot
    resolve within 512 panels: the result is unconverged, at the panel limit; the values
    alone converge."""
    from scatrs_amd import _wire

    items = [(("gamma", 2.0, 0.1), ("sphere", np.array([1.33 + 0j])), 1.0)]
    wl = np.array([0.65])
    names = _wire.ensemble_parameters(items, False)
    d = _wire.ensemble_compute(items, wl, False, None, False, np.zeros(0), True,
                               [names.index("m_imag")], 1e-6, 512, False, False)
    assert not d["converged"][0] and d["panels"][0] == 512
    v = _wire.ensemble_compute(items, wl, False, None, False, np.zeros(0), True, [], 1e-6, 512,
                               False, False)
    assert v["converged"][0] and v["panels"][0] < 512
    assert np.all(np.isfinite(d["values"])) and np.all(np.isfinite(d["d"]))


def test_ensemble_mixture_and_wire_use_the_engine():
    """Ensemble.compute(rtol=...) and the wire format's adaptive option (its differentiable
    pass 2, summed in another order) give the same values;
    a monodisperse law is the sphere; a mixture converges part by part."""
    from scatrs_amd import _wire
    from scatrs_amd.ensemble import Ensemble, Mixture
    from scatrs_amd.particles import Sphere
    from scatrs_amd.size import Gamma, LogNormal, Monodisperse

    wl = np.array([0.5, 1.0])
    m = np.array([1.45 + 0.01j, 1.44 + 0.02j])
    r = Ensemble(Sphere(m), LogNormal(0.2, 1.5)).compute(wl, greek=8, rtol=1e-7, mu=[-1, 0, 1])
    assert r.converged.all() and r.panels.min() >= 4
    raw = _wire.ensemble_compute([(("lognormal", 0.2, 1.5), ("sphere", m), 1.0)], wl, False, 8,
                                 False, np.array([-1.0, 0.0, 1.0]), True, [], 1e-7, 4096, False,
                                 False, adaptive=True)
    assert np.array_equal(raw["converged"], r.converged)
    assert np.array_equal(raw["panels"], r.panels.astype(np.uint64))
    assert np.allclose(raw["values"][0], r.c_ext, rtol=1e-13, atol=0)
    assert np.allclose(raw["phase"][0], r.phase.f11, rtol=1e-13, atol=0)
    assert np.all(raw["errors"][0] <= 1e-6 * raw["values"][0])
    mono = Ensemble(Sphere(m), Monodisperse(0.4)).compute(wl, rtol=1e-6)
    q, _ = oracle.mie_efficiencies_batch(2 * np.pi * 0.4 / wl, m)
    assert np.allclose(mono.c_ext, np.pi * 0.16 * q[0], rtol=1e-12, atol=0)
    mix = Mixture(((2.0, Ensemble(Sphere(m), LogNormal(0.2, 1.5))),
                   (1.0, Ensemble(Sphere(m), Gamma(0.5, 0.1))))).compute(wl, greek=8, rtol=1e-6)
    assert mix.converged.all() and np.all(mix.panels >= 8)
    assert np.allclose(mix.greek.alpha1[:, 0], 1.0, rtol=0, atol=1e-14)


def _nodes_case():
    from scatrs_amd.adaptive import log_density, size_range

    kind, p, n = 1, (0.5, 0.2, 0.0), 24
    lo, hi = size_range(kind, p)
    x, w = np.polynomial.legendre.leggauss(n)
    half, mid = 0.5 * (hi - lo), 0.5 * (hi + lo)
    r = np.exp(mid + half * x)
    lu, d = log_density(kind, p, r, derivatives=True)
    return kind, p, r, w * half * np.exp(lu - lu.max()) * r, d


def test_caller_nodes_vjp_is_the_fixed_range_vjp():
    """scatrs_ensemble_nodes_vjp on the fixed path's nodes, with d ln W / d p from the density,
    equals scatrs_ensemble_fixed_vjp with the range held fixed, every column (spheres, and
    coated spheres with phase cotangents)."""
    from scatrs_amd.ensemble import ensemble_fixed_vjp_device, ensemble_nodes_vjp_device

    kind, p, r, W, d = _nodes_case()
    rng = np.random.default_rng(3)
    L, na = 10, 32
    bs, bg = rng.normal(size=(1, 11)), rng.normal(size=(1, 6, L + 1))
    bs[0, 7] = 0.0
    m = np.array([1.5 + 0.02j])
    g1, t1 = ensemble_nodes_vjp_device(
        _dev(np.stack([r, W], -1)[None]), _dev(d[None]), _dev(np.array([0.7])), _dev(m),
        bar_scalars=_dev(bs), bar_greek=_dev(bg), n_angles=na, order=L)
    g2, t2 = ensemble_fixed_vjp_device(
        _dev(np.array([kind], np.int32)), _dev(np.array([p[0]])), _dev(np.array([p[1]])),
        _dev(np.array([0.7])), _dev(m), p2=_dev(np.array([p[2]])), bar_scalars=_dev(bs),
        bar_greek=_dev(bg), n_radius=r.size, n_angles=na, order=L, fixed_range=True)
    g1, g2 = g1.cpu().numpy(), g2.cpu().numpy()
    assert np.max(np.abs(g1 - g2)) <= 1e-11 * np.max(np.abs(g2)), (g1, g2)
    assert t1 == t2
    # coated spheres: every layer's m and the radius fraction
    ml = np.array([[1.6 + 0.1j, 1.38 + 1e-4j]])
    fr = _dev(np.array([[0.55]]))
    bp = rng.normal(size=(1, 6, na))
    g1, _ = ensemble_nodes_vjp_device(
        _dev(np.stack([r, W], -1)[None]), _dev(d[None]), _dev(np.array([0.9])), _dev(ml), fr,
        bar_scalars=_dev(bs), bar_greek=_dev(bg), bar_phase=_dev(bp), n_angles=na, order=L)
    g2, _ = ensemble_fixed_vjp_device(
        _dev(np.array([kind], np.int32)), _dev(np.array([p[0]])), _dev(np.array([p[1]])),
        _dev(np.array([0.9])), _dev(ml), fr, p2=_dev(np.array([p[2]])), bar_scalars=_dev(bs),
        bar_greek=_dev(bg), bar_phase=_dev(bp), n_radius=r.size, n_angles=na, order=L,
        fixed_range=True)
    g1, g2 = g1.cpu().numpy(), g2.cpu().numpy()
    assert g1.shape == (1, 3 + 4 + 1)
    assert np.max(np.abs(g1 - g2)) <= 1e-11 * np.max(np.abs(g2)), (g1, g2)


@pytest.mark.parametrize("layers", [1, 2, 3])
def test_caller_nodes_forward_mode_is_the_transposed_vjp(layers):
    """Forward mode of ensemble_nodes (scatrs_ensemble_nodes_tangent for the particle's
    directions, a reweighted call for the distribution's) against the reverse mode
    (scatrs_ensemble_nodes_vjp) by the transposition identity c . (J e_j) = (J^T c) . e_j, for
    every parameter e_j alone (the distribution's three, Re m and Im m of each layer, the radius
    fractions) with random cotangents over the scalars, Greek coefficients and phase matrix;
    and replicas with different tangents in one call."""
    from torch.autograd import forward_ad

    from scatrs_amd.ensemble import ensemble_nodes, ensemble_nodes_vjp_device

    kind, p, r, W, d = _nodes_case()
    rng = np.random.default_rng(11 + layers)
    L, na, wl = 10, 32, 0.8
    ms = [1.6 + 0.1j, 1.45 + 0.02j, 1.38 + 1e-3j][-layers:]
    m = np.array([ms]) if layers > 1 else np.array([ms[0]])
    fr = _dev(np.array([[0.35, 0.7][-(layers - 1):]])) if layers > 1 else None
    bs, bg, bp = rng.normal(size=(1, 11)), rng.normal(size=(1, 6, L + 1)), rng.normal(
        size=(1, 6, na))
    bs[0, 7] = 0.0
    nodes, dl, wlt, md = (_dev(np.stack([r, W], -1)[None]), _dev(d[None]),
                          _dev(np.array([wl])), _dev(m))
    g, _ = ensemble_nodes_vjp_device(nodes, dl, wlt, md, fr, bar_scalars=_dev(bs),
                                     bar_greek=_dev(bg), bar_phase=_dev(bp), n_angles=na,
                                     order=L, normalized=False)
    g = g.cpu().numpy()[0]
    P = 3 + 2 * layers + (layers - 1)
    assert g.shape == (P,)
    p0 = _dev(np.array([p]))
    mr, mi = md.real.contiguous(), md.imag.contiguous()

    def tangent(j, k=1):
        """Outputs' tangents along parameter e_j (rows repeated k times, row 0 seeded)."""
        e = np.zeros(P)
        e[j] = 1.0
        rep = lambda a: None if a is None else a.repeat_interleave(k, 0)  # noqa: E731
        seeds = [e[:3], e[3:3 + 2 * layers:2], e[4:4 + 2 * layers:2], e[3 + 2 * layers:]]
        t = [np.zeros((k,) + a.shape) for a in seeds]
        for a, s in zip(t, seeds):
            a[0] = s
        if layers == 1:
            t[1], t[2] = t[1][:, 0], t[2][:, 0]
        with forward_ad.dual_level():
            dual = lambda a, s: forward_ad.make_dual(rep(a), _dev(s))  # noqa: E731
            out = ensemble_nodes(
                dual(p0, t[0]), dual(mr, t[1]), dual(mi, t[2]),
                None if fr is None else dual(fr, t[3]), rep(nodes), rep(dl), rep(wlt),
                n_angles=na, order=L, normalized=False, replicas=k)
            return [forward_ad.unpack_dual(o).tangent.cpu().numpy() for o in out]

    scale = np.abs(g).max()
    for j in range(P):
        ts, tg, tp = tangent(j)
        lhs = (bs * ts).sum() + (bg * tg).sum() + (bp * tp).sum()
        assert abs(lhs - g[j]) <= 1e-9 * scale, (j, lhs, g[j])
    # replicas: row 0 carries the tangent, row 1 none
    ts, tg, tp = tangent(P - 1, k=2)
    lhs = (bs * ts[:1]).sum() + (bg * tg[:1]).sum() + (bp * tp[:1]).sum()
    assert abs(lhs - g[P - 1]) <= 1e-9 * scale
    assert np.all(ts[1] == 0) and np.all(tg[1] == 0) and np.all(tp[1] == 0)


def test_adaptive_forward_and_reverse_derivatives_agree(monkeypatch):
    """The wire's derivative columns in forward mode (one replica per parameter) and in reverse
    mode (one replica per output component) on the same panels: a two-component mixture with a
    coated sphere, the phase matrix at three cosines and Greek coefficients; they agree to
    1e-9 of each column's scale."""
    import functools

    from scatrs_amd import _wire

    wl = np.array([0.55, 0.9])
    items = [(("lognormal", 0.25, 1.7), ("sphere", np.full(2, 1.5 + 0.01j)), 1.5),
             (("gamma", 0.4, 0.12), ("layered", np.array([0.6]),
                                     np.array([[1.6 + 0.1j] * 2, [1.38 + 1e-3j] * 2])), 1.0)]
    names = _wire.ensemble_parameters(items, True)
    idx = list(range(len(names)))

    def run():
        return _wire.ensemble_compute(items, wl, True, 5, False, np.array([-1.0, 0.2, 1.0]),
                                      True, idx, 1e-7, 4096, False, False)

    fwd = run()
    monkeypatch.setattr(_wire, "_run_adaptive",
                        functools.partial(_wire._run_adaptive, forward=False))
    rev = run()
    # the same panels; the values' calls group the rows differently (one call for both rules)
    assert np.array_equal(fwd["panels"], rev["panels"])
    assert np.allclose(fwd["values"], rev["values"], rtol=1e-12, atol=0)
    for key in ("d", "greek_d", "phase_d"):
        a, b = fwd[key], rev[key]
        assert a.shape == b.shape == (len(names),) + a.shape[1:]
        for j in range(len(names)):
            scale = max(np.abs(b[j]).max(), 1e-300)
            assert np.abs(a[j] - b[j]).max() <= 1e-9 * scale, (key, names[j])


def test_adaptive_derivative_columns_match_central_differences():
    """The wire's adaptive option with derivatives: d and greek_d of a lognormal's median
    radius, width, Re m and Im m, and of a mixture's weight, against central differences of the
    adaptive engine itself at rtol 1e-10 (h = 1e-4 relative: the re-refined panels and the
    moving range change the differences by far less than the tolerance)."""
    from scatrs_amd import _wire

    wl = np.array([0.6, 1.1])
    L = 6

    def items(rg=0.3, sg=1.6, mr=1.45, mi=0.01, w=2.0):
        m = np.full(2, mr + 1j * mi)
        return [(("lognormal", rg, sg), ("sphere", m), w),
                (("gamma", 0.5, 0.15), ("sphere", np.full(2, 1.4 + 1e-3j)), 1.0)]

    def run(its, derivs):
        return _wire.ensemble_compute(its, wl, True, L, False, np.zeros(0), True, derivs,
                                      1e-10, 4096, False, False, adaptive=True)

    names = _wire.ensemble_parameters(items(), True)
    want = ["c0.median_radius", "c0.width", "c0.m_real", "c0.m_imag", "c0.weight"]
    raw = run(items(), [names.index(n) for n in want])
    assert raw["converged"].all()
    base = dict(rg=0.3, sg=1.6, mr=1.45, mi=0.01, w=2.0)
    for k, key in enumerate(["rg", "sg", "mr", "mi", "w"]):
        h = 1e-4 * base[key]
        up = run(items(**dict(base, **{key: base[key] + h})), [])
        dn = run(items(**dict(base, **{key: base[key] - h})), [])
        fd = (up["values"] - dn["values"]) / (2 * h)
        sel = [0, 1, 2, 4, 5, 8, 9, 10]
        if key in ("mr", "mi"):
            # the moments do not depend on m: exactly 0 here, while the re-refined panels of
            # the differences move them by ~1e-12
            assert np.all(raw["d"][k][[8, 9, 10]] == 0.0)
            sel = [0, 1, 2, 4, 5]
        scale = np.max(np.abs(fd[sel]), 1, keepdims=True)
        err = np.abs(raw["d"][k][sel] - fd[sel]) / scale
        assert err.max() <= 2e-5, (want[k], err.max())
        fdg = (up["greek"] - dn["greek"]) / (2 * h)
        errg = np.abs(raw["greek_d"][k] - fdg) / np.max(np.abs(fdg))
        assert errg.max() <= 2e-5, (want[k], errg.max())
        assert np.all(raw["d_errors"][k][sel] <= 1e-6 * scale)


def test_adaptive_layered_and_tabulated():
    """Coated spheres over a gamma law and a tabulated law (segments with fixed supports) by
    the adaptive engine, against the fixed path on many nodes: the gamma law's fixed range
    holds all but ~1e-9 of its number and area, so the fixed path on 2048 nodes is the
    converged integral to ~1e-8; a tabulated law is integrated exactly segment by segment."""
    from scatrs_amd.ensemble import Ensemble
    from scatrs_amd.particles import LayeredSphere, Sphere
    from scatrs_amd.size import Gamma, Tabulated

    wl = np.array([0.5, 0.9])
    coated = LayeredSphere([0.6], [1.6 + 0.2j, 1.4 + 1e-3j])
    e = Ensemble(coated, Gamma(0.3, 0.1))
    a = e.compute(wl, greek=12, rtol=1e-8, mu=[-1.0, 0.0, 1.0], n_angles=None)
    f = e.compute(wl, greek=12, n_radius=2048, n_angles=256, mu=[-1.0, 0.0, 1.0])
    assert a.converged.all()
    for k in ("c_ext", "c_sca", "c_abs", "g", "ssa", "effective_radius"):
        assert np.allclose(getattr(a, k), getattr(f, k), rtol=1e-6, atol=0), k
    assert np.allclose(a.greek.alpha1, f.greek.alpha1, rtol=0, atol=1e-6)
    assert np.allclose(a.phase.f11, f.phase.f11, rtol=1e-6, atol=0)
    tab = Tabulated((0.1, 0.3, 0.6, 1.0), (0.5, 2.0, 1.0, 0.2))
    t = Ensemble(Sphere(1.5 + 0.01j), tab)
    a = t.compute(wl, greek=8, rtol=1e-8, n_angles=None, phase=False)
    f = t.compute(wl, greek=8, n_radius=1024, n_angles=128, phase=False)
    assert a.converged.all() and np.all(a.panels >= 3 * 4)
    for k in ("c_ext", "c_sca", "g", "effective_radius", "mean_volume"):
        assert np.allclose(getattr(a, k), getattr(f, k), rtol=1e-7, atol=0), k


def test_refine_backscatter_and_layered_derivatives():
    """refine_backscatter adds C_back to the refined quantities (engine.rs:675-690): it takes
    at least as many panels and its C_back error meets rtol against max(C_back, C_sca). A
    coated-sphere gamma law's adaptive derivative columns (both layers' m, the radius
    fraction, the effective variance) against central differences of the engine at rtol 1e-10."""
    from scatrs_amd import _wire

    wl = np.array([0.7])
    sph = [(("lognormal", 0.4, 1.5), ("sphere", np.array([1.45 + 1e-4j])), 1.0)]
    a = _wire.ensemble_compute(sph, wl, False, None, False, np.zeros(0), True, [], 1e-6, 4096,
                               False, False)
    b = _wire.ensemble_compute(sph, wl, False, None, False, np.zeros(0), True, [], 1e-6, 4096,
                               False, True)
    assert a["converged"].all() and b["converged"].all()
    assert b["panels"][0] >= a["panels"][0]
    assert b["errors"][3, 0] <= 1e-6 * max(b["values"][3, 0], b["values"][1, 0])

    def items(f=0.6, m0=1.6 + 0.1j, m1=1.38 + 1e-3j, v=0.15):
        m = np.array([[m0], [m1]])
        return [(("gamma", 0.4, v), ("layered", np.array([f]), m), 1.0)]

    names = _wire.ensemble_parameters(items(), False)
    want = ["effective_variance", "layer0.m_real", "layer0.m_imag", "layer1.m_real",
            "layer1.m_imag", "layer0.radius_fraction"]
    raw = _wire.ensemble_compute(items(), wl, False, 4, False, np.zeros(0), True,
                                 [names.index(n) for n in want], 1e-10, 4096, False, False)
    assert raw["converged"].all()
    base = dict(f=0.6, m0=1.6 + 0.1j, m1=1.38 + 1e-3j, v=0.15)
    # steps of 1e-4 relative, 1e-5 in Im m (the shell's nearly linear absorption)
    steps = [("v", 1e-4 * 0.15), ("m0", 1e-4 * 1.6), ("m0", 1e-5j), ("m1", 1e-4 * 1.38),
             ("m1", 1e-5j), ("f", 1e-4 * 0.6)]
    sel = [0, 1, 2, 4, 5]
    for k, (key, h) in enumerate(steps):
        up = _wire.ensemble_compute(items(**dict(base, **{key: base[key] + h})), wl, False, 4,
                                    False, np.zeros(0), True, [], 1e-10, 4096, False, False)
        dn = _wire.ensemble_compute(items(**dict(base, **{key: base[key] - h})), wl, False, 4,
                                    False, np.zeros(0), True, [], 1e-10, 4096, False, False)
        fd = (up["values"] - dn["values"]) / (2 * abs(h))
        scale = np.max(np.abs(fd[sel]), 1, keepdims=True)
        err = np.abs(raw["d"][k][sel] - fd[sel]) / scale
        assert err.max() <= 2e-5, (want[k], err.max())
        fdg = (up["greek"] - dn["greek"]) / (2 * abs(h))
        errg = np.abs(raw["greek_d"][k] - fdg) / np.max(np.abs(fdg))
        assert errg.max() <= 2e-5, (want[k], errg.max())


def test_ensemble_api_errors_derivatives_and_grids():
    """Ensemble.compute as the reference's: the adaptive engine by default, result.error,
    result.d(name) (equal to the wire format's columns), a grid of distributions through the
    batch call, and the fixed grid with n_radius."""
    from scatrs_amd import _wire
    from scatrs_amd.ensemble import Ensemble
    from scatrs_amd.particles import Sphere
    from scatrs_amd.size import LogNormal

    wl = np.array([0.5, 0.8])
    m = 1.45 + 0.01j
    r = Ensemble(Sphere(m), LogNormal(0.2, 1.5)).compute(wl, greek=4,
                                                         derivatives=["width", "m_imag"])
    assert r.params == ("width", "m_imag") and r.converged.all()
    assert np.all(r.error.c_ext <= 1e-6 * r.c_ext) and r.error.c_ext.shape == wl.shape
    raw = _wire.ensemble_compute([(("lognormal", 0.2, 1.5), ("sphere", np.full(2, m)), 1.0)],
                                 wl, False, 4, False, np.zeros(0), True, [1, 3], 1e-6, 4096,
                                 False, False)
    assert np.array_equal(r.d("width").c_ext, raw["d"][0][0])
    assert np.array_equal(r.d("m_imag").greek.alpha1, raw["greek_d"][1][0])
    assert r.d("width").error.c_ext.shape == wl.shape
    with pytest.raises(KeyError):
        r.d("median_radius")
    with pytest.raises(InvalidInputError, match="no parameter"):
        Ensemble(Sphere(m), LogNormal(0.2, 1.5)).compute(wl, derivatives=["radius"])
    grid = Ensemble(Sphere(m), LogNormal(np.array([0.1, 0.2, 0.4]), 1.5)).compute(wl, greek=4)
    assert grid.c_ext.shape == (3, 2) and grid.greek.alpha1.shape[:2] == (3, 2)
    assert grid.converged.shape == (3, 2) and grid.converged.all()
    # the same law without derivatives: pass 1 refines fewer quantities, so its panels and
    # values differ within rtol; a grid point equals the single law's run bit for bit
    plain = Ensemble(Sphere(m), LogNormal(0.2, 1.5)).compute(wl, greek=4)
    assert np.allclose(grid.c_ext[1], plain.c_ext, rtol=1e-13, atol=0)
    assert np.allclose(plain.c_ext, r.c_ext, rtol=2e-6, atol=0)
    fixed = Ensemble(Sphere(m), LogNormal(0.2, 1.5)).compute(wl, greek=4, n_radius=256)
    assert fixed.error is None and np.allclose(fixed.c_ext, r.c_ext, rtol=1e-4, atol=0)


def test_batch_members_share_panels_and_match_single_runs():
    """The batch call's shared panels (one panel set and one node solve per wavelength for
    every distribution of the particle, engine.rs:257-296): a lognormal, a gamma law, a
    bimodal law with a power-law mode and a tabulated law, at rtol 1e-8, each within a few rtol
    of its own adaptive run (scalars, Greek coefficients, the phase matrix at given cosines),
    converged, with error estimates within the reference's targets."""
    from scatrs_amd import _wire

    wl = np.array([0.45, 0.9])
    m = np.array([1.45 + 0.01j, 1.40 + 1e-3j])
    specs = [("lognormal", 0.2, 1.6), ("gamma", 0.5, 0.15),
             ("multimodal", [(1.0, ("lognormal", 0.08, 1.4)),
                             (0.3, ("power_law", 3.5, 0.1, 1.5))]),
             ("tabulated", np.array([0.1, 0.3, 0.6, 1.0]), np.array([0.5, 2.0, 1.0, 0.2]))]
    mu = np.array([-1.0, -0.3, 0.0, 0.7, 1.0])
    opts = (1e-8, 4096, False, False)
    b = _wire.ensemble_compute_batch(specs, ("sphere", m), wl, 12, False, mu, True, *opts,
                                     shared_panels=True)
    assert b["converged"].all()
    # C_ext and C_sca to rtol of themselves, C_abs to rtol of C_ext (the reference's targets)
    assert np.all(b["errors"][:2] <= 1e-8 * np.abs(b["values"][:2]))
    assert np.all(b["errors"][2] <= 1e-8 * np.abs(b["values"][0]))
    for s, spec in enumerate(specs):
        one = _wire.ensemble_compute([(spec, ("sphere", m), 1.0)], wl, False, 12, False, mu,
                                     True, [], *opts)
        rows = slice(2 * s, 2 * s + 2)
        for k in (0, 1, 2, 4, 5, 8, 9, 10):
            d = np.abs(b["values"][k, rows] - one["values"][k])
            assert np.all(d <= 5e-8 * np.abs(one["values"][k])), (spec[0], k, d)
        dg = np.abs(b["greek"][:, rows] - one["greek"]).max()
        assert dg <= 5e-8 * np.abs(one["greek"]).max(), (spec[0], dg)
        dp = np.abs(b["phase"][:, rows] - one["phase"]).max()
        assert dp <= 5e-8 * np.abs(one["phase"]).max(), (spec[0], dp)


def test_pass1_sums_on_the_device_equal_the_host_sums(monkeypatch):
    """Large pass-1 rounds of spheres without refined derivatives form their panels' sums on the
    device (_panel_sums_device): the same panels and, to rounding, the same values as the
    host's sums (every round forced to the device, then none)."""
    from scatrs_amd import adaptive
    from scatrs_amd.ensemble import Ensemble
    from scatrs_amd.particles import Sphere
    from scatrs_amd.size import LogNormal

    e = Ensemble(Sphere(1.43 + 1e-8j), LogNormal(np.array([0.1, 0.4, 1.0]),
                                                  np.array([1.4, 1.8, 2.0])))
    wl = np.array([0.35, 0.7])
    out = {}
    for name, limit in (("device", 0), ("host", 1 << 62)):
        monkeypatch.setattr(adaptive, "_DEVICE_SUMS_MIN", limit)
        out[name] = e.compute(wl, rtol=3e-4, greek=16, phase=False, n_angles=None)
    d, h = out["device"], out["host"]
    assert np.array_equal(np.asarray(d.panels), np.asarray(h.panels))
    for k in ("c_ext", "c_sca", "ssa", "g", "c_back"):
        assert np.allclose(getattr(d, k), getattr(h, k), rtol=1e-12, atol=0), k
    scale = np.abs(h.greek.alpha1).max()
    assert np.allclose(d.greek.alpha1, h.greek.alpha1, rtol=0, atol=1e-12 * scale)


def test_pass2_from_pass1_sums_equals_pass2():
    """A call without Greek coefficients, phase or derivatives takes pass 2's sums from pass 1's
    panels (adaptive.pass2_from_panels); the same call with Greek to order 0 runs pass 2 on the
    same panels: the scalars and their error bounds agree to rounding."""
    from scatrs_amd.ensemble import Ensemble
    from scatrs_amd.particles import Sphere
    from scatrs_amd.size import LogNormal

    e = Ensemble(Sphere(1.43 + 1e-3j), LogNormal(np.array([0.2, 0.8]), np.array([1.5, 1.9])))
    wl = np.array([0.4, 1.0])
    a = e.compute(wl, rtol=1e-5, phase=False, n_angles=None)
    b = e.compute(wl, rtol=1e-5, greek=0, phase=False, n_angles=None)
    assert np.array_equal(np.asarray(a.panels), np.asarray(b.panels))
    for k in ("c_ext", "c_sca", "c_abs", "c_back", "ssa", "g", "lidar_ratio", "effective_radius"):
        assert np.allclose(getattr(a, k), getattr(b, k), rtol=1e-12, atol=0), k
        assert np.allclose(getattr(a.error, k), getattr(b.error, k), rtol=1e-6,
                           atol=1e-15 * np.abs(getattr(b, k)).max()), k
