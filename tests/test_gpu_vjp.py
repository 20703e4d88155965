"""GPU parity of the reverse-mode ensemble kernel (scatrs_ensemble_fixed_vjp) through the C ABI:
against the oracle's forward-mode Jacobian (the reference's derivative structure: efficiency
Jacobians, phase-element derivatives, quotient rule) contracted with random cotangents, and
against central differences of the GPU forward at h = 1e-6 relative. The oracle Jacobian is
itself pinned by central differences of the oracle forward (test_oracle_vjp.py)."""

from __future__ import annotations

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def _cotangents(rng, count, order, n_angles, which=("scalars", "greek", "phase")):
    bs = rng.normal(size=(count, 11)) if "scalars" in which else None
    if bs is not None:
        bs[:, 7] = 0.0  # the depolarization of spheres is identically 0
    bg = rng.normal(size=(count, 6, order + 1)) if "greek" in which else None
    bp = rng.normal(size=(count, 6, n_angles)) if "phase" in which else None
    return bs, bg, bp


def _gpu_vjp(kind, p0, p1, wl, m, fr, bs, bg, bp, **kw):
    from scatrs_amd.ensemble import ensemble_fixed_vjp_device

    if kw.get("p2") is not None:
        kw["p2"] = _dev(np.asarray(kw["p2"], np.float64))
    g, terms = ensemble_fixed_vjp_device(
        _dev(np.asarray(kind, np.int32)), _dev(p0), _dev(p1), _dev(wl), _dev(m),
        None if fr is None else _dev(fr), bar_scalars=None if bs is None else _dev(bs),
        bar_greek=None if bg is None else _dev(bg), bar_phase=None if bp is None else _dev(bp),
        **kw)
    return g.cpu().numpy(), terms


def _oracle_vjp(kind, p0, p1, wl, m, fr, bs, bg, bp, n_radius, n_angles, order, normalized,
                fixed_range, p2=0.0):
    """sum over outputs of cotangent x derivative, and the matching scale sum |cot x deriv|."""
    J = oracle.ensemble_fixed_jacobian(kind, p0, p1, wl, m, fr, n_radius, n_angles, order,
                                       normalized, fixed_range, p2=p2)
    terms = []
    if bs is not None:
        terms.append(J["d_scalars"] * bs[None, :])
    if bg is not None:
        terms.append((J["d_greek"] * bg[None]).reshape(J["d_greek"].shape[0], -1))
    if bp is not None:
        terms.append((J["d_phase"] * bp[None]).reshape(J["d_phase"].shape[0], -1))
    t = np.concatenate(terms, axis=1)
    return t.sum(axis=1), np.abs(t).sum(axis=1)


def _assert_close(got, ref, scale, tol, what):
    # a column with no dependence (an unused distribution slot) must be exactly 0
    err = np.abs(got - ref) / np.where(scale > 0, scale, 1.0)
    assert err.max() <= tol, (what, err, got, ref)


@pytest.mark.parametrize("fixed_range", [False, True])
def test_sphere_vjp_matches_oracle_jacobian(fixed_range):
    rng = np.random.default_rng(5)
    kind = np.array([0, 0, 1, 1])
    p0 = np.array([0.12, 0.4, 0.3, 0.8])
    p1 = np.array([1.5, 2.1, 0.1, 0.3])
    wl = np.array([0.55, 0.9, 0.45, 1.6])
    m = np.array([1.43 + 1e-8j, 1.6 + 0.02j, 1.33 + 1e-3j, 1.52 + 0.3j])
    nr, na, order = 40, 70, 24
    bs, bg, bp = _cotangents(rng, 4, order, na)
    g, _ = _gpu_vjp(kind, p0, p1, wl, m, None, bs, bg, bp, n_radius=nr, n_angles=na,
                    order=order, fixed_range=fixed_range)
    for e in range(4):
        ref, scale = _oracle_vjp(kind[e], p0[e], p1[e], wl[e], m[e], None, bs[e], bg[e], bp[e],
                                 nr, na, order, True, fixed_range)
        _assert_close(g[e], ref, scale, 1e-10, f"ensemble {e}")


def test_sphere_vjp_each_output_alone_and_unnormalized():
    rng = np.random.default_rng(6)
    kind, p0, p1, wl = np.array([0, 1]), np.array([0.2, 0.6]), np.array([1.8, 0.2]), \
        np.array([0.5, 1.1])
    m = np.array([1.5 + 0.001j, 1.38 + 1e-6j])
    nr, na, order = 32, 51, 20  # odd n_angles: mu = 0 is a node
    for which in (("scalars",), ("greek",), ("phase",)):
        for normalized in (True, False):
            bs, bg, bp = _cotangents(rng, 2, order, na, which)
            g, _ = _gpu_vjp(kind, p0, p1, wl, m, None, bs, bg, bp, n_radius=nr, n_angles=na,
                            order=order, normalized=normalized)
            for e in range(2):
                ref, scale = _oracle_vjp(kind[e], p0[e], p1[e], wl[e], m[e], None,
                                         None if bs is None else bs[e],
                                         None if bg is None else bg[e],
                                         None if bp is None else bp[e], nr, na, order,
                                         normalized, False)
                _assert_close(g[e], ref, scale, 1e-10, (which, normalized, e))


@pytest.mark.parametrize("fixed_range", [False, True])
def test_new_distributions_vjp_match_oracle_jacobian(fixed_range):
    """Modified gamma, power law and monodisperse ensembles (spheres and a coated sphere):
    the VJP against the oracle's contracted Jacobian, with moving nodes and with the range
    held."""
    rng = np.random.default_rng(15)
    kind = np.array([2, 2, 3, 3, 4])
    p0 = np.array([2.0, 0.6, 3.0, -1.0, 0.45])
    p1 = np.array([6.0, 15.0, 0.05, 0.1, 0.0])
    p2 = np.array([1.5, 0.8, 1.0, 0.6, 0.0])
    wl = np.array([0.6, 0.9, 0.5, 1.2, 0.55])
    m = np.array([1.5 + 0.01j, 1.38 + 1e-4j, 1.45 + 1e-3j, 1.6 + 0.1j, 1.52 + 0.02j])
    nr, na, order = 40, 64, 20
    bs, bg, bp = _cotangents(rng, kind.size, order, na)
    g, _ = _gpu_vjp(kind, p0, p1, wl, m, None, bs, bg, bp, p2=p2, n_radius=nr, n_angles=na,
                    order=order, fixed_range=fixed_range)
    for e in range(kind.size):
        ref, scale = _oracle_vjp(kind[e], p0[e], p1[e], wl[e], m[e], None, bs[e], bg[e], bp[e],
                                 nr, na, order, True, fixed_range, p2=p2[e])
        _assert_close(g[e], ref, scale, 1e-10, f"ensemble {e}")


@pytest.mark.parametrize("layers", [2, 3, 5])
def test_layered_vjp_matches_oracle_jacobian(layers):
    rng = np.random.default_rng(10 + layers)
    n = 3
    kind = np.ones(n, np.int32)
    p0 = np.exp(rng.uniform(np.log(0.05), np.log(0.8), n))
    p1 = rng.uniform(0.05, 0.35, n)
    wl = np.exp(rng.uniform(np.log(0.3), np.log(2.5), n))
    fr = np.cumsum(rng.dirichlet(np.full(layers, 2.0), n), axis=1)[:, :-1]
    m = rng.uniform(1.3, 2.0, (n, layers)) + 1j * np.exp(rng.uniform(np.log(1e-6), 0.0,
                                                                       (n, layers)))
    nr, na, order = 24, 48, 16
    bs, bg, bp = _cotangents(rng, n, order, na)
    g, _ = _gpu_vjp(kind, p0, p1, wl, m, fr, bs, bg, bp, n_radius=nr, n_angles=na, order=order)
    assert g.shape == (n, 3 + 2 * layers + layers - 1)
    for e in range(n):
        ref, scale = _oracle_vjp(1, p0[e], p1[e], wl[e], m[e], fr[e], bs[e], bg[e], bp[e], nr,
                                 na, order, True, False)
        _assert_close(g[e], ref, scale, 1e-9, f"layers {layers} ensemble {e}")


def test_vjp_matches_central_differences_of_the_gpu_forward():
    """Central differences at h = 1e-6 relative of sum(bar . outputs) through the GPU
    forward, for every parameter of a lognormal sphere ensemble, a 3-layer gamma one, a modified
    gamma law of spheres, a power law (its fixed limits get 0) and monodisperse coated spheres."""
    from scatrs_amd.ensemble import ensemble_fixed_device

    rng = np.random.default_rng(21)
    nr, na, order = 48, 64, 24
    cases = [  # kind, (p0, p1, p2), wavelength, m, fractions, free distribution parameters
        (0, (0.25, 1.7, 0.0), 0.6, np.array([1.45 + 0.003j]), None, 2),
        (1, (0.35, 0.2, 0.0), 0.8, np.array([1.7 + 0.2j, 1.5 + 0.01j, 1.35 + 1e-5j]),
         np.array([0.3, 0.65]), 2),
        (2, (2.0, 6.0, 1.5), 0.6, np.array([1.5 + 0.01j]), None, 3),
        (3, (3.0, 0.05, 1.0), 0.7, np.array([1.45 + 1e-3j]), None, 1),
        (4, (0.7, 0.0, 0.0), 0.55, np.array([1.6 + 0.1j, 1.4 + 1e-3j]), np.array([0.5]), 1),
    ]
    for kind, q, wl, m, fr, free in cases:
        L = m.size
        bs, bg, bp = _cotangents(rng, 1, order, na)

        def loss(q, m, fr):
            sc, gr, ph, _ = ensemble_fixed_device(
                _dev(np.array([kind], np.int32)), _dev(np.array([q[0]])), _dev(np.array([q[1]])),
                _dev(np.array([wl])), _dev(m[None] if L > 1 else m),
                None if fr is None else _dev(fr[None]), p2=_dev(np.array([q[2]])), n_radius=nr,
                n_angles=na, order=order)
            return float((sc.cpu().numpy() * bs).sum() + (gr.cpu().numpy() * bg).sum() +
                         (ph.cpu().numpy() * bp).sum())

        from scatrs_amd.ensemble import ensemble_fixed_vjp_device

        g, _ = ensemble_fixed_vjp_device(
            _dev(np.array([kind], np.int32)), _dev(np.array([q[0]])), _dev(np.array([q[1]])),
            _dev(np.array([wl])), _dev(m[None] if L > 1 else m),
            None if fr is None else _dev(fr[None]), p2=_dev(np.array([q[2]])),
            bar_scalars=_dev(bs), bar_greek=_dev(bg), bar_phase=_dev(bp), n_radius=nr,
            n_angles=na, order=order)
        g = g.cpu().numpy()[0]
        assert g.size == 3 + 2 * L + L - 1
        fd = []
        h = 1e-6
        for p in range(g.size):
            qp, qm = list(q), list(q)
            args_p = [qp, m.copy(), None if fr is None else fr.copy()]
            args_m = [qm, m.copy(), None if fr is None else fr.copy()]
            if p < 3:
                if p >= free:  # an unused slot or a power law's fixed limit
                    assert g[p] == 0.0, (kind, p, g)
                    fd.append(0.0)
                    continue
                d = h * abs(q[p])
                qp[p] += d
                qm[p] -= d
            elif p < 3 + 2 * L:
                l, im = (p - 3) // 2, (p - 3) % 2
                d = h * abs(m[l])
                step = 1j * d if im else d
                args_p[1][l] += step
                args_m[1][l] -= step
            else:
                l = p - 3 - 2 * L
                d = h * fr[l]
                args_p[2][l] += d
                args_m[2][l] -= d
            fd.append((loss(*args_p) - loss(*args_m)) / (2 * d))
        fd = np.array(fd)
        scale = np.abs(fd).max()
        err = np.abs(g - fd) / scale
        print(f"kind {kind}: max |vjp - fd| / max|fd| {err.max():.2e}")
        assert err.max() <= 2e-6, (kind, err, g, fd)


def test_vjp_is_linear_in_the_cotangents():
    rng = np.random.default_rng(3)
    kind, p0, p1, wl = np.zeros(3), np.array([0.1, 0.3, 0.9]), np.array([1.4, 1.9, 2.3]), \
        np.array([0.4, 0.7, 2.0])
    m = np.array([1.4 + 1e-4j, 1.55 + 0.05j, 1.62 + 1e-7j])
    kw = dict(n_radius=64, n_angles=96, order=32)
    a = _cotangents(rng, 3, 32, 96)
    b = _cotangents(rng, 3, 32, 96)
    ga, _ = _gpu_vjp(kind, p0, p1, wl, m, None, *a, **kw)
    gb, _ = _gpu_vjp(kind, p0, p1, wl, m, None, *b, **kw)
    gab, _ = _gpu_vjp(kind, p0, p1, wl, m, None, *(2.0 * x + y for x, y in zip(a, b)), **kw)
    scale = np.abs(2.0 * ga).max(axis=1, keepdims=True) + np.abs(gb).max(axis=1, keepdims=True)
    assert np.max(np.abs(gab - (2.0 * ga + gb)) / scale) <= 1e-12


def test_autograd_function_matches_vjp():
    from scatrs_amd.ensemble import ensemble_fixed

    rng = np.random.default_rng(8)
    n = 3
    p0 = _dev(np.array([0.1, 0.5, 0.3])).requires_grad_()
    p1 = _dev(np.array([1.6, 2.0, 1.3])).requires_grad_()
    mr = _dev(np.array([1.4, 1.6, 1.5])).requires_grad_()
    mi = _dev(np.array([1e-3, 0.05, 1e-6])).requires_grad_()
    kind = _dev(np.zeros(n, np.int32))
    wl = _dev(np.array([0.5, 1.0, 0.8]))
    sc, gr, ph = ensemble_fixed(kind, p0, p1, wl, mr, mi, n_radius=48, n_angles=64, order=24)
    ws = _dev(rng.normal(size=(n, 11)))
    wg = _dev(rng.normal(size=(n, 6, 25)))
    loss = (sc * ws).sum() + (gr * wg).sum()
    loss.backward()
    g, _ = _gpu_vjp(np.zeros(n), p0.detach().cpu().numpy(), p1.detach().cpu().numpy(),
                    wl.cpu().numpy(), (mr + 1j * mi).detach().cpu().numpy(), None,
                    ws.cpu().numpy(), wg.cpu().numpy(), None, n_radius=48, n_angles=64,
                    order=24)
    got = np.stack([t.grad.cpu().numpy() for t in (p0, p1, mr, mi)], axis=1)
    assert not g[:, 2].any()  # a lognormal has no third parameter
    assert np.array_equal(got, g[:, [0, 1, 3, 4]])


def test_adjoint_workload_full_size():
    """The adjoint workload at its full size (1024 lognormals x 32 wavelengths, 512 radius
    nodes, 1024 angles, Greek to 256): finite gradients, and the first rows against the
    oracle's Jacobian contracted with the loss's cotangents."""
    from scatrs_amd import workloads
    from scatrs_amd.ensemble import ensemble_fixed_device, ensemble_fixed_vjp_device

    (kind, p0, p1, wl, m), tg = workloads.adjoint()
    nr, na, order = 512, 1024, 256
    args = [_dev(kind), _dev(p0), _dev(p1), _dev(wl), _dev(m)]
    sc, gr, _, _ = ensemble_fixed_device(*args, n_radius=nr, n_angles=na, order=order,
                                         want_phase=False)
    t = {k: _dev(v) for k, v in tg.items()}
    loss, bs, bg = workloads.adjoint_loss(sc, gr, t)
    g, terms = ensemble_fixed_vjp_device(*args, bar_scalars=bs, bar_greek=bg, n_radius=nr,
                                         n_angles=na, order=order)
    g = g.cpu().numpy()
    assert np.all(np.isfinite(g)) and terms > 0
    bs, bg = bs.cpu().numpy(), bg.cpu().numpy()
    for e in (0, 1, 40):
        ref, scale = _oracle_vjp(0, p0[e], p1[e], wl[e], m[e], None, bs[e], bg[e], None, nr, na,
                                 order, True, False)
        _assert_close(g[e], ref, scale, 1e-9, f"row {e}")


def test_adjoint_workload_with_fp16_split_forward_tables():
    """The adjoint workload with the fp16-split angular kernel selected: the forward call and the
    reverse pass's forward tables both come from that kernel (the reverse products stay FP64).
    Finite gradients, and the first rows within 2e-6 of the oracle's Jacobian contracted with
    the loss's cotangents, relative to sum |cotangent x derivative|."""
    from scatrs_amd import _native, workloads
    from scatrs_amd.ensemble import ensemble_fixed_device, ensemble_fixed_vjp_device

    (kind, p0, p1, wl, m), tg = workloads.adjoint()
    nr, na, order = 512, 1024, 256
    args = [_dev(kind), _dev(p0), _dev(p1), _dev(wl), _dev(m)]
    ctx = _native.context(0)
    ctx.set_angular_kernel(2)
    try:
        sc, gr, _, _ = ensemble_fixed_device(*args, n_radius=nr, n_angles=na, order=order,
                                             want_phase=False)
        t = {k: _dev(v) for k, v in tg.items()}
        _, bs, bg = workloads.adjoint_loss(sc, gr, t)
        g, _ = ensemble_fixed_vjp_device(*args, bar_scalars=bs, bar_greek=bg, n_radius=nr,
                                         n_angles=na, order=order)
    finally:
        ctx.set_angular_kernel(0)
    g = g.cpu().numpy()
    assert np.all(np.isfinite(g))
    bs, bg = bs.cpu().numpy(), bg.cpu().numpy()
    for e in (0, 1, 40):
        ref, scale = _oracle_vjp(0, p0[e], p1[e], wl[e], m[e], None, bs[e], bg[e], None, nr, na,
                                 order, True, False)
        err = np.abs(g[e] - ref) / np.where(scale > 0, scale, 1.0)
        print(f"row {e}: max |g - ref| / scale {err.max():.2e}")
        _assert_close(g[e], ref, scale, 2e-6, f"row {e}")


def test_vjp_with_phase_at_the_callers_cosines():
    """Phase cotangents at the caller's cosines (unsorted, repeated, 0 and +-1), alone and with
    scalar and Greek cotangents: the VJP against central differences of the GPU forward at the
    same cosines (h = 1e-6 relative, 2e-6 of the largest), and at the rule's own cosines the
    rule's VJP; through the autograd Function as well."""
    from scatrs_amd.ensemble import ensemble_fixed, ensemble_fixed_device, ensemble_fixed_vjp_device
    from scatrs_amd.quadrature import gauss_legendre

    rng = np.random.default_rng(33)
    mu = np.array([0.3, -1.0, 1.0, 0.0, -0.3, 0.95, 0.3, -0.7071])
    nr, na, order = 40, 64, 16
    kw = dict(n_radius=nr, n_angles=na, order=order)
    cases = [(0, (0.25, 1.7, 0.0), 0.6, 1.45 + 0.003j), (3, (3.0, 0.05, 1.2), 0.7, 1.5 + 0.01j)]
    for kind, q, wl, m in cases:
        for which in (("phase",), ("scalars", "greek", "phase")):
            bs, bg, _ = _cotangents(rng, 1, order, na, which)
            bp = rng.normal(size=(1, 6, mu.size))

            def loss(q, m):
                sc, gr, ph, _ = ensemble_fixed_device(
                    _dev(np.array([kind], np.int32)), _dev(np.array([q[0]])),
                    _dev(np.array([q[1]])), _dev(np.array([wl])), _dev(np.array([m])),
                    p2=_dev(np.array([q[2]])), mu=mu, **kw)
                out = (ph.cpu().numpy() * bp).sum()
                if bs is not None:
                    out += (sc.cpu().numpy() * bs).sum() + (gr.cpu().numpy() * bg).sum()
                return float(out)

            g, _ = ensemble_fixed_vjp_device(
                _dev(np.array([kind], np.int32)), _dev(np.array([q[0]])), _dev(np.array([q[1]])),
                _dev(np.array([wl])), _dev(np.array([m])), p2=_dev(np.array([q[2]])),
                bar_scalars=None if bs is None else _dev(bs),
                bar_greek=None if bg is None else _dev(bg), bar_phase=_dev(bp), mu=mu, **kw)
            g = g.cpu().numpy()[0]
            free = 2 if kind == 0 else 1
            fd = np.zeros_like(g)
            h = 1e-6
            for p in list(range(free)) + [3, 4]:
                qp, qm, mp, mm = list(q), list(q), m, m
                if p < 3:
                    d = h * abs(q[p])
                    qp[p] += d
                    qm[p] -= d
                else:
                    d = h * abs(m)
                    step = d if p == 3 else 1j * d
                    mp, mm = m + step, m - step
                fd[p] = (loss(qp, mp) - loss(qm, mm)) / (2 * d)
            err = np.abs(g - fd) / np.abs(fd).max()
            assert err.max() <= 2e-6, (kind, which, err, g, fd)
    # at the rule's own cosines: the rule's VJP
    rule, _ = gauss_legendre(na)
    bp = rng.normal(size=(1, 6, na))
    args = [_dev(np.array([0], np.int32)), _dev(np.array([0.25])), _dev(np.array([1.7])),
            _dev(np.array([0.6])), _dev(np.array([1.45 + 0.003j]))]
    ga, _ = ensemble_fixed_vjp_device(*args, bar_phase=_dev(bp), **kw)
    gb, _ = ensemble_fixed_vjp_device(*args, bar_phase=_dev(bp), mu=rule, **kw)
    assert float((ga - gb).abs().max() / ga.abs().max()) <= 1e-10
    # the autograd Function with mu
    p0 = _dev(np.array([0.25])).requires_grad_()
    sc, gr, ph = ensemble_fixed(args[0], p0, args[2], args[3], _dev(np.array([1.45])),
                                _dev(np.array([0.003])), mu=mu, **kw)
    w = _dev(rng.normal(size=(1, 6, mu.size)))
    (ph * w).sum().backward()
    gv, _ = ensemble_fixed_vjp_device(*args, bar_phase=w, mu=mu, **kw)
    assert float(p0.grad[0]) == float(gv[0, 0])
