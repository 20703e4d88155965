"""Host logic: workloads, cost accounting and the Python API's checks (CPU only)."""

from __future__ import annotations

import math

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


def test_kronrod15_rule_is_qk15():
    """The 7-point Gauss part equals numpy's Gauss-Legendre rule; the 15-point Kronrod rule
    integrates x^k exactly for k <= 22 (degree 3 n + 1 for n = 7) and not for k = 24."""
    from scatrs_amd.adaptive import kronrod15

    x, wk, wg = kronrod15()
    gx, gw = np.polynomial.legendre.leggauss(7)
    sel = wg > 0
    assert np.allclose(x[sel], gx, rtol=0, atol=1e-15)
    assert np.allclose(wg[sel], gw, rtol=0, atol=1e-15)
    assert np.all(np.diff(x) > 0) and np.allclose(x, -x[::-1], atol=0)
    for k in range(23):
        exact = 2.0 / (k + 1) if k % 2 == 0 else 0.0
        assert abs(np.sum(wk * x ** k) - exact) < 1e-15, k
    assert abs(np.sum(wk * x ** 24) - 2.0 / 25) > 1e-9


def test_integration_range_cuts_the_area_tails_at_rtol_over_10():
    """A lognormal's area-weighted law is a lognormal of median r_g exp(2 s^2): its range ends
    where the normal tail holds rtol/20, within the grid's spacing; the density is the
    closed form's up to a constant."""
    from scipy.stats import norm

    from scatrs_amd.adaptive import _GRID, integration_range, log_density, size_range

    r_g, sigma, rtol = 0.3, 1.6, 1e-6
    s = np.log(sigma)
    lo, hi = size_range(0, (r_g, sigma, 0.0))
    assert np.isclose(lo, np.log(r_g) - 5 * s) and np.isclose(hi, np.log(r_g) + 5 * s)
    t_lo, t_hi, _ = integration_range(0, (r_g, sigma, 0.0), rtol)
    c = np.log(r_g) + 2 * s * s
    z = norm.isf(rtol / 20)
    step = 30 * s / (_GRID - 1)  # the grid spans three times the fixed range, +-15 s
    assert abs(t_lo - (c - z * s)) <= 2 * step and abs(t_hi - (c + z * s)) <= 2 * step
    r = np.array([0.05, 0.3, 1.2])
    lu = log_density(0, (r_g, sigma, 0.0), r)
    ref = -np.log(r / r_g) ** 2 / (2 * s * s) - np.log(r)
    assert np.allclose(lu - lu[1], ref - ref[1], rtol=0, atol=1e-13)
    with pytest.raises(InvalidInputError):
        size_range(0, (0.3, 0.9, 0.0))


def _stub_nodes(q_of_r, dq_of_r=None):
    """A stand-in for pass 1's device call (adaptive._node_values): closed-form per-node
    quantities q_of_r(r) -> (C_ext, C_sca, C_abs, C_back, g), and with dq_of_r(r) -> the same
    five along one direction of the particle, their derivatives."""

    def call(tasks, ti, r, allow_unvalidated, table=None):
        ext, sca, absb, back, g = q_of_r(r)
        Q = np.stack([ext, sca, sca * g, absb, back], 1)
        D = max((len(tasks[k].dirs) for k in np.unique(ti)), default=0)
        dQ = None
        if D:
            de, ds, da, db, dg = dq_of_r(r)
            one = np.stack([de, ds, ds * g + sca * dg, da, db], 1)
            has = np.array([len(tasks[k].dirs) > 0 for k in ti])
            dQ = np.zeros((r.size, D, 5))
            dQ[:, 0] = np.where(has[:, None], one, 0.0)
        return Q, dQ, int(r.size)

    return call


def test_adaptive_pass1_converges_to_the_closed_form(monkeypatch):
    """Pass 1's panel logic on the CPU (the device call replaced by closed-form per-node
    quantities): for a lognormal with C_ext = pi r^2 and C_sca = 0.9 pi r^2 (1 + 0.3 sin(40 ln
    r)) the engine's <C_ext> meets rtol against the closed form r_g^2 exp(2 s^2) pi over the
    whole law (its moment integration's Z), refines where the integrand oscillates, and stops
    unconverged at max_panels."""
    from scatrs_amd import adaptive

    def q(r):
        ext = np.pi * r * r
        sca = 0.9 * ext * (1.0 + 0.3 * np.sin(40.0 * np.log(r)))
        return ext, sca, ext - sca, 0.1 * ext, np.full_like(r, 0.5)

    monkeypatch.setattr(adaptive, "_node_values", _stub_nodes(q))
    r_g, sigma, rtol = 0.3, 1.8, 1e-8
    s = np.log(sigma)

    def task():
        return adaptive.Task(0, np.array([r_g, sigma, 0.0]), 0.5, np.array([1.5 + 0j]),
                             np.zeros(0))

    t = task()
    adaptive.pass1([t], rtol, 4096, False, False)
    assert t.converged and t.a.size > 4 * 4
    assert np.all(np.diff(t.a) > 0) and np.allclose(t.a[1:], t.b[:-1], rtol=0, atol=1e-15)
    s_ext = np.sum(t.K[:, 0])  # C_ext sum over the panels, in the weights' scale
    exact = np.pi * r_g ** 2 * np.exp(2 * s * s)
    assert abs(s_ext / t.moments[0] - exact) <= 5 * rtol * exact
    err = np.abs(t.K - t.G).sum(0)
    assert err[0] <= rtol * abs(t.K[:, 0].sum()) and err[1] <= rtol * abs(t.K[:, 1].sum())
    assert np.all(np.isfinite(t.K)) and np.all(np.isfinite(t.G))
    for cap in (8, 11):  # the splits fill the budget exactly, worst first
        small = task()
        adaptive.pass1([small], rtol, cap, False, False)
        assert not small.converged and small.a.size == cap
        assert np.all(np.isfinite(small.K))


def test_adaptive_pass1_grows_the_range_for_steep_integrands(monkeypatch):
    """With C_sca ~ r^6 (the Rayleigh law) the outer panel holds more than rtol/10 of it at the
    area rule's upper end: the range grows until the r^6 law's tail is below rtol/10 of its
    closed-form total r_g^6 exp(18 s^2); a gamma task in the same batch converges beside it."""
    from scatrs_amd import adaptive

    def q(r):
        ext = r ** 6 + 0.01 * r * r
        return ext, r ** 6, 0.01 * r * r, r ** 6, np.zeros_like(r)

    monkeypatch.setattr(adaptive, "_node_values", _stub_nodes(q))
    r_g, sigma, rtol = 0.01, 2.0, 1e-7
    s = np.log(sigma)
    t = adaptive.Task(0, np.array([r_g, sigma, 0.0]), 1.0, np.array([1.5 + 0j]), np.zeros(0))
    g = adaptive.Task(1, np.array([0.5, 0.2, 0.0]), 1.0, np.array([1.5 + 0j]), np.zeros(0))
    _, t_hi, _ = adaptive.integration_range(0, t.p, rtol)
    adaptive.pass1([t, g], rtol, 4096, False, False)
    assert t.converged and g.converged
    assert t.t_hi > t_hi + 1e-9
    sca = np.sum(t.K[:, 1]) / t.moments[0]
    exact = r_g ** 6 * np.exp(18 * s * s)
    assert abs(sca - exact) <= 5 * rtol * exact


def test_adaptive_pass1_refines_the_requested_derivatives(monkeypatch):
    """Derivative targets (engine.rs:620-716) on the CPU: a direction of the particle whose
    derivative has Lorentzian spikes of width 0.03 in ln r (as d C / d Im m has at the
    resonances of a nonabsorbing sphere) cannot be resolved within 64 panels, so the task ends
    unconverged at exactly 64 panels, while the same values without that derivative converge
    on far fewer; a pass whose budget suffices meets the derivative's target, rtol times the
    integral of its absolute integrand."""
    from scatrs_amd import adaptive

    centers = np.linspace(np.log(0.05), np.log(3.0), 25)

    def q(r):
        ext = np.pi * r * r
        return ext, 0.9 * ext, 0.1 * ext, 0.05 * ext, np.full_like(r, 0.6)

    def dq(r):
        t = np.log(r)[:, None]
        spikes = (0.03 ** 2 / ((t - centers[None, :]) ** 2 + 0.03 ** 2)).sum(1)
        d = np.pi * r * r * spikes
        return d, 0.5 * d, 0.5 * d, 0.1 * d, np.zeros_like(r)

    monkeypatch.setattr(adaptive, "_node_values", _stub_nodes(q, dq))
    rtol = 1e-6

    def task(dirs=()):
        return adaptive.Task(0, np.array([0.3, 1.8, 0.0]), 0.5, np.array([1.33 + 0j]),
                             np.zeros(0), dirs=dirs)

    plain, wanted = task(), task(dirs=(1,))
    adaptive.pass1([plain, wanted], rtol, 64, False, False)
    assert plain.converged and plain.a.size < 32
    assert not wanted.converged and wanted.a.size == 64
    big = task(dirs=(1,))
    adaptive.pass1([big], rtol, 4096, False, False)
    assert big.converged and big.a.size > 64
    blk = slice(5, 10)  # the direction's block: C_ext, C_sca, C_sca g, C_abs, C_back
    err = np.abs(big.K[:, blk] - big.G[:, blk]).sum(0)
    # its absolute integrands are the integrands themselves here (positive spikes)
    absint = np.abs(big.K[:, blk]).sum(0)
    assert np.all(err[:4] <= rtol * absint[[0, 1, 2, 0]])


def test_adaptive_pass1_distribution_parameter_blocks(monkeypatch):
    """The distribution-parameter blocks of pass 1 (engine.rs:528-545): with C_ext = pi r^2 the
    lognormal's integral of u C_ext has the closed form pi r0 sqrt(2 pi) r_g^2 exp(2 s^2) (u at
    its fixed reference radius r0 = r_g, s = ln sigma_g), so the blocks of d/d r_g and d/d
    sigma_g are that integral times 2 / r_g and 4 s / sigma_g; both are refined to rtol of
    their absolute integrals and take at least the panels of the values alone."""
    from scatrs_amd import adaptive

    def q(r):
        ext = np.pi * r * r
        sca = 0.9 * ext * (1.0 + 0.3 * np.sin(12.0 * np.log(r)))
        return ext, sca, ext - sca, 0.1 * ext, np.full_like(r, 0.5)

    monkeypatch.setattr(adaptive, "_node_values", _stub_nodes(q))
    r_g, sigma, rtol = 0.3, 1.8, 1e-9
    s = np.log(sigma)
    plain = adaptive.Task(0, np.array([r_g, sigma, 0.0]), 0.5, np.array([1.5 + 0j]), np.zeros(0))
    t = adaptive.Task(0, np.array([r_g, sigma, 0.0]), 0.5, np.array([1.5 + 0j]), np.zeros(0),
                      theta=(0, 1))
    adaptive.pass1([plain, t], rtol, 4096, False, False)
    assert plain.converged and t.converged and t.a.size >= plain.a.size
    base = t.K[:, 0].sum()
    d_rg, d_sigma = t.K[:, 5].sum(), t.K[:, 10].sum()
    # the range holds all but ~rtol/10 of the area; the derivative blocks' heavier tails
    # (z^2 u) lose a few times that
    assert abs(d_rg / base - 2.0 / r_g) <= 10 * rtol * 2.0 / r_g
    assert abs(d_sigma / base - 4.0 * s / sigma) <= 10 * rtol * 4.0 * s / sigma
    for o in (5, 10):
        err = np.abs(t.K[:, o:o + 4] - t.G[:, o:o + 4]).sum(0)
        absint = np.abs(t.K[:, o:o + 4]).sum(0)  # a lower bound of the absolute integrals
        assert np.all(err <= rtol * absint[[0, 1, 2, 0]] * 1.0001)


def test_adaptive_pass2_rows_recombine_to_the_panel_sums(monkeypatch):
    """pass2_torch on the CPU with the device call replaced by closed-form per-node quantities
    (C_ext = pi r^2, alpha_l = r^l): the rows of 16 panels, grouped by their angle rules and
    recombined by their weights, give each task's sums over all its panels divided by the
    moment integration's Z, for every replica."""
    torch = pytest.importorskip("torch")
    from scatrs_amd import adaptive, ensemble

    def device(nodes, wavelength, m, fractions=None, *, n_angles, order, normalized,
               allow_unvalidated, want_phase, mu):
        r, w = nodes[..., 0], nodes[..., 1]
        z = w.sum(1)
        sc = torch.zeros((r.shape[0], 11), dtype=torch.float64)
        sc[:, 0] = (w * np.pi * r * r).sum(1) / z
        sc[:, 1] = 0.5 * sc[:, 0]
        sc[:, 9] = sc[:, 0]  # pi <r^2>
        gr = torch.stack([torch.stack([(w * r ** l).sum(1) / z for l in range(order + 1)], 1)
                          for _ in range(6)], 1)
        return sc, gr, None, 0

    monkeypatch.setattr(ensemble, "ensemble_nodes_device", device)
    tasks = []
    for p, wl in (((0.3, 1.6, 0.0), 0.5), ((0.5, 0.2, 0.0), 0.9)):
        kind = 0 if p[1] > 1 else 1
        q = adaptive.Task(kind, np.array(p), wl, np.array([1.5 + 0j]), np.zeros(0))
        q.t_lo, q.t_hi, q.c = adaptive.integration_range(kind, q.p, 1e-8)
        q.moments, q.d_moments = adaptive.moment_integrals(kind, q.p, q.c, 1e-8)
        edges = np.linspace(q.t_lo, q.t_hi, 38)  # 37 panels: rows of 16, 16 and 5
        q.a, q.b = edges[:-1], edges[1:]
        tasks.append(q)
    R, order = 2, 3
    params = [(torch.tensor(np.tile(q.p, (R, 1))), torch.full((R, 1), 1.5, dtype=torch.float64),
               torch.zeros((R, 1), dtype=torch.float64), None) for q in tasks]
    avg, gr, ph, _, _, rows = adaptive.pass2_torch(
        tasks, params, R=R, idx=np.arange(15), order=order, n_angles=None, mu=None,
        want_phase=False, allow_unvalidated=False)
    assert ph is None and avg.shape == (2, R, 7) and gr.shape == (2, R, 6, order + 1)
    for t, q in enumerate(tasks):
        r, w = adaptive._nodes(tasks, np.full(q.a.size, t), q.a, q.b, np.arange(15))
        zm = q.moments[0]
        for k in range(R):
            assert np.isclose(avg[t, k, 0].item(), (w * np.pi * r * r).sum() / zm, rtol=1e-13)
            for l in range(order + 1):
                assert np.isclose(gr[t, k, 2, l].item(), (w * r ** l).sum() / zm, rtol=1e-13)
    assert rows["n_rows"] == 6 and rows["n_mono"] == 0


def test_log_density_on_host_threads_equals_the_serial_loop():
    """scatrs_size_log_density splits a long call over host threads: every radius gives the
    same bits as in calls short enough for one thread, and an invalid parameter row in the last
    slice still fails the call."""
    from scatrs_amd.adaptive import log_density

    rng = np.random.default_rng(3)
    n = 300_000
    kind = rng.integers(0, 3, n).astype(np.int32)
    p = np.stack([rng.uniform(0.1, 1.0, n), rng.uniform(1.2, 2.0, n), np.zeros(n)], 1)
    gam = kind != 0
    p[gam, 1] = rng.uniform(0.05, 0.3, gam.sum())  # effective variance
    p[kind == 2, 2] = rng.uniform(0.5, 2.0, (kind == 2).sum())
    p[kind == 2, 1] = rng.uniform(1.0, 3.0, (kind == 2).sum())
    r = np.exp(rng.uniform(-4.0, 1.0, n))
    lu, d = log_density(kind, p, r, derivatives=True)
    for lo in range(0, n, 30_000):
        s = slice(lo, lo + 30_000)
        a, b = log_density(kind[s], p[s], r[s], derivatives=True)
        assert np.array_equal(a, lu[s]) and np.array_equal(b, d[s])
    p[-1, 1] = 0.5  # a lognormal width <= 1 or an effective variance >= 0.5
    kind[-1] = 0
    with pytest.raises(InvalidInputError):
        log_density(kind, p, r)


def test_segment_sum_matches_index_add_and_its_gradient():
    """ensemble.segment_sum, the deterministic stand-in for index_add: the same sums (to
    rounding, in another order), unique indices exact, empty input, and the gradient a gather
    of the output's."""
    import torch

    from scatrs_amd.ensemble import segment_sum

    rng = np.random.default_rng(0)
    idx = torch.from_numpy(rng.integers(0, 7, 50))
    src = torch.from_numpy(rng.normal(size=(50, 3, 2))).requires_grad_(True)
    out = segment_sum(idx, src, 9)
    ref = torch.zeros(9, 3, 2, dtype=torch.float64).index_add(0, idx, src.detach())
    assert torch.allclose(out, ref, rtol=1e-14, atol=1e-14)
    assert torch.all(out[7:] == 0.0)
    w = torch.arange(54.0, dtype=torch.float64).reshape(9, 3, 2)
    (g,) = torch.autograd.grad((out * w).sum(), src)
    assert torch.equal(g, w[idx])
    one = segment_sum(torch.tensor([2, 0]), torch.tensor([1.5, 2.5], dtype=torch.float64), 3)
    assert one.tolist() == [2.5, 0.0, 1.5]
    assert segment_sum(torch.zeros(0, dtype=torch.int64), torch.ones(0), 2).tolist() == [0, 0]


def _toy_ensembles(order, mu):
    """CPU stand-ins for ensemble_nodes_device and ensemble_nodes_tangent_device on a toy
    particle: two complex coefficients per node, analytic in each layer's m and linear in the
    core fraction, with C_ext linear and C_sca, C_sca g, C_back, the Greek rows and the phase
    rows quadratic in them (as Mie sums are), C_abs = C_ext - C_sca; the tangent call moves
    every sum with the coefficients, as the library's does (its absorption sum by the absorption
    identity's derivative)."""
    import torch

    from scatrs_amd.ensemble import assemble

    mu = torch.as_tensor(mu, dtype=torch.float64)
    ell = torch.arange(order + 1, dtype=torch.float64)

    def coeffs(x, m, fr):
        ms = m[:, -1:]
        core = 1.0 + 0.2 * (m[:, :1] - ms) * (fr[:, :1] if m.shape[1] > 1 else 0.0)
        base = (ms * ms - 1.0) / (ms * ms + 2.0)
        c1 = base * x ** 3 * core * (1.0 + 0.1j * x)
        c2 = ms * x * torch.exp(-x / 10.0) * (0.3 + 0.05j) * core
        return c1, c2

    def sums(nodes, wl, c1, c2, wscale):
        r, W = nodes[..., 0], nodes[..., 1] * wscale
        area = math.pi * r * r
        ext = area * (c1 + c2).real
        sca = area * (c1.abs() ** 2 + c2.abs() ** 2)
        absb = area * ((c1 + c2).real - c1.abs() ** 2 - c2.abs() ** 2)
        back = area * (c1 - c2).abs() ** 2
        scag = area * (c1 * c2.conj()).real
        lin = torch.stack([ext, sca, absb, back, scag, r * r, r ** 3], -1)
        pair = torch.stack([c1.abs() ** 2, c2.abs() ** 2, (c1 * c2.conj()).real,
                            (c1 * c2.conj()).imag, (c1 + c2).abs() ** 2, (c1 - 0.5 * c2).real ** 2],
                           -1) * area[..., None]
        gr = pair[..., :, None] / (1.0 + ell)  # (C, N, 6, order + 1)
        ph = (c1[..., None, None] + mu * c2[..., None, None]).abs() ** 2 * (1.0 + torch.arange(
            6.0, dtype=torch.float64))[:, None] * area[..., None, None]
        z = W.sum(1)
        avg = (W[..., None] * lin).sum(1) / z[:, None]
        g = (W[..., None, None] * gr).sum(1) / z[:, None, None]
        p = (W[..., None, None] * ph).sum(1) / z[:, None, None]
        return assemble(avg, g, p, normalized=False)

    def split(m, fr):
        mm = m if m.dim() == 2 else m[:, None]
        f = fr if fr is not None else torch.zeros((mm.shape[0], 0), dtype=torch.float64)
        return mm, f

    def forward(nodes, wl, m, fr=None, **kw):
        mm, f = split(m, fr)
        x = 2.0 * math.pi * nodes[..., 0] / wl[:, None]
        c1, c2 = coeffs(x, mm, f)
        sc, g, p = sums(nodes, wl, c1, c2, 1.0)
        return sc, g, p, 0

    def tangent(nodes, wl, m, fr=None, *, direction, sign, **kw):
        mm, f = split(m, fr)
        L = mm.shape[1]
        x = 2.0 * math.pi * nodes[..., 0] / wl[:, None]
        em, ef = torch.zeros_like(mm), torch.zeros_like(f)
        if direction < 2 * L:
            em[:, direction // 2] = 1.0 if direction % 2 == 0 else 1j
        else:
            ef[:, direction - 2 * L] = 1.0
        c0 = coeffs(x, mm, f)
        hh = 1e-6  # dc by a central difference (the toy is smooth; its error is ~1e-12)
        cp, cm = coeffs(x, mm + hh * em, f + hh * ef), coeffs(x, mm - hh * em, f - hh * ef)
        dc = [(a - b) / (2 * hh) for a, b in zip(cp, cm)]
        step = sign / (1.0 + x)
        c1, c2 = (c + step * d for c, d in zip(c0, dc))
        sc, g, p = sums(nodes, wl, c1, c2, (1.0 + x) / 2.0)
        return sc, g, p, 0

    return forward, tangent


@pytest.mark.parametrize("layers", [1, 2])
def test_nodes_forward_mode_matches_central_differences(monkeypatch, layers):
    """ensemble_nodes' forward mode (the host side: weights' part from a shifted positive
    weighting, particle directions from the library's sign +-1 tangent calls, replica runs
    each with its own tangent) on toy per-node coefficients: every output's tangent within
    1e-7 of a central difference of the same toy's averages, for a distribution direction, the
    particle's directions (Re m, Im m per layer, the core fraction) and a mix."""
    import torch
    from torch.autograd import forward_ad

    from scatrs_amd import ensemble

    order, mu = 3, [-1.0, -0.2, 0.5, 1.0]
    fwd, tan = _toy_ensembles(order, mu)
    monkeypatch.setattr(ensemble, "ensemble_nodes_device", fwd)
    monkeypatch.setattr(ensemble, "ensemble_nodes_tangent_device", tan)
    rng = np.random.default_rng(3)
    C, N, k = 3, 9, 4
    f64 = torch.float64
    r = torch.from_numpy(np.sort(rng.uniform(0.05, 1.5, (C, N)), 1))
    W0 = torch.from_numpy(rng.uniform(0.1, 1.0, (C, N)))
    dlnw = torch.from_numpy(rng.normal(size=(C, N, 3)))
    wl = torch.tensor([0.4, 0.55, 1.0], dtype=f64)
    m = torch.from_numpy(rng.uniform(1.3, 1.6, (C, layers)) + 1j * rng.uniform(1e-3, 0.1,
                                                                               (C, layers)))
    fr = torch.from_numpy(rng.uniform(0.3, 0.8, (C, 1))) if layers > 1 else None
    p0 = torch.from_numpy(rng.normal(size=(C, 3)))
    rep = lambda a: None if a is None else a.repeat_interleave(k, 0)  # noqa: E731
    nodes = torch.stack([r, W0], -1)
    # replica tangents: a distribution direction, Re m (outer layer), Im m (core), a mix
    dp = torch.zeros((C * k, 3), dtype=f64)
    dmr = torch.zeros((C * k, layers), dtype=f64)
    dmi = torch.zeros((C * k, layers), dtype=f64)
    dfr = torch.zeros((C * k, 1), dtype=f64) if layers > 1 else None
    dp[0::k] = torch.tensor([1.0, -0.5, 0.25], dtype=f64)
    dmr[1::k, -1] = 1.0
    dmi[2::k, 0] = 1.0
    dp[3::k, 1] = 0.3
    dmr[3::k, 0] = -0.7
    dmi[3::k, -1] = 0.4
    if layers > 1:
        dfr[3::k, 0] = 1.3
    squeeze = (lambda a: a[:, 0]) if layers == 1 else (lambda a: a)  # noqa: E731
    with forward_ad.dual_level():
        dual = lambda a, t: forward_ad.make_dual(a, t)  # noqa: E731
        out = ensemble.ensemble_nodes(
            dual(rep(p0), dp), dual(squeeze(rep(m.real)), squeeze(dmr)),
            dual(squeeze(rep(m.imag)), squeeze(dmi)),
            None if fr is None else dual(rep(fr), dfr), rep(nodes), rep(dlnw), rep(wl),
            order=order, mu=mu, normalized=False, replicas=k)
        tg = [forward_ad.unpack_dual(o).tangent for o in out]
    h = 1e-5

    def at(s):
        w = W0.repeat_interleave(k, 0) * torch.exp(s * (rep(dlnw) * dp[:, None, :]).sum(-1))
        mm = rep(m) + s * (dmr + 1j * dmi)
        ff = None if fr is None else rep(fr) + s * dfr
        return fwd(torch.stack([rep(r), w], -1), rep(wl), mm, ff)[:3]

    plus, minus = at(h), at(-h)
    for t, a, b in zip(tg, plus, minus):
        fd = (a - b) / (2 * h)
        scale = fd.abs().amax() + 1e-300
        assert torch.allclose(t, fd, rtol=1e-7, atol=1e-7 * float(scale)), (t - fd).abs().max()
    assert torch.all(tg[0][:, 7] == 0) and torch.all(tg[0][1::k, 8:] == 0)


def test_shared_panels_integrate_every_member(monkeypatch):
    """Shared pass-1 panels (scatrs_amd.shared, engine.rs:257-296) on the CPU with closed-form
    per-node quantities: three lognormals of one particle at one wavelength share one panel set;
    each member meets rtol against its own closed form <C_ext> = pi r_g^2 exp(2 s^2), every
    member's panels are panels of the group, and the nodes solved are fewer than three separate
    runs solve. A panel limit leaves the members unconverged within it."""
    from scatrs_amd import adaptive, shared

    def q(r):
        ext = np.pi * r * r
        sca = 0.9 * ext * (1.0 + 0.3 * np.sin(40.0 * np.log(r)))
        return ext, sca, ext - sca, 0.1 * ext, np.full_like(r, 0.5)

    monkeypatch.setattr(adaptive, "_node_values", _stub_nodes(q))
    rtol = 1e-8
    laws = [(0.3, 1.8), (0.5, 1.4), (0.12, 2.2)]

    def tasks():
        return [adaptive.Task(0, np.array([rg, sg, 0.0]), 0.5, np.array([1.5 + 0j]), np.zeros(0))
                for rg, sg in laws]

    ts = tasks()
    n_shared = shared.pass1_groups(ts, [[0, 1, 2]], rtol, 4096, False, False)
    edges = set()
    for t, (rg, sg) in zip(ts, laws):
        assert t.converged
        assert np.all(np.isfinite(t.K)) and np.all(np.diff(t.a) > 0)
        s = np.log(sg)
        exact = np.pi * rg ** 2 * np.exp(2 * s * s)
        assert abs(t.K[:, 0].sum() / t.moments[0] - exact) <= 5 * rtol * exact
        err = np.abs(t.K - t.G).sum(0)
        assert err[0] <= rtol * abs(t.K[:, 0].sum()) and err[1] <= rtol * abs(t.K[:, 1].sum())
        edges |= set(zip(t.a.tolist(), t.b.tolist()))
    # one panel set: any two members' panels either coincide or do not overlap
    e = sorted(edges)
    for (a0, b0), (a1, b1) in zip(e[:-1], e[1:]):
        assert (a0, b0) == (a1, b1) or b0 <= a1
    alone = tasks()
    n_alone = adaptive.pass1(alone, rtol, 4096, False, False)
    assert n_shared < n_alone
    capped = tasks()
    shared.pass1_groups(capped, [[0, 1, 2]], rtol, 24, False, False)
    assert not all(t.converged for t in capped)
    assert len({(x, y) for t in capped for x, y in zip(t.a, t.b)}) <= 24


def test_shared_panels_extend_a_member_with_a_steep_integrand(monkeypatch):
    """With C_sca ~ r^6 a member's range grows past its area rule's end inside the group (onto
    the group's panels, or new ones), until the r^6 tail is below rtol/10; the closed form
    r_g^6 exp(18 s^2) is met; a gamma member beside it converges too."""
    from scatrs_amd import adaptive, shared

    def q(r):
        ext = r ** 6 + 0.01 * r * r
        return ext, r ** 6, 0.01 * r * r, r ** 6, np.zeros_like(r)

    monkeypatch.setattr(adaptive, "_node_values", _stub_nodes(q))
    r_g, sigma, rtol = 0.01, 2.0, 1e-7
    s = np.log(sigma)
    t = adaptive.Task(0, np.array([r_g, sigma, 0.0]), 1.0, np.array([1.5 + 0j]), np.zeros(0))
    g = adaptive.Task(1, np.array([0.5, 0.2, 0.0]), 1.0, np.array([1.5 + 0j]), np.zeros(0))
    _, t_hi, _ = adaptive.integration_range(0, t.p, rtol)
    shared.pass1_groups([t, g], [[0, 1]], rtol, 4096, False, False)
    assert t.converged and g.converged
    assert t.t_hi > t_hi + 1e-9
    sca = t.K[:, 1].sum() / t.moments[0]
    exact = r_g ** 6 * np.exp(18 * s * s)
    assert abs(sca - exact) <= 5 * rtol * exact


def test_pass2_row_rules_vectorized_equal_the_per_row_rules():
    """Pass 2's rows and their Greek rules from arrays (_panel_rows, _rules_for) equal the rules
    of the rows listed one by one (_row_rules), monodisperse tasks included."""
    from scatrs_amd import adaptive

    rng = np.random.default_rng(5)
    tasks = []
    for k in range(40):
        kind = adaptive.MONODISPERSE if k % 7 == 3 else 0
        q = adaptive.Task(kind, np.array([rng.uniform(0.05, 3.0), 1.5, 0.0]),
                          float(rng.uniform(0.3, 2.0)), np.array([1.5 + 0j]), np.zeros(0))
        n = int(rng.integers(1, 70))
        edges = np.sort(rng.uniform(-6.0, 3.0, n + 1))
        q.a, q.b = edges[:-1], edges[1:]
        tasks.append(q)
    rows = [(ti, lo, min(lo + adaptive._ROW_PANELS, q.a.size)) for ti, q in enumerate(tasks)
            if q.kind != adaptive.MONODISPERSE for lo in range(0, q.a.size, adaptive._ROW_PANELS)]
    mono = [ti for ti, q in enumerate(tasks) if q.kind == adaptive.MONODISPERSE]
    ti, end = adaptive._panel_rows(tasks)
    assert np.array_equal(ti, [r[0] for r in rows])
    wl = np.array([q.wl for q in tasks])
    x = np.concatenate([2.0 * np.pi * np.exp(end) / wl[ti],
                        [2.0 * np.pi * tasks[t].p[0] / tasks[t].wl for t in mono]])
    for order, n_angles, single in ((32, None, False), (63, None, True), (None, None, False),
                                    (16, 40, False)):
        old = adaptive._row_rules(tasks, rows + [(t, -1, -1) for t in mono], order, n_angles,
                                  single)
        assert np.array_equal(adaptive._rules_for(x, order, n_angles, single), old)
