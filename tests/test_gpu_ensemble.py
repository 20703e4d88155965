"""GPU parity of the fixed-path ensemble (scatrs_ensemble_fixed) against the oracle, through
the C ABI, and the invariants the reference's ensemble tests pin (scatrs/tests/ensemble.rs)."""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from scatrs_amd.errors import InvalidInputError

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def _gpu(kind, p0, p1, wl, m, fr=None, **kw):
    from scatrs_amd.ensemble import ensemble_fixed_device

    sc, gr, ph, terms = ensemble_fixed_device(_dev(np.asarray(kind, np.int32)), _dev(p0),
                                              _dev(p1), _dev(wl), _dev(m),
                                              None if fr is None else _dev(fr), **kw)
    return (sc.cpu().numpy(), None if gr is None else gr.cpu().numpy(),
            None if ph is None else ph.cpu().numpy(), terms)


def _check(sc, gr, ph, ref, tol):
    rs, rg, rp = ref["scalars"], ref["greek"], ref["phase"]
    rel = np.abs(sc - rs) / np.maximum(np.abs(rs), 1e-300)
    rel[7] = abs(sc[7] - rs[7])
    assert rel.max() <= tol, (rel, sc, rs)
    if gr is not None:
        scale = np.max(np.abs(rg[0]))
        assert np.max(np.abs(gr - rg)) / scale <= tol
    if ph is not None:
        scale = np.max(np.abs(rp[0]))
        assert np.max(np.abs(ph - rp)) / scale <= tol


def test_sulfate_three_wavelengths_matches_oracle():
    """The lognormal r_g = 0.08 um, sigma_g = 1.6, m = 1.43 + 1e-8 i at 0.35, 0.55, 1.0 um,
    128 radius nodes, Greek to 32 (the reference's CPU-runnable case)."""
    from scatrs_amd import workloads

    kind, p0, p1, wl, m = workloads.sulfate()
    sc, gr, ph, terms = _gpu(kind, p0, p1, wl, m, n_radius=128, n_angles=128, order=32)
    for e in range(3):
        ref = oracle.ensemble_fixed(0, p0[e], p1[e], wl[e], m[e], n_radius=128, n_angles=128,
                                    order=32)
        _check(sc[e], gr[e], ph[e], ref, 1e-10)
    # alpha1_0 = 1, alpha1_1 = 3 g (the projection is exact here: 128 >= N + 17)
    assert np.all(gr[:, 0, 0] == 1.0)
    assert np.allclose(gr[:, 0, 1], 3.0 * sc[:, 5], rtol=1e-11, atol=0)


def test_random_lognormals_match_oracle():
    rng = np.random.default_rng(11)
    n = 12
    p0 = np.exp(rng.uniform(np.log(0.01), np.log(0.6), n))
    p1 = rng.uniform(1.2, 2.0, n)
    wl = np.exp(rng.uniform(np.log(0.3), np.log(2.5), n))
    m = rng.uniform(1.3, 1.7, n) + 1j * np.exp(rng.uniform(np.log(1e-8), np.log(1e-1), n))
    sc, gr, ph, _ = _gpu(np.zeros(n), p0, p1, wl, m, n_radius=96, n_angles=200, order=64)
    for e in range(n):
        ref = oracle.ensemble_fixed(0, p0[e], p1[e], wl[e], m[e], n_radius=96, n_angles=200,
                                    order=64)
        _check(sc[e], gr[e], ph[e], ref, 1e-9)


def test_scalars_only_calls_solve_one_node_per_wavefront():
    """A call without phase or Greek outputs takes no angular products; its node solves then run
    one node per wavefront when the launch is small (SCATRS_TUNE_SPHERE_WAVE = 1 forces it, 0
    keeps one node per lane): both match the oracle's scalars and each other to rounding."""
    from scatrs_amd import _native, tune

    rng = np.random.default_rng(23)
    n = 16
    p0 = np.exp(rng.uniform(np.log(0.01), np.log(2.0), n))
    p1 = rng.uniform(1.2, 2.0, n)
    wl = np.exp(rng.uniform(np.log(0.3), np.log(2.5), n))
    m = rng.uniform(1.3, 1.7, n) + 1j * np.exp(rng.uniform(np.log(1e-8), np.log(1e-1), n))
    ctx = _native.context(0)
    out = {}
    try:
        for mode in (1, 0):
            ctx.set_tuning(tune.KEYS["sphere_wave"], mode)
            out[mode] = _gpu(np.zeros(n), p0, p1, wl, m, n_radius=96, n_angles=16, order=None,
                             want_phase=False)
    finally:
        ctx.set_tuning(tune.KEYS["sphere_wave"], tune.DEFAULTS["sphere_wave"])
    sw, sl = out[1][0], out[0][0]
    assert out[1][1] is None and out[1][2] is None
    for e in range(n):
        ref = oracle.ensemble_fixed(0, p0[e], p1[e], wl[e], m[e], n_radius=96, n_angles=16,
                                    order=-1)
        _check(sw[e], None, None, ref, 1e-9)
    rel = np.abs(sw - sl) / np.maximum(np.abs(sl), 1e-300)
    rel[:, 7] = np.abs(sw[:, 7] - sl[:, 7])
    assert rel.max() < 1e-11, rel.max()


@pytest.mark.parametrize("kernel", [0, 2])
def test_node_solves_one_per_wavefront_stage_the_coefficients(kernel):
    """With phase and Greek outputs the node solves of a small launch, one node per wavefront,
    also stage each node's coefficients (FP64 for the FP64 matrix-core kernel, fp32 for the
    fp16-split one) and its coefficient exponent: the outputs match the one-node-per-lane
    solves' (to rounding in FP64; to the split products' 2^-22 with kernel 2) and the oracle."""
    from scatrs_amd import _native, tune

    rng = np.random.default_rng(29)
    n = 8
    p0 = np.exp(rng.uniform(np.log(0.01), np.log(1.0), n))
    p1 = rng.uniform(1.2, 2.0, n)
    wl = np.exp(rng.uniform(np.log(0.3), np.log(2.5), n))
    m = rng.uniform(1.3, 1.7, n) + 1j * np.exp(rng.uniform(np.log(1e-8), np.log(1e-1), n))
    ctx = _native.context(0)
    out = {}
    try:
        ctx.set_angular_kernel(kernel)
        for mode in (1, 0):
            ctx.set_tuning(tune.KEYS["sphere_wave"], mode)
            out[mode] = _gpu(np.zeros(n), p0, p1, wl, m, n_radius=64, n_angles=96, order=32)
    finally:
        ctx.set_tuning(tune.KEYS["sphere_wave"], tune.DEFAULTS["sphere_wave"])
        ctx.set_angular_kernel(0)
    tol = 1e-11 if kernel == 0 else 1e-6
    (sw, gw, pw, _), (sl, gl, pl, _) = out[1], out[0]
    rel = np.abs(sw - sl) / np.maximum(np.abs(sl), 1e-300)
    rel[:, 7] = np.abs(sw[:, 7] - sl[:, 7])
    assert rel.max() < 1e-11, rel.max()
    for e in range(n):
        assert np.max(np.abs(gw[e] - gl[e])) / np.max(np.abs(gl[e, 0])) < tol
        assert np.max(np.abs(pw[e] - pl[e])) / np.max(np.abs(pl[e, 0])) < tol
        ref = oracle.ensemble_fixed(0, p0[e], p1[e], wl[e], m[e], n_radius=64, n_angles=96,
                                    order=32)
        _check(sw[e], gw[e], pw[e], ref, 1e-9 if kernel == 0 else 2e-6)


@pytest.mark.parametrize("kernel", [0, 2])
def test_tangent_node_solves_one_per_wavefront(kernel):
    """scatrs_ensemble_nodes_tangent (the coefficients moved by sign eps dc along Re m or Im m)
    with its node solves one node per wavefront (SCATRS_TUNE_SPHERE_WAVE = 1) against one per
    lane (0), with and without angular products: equal to rounding (to the split products'
    2^-22 with the fp16-split kernel)."""
    from scatrs_amd import _native, tune
    from scatrs_amd.ensemble import ensemble_nodes_tangent_device

    rng = np.random.default_rng(31)
    count, k = 6, 40
    r = np.sort(np.exp(rng.uniform(np.log(0.05), np.log(8.0), (count, k))), 1)
    w = rng.uniform(0.1, 1.0, (count, k))
    nodes = _dev(np.stack([r, w], -1))
    wl = _dev(np.exp(rng.uniform(np.log(0.3), np.log(2.0), count)))
    m = _dev(rng.uniform(1.3, 1.7, count) + 1j * np.exp(rng.uniform(np.log(1e-6), np.log(0.1),
                                                                      count)))
    ctx = _native.context(0)
    tol = 1e-11 if kernel == 0 else 1e-6
    try:
        ctx.set_angular_kernel(kernel)
        for direction in (0, 1):
            for order, phase in ((None, False), (16, True)):
                out = {}
                for mode in (1, 0):
                    ctx.set_tuning(tune.KEYS["sphere_wave"], mode)
                    sc, gr, ph, _ = ensemble_nodes_tangent_device(
                        nodes, wl, m, direction=direction, sign=1, n_angles=48, order=order,
                        want_phase=phase)
                    out[mode] = (sc.cpu().numpy(), None if gr is None else gr.cpu().numpy(),
                                 None if ph is None else ph.cpu().numpy())
                (sw, gw, pw), (sl, gl, pl) = out[1], out[0]
                rel = np.abs(sw - sl) / np.maximum(np.abs(sl), 1e-300)
                rel[:, 7] = np.abs(sw[:, 7] - sl[:, 7])
                assert rel.max() < 1e-11, (direction, order, rel.max())
                if gr is not None:
                    for e in range(count):
                        assert np.max(np.abs(gw[e] - gl[e])) / np.max(np.abs(gl[e, 0])) < tol
                        assert np.max(np.abs(pw[e] - pl[e])) / np.max(np.abs(pl[e, 0])) < tol
    finally:
        ctx.set_tuning(tune.KEYS["sphere_wave"], tune.DEFAULTS["sphere_wave"])
        ctx.set_angular_kernel(0)


def test_gamma_and_unnormalized_match_oracle():
    p0 = np.array([0.2, 0.5, 1.1])
    p1 = np.array([0.05, 0.2, 0.38])
    wl = np.array([0.55, 1.0, 2.0])
    m = np.array([1.5 + 0.01j, 1.33 + 1e-6j, 1.7 + 0.3j])
    sc, gr, ph, _ = _gpu(np.ones(3), p0, p1, wl, m, n_radius=64, n_angles=96, order=40,
                         normalized=False)
    for e in range(3):
        ref = oracle.ensemble_fixed(1, p0[e], p1[e], wl[e], m[e], n_radius=64, n_angles=96,
                                    order=40, normalized=False)
        _check(sc[e], gr[e], ph[e], ref, 1e-9)
    # unnormalized: alpha1_0 = <C_sca>
    assert np.allclose(gr[:, 0, 0], sc[:, 1], rtol=1e-12)


def test_layered_ensembles_match_oracle():
    for L in (2, 3, 5):
        rng = np.random.default_rng(100 + L)
        n = 4
        p0 = np.exp(rng.uniform(np.log(0.05), np.log(0.8), n))
        p1 = rng.uniform(0.05, 0.35, n)
        wl = np.exp(rng.uniform(np.log(0.3), np.log(2.5), n))
        fr = np.cumsum(rng.dirichlet(np.full(L, 2.0), n), axis=1)[:, :-1]
        m = rng.uniform(1.3, 2.0, (n, L)) + 1j * np.exp(rng.uniform(np.log(1e-6), 0.0, (n, L)))
        sc, gr, ph, _ = _gpu(np.ones(n), p0, p1, wl, m, fr, n_radius=48, n_angles=64,
                             order=24)
        for e in range(n):
            ref = oracle.ensemble_fixed(1, p0[e], p1[e], wl[e], m[e], fractions=fr[e],
                                        n_radius=48, n_angles=64, order=24)
            _check(sc[e], gr[e], ph[e], ref, 1e-8)


def test_phase_normalization_and_greek_invariants():
    """1/2 int F11 dmu = 1 and alpha1_0 = 1, alpha1_1 = 3 g (scatrs/tests/ensemble.rs:114-202)."""
    sc, gr, ph, _ = _gpu([0], [0.3], [1.5], [0.6], [1.45 + 0.001j], n_radius=128,
                         n_angles=256, order=64)
    mu, w = np.polynomial.legendre.leggauss(256)
    assert abs(0.5 * np.sum(w * ph[0, 0]) - 1.0) < 1e-12
    assert np.array_equal(ph[0, 0], ph[0, 2]) and np.array_equal(ph[0, 3], ph[0, 5])
    assert gr[0, 0, 0] == 1.0 and abs(gr[0, 0, 1] - 3 * sc[0, 5]) < 1e-11
    # C_ext = C_sca + C_abs; ssa = C_sca / C_ext
    assert abs(sc[0, 0] - sc[0, 1] - sc[0, 2]) <= 1e-14 * sc[0, 0]
    assert sc[0, 4] == sc[0, 1] / sc[0, 0]


def test_errors_map_like_the_reference():
    from scatrs_amd import InvalidInputError, OutsideValidatedDomainError

    with pytest.raises(InvalidInputError, match="ensemble 1: the lognormal width"):
        _gpu([0, 0], [0.1, 0.1], [1.5, 0.9], [0.5, 0.5], [1.5 + 0j, 1.5 + 0j])
    with pytest.raises(InvalidInputError, match="wavelength must be finite"):
        _gpu([0], [0.1], [1.5], [-0.5], [1.5 + 0j])
    with pytest.raises(OutsideValidatedDomainError, match="at wavelength"):
        _gpu([0], [2000.0], [1.5], [0.5], [1.5 + 0j], n_radius=16, n_angles=16, order=4)
    sc, _, _, _ = _gpu([0], [0.1], [1.5], [0.5], [1.5 + 0j], n_radius=0 + 8, n_angles=8,
                       order=4)
    assert np.all(np.isfinite(sc))


def test_fp16_split_angular_kernel_within_1e6_of_the_oracle():
    """The fp16-split matrix-core angular kernel (scatrs_set_angular_kernel 2): alpha1 within
    1e-6 of max|alpha1| and the phase matrix within 1e-6 of max F11 against the FP64 oracle, on
    random lognormals and on the largest-N ensembles of the 4096 x 64 workload at its full
    resolution (512 radius nodes, 1024 angles, Greek to 256); the scalars, which do not go
    through the angular products, stay at FP64 parity."""
    from scatrs_amd import _native, workloads

    ctx = _native.context(0)
    kind, p0, p1, wl, m = workloads.lognormals()
    score = p0 * p1 ** 5 / wl
    big = list(np.argsort(score)[::-1][:2])
    rng = np.random.default_rng(31)
    small = list(rng.choice(kind.size, 4, replace=False))
    idx = np.array(big + small)
    kernel = 2
    ctx.set_angular_kernel(kernel)
    try:
        sc, gr, ph, _ = _gpu(kind[idx], p0[idx], p1[idx], wl[idx], m[idx], n_radius=512,
                             n_angles=1024, order=256)
    finally:
        ctx.set_angular_kernel(0)
    for r, e in enumerate(idx):
        ref = oracle.ensemble_fixed(0, p0[e], p1[e], wl[e], m[e], n_radius=512, n_angles=1024,
                                    order=256)
        rs = ref["scalars"]
        rel = np.abs(sc[r] - rs) / np.maximum(np.abs(rs), 1e-300)
        rel[7] = abs(sc[r][7])
        assert rel.max() <= 1e-10, (e, rel)
        ga = np.max(np.abs(gr[r][0] - ref["greek"][0])) / np.max(np.abs(ref["greek"][0]))
        gall = np.max(np.abs(gr[r] - ref["greek"])) / np.max(np.abs(ref["greek"][0]))
        pe = np.max(np.abs(ph[r] - ref["phase"])) / np.max(np.abs(ref["phase"][0]))
        print(f"kernel {kernel} ensemble {e}: alpha1 {ga:.2e}, greek {gall:.2e}, phase {pe:.2e}")
        assert ga <= 1e-6 and gall <= 1e-6 and pe <= 1e-6, (e, ga, gall, pe)


def test_fp64_matrix_core_kernel_beyond_its_lds_node_list():
    """The FP64 matrix-core angular kernel keeps an ensemble's node list in LDS up to 8192
    radius nodes and reads it from memory beyond: at 8200 nodes (a broad lognormal, so that
    every group of nodes has terms) its tables, Greek coefficients and phase matrix equal the
    plain FP64 vector kernel's to rounding, and at 8192 nodes (in LDS) likewise."""
    from scatrs_amd import _native

    ctx = _native.context(0)
    kind, p0, p1 = np.zeros(2, np.int32), np.array([0.3, 1.0]), np.array([1.8, 1.5])
    wl, m = np.array([0.55, 1.0]), np.array([1.45 + 1e-3j, 1.33 + 1e-6j])
    for nr in (8192, 8200):
        out = {}
        for kernel in (0, 1):
            ctx.set_angular_kernel(kernel)
            try:
                out[kernel] = _gpu(kind, p0, p1, wl, m, n_radius=nr, n_angles=96, order=40)
            finally:
                ctx.set_angular_kernel(0)
        (s0, g0, f0, _), (s1, g1, f1, _) = out[0], out[1]
        assert np.max(np.abs(s0 - s1) / np.maximum(np.abs(s1), 1e-300)) <= 1e-13, nr
        for e in range(2):
            assert np.max(np.abs(g0[e] - g1[e])) / np.max(np.abs(g1[e][0])) <= 1e-12, (nr, e)
            assert np.max(np.abs(f0[e] - f1[e])) / np.max(np.abs(f1[e][0])) <= 1e-12, (nr, e)


def test_mixture_and_multimodal_match_the_oracle():
    """Mixture (a lognormal of spheres, a gamma law of coated spheres, a bimodal lognormal of
    spheres) and a MultiModal Ensemble through the GPU path (one library call per layer count,
    mixed by mix_parts) against the oracle's mixture, normalized and not."""
    from scatrs_amd.ensemble import SCALARS, Ensemble, Mixture
    from scatrs_amd.particles import LayeredSphere, Sphere
    from scatrs_amd.size import Gamma, LogNormal, MultiModal

    wl = np.array([0.45, 0.8, 1.3])
    sph = Sphere(1.45 + 0.01j)
    coated = LayeredSphere([0.7], [1.6 + 0.1j, 1.35 + 1e-5j])
    bimodal = MultiModal(((3.0, LogNormal(0.05, 1.4)), (1.0, LogNormal(0.6, 1.8))))
    mix = Mixture(((2.0, Ensemble(sph, LogNormal(0.3, 1.5))),
                   (1.0, Ensemble(coated, Gamma(0.4, 0.15))),
                   (1.0, Ensemble(sph, bimodal))))
    kw = dict(n_radius=64, n_angles=48)
    parts = [(0.5, 0, 0.3, 1.5, 1.45 + 0.01j, None),
             (0.25, 1, 0.4, 0.15, np.array([1.6 + 0.1j, 1.35 + 1e-5j]), np.array([0.7])),
             (0.25 * 0.75, 0, 0.05, 1.4, 1.45 + 0.01j, None),
             (0.25 * 0.25, 0, 0.6, 1.8, 1.45 + 0.01j, None)]
    for normalized in (True, False):
        r = mix.compute(wl, greek=16, normalized=normalized, **kw)
        b = Ensemble(sph, bimodal).compute(wl, greek=16, normalized=normalized, **kw)
        for j, w in enumerate(wl):
            ref = oracle.ensemble_mixture(parts, w, order=16, normalized=normalized, **kw)
            got = np.array([getattr(r, k)[j] for k in SCALARS])
            rel = np.abs(got - ref["scalars"]) / np.maximum(np.abs(ref["scalars"]), 1e-300)
            assert rel.max() <= 1e-10, (normalized, j, rel)
            a1 = r.greek.alpha1[j]
            assert np.max(np.abs(a1 - ref["greek"][0])) <= 1e-10 * np.max(np.abs(ref["greek"][0]))
            assert np.max(np.abs(r.phase.f11[j] - ref["phase"][0])) <= \
                1e-10 * np.max(np.abs(ref["phase"][0]))
            bref = oracle.ensemble_mixture([(0.75, *parts[2][1:]), (0.25, *parts[3][1:])], w,
                                           order=16, normalized=normalized, **kw)
            assert abs(b.c_sca[j] / bref["scalars"][1] - 1) <= 1e-10
            assert abs(b.effective_radius[j] / bref["scalars"][8] - 1) <= 1e-10
            assert np.max(np.abs(b.greek.beta1[j] - bref["greek"][4])) <= \
                1e-10 * np.max(np.abs(bref["greek"][0]))


def test_mixture_gradient_through_autograd_matches_central_differences():
    """ensemble_fixed(..., normalized=False) for the parts and mix_parts: the gradient of a
    loss on the mixture's ssa, g and alpha1 with respect to the parts' parameters, indices and
    number fractions, against central differences of the same GPU forward (h = 1e-6
    relative)."""
    from scatrs_amd.ensemble import ensemble_fixed, mix_parts

    kw = dict(n_radius=48, n_angles=64, order=12, normalized=False)
    p0 = torch.tensor([0.25, 0.7], dtype=torch.float64, device="cuda")
    p1 = torch.tensor([1.5, 0.2], dtype=torch.float64, device="cuda")
    kind = torch.tensor([0, 1], dtype=torch.int32, device="cuda")
    wl = torch.tensor([0.6, 0.6], dtype=torch.float64, device="cuda")
    mr = torch.tensor([1.45, 1.55], dtype=torch.float64, device="cuda")
    mi = torch.tensor([0.01, 1e-3], dtype=torch.float64, device="cuda")
    f = torch.tensor([0.3, 0.7], dtype=torch.float64, device="cuda")

    def loss(p0, p1, mr, mi, f):
        sc, gr, ph = ensemble_fixed(kind, p0, p1, wl, mr, mi, **kw)
        s, g, _ = mix_parts(sc[:, None], gr[:, None], ph[:, None], f, True)
        return s[0, 4] ** 2 + 3.0 * s[0, 5] + (g[0, 0, :6] * torch.arange(6, device="cuda")).sum()

    xs = [t.clone().requires_grad_(True) for t in (p0, p1, mr, mi, f)]
    loss(*xs).backward()
    for i, x in enumerate(xs):
        for k in range(2):
            h = 1e-6 * abs(float(x[k].detach()))
            up = [t.detach().clone() for t in xs]
            dn = [t.detach().clone() for t in xs]
            up[i][k] += h
            dn[i][k] -= h
            fd = (float(loss(*up)) - float(loss(*dn))) / (2 * h)
            got = float(x.grad[k])
            assert abs(got - fd) <= 1e-6 * max(abs(fd), 1e-3), (i, k, got, fd)


def test_new_distributions_match_the_oracle():
    """Modified gamma, power law and monodisperse ensembles of spheres and coated spheres through
    the C ABI (the third parameter in d_p2) against the oracle; the Python classes give the same
    numbers; a monodisperse ensemble is its sphere."""
    from scatrs_amd.ensemble import Ensemble
    from scatrs_amd.particles import Sphere
    from scatrs_amd.size import ModifiedGamma, Monodisperse, PowerLaw

    kind = np.array([2, 2, 3, 3, 4, 4], np.int32)
    p0 = np.array([2.0, 0.5, 3.0, 0.0, 0.4, 2.5])
    p1 = np.array([3.0, 8.0, 0.05, 0.2, 0.0, 0.0])
    p2 = np.array([1.2, 0.7, 2.0, 0.6, 0.0, 0.0])
    wl = np.array([0.55, 0.8, 0.45, 1.1, 0.55, 0.6])
    m = np.array([1.45 + 0.01j, 1.33 + 1e-6j, 1.5 + 0.001j, 1.6 + 0.2j, 1.5 + 0.01j, 1.4 + 0j])
    kw = dict(n_radius=96, n_angles=64, order=24)
    from scatrs_amd.ensemble import ensemble_fixed_device

    sc, gr, ph, _ = ensemble_fixed_device(_dev(kind), _dev(p0), _dev(p1), _dev(wl), _dev(m),
                                          p2=_dev(p2), **kw)
    sc, gr, ph = sc.cpu().numpy(), gr.cpu().numpy(), ph.cpu().numpy()
    for e in range(kind.size):
        ref = oracle.ensemble_fixed(int(kind[e]), p0[e], p1[e], wl[e], m[e], p2=p2[e], **kw)
        _check(sc[e], gr[e], ph[e], ref, 1e-10)
    # the same through the classes
    ck = dict(greek=24, n_radius=96, n_angles=64)
    r = Ensemble(Sphere(m[0]), ModifiedGamma(2.0, 3.0, 1.2)).compute(0.55, **ck)
    assert abs(r.c_ext / sc[0, 0] - 1) < 1e-14 and abs(r.g / sc[0, 5] - 1) < 1e-14
    r = Ensemble(Sphere(m[2]), PowerLaw(3.0, 0.05, 2.0)).compute(0.45, **ck)
    assert abs(r.c_sca / sc[2, 1] - 1) < 1e-14
    r = Ensemble(Sphere(m[4]), Monodisperse(0.4)).compute(0.55, **ck)
    x = 2 * np.pi * 0.4 / 0.55
    q = oracle.efficiencies(x, oracle.mie_solve(x, m[4]))
    assert abs(r.c_ext / (np.pi * 0.16 * q[0]) - 1) < 1e-12
    assert abs(r.effective_radius / 0.4 - 1) < 1e-14
    with pytest.raises(InvalidInputError, match="gamma must be finite and > 0"):
        _gpu([2], [1.0], [1.0], [0.5], [1.5 + 0j], None, **kw)
    with pytest.raises(InvalidInputError, match="min_radius must be < max_radius"):
        ensemble_fixed_device(_dev(np.array([3], np.int32)), _dev(np.array([1.0])),
                              _dev(np.array([0.5])), _dev(np.array([0.5])),
                              _dev(np.array([1.5 + 0j])), p2=_dev(np.array([0.2])), **kw)


@pytest.mark.parametrize("kernel", [0, 1, 2])
def test_phase_at_the_callers_cosines(kernel):
    """scatrs_set_phase_cosines: the phase matrix at arbitrary cosines (unsorted, repeated, 0 and
    +-1) with each angular kernel against the oracle (1e-10, the fp16-split kernel 1e-6 of max
    F11); at the rule's own cosines the rule's phase matrix; the scalars and Greek
    coefficients unchanged; through Ensemble and Mixture."""
    from scatrs_amd import _native
    from scatrs_amd.ensemble import Ensemble, Mixture, ensemble_fixed_device
    from scatrs_amd.particles import Sphere
    from scatrs_amd.quadrature import gauss_legendre
    from scatrs_amd.size import LogNormal, PowerLaw

    ctx = _native.context(0)
    kind = np.array([0, 1, 3], np.int32)
    p0, p1, p2 = np.array([0.3, 0.6, 3.0]), np.array([1.6, 0.15, 0.05]), np.array([0, 0, 1.5])
    wl = np.array([0.55, 0.9, 0.45])
    m = np.array([1.45 + 0.01j, 1.38 + 1e-5j, 1.5 + 0.002j])
    mu = np.array([0.3, -1.0, 1.0, 0.0, -0.3, 0.95, 0.3, -0.7071])
    kw = dict(n_radius=64, n_angles=96, order=24)
    args = [_dev(kind), _dev(p0), _dev(p1), _dev(wl), _dev(m)]
    ctx.set_angular_kernel(kernel)
    try:
        s0, g0, f0, _ = ensemble_fixed_device(*args, p2=_dev(p2), **kw)
        s1, g1, f1, _ = ensemble_fixed_device(*args, p2=_dev(p2), mu=mu, **kw)
        rule, _ = gauss_legendre(96)
        _, _, f2, _ = ensemble_fixed_device(*args, p2=_dev(p2), mu=rule, **kw)
    finally:
        ctx.set_angular_kernel(0)
    assert torch.equal(s0, s1) and torch.equal(g0, g1)
    # the library's own cosines label the outputs: at them the caller's-cosine path gives the
    # rule's phase matrix, its angular tables built apart (rounding only)
    assert float((f0 - f2).abs().max() / f0[:, 0].abs().max()) <= 1e-13
    f1 = f1.cpu().numpy()
    tol = 1e-6 if kernel == 2 else 1e-10
    for e in range(3):
        ref = oracle.ensemble_fixed(int(kind[e]), p0[e], p1[e], wl[e], m[e], p2=p2[e], mu=mu,
                                    **kw)["phase"]
        err = np.max(np.abs(f1[e] - ref)) / np.max(np.abs(ref[0]))
        assert err <= tol, (kernel, e, err)
    r = Ensemble(Sphere(m[0]), LogNormal(0.3, 1.6)).compute(wl[:1], greek=24, mu=mu, n_radius=64,
                                                             n_angles=96)
    assert r.phase.f11.shape == (1, mu.size) and np.array_equal(r.phase.mu, mu)
    mix = Mixture(((1.0, Ensemble(Sphere(m[0]), LogNormal(0.3, 1.6))),
                   (2.0, Ensemble(Sphere(m[2]), PowerLaw(3.0, 0.05, 1.5)))))
    rm = mix.compute(wl[:1], greek=24, mu=mu, n_radius=64, n_angles=96)
    assert rm.phase.f33.shape == (1, mu.size) and np.all(np.isfinite(rm.phase.f11))
    with pytest.raises(InvalidInputError, match=r"mu must be in \[-1, 1\]"):
        ctx.set_phase_cosines([0.5, 1.5])


def test_tabulated_ensembles_match_the_oracle():
    """A tabulated law (its segments on the device, mixed by their number integrals) alone, as
    a mode of a multimodal law and in a mixture, against the oracle's mixture of the same
    segments; the segment kind's errors."""
    from scatrs_amd.ensemble import SCALARS, Ensemble, Mixture
    from scatrs_amd.particles import Sphere
    from scatrs_amd.size import LogNormal, MultiModal, Tabulated

    radius, density = [0.05, 0.2, 0.6, 1.5], [0.0, 3.0, 1.0, 0.2]
    tab = Tabulated(radius, density)
    m = 1.47 + 0.004j
    wl = np.array([0.45, 0.9])
    kw = dict(n_radius=64, n_angles=64)
    segs, w = tab.segments()
    f = w / w.sum()
    parts = [(fi, 5, r0, r1, m, None, tilt) for fi, (r0, r1, tilt) in zip(f, segs)]
    r = Ensemble(Sphere(m), tab).compute(wl, greek=16, **kw)
    mm = Ensemble(Sphere(m), MultiModal(((1.0, tab), (1.0, LogNormal(0.3, 1.5))))).compute(
        wl, greek=16, **kw)
    mix = Mixture(((3.0, Ensemble(Sphere(m), tab)),)).compute(wl, greek=16, **kw)
    for j, lam in enumerate(wl):
        ref = oracle.ensemble_mixture(parts, lam, order=16, **kw)
        got = np.array([getattr(r, k)[j] for k in SCALARS])
        rel = np.abs(got - ref["scalars"]) / np.maximum(np.abs(ref["scalars"]), 1e-300)
        assert rel.max() <= 1e-10, (j, rel)
        assert np.max(np.abs(r.greek.alpha1[j] - ref["greek"][0])) <= 1e-10 * abs(ref["greek"][0]).max()
        assert abs(mix.c_ext[j] / r.c_ext[j] - 1) < 1e-14
        both = [(0.5 * p[0], *p[1:]) for p in parts] + [(0.5, 0, 0.3, 1.5, m, None)]
        rb = oracle.ensemble_mixture(both, lam, order=16, **kw)
        assert abs(mm.c_sca[j] / rb["scalars"][1] - 1) <= 1e-10
    with pytest.raises(InvalidInputError, match="tilt must be in"):
        _gpu([5], [0.1], [0.2], [0.5], [1.5 + 0j], None, n_radius=16, n_angles=16, order=4,
             p2=_dev(np.array([1.5])))
