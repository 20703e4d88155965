"""The layered-sphere workload at its own size and resolution on the GPU: 2048 gamma
distributions (r_eff log-uniform on [0.05, 3] um, v_eff uniform on [0.05, 0.4]) of a core and
1-4 shells (fractions from Dirichlet(2), Re m uniform on [1.3, 2.0], Im m log-uniform on
[1e-6, 1]) x 32 wavelengths, 256 radius nodes, 512 angles, Greek to 128, forward and VJP for
every layer index and radius fraction.

Every layer count (2, 3, 4, 5 layers) runs over all of its rows, with:
  * size-independent properties on every row: finite outputs and gradients,
    C_ext = C_sca + C_abs, ssa = C_sca / C_ext, alpha1_0 = 1, and, on the rows whose every node
    has N + 1 <= 512 (where the 512-point rule integrates F11 mu exactly), alpha1_1 = 3 g and
    1/2 sum w F11 = 1 (scatrs/tests/ensemble_layered.rs:188-274 pins the same invariants);
  * the forward against the oracle (oracle.ensemble_fixed) on sampled rows, the largest
    particle of each layer count among them;
  * the VJP against the oracle's forward-mode Jacobian contracted with the cotangents the
    benchmark uses, on sampled rows.
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

NR, NA, ORDER = 256, 512, 128


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def _workload():
    from scatrs_amd import workloads

    return sorted(workloads.layered().items())


def _n_max_per_row(p0, p1, wl):
    """The largest N over each row's radius nodes: N at the top of the fixed range (the
    truncation grows with x)."""
    from scatrs_amd.adaptive import size_range
    from scatrs_amd.mie import default_n_max

    out = np.zeros(p0.size, np.int64)
    # rows of one distribution share (p0, p1): the range once per distribution
    cache = {}
    for i in range(p0.size):
        key = (p0[i], p1[i])
        if key not in cache:
            cache[key] = size_range(1, [p0[i], p1[i], 0.0])[1]
        x = 2.0 * np.pi * np.exp(cache[key]) / wl[i]
        out[i] = default_n_max(x)
    return out


def _sampled_rows(p0, wl, rng, k):
    """The row with the largest r_eff / wavelength (the largest particles) and k - 1 others."""
    big = int(np.argmax(p0 / wl))
    rest = rng.choice(np.setdiff1d(np.arange(p0.size), [big]), k - 1, replace=False)
    return [big, *[int(r) for r in rest]]


@pytest.mark.parametrize("kernel", [0, 2])
def test_layered_workload_forward(kernel):
    """Every row of every layer count, with the FP64 matrix-core angular kernel (0) and the
    fp16-split one (2, the fused path)."""
    from scatrs_amd import _native
    from scatrs_amd.ensemble import ensemble_fixed_device

    ctx = _native.context(0)
    # kernel 2 carries the fp16 split's ~1e-7 of the largest table element
    tol_tables = 1e-9 if kernel == 0 else 2e-6
    ctx.set_angular_kernel(kernel)
    try:
        seen = set()
        for L, (kind, p0, p1, wl, fr, m) in _workload():
            print(f"layered workload: L={L}, {kind.size} rows, kernel {kernel}", flush=True)
            seen.add(L)
            sc, gr, ph, terms = ensemble_fixed_device(
                _dev(kind), _dev(p0), _dev(p1), _dev(wl), _dev(m), _dev(fr), n_radius=NR,
                n_angles=NA, order=ORDER)
            assert terms > 0
            sc_h = sc.cpu().numpy()
            gr_h = gr.cpu().numpy()
            # the size-independent properties, on every row
            assert torch.isfinite(sc).all() and torch.isfinite(gr).all()
            assert torch.isfinite(ph).all()
            bal = np.abs(sc_h[:, 0] - sc_h[:, 1] - sc_h[:, 2]) / sc_h[:, 0]
            assert bal.max() <= 1e-13, (L, bal.max())
            assert np.array_equal(sc_h[:, 4], sc_h[:, 1] / sc_h[:, 0])
            assert np.all(gr_h[:, 0, 0] == 1.0)
            exact = _n_max_per_row(p0, p1, wl) + 1 <= NA
            assert exact.sum() > kind.size // 4, (L, exact.sum())
            a11 = np.abs(gr_h[exact, 0, 1] - 3.0 * sc_h[exact, 5])
            assert a11.max() <= 3.0 * tol_tables + 1e-12, (L, a11.max())
            mu, w = oracle.gauss_legendre(NA)
            norm = 0.5 * (ph[:, 0, :].cpu().numpy() @ w)
            assert np.abs(norm[exact] - 1.0).max() <= tol_tables + 1e-12, (L,)
            # the forward against the oracle on sampled rows
            rng = np.random.default_rng(40 + L)
            ph_h = ph.cpu().numpy()
            for e in _sampled_rows(p0, wl, rng, 4):
                ref = oracle.ensemble_fixed(1, p0[e], p1[e], wl[e], m[e], fractions=fr[e],
                                            n_radius=NR, n_angles=NA, order=ORDER)
                rs = ref["scalars"]
                rel = np.abs(sc_h[e] - rs) / np.maximum(np.abs(rs), 1e-300)
                rel[7] = abs(sc_h[e, 7] - rs[7])
                assert rel.max() <= 1e-9, (L, e, rel)
                ga = np.max(np.abs(gr_h[e] - ref["greek"])) / np.max(np.abs(ref["greek"][0]))
                assert ga <= tol_tables, (L, e, ga)
                pa = np.max(np.abs(ph_h[e] - ref["phase"])) / np.max(np.abs(ref["phase"][0]))
                assert pa <= tol_tables, (L, e, pa)
            del sc, gr, ph
        assert seen == {2, 3, 4, 5}
    finally:
        ctx.set_angular_kernel(0)


def test_layered_workload_vjp():
    """The reverse pass over every row of every layer count with the benchmark's cotangents
    (standard normal per output, seeded by the layer count), with the fused forward tables:
    finite gradients everywhere, and on sampled rows the oracle's Jacobian contracted with the
    same cotangents, within 1e-8 of sum |cotangent x derivative| (the fp16-split forward tables
    carry ~1e-7 of the largest table element into the Greek cotangents' part, whose share of
    that sum is small)."""
    from scatrs_amd import _native
    from scatrs_amd.ensemble import ensemble_fixed_vjp_device

    ctx = _native.context(0)
    for L, (kind, p0, p1, wl, fr, m) in _workload():
        rng = np.random.default_rng(L)
        bs = rng.normal(size=(kind.size, 11))
        bg = rng.normal(size=(kind.size, 6, ORDER + 1))
        grads = {}
        for kernel in (0, 2):
            print(f"layered workload VJP: L={L}, {kind.size} rows, kernel {kernel}", flush=True)
            ctx.set_angular_kernel(kernel)
            try:
                g, _ = ensemble_fixed_vjp_device(
                    _dev(kind), _dev(p0), _dev(p1), _dev(wl), _dev(m), _dev(fr),
                    bar_scalars=_dev(bs), bar_greek=_dev(bg), n_radius=NR, n_angles=NA,
                    order=ORDER)
            finally:
                ctx.set_angular_kernel(0)
            assert g.shape == (kind.size, 3 + 2 * L + L - 1)
            assert torch.isfinite(g).all(), (L, kernel)
            grads[kernel] = g.cpu().numpy()
            del g
            # the gamma law has two parameters: the third column is exactly 0
            assert np.all(grads[kernel][:, 2] == 0.0)
        for e in _sampled_rows(p0, wl, np.random.default_rng(50 + L), 2):
            J = oracle.ensemble_fixed_jacobian(1, p0[e], p1[e], wl[e], m[e], fr[e], NR, NA,
                                               ORDER)
            t = np.concatenate([J["d_scalars"] * bs[e][None, :],
                                (J["d_greek"] * bg[e][None]).reshape(J["d_greek"].shape[0], -1)],
                               axis=1)
            ref, scale = t.sum(axis=1), np.abs(t).sum(axis=1)
            for kernel, tol in ((0, 1e-9), (2, 2e-6)):
                err = np.abs(grads[kernel][e] - ref) / np.where(scale > 0, scale, 1.0)
                assert err.max() <= tol, (L, e, kernel, err)
