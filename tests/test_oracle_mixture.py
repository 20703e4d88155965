"""Mixtures of parts (multimodal distributions, external mixtures) on the fixed path: the
oracle's restatement of the reference's mixing (scatrs/src/ensemble/engine.rs:2120-2150) pinned
by invariants and closed forms, and the library's host-side mixing (scatrs_amd.ensemble.mix_parts)
against it."""

from __future__ import annotations

import numpy as np
import pytest

import oracle

SPH = (0, 0.3, 1.5, 1.45 + 0.01j, None)          # lognormal r_g = 0.3, sigma_g = 1.5
SPH2 = (0, 0.9, 1.3, 1.52 + 1e-4j, None)
GAM = (1, 0.5, 0.2, 1.38 + 1e-3j, None)          # gamma a = 0.5, b = 0.2
LAY = (1, 0.4, 0.15, np.array([1.6 + 0.1j, 1.35 + 1e-5j]), np.array([0.6]))
KW = dict(n_radius=64, n_angles=48, order=16)


def _parts(*fp):
    return [(f, k, a, b, m, fr) for f, (k, a, b, m, fr) in fp]


def _close(a, b, tol):
    a, b = np.asarray(a), np.asarray(b)
    assert np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)) <= tol, (a, b)


@pytest.mark.parametrize("normalized", [True, False])
def test_one_part_is_the_part(normalized):
    for part in (SPH, GAM, LAY):
        k, a, b, m, fr = part
        mix = oracle.ensemble_mixture(_parts((1.0, part)), 0.55, normalized=normalized, **KW)
        one = oracle.ensemble_fixed(k, a, b, 0.55, m, fr, normalized=normalized, **KW)
        _close(mix["scalars"], one["scalars"], 1e-14)
        assert np.max(np.abs(mix["phase"] - one["phase"])) <= 1e-14 * np.max(np.abs(one["phase"]))
        assert np.max(np.abs(mix["greek"] - one["greek"])) <= 1e-14 * np.max(np.abs(one["greek"]))


def test_a_part_mixed_with_itself_is_the_part():
    mix = oracle.ensemble_mixture(_parts((0.3, SPH), (0.7, SPH)), 0.8, **KW)
    one = oracle.ensemble_fixed(*SPH[:3], 0.8, SPH[3], **KW)
    _close(mix["scalars"], one["scalars"], 1e-14)
    assert np.max(np.abs(mix["phase"] - one["phase"])) <= 1e-14 * np.max(np.abs(one["phase"]))


def test_bimodal_lognormal_moments_match_closed_forms():
    """<r^k> of sum_j w_j n_j (modes normalized) is sum_j f_j <r^k>_j (the reference's
    MultiModal moment, distribution.rs:254-264), each mode's the lognormal's truncated to
    ln r_g -+ 5 ln sigma_g: r_g^k exp(k^2 s^2/2) [Phi(5 - k s) - Phi(-5 - k s)] / [Phi(5) - Phi(-5)]."""
    from scipy.special import ndtr

    modes = [(0.8, 0.05, 1.4), (0.2, 0.6, 1.8)]
    parts = [(f, 0, rg, sg, 1.5 + 0j, None) for f, rg, sg in modes]
    r = oracle.ensemble_mixture(parts, 0.5, n_radius=256, n_angles=8, order=-1)

    def mom(k):
        z = ndtr(5.0) - ndtr(-5.0)
        return sum(f * rg ** k * np.exp(0.5 * k * k * np.log(sg) ** 2)
                   * (ndtr(5 - k * np.log(sg)) - ndtr(-5 - k * np.log(sg))) / z
                   for f, rg, sg in modes)

    assert abs(r["scalars"][9] / (np.pi * mom(2)) - 1) < 1e-12
    assert abs(r["scalars"][10] / (4 * np.pi / 3 * mom(3)) - 1) < 1e-12
    assert abs(r["scalars"][8] / (mom(3) / mom(2)) - 1) < 1e-12


def test_mixture_weights_phase_by_c_sca_and_greek_by_alpha1_0():
    """From the parts' normalized outputs: <C_ext> and <C_sca> are number-fraction averages,
    g is the C_sca-weighted mean of the parts' g, the phase matrix the C_sca-weighted mean of
    their phase matrices (the Mixture docstring, python/scatrs/ensemble.py:616-620), the Greek
    coefficients the mean weighted by each part's unnormalized alpha1_0."""
    fp = ((0.5, SPH), (0.2, GAM), (0.3, LAY))
    mix = oracle.ensemble_mixture(_parts(*fp), 0.6, **KW)
    norm = [oracle.ensemble_fixed(k, a, b, 0.6, m, fr, **KW) for _, (k, a, b, m, fr) in fp]
    raw = [oracle.ensemble_fixed(k, a, b, 0.6, m, fr, normalized=False, **KW)
           for _, (k, a, b, m, fr) in fp]
    f = np.array([x for x, _ in fp])
    sca = np.array([n["scalars"][1] for n in norm])
    ext = np.array([n["scalars"][0] for n in norm])
    g = np.array([n["scalars"][5] for n in norm])
    _close(mix["scalars"][0], f @ ext, 1e-14)
    _close(mix["scalars"][1], f @ sca, 1e-14)
    _close(mix["scalars"][5], (f * sca) @ g / (f @ sca), 1e-14)
    ph = sum(w * s * n["phase"] for w, s, n in zip(f, sca, norm)) / (f @ sca)
    assert np.max(np.abs(mix["phase"] - ph)) <= 1e-13 * np.max(np.abs(ph))
    a0 = np.array([r["greek"][0, 0] for r in raw])
    gr = sum(w * a * n["greek"] for w, a, n in zip(f, a0, norm)) / (f @ a0)
    assert np.max(np.abs(mix["greek"] - gr)) <= 1e-13 * np.max(np.abs(gr))


@pytest.mark.parametrize("normalized", [True, False])
def test_library_mixing_matches_the_oracle(normalized):
    """scatrs_amd.ensemble.mix_parts (torch, here on the CPU) on the oracle's per-part outputs
    against the oracle's mixture, at two wavelengths at once."""
    torch = pytest.importorskip("torch")
    from scatrs_amd.ensemble import mix_parts

    fp = ((0.25, SPH), (0.35, SPH2), (0.4, LAY))
    wls = (0.45, 1.1)
    raw = [[oracle.ensemble_fixed(k, a, b, w, m, fr, normalized=False, **KW) for w in wls]
           for _, (k, a, b, m, fr) in fp]
    t = lambda key: torch.tensor(np.array([[r[key] for r in row] for row in raw]))  # noqa: E731
    f = torch.tensor([x for x, _ in fp], dtype=torch.float64)
    sc, gr, ph = mix_parts(t("scalars"), t("greek"), t("phase"), f, normalized)
    # fractions given per row (parts, wavelengths) give the same
    sc2, gr2, ph2 = mix_parts(t("scalars"), t("greek"), t("phase"),
                              f[:, None].expand(-1, len(wls)), normalized)
    assert torch.equal(sc, sc2) and torch.equal(gr, gr2) and torch.equal(ph, ph2)
    for j, w in enumerate(wls):
        ref = oracle.ensemble_mixture(_parts(*fp), w, normalized=normalized, **KW)
        _close(sc[j].numpy(), ref["scalars"], 1e-14)
        assert np.max(np.abs(gr[j].numpy() - ref["greek"])) <= 1e-14 * np.max(np.abs(ref["greek"]))
        assert np.max(np.abs(ph[j].numpy() - ref["phase"])) <= 1e-14 * np.max(np.abs(ref["phase"]))


def test_multimodal_and_mixture_validate_like_the_reference():
    """Errors of python/scatrs/ensemble.py:637-657 and distribution.rs:172-180."""
    from scatrs_amd.errors import InvalidInputError
    from scatrs_amd.ensemble import Ensemble, Mixture
    from scatrs_amd.particles import Sphere
    from scatrs_amd.size import Gamma, LogNormal, MultiModal

    with pytest.raises(InvalidInputError, match="needs a mode"):
        MultiModal(())
    with pytest.raises(InvalidInputError, match="weight of mode 1"):
        MultiModal(((1.0, LogNormal(0.1, 1.5)), (0.0, Gamma(0.2, 0.1))))
    with pytest.raises(TypeError, match="mode 0"):
        MultiModal((LogNormal(0.1, 1.5),))
    with pytest.raises(InvalidInputError, match="scalar parameters"):
        MultiModal(((1.0, LogNormal(np.array([0.1, 0.2]), 1.5)),))
    sph = Sphere(1.5 + 0.01j)
    with pytest.raises(TypeError, match="component 0"):
        Mixture((sph,))
    with pytest.raises(InvalidInputError, match="grid"):
        Mixture(((1.0, Ensemble(sph, LogNormal(np.array([0.1, 0.2]), 1.5))),))
    with pytest.raises(InvalidInputError, match="not all be 0"):
        Mixture(((0.0, Ensemble(sph, LogNormal(0.1, 1.5))),))
    mm = MultiModal(((3.0, LogNormal(0.1, 1.5)), (1.0, Gamma(0.2, 0.1))))
    assert np.allclose(mm.fractions(), [0.75, 0.25])
    assert Ensemble(sph, mm).grid_shape == ()


def test_tabulated_law_is_its_segments_mixed():
    """A tabulated law (n linear between its points, distribution.rs:55-61): its segments, each
    a linear density on its interval, mixed by their number integrals, give the moments of the
    piecewise-linear density in closed form (sum_i int r^k n dr / sum_i int n dr)."""
    from scatrs_amd.size import Tabulated

    radius, density = [0.05, 0.2, 0.6, 1.5], [0.0, 3.0, 1.0, 0.2]
    segs, w = Tabulated(radius, density).segments()
    f = w / w.sum()
    parts = [(fi, oracle.SEGMENT, r0, r1, 1.5 + 0j, None, tilt)
             for fi, (r0, r1, tilt) in zip(f, segs)]
    r = oracle.ensemble_mixture(parts, 0.5, n_radius=64, n_angles=8, order=-1)

    def integral(k):
        out = 0.0
        for r0, r1, d0, d1 in zip(radius, radius[1:], density, density[1:]):
            b = (d1 - d0) / (r1 - r0)
            a = d0 - b * r0
            out += a * (r1 ** (k + 1) - r0 ** (k + 1)) / (k + 1) + \
                b * (r1 ** (k + 2) - r0 ** (k + 2)) / (k + 2)
        return out

    mom = lambda k: integral(k) / integral(0)  # noqa: E731
    assert abs(r["scalars"][9] / (np.pi * mom(2)) - 1) < 1e-13
    assert abs(r["scalars"][8] / (mom(3) / mom(2)) - 1) < 1e-13
    # one flat segment is the uniform law on its interval
    u = oracle.ensemble_fixed(oracle.SEGMENT, 0.3, 0.9, 0.5, 1.5 + 0j, n_radius=32, n_angles=8,
                              order=-1, p2=0.5)
    assert abs(u["scalars"][9] / (np.pi * (0.9 ** 3 - 0.3 ** 3) / (3 * 0.6)) - 1) < 1e-14


def test_tabulated_validates_like_the_reference():
    """distribution.rs:148-170."""
    from scatrs_amd.errors import InvalidInputError
    from scatrs_amd.size import Tabulated

    with pytest.raises(InvalidInputError, match="at least 2 points"):
        Tabulated([0.1], [1.0])
    with pytest.raises(InvalidInputError, match="must increase"):
        Tabulated([0.2, 0.1], [1.0, 1.0])
    with pytest.raises(InvalidInputError, match="finite and >= 0"):
        Tabulated([0.1, 0.2], [1.0, -1.0])
    with pytest.raises(InvalidInputError, match="not all be 0"):
        Tabulated([0.1, 0.2], [0.0, 0.0])
