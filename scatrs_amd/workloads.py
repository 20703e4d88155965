"""Seeded synthetic workloads of the ensemble path, and their algorithmic cost accounting.

Every generator takes a seed and returns host numpy arrays; the same seed gives the same
inputs on any machine (numpy's PCG64).

Workloads:
  * ``spheres``: 2^20 independent spheres, x log-uniform on [0.1, 1000], Re m uniform on
    [1.30, 1.70], Im m log-uniform on [1e-8, 1e-1];
  * ``sulfate``: a lognormal r_g = 0.08 um, sigma_g = 1.6 of m = 1.43 + 1e-8 i at
    wavelengths 0.35, 0.55, 1.0 um, 128 radius nodes, Greek to 32;
  * ``lognormals``: 4096 lognormals (r_g log-uniform on [0.01, 5] um, sigma_g uniform on
    [1.2, 2.5], m as for ``spheres``) at 64 wavelengths log-spaced over 0.3-2.5 um, 512 radius
    nodes, 1024 angles, Greek to 256;
  * ``adjoint``: the first 1024 distributions of ``lognormals`` at every other of its
    wavelengths (32), with a seeded loss sum w [(ssa - t)^2 + (g - t)^2 +
    sum_{l<64} (alpha1_l - t_l)^2] for the reverse-mode path;
  * ``layered``: gamma distributions (r_eff log-uniform on [0.05, 3] um, v_eff uniform on
    [0.05, 0.4]) of a core and 1-4 shells (fractions from Dirichlet(2) cumulative sums,
    Re m uniform on [1.3, 2.0], Im m log-uniform on [1e-6, 1]) at 32 wavelengths;
  * ``distributions``: modified gamma, power-law and monodisperse spheres at 32 wavelengths;
  * ``bimodal``: bimodal lognormals of spheres (a fine and a coarse mode) at 32 wavelengths.
"""

from __future__ import annotations

import numpy as np

SEED = 20261203


def spheres(n: int = 1 << 20, seed: int = SEED):
    """x (n,), m (n,) complex."""
    rng = np.random.default_rng(seed)
    x = np.exp(rng.uniform(np.log(0.1), np.log(1000.0), n))
    m = rng.uniform(1.30, 1.70, n) + 1j * np.exp(rng.uniform(np.log(1e-8), np.log(1e-1), n))
    return x, m


def sulfate():
    """One distribution at three wavelengths: (kind, p0, p1, wavelength, m) arrays of 3."""
    wl = np.array([0.35, 0.55, 1.0])
    kind = np.zeros(3, np.int32)
    p0 = np.full(3, 0.08)
    p1 = np.full(3, 1.6)
    m = np.full(3, 1.43 + 1e-8j)
    return kind, p0, p1, wl, m


def lognormals(n_dist: int = 4096, n_wl: int = 64, seed: int = SEED):
    """n_dist x n_wl ensembles, distribution-major: (kind, p0, p1, wavelength, m)."""
    rng = np.random.default_rng(seed + 1)
    rg = np.exp(rng.uniform(np.log(0.01), np.log(5.0), n_dist))
    sg = rng.uniform(1.2, 2.5, n_dist)
    m = rng.uniform(1.30, 1.70, n_dist) + 1j * np.exp(rng.uniform(np.log(1e-8), np.log(1e-1),
                                                                  n_dist))
    wl = np.exp(np.linspace(np.log(0.3), np.log(2.5), n_wl))
    kind = np.zeros(n_dist * n_wl, np.int32)
    return (kind, np.repeat(rg, n_wl), np.repeat(sg, n_wl), np.tile(wl, n_dist),
            np.repeat(m, n_wl))


def adjoint(n_dist: int = 1024, n_wl: int = 32, seed: int = SEED):
    """The first n_dist distributions of ``lognormals`` at n_wl of its 64 wavelengths (every
    64 // n_wl-th), distribution-major: (kind, p0, p1, wavelength, m) and the loss's seeded
    targets and weights {"ssa", "g", "alpha1" (count, 64), "w"}. The alpha1 targets are a
    Henyey-Greenstein phase function's, (2l + 1) t_g^l."""
    kind, p0, p1, wl, m = lognormals(4096, 64, seed)
    step = 64 // n_wl
    sel = (np.arange(n_dist)[:, None] * 64 + np.arange(0, 64, step)[None, :n_wl]).reshape(-1)
    rng = np.random.default_rng(seed + 3)
    cnt = sel.size
    t_g = rng.uniform(0.3, 0.9, cnt)
    targets = {"ssa": rng.uniform(0.5, 1.0, cnt), "g": t_g,
               "alpha1": (2.0 * np.arange(64) + 1.0)[None, :] * t_g[:, None] ** np.arange(64),
               "w": rng.uniform(0.5, 1.5, cnt)}
    return (kind[sel], p0[sel], p1[sel], wl[sel], m[sel]), targets


def adjoint_loss(scalars, greek, targets):
    """The loss of ``adjoint`` and its cotangents (bar_scalars (count, 11), bar_greek
    (count, 6, order + 1)) from the outputs; numpy or torch arrays alike."""
    w = targets["w"]
    d_ssa = scalars[:, 4] - targets["ssa"]
    d_g = scalars[:, 5] - targets["g"]
    d_a = greek[:, 0, :64] - targets["alpha1"]
    loss = (w * (d_ssa ** 2 + d_g ** 2 + (d_a ** 2).sum(1))).sum()
    bar_s = 0.0 * scalars
    bar_g = 0.0 * greek
    bar_s[:, 4] = 2.0 * w * d_ssa
    bar_s[:, 5] = 2.0 * w * d_g
    bar_g[:, 0, :64] = 2.0 * w[:, None] * d_a
    return loss, bar_s, bar_g


def layered(n_dist: int = 2048, n_wl: int = 32, layers: int | None = None, seed: int = SEED):
    """Gamma distributions of layered spheres. With ``layers`` None each distribution draws
    1-4 shells; the library takes one layer count per call, so the result is a dict keyed by
    the layer count L of (kind, p0, p1, wavelength, fractions (count, L-1), m (count, L))."""
    rng = np.random.default_rng(seed + 2)
    reff = np.exp(rng.uniform(np.log(0.05), np.log(3.0), n_dist))
    veff = rng.uniform(0.05, 0.4, n_dist)
    shells = rng.integers(1, 5, n_dist) if layers is None else np.full(n_dist, layers - 1)
    wl = np.exp(np.linspace(np.log(0.3), np.log(2.5), n_wl))
    out = {}
    for s in np.unique(shells):
        idx = np.nonzero(shells == s)[0]
        L = int(s) + 1
        # Interface radii: cumulative sums of Dirichlet(2, ..., 2) volume-like shares.
        shares = rng.dirichlet(np.full(L, 2.0), idx.size)
        fr = np.cumsum(shares, axis=1)[:, :-1]
        mre = rng.uniform(1.3, 2.0, (idx.size, L))
        mim = np.exp(rng.uniform(np.log(1e-6), np.log(1.0), (idx.size, L)))
        cnt = idx.size * n_wl
        out[L] = (np.ones(cnt, np.int32), np.repeat(reff[idx], n_wl),
                  np.repeat(veff[idx], n_wl), np.tile(wl, idx.size),
                  np.repeat(fr, n_wl, axis=0), np.repeat(mre + 1j * mim, n_wl, axis=0))
    return out


def distributions(n_dist: int = 1536, n_wl: int = 32, seed: int = SEED):
    """Spheres over the other distribution kinds, a third each: modified gamma (alpha uniform
    on [1, 8], gamma on [0.5, 1.5], b set by a mode radius log-uniform on [0.05, 2] um), power
    law (exponent uniform on [2, 4.5], r_min log-uniform on [0.01, 0.1] um, r_max on
    [1, 5] um), monodisperse (radius log-uniform on [0.05, 5] um); m as for ``spheres``; 32
    wavelengths as for ``layered``. (kind, p0, p1, p2, wavelength, m) arrays of
    n_dist * n_wl."""
    rng = np.random.default_rng(seed + 3)
    kind = np.repeat(np.array([2, 3, 4], np.int32), n_dist // 3)
    n = kind.size
    p0, p1, p2 = np.zeros(n), np.zeros(n), np.zeros(n)
    mg, pl, mono = kind == 2, kind == 3, kind == 4
    alpha = rng.uniform(1.0, 8.0, n)
    gam = rng.uniform(0.5, 1.5, n)
    rmode = np.exp(rng.uniform(np.log(0.05), np.log(2.0), n))
    p0[mg], p2[mg] = alpha[mg], gam[mg]
    p1[mg] = alpha[mg] / (gam[mg] * rmode[mg] ** gam[mg])  # the mode (alpha / (b gamma))^(1/gamma)
    p0[pl] = rng.uniform(2.0, 4.5, pl.sum())
    p1[pl] = np.exp(rng.uniform(np.log(0.01), np.log(0.1), pl.sum()))
    p2[pl] = np.exp(rng.uniform(np.log(1.0), np.log(5.0), pl.sum()))
    p0[mono] = np.exp(rng.uniform(np.log(0.05), np.log(5.0), mono.sum()))
    m = rng.uniform(1.30, 1.70, n) + 1j * np.exp(rng.uniform(np.log(1e-8), np.log(1e-1), n))
    wl = np.exp(np.linspace(np.log(0.3), np.log(2.5), n_wl))
    r = lambda a: np.repeat(a, n_wl)  # noqa: E731
    return r(kind), r(p0), r(p1), r(p2), np.tile(wl, n), r(m)


def bimodal(n_dist: int = 1024, n_wl: int = 32, seed: int = SEED):
    """Bimodal lognormals of spheres: a fine mode (r_g log-uniform on [0.02, 0.2] um,
    sigma_g uniform on [1.3, 1.8]) and a coarse one (r_g on [0.5, 3] um, sigma_g on
    [1.5, 2.2]) with the coarse number fraction log-uniform on [1e-4, 1e-1]; m as for
    ``spheres``; 32 wavelengths. Returns the parts (kind, p0, p1, wavelength, m) of
    2 * n_dist * n_wl rows, parts first, and the number fractions (2, n_dist * n_wl)."""
    rng = np.random.default_rng(seed + 4)
    rf = np.exp(rng.uniform(np.log(0.02), np.log(0.2), n_dist))
    sf = rng.uniform(1.3, 1.8, n_dist)
    rc = np.exp(rng.uniform(np.log(0.5), np.log(3.0), n_dist))
    sc = rng.uniform(1.5, 2.2, n_dist)
    fc = np.exp(rng.uniform(np.log(1e-4), np.log(1e-1), n_dist))
    m = rng.uniform(1.30, 1.70, n_dist) + 1j * np.exp(rng.uniform(np.log(1e-8), np.log(1e-1),
                                                                   n_dist))
    wl = np.exp(np.linspace(np.log(0.3), np.log(2.5), n_wl))
    r = lambda a: np.repeat(a, n_wl)  # noqa: E731
    cnt = n_dist * n_wl
    kind = np.zeros(2 * cnt, np.int32)
    p0 = np.concatenate([r(rf), r(rc)])
    p1 = np.concatenate([r(sf), r(sc)])
    wls = np.tile(np.tile(wl, n_dist), 2)
    ms = np.tile(r(m), 2)
    frac = np.stack([1.0 - r(fc), r(fc)])
    return (kind, p0, p1, wls, ms), frac


# ---- algorithmic cost of the Mie recurrences (FP64 operations; a division counts 1) ----
# Tallied from the recurrences as written (scatrs/src/mie/kernel.rs:116-136,378-556 and the
# efficiency sums of efficiencies.rs:64-84): one complex downward step (c = (2k+1)/z, c - u,
# the one-division complex inverse) is 10; one real downward step 3; one order of the upward
# region (psi/chi upward, two complex inverses, a_n, b_n, the absorption term and the
# Q_sca/Q_back/g sums) 118; one order of the ratio region (R_n, T_n, 1/xi_n, the
# cancellation-free numerators and two complex inverses in place of the upward step) 160.
FLOPS_DOWN_COMPLEX = 10
FLOPS_DOWN_REAL = 3
FLOPS_ORDER_UPWARD = 118
FLOPS_ORDER_RATIO = 160


def default_n_max(x):
    return np.ceil(x + 6.0 * np.cbrt(x) + 3.0).astype(np.int64)


def start_index(n_top, rho):
    """The downward recurrences' start (scatrs/src/mie/kernel.rs:99-112), vectorized."""
    n_top = np.asarray(n_top, np.int64)
    rho = np.asarray(rho, np.float64)
    k0 = np.maximum(n_top, np.ceil(1.5 * rho).astype(np.int64))
    q = (2 * k0 + 1) / rho - 0.5
    enclosed = k0 + np.ceil(52.0 / (2.0 * np.log2(q))).astype(np.int64) + 2
    s0 = n_top - rho
    s0_32 = np.where(s0 > 0, np.abs(s0) * np.sqrt(np.abs(s0)), 0.0)
    s = (s0_32 + 1.4 * 0.37 * 52.0 * np.sqrt(rho)) ** (2.0 / 3.0)
    asym = np.ceil(rho + s).astype(np.int64) + 30
    return np.where(n_top >= 1.5 * rho, enclosed, np.minimum(asym, enclosed))


def mie_cost(x, m):
    """(terms, flops) of solving spheres (x, m) with the default truncation."""
    x = np.asarray(x, np.float64)
    m = np.asarray(m, np.complex128)
    n = default_n_max(x)
    n_s = np.where(x >= 1.0, np.minimum(np.floor(x).astype(np.int64), n), 0)
    steps_z = start_index(n + 1, np.abs(m * x))
    steps_x = np.where(n_s < n, start_index(n + 1, x) - np.maximum(n_s + 1, 1) + 1, 0)
    down = np.where(m.imag == 0, FLOPS_DOWN_REAL, FLOPS_DOWN_COMPLEX)
    flops = (FLOPS_ORDER_UPWARD * n_s + FLOPS_ORDER_RATIO * (n - n_s) + down * steps_z
             + FLOPS_DOWN_REAL * steps_x)
    return int(n.sum()), float(flops.sum())
