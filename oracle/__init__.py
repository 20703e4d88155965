"""ctypes bindings of the CPU FP64 oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg may import this
package, and only to check or time the GPU library against it: the product (``scatrs_amd``)
never imports it. Build it with ``make -C oracle``.
"""

from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB = None


class _Cx(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_cp = np.ctypeslib.ndpointer(dtype=np.complex128, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def lib() -> C.CDLL:
    """The oracle library, built on first use if absent."""
    global _LIB
    if _LIB is None:
        path = _HERE / "liboracle.so"
        if not path.exists():
            os.system(f"make -s -C {_HERE}")
        L = C.CDLL(str(path))
        L.orc_default_n_max.restype = C.c_int64
        L.orc_default_n_max.argtypes = [C.c_double]
        L.orc_start_index.restype = C.c_int64
        L.orc_start_index.argtypes = [C.c_int64, C.c_double]
        L.orc_mie_solve.argtypes = [C.c_double, _Cx, C.c_int64, C.c_int,
                                    C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_efficiencies.argtypes = [C.c_double, C.c_int64, C.c_void_p, C.c_void_p,
                                       C.c_double, C.c_void_p]
        L.orc_efficiencies_jacobian.argtypes = [C.c_double, C.c_int64] + [C.c_void_p] * 6 + [
            C.c_double, C.c_void_p, C.c_int, C.c_void_p]
        L.orc_mie_efficiencies_batch.argtypes = [_dp, _cp, C.c_int64, C.c_int64, _dp,
                                                 C.c_void_p, C.c_int]
        L.orc_amplitudes.argtypes = [C.c_int64, _cp, _cp, _dp, C.c_int64, _cp, _cp]
        L.orc_gauss_legendre.argtypes = [C.c_int64, _dp, _dp]
        L.orc_wigner_d.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int64, _dp]
        L.orc_greek_project.argtypes = [C.c_int64, _dp, _dp, _dp, _dp, _dp, _dp, C.c_int64, _dp]
        L.orc_layered_solve.argtypes = [C.c_int, _dp, _cp, C.c_int64, C.c_int, _cp, _cp,
                                        C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_density.restype = C.c_double
        L.orc_density.argtypes = [C.c_int, _dp, C.c_double, C.c_void_p]
        L.orc_radius_range.argtypes = [C.c_int, _dp, C.POINTER(C.c_double),
                                       C.POINTER(C.c_double)]
        L.orc_ensemble_fixed.argtypes = [C.c_int, _dp, C.c_double, C.c_int, C.c_void_p, _cp,
                                         C.c_int64, C.c_int64, C.c_int64, C.c_int, _dp,
                                         C.c_void_p, C.c_void_p]
        L.orc_ensemble_fixed_ex.argtypes = [C.c_int, _dp, C.c_double, C.c_int, C.c_void_p, _cp,
                                            C.c_int64, C.c_int64, C.c_int64, C.c_int, C.c_void_p,
                                            C.c_void_p, _dp, C.c_void_p, C.c_void_p]
        L.orc_ensemble_fixed_mu.argtypes = [C.c_int, _dp, C.c_double, C.c_int, C.c_void_p, _cp,
                                            C.c_int64, C.c_int64, C.c_int64, C.c_int, C.c_void_p,
                                            C.c_void_p, C.c_void_p, C.c_int64, _dp, C.c_void_p,
                                            C.c_void_p]
        L.orc_ensemble_fixed_jacobian.argtypes = [C.c_int, _dp, C.c_double, C.c_int, C.c_void_p,
                                                  _cp, C.c_int64, C.c_int64, C.c_int64, C.c_int,
                                                  C.c_int, _dp, C.c_void_p, C.c_void_p, _dp,
                                                  C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_ensemble_batch.argtypes = [C.c_int64, _ip, _dp, _dp, C.c_void_p, _dp, _cp,
                                         C.c_int64, C.c_int64, C.c_int64, C.c_int, _dp,
                                         C.c_void_p, C.c_void_p, C.c_int]
        _LIB = L
    return _LIB


def _ptr(a: np.ndarray | None) -> int | None:
    return None if a is None else a.ctypes.data


def default_n_max(x: float) -> int:
    return int(lib().orc_default_n_max(float(x)))


def start_index(n_top: int, rho_abs: float) -> int:
    return int(lib().orc_start_index(int(n_top), float(rho_abs)))


def mie_solve(x: float, m: complex, n: int | None = None, derivatives: bool = False) -> dict:
    """a, b (n,), abs_sum and, with derivatives, da_dx, da_dm, db_dx, db_dm, abs_jac (3,)."""
    n = default_n_max(x) if n is None else int(n)
    a = np.zeros(n, np.complex128)
    b = np.zeros(n, np.complex128)
    d = [np.zeros(n, np.complex128) for _ in range(4)] if derivatives else [None] * 4
    abs_sum = np.zeros(1)
    abs_jac = np.zeros(3)
    mm = _Cx(m.real, m.imag)
    lib().orc_mie_solve(float(x), mm, n, int(derivatives), _ptr(a), _ptr(b), *map(_ptr, d),
                        _ptr(abs_sum), _ptr(abs_jac))
    out = {"a": a, "b": b, "abs_sum": float(abs_sum[0]), "n": n}
    if derivatives:
        out.update(da_dx=d[0], da_dm=d[1], db_dx=d[2], db_dm=d[3], abs_jac=abs_jac)
    return out


def efficiencies(x: float, s: dict) -> np.ndarray:
    q = np.zeros(5)
    lib().orc_efficiencies(float(x), s["n"], _ptr(s["a"]), _ptr(s["b"]), s["abs_sum"], _ptr(q))
    return q


def efficiencies_jacobian(x: float, s: dict, outer: bool = True) -> np.ndarray:
    j = np.zeros(15)
    lib().orc_efficiencies_jacobian(float(x), s["n"], _ptr(s["a"]), _ptr(s["b"]),
                                    _ptr(s["da_dx"]), _ptr(s["da_dm"]), _ptr(s["db_dx"]),
                                    _ptr(s["db_dm"]), s["abs_sum"], _ptr(s["abs_jac"]),
                                    int(outer), _ptr(j))
    return j.reshape(3, 5)


def mie_efficiencies_batch(x: np.ndarray, m: np.ndarray, derivatives: bool = False,
                           n_max: int = 0, threads: int = 0):
    x = np.ascontiguousarray(x, np.float64)
    m = np.ascontiguousarray(m, np.complex128)
    n = x.size
    values = np.zeros((5, n))
    jac = np.zeros((3, 5, n)) if derivatives else None
    rc = lib().orc_mie_efficiencies_batch(x, m, n, int(n_max), values, _ptr(jac), int(threads))
    if rc != 0:
        raise ValueError(f"invalid input at element {-rc - 1}")
    return values, jac


def amplitudes(s: dict, mu: np.ndarray):
    mu = np.ascontiguousarray(mu, np.float64)
    s1 = np.zeros(mu.size, np.complex128)
    s2 = np.zeros(mu.size, np.complex128)
    lib().orc_amplitudes(s["n"], s["a"], s["b"], mu, mu.size, s1, s2)
    return s1, s2


def gauss_legendre(n: int):
    nodes = np.zeros(n)
    weights = np.zeros(n)
    lib().orc_gauss_legendre(int(n), nodes, weights)
    return nodes, weights


def wigner_d(m: int, n: int, x: float, l_max: int) -> np.ndarray:
    out = np.zeros(l_max + 1)
    if lib().orc_wigner_d(int(m), int(n), float(x), int(l_max), out) != 0:
        raise ValueError("pair not supported by the oracle")
    return out


def greek_project(nodes, weights, f11, f12, f33, f34, order: int) -> np.ndarray:
    out = np.zeros(6 * (order + 1))
    arrs = [np.ascontiguousarray(v, np.float64) for v in (nodes, weights, f11, f12, f33, f34)]
    lib().orc_greek_project(arrs[0].size, *arrs, int(order), out)
    return out.reshape(6, order + 1)


def layered_solve(x, m, n: int | None = None, derivatives: bool = False) -> dict:
    x = np.ascontiguousarray(x, np.float64)
    m = np.ascontiguousarray(m, np.complex128)
    L = x.size
    n = default_n_max(x[-1]) if n is None else int(n)
    a = np.zeros(n, np.complex128)
    b = np.zeros(n, np.complex128)
    dgrp = np.zeros(L * 4 * n, np.complex128) if derivatives else None
    abs_sum = np.zeros(1)
    abs_jac = np.zeros(3 * L) if derivatives else None
    rc = lib().orc_layered_solve(L, x, m, n, int(derivatives), a, b, _ptr(dgrp),
                                 _ptr(abs_sum), _ptr(abs_jac))
    if rc != 0:
        raise ValueError("invalid layered input")
    out = {"a": a, "b": b, "abs_sum": float(abs_sum[0]), "n": n}
    if derivatives:
        out["d"] = dgrp.reshape(L, 4, n)
        out["abs_jac"] = abs_jac.reshape(L, 3)
    return out


def _params(p0, p1, p2) -> np.ndarray:
    return np.array([p0, p1, p2], np.float64)


def density(kind: int, p0: float, p1: float, r: float, p2: float = 0.0):
    g = np.zeros(3)
    u = lib().orc_density(kind, _params(p0, p1, p2), r, g.ctypes.data)
    return u, g


def radius_range(kind: int, p0: float, p1: float, p2: float = 0.0):
    lo, hi = C.c_double(), C.c_double()
    if lib().orc_radius_range(kind, _params(p0, p1, p2), C.byref(lo), C.byref(hi)) != 0:
        raise ValueError("invalid distribution")
    return lo.value, hi.value


LOGNORMAL, GAMMA, MODIFIED_GAMMA, POWER_LAW, MONODISPERSE, SEGMENT = 0, 1, 2, 3, 4, 5


def ensemble_fixed(kind: int, p0: float, p1: float, wavelength: float, m, fractions=None,
                   n_radius: int = 128, n_angles: int = 128, order: int = 32,
                   normalized: bool = True, t_range=None, n_nodes=None, p2: float = 0.0,
                   mu=None) -> dict:
    """One ensemble on the fixed discrete path: scalars (11,), greek (6, L+1), phase (6, n_q),
    or (6, len(mu)) at the cosines mu when given. p2 is the third parameter of a modified gamma
    or power law. t_range = (t_lo, t_hi) replaces the distribution's ln r range and n_nodes the
    per-node N (for finite differences at a fixed range or fixed truncations)."""
    m = np.atleast_1d(np.asarray(m, np.complex128)).copy()
    layers = m.size
    fr = None if fractions is None else np.ascontiguousarray(fractions, np.float64)
    scalars = np.zeros(11)
    greek = np.zeros(6 * (order + 1)) if order >= 0 else None
    mu = None if mu is None else np.ascontiguousarray(mu, np.float64).reshape(-1)
    nq = n_angles if mu is None else mu.size
    phase = np.zeros(6 * nq)
    rng = None if t_range is None else np.ascontiguousarray(t_range, np.float64)
    nn = None if n_nodes is None else np.ascontiguousarray(n_nodes, np.int64)
    rc = lib().orc_ensemble_fixed_mu(kind, _params(p0, p1, p2), wavelength, layers, _ptr(fr), m,
                                     n_radius, n_angles, order, int(normalized), _ptr(rng),
                                     _ptr(nn), _ptr(mu), 0 if mu is None else mu.size, scalars,
                                     _ptr(greek), _ptr(phase))
    if rc != 0:
        raise ValueError(f"invalid ensemble input ({rc})")
    return {"scalars": scalars,
            "greek": None if greek is None else greek.reshape(6, order + 1),
            "phase": phase.reshape(6, nq)}


def ensemble_fixed_jacobian(kind: int, p0: float, p1: float, wavelength: float, m,
                            fractions=None, n_radius: int = 128, n_angles: int = 128,
                            order: int = 32, normalized: bool = True,
                            fixed_range: bool = False, p2: float = 0.0) -> dict:
    """Values and forward-mode derivatives of one ensemble along [p0, p1, p2, Re m_0, Im m_0,
    ..., fraction_0, ...]: d_scalars (P, 11), d_greek (P, 6, L+1), d_phase (P, 6, n_q),
    n_nodes."""
    m = np.atleast_1d(np.asarray(m, np.complex128)).copy()
    layers = m.size
    P = 3 + 2 * layers + (layers - 1)
    fr = None if fractions is None else np.ascontiguousarray(fractions, np.float64)
    scalars = np.zeros(11)
    greek = np.zeros(6 * (order + 1))
    phase = np.zeros(6 * n_angles)
    ds = np.zeros(P * 11)
    dg = np.zeros(P * 6 * (order + 1))
    dp = np.zeros(P * 6 * n_angles)
    nn = np.zeros(n_radius, np.int64)
    rc = lib().orc_ensemble_fixed_jacobian(kind, _params(p0, p1, p2), wavelength, layers,
                                           _ptr(fr), m,
                                           n_radius, n_angles, order, int(normalized),
                                           int(fixed_range), scalars, _ptr(greek), _ptr(phase),
                                           ds, _ptr(dg), _ptr(dp), _ptr(nn))
    if rc != 0:
        raise ValueError(f"invalid ensemble input ({rc})")
    return {"scalars": scalars, "greek": greek.reshape(6, order + 1),
            "phase": phase.reshape(6, n_angles), "d_scalars": ds.reshape(P, 11),
            "d_greek": dg.reshape(P, 6, order + 1), "d_phase": dp.reshape(P, 6, n_angles),
            "n_nodes": nn}


def ensemble_batch(kind, p0, p1, wavelength, m, n_radius: int, n_angles: int, order: int,
                   normalized: bool = True, threads: int = 0, want_tables: bool = True,
                   p2=None):
    kind = np.ascontiguousarray(kind, np.int32)
    cnt = kind.size
    p0 = np.ascontiguousarray(p0, np.float64)
    p1 = np.ascontiguousarray(p1, np.float64)
    wl = np.ascontiguousarray(wavelength, np.float64)
    m = np.ascontiguousarray(m, np.complex128)
    scalars = np.zeros((cnt, 11))
    greek = np.zeros((cnt, 6, order + 1)) if want_tables else None
    phase = np.zeros((cnt, 6, n_angles)) if want_tables else None
    p2 = None if p2 is None else np.ascontiguousarray(p2, np.float64)
    rc = lib().orc_ensemble_batch(cnt, kind, p0, p1, _ptr(p2), wl, m, n_radius, n_angles, order,
                                  int(normalized), scalars, _ptr(greek), _ptr(phase),
                                  int(threads))
    if rc != 0:
        raise ValueError("invalid ensemble input")
    return scalars, greek, phase


def ensemble_mixture(parts, wavelength: float, n_radius: int = 128, n_angles: int = 128,
                     order: int = 32, normalized: bool = True) -> dict:
    """A mixture of parts on the fixed discrete path, as the reference's engine mixes them
    (scatrs/src/ensemble/engine.rs:2120-2150, parts from engine.rs:143-160): each part
    (f, kind, p0, p1, m, fractions) integrated on its own, its number averages (<C_ext>,
    <C_sca>, <C_abs>, <C_back>, <C_sca g>, <r^2>, <r^3> and the averaged tables, the outputs of
    ensemble_fixed with normalized=False) weighted by its number fraction f (the component's
    weight over the total times the mode's weight over its component's total), then the
    outputs assembled from the mixed averages as for one part (engine.rs:2396-2497): ratios
    ssa = <C_sca>/<C_ext>, g = <C_sca g>/<C_sca>, lidar ratio 4 pi <C_ext>/<C_back>, effective
    radius <r^3>/<r^2>; the phase matrix divided by <C_sca> and the Greek coefficients by their
    alpha1_0 when normalized (engine.rs:2447-2490)."""
    acc = np.zeros(7)
    greek = np.zeros((6, order + 1)) if order >= 0 else None
    phase = np.zeros((6, n_angles))
    for (f, kind, p0, p1, m, fr, *rest) in parts:
        r = ensemble_fixed(kind, p0, p1, wavelength, m, fr, n_radius=n_radius,
                           n_angles=n_angles, order=order, normalized=False,
                           p2=rest[0] if rest else 0.0)
        s = r["scalars"]
        avg = np.array([s[0], s[1], s[2], s[3], s[5] * s[1], s[9] / np.pi,
                        s[10] * 3.0 / (4.0 * np.pi)])
        acc += f * avg
        phase += f * r["phase"]
        if greek is not None:
            greek += f * r["greek"]
    ext, sca, absb, back, scag, r2, r3 = acc
    scalars = np.array([ext, sca, absb, back, sca / ext, scag / sca, 4.0 * np.pi * ext / back,
                        0.0, r3 / r2, np.pi * r2, 4.0 * np.pi / 3.0 * r3])
    if normalized:
        phase = phase / sca
        if greek is not None:
            greek = greek / greek[0, 0]
            greek[0, 0] = 1.0
    return {"scalars": scalars, "greek": greek, "phase": phase}
