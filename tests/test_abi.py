"""The C-ABI library loads and exports every symbol include/scatrs_hip.h declares (no GPU
needed: nothing here launches a kernel)."""

from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

import numpy as np

import oracle

ROOT = Path(__file__).resolve().parent.parent


def _declared() -> list[str]:
    text = (ROOT / "include" / "scatrs_hip.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|const char \*|double)\s*(scatrs_\w+)\(",
                                 text, re.M)))


def test_library_exports_every_declared_symbol():
    from scatrs_amd import _native

    lib = _native.load()
    names = _declared()
    assert len(names) >= 8
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_native.SYMBOLS)


def test_abi_version_and_host_truncation():
    from scatrs_amd import _native

    lib = _native.load()
    assert lib.scatrs_abi_version() == 3
    for x in (1e-6, 0.1, 1.0, 7.3, 100.0, 999.9, 2e4):
        assert lib.scatrs_default_n_max(x) == oracle.default_n_max(x)


def test_context_rejects_a_missing_device():
    from scatrs_amd import _native

    lib = _native.load()
    h = C.c_void_p()
    assert lib.scatrs_create(1 << 20, None, C.byref(h)) != 0
    assert not h.value


def test_rust_binding_declares_every_symbol():
    """rust/scatrs-hip-sys/src/lib.rs declares exactly the symbols of include/scatrs_hip.h,
    and maps every status (3 -> NotConverged) in its check."""
    text = (ROOT / "rust" / "scatrs-hip-sys" / "src" / "lib.rs").read_text()
    rust = set(re.findall(r"pub fn (scatrs_\w+)\(", text))
    assert rust == set(_declared())
    for code, variant in (("SCATRS_INVALID_INPUT", "InvalidInput"),
                          ("SCATRS_OUTSIDE_VALIDATED_DOMAIN", "OutsideValidatedDomain"),
                          ("SCATRS_NOT_CONVERGED", "NotConverged")):
        assert re.search(code + r" => Err\(Error::" + variant + r"\(", text), code


def test_library_gauss_legendre_rule():
    """scatrs_gauss_legendre, the rule the kernels use (host only): nodes as numpy's
    Golub-Welsch rule's to 2^-52 (whose weights are the less accurate ones: 3e-11 relative at
    n = 200), exactly symmetric, and the weights pinned by exactness: x^k integrated to
    2/(k+1) for even k up to 2n - 2; n outside [1, 65536] rejected."""
    import pytest

    from scatrs_amd.errors import InvalidInputError
    from scatrs_amd.quadrature import gauss_legendre

    for n in (1, 2, 7, 64, 200):
        x, w = gauss_legendre(n)
        xr, wr = np.polynomial.legendre.leggauss(n)
        assert np.max(np.abs(x - xr)) <= 4e-16
        assert np.array_equal(x, -x[::-1]) and np.array_equal(w, w[::-1])
        for k in sorted({0, min(2, 2 * n - 2), n // 2 * 2, 2 * n - 2}):
            assert abs((w * x ** k).sum() - 2.0 / (k + 1)) <= 1e-14, (n, k)
    for bad in (0, 65537):
        with pytest.raises(InvalidInputError):
            gauss_legendre(bad)
