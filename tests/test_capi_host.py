"""Host-side checks of the C ABI: the library loads, exports every declared symbol, and the
host seed path matches the reference vectors.  No GPU compute is called here."""

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from jathenax_amd import _lib

ROOT = Path(__file__).resolve().parents[1]


def _declared_symbols():
    text = (ROOT / "include" / "jathenax_amd.h").read_text()
    return sorted(set(re.findall(r"\b(jx_[a-z_0-9]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    if not _lib.LIB_PATH.exists():
        pytest.fail("libjathenax_amd.so is not built (python -m jathenax_amd.build)")
    return _lib.load()


def test_library_exports_every_declared_symbol(lib):
    declared = _declared_symbols()
    assert len(declared) >= 18
    for name in declared:
        assert hasattr(lib, name), name
    assert set(declared) == set(_lib.SIGNATURES), "ctypes prototypes out of sync with header"


def test_run_params_layout_matches_header():
    # 8 ints, 12 doubles, then the RK-table pointer; no padding surprises.
    assert ctypes.sizeof(_lib.RunParams) == 8 * 4 + 12 * 8 + 8
    assert _lib.RunParams.dx.offset == 32
    assert _lib.RunParams.rk_rows.offset == 8 * 4 + 12 * 8


def test_seed_streams_match_reference(lib, ref_vectors):
    from jathenax_amd import forcing
    u = forcing.ran_st(1618)
    np.testing.assert_array_equal([next(u) for _ in range(256)], ref_vectors["ran_st_1618"])
    u = forcing.ran_st(7)
    np.testing.assert_array_equal([next(u) for _ in range(64)], ref_vectors["ran_st_7"])
    g = forcing.ran_gaussian_st(1618)
    np.testing.assert_array_equal([next(g) for _ in range(800)], ref_vectors["gauss_1618"])


@pytest.mark.parametrize("lo,hi,peak,kb", [(3, 5, 4.0, 5), (1, 3, 2.0, 3)])
def test_modal_table_matches_reference(lib, ref_vectors, lo, hi, peak, kb):
    from jathenax_amd import forcing
    modes = forcing.cartesian_modes(lo, hi, (-kb, kb), (-kb, kb))
    np.testing.assert_array_equal(modes, ref_vectors[f"modes_{lo}_{hi}"])
    coef = forcing.modal_coefficients(modes, (1.0, 1.0), lo, hi, peak, 1.0, 1618, 0.75, 0.0075, 6)
    np.testing.assert_array_equal(coef, ref_vectors[f"coef_{lo}_{hi}"])
    coef = forcing.modal_coefficients(forcing.cartesian_modes(1, 3, (-3, 3), (-3, 3)),
                                      (1.0, 1.0), 1, 3, 2.0, 0.7, 99, 0.0, 0.01, 3)
    np.testing.assert_array_equal(coef, ref_vectors["coef_1_3_sol07_tcorr0"])


def test_seed_path_error_contracts(lib):
    from jathenax_amd import forcing
    with pytest.raises(ValueError):
        forcing.cartesian_modes(3, 3, (-5, 5), (-5, 5))
    with pytest.raises(ValueError):
        forcing.cartesian_modes(30, 40, (-5, 5), (-5, 5))
    modes = forcing.cartesian_modes(1, 3, (-3, 3), (-3, 3))
    with pytest.raises(ValueError):
        forcing.modal_coefficients(modes, (1.0, 1.0), 1, 3, 5.0, 1.0, 1, 0.75, 0.01, 2)
    with pytest.raises(ValueError):
        forcing.modal_coefficients(modes, (1.0, 1.0), 1, 3, 2.0, 1.0, 1, 0.75, 0.0, 2)


def test_option_validation_before_device_work():
    import jathenax_amd as jx
    U = np.ones((6, 8, 3))
    with pytest.raises(ValueError, match="viscosity"):
        jx.evolve_isothermal(U, 0.1, 0.1, 1.0, 1e-10, "rk3", 0.4, 1.0, 10, viscosity=-1.0)
    with pytest.raises(ValueError, match="dedt"):
        jx.evolve_isothermal(U, 0.1, 0.1, 1.0, 1e-10, "rk3", 0.4, 1.0, 10, dedt=float("nan"))
    with pytest.raises(ValueError, match="dt_update"):
        jx.evolve_isothermal(U, 0.1, 0.1, 1.0, 1e-10, "rk3", 0.4, 1.0, 10, forcing=((), ()),
                             dt_update=0.0)
