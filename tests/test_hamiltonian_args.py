"""Argument checks of the Python mirror (effmass_amd/hamiltonian.py) that run
before any library call: a band count or array shape that does not fit
psi / the outputs raises ValueError instead of reaching the device or the
host staging copies."""
import numpy as np
import pytest

from effmass_amd import _lib
from effmass_amd.hamiltonian import _arr_info, _bands

P64 = _lib.EFFM_PREC_FP64


def info(nb, npw, dt=np.complex128):
    return _arr_info(np.zeros((nb, npw), dt), P64 if dt == np.complex128 else _lib.EFFM_PREC_MIXED)


def test_band_count_defaults_and_override():
    assert _bands(None, 100, info(4, 100), [info(4, 100)]) == 4
    assert _bands(2, 100, info(4, 100), [info(3, 100)]) == 2
    assert _bands(0, 100, info(4, 100), [info(4, 100)]) == 0


def test_more_psi_bands_than_output_raises():
    with pytest.raises(ValueError):
        _bands(None, 100, info(4, 100), [info(3, 100)])


@pytest.mark.parametrize("n_bands", [5, -1])
def test_n_bands_outside_arrays_raises(n_bands):
    with pytest.raises(ValueError):
        _bands(n_bands, 100, info(4, 100), [info(4, 100)])


def test_short_rows_raise():
    with pytest.raises(ValueError):
        _bands(None, 101, info(4, 100), [info(4, 101)])
    with pytest.raises(ValueError):
        _bands(None, 101, info(4, 101), [info(4, 100)])
    with pytest.raises(ValueError):                       # one of several outputs
        _bands(None, 50, info(2, 50), [info(2, 50), info(2, 49)])


def test_dtype_and_layout():
    with pytest.raises(TypeError):
        _arr_info(np.zeros((2, 10), np.complex64), P64)
    with pytest.raises(ValueError):
        _arr_info(np.zeros((2, 10), np.complex128)[:, ::2], P64)
    with pytest.raises(ValueError):
        _arr_info(np.zeros((2, 2, 10), np.complex128), P64)
    p, nb, ld, host, cols = _arr_info(np.zeros(7, np.complex128), P64)
    assert (nb, ld, host, cols) == (1, 7, True, 7)
