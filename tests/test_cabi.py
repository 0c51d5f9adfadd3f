"""The C-ABI library loads and exports every symbol include/effm_hpsi.h
declares (no GPU calls)."""
import ctypes
import os
import re

from effmass_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_header_and_exports_agree():
    hdr = open(os.path.join(ROOT, "include", "effm_hpsi.h")).read()
    declared = set(re.findall(r"^\s*(?:int|const char\*)\s+(effm_\w+)\s*\(", hdr, re.M))
    assert declared == set(_lib.EXPORTS)


def test_library_loads_and_exports():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in _lib.EXPORTS:
        assert hasattr(lib, name), name
    _lib.load()


def test_status_codes_match_header():
    hdr = open(os.path.join(ROOT, "include", "effm_hpsi.h")).read()
    for name in ("EFFM_ERR_RUNTIME", "EFFM_ERR_OUT_OF_RANGE", "EFFM_ERR_DOMAIN", "EFFM_ERR_INVALID",
                 "EFFM_ERR_NOMEM", "EFFM_ERR_HIP", "EFFM_PREC_FP64", "EFFM_PREC_MIXED",
                 "EFFM_F_OVERWRITE", "EFFM_F_HOST"):
        m = re.search(rf"#define {name}\s+(-?\w+)", hdr)
        assert m and int(m.group(1), 0) == getattr(_lib, name), name
