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


def test_struct_fields_match_header():
    """The ctypes mirrors of the header's structs list the same fields in
    the same order (effm_shapes is passed by pointer to effm_get_shapes /
    effm_set_shapes)."""
    hdr = open(os.path.join(ROOT, "include", "effm_hpsi.h")).read()
    body = re.search(r"typedef struct effm_shapes \{(.*?)\} effm_shapes;", hdr, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = re.findall(r"int\s+(\w+)\s*;", body)
    assert fields == [f for f, _ in _lib.Shapes._fields_]
    assert ctypes.sizeof(_lib.Shapes) == 4 * len(fields)
