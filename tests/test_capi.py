"""The C-ABI library loads, exports every entry point include/lbm.h declares,
and rejects invalid configurations before touching the GPU (CPU only)."""
import ctypes
import os
import re

import pytest

import d3q19

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "lbm.h")).read()
    return sorted(set(re.findall(r"^(?:int|const char \*)\s*(lbm_\w+)\s*\(", text, flags=re.M)))


def test_header_and_binding_agree():
    assert declared_functions() == d3q19.exported_symbols()


def test_library_exports_every_declared_symbol():
    lib = d3q19.load_library()
    for name in declared_functions():
        assert hasattr(lib, name), name


def _cfg(**kw):
    p = d3q19.para(32)
    cfg = d3q19.LbmConfig()
    cfg.lx, cfg.ly, cfg.lz, cfg.precision, cfg.device, cfg.reserved = 32, 8, 8, 0, 0, 0
    for name, _ in d3q19.LbmConfig._fields_[6:]:
        setattr(cfg, name, p[name])
    for k, v in kw.items():
        setattr(cfg, k, v)
    return cfg


@pytest.mark.parametrize("bad", [dict(lx=1), dict(ly=0), dict(precision=7), dict(reserved=1),
                                 dict(val2i=1.0 / 2393.0), dict(coef1=-0.6666), dict(lx=2048, ly=2048, lz=1024)])
def test_create_rejects_invalid_config(bad):
    lib = d3q19.load_library()
    h = ctypes.c_void_p()
    rc = lib.lbm_create(ctypes.byref(_cfg(**bad)), ctypes.byref(h))
    assert rc == 1  # LBM_ERR_ARG
    assert h.value is None
    assert lib.lbm_last_error().decode()


def test_null_handle_is_an_error():
    lib = d3q19.load_library()
    assert lib.lbm_collision_mrt(None) == 1
    assert b"null handle" in lib.lbm_last_error()
