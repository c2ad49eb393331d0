"""The C-ABI library loads and exports every entry point include/mdas.h
declares; argument validation happens before any device call."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    with open(os.path.join(ROOT, "include", "mdas.h")) as f:
        src = f.read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mdas_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    from pmdas_amd import _lib
    lib = _lib.load()
    names = _declared()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)


@pytest.mark.parametrize("field,value", [("dim", 2), ("dim", 7), ("max_samples", 4),
                                         ("candidates_num", 0), ("alpha", -1.0),
                                         ("scale_x", 0.0), ("bits_per_dim", -1),
                                         ("extra_img_bits", 20), ("max_samples", 2**31)])
def test_create_rejects_bad_config(field, value):
    from pmdas_amd import _lib
    lib = _lib.load()
    cfg = _lib.MdasConfig(3, 1 << 20, 4, 1, 2, 0.25, 64.0, 64.0, None)
    setattr(cfg, field, value)
    h = ctypes.c_void_p()
    rc = lib.mdas_create(ctypes.byref(cfg), ctypes.byref(h))
    assert rc == _lib.MDAS_ERR_INVALID
    assert h.value is None
    assert lib.mdas_last_error()


def test_null_arguments_are_errors():
    from pmdas_amd import _lib
    lib = _lib.load()
    assert lib.mdas_initial_sampling(None, None) == _lib.MDAS_ERR_INVALID
    assert lib.mdas_integrate(None, None, 4, 4, None) == _lib.MDAS_ERR_INVALID
    assert lib.mdas_destroy(None) == 0


def test_every_handle_entry_point_rejects_null():
    """A NULL handle (all other arguments zero / NULL) is an error status, a
    zero getter value or a no-op destroy -- never a device call or a crash."""
    from pmdas_amd import _lib
    lib = _lib.load()
    H = _lib.SIGNATURES["mdas_destroy"][1][0]
    seen = 0
    for name, (res, args) in _lib.SIGNATURES.items():
        if not args or args[0] is not H:
            continue
        vals = [0 if a in (ctypes.c_int, ctypes.c_longlong, ctypes.c_uint32) else
                0.0 if a is ctypes.c_float else None for a in args]
        r = getattr(lib, name)(*vals)
        seen += 1
        if name.endswith("_destroy"):
            assert r == 0, name
        elif name.startswith("mdas_get_") and res is not ctypes.c_int or name in (
                "mdas_get_dim", "mdas_get_candidates_num", "mdas_get_max_node_size",
                "mdas_get_new_samples", "mdas_get_new_nodes", "mdas_get_number_of_samples",
                "mdas_get_number_of_nodes", "mdas_get_max_samples", "mdas_get_passes",
                "mdas_scene_params"):
            assert r in (0, 0.0, None), name
            assert lib.mdas_last_error(), name
        else:
            assert r == _lib.MDAS_ERR_INVALID, (name, r)
    assert seen > 50


def test_product_package_does_not_reference_the_oracle():
    pkg = os.path.join(ROOT, "pmdas_amd")
    for dirpath, _, files in os.walk(pkg):
        for fn in files:
            if fn.endswith((".py", ".hip", ".cpp", ".hpp", ".h", "Makefile")):
                with open(os.path.join(dirpath, fn), errors="replace") as f:
                    text = f.read()
                assert "oracle" not in text.lower(), fn
