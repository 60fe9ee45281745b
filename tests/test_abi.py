"""The C-ABI library loads and exports every function include/*.h declares
(no GPU needed: symbols are resolved, nothing is called)."""

import ctypes
import os
import re

import pytest

import hyperalignment_amd as ha

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DECL = re.compile(r"^\s*(?:[A-Za-z_][\w ]*?[\s*]+)(ha_\w+)\s*\(", re.M)


def declared():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in sorted(os.listdir(inc)):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(inc, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)  # drop comments
        for m in DECL.finditer(text):
            names.add(m.group(1))
    return names


def test_headers_declare_the_reference_entry_points():
    names = declared()
    for n in ["ha_searchlight_procrustes", "ha_searchlight_procrustes_dense", "ha_searchlight_weights",
              "ha_sparse_init", "ha_sparse_scatter_add", "ha_sparse_free", "ha_procrustes", "ha_procrustes_f32",
              "ha_hip_init", "ha_hip_cleanup", "ha_hip_available", "ha_searchlight_procrustes_hip",
              "ha_template_procrustes", "ha_set_num_threads", "ha_get_num_threads"]:
        assert n in names, n


def test_library_exports_every_declared_function():
    lib = ctypes.CDLL(ha.LIB_PATH)
    missing = [n for n in sorted(declared()) if not hasattr(lib, n)]
    assert not missing, missing


def test_options_struct_matches_the_library():
    """The ctypes mirror of HaHipOptions (include/ha_hip.h) has the library's
    layout: the defaults of its first and last fields come back as the C
    initialiser sets them, and an invalid value of the last field is rejected
    by ha_hip_set_options (no device is needed for either call)."""
    o = ha.get_options()
    assert o.method == 0 and o.fp64_fallback == 1 and o.max_polar_steps == 28
    assert o.plain_ell == 1e-7 and o.orth_tol == 1e-3 and o.template_ell == 3e-4
    with pytest.raises(ha.HaError):
        ha.set_options(template_ell=0.0)
    assert ha.get_options().template_ell == 3e-4
