"""CPU checks of the C ABI: libwmd.so loads and exports every function that
include/wmd.h declares; host-only entry points behave; the product path has
no CPU fallback.  No kernel runs here."""
import ctypes as C
import os
import re
import subprocess

import pytest

from conftest import REPO, pkg

HEADER = os.path.join(REPO, "include", "wmd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(wmd_[a-z0-9_]+)\s*\(", text)))


def test_header_lists_the_exports():
    nat = pkg()._native
    assert sorted(nat.EXPORTS) == declared_functions()


def test_library_exports_every_declared_symbol():
    nat = pkg()._native
    assert os.path.exists(nat.LIB_PATH), "libwmd.so not built"
    out = subprocess.run(["nm", "-D", "--defined-only", nat.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (wmd_\w+)", out))
    missing = set(declared_functions()) - exported
    assert not missing, missing
    L = nat.load()
    for name in declared_functions():
        assert hasattr(L, name)


def test_library_is_gfx950_code_object():
    nat = pkg()._native
    with open(nat.LIB_PATH, "rb") as f:
        blob = f.read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob  # offload bundle target id


def test_host_only_entry_points():
    nat = pkg()._native
    L = nat.load()
    o = nat.default_options()
    assert o.max_iter == 15 and o.variant == nat.VARIANT_RESIDENT and o.tol == 0.0
    assert o.shape == -1
    shapes = pkg().shapes()
    assert len(shapes) == L.wmd_num_shapes() >= 1
    caps = [s[1] for s in shapes]
    assert caps == sorted(caps) and caps[-1] >= 300  # every batched query length fits
    assert L.wmd_shape_info(99, None, None, None) == nat.WMD_ERR_ARG
    assert b"out of range" in L.wmd_last_error()
    # query capacities step by one slot per lane group: at most B - 1 padded words
    for (w, cap, _), (_, prev, _) in zip(shapes[1:], shapes[:-1]):
        assert cap - prev <= 64 * w // 8 or cap - prev <= 16
    n_cls, max_q = pkg().api.tuning_dims()
    assert n_cls == 20 and max_q == caps[-1]
    assert L.wmd_context_autotune(None, -1.0, None, 8, None) == nat.WMD_ERR_ARG
    assert L.wmd_context_get_tuning(None, 0, None, 0) == nat.WMD_ERR_ARG
    assert L.wmd_context_set_tuning(None, 0, None, 0) == nat.WMD_ERR_ARG


def test_null_arguments_rejected_before_device_use():
    nat = pkg()._native
    L = nat.load()
    st = L.wmd_sinkhorn(None, None, None, -1.0, 15, 1, 1, 300, None, None, None)
    assert st == nat.WMD_ERR_ARG
    assert L.wmd_context_create(0, None) == nat.WMD_ERR_ARG
    assert L.wmd_context_run(None, -1.0, None, None) == nat.WMD_ERR_ARG


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    nat = pkg()._native
    monkeypatch.setattr(nat, "_lib", None)
    monkeypatch.setattr(nat, "LIB_PATH", str(tmp_path / "absent.so"))
    with pytest.raises(ImportError):
        nat.load()
