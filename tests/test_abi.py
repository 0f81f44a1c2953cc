"""The C ABI library (no GPU needed): it loads, exports every function
include/lbmc.h declares, and validates descriptors before touching a device.
The product path has no CPU fallback."""
import ctypes as ct
import os
import re

import pytest

import lbmc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lbmc.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lbmc_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_expected_entry_points():
    names = declared_functions()
    assert set(names) == set(lbmc.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ct.CDLL(lbmc.LIB_PATH)
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_exports_are_c_symbols():
    out = os.popen(f"nm -D --defined-only {lbmc.LIB_PATH}").read()
    for name in declared_functions():
        assert re.search(rf"\bT {name}$", out, flags=re.M), name


def _create(**kw):
    d = lbmc.Desc()
    d.lattice = kw.get("lattice", lbmc.D2Q9)
    L = kw.get("length", (8, 8, 1))
    for k in range(3):
        d.length[k] = L[k]
    d.collision_type = kw.get("collision_type", lbmc.BGK)
    d.storage = kw.get("storage", lbmc.FP32)
    d.tau = 1.0
    h = ct.c_void_p()
    st = lbmc.lib().lbmc_create(ct.byref(d), ct.byref(h))
    return st, h, lbmc.lib().lbmc_last_error(None).decode()


@pytest.mark.parametrize("kw,status", [
    (dict(lattice=7), 1),
    (dict(collision_type=4), 1),
    (dict(collision_type=-1), 1),
    (dict(storage=9), 1),
    (dict(length=(0, 8, 1)), 1),
    (dict(lattice=lbmc.D3Q15, length=(8, 8, -2)), 1),
    (dict(lattice=lbmc.D3Q15, length=(2048, 2048, 1024)), 5),
])
def test_descriptor_validation_before_device(kw, status):
    st, h, msg = _create(**kw)
    assert st == status, msg
    assert not h.value
    assert msg


def test_null_arguments():
    L = lbmc.lib()
    assert L.lbmc_create(None, None) == 1
    assert L.lbmc_step(None, 0, 0) == 1
    assert L.lbmc_sync(None) == 1
    assert L.lbmc_destroy(None) == 0


def test_no_cpu_fallback_in_product_sources():
    """Nothing under lbm-c_amd/ imports, links or loads the oracle."""
    pkg = os.path.join(ROOT, "lbm-c_amd")
    for dirpath, _, files in os.walk(pkg):
        if "_build" in dirpath:
            continue
        for fn in files:
            if fn.endswith((".py", ".cpp", ".hip", ".hpp", ".h", "Makefile")):
                text = open(os.path.join(dirpath, fn)).read()
                for token in ("pyoracle", "lbmc_oracle", "orc_", "oracle/"):
                    assert token not in text, (fn, token)


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    import importlib.util

    src = os.path.join(ROOT, "lbm-c_amd", "lbmc.py")
    dst = tmp_path / "lbmc_copy.py"
    dst.write_text(open(src).read())
    spec = importlib.util.spec_from_file_location("lbmc_copy", dst)
    mod = importlib.util.module_from_spec(spec)
    with pytest.raises(ImportError):
        spec.loader.exec_module(mod)
