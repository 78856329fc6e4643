"""CPU checks of the C-ABI boundary: the library builds for gfx950, loads,
exports every function include/*.h declares, and rejects bad arguments
without touching a GPU."""
import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("wrs.h", "wrs_runtime.h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(re.findall(r"\b(wrs_\w+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    import pywrs
    L = pywrs.lib()
    declared = declared_functions()
    assert len(declared) >= 40
    missing = [n for n in sorted(declared) if not hasattr(L, n)]
    assert not missing, missing
    assert declared == set(pywrs.EXPORTS), declared ^ set(pywrs.EXPORTS)


def test_exported_symbols_are_c_linkage():
    import pywrs
    out = subprocess.run(["nm", "-D", "--defined-only", pywrs.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    syms = {l.split()[-1] for l in out.splitlines() if l.strip()}
    for n in declared_functions():
        assert n in syms, n


def test_code_object_targets_gfx950():
    import pywrs
    blob = open(pywrs.LIB_PATH, "rb").read()
    assert b"gfx950" in blob


def test_bad_arguments_are_rejected():
    import pywrs
    L = pywrs.lib()
    plan = C.c_void_p()
    # null out-pointer and max_n == 0 are argument errors, before any device query
    assert L.wrs_plan_create(None, 16, None) == pywrs.WRS_EINVAL
    assert L.wrs_plan_create(None, 0, C.byref(plan)) == pywrs.WRS_EINVAL
    assert L.wrs_plan_create(None, 1 << 32, C.byref(plan)) == pywrs.WRS_EINVAL
    assert L.wrs_psa_build(None, None, 16, None, None) == pywrs.WRS_EINVAL
    assert b"plan" in L.wrs_last_error()
    assert L.wrs_alias_sample(None, 16, 16, 0, 0, None, None) == pywrs.WRS_EINVAL
    assert L.wrs_generate(7, 0.0, 0.0, 0, 16, C.c_void_p(16), None) == pywrs.WRS_EINVAL
    assert L.wrs_batch_plan_create(None, None, 0, C.byref(plan)) == pywrs.WRS_EINVAL


def test_version_string():
    import pywrs
    assert "gfx950" in pywrs.version()


def test_missing_library_fails_loudly(monkeypatch):
    import pywrs
    monkeypatch.setattr(pywrs, "_lib", None)
    monkeypatch.setattr(pywrs, "LIB_PATH", "/nonexistent/libwrs_amd.so")
    with pytest.raises(ImportError):
        pywrs.lib()


def test_bad_tuning_is_rejected():
    import pywrs
    L = pywrs.lib()
    plan = C.c_void_p()
    for t in (dict(merge_tile=3000), dict(merge_tile=4096, merge_per=2), dict(merge_per=3),
              dict(partition=2), dict(direct=7)):
        cfg = pywrs._config(32, pywrs.PREC_COMPENSATED, t)
        assert L.wrs_plan_create(C.byref(cfg), 1024, C.byref(plan)) == pywrs.WRS_EINVAL, t
        msg = L.wrs_last_error()
        assert b"merge shape" in msg or b"partition" in msg or b"direct" in msg
    assert C.sizeof(pywrs.Config) == 24 and C.sizeof(pywrs.Tuning) == 16
