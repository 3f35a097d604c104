"""The C-ABI library loads on a CPU-only host and exports every function the public
header declares (no compute calls: there is no GPU here)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "chasqui_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(chq_\w+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ("chq_create", "chq_fluxes", "chq_corner_emf", "chq_update", "chq_boundaries",
              "chq_cons_to_prim", "chq_fofc_flag", "chq_fofc_repair", "chq_stage", "chq_step"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from chasqui_amd import LIB_PATH, lib
    L = lib()
    out = subprocess.run(["nm", "-D", "--defined-only", LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing
    for n in declared_functions():
        getattr(L, n)
    assert L.chq_version().decode().startswith("chasqui_amd")


def test_python_binding_covers_header():
    from chasqui_amd._capi import exported_symbols
    assert set(declared_functions()) <= set(exported_symbols()) | {"chq_sim_create"} | set(
        n for n in declared_functions() if n.startswith("chq_sim_") or n.startswith("chq_batch_"))


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    import chasqui_amd._capi as capi
    monkeypatch.setattr(capi, "_lib", None)
    monkeypatch.setattr(capi, "LIB_PATH", str(tmp_path / "absent.so"))
    with pytest.raises(capi.NativeLibraryMissing):
        capi.lib()


def test_null_handles_are_rejected_without_a_device():
    """Every entry point that takes a context, batch or server handle and returns a status
    rejects a NULL handle with CHQ_EINVAL and a message of its own, before any device call (this
    runs with no GPU); chq_batch_size, a count, gives 0. Before each call the thread's message is
    set to a known one (chq_create's), so a call that returns without setting its own fails."""
    import ctypes as C
    from chasqui_amd import _capi
    L = _capi.lib()
    checked = 0
    for n in _capi.exported_symbols():
        f = getattr(L, n)
        if f.restype is not C.c_int or not f.argtypes or f.argtypes[0] is not C.c_void_p:
            continue
        if n.startswith("chq_deck") or n.endswith("_destroy"):
            continue   # deck calls report through chq_deck_last_error; destroy(NULL) is a no-op
        assert L.chq_create(None, None, None) == -1
        marker = L.chq_last_error()
        assert marker
        args = [0 if t in (C.c_int, C.c_long, C.c_size_t) else 0.0 if t is C.c_double else None for t in f.argtypes]
        rc = f(*args)
        if n == "chq_batch_size":
            assert rc == 0
            continue
        assert rc == -1, (n, rc)          # CHQ_EINVAL
        msg = L.chq_last_error()
        assert msg and msg != marker, (n, msg)
        checked += 1
    assert checked >= 40
