"""CPU checks of the C-ABI boundary: the in-tree library loads, exports every
symbol that include/cabanapic_amd.h declares, and the Python binding declares
each of them.  No device call is made."""
import ctypes
import os
import re

import pytest

from cabanapic_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cabanapic_amd.h")


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if fn.endswith(".h"):
            text = open(os.path.join(inc, fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(cpic_[a-z_0-9]+)\s*\(", text))
    return names


def test_header_declares_entry_points():
    names = declared_symbols()
    for n in ("cpic_create", "cpic_push", "cpic_load_interpolator_array",
              "cpic_unload_accumulator_array", "cpic_advance_b", "cpic_advance_e",
              "cpic_energies", "cpic_step"):
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_capi.LIB_PATH)
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert declared_symbols() == set(_capi.SIGNATURES)


def test_struct_layouts_match_header(tmp_path):
    """Every struct of the header, compiled by gcc: its size and the offset
    of every member equal those of the ctypes binding."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    structs = {"cpic_config": _capi.Config, "cpic_step_params": _capi.StepParams,
               "cpic_plasma_spec": _capi.PlasmaSpec}
    lines = ["#include <stdio.h>", "#include <stddef.h>", f'#include "{HEADER}"',
             "int main(void) {"]
    for name, cls in structs.items():
        lines.append(f'  printf("{name} %zu\\n", sizeof({name}));')
        for field, _ in cls._fields_:
            lines.append(f'  printf("{name}.{field} %zu\\n", offsetof({name}, {field}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-o", str(exe), str(src)], check=True)
    out = dict(l.split() for l in subprocess.run([str(exe)], check=True, capture_output=True,
                                                 text=True).stdout.splitlines())
    for name, cls in structs.items():
        assert int(out[name]) == ctypes.sizeof(cls), name
        for field, _ in cls._fields_:
            assert int(out[f"{name}.{field}"]) == getattr(cls, field).offset, (name, field)


def test_last_error_without_device_call():
    lib = _capi.load()
    assert isinstance(lib.cpic_last_error(), bytes)


def test_invalid_config_rejected_before_any_device_call():
    lib = _capi.load()
    cfg = _capi.Config(3, 1, 4, 4, 4, 10)
    h = ctypes.c_void_p()
    assert lib.cpic_create(ctypes.byref(cfg), ctypes.byref(h)) == _capi.CPIC_ERR_INVALID
    assert b"real_size" in lib.cpic_last_error()
    cfg = _capi.Config(4, 2, 4, 4, 4, 10)
    assert lib.cpic_create(ctypes.byref(cfg), ctypes.byref(h)) == _capi.CPIC_ERR_UNSUPPORTED
    with pytest.raises(_capi.CpicError):
        _capi.check(lib.cpic_destroy(None) or _capi.CPIC_ERR_INVALID)


def test_column_loaders_reject_invalid_specs():
    """cpic_init_plasma and cpic_init_column_fields (host-side loaders, no
    device call) refuse an unknown profile, a column without a ring width or
    guide field, and a field request for a uniform plasma."""
    import numpy as np
    lib = _capi.load()

    def spec(**kw):
        s = _capi.PlasmaSpec()
        s.seed, s.nx, s.ny, s.nz, s.nppc, s.weight = 3, 8, 8, 2, 2, 1.0
        s.profile, s.len_x, s.len_y = 1, 0.8, 0.8
        s.ring_radius, s.ring_width, s.column_charge, s.guide_bz = 0.24, 0.04, -0.25, 1.0
        for k, v in kw.items():
            setattr(s, k, v)
        return s

    n = 8 * 8 * 2 * 2
    arrs = [np.zeros(n) for _ in range(7)]
    cell = np.zeros(n, np.int32)
    ptrs = [a.ctypes.data_as(ctypes.c_void_p) for a in arrs] + [cell.ctypes.data_as(ctypes.c_void_p)]
    f = np.zeros((9, 10 * 10 * 4))
    fp = f.ctypes.data_as(ctypes.c_void_p)
    assert lib.cpic_init_plasma(ctypes.byref(spec()), 8, n, *ptrs) == _capi.CPIC_OK
    assert lib.cpic_init_column_fields(ctypes.byref(spec()), 8, fp) == _capi.CPIC_OK
    assert np.all(f[5] == 1.0) and np.any(f[0]) and np.any(f[1])
    for bad in (dict(profile=2), dict(ring_width=0.0), dict(guide_bz=0.0), dict(len_y=0.0)):
        assert lib.cpic_init_plasma(ctypes.byref(spec(**bad)), 8, n, *ptrs) == _capi.CPIC_ERR_INVALID, bad
    assert lib.cpic_init_column_fields(ctypes.byref(spec(profile=0)), 8, fp) == _capi.CPIC_ERR_INVALID
    assert lib.cpic_init_column_fields(ctypes.byref(spec()), 3, fp) == _capi.CPIC_ERR_INVALID
