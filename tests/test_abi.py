"""C ABI of libqpad_hip.so (CPU-only checks: no kernel is launched).

* every entry point declared in include/qpad_hip.h is exported;
* the ctypes structures match the C layout (a gcc-compiled probe prints
  sizeof/offsetof of each struct);
* argument validation and error reporting work without a GPU.
"""
import ctypes as C
import os
import re
import subprocess

import pytest

import qpad_amd._lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "qpad_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(qpad_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    lib = L.load()
    names = declared_functions()
    assert len(names) >= 35
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the ctypes table covers exactly the declared set
    assert sorted(L.SIGNATURES) == names


def test_struct_layout_matches_header(tmp_path):
    src = tmp_path / "probe.c"
    fields = {
        "qpad_grid": [f for f, _ in L.Grid._fields_],
        "qpad_species_cfg": [f for f, _ in L.SpeciesCfg._fields_],
        "qpad_beam_cfg": [f for f, _ in L.BeamCfg._fields_],
        "qpad_launch_shape": [f for f, _ in L.LaunchShape._fields_],
        "qpad_neutral_cfg": [f for f, _ in L.NeutralCfg._fields_],
    }
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){"]
    for st, fs in fields.items():
        lines.append(f'printf("{st} %zu\\n", sizeof({st}));')
        for f in fs:
            lines.append(f'printf("{st}.{f} %zu\\n", offsetof({st}, {f}));')
    lines.append("return 0;}")
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c11", str(src), "-o", str(exe)], check=True)
    out = dict(line.rsplit(" ", 1) for line in subprocess.run([str(exe)], check=True, capture_output=True,
                                                               text=True).stdout.split("\n") if line)
    for st, cls in (("qpad_grid", L.Grid), ("qpad_species_cfg", L.SpeciesCfg),
                    ("qpad_beam_cfg", L.BeamCfg), ("qpad_launch_shape", L.LaunchShape),
                    ("qpad_neutral_cfg", L.NeutralCfg)):
        assert int(out[st]) == C.sizeof(cls), st
        for f, _ in cls._fields_:
            assert int(out[f"{st}.{f}"]) == getattr(cls, f).offset, f"{st}.{f}"


def test_abi_version():
    """2: qpad_launch_shape gained dep_mfma (the struct grew)"""
    assert L.load().qpad_abi_version() == 2


def test_invalid_grid_is_rejected_with_message():
    lib = L.load()
    ctx = C.c_void_p()
    g = L.Grid(4, 16, 1, 0.1, 0.05, 1e-3, L.BND_OPEN, 10, 1e-3, 1e-3, 10.0)   # nr < 8
    rc = lib.qpad_create(C.byref(g), C.byref(ctx))
    assert rc == -1
    assert b"nr" in lib.qpad_last_error()
    g = L.Grid(64, 16, 7, 0.1, 0.05, 1e-3, L.BND_OPEN, 10, 1e-3, 1e-3, 10.0)   # max_mode > 4
    assert lib.qpad_create(C.byref(g), C.byref(ctx)) == -1
    g = L.Grid(64, 16, 1, 0.1, 0.05, 1e-3, 7, 10, 1e-3, 1e-3, 10.0)            # bad boundary
    assert lib.qpad_create(C.byref(g), C.byref(ctx)) == -1
    assert lib.qpad_create(None, C.byref(ctx)) == -1


def test_null_context_calls_fail_cleanly():
    lib = L.load()
    assert lib.qpad_slice(None, 1, None) == -1
    assert lib.qpad_step3d_begin(None) == -1
    assert lib.qpad_destroy(None) == 0


def test_simulation_raises_without_device_or_library(monkeypatch):
    from qpad_amd import QpadError, Simulation, configs
    import qpad_amd._lib as lib_mod
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if not has_gpu:
        with pytest.raises(QpadError):
            Simulation(configs.small())
    # a missing library is an error, never a fallback
    monkeypatch.setattr(lib_mod, "_lib", None)
    with pytest.raises(QpadError, match="missing"):
        lib_mod.load("/nonexistent/libqpad_hip.so")


def test_adk_table_matches_oracle():
    """qpad_neutral_adk (a host table read, no GPU) returns the oracle's ADK
    coefficients for every element and level the path follows; levels
    beyond an element's charge are refused"""
    import pyoracle as po
    from qpad_amd.input import NEUTRAL_ELEMENTS
    lib = L.load()
    ol = po.lib()
    out = (C.c_double * 3)()
    n = 0
    for elem, nlev in NEUTRAL_ELEMENTS.items():
        for lev in range(1, nlev + 1):
            assert lib.qpad_neutral_adk(elem, lev, out) == 0, (elem, lev)
            ref = ol.ora_adk(elem, lev)
            assert [out[t] for t in range(3)] == [ref[t] for t in range(3)], (elem, lev)
            n += 1
        assert lib.qpad_neutral_adk(elem, nlev + 1, out) != 0
    assert n == 55
