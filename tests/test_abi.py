"""The C ABI library on the CPU: it loads, exports every entry point include/tdglhip.h declares,
its structs have the header's layout, and tdgl_supported names what it refuses (no GPU call)."""
import ctypes as ct
import os
import re
import subprocess

import pytest

from juliatdgl_amd import abi, hip
from juliatdgl_amd.scheme import Grid, box_masks, make_stepper
from cases import make_case, mgb2

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tdglhip.h")


def _declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(tdgl_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    L = hip.load()
    names = _declared()
    assert len(names) >= 10
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(hip.SYMBOLS)


def test_struct_layout_matches_header(tmp_path):
    src = tmp_path / "sz.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "tdglhip.h"\n'
                   'int main(){printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\\n", sizeof(tdgl_desc), sizeof(tdgl_state),'
                   ' sizeof(tdgl_pcg), sizeof(tdgl_stats), offsetof(tdgl_desc, f), offsetof(tdgl_desc, pcg_tol),'
                   ' sizeof(tdgl_meter), sizeof(tdgl_cycle_record), offsetof(tdgl_cycle_record, It),'
                   ' offsetof(tdgl_cycle_record, diffh));}\n')
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    got = [int(x) for x in out]
    want = [ct.sizeof(abi.DescC), ct.sizeof(abi.StateC), ct.sizeof(abi.PcgC), ct.sizeof(abi.StatsC),
            abi.DescC.f.offset, abi.DescC.pcg_tol.offset, ct.sizeof(abi.MeterC), ct.sizeof(abi.CycleRecordC),
            abi.CycleRecordC.It.offset, abi.CycleRecordC.diffh.offset]
    assert got == want


def test_supported_accepts_production_and_names_refusals():
    st, _ = make_case("production")
    ok, why = hip.supported_desc(st)
    assert ok and why == ""
    pd = abi.PackedDesc(st)
    pd.c.nbands = 7
    buf = ct.create_string_buffer(128)
    assert hip.load().tdgl_supported(ct.byref(pd.c), buf, 128) == 0
    assert b"bands" in buf.value
    pd = abi.PackedDesc(st)
    pd.c.abi_version = 99
    assert hip.load().tdgl_supported(ct.byref(pd.c), buf, 128) == 0
    assert b"ABI" in buf.value


def test_create_without_device_or_with_bad_desc_raises():
    st, _ = make_case("production")
    pd = abi.PackedDesc(st)
    pd.c.Nx = 1
    h = ct.c_void_p()
    rc = hip.load().tdgl_create(ct.byref(pd.c), ct.byref(h))
    assert rc != 0 and not h.value
    assert hip.load().tdgl_last_error()


def test_stepper_options_outside_the_path_are_refused():
    gr = Grid.make(8, 6, 4, dx=0.5)
    with pytest.raises(NotImplementedError):
        make_stepper(mgb2(), gr, box_masks(gr), materials=["metal"])
    with pytest.raises(NotImplementedError):
        make_stepper(mgb2(), gr, box_masks(gr), heat=1.0)
    with pytest.raises(NotImplementedError):
        make_stepper(mgb2(), gr, box_masks(gr), drag=0.1)
