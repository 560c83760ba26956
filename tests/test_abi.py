"""C ABI of libmuhrec_hip.so: loads, exports every declared symbol, and
rejects invalid configurations before touching a device (CPU only)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(mr_\w+)\s*\(", txt, re.M)))


def test_header_symbols_exported():
    import imagingsuite_amd._lib as L
    lib = L.load()
    declared = _declared("muhrec_hip.h")
    assert len(declared) >= 15
    assert set(declared) == set(L.EXPORTED_SYMBOLS)
    for name in declared:
        assert hasattr(lib, name), name


def test_last_error_and_version():
    import imagingsuite_amd._lib as L
    lib = L.load()
    assert lib.mr_version() >= 1
    assert isinstance(lib.mr_last_error(), bytes)


def test_invalid_cutoff_rejected_without_device():
    import imagingsuite_amd._lib as L
    lib = L.load()
    plan = ctypes.c_void_p()
    rc = lib.mr_filter_plan_create(ctypes.byref(plan), 256, 3, 0.7, 1.0, 1, 0.1, 2)
    assert rc == L.MR_ERR_INVALID
    assert b"Cut off" in lib.mr_last_error()
    assert not plan.value


def test_filter_width_limit_without_device():
    """Widths up to 8192 run through the short form (a Q <= 16384 point
    convolution even where the padded length exceeds 16384); wider rows,
    whose short form would need 32768 points, are refused before any device
    call."""
    import imagingsuite_amd._lib as L
    lib = L.load()
    plan = ctypes.c_void_p()
    rc = lib.mr_filter_plan_create(ctypes.byref(plan), 8193, 3, 0.5, 1.0, 1, 0.1, 2)
    assert rc == L.MR_ERR_UNSUPPORTED
    assert b"outside the supported" in lib.mr_last_error()
    assert not plan.value


def test_bp_plan_rejects_bad_params():
    import imagingsuite_amd._lib as L
    lib = L.load()
    p = L.BpParams()
    p.width, p.slices, p.interp, p.direction = 1, 4, 0, 0
    plan = ctypes.c_void_p()
    assert lib.mr_bp_plan_create(ctypes.byref(plan), ctypes.byref(p)) == L.MR_ERR_INVALID
    p.width, p.use_roi = 64, 1
    p.roi[0], p.roi[1], p.roi[2], p.roi[3] = 10, 0, 5, 63   # x1 < x0
    assert lib.mr_bp_plan_create(ctypes.byref(plan), ctypes.byref(p)) == L.MR_ERR_INVALID
    assert b"ROI" in lib.mr_last_error()


def test_null_arguments_rejected():
    import imagingsuite_amd._lib as L
    lib = L.load()
    assert lib.mr_filter_rows(None, None, 0, None) == L.MR_ERR_INVALID
    assert lib.mr_bp_parallel(None, None, 0, None, None, None, 0, None) == L.MR_ERR_INVALID
    assert lib.mr_fdk_bp(None, None, 0, 0, None, None, None, None) == L.MR_ERR_INVALID


def test_filter_type_string_rejected():
    from imagingsuite_amd import MrError
    from imagingsuite_amd.recon import ProjectionFilterSingle
    with pytest.raises(MrError):
        ProjectionFilterSingle(256, filtertype="gaussian")


def test_batch_rejects_unknown_flags():
    import imagingsuite_amd._lib as L
    lib = L.load()
    plan = ctypes.c_void_p()
    arr = (L.BatchStack * 1)()
    assert lib.mr_batch_plan_create(ctypes.byref(plan), arr, 1, 4) == L.MR_ERR_INVALID
    assert not plan.value
    assert b"flags" in lib.mr_last_error()


def test_tilt_refusals_without_device():
    """The no-allocation entry point refuses a tilted configuration and the
    FDK plan refuses fused mode with tilt, both before any device call (the
    volume pointer is never dereferenced)."""
    import imagingsuite_amd._lib as L
    lib = L.load()
    p = L.BpParams()
    p.width, p.slices, p.interp, p.direction = 64, 8, 0, 0
    p.correct_tilt, p.tilt_angle = 1, 2.0
    rc = lib.mr_bp_parallel_noalloc(ctypes.byref(p), None, 0, None, None, ctypes.c_void_p(16), 0, None)
    assert rc == L.MR_ERR_UNSUPPORTED
    assert b"tilt" in lib.mr_last_error()
    q = L.FdkParams()
    q.sod, q.sdd = 500.0, 1000.0
    q.pixel[0] = q.pixel[1] = 0.2
    q.roi[0], q.roi[1], q.roi[2], q.roi[3] = 0, 0, 64, 32
    q.fused, q.correct_tilt, q.tilt_angle = 1, 1, 2.0
    plan = ctypes.c_void_p()
    assert lib.mr_fdk_plan_create(ctypes.byref(plan), ctypes.byref(q)) == L.MR_ERR_INVALID
    assert b"tilt" in lib.mr_last_error() and not plan.value


def test_struct_layouts_match_header(tmp_path):
    """The ctypes mirrors of the parameter structs have the C header's size
    and field offsets (gcc compiles a probe against include/muhrec_hip.h)."""
    import shutil
    import subprocess
    import imagingsuite_amd._lib as L
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("no C compiler")
    structs = {"mr_bp_params": L.BpParams, "mr_fdk_params": L.FdkParams, "mr_batch_stack": L.BatchStack,
               "mr_joseph_geom": L.JosephGeom}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "muhrec_hip.h"', "int main(void) {"]
    for cname, cls in structs.items():
        lines.append(f'printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run([gcc, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for ln in out:
        if ln:
            cname, key, val = ln.split()
            got[(cname, key)] = int(val)
    for cname, cls in structs.items():
        assert got[(cname, "sizeof")] == ctypes.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert got[(cname, fname)] == getattr(cls, fname).offset, (cname, fname)
