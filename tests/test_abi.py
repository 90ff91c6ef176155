"""The C ABI library without a GPU: it loads, exports every symbol that
include/bh_api.h declares, its structures have the same layout in C and in
the ctypes mirror, and calls that need a device fail with a status code."""
import ctypes as C
import re
import subprocess
import tempfile
from pathlib import Path

import pytest

from blackhole_8_amd import abi

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "bh_api.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w\s\*]*?\b(bh_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = abi.load()
    names = declared_functions()
    assert len(names) >= 25
    out = subprocess.run(["nm", "-D", "--defined-only", str(abi.LIB_PATH)], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (bh_\w+)$", out, flags=re.M))
    missing = [n for n in names if n not in exported]
    assert not missing, f"declared but not exported: {missing}"
    assert set(names) == set(abi.SIGNATURES), "ctypes signatures out of step with the header"
    for n in names:
        assert getattr(lib, n) is not None


STRUCTS = {
    "bh_image": abi.Image, "bh_camera": abi.Camera, "bh_hole": abi.Hole,
    "bh_annulus": abi.Annulus, "bh_sphere": abi.Sphere, "bh_scene": abi.Scene,
    "bh_options": abi.Options, "bh_outputs": abi.Outputs, "bh_rect": abi.Rect,
    "bh_plane": abi.Plane, "bh_flat_scene": abi.FlatScene, "bh_diff_params": abi.DiffParams,
    "bh_batch_tuning": abi.BatchTuning, "bh_batch_stats": abi.BatchStats,
}


def test_struct_layout_matches_c():
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', 'int main(void) {']
    for cname, cls in STRUCTS.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    with tempfile.TemporaryDirectory() as d:
        src = Path(d) / "layout.c"
        exe = Path(d) / "layout"
        src.write_text("\n".join(lines))
        subprocess.run(["gcc", "-std=c11", str(src), "-o", str(exe)], check=True)
        out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    got = dict(line.rsplit(" ", 1) for line in out.strip().splitlines())
    for cname, cls in STRUCTS.items():
        assert int(got[f"{cname} size"]) == C.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert int(got[f"{cname}.{fname}"]) == getattr(cls, fname).offset, f"{cname}.{fname}"


def _has_device():
    n = C.c_int(0)
    try:
        hip = C.CDLL("libamdhip64.so")
        return hip.hipGetDeviceCount(C.byref(n)) == 0 and n.value > 0
    except OSError:
        return False


def test_calls_without_a_device_or_handle_fail_with_status():
    lib = abi.load()
    assert lib.bh_version() == 1
    ctx = C.c_void_p()
    rc = lib.bh_ctx_create(0, C.byref(ctx))
    if _has_device():
        assert rc == abi.BH_OK
        lib.bh_ctx_destroy(ctx)
    else:
        assert rc == abi.BH_ERR_NO_DEVICE and not ctx.value
        assert b"device" in lib.bh_last_error(None)
    assert lib.bh_ctx_create(0, None) == abi.BH_ERR_INVALID_ARG
    s, o, out = abi.Scene(), abi.Options(), abi.Outputs()
    assert lib.bh_render(None, C.byref(s), C.byref(o), C.byref(out)) == abi.BH_ERR_BAD_HANDLE
    assert lib.bh_render_device(None, C.byref(s), C.byref(o), C.byref(out)) == abi.BH_ERR_BAD_HANDLE
    assert lib.bh_texture_set(None, 0, None) == abi.BH_ERR_BAD_HANDLE
    assert lib.bh_scene_check(None, C.byref(s), C.byref(o)) == abi.BH_ERR_BAD_HANDLE
    assert lib.bh_path_launch(None, None) == abi.BH_ERR_BAD_HANDLE
    lib.bh_ctx_destroy(None)
    lib.bh_path_destroy(None)


def test_oracle_library_is_separate_from_the_product():
    """The product library never links the oracle."""
    out = subprocess.run(["nm", "-D", str(abi.LIB_PATH)], capture_output=True, text=True,
                         check=True).stdout
    assert "or_render" not in out and "or_trace_pixel" not in out
    deps = subprocess.run(["ldd", str(abi.LIB_PATH)], capture_output=True, text=True).stdout
    assert "oracle" not in deps
