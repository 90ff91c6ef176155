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


# ---------------------------------------------------------------------------
# Register budget of the FP32 march kernels, read from the library's own gfx950
# code objects (the .hip_fatbin section holds one offload bundle per source
# file).  The march is issue-bound at 8 waves per SIMD, which needs <= 64
# VGPRs, and a kernel with a private segment (scratch) makes the runtime
# allocate scratch memory at its first launch.
# ---------------------------------------------------------------------------
_LLVM = Path("/opt/rocm/llvm/bin")
_BUNDLE_MAGIC = b"__CLANG_OFFLOAD_BUNDLE__"


def _kernel_metadata():
    """{kernel symbol: (vgpr_count, private_segment_fixed_size)} over every
    gfx950 code object of the built library."""
    bundler, readelf = _LLVM / "clang-offload-bundler", _LLVM / "llvm-readelf"
    if not (bundler.exists() and readelf.exists()):
        pytest.skip("ROCm LLVM tools not present")
    meta = {}
    with tempfile.TemporaryDirectory() as d:
        fat = Path(d) / "fat.bin"
        # (objcopy given one file rewrites it in place: always name an output)
        subprocess.run(["objcopy", "-O", "binary", "--only-section=.hip_fatbin",
                        str(abi.LIB_PATH), str(fat)], check=True, capture_output=True)
        blob = fat.read_bytes()
        starts = [m.start() for m in re.finditer(re.escape(_BUNDLE_MAGIC), blob)]
        assert starts, "no offload bundle in the library"
        for i, s in enumerate(starts):
            part = Path(d) / f"b{i}.bin"
            part.write_bytes(blob[s:starts[i + 1] if i + 1 < len(starts) else len(blob)])
            co = Path(d) / f"b{i}.co"
            subprocess.run([str(bundler), "--unbundle", "--type=o",
                            "--targets=hipv4-amdgcn-amd-amdhsa--gfx950", f"--input={part}",
                            f"--output={co}"], check=True, capture_output=True)
            notes = subprocess.run([str(readelf), "--notes", str(co)], check=True,
                                   capture_output=True, text=True).stdout
            for block in notes.split("  - .")[1:]:
                name = re.search(r"\.?name:\s+(\S+)", block)
                vg = re.search(r"\.vgpr_count:\s+(\d+)", block)
                ps = re.search(r"\.private_segment_fixed_size:\s+(\d+)", block)
                if name and vg and ps:
                    meta[name.group(1)] = (int(vg.group(1)), int(ps.group(1)))
    return meta


def test_fp32_march_kernels_fit_eight_waves_without_scratch():
    # every instance (runtime step count and the one compiled for the
    # reference's 20 steps) of the path, single-frame and batch-grid march of
    # scenes without escape or spheres
    meta = _kernel_metadata()
    hot = {k: v for k, v in meta.items()
           if re.search(r"(march_frames_kernel|march_kernel|batch_grid_kernel)ILb0ELb0E", k)}
    for family in ("march_frames_kernel", "12march_kernel", "batch_grid_kernel"):
        assert any(family in k for k in hot), (family, sorted(meta))
    assert any("Li20E" in k for k in hot), "no instance for the reference's step count"
    for name, (vgpr, scratch) in hot.items():
        assert vgpr <= 64, (name, vgpr)
        assert scratch == 0, (name, scratch)
