"""The C-ABI library loads and exports every entry point include/*.h declares
(no compute calls: these run without a GPU)."""
import ctypes
import os
import re

from dort_amd import abi, gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        src = open(os.path.join(ROOT, "include", h)).read()
        for m in re.finditer(r"^\s*(?:dgpu_status|void|uint32_t|const char\*)\s+(dgpu_\w+)\s*\(", src, re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    lib = gpu.lib()
    declared = _declared()
    assert "dgpu_render" in declared and "dgpu_trace_closest" in declared
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert not missing, missing
    assert set(gpu.EXPORTS) <= declared


def test_abi_version():
    assert gpu.lib().dgpu_abi_version() == 4


def test_struct_layouts_match_header(tmp_path):
    """The ctypes mirror has the sizes and field offsets the C compiler gives the header."""
    import subprocess
    structs = {"dgpu_material": abi.Material, "dgpu_area_light": abi.AreaLight,
               "dgpu_env_light": abi.EnvLight, "dgpu_texture": abi.Texture, "dgpu_camera": abi.Camera,
               "dgpu_scene_desc": abi.SceneDesc, "dgpu_bvh_opts": abi.BvhOpts, "dgpu_rays": abi.Rays,
               "dgpu_hits": abi.Hits, "dgpu_render_opts": abi.RenderOpts, "dgpu_stats": abi.RenderStats,
               "dgpu_scene_info": abi.SceneInfo}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "dortgpu.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)])
    got = dict(l.split() for l in subprocess.check_output([str(exe)]).decode().splitlines())
    for cname, py in structs.items():
        assert int(got[cname]) == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(got[f"{cname}.{f}"]) == getattr(py, f).offset, (cname, f)


def test_errors_without_device_are_reported():
    """With no HIP device (this container) dgpu_create fails cleanly with a message."""
    import pytest
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    h = ctypes.c_void_p()
    st = gpu.lib().dgpu_create(0, ctypes.byref(h))
    assert st in (abi.DGPU_EDEVICE, abi.DGPU_EINVAL)
    assert gpu.lib().dgpu_last_error()
