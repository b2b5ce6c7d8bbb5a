"""CPU tests of the C-ABI boundary: the library loads, exports every entry
point include/dayflower_hip.h declares, and the ctypes records match the C
layouts.  No compute call is made (no GPU here)."""
import ctypes
import os
import re
import subprocess

import pytest

from dayflower_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dayflower_hip.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dfr_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = abi.load()
    names = _declared()
    assert len(names) >= 30
    for n in names:
        assert hasattr(lib, n), f"{n} not exported"


def test_abi_table_matches_header():
    assert sorted(n for n, _, _ in abi.ENTRY_POINTS) == _declared()


def test_ordinals_match_header():
    """Rendering-algorithm, tone-mapper and material ordinals of the header
    (the first two those of AbstractGPURenderer.java:50-87) are the ones the
    Python mirror passes."""
    text = open(HEADER).read()
    defs = dict((m.group(1), int(m.group(2))) for m in re.finditer(r"#define\s+DFR_((?:ALGO|TONE_MAPPER|MAT)_\w+)\s+(\d+)", text))
    assert len(defs) == 22
    for name, value in defs.items():
        assert getattr(abi, name) == value, name


@pytest.fixture(scope="module")
def c_layout(tmp_path_factory):
    d = tmp_path_factory.mktemp("layout")
    src = d / "layout.c"
    exe = d / "layout"
    fields = {
        "dfr_material": ["type", "emission", "kd", "eta", "k", "angle", "roughness_u", "disney_metallic",
                         "disney_flatness"],
        "dfr_primitive": ["shape", "material", "area_light", "first_triangle", "triangle_count", "data"],
        "dfr_light": ["type", "primitive", "triangle", "two_sided", "radiance", "position", "direction",
                      "turbidity", "cone_angle", "cone_delta", "radius"],
        "dfr_scene_desc": ["n_vertices", "positions", "normals", "uvs", "n_triangles", "indices", "n_primitives",
                           "primitives", "n_materials", "materials", "n_lights", "lights", "camera"],
        "dfr_stats": ["seconds", "path_samples", "rays_closest", "rays_any", "bounces", "extend_seconds",
                      "shadow_seconds", "extend_launches", "extend_rays", "extend_nodes", "extend_tris",
                      "shadow_rays", "shadow_nodes", "shadow_tris"],
        "dfr_job": ["scene", "width", "height", "spp", "maximum_bounce", "minimum_bounce_rr", "seed", "camera"],
        "dfr_server_opts": ["hbm_budget_bytes", "max_active", "pool_paths", "spp_quantum"],
        "dfr_server_stats": ["completed", "rejected", "path_samples", "busy_seconds", "peak_bytes", "peak_active"],
    }
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void) {"]
    for s, fs in fields.items():
        lines.append(f'printf("{s} size %zu\\n", sizeof({s}));')
        for f in fs:
            lines.append(f'printf("{s} {f} %zu\\n", offsetof({s}, {f}));')
    lines.append("return 0; }")
    src.write_text("\n".join(lines))
    subprocess.run(["gcc", "-std=c99", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    table = {}
    for line in out.splitlines():
        s, f, v = line.split()
        table[(s, f)] = int(v)
    return fields, table


def test_record_layouts_match_c(c_layout):
    fields, table = c_layout
    py = {"dfr_material": abi.Material, "dfr_primitive": abi.Primitive, "dfr_light": abi.Light,
          "dfr_scene_desc": abi.SceneDesc, "dfr_stats": abi.Stats, "dfr_job": abi.Job,
          "dfr_server_opts": abi.ServerOpts, "dfr_server_stats": abi.ServerStats}
    for s, cls in py.items():
        assert ctypes.sizeof(cls) == table[(s, "size")], s
        for f in fields[s]:
            assert getattr(cls, f).offset == table[(s, f)], (s, f)


def test_create_without_device_fails_cleanly():
    """No HIP device in this container: dfr_create returns NULL, and every
    entry point rejects a NULL context instead of crashing."""
    lib = abi.load()
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    assert not lib.dfr_create(0)
    assert lib.dfr_render(None, 1) == abi.E_NULL
    assert lib.dfr_setup(None, 1) == abi.E_NULL
    assert lib.dfr_set_i32(None, abi.KEY_MAXIMUM_BOUNCE, 3) == abi.E_NULL
    assert lib.dfr_request_stop(None) == abi.E_NULL
    assert lib.dfr_batch_render(None, None, 0, 0, None) == abi.E_NULL
