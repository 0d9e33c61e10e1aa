"""The C ABI without a GPU: the library loads, exports every entry point include/ge_rt.h declares,
its records have the engine's sizes, and calls that need no device fail cleanly."""
import ctypes as C
import os
import re
import subprocess

import pytest

from game_engine_amd import abi, rt
from oracle import binding as ob

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "ge_rt.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int32_t|void|uint64_t)\s+(ge_(?:rt|clusters)_\w+)\(", text, re.M)))


def test_library_exports_every_declared_entry_point():
    names = declared()
    assert len(names) >= 18
    lib = rt.lib()
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(rt.EXPORTS) == names


def test_record_sizes_match_the_engine():
    assert C.sizeof(abi.PathTraceParams) == 368 and C.sizeof(abi.RayVisibilityParams) == 256
    assert C.sizeof(abi.ClusterDesc) == 48 and C.sizeof(abi.InstanceDesc) == 96 and C.sizeof(abi.MeshDesc) == 64
    assert C.sizeof(abi.ResolveMaterial) == 128 and C.sizeof(abi.ResolveLight) == 64 and C.sizeof(abi.FrameEye) == 48
    text = open(HEADER).read()
    for struct, size in re.findall(r"static_assert\(sizeof\((\w+)\) == (\d+)", text):
        mirror = {"ge_cluster_desc": abi.ClusterDesc, "ge_mesh_desc": abi.MeshDesc, "ge_instance_desc": abi.InstanceDesc,
                  "ge_frame_eye": abi.FrameEye, "ge_resolve_material": abi.ResolveMaterial,
                  "ge_resolve_light": abi.ResolveLight, "ge_path_trace_params": abi.PathTraceParams,
                  "ge_ray_visibility_params": abi.RayVisibilityParams, "ge_vertex_attributes": abi.VertexAttributes,
                  "ge_texture_desc": abi.TextureDesc, "ge_sampler_desc": abi.SamplerDesc,
                  "ge_scene_desc": abi.SceneDesc, "ge_rt_counters": abi.Counters}
        if struct in mirror:
            assert C.sizeof(mirror[struct]) == int(size), struct
        dtypes = {"ge_ray": abi.RAY_DTYPE, "ge_hit": abi.HIT_DTYPE, "ge_bvh_node": abi.NODE_DTYPE,
                  "ge_bvh_wide_node": abi.WIDE_DTYPE, "ge_bvh_triangle": abi.TRI_DTYPE,
                  "ge_shadow_receiver": abi.RECEIVER_DTYPE}
        if struct in dtypes:
            assert dtypes[struct].itemsize == int(size), struct


@pytest.mark.skipif(not os.path.isdir(ob.REFERENCE), reason="the reference's headers are not on this machine")
def test_record_fields_match_the_engines_headers():
    """oracle/abi_pins.cpp compiles include/ge_rt.h beside the reference's path_trace.h,
    ray_visibility.h, cluster_cull.h, visibility_resolve.h and geometry/cluster.h, with a
    static_assert on the offset and size of every mirrored field and the value of every mirrored
    constant; the same file with one field compared against its neighbour must not compile."""
    run = lambda *extra: subprocess.run(["make", "-s", "-C", os.path.dirname(ob.__file__), "abi",  # noqa: E731
                                         f"REFERENCE={ob.REFERENCE}", *extra], capture_output=True, text=True)
    good = run()
    assert good.returncode == 0, good.stderr
    bad = run("ABI_FLAGS=-DGE_ABI_PINS_PLANT_MISMATCH")
    assert bad.returncode != 0 and "ge_path_trace_params::width is not at" in bad.stderr
    pins = open(os.path.join(os.path.dirname(ob.__file__), "abi_pins.cpp")).read()
    fields = re.findall(r"^GE_PIN\((\w+), [\w:]+, (\w+)\);", pins, re.M)
    for struct, mirror in (("ge_path_trace_params", abi.PathTraceParams), ("ge_ray_visibility_params", abi.RayVisibilityParams),
                           ("ge_instance_desc", abi.InstanceDesc), ("ge_mesh_desc", abi.MeshDesc),
                           ("ge_frame_eye", abi.FrameEye), ("ge_resolve_material", abi.ResolveMaterial),
                           ("ge_resolve_light", abi.ResolveLight), ("ge_cluster_desc", abi.ClusterDesc)):
        assert sorted(f for s, f in fields if s == struct) == sorted(f[0] for f in mirror._fields_), struct


MIRRORS = {"ge_path_trace_params": abi.PathTraceParams, "ge_ray_visibility_params": abi.RayVisibilityParams,
           "ge_instance_desc": abi.InstanceDesc, "ge_mesh_desc": abi.MeshDesc, "ge_frame_eye": abi.FrameEye,
           "ge_resolve_material": abi.ResolveMaterial, "ge_resolve_light": abi.ResolveLight,
           "ge_cluster_desc": abi.ClusterDesc, "ge_texture_desc": abi.TextureDesc, "ge_sampler_desc": abi.SamplerDesc,
           "ge_scene_desc": abi.SceneDesc, "ge_rt_counters": abi.Counters}


def test_python_records_match_the_header_field_by_field(tmp_path):
    """The ctypes mirrors against the C compiler's own offsets of include/ge_rt.h's fields."""
    lines = ["#include <stddef.h>", "#include <stdio.h>", f'#include "{HEADER}"', "int main(void) {"]
    for struct, mirror in MIRRORS.items():
        for name, *_ in mirror._fields_:
            lines.append(f'  printf("{struct} {name} %zu\\n", offsetof({struct}, {name}));')
    lines.append("  return 0;\n}")
    src, exe = tmp_path / "offsets.c", tmp_path / "offsets"
    src.write_text("\n".join(lines))
    subprocess.run(["gcc", "-std=c11", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    seen = 0
    for line in filter(None, out):
        struct, name, off = line.split()
        assert getattr(MIRRORS[struct], name).offset == int(off), (struct, name)
        seen += 1
    assert seen == sum(len(m._fields_) for m in MIRRORS.values())


def test_record_dtypes_match_the_header_field_by_field(tmp_path):
    """The numpy views of the walks' records against the C compiler's offsets of their fields."""
    dtypes = {"ge_bvh_node": abi.NODE_DTYPE, "ge_bvh_wide_node": abi.WIDE_DTYPE, "ge_bvh_triangle": abi.TRI_DTYPE}
    lines = ["#include <stddef.h>", "#include <stdio.h>", f'#include "{HEADER}"', "int main(void) {"]
    for struct, dt in dtypes.items():
        for name in dt.names:
            lines.append(f'  printf("{struct} {name} %zu\\n", offsetof({struct}, {name}));')
    lines.append("  return 0;\n}")
    src, exe = tmp_path / "offsets.c", tmp_path / "offsets"
    src.write_text("\n".join(lines))
    subprocess.run(["gcc", "-std=c11", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    for line in filter(None, out):
        struct, name, off = line.split()
        assert dtypes[struct].fields[name][1] == int(off), (struct, name)
    assert len(list(filter(None, out))) == sum(len(d.names) for d in dtypes.values())


def test_null_handles_fail_cleanly():
    lib = rt.lib()
    lib.ge_rt_destroy(None)  # idempotent on null (ReferenceRenderer::destroy)
    assert lib.ge_rt_last_error(None, None, 0) == 1
    assert lib.ge_rt_synchronize(None) == 1
    assert lib.ge_rt_create(None, 0) == 1
    assert lib.ge_rt_stats(None, None) == 1
