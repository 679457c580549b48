"""The C-ABI library loads, exports every entry point include/tc_abi.h
declares, its records have the C sizes, and argument errors come back as
status codes (no GPU work is issued by these calls)."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

from harness import oracle, ROOT
from vk_tessellated_clusters_amd import abi

HEADER = ROOT / "include" / "tc_abi.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:tc_status|uint32_t|uint64_t)\s+(tc_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    L = abi.lib()
    names = declared_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(L, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", str(abi.LIB_PATH)], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (tc_\w+)", out))
    assert set(names) <= exported
    assert set(abi.EXPORTS) <= exported


def test_record_sizes():
    for name, size in abi.EXPECTED_SIZES.items():
        assert C.sizeof(getattr(abi, name)) == size, name
    assert abi.TESS_TRIANGLE_INFO.itemsize == 24 and abi.TEMPLATE_INSTANTIATE_INFO.itemsize == 32
    assert abi.BLAS_BUILD_INFO.itemsize == 16 and abi.CLUSTER_DT.itemsize == 16 and abi.BBOX_DT.itemsize == 32
    assert C.sizeof(abi.Bindings) == 120
    assert C.sizeof(abi.Hiz) == 32
    assert C.sizeof(abi.AdjointParams) == 80
    assert abi.VERTEX_TAPE_DT.itemsize == 48


def test_header_compiles_as_c_and_cxx(tmp_path):
    src = tmp_path / "t.c"
    src.write_text('#include "tc_abi.h"\nint main(void){return sizeof(tc_scene_building) == 368 ? 0 : 1;}\n')
    for cc, f in (("gcc", src), ("g++", src)):
        exe = tmp_path / ("a_" + cc)
        r = subprocess.run([cc, "-x", "c" if cc == "gcc" else "c++", "-I", str(ROOT / "include"), str(f), "-o", str(exe)],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        assert subprocess.run([str(exe)]).returncode == 0


def test_host_queries():
    L = abi.lib()
    cfg = abi.default_config()
    assert L.tc_workspace_size(C.byref(cfg)) > 0
    bad = abi.default_config(maxSplitFactor=1)
    assert L.tc_workspace_size(C.byref(bad)) == 0
    for f in range(2, 12):
        assert L.tc_num_split_passes(f) == oracle().oracle_num_split_passes(f)
    assert L.tc_num_split_passes(8) == 4  # renderer_raytrace_clusters_tess.cpp:516-543 at splitFactor 8


def test_clas_size_formula():
    L = abi.lib()
    r16 = lambda v: (v + 15) & ~15
    for nt in (1, 4, 5, 64, 121, 128):
        for nv in (3, 45, 78, 81, 128):
            ng = (nt + 3) // 4
            assert L.tc_clas_size(nt, nv) == 48 + r16(24 * ng) + r16(4 * nt) + r16(12 * nv)


def test_argument_errors_without_gpu_work():
    L = abi.lib()
    cfg = abi.default_config()
    assert L.tc_cluster_classify(None, cfg, None) == 1
    b = abi.Bindings()
    assert L.tc_triangle_split(C.byref(b), cfg, None) == 1
    assert L.tc_build_setup(None, cfg, 0, None) == 1
    assert L.tc_frame_tessellate(None, cfg, 0, None) == 1
    assert L.tc_memcpy(None, None, 16) == 1
    assert L.tc_cluster_template_order(0, 0, None, None, None) == 1
    b.view = b.tessTable = b.readback = b.instances = b.build = b.workspace = b.vertexTape = 16
    b.numRenderInstances = 1
    # an unknown flag bit, and the vertex tape with flags it does not cover
    for flags in (1 << 20, abi.FLAG_VERTEX_TAPE | abi.FLAG_2X_TRANSIENT, abi.FLAG_VERTEX_TAPE | abi.FLAG_ANIMATION):
        assert L.tc_cluster_classify(C.byref(b), abi.default_config(flags=flags), None) == 5, hex(flags)


def test_cluster_template_order_is_a_permutation():
    import numpy as np
    from vk_tessellated_clusters_amd import scene
    from vk_tessellated_clusters_amd.device import cluster_templates
    geo = scene.icosphere(16)
    words, sizes = cluster_templates(geo)
    for c in geo.clusters[:20]:
        t0, nt = int(c["firstTriangle"]), int(c["numTriangles"])
        w = words[t0:t0 + nt]
        assert sorted((w >> 24).tolist()) == list(range(nt))
        src = geo.local_triangles[3 * t0:3 * (t0 + nt)].reshape(-1, 3)[w >> 24]
        assert np.array_equal(w & 0xFF, src[:, 0]) and np.array_equal((w >> 8) & 0xFF, src[:, 1])


def test_tuning_roundtrip_and_validation():
    import ctypes as C
    L = abi.lib()
    t = abi.Tuning()
    abi.check(L.tc_get_tuning(C.byref(t)))
    default = t.as_dict()
    assert default["traceThreads"] in (64, 128, 256) and all(v > 0 for v in default.values())
    try:
        t.classifyBlocksPerCU, t.traceThreads = 7, 256
        abi.check(L.tc_set_tuning(C.byref(t)))
        u = abi.Tuning()
        abi.check(L.tc_get_tuning(C.byref(u)))
        assert u.classifyBlocksPerCU == 7 and u.traceThreads == 256
        bad = abi.Tuning(**dict(u.as_dict(), traceThreads=96))
        assert L.tc_set_tuning(C.byref(bad)) == abi.TC_ERROR_INVALID_ARGUMENT
        bad = abi.Tuning(**dict(u.as_dict(), splitBlocksPerCU=0))
        assert L.tc_set_tuning(C.byref(bad)) == abi.TC_ERROR_INVALID_ARGUMENT
    finally:
        abi.check(L.tc_set_tuning(None))
    abi.check(L.tc_get_tuning(C.byref(t)))
    assert t.as_dict() == default


def test_noalloc_tessellate_rejects_oversized_configs():
    """tc_frame_tessellate_noalloc refuses configs whose tile counts exceed
    its fixed device array, before any device call."""
    import ctypes as C
    L = abi.lib()
    b = abi.Bindings()
    for f in ("view", "tessTable", "readback", "instances", "build"):
        setattr(b, f, 0x1000)
    b.numRenderInstances = 1
    for field, value in (("maxPartTriangles", (1 << 24) + 1024), ("maxVisibleClusters", (1 << 20) + 64),
                         ("maxSplitTriangles", (1 << 20) + 64)):
        cfg = abi.default_config(**{field: value})
        assert L.tc_frame_tessellate_noalloc(C.byref(b), cfg, 0, None) == abi.TC_ERROR_CAPACITY, field
    bad = abi.default_config(clusterTriangleCount=0)
    assert L.tc_frame_tessellate_noalloc(C.byref(b), bad, 0, None) == abi.TC_ERROR_INVALID_ARGUMENT
