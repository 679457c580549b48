"""GPU parity at the benchmark workloads themselves, at full size.

Each test renders one whole frame of a benchmark workload through the C-ABI
(instance culling, classify, split, instantiate, CLAS build, BLAS
insertion, BVH build and a full-viewport trace of primary and shadow rays)
and runs the threaded CPU oracle on the same inputs:

* counters, readback, every list, record, CLAS block and BLAS reference list
  bit-exact, generated positions within 2 ULP (``harness.compare``);
* the BVH leaves of every instance equal the stable sort of the CLAS
  references by their 64-bit keys, recomputed on the host from the CLAS
  headers (``harness.check_blas_key_order``);
* a seeded sample of the viewport's pixels is cast by the oracle's
  hierarchy-free caster: hit identities agree on >= 99.99 % of the rays,
  hit t within 1e-5 relative, decoded hits and shadow flags likewise.

Workloads: 64 instances of a 256 x 256-quad grid with a 2048^2 R16 map at
1920 x 1080 and 4 px (plain flags, and the reference renderer's default
flags with PN and both transient build forms); the 500-mesh batched frame
with its own 1024^2 map per mesh; one frame of the 3840 x 2160, 2 px
spline sequence over 16.8M base triangles with a 4096^2 map.  The
reference path is the frame of renderer_raytrace_clusters_tess.cpp:410-780.
"""
import os

import numpy as np
import pytest

import harness
from vk_tessellated_clusters_amd import abi, scene

pytestmark = pytest.mark.gpu

TRACE = abi.TRACE_CULL_BACK | abi.TRACE_SHADOW


def host_threads():
    return max(1, min(16, os.cpu_count() or 1))


def run_workload(host, cfg, fc, n_pixels, seed, device_textures=False):
    """GPU frame and oracle frame on the same inputs, compared as the
    module docstring states.  device_textures: the oracle samples the height
    maps the device generated (tc_fbm_height, pinned against numpy by
    test_device_fbm_matches_numpy).  Returns (oracle results, oracle hits,
    hit agreement, checked BVH references)."""
    W, H = host.camera.width, host.camera.height
    print("workload %s: %d instances, %d base tris, %dx%d, flags 0x%x" % (
        host.name, len(host.instances), host.num_base_triangles, W, H, cfg.flags), flush=True)
    g = harness.GpuFrame(host, cfg)
    try:
        g.run(fc, stages="accel")
        res = g.results()
        arr = g.accel_arrays()
        nref = harness.check_blas_key_order(g, arr)
        gh, gd = g.trace(W, H, TRACE)
        tex = g.device_textures() if device_textures else None
    finally:
        g.close()
    print("  gpu frame done: %d CLAS references in key order" % nref, flush=True)
    of = harness.OracleFrame(host, cfg, textures=tex)
    of.run(fc, nthreads=host_threads())
    ref = of.results()
    harness.compare(res, ref)
    print("  tessellation, CLAS and insertion bit-exact: %d displaced tris, %d CLAS" % (
        ref.readback["numTotalTriangles"], len(ref.arrays["tempClusterSizes"])), flush=True)
    px = np.sort(np.random.default_rng(seed).choice(W * H, n_pixels, replace=False)).astype(np.uint32)
    oh, od = of.trace_pixels(W, H, TRACE, px, nthreads=host_threads())
    agree = harness.compare_hits(gh[px], oh, gd[px], od)
    print("  trace: %d sampled pixels, agreement %.6f, %d hits, %d shadowed" % (
        n_pixels, agree, int((oh["t"] > 0).sum()), int(od["shadowed"].sum())), flush=True)
    return ref, oh, od, agree, nref


def test_grid64_frame_parity():
    """64 x 256^2-quad grids, 2048^2 R16 map, 1920 x 1080 at 4 px, the
    benchmark's flags (displacement, culling)."""
    host = scene.scene_grid_instances()
    ref, oh, od, agree, nref = run_workload(host, scene.workload_config(), host.frame(), 16384, 11)
    assert ref.readback["numTotalTriangles"] > 9_000_000
    assert (oh["t"] > 0).mean() > 0.3
    assert nref == len(ref.arrays["blasClusterOffsets"])


def test_grid64_reference_default_flags_parity():
    """The same frame with the reference renderer's default flags
    (renderer.hpp:43-48): PN base surface, 1X transient clusters and 2X
    mini batches on."""
    host = scene.scene_grid_instances()
    cfg = scene.workload_config(flags=scene.REFERENCE_DEFAULT_FLAGS)
    ref, oh, od, agree, nref = run_workload(host, cfg, host.frame(), 8192, 12)
    modes = ref.arrays["transBuilds"]["clusterID"] >> 30
    assert (modes == abi.RT_CLUSTER_MODE_1X_SUBSET_CLUSTER).sum() > 0
    assert (modes == abi.RT_CLUSTER_MODE_2X_BATCHED_TESSELLATED).sum() > 0
    hit_modes = oh["clusterID"][oh["t"] > 0] >> 30
    assert len(np.unique(hit_modes)) >= 2


def test_batched500_frame_parity():
    """500 meshes (log-uniform 200..200000 triangles, grids and
    icospheres), each with its own 1024^2 R16 map: more than 256 instances
    with many clusters each, so the BVH key sort runs its instance digits."""
    host = scene.scene_batched()
    ref, oh, od, agree, nref = run_workload(host, scene.workload_config(), host.frame(), 8192, 13,
                                            device_textures=True)
    infos = ref.arrays["blasBuildInfos"]
    busy = np.nonzero(infos["clusterReferencesCount"] > 1)[0]
    assert busy.size > 50 and busy.max() >= 256
    assert len(np.unique(oh["instanceID"][oh["t"] > 0])) > 20


def test_large_spline_frame_parity():
    """One frame of the 3840 x 2160, 2 px spline sequence: 8 x 8 patches
    of 512 x 256 quads (16.8M base triangles), 4096^2 R16 map."""
    host = scene.scene_large()
    span_x, span_z, seed = host.spline
    host.camera.eye, host.camera.center = scene.camera_spline(37, 100, span_x, span_z, seed)
    cfg = scene.workload_config(large=True)
    ref, oh, od, agree, nref = run_workload(host, cfg, host.frame(), 4096, 14, device_textures=True)
    assert ref.readback["numTotalTriangles"] > 40_000_000
    assert ref.counters["splitWriteCounter"] > 0
