#!/usr/bin/env python3
"""Frame benchmark of the tessellated-cluster path on one MI355X.

Workload (default): 64 instances of a 256 x 256-quad grid patch (131072
triangles each, FP32 positions) with seeded rigid transforms and frustum
culling, a 2048^2 R16_UNORM fBm height map (seed 1, 6 octaves, bilinear),
1920 x 1080, 4 px per segment, at most 11 segments per edge.

A step is one whole frame with every input resident in HBM:
instances_classify -> clusters_cull -> build_setup -> cluster_classify ->
4 triangle_split passes -> triangle_tess_template_instantiate -> CLAS build
-> blas_setup_insertion -> blas_clusters_insert -> BLAS + TLAS build ->
one primary ray per pixel (back faces culled) and one shadow ray per hit.
`value` is displaced triangles per second of whole frames.  Further frames
are run stage by stage with HIP events on the same stream to time every
stage; the slowest stage is the roofline kernel, priced with the
algorithmic bytes of its work.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0

# stages that are one memory-bound kernel, and that kernel
ROOFLINE_KERNELS = {"cluster_classify": "k_cluster_classify", "triangle_tess_template_instantiate": "k_instantiate",
                    "clas_instantiate": "k_clas_instantiate"}


def profiled_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC
    summary profiles/<kernel>_<workload>_traffic.txt (FETCH_SIZE and
    WRITE_SIZE passes scaled by a copy kernel of known size), or None."""
    path = ROOT / "profiles" / ("%s_%s_traffic.txt" % (kernel, workload))
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch="):
            return int(line.split("=", 1)[1])
    return None


def bench_config(abi, transient_1x=False, large=False, animation=False, transient_2x=False):
    from vk_tessellated_clusters_amd import scene
    c = scene.workload_config(large)
    if transient_1x:
        c.flags |= abi.FLAG_1X_TRANSIENT
    if transient_2x:
        c.flags |= abi.FLAG_2X_TRANSIENT
    if animation:
        c.flags |= abi.FLAG_ANIMATION
    return c


def make_scene(name):
    from vk_tessellated_clusters_amd import scene
    if name == "grid64":
        return scene.scene_grid_instances(), "64x grid 256x256 quads (8.4M tris), 2048^2 R16 fBm, 1920x1080 @ 4px"
    if name == "icosphere":
        return scene.scene_icosphere(), "icosphere 1280 tris, 4-oct fBm amp 0.05, 1280x720 @ 8px"
    if name == "batched500":
        return scene.scene_batched(), ("500 meshes log-uniform 200..200000 tris (seed 2, 14.0M tris), grids + "
                                       "icospheres, own 1024^2 R16 fBm each, 1920x1080 @ 4px")
    if name == "large":
        return scene.scene_large(), ("8x8 patches of 512x256 quads (16.8M tris), 4096^2 R16 fBm, 3840x2160 @ 2px, "
                                     "camera on a seeded spline over 100 frames")
    raise SystemExit("unknown workload " + name)


def frame_sequence(host, n=100, animation=False):
    """Frame constants of every frame: the spline path for the large
    workload, else the scene's one camera; with the ripple animation the
    animation clock runs at half the time of 60 Hz frames
    (tessellatedclusters.cpp:556)."""
    from vk_tessellated_clusters_amd import scene
    if not hasattr(host, "spline") and not animation:
        return [host.frame()]
    out = []
    for k in range(n):
        if hasattr(host, "spline"):
            span_x, span_z, seed = host.spline
            host.camera.eye, host.camera.center = scene.camera_spline(k, n, span_x, span_z, seed)
        out.append(host.frame(animation_time=k * 0.5 / 60.0) if animation else host.frame())
    return out


def per_mesh_ms(torch, ds, cfg, fc, stream, reps):
    """Frame time of the per-mesh form: one renderer sized for one mesh,
    re-pointed at each mesh in turn, the whole tessellation + CLAS + BLAS
    chain launched once per mesh (eager launches, no graph)."""
    from vk_tessellated_clusters_amd import abi
    from vk_tessellated_clusters_amd.device import Renderer
    L = abi.lib()
    n = len(ds.instances)
    one = (abi.RenderInstance * 1)(ds.instances[0])

    class One:
        instances, textures, num_textures = one, ds.textures, ds.num_textures

    r1 = Renderer(One, cfg)
    xf = np.zeros(24 * n, np.float32)
    abi.check(L.tc_instance_xforms_host(C.addressof(ds.instances), n, xf.ctypes.data))
    inst_dev = torch.from_numpy(np.frombuffer(bytes(ds.instances), np.uint8).copy()).cuda()
    xf_dev = torch.from_numpy(xf.view(np.uint8)).cuda()
    s = C.c_void_p(stream.cuda_stream)

    def frame():
        for i in range(n):
            abi.check(L.tc_renderer_rebind_instances(r1.handle, C.c_void_p(inst_dev.data_ptr() + 192 * i),
                                                     C.c_void_p(xf_dev.data_ptr() + 96 * i), 1))
            abi.check(L.tc_renderer_frame(r1.handle, C.byref(fc), None, s))

    frame()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        frame()
    e1.record(stream)
    torch.cuda.synchronize()
    r1.close()
    return e0.elapsed_time(e1) / reps


class StageTimer:
    """Records HIP events around stage calls on one stream."""

    def __init__(self, torch, stream):
        self.torch, self.stream, self.marks = torch, stream, []

    def mark(self, name):
        e = self.torch.cuda.Event(enable_timing=True)
        e.record(self.stream)
        self.marks.append((name, e))

    def durations(self):
        self.torch.cuda.synchronize()
        out = {}
        for (n0, e0), (n1, e1) in zip(self.marks[:-1], self.marks[1:]):
            out[n1] = out.get(n1, 0.0) + e0.elapsed_time(e1)
        return out


def staged_frame(r, fc, stream, timer, trace_params=None):
    """One frame through the stage entry points, with an event after each."""
    from vk_tessellated_clusters_amd import abi
    L, b, cfg = r.L, C.byref(r.bindings), r.config
    s = C.c_void_p(stream.cuda_stream)
    timer.mark("start")
    abi.check(L.tc_renderer_begin_frame(r.handle, C.byref(fc), s))
    abi.check(L.tc_instances_classify(b, cfg, s))
    abi.check(L.tc_clusters_cull(b, cfg, s))
    timer.mark("frame_setup_cull")
    abi.check(L.tc_build_setup(b, cfg, abi.BUILD_SETUP_CLASSIFY, s))
    abi.check(L.tc_cluster_classify(b, cfg, s))
    timer.mark("cluster_classify")
    abi.check(L.tc_build_setup(b, cfg, abi.BUILD_SETUP_SPLIT, s))
    passes = L.tc_num_split_passes(cfg.maxSplitFactor)
    for p in range(passes):
        abi.check(L.tc_triangle_split(b, cfg, s))
        if p + 1 < passes:
            abi.check(L.tc_build_setup(b, cfg, abi.BUILD_SETUP_SPLIT_PASS, s))
    timer.mark("triangle_split")
    abi.check(L.tc_build_setup(b, cfg, abi.BUILD_SETUP_INSTANTIATE_TESS, s))
    abi.check(L.tc_triangle_tess_template_instantiate(b, cfg, s))
    timer.mark("triangle_tess_template_instantiate")
    abi.check(L.tc_build_setup(b, cfg, abi.BUILD_SETUP_BUILD_BLAS, s))
    abi.check(L.tc_clas_instantiate(b, cfg, s))
    abi.check(L.tc_clas_build_transient(b, cfg, s))
    timer.mark("clas_instantiate")
    abi.check(L.tc_blas_setup_insertion(b, cfg, s))
    abi.check(L.tc_blas_clusters_insert(b, cfg, 0, s))
    abi.check(L.tc_blas_clusters_insert(b, cfg, 1, s))
    timer.mark("blas_insert")
    accel, xf = r.accel()
    abi.check(L.tc_accel_build(b, cfg, C.byref(accel), C.c_void_p(xf), s))
    timer.mark("blas_tlas_build")
    if trace_params is not None:
        abi.check(L.tc_renderer_trace(r.handle, C.byref(trace_params), s))
        timer.mark("trace_rays")


def algorithmic_bytes(r, host, cfg):
    """Compulsory bytes of each stage of the last frame (reads + writes of
    the records each stage must touch; height map counted once per launch
    up to its size)."""
    from vk_tessellated_clusters_amd import abi
    sb = r.scene_building()
    rb = r.readback()
    geo_of = [host.geometries[i.geometry] for i in host.instances]
    nvis = sb.visibleClusterCounter
    vis = r.copy(sb.visibleClusters, 8 * nvis).view(np.uint32).reshape(-1, 2)
    V = np.array([int(geo_of[i].clusters["numVertices"][c]) for i, c in vis], np.int64)
    T = np.array([int(geo_of[i].clusters["numTriangles"][c]) for i, c in vis], np.int64)
    ntemp = sb.tempInstantiateCounter
    ti = r.copy(sb.tempInstantiations, 32 * ntemp).view(abi.TEMPLATE_INSTANTIATE_INFO)
    full_mask = (ti["clusterIdOffset"] >> 30) == 0
    nparts = sb.partTriangleCounter
    parts = r.copy(sb.partTriangles, 24 * nparts).view(abi.TESS_TRIANGLE_INFO)
    lk = np.zeros(4096, abi.TABLE_ENTRY_DT)
    abi.check(abi.lib().tc_table_lookup_host(lk.ctypes.data))
    cfgs = (parts["triangleID_config"] >> 16) & 0x7FFF
    nv_parts = lk["numVertices"][cfgs].astype(np.int64)
    nt_parts = lk["numTriangles"][cfgs].astype(np.int64)
    sizes = r.copy(sb.tempClusterSizes, 4 * ntemp).view(np.uint32).astype(np.int64)
    tex_bytes = sum(t.nbytes if hasattr(t, "seed") else t[0].nbytes for t in host.textures)
    full_verts = int(sb.genVertexCounter - nv_parts.sum())
    n_split_items = sb.splitWriteCounter
    classify = int((8 + 16 + 4 + 12 * V + 3 * T).sum() + (12 + 8 + 12) * full_verts + 44 * int(full_mask.sum())
                   + 24 * (nparts + min(sb.splitWriteCounter, cfg.maxSplitTriangles)) + min(tex_bytes, 8 * full_verts))
    split = int((24 + 36 + 3 + 16) * n_split_items + 24 * n_split_items * 8)
    instantiate = int((24 + 96 + 3 + 16 + 44) * nparts + 12 * nv_parts.sum() + min(tex_bytes, 8 * int(nv_parts.sum())))
    clas = int((32 + 32 + 4 + 8) * ntemp + sizes.sum() + 12 * sb.genVertexCounter + 4 * (nt_parts.sum()))
    insert = int(ntemp * (4 + 8 + 4 + 8) + 2 * 12 * ntemp * 2)
    # keys from CLAS headers (48 B), 5 radix passes of (8 B key + 4 B value)
    # read and written, hierarchy + refit writing (N - 1) binary nodes of
    # 64 B + 24 B boxes and the 128 B four-wide records
    accel = int(ntemp * (48 + 12) + 5 * 2 * 12 * ntemp + ntemp * (64 + 24 + 128 + 16 + 8))
    return {"cluster_classify": classify, "triangle_split": split, "triangle_tess_template_instantiate": instantiate,
            "clas_instantiate": clas, "blas_insert": insert, "blas_tlas_build": accel}, rb, sb


def cpu_baseline(host, cfg, sample_instances, budget_s):
    """The oracle (C++ port of the reference stages) on a bounded sample of
    the workload: the first `sample_instances` instances, all host threads."""
    sys.path.insert(0, str(ROOT / "tests"))
    import harness
    from vk_tessellated_clusters_amd.scene import HostScene
    sub = HostScene(host.geometries, host.instances[:sample_instances], host.textures, host.camera,
                    host.tess_rate_pixels, host.name)
    threads = min(16, os.cpu_count() or 1)
    of = harness.OracleFrame(sub, cfg)
    of.run(sub.frame(), nthreads=threads, stages="tess")  # warm-up
    times, tris = [], 0
    t_end = time.time() + budget_s
    while time.time() < t_end or len(times) < 3:
        t0 = time.perf_counter()
        of.run(sub.frame(), nthreads=threads, stages="all")
        times.append(time.perf_counter() - t0)
        tris = of.rb.numTotalTriangles
        if len(times) >= 20:
            break
    med = float(np.median(times))
    return {"value": tris / med, "unit": "displaced tris/s", "cores": threads, "kind": "port",
            "sample": "first %d of %d instances, whole frame (classify..insert), median of %d frames, %d tris/frame"
                      % (sample_instances, len(host.instances), len(times), tris)}


def bench_diff(args):
    """Differentiable form: one step = depth render of the displaced
    64x64-quad grid (tessellate, CLAS / BLAS / TLAS, 512^2 orthographic
    rays), L2 loss, tc_depth_adjoint backward and an Adam update of the
    512^2 FP32 height map."""
    import torch
    from vk_tessellated_clusters_amd import diff
    target = diff.target_depth(512, 3)
    heights = torch.zeros(512, 512, dtype=torch.float32, device="cuda", requires_grad=True)
    r = diff.DepthRenderer(heights)
    opt = torch.optim.Adam([heights], lr=0.02)
    stream = torch.cuda.current_stream()

    def step():
        opt.zero_grad(set_to_none=True)
        loss = diff.depth_loss(diff.displaced_depth(heights, r), target)
        loss.backward()
        opt.step()
        return loss

    for _ in range(args.warmup):
        l0 = step()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(args.steps):
        loss = step()
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    # forward alone and adjoint alone
    f0, f1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    f0.record(stream)
    for _ in range(args.steps):
        t = r.render()
    f1.record(stream)
    g = torch.randn_like(t) * 1e-6
    a0, a1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a0.record(stream)
    for _ in range(args.steps):
        r.adjoint(g)
    a1.record(stream)
    torch.cuda.synchronize()
    rb = r.readback()
    out = {
        "metric": "displaced tris/s, cluster-BVH build ms/frame, Mrays/s, HBM GB/s (1 GPU)",
        "value": 1000.0 / ms, "unit": "adam steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "differentiable depth: 64x64-quad grid, learnable 512^2 R32 height map, 512^2 "
                               "orthographic depth rays, L2 vs seed-3 fBm depth, Adam",
                   "displaced_tris_per_frame": int(rb.numTotalTriangles), "gen_vertices": int(rb.numGenVertices),
                   "rays": r.num_rays},
        "forward_ms": f0.elapsed_time(f1) / args.steps,
        "adjoint_ms": a0.elapsed_time(a1) / args.steps,
        "loss_first_last": [float(l0), float(loss)],
    }
    print(json.dumps(out))
    r.close()
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="grid64")
    ap.add_argument("--cpu-sample-instances", type=int, default=8)
    ap.add_argument("--cpu-budget-s", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", dest="graph", action="store_false")
    ap.add_argument("--per-mesh-reps", type=int, default=2)
    ap.add_argument("--autotune", dest="autotune", action="store_true", default=None,
                    help="pick launch shapes by measurement first (default for batched500)")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false")
    ap.add_argument("--transient-1x", action="store_true",
                    help="1X transient cluster builds (the reference renderer's default, renderer.hpp:47)")
    ap.add_argument("--transient-2x", action="store_true",
                    help="2X transient mini batches (the reference renderer's default, renderer.hpp:48)")
    ap.add_argument("--animation", action="store_true",
                    help="ripple animation of the generated vertices, clock advancing every frame")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="after timing, copy a buffer of this size once (scale reference for PMC byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the frame shards by instance; "
              "see the sharding notes); run with --gpus 1", file=sys.stderr)
        return 2

    if args.workload == "diff":
        return bench_diff(args)

    import torch
    from vk_tessellated_clusters_amd import abi
    from vk_tessellated_clusters_amd.device import DeviceScene, Renderer

    host, desc = make_scene(args.workload)
    cfg = bench_config(abi, args.transient_1x, args.workload == "large", args.animation, args.transient_2x)
    ds = DeviceScene(host)
    r = Renderer(ds, cfg)
    frames = frame_sequence(host, animation=args.animation)
    fc = frames[0]
    # a created stream (default priority): graphs cannot be captured on the
    # null stream
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)

    W, H = host.camera.width, host.camera.height
    hits = torch.empty(W * H * abi.HIT_DT.itemsize, dtype=torch.uint8, device="cuda")
    dec = torch.empty(W * H * abi.HIT_DECODED_DT.itemsize, dtype=torch.uint8, device="cuda")
    tp = abi.TraceParams()
    tp.width, tp.height, tp.flags = W, H, abi.TRACE_CULL_BACK | abi.TRACE_SHADOW
    tp.hits, tp.decoded = hits.data_ptr(), dec.data_ptr()

    def frame(k=0):
        r.tessellate(frames[k % len(frames)], stream)
        r.build_accel(stream)
        r.trace(tp, stream)

    for _ in range(args.warmup):
        frame()
    torch.cuda.synchronize()
    from vk_tessellated_clusters_amd import autotune
    autotuned = args.autotune if args.autotune is not None else args.workload == "batched500"
    if autotuned:
        autotune.autotune(r, fc, tp, stream, reps=3, log=lambda m: print(m, file=sys.stderr, flush=True))
    if args.graph:
        abi.check(r.L.tc_renderer_graph_capture(r.handle, C.byref(fc), C.byref(tp), C.c_void_p(stream.cuda_stream)),
                  "graph capture")

        def frame(k=0):  # noqa: F811
            abi.check(r.L.tc_renderer_graph_launch(r.handle, C.byref(frames[k % len(frames)]),
                                                   C.c_void_p(stream.cuda_stream)), "graph launch")

        frame()
        torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for k in range(args.steps):
        frame(k)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    # displaced triangles of every timed frame (the frames are deterministic;
    # counted in an untimed pass when the camera moves)
    if len(frames) > 1:
        frame_tris = []
        for k in range(args.steps):
            r.tessellate(frames[k % len(frames)], stream)
            frame_tris.append(int(r.readback().numTotalTriangles))
        mean_tris = sum(frame_tris) / len(frame_tris)
    else:
        frame_tris = None

    # stage-by-stage frames for the stage times and the roofline kernel
    stage_ms = {}
    nstage = max(3, min(args.steps, 10))
    for _ in range(nstage):
        t = StageTimer(torch, stream)
        staged_frame(r, fc, stream, t, tp)
        for k, v in t.durations().items():
            stage_ms[k] = stage_ms.get(k, 0.0) + v / nstage
    alg, rb, sb = algorithmic_bytes(r, host, cfg)
    tris = rb.numTotalTriangles if frame_tris is None else mean_tris
    torch.cuda.synchronize()
    dec_np = dec.cpu().numpy().view(abi.HIT_DECODED_DT)
    n_primary = W * H
    n_shadow = int((dec_np["instanceID"] != 0xFFFFFFFF).sum())
    bvh_ms = stage_ms.get("clas_instantiate", 0) + stage_ms.get("blas_insert", 0) + stage_ms.get("blas_tlas_build", 0)
    trace_ms = stage_ms.get("trace_rays", 0.0)
    # roofline kernel: the slowest stage that is one HBM-bound kernel
    # (traversal is latency-bound and reported as Mrays/s)
    dom = max(ROOFLINE_KERNELS, key=lambda k: stage_ms.get(k, 0.0))
    dom_ms = stage_ms[dom]
    achieved = alg[dom] / (dom_ms * 1e-3) / 1e9
    frame_bytes = sum(alg.values())
    traffic = profiled_traffic(ROOFLINE_KERNELS[dom], args.workload)

    out = {
        "metric": "displaced tris/s, cluster-BVH build ms/frame, Mrays/s, HBM GB/s (1 GPU)",
        "value": tris / (ms * 1e-3),
        "unit": "displaced tris/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": desc + (", 1X transient builds" if args.transient_1x else "") +
                   (", 2X transient batches" if args.transient_2x else "") +
                   (", ripple animation" if args.animation else ""),
                   "displaced_tris_per_frame": int(tris), "base_tris": host.num_base_triangles,
                   "frames_on_path": len(frames),
                   "clusters": host.num_clusters, "full_clusters": int(rb.numFullClusters),
                   "part_triangles": int(rb.numPartTriangles), "split_triangles": int(rb.numSplitTriangles),
                   "clas_per_frame": int(sb.tempInstantiateCounter + (sb.transBuildCounter if (args.transient_1x or args.transient_2x) else 0)),
                   "transient_clas": int(sb.transBuildCounter if (args.transient_1x or args.transient_2x) else 0),
                   "gen_vertices": int(sb.genVertexCounter)},
        "capacity_ok": bool(rb.numVisibleClusters <= cfg.maxVisibleClusters and
                            rb.numPartTriangles <= cfg.maxPartTriangles and
                            rb.numSplitTriangles <= cfg.maxSplitTriangles and
                            rb.numGenVertices <= cfg.maxGeneratedVertices and
                            rb.numBlasClusters <= cfg.maxGeneratedClusters and
                            rb.numGenDatas <= cfg.maxGeneratedClusterMegs * (1 << 20)),
        "stage_ms": {k: round(v, 4) for k, v in stage_ms.items()},
        "tuning": dict(autotune.get_tuning(), autotuned=bool(autotuned)),
        "cluster_bvh_build_ms_per_frame": bvh_ms,
        "mrays_per_s": (n_primary + n_shadow) / (trace_ms * 1e-3) / 1e6 if trace_ms else None,
        "rays_per_frame": {"primary": n_primary, "shadow": n_shadow},
        "frame_hbm_gbs": frame_bytes / (ms * 1e-3) / 1e9,
        "roofline": {"bound": "hbm", "kernel": ROOFLINE_KERNELS[dom], "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "alg_bytes": alg[dom],
                     "launch_ms": dom_ms},
    }
    if args.workload == "batched500":
        # tessellation + CLAS + BLAS + TLAS of all meshes: one graph replay vs
        # per-mesh launches (rays excluded from both)
        abi.check(r.L.tc_renderer_graph_capture(r.handle, C.byref(fc), None, C.c_void_p(stream.cuda_stream)))
        for _ in range(2):
            abi.check(r.L.tc_renderer_graph_launch(r.handle, None, C.c_void_p(stream.cuda_stream)))
        torch.cuda.synchronize()
        b0, b1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        b0.record(stream)
        for _ in range(args.steps):
            abi.check(r.L.tc_renderer_graph_launch(r.handle, None, C.c_void_p(stream.cuda_stream)))
        b1.record(stream)
        torch.cuda.synchronize()
        batched = b0.elapsed_time(b1) / args.steps
        pm = per_mesh_ms(torch, ds, cfg, fc, stream, args.per_mesh_reps)
        out["batched_vs_per_mesh"] = {"batched_graph_ms": batched, "per_mesh_ms": pm, "speedup": pm / batched}
    if args.calibrate_copy_mib:
        src = torch.ones(args.calibrate_copy_mib << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        out["calibration_copy_bytes"] = src.numel() * 4
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(host, cfg, args.cpu_sample_instances, args.cpu_budget_s)
    print(json.dumps(out))
    r.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
