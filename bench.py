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
    """Frame time of the per-mesh form: a renderer for one mesh runs the
    tessellation + CLAS + BLAS + TLAS chain once per mesh, all 500 chains in
    one graph (tc_renderer_graph_capture_meshwise), so no host launch cost
    is timed.  Its capacities, which set its grid sizes, are the largest
    single mesh's needs, measured by running every mesh once in a renderer
    with the batched frame's capacities.  Returns (ms, displaced triangles
    summed over the meshes, capacities used)."""
    from vk_tessellated_clusters_amd import abi
    from vk_tessellated_clusters_amd.device import Renderer
    L = abi.lib()
    n = len(ds.instances)
    one = (abi.RenderInstance * 1)(ds.instances[0])

    class One:
        instances, textures, num_textures = one, ds.textures, ds.num_textures

    xf = np.zeros(24 * n, np.float32)
    abi.check(L.tc_instance_xforms_host(C.addressof(ds.instances), n, xf.ctypes.data))
    inst_dev = torch.from_numpy(np.frombuffer(bytes(ds.instances), np.uint8).copy()).cuda()
    xf_dev = torch.from_numpy(xf.view(np.uint8)).cuda()
    s = C.c_void_p(stream.cuda_stream)
    # each mesh's needs, in a renderer with the batched capacities
    need = {k: 0 for k in ("vis", "part", "split", "vert", "clas", "data")}
    tris = 0
    big = Renderer(One, cfg)
    for i in range(n):
        abi.check(L.tc_renderer_rebind_instances(big.handle, C.c_void_p(inst_dev.data_ptr() + 192 * i),
                                                 C.c_void_p(xf_dev.data_ptr() + 96 * i), 1))
        abi.check(L.tc_renderer_frame(big.handle, C.byref(fc), None, s))
        torch.cuda.synchronize()
        rb = big.readback()
        tris += rb.numTotalTriangles
        for k, v in (("vis", rb.numVisibleClusters), ("part", rb.numPartTriangles), ("split", rb.numSplitTriangles),
                     ("vert", rb.numGenVertices), ("clas", rb.numBlasClusters), ("data", rb.numGenDatas)):
            need[k] = max(need[k], int(v))
    big.close()

    def up(v, q):
        return max(q, (int(v * 1.25) + q - 1) // q * q)

    c1 = abi.Config.from_buffer_copy(bytes(cfg))
    c1.maxVisibleClusters = up(need["vis"], 64)
    c1.maxPartTriangles = up(need["part"], 256)
    c1.maxSplitTriangles = up(need["split"], 64)
    c1.maxGeneratedVertices = up(need["vert"], 1024)
    c1.maxGeneratedClusters = up(need["clas"], 256)
    c1.maxGeneratedClusterMegs = up(need["data"] / (1 << 20), 1)
    r1 = Renderer(One, c1)
    abi.check(L.tc_renderer_graph_capture_meshwise(r1.handle, C.c_void_p(inst_dev.data_ptr()),
                                                   C.c_void_p(xf_dev.data_ptr()), n, C.byref(fc), s), "meshwise capture")
    for _ in range(2):
        abi.check(L.tc_renderer_graph_launch(r1.handle, None, s))
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        abi.check(L.tc_renderer_graph_launch(r1.handle, None, s))
    e1.record(stream)
    torch.cuda.synchronize()
    r1.close()
    caps = {"maxVisibleClusters": c1.maxVisibleClusters, "maxPartTriangles": c1.maxPartTriangles,
            "maxSplitTriangles": c1.maxSplitTriangles, "maxGeneratedVertices": c1.maxGeneratedVertices,
            "maxGeneratedClusters": c1.maxGeneratedClusters, "maxGeneratedClusterMegs": c1.maxGeneratedClusterMegs}
    return e0.elapsed_time(e1) / reps, tris, caps


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
    the workload: the first `sample_instances` instances, whole frames from
    instance culling through BLAS insertion, timed on every host thread and
    on one thread."""
    sys.path.insert(0, str(ROOT / "tests"))
    import harness
    from vk_tessellated_clusters_amd.scene import HostScene
    sub = HostScene(host.geometries, host.instances[:sample_instances], host.textures, host.camera,
                    host.tess_rate_pixels, host.name)
    cores = os.cpu_count() or 1
    of = harness.OracleFrame(sub, cfg)

    def rate(threads, budget, max_frames):
        of.run(sub.frame(), nthreads=threads, stages="tess")  # warm-up
        times = []
        t_end = time.time() + budget
        while time.time() < t_end or len(times) < 3:
            t0 = time.perf_counter()
            of.run(sub.frame(), nthreads=threads, stages="all")
            times.append(time.perf_counter() - t0)
            if len(times) >= max_frames:
                break
        return of.rb.numTotalTriangles / float(np.median(times)), len(times)

    multi, nm = rate(cores, budget_s * 0.4, 20)
    single, ns = rate(1, budget_s * 0.6, 7)
    tris = of.rb.numTotalTriangles
    return {"value": multi, "unit": "displaced tris/s", "cores": cores, "kind": "port",
            "value_1thread": single, "threads": {"all": cores, "one": 1},
            "sample": "first %d of %d instances, whole frame (cull..insert), %d tris/frame; median of %d frames on "
                      "%d threads (os.cpu_count()) and of %d frames on 1 thread"
                      % (sample_instances, len(host.instances), tris, nm, cores, ns)}


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


def profiled_fields(kernel, workload):
    """key=value lines of profiles/<kernel>_<workload>_traffic.txt."""
    path = ROOT / "profiles" / ("%s_%s_traffic.txt" % (kernel, workload))
    out = {}
    if path.exists():
        for line in path.read_text().splitlines():
            if "=" in line and not line.startswith("#"):
                k, v = line.split("=", 1)
                try:
                    out[k.strip()] = float(v)
                except ValueError:
                    pass
    return out


def measure(args, torch, workload, flags=0, steps=None, warmup=None, cpu=False, per_mesh=False, autotuned=None):
    """Times `steps` whole frames of `workload` (graph replays), then stage
    by stage; returns the result dict of the bench line."""
    from vk_tessellated_clusters_amd import abi, autotune
    from vk_tessellated_clusters_amd.device import DeviceScene, Renderer

    steps = args.steps if steps is None else steps
    warmup = args.warmup if warmup is None else warmup
    host, desc = make_scene(workload)
    cfg = bench_config(abi, args.transient_1x, workload == "large", args.animation, args.transient_2x)
    cfg.flags |= flags
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

    for _ in range(warmup):
        frame()
    torch.cuda.synchronize()
    if autotuned is None:
        autotuned = args.autotune if args.autotune is not None else workload == "batched500"
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
    for k in range(steps):
        frame(k)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    # displaced triangles of every timed frame (the frames are deterministic;
    # counted in an untimed pass when the camera moves)
    if len(frames) > 1:
        frame_tris = []
        for k in range(steps):
            r.tessellate(frames[k % len(frames)], stream)
            frame_tris.append(int(r.readback().numTotalTriangles))
        mean_tris = sum(frame_tris) / len(frame_tris)
    else:
        frame_tris = None

    # stage-by-stage frames for the stage times and the roofline kernel
    stage_ms = {}
    nstage = max(3, min(steps, 10))
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
    traffic = profiled_traffic(ROOFLINE_KERNELS[dom], workload)
    transient = bool(cfg.flags & abi.TRANSIENT_FLAGS)
    flag_names = [n for n, f in (("PN", abi.FLAG_PN), ("1X transient builds", abi.FLAG_1X_TRANSIENT),
                                 ("2X transient batches", abi.FLAG_2X_TRANSIENT), ("ripple animation", abi.FLAG_ANIMATION))
                  if cfg.flags & f]
    trace_prof = profiled_fields("k_trace", workload)
    out = {
        "metric": "displaced tris/s, cluster-BVH build ms/frame, Mrays/s, HBM GB/s (1 GPU)",
        "value": tris / (ms * 1e-3),
        "unit": "displaced tris/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": desc + "".join(", " + f for f in flag_names),
                   "displaced_tris_per_frame": int(tris), "base_tris": host.num_base_triangles,
                   "frames_on_path": len(frames),
                   "clusters": host.num_clusters, "full_clusters": int(rb.numFullClusters),
                   "part_triangles": int(rb.numPartTriangles), "split_triangles": int(rb.numSplitTriangles),
                   "clas_per_frame": int(sb.tempInstantiateCounter + (sb.transBuildCounter if transient else 0)),
                   "transient_clas": int(sb.transBuildCounter if transient else 0),
                   "gen_vertices": int(sb.genVertexCounter)},
        "frame_ok": bool(rb.numVisibleClusters <= cfg.maxVisibleClusters and
                         rb.numPartTriangles <= cfg.maxPartTriangles and
                         rb.numSplitTriangles <= cfg.maxSplitTriangles and
                         rb.numGenVertices <= cfg.maxGeneratedVertices and
                         rb.numBlasClusters <= cfg.maxGeneratedClusters and
                         rb.numGenDatas <= cfg.maxGeneratedClusterMegs * (1 << 20) and rb.errorFlags == 0),
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
    if trace_prof:
        # traversal: HBM bytes per launch from the committed PMC summary over
        # this run's trace time, with the occupancy that summary measured
        tb = trace_prof.get("traffic_bytes_per_launch")
        out["trace"] = {"ms": trace_ms, "hbm_gbs": tb / (trace_ms * 1e-3) / 1e9 if tb and trace_ms else None,
                        "traffic_bytes_per_launch": tb, "l2_hit_pct": trace_prof.get("l2_hit_pct"),
                        "waves_per_simd": trace_prof.get("mean_waves_per_simd"),
                        "occupancy_pct": trace_prof.get("occupancy_pct")}
    if per_mesh:
        # tessellation + CLAS + BLAS + TLAS of all meshes: one graph replay vs
        # the per-mesh form (rays excluded from both)
        abi.check(r.L.tc_renderer_graph_capture(r.handle, C.byref(fc), None, C.c_void_p(stream.cuda_stream)))
        for _ in range(2):
            abi.check(r.L.tc_renderer_graph_launch(r.handle, None, C.c_void_p(stream.cuda_stream)))
        torch.cuda.synchronize()
        b0, b1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        b0.record(stream)
        for _ in range(steps):
            abi.check(r.L.tc_renderer_graph_launch(r.handle, None, C.c_void_p(stream.cuda_stream)))
        b1.record(stream)
        torch.cuda.synchronize()
        batched = b0.elapsed_time(b1) / steps
        batched_tris = int(r.readback().numTotalTriangles)
        pm, pm_tris, caps = per_mesh_ms(torch, ds, cfg, fc, stream, args.per_mesh_reps)
        out["batched_vs_per_mesh"] = {"batched_graph_ms": batched, "per_mesh_graph_ms": pm, "speedup": pm / batched,
                                      "same_triangles": pm_tris == batched_tris, "displaced_tris": batched_tris,
                                      "per_mesh_capacities": caps}
    if args.calibrate_copy_mib:
        src = torch.ones(args.calibrate_copy_mib << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        out["calibration_copy_bytes"] = src.numel() * 4
    if cpu:
        out["cpu_baseline"] = cpu_baseline(host, cfg, args.cpu_sample_instances, args.cpu_budget_s)
    r.close()
    del ds
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    return out


def secondary(res):
    """The few figures of a secondary workload carried on the bench line."""
    keep = {"ms_per_step", "value", "config", "frame_ok", "stage_ms", "cluster_bvh_build_ms_per_frame",
            "mrays_per_s", "rays_per_frame", "batched_vs_per_mesh", "steps"}
    out = {k: v for k, v in res.items() if k in keep}
    out["displaced_tris_per_s"] = out.pop("value")
    return out


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
    ap.add_argument("--no-secondary", dest="secondary", action="store_false",
                    help="skip the secondary workloads the default grid64 line carries")
    ap.add_argument("--per-mesh-reps", type=int, default=3)
    ap.add_argument("--autotune", dest="autotune", action="store_true", default=None,
                    help="pick launch shapes by measurement first (default for batched500)")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false")
    ap.add_argument("--transient-1x", action="store_true",
                    help="1X transient cluster builds (the reference renderer's default, renderer.hpp:47)")
    ap.add_argument("--transient-2x", action="store_true",
                    help="2X transient mini batches (the reference renderer's default, renderer.hpp:48)")
    ap.add_argument("--reference-defaults", action="store_true",
                    help="the reference renderer's default flags: PN, 1X and 2X transient builds (renderer.hpp:43-48)")
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
    from vk_tessellated_clusters_amd import abi, scene

    flags = scene.REFERENCE_DEFAULT_FLAGS if args.reference_defaults else 0
    out = measure(args, torch, args.workload, flags=flags, cpu=not args.no_cpu,
                  per_mesh=args.workload == "batched500")
    if args.workload == "grid64" and args.secondary and not (args.transient_1x or args.transient_2x or args.animation
                                                             or args.reference_defaults):
        # the same frame with the reference renderer's default flags, the
        # 500-mesh batched frame against its per-mesh form, and the 4K
        # spline sequence, each as a few timed frames
        sec = {}
        for name, wl, fl, kw in (("reference_default_flags", "grid64", scene.REFERENCE_DEFAULT_FLAGS, {}),
                                 ("batched500", "batched500", 0, {"per_mesh": True, "autotuned": False}),
                                 ("large", "large", 0, {})):
            print("secondary workload:", name, file=sys.stderr, flush=True)
            sec[name] = secondary(measure(args, torch, wl, flags=fl, steps=min(args.steps, 10), warmup=2, **kw))
        out["secondary"] = sec
    print(json.dumps(out))
    return 0

if __name__ == "__main__":
    sys.exit(main())
