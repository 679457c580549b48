"""Launch-shape autotuner for the frame.

The work-queue kernels produce identical results for every launch shape
(tiles are taken by ticket and ordered by look-back), so the shapes in
``tc_tuning`` are a pure speed choice.  This tuner measures whole frames as
graph replays on the caller's workload and keeps, parameter by parameter
(coordinate descent), the fastest value.  No reference counterpart: the
reference's dispatch sizes are fixed by its shaders.
"""
import ctypes as C

from . import abi

CANDIDATES = {
    "classifyBlocksPerCU": [2, 3, 4, 5, 8],
    "splitBlocksPerCU": [2, 4, 8, 16],
    "instantiateBlocksPerCU": [2, 3, 4, 6, 8],
    "clasWavesPerCU": [16, 32, 64, 128],
    "traceThreads": [64, 128, 256],
    "traceBlocksPerCU": [32, 64, 128, 256, 512],
    "bvhBlocksPerCU": [4, 8, 16],
}


def get_tuning() -> dict:
    t = abi.Tuning()
    abi.check(abi.lib().tc_get_tuning(C.byref(t)), "tc_get_tuning")
    return t.as_dict()


def set_tuning(values: dict = None):
    L = abi.lib()
    if values is None:
        abi.check(L.tc_set_tuning(None), "tc_set_tuning")
        return
    t = abi.Tuning(**values)
    abi.check(L.tc_set_tuning(C.byref(t)), "tc_set_tuning")


def frame_ms(renderer, fc, trace_params, stream, reps: int = 5) -> float:
    """Mean time of `reps` graph replays of the whole frame (captured with the
    current tuning)."""
    import torch
    L = renderer.L
    s = C.c_void_p(stream.cuda_stream)
    tp = C.byref(trace_params) if trace_params is not None else None
    abi.check(L.tc_renderer_graph_capture(renderer.handle, C.byref(fc), tp, s), "graph capture")
    abi.check(L.tc_renderer_graph_launch(renderer.handle, None, s), "graph launch")
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        abi.check(L.tc_renderer_graph_launch(renderer.handle, None, s), "graph launch")
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def autotune(renderer, fc, trace_params, stream, reps: int = 5, rounds: int = 2, candidates: dict = None,
             log=None):
    """Coordinate descent over `candidates` (parameter -> values); leaves the
    best tuning set and returns (tuning dict, frame ms, history)."""
    candidates = candidates or CANDIDATES
    best = get_tuning()
    best_ms = frame_ms(renderer, fc, trace_params, stream, reps)
    history = [(dict(best), best_ms)]
    for _ in range(rounds):
        improved = False
        for name, values in candidates.items():
            for v in values:
                if v == best[name]:
                    continue
                trial = dict(best, **{name: v})
                set_tuning(trial)
                ms = frame_ms(renderer, fc, trace_params, stream, reps)
                history.append((trial, ms))
                if log:
                    log("autotune %s=%d: %.3f ms (best %.3f)" % (name, v, ms, best_ms))
                if ms < best_ms * 0.99:
                    best, best_ms, improved = trial, ms, True
            set_tuning(best)
        if not improved:
            break
    set_tuning(best)
    return best, best_ms, history
