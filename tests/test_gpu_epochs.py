"""Look-back epochs across their wrap.

Stage tiles and radix-sort tiles publish status words tagged with a 20-bit
epoch and never clear them.  When an epoch counter would pass 2^20 - 1, the
words are cleared and the counter restarts at 1: the stage prologue
(csrc/tc_pipeline.hip) and the sort's upsweep (csrc/tc_sort.hip) do this.
A frame runs about seven stage prologues and five sort passes, so a
renderer crosses both wraps every few minutes at a thousand frames a
second.  This test starts both counters just below the wrap and requires
frames identical to a fresh renderer's.
"""
import ctypes as C

import numpy as np
import pytest

import harness
from test_gpu_parity import small_grid
from vk_tessellated_clusters_amd import abi

pytestmark = pytest.mark.gpu

# Workspace offsets of the two counters (ws_layout in csrc/tc_abi.hip): the
# first word of the stage header, and the first word of the sort state,
# which follows the 1024-byte stage header.
STAGE_EPOCH = 0
SORT_EPOCH = 1024
EPOCH_LIMIT = 1 << 20

TRACE = abi.TRACE_CULL_BACK | abi.TRACE_SHADOW


def workspace_word(g, offset, value=None):
    """Reads (or, given a value, writes) one u32 of the renderer's workspace."""
    addr = int(g.r.bindings.workspace) + offset
    w = np.zeros(1, np.uint32)
    L = abi.lib()
    if value is None:
        abi.check(L.tc_memcpy(w.ctypes.data, C.c_void_p(addr), 4), "tc_memcpy")
        return int(w[0])
    w[0] = value
    abi.check(L.tc_memcpy(C.c_void_p(addr), w.ctypes.data, 4), "tc_memcpy")


def frame(g, host, frames):
    for _ in range(frames):
        g.run(host.frame(), stages="accel")
    res = g.results()
    arr = g.accel_arrays()
    hits, dec = g.trace(96, 54, TRACE)
    return res, arr, hits, dec


def written(a):
    """BLAS / TLAS node records the build writes (n - 1 internal nodes per
    instance) and the leaves as offsets into the CLAS pool."""
    sb, n = a["sb"], len(a["infos"])
    keep = np.zeros(len(a["blasNodes"]), bool)
    for info in a["infos"]:
        first = int(info["clusterReferences"] - sb.blasClusterAddresses) // 8
        keep[first:first + max(0, int(info["clusterReferencesCount"]) - 1)] = True
    return (a["blasNodes"][keep].view(np.uint8), a["blasNodes4"][keep].view(np.uint8),
            a["tlasNodes"][:n - 1].view(np.uint8), a["tlasNodes4"][:n - 1].view(np.uint8),
            a["blasLeaves"] - np.uint64(sb.genClusterData))


def test_epoch_wraps_keep_frames_identical():
    host = small_grid()
    cfg = harness.small_config()
    g = harness.GpuFrame(host, cfg)
    try:
        r0, a0, h0, d0 = frame(g, host, 1)
        stage0, sort0 = workspace_word(g, STAGE_EPOCH), workspace_word(g, SORT_EPOCH)
    finally:
        g.close()
    assert 0 < stage0 < 1000 and 0 < sort0 < 1000, (stage0, sort0)

    g = harness.GpuFrame(host, cfg)
    try:
        # the first frame's prologues cross the stage wrap; the second
        # frame's first sort crosses the sort wrap (a frame takes five
        # sort epochs)
        workspace_word(g, STAGE_EPOCH, EPOCH_LIMIT - 3)
        workspace_word(g, SORT_EPOCH, EPOCH_LIMIT - 6)
        r1, a1, h1, d1 = frame(g, host, 2)
        stage1, sort1 = workspace_word(g, STAGE_EPOCH), workspace_word(g, SORT_EPOCH)
    finally:
        g.close()
    # both counters restarted: the second run's epochs are small again
    assert 0 < stage1 < 1000 and 0 < sort1 < 1000, (stage1, sort1)

    harness.compare(r1, r0, max_ulp=0)
    for x, y in zip(written(a1), written(a0)):
        assert np.array_equal(x, y)
    assert np.array_equal(h1.view(np.uint8), h0.view(np.uint8))
    assert np.array_equal(d1.view(np.uint8), d0.view(np.uint8))
