"""ctypes access to the FP64 depth-adjoint checker in oracle/liboracle.so
(test infrastructure)."""
import ctypes as C

import numpy as np

from harness import oracle
from vk_tessellated_clusters_amd import abi

P = C.c_void_p


class HeightMap(C.Structure):
    _fields_ = [("texels", P), ("width", C.c_uint32), ("height", C.c_uint32)]


_ready = False


def _lib():
    global _ready
    L = oracle()
    if not _ready:
        common = [P, P, P, C.c_uint32, P, P, C.c_uint32, P]
        L.oracle_depth_t64.argtypes = common + [C.c_uint32, C.c_uint64, C.c_double, P]
        L.oracle_depth_t64.restype = None
        L.oracle_depth_adjoint64.argtypes = common + [P, P]
        L.oracle_depth_adjoint64.restype = None
        L.oracle_depth_loss_fd64.argtypes = common + [P, C.c_double, C.c_uint32, P, C.c_uint32, C.c_double, P]
        L.oracle_depth_loss_fd64.restype = None
        _ready = True
    return L


class Problem:
    """Frozen forward state of a depth frame: rays (RAY_DT), hits (HIT_DT),
    hitVertices (u32 x4 per ray), tape (VERTEX_TAPE_DT), height maps
    (float32 2-D arrays, None: not differentiated) and instance xforms (24
    floats per instance)."""

    def __init__(self, rays, hits, hit_vertices, tape, maps, xforms):
        self.rays = np.ascontiguousarray(rays, abi.RAY_DT)
        self.hits = np.ascontiguousarray(hits, abi.HIT_DT)
        self.hv = np.ascontiguousarray(hit_vertices, np.uint32)
        self.tape = np.ascontiguousarray(tape, abi.VERTEX_TAPE_DT)
        self.maps = [None if m is None else np.ascontiguousarray(m, np.float32) for m in maps]
        self.xforms = np.ascontiguousarray(xforms, np.float32)
        self.cmaps = (HeightMap * max(1, len(self.maps)))()
        for i, m in enumerate(self.maps):
            if m is not None:
                self.cmaps[i].texels = m.ctypes.data
                self.cmaps[i].height, self.cmaps[i].width = m.shape
        self.n = len(self.rays)

    def _common(self):
        return (self.rays.ctypes.data, self.hits.ctypes.data, self.hv.ctypes.data, self.n, self.tape.ctypes.data,
                C.addressof(self.cmaps), len(self.maps), self.xforms.ctypes.data)

    def t64(self, perturb_map=0xFFFFFFFF, texel=0, delta=0.0):
        out = np.empty(self.n, np.float64)
        _lib().oracle_depth_t64(*self._common(), perturb_map, texel, delta, out.ctypes.data)
        return out

    def adjoint64(self, grad_t):
        g = np.ascontiguousarray(grad_t, np.float64)
        grads = [None if m is None else np.zeros(m.shape, np.float64) for m in self.maps]
        ptrs = (C.c_void_p * max(1, len(grads)))(*[0 if x is None else x.ctypes.data for x in grads])
        _lib().oracle_depth_adjoint64(*self._common(), g.ctypes.data, ptrs)
        return grads

    def loss_fd64(self, target, scale, map_index, texels, eps):
        tg = np.ascontiguousarray(target, np.float32)
        tx = np.ascontiguousarray(texels, np.uint64)
        out = np.empty(len(tx), np.float64)
        _lib().oracle_depth_loss_fd64(*self._common(), tg.ctypes.data, scale, map_index, tx.ctypes.data, len(tx), eps,
                                      out.ctypes.data)
        return out


def identity_xforms(n=1):
    m = np.zeros((n, 24), np.float32)
    for k in (0, 12):
        m[:, k + 0] = m[:, k + 5] = m[:, k + 10] = 1.0
    return m.reshape(-1)
