"""Differentiable depth of a displaced, tessellated surface.

A height map displaces a tessellated grid; orthographic depth rays are
traced through the cluster BVH (the forward is the ordinary frame:
tessellate, CLAS / BLAS / TLAS build, trace), and the backward pass is the
hand-written adjoint ``tc_depth_adjoint``: d(hit distance)/d(texel) through
bilinear sampling, vertex generation and ray / triangle intersection, with
the tessellation pattern and the hit triangles of the forward pass frozen.
PyTorch provides the device memory, the stream and the autograd plumbing.
No reference counterpart: the reference renderer is not differentiable.
"""
import ctypes as C

import numpy as np

from . import abi
from .device import DeviceScene, Renderer, _torch
from .scene import Camera, HostScene, Instance, fbm_height, grid_patch


def depth_config() -> abi.Config:
    """Capacities for a 64x64-quad grid at up to 11 segments per edge, with
    the vertex tape on."""
    c = abi.default_config()
    c.flags = abi.FLAG_DISPLACEMENT | abi.FLAG_CULLING | abi.FLAG_VERTEX_TAPE
    c.maxVisibleClusters = 1 << 12
    c.maxPartTriangles = 1 << 15
    c.maxSplitTriangles = 1 << 12
    c.maxGeneratedClusters = (1 << 12) + (1 << 15)
    c.maxGeneratedClusterMegs = 128
    c.maxGeneratedVertices = 1 << 22
    return c


def ortho_rays(n: int, height: float, extent: float = 1.0) -> np.ndarray:
    """n x n rays straight down (-Y) from y = height through pixel centres
    of [0, extent]^2 in XZ; row-major with z along rows."""
    c = (np.arange(n, dtype=np.float64) + 0.5) / n * extent
    x, z = np.meshgrid(c, c, indexing="xy")
    rays = np.zeros(n * n, abi.RAY_DT)
    rays["origin"][:, 0] = x.reshape(-1)
    rays["origin"][:, 1] = height
    rays["origin"][:, 2] = z.reshape(-1)
    rays["direction"][:, 1] = -1.0
    rays["tMin"] = 0.0
    rays["tMax"] = 100.0
    return rays


class DepthRenderer:
    """Grid of `quads` x `quads` quads over [0,1]^2 (XZ plane, normal +Y)
    displaced by `height_scale` * H, H the bilinear REPEAT sample of the
    R32_FLOAT tensor `heights` (size x size, device, contiguous; the renderer
    reads its storage in place, so in-place optimiser updates are seen by
    the next forward).  `rays` x `rays` orthographic depth rays start at
    y = eye_height.  Tessellation factors come from a camera at the same
    height above the grid centre with the given pixel target."""

    def __init__(self, heights, quads: int = 64, rays: int = 512, height_scale: float = 0.1, eye_height: float = 2.0,
                 tess_rate_pixels: float = 0.5, config: abi.Config = None):
        torch = _torch()
        if heights.dtype != torch.float32 or heights.dim() != 2 or not heights.is_contiguous() or not heights.is_cuda:
            raise ValueError("heights must be a contiguous 2-D float32 device tensor")
        self.torch = torch
        self.L = abi.lib()
        self.heights = heights.detach()
        geo = grid_patch(quads, quads, 1.0, 1.0)
        inst = Instance(0, np.eye(4), displacement_index=0, displacement_scale=height_scale)
        cam = Camera(eye=(0.5, eye_height, 0.5), center=(0.5, 0.0, 0.5), up=(0.0, 0.0, -1.0), width=rays, height=rays,
                     near=0.01, far=100.0)
        self.host = HostScene([geo], [inst], [(self.heights, abi.TEXTURE_R32_FLOAT)], cam, tess_rate_pixels,
                              "depth_grid")
        self.dscene = DeviceScene(self.host)
        self.config = config or depth_config()
        if not self.config.flags & abi.FLAG_VERTEX_TAPE:
            raise ValueError("the depth renderer needs FLAG_VERTEX_TAPE")
        self.r = Renderer(self.dscene, self.config)
        self.fc = self.host.frame()
        dev = torch.device("cuda:0")
        self.rays_host = ortho_rays(rays, eye_height)
        self.rays = torch.from_numpy(self.rays_host.view(np.uint8).copy()).to(dev)
        self.num_rays = rays * rays
        self.hits = torch.empty(self.num_rays * abi.HIT_DT.itemsize, dtype=torch.uint8, device=dev)
        self.hit_vertices = torch.empty(self.num_rays * 4, dtype=torch.int32, device=dev)
        self.grad_ptrs = torch.zeros(1, dtype=torch.int64, device=dev)
        self.frame_id = 0
        accel, self.xforms_addr = self.r.accel()
        self.tlas_instances = accel.tlasInstances

    def close(self):
        self.r.close()

    def _stream(self):
        return self.torch.cuda.current_stream()

    def render(self):
        """Depth (hit distance) per ray, -1 for a miss; float32 (rays,)."""
        s = self._stream()
        self.r.tessellate(self.fc, s)
        self.r.build_accel(s)
        p = abi.TraceParams(self.rays.data_ptr(), self.num_rays, 0, 0, 0, self.hits.data_ptr(), None,
                            self.hit_vertices.data_ptr())
        self.r.trace(p, s)
        self.frame_id += 1
        return self.hits.view(self.torch.float32).view(self.num_rays, 6)[:, 0].clone()

    def adjoint(self, grad_t):
        """d(sum grad_t * t)/d(heights) for the last render()."""
        torch = self.torch
        grad_t = grad_t.detach().to(torch.float32).contiguous()
        if grad_t.numel() != self.num_rays or not grad_t.is_cuda:
            raise ValueError("grad_t must hold one float per ray on the device")
        grad = torch.zeros_like(self.heights)
        self.grad_ptrs.copy_(torch.tensor([grad.data_ptr()], dtype=torch.int64))
        p = abi.AdjointParams()
        p.rays, p.hits, p.hitVertices = self.rays.data_ptr(), self.hits.data_ptr(), self.hit_vertices.data_ptr()
        p.gradT, p.numRays, p.numTextures = grad_t.data_ptr(), self.num_rays, 1
        p.genVertices = self.r.building.genVertices
        p.tape = self.r.bindings.vertexTape
        p.tlasInstances = self.tlas_instances
        p.textures = self.r.bindings.displacementTextures
        p.gradTexels = self.grad_ptrs.data_ptr()
        abi.check(self.L.tc_depth_adjoint(C.byref(p), C.c_void_p(self._stream().cuda_stream)), "tc_depth_adjoint")
        return grad

    # ---- host copies of the last frame (checkers, diagnostics)

    def readback(self) -> abi.Readback:
        self.torch.cuda.synchronize()
        return self.r.readback()

    def host_state(self) -> dict:
        """Rays, hits, hit vertex indices, vertex tape and instance
        transforms of the last frame as numpy arrays."""
        rb = self.readback()
        nv = int(rb.numGenVertices)
        return {
            "rays": self.rays_host,
            "hits": self.hits.cpu().numpy().view(abi.HIT_DT),
            "hitVertices": self.hit_vertices.cpu().numpy().view(np.uint32),
            "tape": self.r.copy(self.r.bindings.vertexTape, nv * abi.VERTEX_TAPE_DT.itemsize).view(abi.VERTEX_TAPE_DT),
            "genVertices": self.r.copy(self.r.building.genVertices, nv * 12).view(np.float32).reshape(-1, 3),
            "xforms": self.r.copy(self.xforms_addr, 24 * 4 * len(self.host.instances)).view(np.float32),
            "readback": rb,
        }


def _autograd():
    torch = _torch()

    class DisplacedDepth(torch.autograd.Function):
        """t = depth(heights); backward through tc_depth_adjoint.  The
        backward of a forward must run before the next forward of the same
        renderer (the tape and hits are per frame)."""

        @staticmethod
        def forward(ctx, heights, renderer):
            if heights.data_ptr() != renderer.heights.data_ptr():
                raise ValueError("heights is not the tensor the renderer was built over")
            t = renderer.render()
            ctx.renderer = renderer
            ctx.frame = renderer.frame_id
            return t

        @staticmethod
        def backward(ctx, grad_t):
            r = ctx.renderer
            if r.frame_id != ctx.frame:
                raise RuntimeError("backward after a later forward of the same renderer: its tape was overwritten")
            return r.adjoint(grad_t), None

    return DisplacedDepth


_FN = None


def displaced_depth(heights, renderer: DepthRenderer):
    """Differentiable depth per ray (autograd through tc_depth_adjoint)."""
    global _FN
    if _FN is None:
        _FN = _autograd()
    return _FN.apply(heights, renderer)


def depth_loss(t, target):
    """Mean squared depth error over rays that hit in both renders."""
    mask = (t.detach() >= 0) & (target >= 0)
    return ((t - target)[mask] ** 2).sum() / mask.sum().clamp(min=1)


def target_depth(size: int = 512, seed: int = 3, octaves: int = 4, **kw):
    """Depth rendered with a seed-`seed` fBm height map (the fitting target)."""
    torch = _torch()
    h = torch.from_numpy(fbm_height(size, seed, octaves)).cuda()
    r = DepthRenderer(h, **kw)
    t = r.render()
    torch.cuda.synchronize()
    r.close()
    return t


def fit_heights(steps: int = 200, lr: float = 0.02, size: int = 512, seed: int = 3, callback=None, **kw):
    """Adam on a zero-initialised size x size height map so that its depth
    render matches the render of a seed-`seed` fBm map.  Returns the fitted
    heights and the per-step losses."""
    torch = _torch()
    target = target_depth(size, seed, **kw)
    heights = torch.zeros(size, size, dtype=torch.float32, device="cuda", requires_grad=True)
    r = DepthRenderer(heights, **kw)
    opt = torch.optim.Adam([heights], lr=lr)
    losses = []
    for step in range(steps):
        opt.zero_grad(set_to_none=True)
        loss = depth_loss(displaced_depth(heights, r), target)
        loss.backward()
        opt.step()
        losses.append(loss.detach())
        if callback:
            callback(step, loss)
    torch.cuda.synchronize()
    r.close()
    return heights.detach(), [float(x) for x in torch.stack(losses).cpu()]
