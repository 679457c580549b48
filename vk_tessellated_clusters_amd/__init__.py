"""MI355X-native per-frame tessellated-cluster path: classify, split,
template instantiation, CLAS / BLAS build and ray traversal as gfx950 HIP
kernels behind the C ABI of include/tc_abi.h."""
from . import abi  # noqa: F401

__all__ = ["abi"]
