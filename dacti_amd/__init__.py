"""MI355X-native octree leaf update (DACTI hot path) over a HIP C ABI."""
from . import case, scene, workloads  # noqa: F401
