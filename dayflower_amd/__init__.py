"""MI355X-native path-tracing hot path of Dayflower: HIP kernels behind a C-ABI
(libdayflower_hip.so) with a Python mirror of the GPURenderer API."""
from . import abi, scene  # noqa: F401

__all__ = ["abi", "scene"]
