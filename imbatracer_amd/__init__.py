"""imbatracer_amd: MI355X (gfx950) ray-queue traversal, queue compaction and photon hash grid.

The compute path is libimba_hip.so (hand-written HIP kernels + C ABI, include/imba_hip.h);
this package holds its build script and ctypes bindings.
"""
from .native import DeviceBuffer, Event, ImbaError, Stream, device_count  # noqa: F401
from .scene import HashGrid, Scene, SceneBuild  # noqa: F401

__all__ = ["DeviceBuffer", "Event", "ImbaError", "Stream", "device_count", "HashGrid", "Scene", "SceneBuild"]
