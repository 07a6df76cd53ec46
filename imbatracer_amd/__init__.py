"""imbatracer_amd: MI355X (gfx950) ray-queue traversal, queue compaction and photon hash grid.

The compute path is libimba_hip.so (hand-written HIP kernels + C ABI, include/imba_hip.h);
this package holds its build script and ctypes bindings.  The library is loaded on first use
of a binding (importing imbatracer_amd.build does not need it).
"""
_EXPORTS = {
    "DeviceBuffer": "native", "Event": "native", "ImbaError": "native", "Stream": "native",
    "device_count": "native", "HashGrid": "scene", "Scene": "scene", "SceneBuild": "scene",
}

__all__ = sorted(_EXPORTS)


def __getattr__(name):
    if name in _EXPORTS:
        import importlib

        mod = importlib.import_module("." + _EXPORTS[name], __name__)
        return getattr(mod, name)
    raise AttributeError(name)
