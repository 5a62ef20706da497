"""MI355X-native radiation-transport scattering integral.

The product is librsh_scatter.so (hand-written HIP for gfx950 behind the C ABI
in include/rsh_scatter.h); this package is its Python host mirror. Submodules
load the library on first use, so `python -m relativistic_solver_hllc_amd.build`
works before anything is built.
"""
import importlib

_EXPORTS = {
    "DirectionSet": "scatter",
    "Problem": "scatter",
    "device_count": "scatter",
    "fibonacci_directions": "scatter",
    "fill_lognormal_dev": "scatter",
    "partition": "scatter",
    "read_directions": "scatter",
    "read_simple_bin": "scatter",
    "compare_files": "scatter",
    "selftest_mfma_f64": "scatter",
    "write_directions": "scatter",
    "write_simple_bin": "scatter",
    "DeviceArray": "device",
    "PinnedArray": "device",
    "Stream": "device",
    "Event": "device",
}

__all__ = sorted(_EXPORTS)


def __getattr__(name):
    if name in _EXPORTS:
        mod = importlib.import_module(f".{_EXPORTS[name]}", __name__)
        return getattr(mod, name)
    raise AttributeError(name)
