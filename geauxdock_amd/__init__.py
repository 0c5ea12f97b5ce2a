"""MI355X (gfx950) GeauxDock Monte Carlo hot path.

The product is ``lib/libgeauxdock_amd.so``: C++ host code (loaders, job
packing, docking drivers, batching, serving) and hand-written HIP kernels,
behind the C ABI of ``include/geauxdock_amd.h``.  This package only locates
and loads that library (``lib()``) and offers thin numpy-friendly wrappers.
There is no fallback: without the built library every call raises.
"""
import ctypes as _C
import os as _os

from . import abi

HERE = _os.path.dirname(_os.path.abspath(__file__))
LIB_PATH = _os.path.join(HERE, "lib", "libgeauxdock_amd.so")
PARAM_DIR = _os.path.join(HERE, "data", "parameters")
_lib = None


class GDError(RuntimeError):
    pass


def lib():
    """Load the C-ABI library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not _os.path.exists(LIB_PATH):
            raise GDError("geauxdock_amd native library missing: %s (run python -m "
                          "geauxdock_amd.build)" % LIB_PATH)
        _lib = abi.declare(_C.CDLL(LIB_PATH), optional=True)
    return _lib


def check(status, what=""):
    if status != abi.GD_OK:
        msg = lib().gd_last_error()
        raise GDError("%s failed: %s %s" % (what, abi.STATUS_NAMES.get(status, status),
                                            msg.decode() if msg else ""))


from .host import (HostJob, load_enepara, load_complex, synth_complex, synth_screen, DeviceJob,  # noqa: E402
                   make_mc)
