"""MI355X (gfx950) implementation of the SNAP force path of MDP.

``libmdpsnap.so`` (built in-tree by ``python -m mdp_amd.build``) exports

* the 23 C entry points of the upstream SNAP library (src/Julia/Potential/cpuSnap.h) with their
  signatures and host-pointer semantics, computed by HIP kernels -- see :mod:`mdp_amd.snapcpp` for
  the Python mirror of the Julia driver that calls them, and
* a device-resident velocity-Verlet MD engine (:class:`mdp_amd.md.SnapMD`).

There is no CPU fallback: :func:`library` raises if the HIP library is missing or cannot be loaded.
"""
import ctypes
import os

from . import _abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libmdpsnap.so")

_lib = None


class LibraryMissing(RuntimeError):
    pass


def library():
    """The GPU library bound to the SNAP C ABI (loaded once)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise LibraryMissing(f"{LIBPATH} is not built; run `python -m mdp_amd.build`")
        abi = _abi.bind(LIBPATH)
        c = abi.cdll
        c.mdp_snap_last_error.restype = ctypes.c_char_p
        c.mdp_snap_last_error.argtypes = []
        _lib = abi
    return _lib


def last_error():
    """Message of the last failed library call on this thread ('' if it succeeded)."""
    return library().cdll.mdp_snap_last_error().decode()


def check(where=""):
    """Raise RuntimeError if the last drop-in call recorded an error."""
    msg = last_error()
    if msg:
        raise RuntimeError(f"{where}: {msg}" if where else msg)
