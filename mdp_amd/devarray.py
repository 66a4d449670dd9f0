"""Device arrays owned through libmdpsnap (mdp_dev_* in include/mdpsnap.h).

The learned-surrogate kernels take caller-owned device buffers.  DevArray allocates them with the
library's own HIP runtime, so no second GPU runtime (such as the one bundled with a deep-learning
framework) has to share the process.
"""
import ctypes

import numpy as np

from . import library

_bound = False


def _lib():
    global _bound
    c = library().cdll
    if not _bound:
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        c.mdp_dev_malloc.argtypes = [ctypes.POINTER(vp), sz]
        c.mdp_dev_free.argtypes = [vp]
        c.mdp_dev_h2d.argtypes = [vp, vp, sz]
        c.mdp_dev_d2h.argtypes = [vp, vp, sz]
        c.mdp_dev_zero.argtypes = [vp, sz]
        c.mdp_dev_sync.argtypes = []
        _bound = True
    return c


def _chk(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what}: {library().cdll.mdp_snap_last_error().decode()}")


class DevArray:
    """A contiguous device buffer of ``size`` elements of ``dtype``."""

    def __init__(self, size, dtype=np.float64):
        self.c = _lib()
        self.size = int(size)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.size * self.dtype.itemsize
        p = ctypes.c_void_p()
        _chk(self.c.mdp_dev_malloc(ctypes.byref(p), max(self.nbytes, 8)), "mdp_dev_malloc")
        self._p = p

    @classmethod
    def from_numpy(cls, a, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype)
        d = cls(a.size, a.dtype)
        d.upload(a)
        return d

    @classmethod
    def zeros(cls, size, dtype=np.float64):
        d = cls(size, dtype)
        d.zero()
        return d

    @property
    def ptr(self):
        return self._p

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype).ravel()
        if a.size != self.size:
            raise ValueError(f"DevArray.upload: {a.size} elements into {self.size}")
        _chk(self.c.mdp_dev_h2d(self._p, a.ctypes.data, self.nbytes), "mdp_dev_h2d")

    def numpy(self):
        out = np.empty(self.size, self.dtype)
        _chk(self.c.mdp_dev_d2h(out.ctypes.data, self._p, self.nbytes), "mdp_dev_d2h")
        return out

    def zero(self):
        _chk(self.c.mdp_dev_zero(self._p, self.nbytes), "mdp_dev_zero")

    def free(self):
        if getattr(self, "_p", None) and self._p.value:
            self.c.mdp_dev_free(self._p)
            self._p = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def sync():
    _chk(_lib().mdp_dev_sync(), "mdp_dev_sync")
