"""ctypes prototypes of the SNAP C ABI (the 23 entry points of src/Julia/Potential/cpuSnap.h:4-88).

The same table binds any shared library that exports that interface: the GPU library shipped by this
package, or a CPU build used as a checker by the tests.  Every pointer argument is a C-contiguous
numpy array of float64 (``d``) or int32 (``i``); scalars are C ``double``/``int``.
"""
import ctypes

import numpy as np

_D = "d"   # double*
_I = "i"   # int*
_f = "f"   # double (scalar)
_n = "n"   # int (scalar)

# name -> (restype, argument kinds); argument order and meaning follow cpuSnap.h
PROTOTYPES = {
    "idxcgcount": (ctypes.c_int, (_n,)),
    "idxucount": (ctypes.c_int, (_n,)),
    "idxbcount": (ctypes.c_int, (_n,)),
    "idxzcount": (ctypes.c_int, (_n,)),
    "cpuInitSna": (None, (_D, _D, _D, _I, _I, _I, _I, _I, _I, _I, _n)),
    "cpuNeighPairList": (None, (_I, _I, _D, _D, _I, _I, _I, _I, _I, _n, _n, _n)),
    "cpuNeighPairs": (None, (_D, _D, _I, _I, _I, _I, _I, _I, _I, _I, _I, _I, _n, _n)),
    "cpuSnapComputeUi": (None, (_D, _D, _D, _D, _D, _D, _f, _f, _f, _I, _I, _I, _I,
                                _n, _n, _n, _n, _n, _n)),
    "cpuAddWself2Ui": (None, (_D, _D, _f, _I, _I, _I, _n, _n, _n, _n, _n, _n)),
    "cpuSnapComputeEi": (None, (_D, _D, _D, _D, _D, _D, _I, _I, _I, _I, _I, _I,
                                _n, _n, _n, _n, _n, _n, _n, _n, _n, _n)),
    "cpuSnapComputeBi": (None, (_D, _D, _D, _D, _D, _I, _I, _I, _I,
                                _n, _n, _n, _n, _n, _n, _n, _n, _n)),
    "cpuSnapComputeYi": (None, (_D, _D, _D, _D, _D, _D, _I, _I, _I, _I, _I, _I,
                                _n, _n, _n, _n, _n, _n, _n, _n, _n)),
    "cpuSnapComputeFi": (None, (_D, _D, _D, _D, _D, _D, _D, _D, _f, _f, _f, _I, _I, _I, _I, _I, _I,
                                _n, _n, _n, _n, _n, _n, _n)),
    "cpuComputeSij": (None, (_D, _D, _D, _D, _D, _D, _D, _D, _D, _D, _D, _D, _f, _f, _f, _I, _I, _I,
                             _n, _n, _n, _n)),
    "cpuComputeUij": (None, (_D, _D, _D, _D, _D, _D, _D, _D, _f, _f, _f, _I, _I, _I, _n, _n, _n, _n)),
    "cpuZeroUarraytot2": (None, (_D, _D, _f, _I, _I, _I, _I, _n, _n, _n, _n, _n, _n)),
    "cpuAddUarraytot": (None, (_D, _D, _D, _D, _I, _I, _I, _n, _n, _n, _n)),
    "cpuComputeZi2": (None, (_D, _D, _D, _D, _D, _I, _I, _I, _n, _n, _n, _n, _n, _n)),
    "cpuComputeBi2": (None, (_D, _D, _D, _D, _D, _D, _I, _I, _I, _I, _I, _I,
                             _n, _n, _n, _n, _n, _n, _n, _n, _n)),
    "cpuComputeDbidrj": (None, (_D, _D, _D, _D, _D, _I, _I, _I, _I, _I, _I,
                                _n, _n, _n, _n, _n, _n, _n, _n, _n)),
    "cpuSnapTallyBispectrum": (None, (_D, _D, _I, _I, _n, _n, _n, _n, _n)),
    "cpuSnapTallyBispectrumDeriv": (None, (_D, _D, _D, _I, _I, _I, _I, _n, _n, _n, _n, _n, _n)),
    "cpuSnapTallyBispectrumVirial": (None, (_D, _D, _D, _D, _I, _I, _n, _n, _n, _n, _n, _n)),
}

_CTYPE = {
    _D: np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS"),
    _I: np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS"),
    _f: ctypes.c_double,
    _n: ctypes.c_int,
}


class SnapABI:
    """A shared library bound to the SNAP C ABI.  Attribute access returns the typed function."""

    def __init__(self, cdll, name=""):
        self.cdll = cdll
        self.name = name or getattr(cdll, "_name", "")
        for fn, (res, kinds) in PROTOTYPES.items():
            f = getattr(cdll, fn)
            f.restype = res
            f.argtypes = [_CTYPE[k] for k in kinds]
            setattr(self, fn, f)

    def __repr__(self):
        return f"SnapABI({self.name})"


def bind(path):
    """dlopen ``path`` (RTLD_LOCAL) and bind the SNAP ABI; raises OSError/AttributeError if absent."""
    return SnapABI(ctypes.CDLL(str(path), mode=ctypes.RTLD_LOCAL), str(path))
