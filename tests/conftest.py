import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libcpusnap_ref.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def _ensure_oracle():
    if not os.path.exists(ORACLE_SO):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "liboracle.so"], check=True,
                       stdout=subprocess.DEVNULL)
    return ORACLE_SO


@pytest.fixture(scope="session")
def oracle():
    """The CPU oracle bound to the SNAP C ABI (test infrastructure)."""
    from mdp_amd import _abi
    return _abi.bind(_ensure_oracle())


@pytest.fixture(scope="session")
def reflib():
    """The compiled upstream SNAP library, when it was built in this tree; else skip."""
    if not os.path.exists(REF_SO):
        pytest.skip("upstream SNAP library not built (oracle/_ref)")
    from mdp_amd import _abi
    return _abi.bind(REF_SO)


def golden_names():
    return sorted(f[5:-4] for f in os.listdir(GOLDEN) if f.startswith("snap_") and f.endswith(".npz"))


def load_golden(name):
    return dict(np.load(os.path.join(GOLDEN, f"snap_{name}.npz"), allow_pickle=False))


def sna_and_pairs(lib, g):
    """Build the SNAP structure on ``lib`` and the pair dict from a golden fixture."""
    from mdp_amd import snapcpp
    s = snapcpp.initsna(lib, g["param"], g["elemradius"], g["elemweight"], g["coeff"])
    n = len(g["x"])
    p = dict(rij=g["rij"], ai=g["ai"], aj=g["aj"], ti=g["ti"], tj=g["tj"],
             atomtype=g["types"].astype(np.int32), ilist=np.arange(n, dtype=np.int32), n=n,
             ijnum=len(g["ai"]))
    return s, p


def relerr(a, b):
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    scale = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / scale)
