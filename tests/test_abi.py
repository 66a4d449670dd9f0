"""The C-ABI library loads here (no GPU needed) and exports every symbol include/mdpsnap.h declares."""
import ctypes
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "mdpsnap.h")
LIB = os.path.join(ROOT, "mdp_amd", "libmdpsnap.so")
UPSTREAM_23 = ["idxcgcount", "idxucount", "idxbcount", "idxzcount", "cpuInitSna", "cpuNeighPairList",
               "cpuNeighPairs", "cpuSnapComputeUi", "cpuAddWself2Ui", "cpuSnapComputeEi", "cpuSnapComputeBi",
               "cpuSnapComputeYi", "cpuSnapComputeFi", "cpuComputeSij", "cpuComputeUij", "cpuZeroUarraytot2",
               "cpuAddUarraytot", "cpuComputeZi2", "cpuComputeBi2", "cpuComputeDbidrj", "cpuSnapTallyBispectrum",
               "cpuSnapTallyBispectrumDeriv", "cpuSnapTallyBispectrumVirial"]


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\([^;{]*\)\s*;", text)) - {"if", "while"})


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.fail("libmdpsnap.so is not built (python -m mdp_amd.build)")
    return ctypes.CDLL(LIB)


def test_header_declares_the_upstream_interface():
    names = declared()
    for n in UPSTREAM_23:
        assert n in names, n


def test_every_declared_symbol_is_exported(lib):
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_index_counts_without_gpu(lib):
    # pure host integer functions of the drop-in (cpuSnap.cpp:23-69)
    assert (lib.idxucount(8), lib.idxzcount(8), lib.idxbcount(8), lib.idxcgcount(8)) == (285, 2386, 55, 4098)


def test_abi_binding_table_matches_header():
    from mdp_amd import _abi
    assert sorted(_abi.PROTOTYPES) == sorted(UPSTREAM_23)
