"""Shared pytest configuration: the `gpu` marker and golden-fixture loaders."""
import ctypes
import json
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "fk20-cuda_amd"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


def load_golden(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def hex_fr(h):
    v = int(h, 16)
    return np.array([(v >> (64 * i)) & 0xFFFFFFFFFFFFFFFF for i in range(4)], dtype=np.uint64)


def hex_fp(h):
    v = int(h, 16)
    return np.array([(v >> (64 * i)) & 0xFFFFFFFFFFFFFFFF for i in range(6)], dtype=np.uint64)


def hex_g1p(triple):
    return np.concatenate([hex_fp(h) for h in triple])


@pytest.fixture(scope="session")
def golden():
    return load_golden


HOST_SRC = os.path.join(ROOT, "tests", "native", "arith_host.cpp")
HOST_LIB = os.path.join(ROOT, "tests", "native", "_build", "libarith_host.so")


@pytest.fixture(scope="session")
def arith_host():
    """The device arithmetic and index-map headers compiled by g++ (tests/native/arith_host.cpp),
    rebuilt when a source is newer than the library."""
    csrc = os.path.join(ROOT, "fk20-cuda_amd", "csrc")
    deps = [HOST_SRC] + [os.path.join(csrc, f) for f in
                         ("fp.hip.h", "fr.hip.h", "g1.hip.h", "ladder.hip.h", "schedule.hip.h", "constants.hip.h")]
    if not os.path.exists(HOST_LIB) or os.path.getmtime(HOST_LIB) < max(os.path.getmtime(d) for d in deps):
        os.makedirs(os.path.dirname(HOST_LIB), exist_ok=True)
        subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-o", HOST_LIB, HOST_SRC], check=True)
    return ctypes.CDLL(HOST_LIB)
