"""The branch-and-bound queue against the reference's MinHeap semantics.

tests/heap_model.py restates helper/minheap.cpp step by step in Python. The
oracle's heap (oracle/minheap_oracle.hpp) and the product's host heap
(product-graph-cycles_amd/csrc/bnb_heap.h, compiled here with g++ into a
harness, tests/heap_harness.cpp) run the same random operation sequences and
must give the same outputs: pop order with equal keys (a newer entry bubbles
above older ones), pushes that are ignored (ids outside [0, capacity), ids
already queued, a full heap), peak of an empty heap, decrease and peakKey.
The device heap (csrc/bnb_device.h) is exercised by the batched solves that
branch (tests/test_gpu_bnb.py) against the oracle.
"""
import ctypes as C
import os
import subprocess
import tempfile

import numpy as np
import pytest

import heap_model as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = {
    "product": (os.path.join(ROOT, "product-graph-cycles_amd", "csrc", "bnb_heap.h"), "pgc"),
    "oracle": (os.path.join(ROOT, "oracle", "minheap_oracle.hpp"), "oracle"),
}
_LIBS = {}


def _lib(which):
    if which not in _LIBS:
        hdr, ns = HEADERS[which]
        d = tempfile.mkdtemp(prefix="pgc_heap_")
        so = os.path.join(d, "heap_%s.so" % which)
        subprocess.run(["g++", "-O1", "-std=c++17", "-shared", "-fPIC", '-DHEAP_HEADER="%s"' % hdr,
                        "-DHEAP_NS=%s" % ns, "-o", so, os.path.join(ROOT, "tests", "heap_harness.cpp")],
                       check=True)
        L = C.CDLL(so)
        L.heap_run.restype = C.c_long
        L.heap_run.argtypes = [C.c_long] + [C.c_void_p] * 6
        _LIBS[which] = L
    return _LIBS[which]


def _run_cpp(which, cap, ops):
    op = np.array([o for o, _, _ in ops], np.int32)
    key = np.array([k for _, k, _ in ops], np.int64)
    val = np.array([v for _, _, v in ops], np.float64)
    ov = np.zeros(len(ops) + 1, np.float64)
    ok = np.zeros(len(ops) + 1, np.int64)
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    m = _lib(which).heap_run(cap, p(op), p(key), p(val), len(ops), p(ov), p(ok))
    return [(float(ov[i]), int(ok[i])) for i in range(m)]


def _random_ops(rng, cap, n):
    ops = []
    values = [0.0, 0.25, 0.5, 0.5, 1.0, 2.0, 3.5]  # few distinct values: many ties
    for _ in range(n):
        r = rng.random()
        key = int(rng.integers(-2, cap + 3))  # some out of range
        val = float(rng.choice(values)) if rng.random() < 0.8 else float(rng.uniform(0, 4))
        if r < 0.5:
            ops.append((0, key, val))
        elif r < 0.8:
            ops.append((1, 0, 0.0))
        elif r < 0.87:
            ops.append((2, 0, 0.0))
        elif r < 0.94:
            ops.append((3, key, val))
        else:
            ops.append((4, key, 0.0))
    return ops


def test_model_known_answers():
    # equal keys: the newer push bubbles above (bubbleUp swaps on >=)
    assert H.run(4, [(0, 1, 0.5), (0, 2, 0.5), (1, 0, 0), (1, 0, 0)]) == [(0.5, 2), (0.5, 1)]
    # ignored pushes: out of range, already present (key kept), full heap
    ops = [(0, 5, 0.1), (0, -1, 0.1), (0, 0, 2.0), (0, 0, 0.1), (0, 1, 1.0), (0, 2, 3.0), (0, 3, 0.0), (1, 0, 0),
           (1, 0, 0), (1, 0, 0), (1, 0, 0), (2, 0, 0)]
    assert H.run(3, ops) == [(1.0, 1), (2.0, 0), (3.0, 2), (-1.0, -1), (float("inf"), -2)]


@pytest.mark.parametrize("which", ["oracle", "product"])
@pytest.mark.parametrize("seed", range(12))
def test_cpp_heap_matches_reference_model(which, seed):
    rng = np.random.default_rng(seed)
    cap = int(rng.integers(1, 24))
    ops = _random_ops(rng, cap, 400)
    assert _run_cpp(which, cap, ops) == H.run(cap, ops)
