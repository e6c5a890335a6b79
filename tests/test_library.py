"""CPU: the C-ABI library loads, exports every declared entry point, and the
host-only pieces (workload generators, host mirror) behave."""
import ctypes as C
import os
import re

import numpy as np

import strata_k as sk
from strata_k import workloads as wl
from strata_k.engine import LIB_PATH, SIGNATURES

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "strata_k.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(sk_\w+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB_PATH), "run __graft_entry__.build() first"
    lib = C.CDLL(LIB_PATH)
    decl = declared_functions()
    assert len(decl) > 40
    for name in decl:
        assert hasattr(lib, name), name


def test_binding_covers_header():
    bound = {n for n, _, _ in SIGNATURES}
    for name in declared_functions():
        assert name in bound, name


def test_library_loads_without_device():
    L = sk.load_library()
    assert b"gfx950" in L.sk_build_info()


def test_generators_deterministic():
    a1, b1 = wl.uniform_graph(7, 500, 2000)
    a2, b2 = wl.uniform_graph(7, 500, 2000)
    assert (a1 == a2).all() and (b1 == b2).all()
    assert len(a1) == 2000
    assert not (a1 == b1).any()
    assert len(set(zip(a1.tolist(), b1.tolist()))) == 2000
    s1, d1 = wl.rmat_graph(3, 12, 40000)
    s2, d2 = wl.rmat_graph(3, 12, 40000)
    assert (s1 == s2).all() and (d1 == d2).all()
    k = s1.astype(np.uint64) << np.uint64(32) | d1.astype(np.uint64)
    assert (np.diff(k.astype(np.int64)) > 0).all()  # sorted, unique
    assert not (s1 == d1).any()
    w = wl.weights(5, 100000)
    assert w.min() >= 1 and w.max() <= 1000
    assert abs(w.mean() - 500.5) < 5
    src = wl.sample_distinct(9, 1000, 64)
    assert len(set(src.tolist())) == 64


def test_degree_orientation():
    a, b = wl.rmat_graph(4, 10, 8000, undirected=True)
    x, y = wl.orient_by_degree(a, b, 1 << 10)
    deg = np.bincount(np.concatenate([a, b]), minlength=1 << 10)
    assert len(x) == len(a)
    assert ((deg[x] < deg[y]) | ((deg[x] == deg[y]) & (x < y))).all()


def test_binary_plan_intermediate_host():
    # K_4 oriented by id: out-degrees 3, 2, 1 -> 9 + 4 + 1
    edges = [(a, b) for a in range(4) for b in range(a + 1, 4)]
    assert sk.binary_plan_intermediate(edges) == 14
