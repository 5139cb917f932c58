"""C-ABI checks that need no GPU: the library loads, exports every declared symbol,
validates arguments, and the host-side builder produces the documented layout."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import neuronbridge_amd as nb
from neuronbridge_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(nbx_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = nb.lib()
    names = declared_symbols()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding declares exactly those entry points
    assert set(L.SIGNATURES) == set(names)


def test_version_and_error_string():
    lib = nb.lib()
    assert lib.nbx_version() == 10000
    assert isinstance(lib.nbx_last_error(), bytes)


def test_argument_validation_without_device():
    md = L.ModelDesc(model=9, integrator=0, n=10, dt=0.1, tau=20.0)
    h = C.c_void_p()
    with pytest.raises(L.NbxError) as e:
        L.call("nbx_net_create", C.byref(md), 0, C.byref(h))
    assert e.value.status == L.NBX_ERR_INVALID and "model" in str(e.value)
    md = L.ModelDesc(model=L.LIF_VJ, integrator=0, n=0, dt=0.1, tau=20.0)
    with pytest.raises(L.NbxError):
        L.call("nbx_net_create", C.byref(md), 0, C.byref(h))
    with pytest.raises(L.NbxError):
        L.call("nbx_net_finalize", None)


def test_builder_layout_invariants():
    n, k, dmin, dmax, n_exc = 3000, 150, 1, 20, 2400
    syn = nb.build_fixed_indegree(n, k, dmin, dmax, 0xABCDEF12345, n_exc, threads=3)
    s2 = nb.build_fixed_indegree(n, k, dmin, dmax, 0xABCDEF12345, n_exc, threads=7)
    # deterministic for any thread count
    for a, b in ((syn.row_ptr, s2.row_ptr), (syn.post, s2.post), (syn.delay, s2.delay), (syn.type, s2.type)):
        assert np.array_equal(a, b)
    assert syn.row_ptr[0] == 0 and syn.row_ptr[-1] == n * k
    assert np.all(np.bincount(syn.post, minlength=n) == k)  # fixed in-degree
    pre = np.repeat(np.arange(n), np.diff(syn.row_ptr))
    assert np.array_equal(syn.type, (pre >= n_exc).astype(np.uint8))
    key = pre.astype(np.int64) * n + syn.post
    assert np.unique(key).size == key.size  # no multapses: sampling without replacement
    rowkey = pre.astype(np.int64) * 10 ** 9 + syn.delay.astype(np.int64) * 10 ** 6 + syn.post
    assert np.all(np.diff(rowkey) > 0)  # rows sorted by (delay, post)
    assert syn.delay.min() == dmin and syn.delay.max() == dmax
    hist = np.bincount(syn.delay)[dmin:]
    assert hist.min() > 0.9 * hist.mean()
    # sources of each post are spread uniformly over the population
    assert abs(np.mean(pre) - (n - 1) / 2) < 0.01 * n
