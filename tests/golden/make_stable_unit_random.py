"""Writes tests/golden/stable_unit_random.npz: input/output vectors of the reference's
StableUnitRandom (src/native/legacy/source_file_realtime_v1_async/LearningRule/src/Rule/
CerebullarLearningRule.cpp:10-24), computed by the reference's own code compiled from its source
into oracle/_ref/libref_sur.so (`make -C oracle ref`, which needs the reference tree).

Inputs: every combination of edge-case seeds / rule ids / synapse ids / spike steps, plus seeded
random draws over the full int32 / uint32 ranges. Outputs are stored as float32 bit patterns.
"""
from __future__ import annotations

import ctypes as C
import itertools
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
LIB = os.path.join(ROOT, "oracle", "_ref", "libref_sur.so")
OUT = os.path.join(HERE, "stable_unit_random.npz")


def vectors():
    seeds = [0, 1, 17, 42, 0x9E3779B9, 0x7FFFFFFF, 0x80000000, 0xFFFFFFFF]
    rules = [0, 1, 2, 7, 63, -1, 0x7FFFFFFF, -0x80000000]
    syns = [0, 1, 2, 999, 123456, 99999999, -1, 0x7FFFFFFF]
    steps = [0, 1, 119, 1200, 20000, 65535, -1, 0x7FFFFFFF]
    grid = np.array(list(itertools.product(seeds, rules, syns, steps)), dtype=np.int64)
    rng = np.random.default_rng(0x5EED0003)
    m = 2048
    rand = np.stack([rng.integers(0, 2 ** 32, m), rng.integers(-2 ** 31, 2 ** 31, m),
                     rng.integers(-2 ** 31, 2 ** 31, m), rng.integers(-2 ** 31, 2 ** 31, m)], axis=1)
    # the cerebellar rule's own range: seed 17 (CerebullarLearningRule.h), small rule ids, CSR-sized
    # connection indices, steps of a long run
    small = np.stack([np.full(m, 17), rng.integers(0, 64, m), rng.integers(0, 10 ** 8, m),
                      rng.integers(0, 10 ** 6, m)], axis=1)
    return np.concatenate([grid, rand, small]).astype(np.int64)


def main() -> int:
    if not os.path.exists(LIB):
        print(f"{LIB} is missing: run `make -C oracle ref` in a tree that has the reference", file=sys.stderr)
        return 1
    lib = C.CDLL(LIB)
    f = lib.ref_stable_unit_random
    f.restype = C.c_float
    f.argtypes = [C.c_uint32, C.c_int32, C.c_int32, C.c_int32]
    x = vectors()
    seed = (x[:, 0] & 0xFFFFFFFF).astype(np.uint32)
    rule, syn, step = (x[:, k].astype(np.int32) for k in (1, 2, 3))
    out = np.array([f(int(a), int(b), int(c), int(d)) for a, b, c, d in zip(seed, rule, syn, step)], np.float32)
    np.savez_compressed(OUT, seed=seed, rule=rule, synapse=syn, step=step, value_bits=out.view(np.uint32))
    print(f"wrote {OUT}: {len(out)} vectors")
    return 0


if __name__ == "__main__":
    sys.exit(main())
