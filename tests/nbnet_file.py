"""Writes .nbnet files for tests, from the format the reference defines
(python/src/neuronbridge/nbnet.py:34-38 and its finalize(): 68-byte header
"<8sHHIIQQ32s", metadata JSON, 32-byte records "<IIiffIii", SHA-256 over metadata + records)."""
from __future__ import annotations

import hashlib
import json
import struct

import numpy as np

HEADER = struct.Struct("<8sHHIIQQ32s")
RECORD_DTYPE = np.dtype([("src", "<u4"), ("dst", "<u4"), ("type", "<i4"), ("weight", "<f4"),
                         ("max_weight", "<f4"), ("delay", "<u4"), ("rule", "<i4"), ("trigger", "<i4")])
assert RECORD_DTYPE.itemsize == 32


def records(src, dst, syn_type, weight, delay, max_weight=None, rule=-1, trigger=-1) -> np.ndarray:
    n = len(src)
    r = np.zeros(n, dtype=RECORD_DTYPE)
    r["src"], r["dst"], r["type"], r["weight"] = src, dst, syn_type, weight
    r["max_weight"] = weight if max_weight is None else max_weight
    r["delay"], r["rule"], r["trigger"] = delay, rule, trigger
    return r


def write_nbnet(path, metadata: dict, recs: np.ndarray, *, magic=b"NBNET01\0", major=1, minor=0, flags=1,
                record_size=32, digest: bytes | None = None, truncate: int = 0) -> bytes:
    """Writes the file; returns the SHA-256 digest stored in the header."""
    meta = json.dumps(metadata, ensure_ascii=False, sort_keys=True, separators=(",", ":")).encode("utf-8")
    body = recs.tobytes()
    if digest is None:
        digest = hashlib.sha256(meta + body).digest()
    header = HEADER.pack(magic, major, minor, flags, record_size, len(meta), len(recs), digest)
    data = header + meta + body
    with open(path, "wb") as f:
        f.write(data[:len(data) - truncate] if truncate else data)
    return digest
