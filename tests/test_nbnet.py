"""The .nbnet reader's host side (header, size and checksum validation, metadata) without a GPU.

SHA-256 is pinned against Python's hashlib: files are written with hashlib's digest and must
verify; a single flipped byte must not. Error behaviour follows the reference reader
(src/native/shared/src/streaming_build/NbnetReader.cpp:533-580)."""
import json

import numpy as np
import pytest

from nbnet_file import records, write_nbnet
from neuronbridge_amd import _lib as L
from neuronbridge_amd.network import nbnet_info, nbnet_metadata

META = {"layers": [{"count": 40, "model": "TimeDrivenLIF_Voltage_jump"}], "learning_rules": [],
        "outer_dynamics": [], "outer_dynamic_connections": [], "input_convs": []}


def some_records(n=40, count=300, seed=0):
    rng = np.random.default_rng(seed)
    return records(rng.integers(0, n, count), rng.integers(0, n, count), rng.integers(0, 2, count),
                   rng.normal(0, 1, count).astype(np.float32), rng.integers(1, 21, count))


def test_header_fields_and_checksum(tmp_path):
    p = tmp_path / "a.nbnet"
    recs = some_records()
    write_nbnet(p, META, recs)
    h = nbnet_info(str(p), True)
    meta = json.dumps(META, sort_keys=True, separators=(",", ":")).encode()
    assert (h.version_major, h.version_minor, h.flags, h.checksum_verified) == (1, 0, 1, 1)
    assert h.connection_count == len(recs) and h.metadata_bytes == len(meta)
    assert h.file_bytes == 68 + len(meta) + 32 * len(recs)
    assert nbnet_metadata(str(p)) == meta


@pytest.mark.parametrize("meta_pad", [0, 1, 7, 8, 9, 63, 64, 65, 1000])
def test_sha256_matches_hashlib_across_block_boundaries(tmp_path, meta_pad):
    # metadata + records lengths walk across the 64-byte block and 56-byte padding boundaries
    p = tmp_path / "b.nbnet"
    meta = dict(META, pad="x" * meta_pad)
    recs = some_records(count=37 + meta_pad)
    write_nbnet(p, meta, recs)
    assert nbnet_info(str(p), True).checksum_verified == 1


def test_large_file_checksum(tmp_path):
    # more than one 8 MiB read buffer
    p = tmp_path / "c.nbnet"
    write_nbnet(p, META, some_records(count=300_001))
    assert nbnet_info(str(p), True).connection_count == 300_001


def flip_byte(path, offset):
    data = bytearray(open(path, "rb").read())
    data[offset] ^= 0x01
    open(path, "wb").write(bytes(data))


@pytest.mark.parametrize("where", ["metadata", "record"])
def test_checksum_mismatch_is_rejected(tmp_path, where):
    p = tmp_path / "d.nbnet"
    write_nbnet(p, META, some_records())
    flip_byte(p, 70 if where == "metadata" else -5)
    with pytest.raises(L.NbxError, match="checksum mismatch"):
        nbnet_info(str(p), True)
    nbnet_info(str(p), False)  # structure is intact: accepted without the checksum


@pytest.mark.parametrize("kw,msg", [
    (dict(magic=b"NBNET02\0"), "invalid nbnet magic"),
    (dict(major=2), "unsupported nbnet major version"),
    (dict(record_size=24), "unsupported nbnet connection record size"),
    (dict(truncate=5), "nbnet file size mismatch"),
    (dict(truncate=10**6), "truncated nbnet header"),
])
def test_structural_errors(tmp_path, kw, msg):
    p = tmp_path / "e.nbnet"
    write_nbnet(p, META, some_records(), **kw)
    with pytest.raises(L.NbxError, match=msg):
        nbnet_info(str(p), False)


def test_missing_checksum_flag(tmp_path):
    p = tmp_path / "f.nbnet"
    write_nbnet(p, META, some_records(), flags=0)
    nbnet_info(str(p), False)
    with pytest.raises(L.NbxError, match="checksum is missing"):
        nbnet_info(str(p), True)


def test_missing_file(tmp_path):
    with pytest.raises(L.NbxError, match="cannot open nbnet file"):
        nbnet_info(str(tmp_path / "none.nbnet"), False)
