"""Checkpoint files in the ARCW v1 layout (include/pic/checkpoint.hpp:1-131),
written and read by the library with host arrays (no GPU needed):
tests/test_checkpoint_format.cu's round trip of the header, deck text,
git hash and manifest, its bad-magic rejection, plus the version rule, the
byte layout of the header, payload round trips, and schema-only files."""
import ctypes as C
import struct

import numpy as np
import pytest

from arcwarden_amd import _lib as L


def _write(path, header, arrays, deck=b"", git=b""):
    arr = (L.AwCheckpointArray * max(1, len(arrays)))()
    keep = []
    for i, (name, dtype, count, data) in enumerate(arrays):
        arr[i].name = name.encode()
        arr[i].dtype = dtype
        arr[i].count = count
        if data is not None:
            keep.append(data)
            arr[i].data = data.ctypes.data
        arr[i].on_device = 0
    return L.lib().aw_checkpoint_write(str(path).encode(), C.byref(header), deck, git, arr,
                                       len(arrays))


def _open(path):
    h = C.c_void_p()
    L.call("aw_checkpoint_open", str(path).encode(), C.byref(h))
    return h


def _header(**kw):
    h = L.AwCheckpointHeader()
    for k, v in kw.items():
        setattr(h, k, v)
    return h


def test_header_round_trip_and_layout(tmp_path):
    path = tmp_path / "format.arcw"
    deck = "[chirp]\nnx = 6554\n# snapshot text with\nunicode Ωe0 ✓\n".encode()
    git = b"59d9eb07e56e79eed94ed10ee9f50cab0fbcf4ab"
    h = _header(rng_seed=0xDEADBEEFCAFEBABE, step=123456789012345, time=9876.5432, sd=250.0,
                eps_l=0.37)
    man = [("xp", 1, 11404826, None), ("ux", 0, 11404826, None), ("cell", 2, 11404826, None),
           ("rng_state", 3, 1, None)]
    assert _write(path, h, man, deck, git) == 0
    raw = path.read_bytes()
    magic, ver, seed, step, t, sd, el = struct.unpack_from("<IIQqddd", raw, 0)
    assert (magic, ver, seed, step, t, sd, el) == (0x57435241, 1, 0xDEADBEEFCAFEBABE,
                                                  123456789012345, 9876.5432, 250.0, 0.37)
    assert raw[:4] == b"ARCW"
    (n,) = struct.unpack_from("<Q", raw, 48)
    assert raw[56:56 + n] == deck
    ck = _open(path)
    r = L.AwCheckpointHeader()
    na, dl, gl = C.c_int32(), C.c_int64(), C.c_int64()
    L.call("aw_checkpoint_info", ck, C.byref(r), C.byref(na), C.byref(dl), C.byref(gl))
    assert (r.version, r.rng_seed, r.step, r.time, r.sd, r.eps_l) == (
        1, h.rng_seed, h.step, h.time, h.sd, h.eps_l)
    buf = C.create_string_buffer(dl.value + 1)
    L.call("aw_checkpoint_text", ck, 0, buf, dl.value + 1)
    assert buf.raw[:dl.value] == deck
    buf = C.create_string_buffer(gl.value + 1)
    L.call("aw_checkpoint_text", ck, 1, buf, gl.value + 1)
    assert buf.value == git
    assert na.value == 4
    for i, (name, dtype, count, _) in enumerate(man):
        nm = C.create_string_buffer(64)
        t, cnt, present = C.c_uint32(), C.c_uint64(), C.c_int32()
        L.call("aw_checkpoint_array_info", ck, i, nm, 64, C.byref(t), C.byref(cnt),
               C.byref(present))
        assert (nm.value.decode(), t.value, cnt.value, present.value) == (name, dtype, count, 0)
    L.call("aw_checkpoint_close", ck)
    # corrupt magic -> rejected with the reference's message
    with open(path, "r+b") as f:
        f.write(struct.pack("<I", 0x12345678))
    h2 = C.c_void_p()
    assert L.lib().aw_checkpoint_open(str(path).encode(), C.byref(h2)) != 0
    assert "bad magic" in L.last_error()


def test_newer_version_and_truncation_rejected(tmp_path):
    path = tmp_path / "v2.arcw"
    assert _write(path, _header(step=1), [], b"d", b"g") == 0
    orig = path.read_bytes()
    raw = bytearray(orig)
    raw[4:8] = struct.pack("<I", 2)
    path.write_bytes(bytes(raw))
    h = C.c_void_p()
    assert L.lib().aw_checkpoint_open(str(path).encode(), C.byref(h)) != 0
    assert "version 2 newer than reader" in L.last_error()
    path.write_bytes(orig[:20])
    assert L.lib().aw_checkpoint_open(str(path).encode(), C.byref(h)) != 0
    assert "short read" in L.last_error()


def test_payload_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    a = rng.normal(size=1001).astype(np.float32)
    b = rng.normal(size=17)
    c = rng.integers(-5, 5, 33).astype(np.int32)
    d = rng.integers(0, 2 ** 63, 9, dtype=np.uint64)
    path = tmp_path / "payload.arcw"
    assert _write(path, _header(step=7, time=0.7),
                  [("a", 0, a.size, a), ("b", 1, b.size, b), ("c", 2, c.size, c),
                   ("d", 3, d.size, d)], b"deck", b"") == 0
    ck = _open(path)
    for i, src in enumerate((a, b, c, d)):
        out = np.zeros_like(src)
        L.call("aw_checkpoint_read_array", ck, i, out.ctypes.data_as(C.c_void_p), 0)
        assert np.array_equal(out, src)
    L.call("aw_checkpoint_close", ck)
    # the file is the header plus the payloads, nothing else
    head = 4 + 4 + 8 * 5 + 8 + 4 + 8 + 0 + 4 + sum(4 + 1 + 4 + 8 for _ in range(4))
    assert path.stat().st_size == head + a.nbytes + b.nbytes + c.nbytes + d.nbytes
    # truncated payloads read as absent
    raw = path.read_bytes()
    path.write_bytes(raw[:-d.nbytes])
    ck = _open(path)
    present = C.c_int32()
    L.call("aw_checkpoint_array_info", ck, 3, None, 0, None, None, C.byref(present))
    assert present.value == 0
    out = np.zeros_like(d)
    assert L.lib().aw_checkpoint_read_array(ck, 3, out.ctypes.data_as(C.c_void_p), 0) != 0
    assert "absent" in L.last_error()
    L.call("aw_checkpoint_close", ck)
