"""CountedState container codec on the host (no GPU): the byte format of the
reference's snapshot file (cpp/src/counted_state.cpp:30-56 writes it,
:122-147 reads it, :57-76 validates it, :149-165 cuts one colony out).

The encoder below writes the format field by field from that description,
independently of the library, so a round trip through the library pins the
field order, widths and the legacy-magic rule; the rejections mirror the
reference's own corruption cases (tests/python/test_state_fork_native.py:113-130).
"""
import struct

import numpy as np
import pytest

from gplsaco_amd import gp_faco_ext as ext

MAGIC = 0x3254534F43414647          # "GFACOST2"
LEGACY = 0x3154534F43414647         # "GFACOST1": 8 trailing bytes, skipped


def _word(v):
    return struct.pack("<Q", v)


def _real(v):
    return struct.pack("<d", v)


def encode(*, dimension=5, colonies=2, ants=4, widths=(3, 2, 2), reals=(2.0, 0.5, 0.05, 0.5), limits=(4, 6),
           mask=0xFFFFFFFF, completed=1, progress=8, tasks=((7, 1), (7, 2)), incumbents=None, buffers=None,
           magic=MAGIC):
    if incumbents is None:
        incumbents = [(10.0 + c, list(range(dimension))[c:] + list(range(dimension))[:c]) for c in range(colonies)]
    if buffers is None:
        buffers = [("keys", 8, 1, bytes(range(8 * colonies))),
                   ("plane", 2, 3, bytes(range(2 * 3 * colonies * 2)))]
    out = _word(magic) + _word(dimension) + _word(colonies) + _word(ants)
    out += b"".join(_word(w) for w in widths) + b"".join(_real(r) for r in reals)
    out += _word(limits[0]) + _word(limits[1]) + _word(mask) + _word(completed) + _word(progress)
    out += _word(len(tasks)) + b"".join(_word(k) + _word(s) for k, s in tasks)
    out += _word(len(incumbents))
    for cost, tour in incumbents:
        out += _real(cost) + _word(len(tour)) + b"".join(_word(v) for v in tour)
    out += _word(len(buffers))
    for name, element, planes, data in buffers:
        out += _word(len(name)) + name.encode() + _word(element) + _word(planes) + _word(len(data)) + data
    return out


def decode(data):
    """The fields of a GFACOST2 container as a dict of encode() arguments."""
    words = lambda at, k: struct.unpack_from("<%dQ" % k, data, at)  # noqa: E731
    (magic, dimension, colonies, ants, pw, bw, lw), at = words(0, 7), 56
    reals = struct.unpack_from("<4d", data, at)
    at += 32
    l0, l1, mask, completed, progress, count = words(at, 6)
    at += 48
    tasks = [words(at + 16 * i, 2) for i in range(count)]
    at += 16 * count
    (count,), at = words(at, 1), at + 8
    incumbents = []
    for _ in range(count):
        (cost,), (length,) = struct.unpack_from("<d", data, at), words(at + 8, 1)
        incumbents.append((cost, list(words(at + 16, length))))
        at += 16 + 8 * length
    (count,), at = words(at, 1), at + 8
    buffers = []
    for _ in range(count):
        (length,), at = words(at, 1), at + 8
        name, at = data[at:at + length].decode(), at + length
        element, planes, size = words(at, 3)
        at += 24
        buffers.append((name, element, planes, data[at:at + size]))
        at += size
    assert at == len(data) and magic == MAGIC
    return dict(dimension=dimension, colonies=colonies, ants=ants, widths=(pw, bw, lw), reals=reals, limits=(l0, l1),
                mask=mask, completed=completed, progress=progress, tasks=tasks, incumbents=incumbents,
                buffers=buffers)


def test_round_trip_keeps_every_byte():
    data = encode()
    state = ext.CountedState.from_bytes(data)
    assert state.to_bytes() == data
    assert encode(**decode(data)) == data
    assert ext.CountedState.from_bytes(bytearray(data)).to_bytes() == data


def test_legacy_magic_skips_its_trailing_metadata():
    body = encode(magic=LEGACY)
    state = ext.CountedState.from_bytes(body + b"\x01" * 8)
    assert state.to_bytes() == encode()
    with pytest.raises(ValueError, match="snapshot"):
        ext.CountedState.from_bytes(body)       # the trailing 8 bytes are part of the legacy format


@pytest.mark.parametrize("cut", ["empty", "short", "long", "magic", "header"])
def test_truncated_extended_or_foreign_bytes_are_rejected(cut):
    data = encode()
    value = {"empty": b"", "short": data[:-1], "long": data + b"extra", "magic": _word(MAGIC + 1) + data[8:],
             "header": data[:12]}[cut]
    with pytest.raises(ValueError, match="snapshot"):
        ext.CountedState.from_bytes(value)


@pytest.mark.parametrize("change", [
    dict(dimension=2, incumbents=[(1.0, [0, 1])] * 2),                 # dimension below 3
    dict(colonies=129, buffers=[("keys", 8, 1, bytes(8 * 129))]),      # more than 128 colonies
    dict(tasks=((7, 1),)),                                              # a task missing
    dict(incumbents=[(10.0, [0, 1, 2, 3, 3]), (11.0, [0, 1, 2, 3, 4])]),  # not a permutation
    dict(incumbents=[(0.0, [0, 1, 2, 3, 4]), (11.0, [0, 1, 2, 3, 4])]),   # cost not positive
    dict(incumbents=[(float("nan"), [0, 1, 2, 3, 4]), (11.0, [0, 1, 2, 3, 4])]),
    dict(progress=6),                                                   # not whole batches of 4 ants
    dict(completed=3),                                                  # past the progress budget
    dict(mask=0xFFFF0000),                                              # no hold action left
    dict(buffers=[]),                                                   # no buffers
    dict(buffers=[("keys", 8, 1, bytes(16)), ("keys", 8, 1, bytes(16))]),  # duplicate name
    dict(buffers=[("keys", 8, 13, bytes(8 * 13 * 2))]),                 # more than 12 planes
    dict(buffers=[("keys", 8, 1, bytes(24))]),                          # not a whole slice per colony
    dict(buffers=[("", 8, 1, bytes(16))]),                              # empty name
])
def test_structure_rules_of_the_reference(change):
    with pytest.raises(ValueError, match="snapshot"):
        ext.CountedState.from_bytes(encode(**change))


def test_select_colony_cuts_each_plane():
    colonies, dimension = 3, 4
    values = np.arange(3 * colonies * 5, dtype=np.uint16).reshape(3, colonies, 5)   # [plane][colony][5]
    planes = values.tobytes()
    keys = np.array([11, 22, 33], dtype=np.uint64).tobytes()
    data = encode(dimension=dimension, colonies=colonies, tasks=((1, 5), (2, 6), (3, 7)),
                  incumbents=[(float(c + 1), [(i + c) % dimension for i in range(dimension)]) for c in range(colonies)],
                  buffers=[("keys", 8, 1, keys), ("plane", 2, 3, planes)])
    state = ext.CountedState.from_bytes(data)
    one = state.select_colony(1)
    want_planes = np.ascontiguousarray(values[:, 1, :]).tobytes()
    want = encode(dimension=dimension, colonies=1, tasks=((2, 6),),
                  incumbents=[(2.0, [1, 2, 3, 0])],
                  buffers=[("keys", 8, 1, np.array([22], dtype=np.uint64).tobytes()), ("plane", 2, 3, want_planes)])
    assert one.to_bytes() == want
    assert state.to_bytes() == data                 # the source is unchanged
    for bad in (3, True, -1, 1.0):
        with pytest.raises(ValueError):
            state.select_colony(bad)
