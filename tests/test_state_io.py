"""The reference's simulation state file (TSPSIM2) read and written by libtsp_host.

Follows the reference's own state-serialization tests (tests/unit/state_serialization.cpp): the
same mixed state (2 fluid, 1 boundary, 3 rigid particles, 2 bodies, every float field set to
base + 16 * field + i + 0.25, wide ids, signed zero), its 604-byte encoding built here byte by
byte from the documented layout (eval/serial/state_format.h:10-120) independently of the
library, and the same rejections: truncated headers, damaged tags, wrong counts and lengths,
invalid step/time, the invalid-id sentinel, and rigid memberships that do not match the bodies.
"""
import os
import struct

import numpy as np
import pytest

from turbospharticles_amd import _capi as C

INVALID_ID = np.iinfo(np.uint64).max


def make_state(nf=2, nb=1, nr=3, nrb=2):
    """MakeState of tests/unit/state_serialization.cpp:71-101."""
    st = C.HostState(nf, nb, nr, nrb)
    st.step = 0x0102030405060708
    st.time = 12.5
    for group, base in ((st.fluid, 100.0), (st.boundary, 1000.0), (st.rigid, 2000.0), (st.body, 3000.0)):
        for f, k in enumerate(group):
            group[k][:] = (np.float32(base) + (16 * f + np.arange(group[k].size)).astype(np.float32)) + np.float32(0.25)
    st.fluid_id[:] = 5_000_000_000 + 17 * np.arange(nf, dtype=np.uint64)
    st.boundary_id[:] = 6_000_000_000 + 19 * np.arange(nb, dtype=np.uint64)
    st.body_id[:] = 100 + 7 * np.arange(nrb, dtype=np.uint32)
    st.rigid_id[:] = 7_000_000_000 + 23 * np.arange(nr, dtype=np.uint64)
    if nrb:
        st.rigid_body[:] = np.arange(nr, dtype=np.uint32) % nrb
        st.body_count[:] = np.bincount(st.rigid_body, minlength=nrb).astype(np.uint32)
    if nf:
        st.fluid_id[0] = INVALID_ID - 1
        st.fluid["vy"][0] = -123.5
        st.fluid["vz"][0] = -0.0
    return st


def encode(st):
    """The file's bytes from the documented layout, without the library."""
    nf, nb, nr, nrb = st.n
    out = [b"TSPSIM2\0", struct.pack("<IIQdQQQQ", 0x01020304, 64, st.step, st.time, nf, nb, nr, nrb)]
    le = lambda a, t: np.asarray(a, dtype=t).astype(np.dtype(t).newbyteorder("<")).tobytes()
    out.append(le(st.fluid_id, np.uint64))
    out += [le(st.fluid[k], np.float32) for k in C.F_FIELDS]
    out.append(le(st.boundary_id, np.uint64))
    out += [le(st.boundary[k], np.float32) for k in C.B_FIELDS]
    out.append(le(st.rigid_id, np.uint64))
    out.append(le(st.rigid_body, np.uint32))
    out += [le(st.rigid[k], np.float32) for k in C.R_FIELDS]
    out.append(le(st.body_id, np.uint32))
    out.append(le(st.body_count, np.uint32))
    out += [le(st.body[k], np.float32) for k in C.RB_FIELDS]
    return b"".join(out)


def assert_same(a, b):
    assert a.n == b.n
    assert a.step == b.step
    assert struct.pack("<d", a.time) == struct.pack("<d", b.time)
    for k in ("fluid_id", "boundary_id", "rigid_id", "rigid_body", "body_id", "body_count"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k
    for g in ("fluid", "boundary", "rigid", "body"):
        for k in getattr(a, g):
            assert np.array_equal(getattr(a, g)[k].view(np.uint32), getattr(b, g)[k].view(np.uint32)), (g, k)


@pytest.fixture(scope="module", autouse=True)
def _lib():
    C.host_lib()


def test_field_orders_match_the_reference_lists():
    # the independent field lists of tests/unit/state_serialization.cpp:28-51
    assert list(C.F_FIELDS) == ["x", "y", "z", "vx", "vy", "vz", "vpx", "vpy", "vpz", "apx", "apy", "apz",
                                "source", "diag", "pressure", "density"]
    assert len(C.B_FIELDS) == 5 and len(C.R_FIELDS) == 5 and len(C.RB_FIELDS) == 32


def test_writes_the_expected_604_bytes(tmp_path):
    st = make_state()
    expected = encode(st)
    assert len(expected) == 604
    # the header bytes as the reference's test states them
    assert expected[:64] == bytes([ord(c) for c in "TSPSIM2"] + [0, 4, 3, 2, 1, 64, 0, 0, 0, 8, 7, 6, 5, 4, 3, 2, 1,
                                   0, 0, 0, 0, 0, 0, 0x29, 0x40, 2] + [0] * 7 + [1] + [0] * 7 + [3] + [0] * 7
                                  + [2] + [0] * 7)
    p = tmp_path / "state.bin"
    C.save_state(p, st)
    assert p.read_bytes() == expected
    assert not (tmp_path / "state.bin.tmp").exists()


def test_round_trips_every_field(tmp_path):
    st = make_state()
    C.save_state(tmp_path / "s.bin", st)
    assert_same(C.load_state(tmp_path / "s.bin"), st)


def test_loads_an_independently_encoded_state(tmp_path):
    st = make_state()
    (tmp_path / "s.bin").write_bytes(encode(st))
    assert_same(C.load_state(tmp_path / "s.bin"), st)


@pytest.mark.parametrize("counts", [(0, 0, 0, 0), (2, 0, 0, 0), (0, 2, 0, 0), (0, 0, 3, 2), (0, 1, 3, 2),
                                    (2, 0, 3, 2), (2, 1, 0, 0), (5, 3, 4, 3)])
def test_round_trips_empty_populations(tmp_path, counts):
    st = make_state(*counts)
    p = tmp_path / "s.bin"
    C.save_state(p, st)
    nf, nb, nr, nrb = counts
    assert os.path.getsize(p) == 64 + 72 * nf + 28 * nb + 32 * nr + 136 * nrb
    assert p.read_bytes() == encode(st)
    assert_same(C.load_state(p), st)


def _rejects(tmp_path, data):
    p = tmp_path / "bad.bin"
    p.write_bytes(bytes(data))
    with pytest.raises(RuntimeError):
        C.load_state(p)


@pytest.mark.parametrize("length", [0, 1, 63])
def test_rejects_truncated_headers(tmp_path, length):
    _rejects(tmp_path, encode(make_state())[:length])


@pytest.mark.parametrize("offset", [0, 6, 7, 8, 12])
def test_rejects_damaged_header_tags(tmp_path, offset):
    b = bytearray(encode(make_state()))
    b[offset] ^= 0xFF
    _rejects(tmp_path, b)


@pytest.mark.parametrize("offset", [32, 40, 48, 56])
def test_rejects_mismatched_population_counts(tmp_path, offset):
    b = bytearray(encode(make_state()))
    b[offset] += 1
    _rejects(tmp_path, b)


@pytest.mark.parametrize("length", [64, 208, 603, 605])
def test_rejects_incorrect_payload_lengths(tmp_path, length):
    b = bytearray(encode(make_state()))
    b = b[:length] if length <= len(b) else b + bytes(length - len(b))
    _rejects(tmp_path, b)


@pytest.mark.parametrize("time", [-1.0, float("inf"), float("nan")])
def test_rejects_invalid_time(tmp_path, time):
    b = bytearray(encode(make_state()))
    b[24:32] = struct.pack("<d", time)
    _rejects(tmp_path, b)


def test_rejects_the_largest_step(tmp_path):
    b = bytearray(encode(make_state()))
    b[16:24] = struct.pack("<Q", INVALID_ID)
    _rejects(tmp_path, b)


@pytest.mark.parametrize("offset", [64, 208, 236])
def test_rejects_the_invalid_id_sentinel(tmp_path, offset):
    b = bytearray(encode(make_state()))
    b[offset:offset + 8] = struct.pack("<Q", INVALID_ID)
    _rejects(tmp_path, b)


@pytest.mark.parametrize("offset,value", [(260, 2), (340, 1)])
def test_rejects_invalid_rigid_attachments(tmp_path, offset, value):
    b = bytearray(encode(make_state()))
    b[offset] = value
    _rejects(tmp_path, b)


@pytest.mark.parametrize("case", ["fluid_id", "boundary_id", "rigid_id", "body_index", "body_count"])
def test_refuses_to_save_invalid_contents(tmp_path, case):
    st = make_state()
    if case.endswith("_id"):
        getattr(st, case)[0] = INVALID_ID
    elif case == "body_index":
        st.rigid_body[0] = st.n[3]
    else:
        st.body_count[0] -= 1
    p = tmp_path / "s.bin"
    with pytest.raises(RuntimeError):
        C.save_state(p, st)
    assert not p.exists() and not (tmp_path / "s.bin.tmp").exists()


def test_info_reports_counts_step_and_time(tmp_path):
    import ctypes
    st = make_state(5, 3, 4, 3)
    p = tmp_path / "s.bin"
    C.save_state(p, st)
    counts = (ctypes.c_uint32 * 4)()
    step = ctypes.c_uint64()
    time = ctypes.c_double()
    assert C.host_lib().tsp_state_file_info(os.fsencode(p), counts, ctypes.byref(step), ctypes.byref(time)) == 0
    assert tuple(counts) == (5, 3, 4, 3) and step.value == st.step and time.value == st.time


def test_scene_states_round_trip(tmp_path):
    from turbospharticles_amd import scenes
    st = scenes.build_scene(scenes.rigid_cube_spec(8, 0.02, side=0.06, density=500.0, center=(0.5, 0.3, 0.25),
                                                   rotation_seed=3))
    st.step, st.time = 17, 0.0425
    C.save_state(tmp_path / "s.bin", st)
    assert_same(C.load_state(tmp_path / "s.bin"), st)
