"""Export snapshots (Sync::ExportSnapshot) on the CPU side: the oracle's surface classification
pinned by a closed form, and the TSPSNP4 file writers pinned by the reference's format tests.

Classification (IsNearSurfaceFluidParticle, core/backend_sim.h:109-117; cpu_backend.cpp:910-981):
on an exact cubic lattice of spacing h (a power of two, so every pair distance is exact and the
pairs at exactly 2h are neighbors) an interior particle has 32 fluid neighbors, one on a face
at most 22, one a layer below the face 31; with the default threshold 27 the near-surface
particles are exactly the outer layer, and the export carries them and their neighbors: the
three outer layers, n^3 - (n - 6)^3 particles.

Files (eval/serial/snapshot_format.h, tests/unit/snapshot_format.cpp:39-84): a particle data
file of 2 fluid and 3 rigid particles is 156 bytes with the rigid positions at byte 96; a static
boundary file of 5 particles is 149 bytes with the volumes at 124 and the mesh flags at 144.
"""
import struct

import numpy as np
import pytest

from turbospharticles_amd import _capi, scenes

import oracle_lib as O

H = 2.0 ** -7


def lattice_state(n, origin=0.5):
    st = _capi.HostState(n ** 3, 0, 0, 0)
    idx = np.arange(n, dtype=np.float32)
    z, y, x = np.meshgrid(idx, idx, idx, indexing="ij")
    st.fluid["x"][:] = np.float32(origin) + x.ravel() * np.float32(H)
    st.fluid["y"][:] = np.float32(origin) + y.ravel() * np.float32(H)
    st.fluid["z"][:] = np.float32(origin) + z.ravel() * np.float32(H)
    st.fluid["density"][:] = 1000.0
    st.fluid_id[:] = np.arange(n ** 3, dtype=np.uint64)
    st.domain = (np.zeros(3, np.float32), np.ones(3, np.float32))
    return st


def classify(st, threshold=27):
    o = O.Oracle(scenes.make_params(H), threads=4)
    o.pull(st)
    o.neighbor_search()
    flags = o.export_surface(threshold)
    return flags, o.push()


@pytest.mark.parametrize("n", [7, 10])
def test_lattice_export_is_the_three_outer_layers(n):
    flags, s = classify(lattice_state(n))
    layer = np.stack([np.rint((s.fluid[c].astype(np.float64) - 0.5) / H) for c in ("x", "y", "z")], 1)
    depth = np.minimum(layer, n - 1 - layer).min(1)
    assert np.array_equal(flags.astype(bool), depth <= 2)
    assert int(flags.sum()) == n ** 3 - max(n - 6, 0) ** 3


def test_threshold_bounds():
    st = lattice_state(8)
    assert classify(st, threshold=0)[0].sum() == 0            # no particle has fewer than 0 neighbors
    assert classify(st, threshold=33)[0].sum() == 8 ** 3      # every particle has at most 32


def test_particles_touching_walls_are_exported():
    st = scenes.build_scene(scenes.dam_break_spec(12, 0.01))
    o = O.Oracle(scenes.make_params(0.01), threads=4)
    o.pull(st)
    o.neighbor_search()
    flags = o.export_surface(0)   # only the boundary rule can mark a particle
    cfb, _, _ = o.lists(O.FB)
    assert cfb.sum() > 0
    near = cfb > 0
    assert np.all(flags[near] == 1)
    # and the marks spread exactly to their FF neighbors
    cff, iff, _ = o.lists(O.FF)
    off = np.concatenate([[0], np.cumsum(cff.astype(np.int64))])
    expect = near.copy()
    for i in np.nonzero(near)[0]:
        expect[iff[off[i]:off[i + 1]]] = True
    assert np.array_equal(flags.astype(bool), expect)


def _snap(nf, nr, nb=0):
    rng = np.random.default_rng(1)
    f32 = np.float32
    return _capi.ExportSnapshot(step=42, time=0.125, fluid_xyz=rng.random((nf, 3), f32),
                                fluid_vel_mag=rng.random(nf, f32), rigid_xyz=rng.random((nr, 3), f32),
                                rigid_vol=rng.random(nr, f32), rigid_body=np.arange(nr, dtype=np.uint32) % 2,
                                boundary_xyz=rng.random((nb, 3), f32) if nb else None,
                                boundary_vol=rng.random(nb, f32) if nb else None,
                                boundary_is_mesh=(np.arange(nb) % 2).astype(np.uint8) if nb else None)


def test_particle_data_file_layout(tmp_path):
    s = _snap(2, 3)
    p = tmp_path / "particle_data_step_000042.bin"
    scenes.save_snapshot(p, s)
    b = p.read_bytes()
    assert len(b) == 156
    assert b[:8] == b"TSPSNP4\0"
    assert struct.unpack("<IIQdQQQQ", b[8:64]) == (0x01020304, 64, 42, 0.125, 2, 0, 3, 0)
    assert np.array_equal(np.frombuffer(b[64:88], "<f4").reshape(2, 3), s.fluid_xyz)
    assert np.array_equal(np.frombuffer(b[88:96], "<f4"), s.fluid_vel_mag)
    assert np.array_equal(np.frombuffer(b[96:132], "<f4").reshape(3, 3), s.rigid_xyz)
    assert np.array_equal(np.frombuffer(b[132:144], "<f4"), s.rigid_vol)
    assert np.array_equal(np.frombuffer(b[144:156], "<u4"), s.rigid_body)
    assert not (tmp_path / "particle_data_step_000042.bin.tmp").exists()


def test_empty_particle_data_file_is_the_header(tmp_path):
    p = tmp_path / "e.bin"
    scenes.save_snapshot(p, _capi.ExportSnapshot(step=1, time=0.0))
    assert p.stat().st_size == 64


def test_static_boundary_file_layout(tmp_path):
    s = _snap(0, 0, nb=5)
    p = tmp_path / "static_boundary.bin"
    scenes.save_boundary_file(p, s)
    b = p.read_bytes()
    assert len(b) == 149
    assert struct.unpack("<IIQdQQQQ", b[8:64]) == (0x01020304, 64, 42, 0.125, 0, 5, 0, 0)
    assert np.array_equal(np.frombuffer(b[64:124], "<f4").reshape(5, 3), s.boundary_xyz)
    assert np.array_equal(np.frombuffer(b[124:144], "<f4"), s.boundary_vol)
    assert np.array_equal(np.frombuffer(b[144:149], np.uint8), s.boundary_is_mesh)


def test_writers_refuse_missing_arrays(tmp_path):
    s = _snap(2, 0)
    v = s.view()
    v.fluid_vel_mag = None
    assert _capi.host_lib().tsp_snapshot_file_write(str(tmp_path / "x.bin").encode(), v) != 0
    with pytest.raises(RuntimeError):
        scenes.save_boundary_file(tmp_path / "b.bin", _snap(1, 0))   # no boundary in the snapshot
