"""Export snapshots on the device against the oracle: after the same steps the device's surface
set, order, positions and speeds are the oracle's (exact: the fluid state is bit-identical and
the classification reads the same lists), the rigid and boundary arrays are the pushed state's,
a scene exported from a batch equals the same scene alone, and an export before any neighbor
search is refused.
"""
import numpy as np
import pytest

from turbospharticles_amd import scenes
from turbospharticles_amd.backend import HipBackend, HipBatch

import oracle_lib as O
from helpers import cube_scene, dam_break

pytestmark = pytest.mark.gpu

H = 0.01


def params():
    return scenes.make_params(H, termination_cond=scenes.TERM_POS)


def expected(o, threshold):
    flags = o.export_surface(threshold).astype(bool)
    s = o.push()
    xyz = np.stack([s.fluid[c] for c in ("x", "y", "z")], 1)[flags]
    v = [s.fluid[c][flags] for c in ("vx", "vy", "vz")]
    speed = np.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])
    return xyz, speed, s


@pytest.mark.parametrize("case,threshold", [("dam40", 27), ("cube", 27), ("cube", 30)])
def test_export_matches_the_oracle(case, threshold):
    print(f"case: export snapshot, {case}, threshold {threshold}", flush=True)
    st = dam_break(40, H) if case == "dam40" else cube_scene(center=(0.15, 0.04, 0.15), vel=(0.0, -1.0, 0.0),
                                                            fluid=(12, 12, 12))
    p = params()
    g = HipBackend(0)
    g.set_params(p)
    g.pull(st)
    o = O.Oracle(p, threads=4)
    o.pull(st)
    g.step(3)
    for _ in range(3):
        o.step()
    mesh_begin = int(st.boundary_id[st.n[1] // 2]) if st.n[1] else None
    snap = g.export_snapshot(threshold, include_boundary=True, mesh_id_begin=mesh_begin)
    xyz, speed, s = expected(o, threshold)
    assert 0 < snap.fluid_vel_mag.size <= st.n[0]
    if case == "dam40":
        assert snap.fluid_vel_mag.size < st.n[0] // 2   # the interior of the block stays out
    assert np.array_equal(snap.fluid_xyz, xyz)
    assert np.array_equal(snap.fluid_vel_mag, speed.astype(np.float32))
    assert snap.step == 3
    gs = g.push()
    bxyz = np.stack([gs.boundary[c] for c in ("x", "y", "z")], 1)
    assert np.array_equal(snap.boundary_xyz, bxyz)
    assert np.array_equal(snap.boundary_vol, gs.boundary["vol"])
    assert np.array_equal(snap.boundary_is_mesh.astype(bool), gs.boundary_id >= np.uint64(mesh_begin))
    if st.n[2]:
        gx, _ = g.rigid_world()
        assert np.array_equal(snap.rigid_xyz, gx)
        assert np.array_equal(snap.rigid_vol, gs.rigid["vol"])
        assert np.array_equal(snap.rigid_body, gs.rigid_body)
    g.close()


def test_batch_export_equals_alone():
    print("case: export from a batch of three scenes", flush=True)
    p = params()
    states = [dam_break(14, H), cube_scene(center=(0.15, 0.04, 0.15), vel=(0.0, -1.0, 0.0)), dam_break(9, H)]
    alone = []
    for st in states:
        g = HipBackend(0)
        g.set_params(p)
        g.pull(st)
        g.step(4)
        alone.append(g.export_snapshot(27))
        g.close()
    b = HipBatch(0)
    ids = [b.add(p, st) for st in states]
    b.step(4)
    for sid, ref in zip(ids, alone):
        got = b.export_snapshot(sid, 27)
        assert np.array_equal(got.fluid_xyz, ref.fluid_xyz)
        assert np.array_equal(got.fluid_vel_mag, ref.fluid_vel_mag)
        assert np.array_equal(got.rigid_xyz, ref.rigid_xyz)
        assert np.array_equal(got.rigid_body, ref.rigid_body)
    b.close()


def test_export_needs_a_search():
    print("case: export right after pull", flush=True)
    g = HipBackend(0)
    g.set_params(params())
    g.pull(dam_break(8, H))
    with pytest.raises(RuntimeError, match="search"):
        g.export_snapshot()
    g.neighbor_search()
    assert g.export_snapshot().fluid_vel_mag.size > 0
    g.close()
