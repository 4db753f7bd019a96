"""Shared helpers of the tests: scenes, oracle runs and comparisons."""
import numpy as np

from turbospharticles_amd import scenes

import oracle_lib as O

FIELDS_STEP = ["x", "y", "z", "vx", "vy", "vz", "density", "pressure", "source", "diag", "vpx", "vpy", "vpz",
               "apx", "apy", "apz"]


def dam_break(n_side=40, h=0.01, **kw):
    return scenes.build_scene(scenes.dam_break_spec(n_side, h, **kw))


def rel_err(a, b):
    """max |a - b| / max |b| (the relative error of a field, normwise)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.max(np.abs(b)) if b.size else 0.0
    if den == 0.0:
        return float(np.max(np.abs(a - b))) if a.size else 0.0
    return float(np.max(np.abs(a - b)) / den)


def split_lists(counts, idx):
    off = np.concatenate([[0], np.cumsum(counts.astype(np.int64))])
    return [idx[off[i]:off[i + 1]] for i in range(len(counts))]


def cube_scene(h=0.01, side=0.06, center=(0.15, 0.04, 0.15), vel=(0.0, 0.0, 0.0), omega=(0.0, 0.0, 0.0),
               quat=(1.0, 0.0, 0.0, 0.0), tank=(0.3, 0.3, 0.3), fluid=(4, 4, 4), density=500.0):
    """A hollow cube body in a small tank with a small fluid block in the far corner; placed
    near a wall it is a rigid-boundary contact case."""
    spec = scenes.rigid_cube_spec(fluid[0], h, side=side, density=density, center=center, tank=tank)
    for a in range(3):
        spec.fluid_counts[a] = fluid[a]
        spec.rigid_vel[a] = vel[a]
        spec.rigid_omega[a] = omega[a]
    for a in range(4):
        spec.rigid_quat[a] = quat[a]
    return scenes.build_scene(spec)


def merge_bodies(a, b):
    """State a with the bodies of state b appended (b's rigid particles follow a's and name the
    bodies after a's); fluid and boundary are a's."""
    from turbospharticles_amd import _capi
    nF, nB, nRa, nRBa = a.n
    nRb, nRBb = b.n[2], b.n[3]
    st = _capi.HostState(nF, nB, nRa + nRb, nRBa + nRBb)
    st.fluid_id[:] = a.fluid_id
    st.boundary_id[:] = a.boundary_id
    for k in a.fluid:
        st.fluid[k][:] = a.fluid[k]
    for k in a.boundary:
        st.boundary[k][:] = a.boundary[k]
    st.rigid_id[:] = np.concatenate([a.rigid_id, b.rigid_id + np.uint64(10_000_000)])
    st.rigid_body[:] = np.concatenate([a.rigid_body, b.rigid_body + np.uint32(nRBa)])
    for k in a.rigid:
        st.rigid[k][:] = np.concatenate([a.rigid[k], b.rigid[k]])
    st.body_id[:] = np.concatenate([a.body_id, b.body_id + np.uint32(nRBa)])
    st.body_count[:] = np.concatenate([a.body_count, b.body_count])
    for k in a.body:
        st.body[k][:] = np.concatenate([a.body[k], b.body[k]])
    st.domain = a.domain
    return st


def state_digest(st):
    """SHA-256 of a state's fluid (ids, then every field in name order) and body records."""
    import hashlib
    h = hashlib.sha256(np.ascontiguousarray(st.fluid_id).tobytes())
    for f in sorted(st.fluid):
        h.update(np.ascontiguousarray(st.fluid[f]).tobytes())
    for f in sorted(st.body):
        h.update(np.ascontiguousarray(st.body[f]).tobytes())
    return h.hexdigest()


def drop_record(st, params):
    """Per-step figures of the 1M cube drop record: mean density error (rho / rho0 - 1 averaged
    over the fluid, cpu_backend.cpp:1509-1514), kinetic energy 1/2 m sum |v|^2 and the body's
    centre of mass, all in float64 from the float32 state."""
    rho = st.fluid["density"].astype(np.float64)
    v2 = sum(st.fluid[c].astype(np.float64) ** 2 for c in ("vx", "vy", "vz"))
    return {"mean_density_error": float(np.mean(rho / np.float64(params.rest_density) - 1.0)),
            "kinetic_energy": float(0.5 * np.float64(params.mass) * np.sum(v2)),
            "xcm": [float(st.body[c][0]) for c in ("xcm", "ycm", "zcm")]}
