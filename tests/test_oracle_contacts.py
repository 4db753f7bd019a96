"""The oracle's rigid contact solver (rigid_solver.cpp:23-799) pinned by what the mathematics
fixes, since the reference's tests hold no contact case:

- the diagonal is the derivative of a particle's own residual with respect to its own contact
  pressure: with one particle's pressure set to d and every other pressure zero, the residual
  changes by -d * diag (exactly linear; checked in FP64 through the force, body-sum, velocity
  and residual code, which the diagonal formula does not share);
- pairwise contact forces between two bodies conserve linear and angular momentum for any
  pressures (the kernel gradient is radial and antisymmetric);
- the relaxation factors of a body's contacting particles sum to the base factor 0.5;
- a cube falling onto the floor is stopped: its contact loads push it up, it comes to rest
  above the wall, and every step that converged meets the density-error target.
"""
import numpy as np
import pytest

from turbospharticles_amd import scenes

import oracle_lib as O
from helpers import cube_scene, merge_bodies

H = 0.01


def params(**kw):
    return scenes.make_params(H, termination_cond=scenes.TERM_POS, **kw)


def pressed_cube():
    # bottom face 0.5 h above the floor's first layer's reach: in contact from the first step
    return cube_scene(center=(0.15, 0.028, 0.15), vel=(0.0, -1.0, 0.0), omega=(0.0, 0.0, 3.0))


def two_cubes():
    a = cube_scene(center=(0.12, 0.15, 0.15), vel=(0.5, 0.0, 0.0), omega=(0.0, 2.0, 0.0))
    b = cube_scene(center=(0.1855, 0.155, 0.152), vel=(-0.5, 0.0, 0.0), omega=(1.0, 0.0, 0.0),
                   quat=(0.99, 0.1, 0.05, 0.0))
    return merge_bodies(a, b)


def probe_oracle(st, precision=64):
    o = O.Oracle(params(), precision=precision, threads=4)
    o.pull(st)
    o.neighbor_search()
    return o


@pytest.mark.parametrize("case", ["wall", "two_bodies"])
def test_diagonal_is_the_derivative_of_the_own_residual(case):
    st = pressed_cube() if case == "wall" else two_cubes()
    o = probe_oracle(st)
    base = o.contact_probe()
    assert base["diag_r"].max() > 0
    if case == "wall":
        assert base["diag_b"].size and base["diag_b"].max() > 0
    nR, nCB = st.n[2], base["diag_b"].size
    rng = np.random.default_rng(3)
    rig = rng.choice(np.nonzero(base["diag_r"] > 0)[0], 12, replace=False)
    bnd = rng.choice(np.nonzero(base["diag_b"] > 0)[0], 8, replace=False) if nCB else []
    d = 1.0e6   # the residual is linear in the pressures; a large step keeps the float32 outputs exact enough
    for i in rig:
        pr = np.zeros(nR, np.float32)
        pr[i] = d
        out = o.contact_probe(pr, None)
        slope = (base["err_r"][i] - out["err_r"][i]) / d
        assert slope == pytest.approx(base["diag_r"][i], rel=1e-5), i
    for i in bnd:
        pb = np.zeros(nCB, np.float32)
        pb[i] = d
        out = o.contact_probe(None, pb)
        slope = (base["err_b"][i] - out["err_b"][i]) / d
        assert slope == pytest.approx(base["diag_b"][i], rel=1e-5), i


def test_rigid_rigid_contact_forces_conserve_momentum():
    st = two_cubes()
    o = probe_oracle(st)
    base = o.contact_probe()
    assert o.info().n_rr > 0 and base["diag_b"].size == 0
    rng = np.random.default_rng(11)
    pr = rng.uniform(0.0, 200.0, st.n[2]).astype(np.float32)
    loads = o.contact_probe(pr, None)["loads"].astype(np.float64)
    f, tau = loads[:, :3], loads[:, 3:]
    scale = np.abs(f).sum()
    assert scale > 0
    assert np.abs(f.sum(0)).max() <= 1e-6 * scale
    xcm = np.stack([st.body[c] for c in ("xcm", "ycm", "zcm")], 1).astype(np.float64)
    ang = (tau + np.cross(xcm, f)).sum(0)
    arm = np.abs(xcm).max() + 0.05
    assert np.abs(ang).max() <= 1e-6 * scale * arm


def test_relaxation_factors_share_the_base_factor_per_body():
    st = two_cubes()
    o = probe_oracle(st)
    out = o.contact_probe()
    for b in range(st.n[3]):
        om = out["omega_r"][st.rigid_body == b]
        if np.any(om > 0):
            assert om[om > 0].sum() == pytest.approx(0.5, rel=1e-6)
    st = pressed_cube()
    o = probe_oracle(st)
    out = o.contact_probe()
    touching = out["omega_b"][out["omega_b"] > 0]
    # one movable body: each touching wall particle divides 0.5 by that body's wall contacts
    n = np.count_nonzero((out["omega_b"] > 0) & (out["diag_b"] > 0))
    assert touching.size and np.allclose(touching[out["diag_b"][out["omega_b"] > 0] > 0], 0.5 / n)


@pytest.mark.parametrize("precision", [32, 64])
def test_falling_cube_comes_to_rest_on_the_floor(precision):
    st = cube_scene(center=(0.15, 0.035, 0.15), vel=(0.0, -1.0, 0.0))
    p = params()
    o = O.Oracle(p, precision=precision, threads=4)
    o.pull(st)
    vy, ycm, solved = [], [], 0
    for _ in range(30):
        o.step()
        i = o.info()
        s = o.push()
        vy.append(float(s.body["vy"][0]))
        ycm.append(float(s.body["ycm"][0]))
        if i.contact_iters > 0:
            solved += 1
        if i.contact_converged:
            assert i.contact_avg_err <= p.rigid_contact_density_dev_target
        assert i.contact_iters <= p.rigid_contact_max_solver_iters
    assert solved >= 2
    # without contacts the body would fall 30 mm in 30 steps; it stops within a few mm
    assert min(ycm) > 0.035 - 0.012
    assert abs(vy[-1]) < 0.1
    half = 0.03
    assert min(ycm) - half > -0.5 * H   # the bottom face stays above the wall's surface layer


def test_contact_parameters_are_checked_when_contacts_exist():
    st = pressed_cube()
    o = O.Oracle(params(rigid_contact_max_solver_iters=0), threads=4)
    o.pull(st)
    assert o.fn("step")(o.h) == -3


def test_fp32_and_fp64_oracles_agree_through_a_landing():
    st = cube_scene(center=(0.15, 0.035, 0.15), vel=(0.0, -1.0, 0.0))
    out = {}
    for prec in (32, 64):
        o = O.Oracle(params(), precision=prec, threads=4)
        o.pull(st)
        for _ in range(12):
            o.step()
        out[prec] = o.push().body
    for k in ("xcm", "ycm", "zcm", "vx", "vy", "vz"):
        assert abs(float(out[32][k][0]) - float(out[64][k][0])) <= 2e-4 * (1.0 + abs(float(out[64][k][0]))), k
