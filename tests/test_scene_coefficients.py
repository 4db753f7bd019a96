"""Boundary and rigid coefficients of the scene builder (ComputeParticleCoefficients,
scene/builder.cpp:174-376): V = 1 / sum_j W(x_i - x_j) over the particles of the same build region
within the support (the particle itself included), psi = rho0 V, the cubic spline with
alpha = 1 / (4 pi h^3).

Pinned two ways that do not share the builder's code:
  - a closed form: a wall particle far from the container's edges sees an ideal slab of L layers
    of the cubic lattice at spacing h, so its sum is (1 / (4 pi h^3)) * sum over lattice offsets
    (a, b, c) with c in the layers of w(|(a, b, c)|), w(q) = (2 - q)^3_+ - 4 (1 - q)^3_+,
    summed in FP64;
  - an FP64 sum over the f32 positions the builder wrote, for every particle of the container
    (faces, edges and corners) and of a rotated hollow cube.
"""
import math

import numpy as np
import pytest
from scipy.spatial import cKDTree

from turbospharticles_amd import scenes

H = 0.01


def w(q):
    return max(2.0 - q, 0.0) ** 3 - 4.0 * max(1.0 - q, 0.0) ** 3


def slab_sum(layers, k):
    """sum of w over the lattice offsets of an infinite slab of `layers` layers, seen from layer k"""
    s = 0.0
    for c in range(-k, layers - k):
        for a in range(-2, 3):
            for b in range(-2, 3):
                s += w(math.sqrt(a * a + b * b + c * c))
    return s


def test_slab_sum_of_three_layers_middle_is_four_pi():
    """The middle layer of three sees the whole 3x3x3-cell neighbourhood of a cubic lattice, whose
    cubic-spline sum is 4 pi (to 1e-3: the kernel's lattice sum, not an identity), so V = h^3."""
    assert abs(slab_sum(3, 1) - 4.0 * math.pi) < 1e-3 * 4.0 * math.pi


@pytest.mark.parametrize("layers", [2, 3])
def test_interior_wall_particles_match_the_slab_closed_form(layers):
    st = scenes.build_scene(scenes.dam_break_spec(10, H, tank=(0.3, 0.2, 0.2), layers=layers))
    b = st.boundary
    p = np.stack([b["x"], b["y"], b["z"]], 1).astype(np.float64)
    alpha = 1.0 / (4.0 * math.pi * H ** 3)
    for k in range(layers):
        # the floor's layer k lies at y = -(k + 0.5) h; its particle nearest (0.15, ., 0.1) is
        # 0.1 m from every other face, and its neighbourhood is the ideal lattice slab
        target = np.array([0.15, -(k + 0.5) * H, 0.1])
        i = int(np.argmin(((p - target) ** 2).sum(1)))
        assert abs(p[i, 1] - target[1]) < 1e-6
        d = (p - p[i]) / H
        near = d[(np.abs(d) < 2.5).all(1)]
        assert np.abs(near - np.round(near)).max() < 1e-3
        assert len(near) == 25 * layers
        vol = 1.0 / (alpha * slab_sum(layers, k))
        assert abs(b["vol"][i] - vol) <= 2e-6 * vol, (k, b["vol"][i], vol)
        assert abs(b["psi"][i] - 1000.0 * vol) <= 2e-6 * 1000.0 * vol, k


def _fp64_coefficients(pos, h):
    alpha = 1.0 / (4.0 * math.pi * h ** 3)
    r = 2.0 * h
    tree = cKDTree(pos)
    vol = np.zeros(len(pos))
    for i, nb in enumerate(tree.query_ball_point(pos, r * (1 + 1e-6))):
        d = np.sqrt(((pos[nb] - pos[i]) ** 2).sum(1)) / h
        vol[i] = 1.0 / (alpha * sum(w(q) for q in d))
    return vol


def test_every_container_particle_matches_an_fp64_sum():
    """All 12,798 wall particles of a small container (faces, edges and corners, 3 layers)."""
    st = scenes.build_scene(scenes.dam_break_spec(10, H, tank=(0.3, 0.2, 0.2), layers=3))
    b = st.boundary
    pos = np.stack([b["x"], b["y"], b["z"]], 1).astype(np.float64)
    vol = _fp64_coefficients(pos, H)
    rel = np.abs(b["vol"].astype(np.float64) - vol) / vol
    assert rel.max() < 5e-6, rel.max()
    assert np.allclose(b["psi"], 1000.0 * b["vol"], rtol=1e-6, atol=0)
    assert vol.max() / vol.min() > 1.5   # corners and edges see fewer neighbours than faces


def test_rotated_cube_particles_match_an_fp64_sum():
    """The hollow cube's particles in its body frame (rotation does not change distances beyond
    rounding; the builder sums in world positions)."""
    st = scenes.build_scene(scenes.rigid_cube_spec(10, H, side=0.06, center=(0.2, 0.1, 0.1), tank=(0.4, 0.2, 0.2)))
    r = st.rigid
    pos = np.stack([r["xrel"], r["yrel"], r["zrel"]], 1).astype(np.float64)
    vol = _fp64_coefficients(pos, H)
    rel = np.abs(r["vol"].astype(np.float64) - vol) / vol
    assert rel.max() < 5e-5, rel.max()
    assert np.allclose(r["psi"], 1000.0 * r["vol"], rtol=1e-6, atol=0)
