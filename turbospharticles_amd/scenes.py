"""Solver parameters and benchmark scenes.

make_params() fills the SimParams fields the step reads with the reference's derived values
(core/parameters_sim.cpp:1478-1483: m = rho0*h*h*h, r = 2h, r^2 = r*r, all in float32) and the
fixed physics of the benchmark configurations: dt 1e-3, gamma 0.5, omega 0.5, 0.1% average
density error on the absolute criterion with 2..100 solver iterations, nu_F 4e-5, nu_B 8e-5
(configs/params_def.json:13-22).

The scenes are generated on the host by libtsp_host.so (include/tsp_scene.h).
"""
import ctypes

import numpy as np

from . import _capi

TERM_FIXED, TERM_RAW, TERM_POS, TERM_ABS = 0, 1, 2, 3


def make_params(h, *, time_step=1e-3, rest_density=1000.0, viscosity_fluid=4e-5,
                viscosity_boundary=8e-5, gravity=(0.0, -9.81, 0.0), press_gamma=0.5,
                jacobi_omega=0.5, jacobi_warm_start=0.0, density_dev_target=0.001,
                fixed_iters=5, min_solver_iters=2, max_solver_iters=100,
                termination_cond=TERM_ABS, grid_cell_size_ratio=1.0,
                rigid_contact_density_dev_target=0.001, rigid_contact_max_solver_iters=100):
    f32 = np.float32
    h = f32(h)
    rho0 = f32(rest_density)
    p = _capi.TspParams()
    p.time_step = f32(time_step)
    p.particle_size = h
    p.kernel_radius = f32(2.0) * h
    p.kernel_radius_sq = f32(p.kernel_radius) * f32(p.kernel_radius)
    p.rest_density = rho0
    p.mass = rho0 * h * h * h
    p.viscosity_fluid = f32(viscosity_fluid)
    p.viscosity_boundary = f32(viscosity_boundary)
    for a in range(3):
        p.gravity[a] = f32(gravity[a])
    p.press_gamma = f32(press_gamma)
    p.jacobi_omega = f32(jacobi_omega)
    p.jacobi_warm_start = f32(jacobi_warm_start)
    p.density_dev_target = f32(density_dev_target)
    p.grid_cell_size_ratio = f32(grid_cell_size_ratio)
    p.fixed_iters = fixed_iters
    p.min_solver_iters = min_solver_iters
    p.max_solver_iters = max_solver_iters
    p.termination_cond = termination_cond
    p.rigid_contact_density_dev_target = f32(rigid_contact_density_dev_target)
    p.rigid_contact_max_solver_iters = rigid_contact_max_solver_iters
    return p


def dam_break_spec(n_side, h, *, tank=(1.0, 0.5, 0.5), jitter_frac=0.01, jitter_seed=67,
                   counts=None, layers=2):
    """Fluid lattice of n_side^3 (or counts) in the tank corner, wall layers on the faces."""
    s = _capi.TspSceneSpec()
    s.particle_size = h
    s.rest_density = 1000.0
    s.boundary_layers = layers
    for a in range(3):
        s.container_origin[a] = 0.0
        s.container_size[a] = tank[a]
        s.fluid_origin[a] = 0.0
        s.fluid_counts[a] = (counts[a] if counts is not None else n_side)
    s.jitter_frac = jitter_frac
    s.jitter_seed = jitter_seed
    s.rigid_enable = 0
    return s


def rigid_cube_spec(n_side, h, *, side=0.1, density=500.0, center=(0.7, 0.3, 0.25),
                    rotation_seed=7, tank=(1.0, 0.5, 0.5), jitter_frac=0.01, jitter_seed=67):
    """Dam break plus a hollow cube body released at rest with a seeded random rotation."""
    s = dam_break_spec(n_side, h, tank=tank, jitter_frac=jitter_frac, jitter_seed=jitter_seed)
    s.rigid_enable = 1
    for a in range(3):
        s.rigid_center[a] = center[a]
        s.rigid_size[a] = side
        s.rigid_vel[a] = 0.0
        s.rigid_omega[a] = 0.0
    s.rigid_density = density
    q = (ctypes.c_float * 4)()
    _capi.host_lib().tsp_scene_random_quat(rotation_seed, q)
    for a in range(4):
        s.rigid_quat[a] = q[a]
    return s


def build_scene(spec):
    """Generates the scene on the host and returns a HostState with .domain set."""
    lib = _capi.host_lib()
    handle = ctypes.c_void_p()
    rc = lib.tsp_scene_build(ctypes.byref(spec), ctypes.byref(handle))
    if rc != 0:
        raise RuntimeError(f"tsp_scene_build failed ({rc})")
    try:
        counts = (ctypes.c_uint32 * 4)()
        lib.tsp_scene_counts(handle, counts)
        st = _capi.HostState(*counts)
        v = st.view()
        rc = lib.tsp_scene_export(handle, ctypes.byref(v))
        if rc != 0:
            raise RuntimeError(f"tsp_scene_export failed ({rc})")
        mn = (ctypes.c_float * 3)()
        mx = (ctypes.c_float * 3)()
        lib.tsp_scene_domain(handle, mn, mx)
        st.domain = (np.array(mn[:], np.float32), np.array(mx[:], np.float32))
    finally:
        lib.tsp_scene_free(handle)
    return st


def save_state(path, state):
    """Writes a state as the reference's TSPSIM2 restart file (SaveUniversalSimState,
    eval/serial/serialization.cpp:830-895)."""
    _capi.save_state(path, state)


def load_state(path):
    """Reads a TSPSIM2 restart file (LoadUniversalSimState, serialization.cpp:897-935); the
    returned state has no domain (the file holds none)."""
    return _capi.load_state(path)


# --- mixed-size scene sets: each scene's tank scaled to its own fluid block ---------------------

TANK_RATIO = (2.5, 1.25, 1.25)   # tank / fluid block edge, as the 0.4 m block in the 1.0 x 0.5 x 0.5 tank


def scaled_dam_break_spec(n_side, h, jitter_seed=67):
    """n_side^3 fluid lattice at spacing h in a tank of TANK_RATIO x the block."""
    block = n_side * h
    return dam_break_spec(n_side, h, tank=tuple(r * block for r in TANK_RATIO), jitter_seed=jitter_seed)


def scaled_cube_drop_spec(n_side, h, rotation_seed=7, jitter_seed=67):
    """The rigid-cube scene scaled to an n_side^3 block: cube side block/4 and density 500, released
    at rest at (1.75, 0.75, 0.625) x block (the 1M scene's 0.1 m cube at (0.7, 0.3, 0.25))."""
    block = n_side * h
    return rigid_cube_spec(n_side, h, side=0.25 * block, density=500.0,
                           center=(1.75 * block, 0.75 * block, 0.625 * block), rotation_seed=rotation_seed,
                           tank=tuple(r * block for r in TANK_RATIO), jitter_seed=jitter_seed)


def log_uniform_sides(n, lo, hi, seed):
    """n lattice sides round(N^(1/3)) for fluid counts N log-uniform in [lo, hi] (seeded)."""
    rng = np.random.default_rng(seed)
    counts = np.exp(rng.uniform(np.log(lo), np.log(hi), n))
    return [max(2, int(round(c ** (1.0 / 3.0)))) for c in counts]
