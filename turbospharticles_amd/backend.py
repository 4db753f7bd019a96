"""Python mirror of Sim::I_SimBackend (core/backend_sim.h:194-261) over the C ABI of libtsp_hip.so.

HipBackend is what tests, the smoke check and bench.py drive; the C++ SimBackend_HIP
(host/sim_backend_hip.h) is the same adapter for C++ callers of the reference interface.
Every method maps to one tsp_* entry point and raises RuntimeError with tsp_last_error() on
failure, as the reference backends throw std::runtime_error (GPU/gpu_utils.h:24-57).
"""
import ctypes

import numpy as np

from . import _capi


class TspError(RuntimeError):
    """A failed tsp_* call: `code` is its tsp_status (TSP_ERR_CAPACITY = -4 for a count beyond
    limit, e.g. a context whose scenes' cell frames exceed its cell limit)."""

    def __init__(self, what, code, msg):
        super().__init__(f"{what} failed ({code}): {msg}")
        self.code = code


class HipBackend:
    def __init__(self, device=0):
        self.lib = _capi.hip_lib()
        h = ctypes.c_void_p()
        rc = self.lib.tsp_create(device, ctypes.byref(h))
        if rc != 0:
            raise RuntimeError(f"tsp_create failed ({rc}): no HIP device {device}")
        self.h = h
        self.n = None

    def close(self):
        if self.h:
            self.lib.tsp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            raise TspError(what, rc, self.lib.tsp_last_error(self.h).decode())

    # I_SimBackend::Name
    def name(self):
        return self.lib.tsp_name().decode()

    def set_params(self, params):
        self._check(self.lib.tsp_set_params(self.h, ctypes.byref(params)), "tsp_set_params")

    def set_step_graph(self, enable):
        """Steps as captured graphs or launch by launch (tsp_set_step_graph)."""
        self._check(self.lib.tsp_set_step_graph(self.h, 1 if enable else 0), "tsp_set_step_graph")

    def set_cell_limit(self, max_cells):
        """The most cells the context's cell frames may hold (tsp_set_cell_limit)."""
        self._check(self.lib.tsp_set_cell_limit(self.h, int(max_cells)), "tsp_set_cell_limit")

    # I_SimBackend::SetSimDomainBounds
    def set_domain(self, mn, mx):
        a = (ctypes.c_float * 3)(*[float(v) for v in mn])
        b = (ctypes.c_float * 3)(*[float(v) for v in mx])
        self._check(self.lib.tsp_set_domain(self.h, a, b), "tsp_set_domain")

    # I_SimBackend::PullFromUniversalSimState
    def pull(self, state):
        if state.domain is not None:
            self.set_domain(*state.domain)
        v = state.view()
        self._check(self.lib.tsp_pull(self.h, ctypes.byref(v)), "tsp_pull")
        self.n = state.n

    def pull_spec(self, spec):
        """The scene of a tsp_scene_spec generated on the device (as scenes.build_scene + pull)"""
        self._check(self.lib.tsp_pull_spec(self.h, ctypes.byref(spec)), "tsp_pull_spec")
        c = (ctypes.c_uint32 * 4)()
        self._check(self.lib.tsp_get_counts(self.h, c), "tsp_get_counts")
        self.n = tuple(c)

    # I_SimBackend::PushToUniversalSimState
    def push(self):
        st = _capi.HostState(*self.n)
        v = st.view()
        self._check(self.lib.tsp_push(self.h, ctypes.byref(v)), "tsp_push")
        st.step, st.time = v.step, v.time
        return st

    # I_SimBackend::ExecuteNeighborSearch
    def neighbor_search(self):
        self._check(self.lib.tsp_neighbor_search(self.h), "tsp_neighbor_search")

    # I_SimBackend::ExecuteFullSimStep
    def step(self, n_steps=1):
        self._check(self.lib.tsp_step(self.h, n_steps), "tsp_step")

    def info(self):
        i = _capi.TspStepInfo()
        self._check(self.lib.tsp_get_step_info(self.h, ctypes.byref(i)), "tsp_get_step_info")
        return i

    # I_SimBackend::PackExportSnapshot
    def export_snapshot(self, surface_nbr_count_threshold=27, include_boundary=False, mesh_id_begin=None):
        """The surface fluid particles (and rigid, optionally boundary, particles) as an
        ExportSnapshot, classified and compacted on the device."""
        mesh = 0xFFFFFFFFFFFFFFFF if mesh_id_begin is None else int(mesh_id_begin)
        nF, nB, nR, _ = self.n
        rc, snap = _capi.export_from(
            lambda v: self.lib.tsp_export_snapshot(self.h, surface_nbr_count_threshold, mesh, ctypes.byref(v)),
            max(nF // 4, 1024), nB, nR, include_boundary)
        self._check(rc, "tsp_export_snapshot")
        return snap

    # I_SimBackend::GetNumFluidParticles / Boundary / Rigid (+ bodies)
    def counts(self):
        c = (ctypes.c_uint32 * 4)()
        self._check(self.lib.tsp_get_counts(self.h, c), "tsp_get_counts")
        return tuple(c)

    def neighbors(self, rel):
        """(counts, concatenated indices) of relation rel (0 FF .. 5 RR) in iteration order."""
        n = self.n[0] if rel < 3 else self.n[2]
        counts = np.zeros(n, np.uint32)
        total = ctypes.c_uint64()
        self._check(self.lib.tsp_get_neighbors(self.h, rel, _capi.u32ptr(counts), None, 0, ctypes.byref(total)),
                    "tsp_get_neighbors")
        idx = np.zeros(max(total.value, 1), np.uint32)
        self._check(self.lib.tsp_get_neighbors(self.h, rel, _capi.u32ptr(counts), _capi.u32ptr(idx), total.value,
                                               ctypes.byref(total)), "tsp_get_neighbors")
        return counts, idx[:total.value]

    def sort_state(self):
        nF, nB = self.n[0], self.n[1]
        of = np.zeros(nF, np.uint32)
        kf = np.zeros(nF, np.uint64)
        kb = np.zeros(nB, np.uint64)
        self._check(self.lib.tsp_get_sort(self.h, _capi.u32ptr(of), _capi.u64ptr(kf), None, _capi.u64ptr(kb)),
                    "tsp_get_sort")
        return of, kf, kb

    def rigid_world(self):
        n = self.n[2]
        x = np.zeros(3 * max(n, 1), np.float32)
        v = np.zeros(3 * max(n, 1), np.float32)
        self._check(self.lib.tsp_get_rigid_world(self.h, _capi.fptr(x), _capi.fptr(v)), "tsp_get_rigid_world")
        return x[:3 * n].reshape(-1, 3), v[:3 * n].reshape(-1, 3)

    def profile_enable(self, on=True):
        self._check(self.lib.tsp_profile_enable(self.h, 1 if on else 0), "tsp_profile_enable")

    def profile_read(self, cap=64):
        names = ctypes.create_string_buffer(48 * cap)
        launches = np.zeros(cap, np.uint64)
        ms = np.zeros(cap, np.float64)
        n = self.lib.tsp_profile_read(self.h, names, _capi.u64ptr(launches),
                                      ms.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), cap)
        if n < 0:
            self._check(n, "tsp_profile_read")
        out = {}
        for q in range(min(n, cap)):
            nm = names.raw[48 * q:48 * (q + 1)].split(b"\0", 1)[0].decode()
            out[nm] = (int(launches[q]), float(ms[q]))
        return out

    def synchronize(self):
        self._check(self.lib.tsp_synchronize(self.h), "tsp_synchronize")

    # I_SimBackend::FillMemoryRecord (device bytes by what they store)
    def memory(self):
        m = _capi.TspMemory()
        self._check(self.lib.tsp_get_memory(self.h, ctypes.byref(m)), "tsp_get_memory")
        return m

    # I_SimBackend::FillStatisticsRecord
    def statistics(self):
        st = _capi.TspStatistics()
        self._check(self.lib.tsp_get_statistics(self.h, ctypes.byref(st)), "tsp_get_statistics")
        return st


class HipBatch:
    """Many independent scenes in one context, advanced together (tsp_batch_* of libtsp_hip.so).

    Each scene keeps its own parameters, domain, frame and solver state; a scene's results are
    bit-identical to the same scene run alone in a HipBackend. Scenes are named by the ids the
    context assigns when they are added.
    """

    def __init__(self, device=0):
        self.lib = _capi.hip_lib()
        h = ctypes.c_void_p()
        rc = self.lib.tsp_create(device, ctypes.byref(h))
        if rc != 0:
            raise RuntimeError(f"tsp_create failed ({rc}): no HIP device {device}")
        self.h = h
        self.counts = {}

    def close(self):
        if self.h:
            self.lib.tsp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            raise TspError(what, rc, self.lib.tsp_last_error(self.h).decode())

    def set_step_graph(self, enable):
        """Steps as captured graphs or launch by launch (tsp_set_step_graph)."""
        self._check(self.lib.tsp_set_step_graph(self.h, 1 if enable else 0), "tsp_set_step_graph")

    def set_cell_limit(self, max_cells):
        """The most cells the cell frames of all resident scenes may hold together
        (tsp_set_cell_limit); an admission beyond it is refused with TSP_ERR_CAPACITY."""
        self._check(self.lib.tsp_set_cell_limit(self.h, int(max_cells)), "tsp_set_cell_limit")

    def update(self, add=(), remove=()):
        """Removes the scenes `remove` (ids), then appends `add`, a sequence of (params, state)
        pairs (state.domain, when set, is the scene's domain); returns the new ids."""
        add = list(add)
        rm = np.asarray(list(remove), np.uint32)
        n = len(add)
        params = (_capi.TspParams * max(n, 1))()
        views = [st.view() for _, st in add]
        vptrs = (ctypes.POINTER(_capi.TspStateView) * max(n, 1))()
        # a scene without a domain gets an empty one (min > max), which the context treats the same
        dom = np.tile(np.array([np.inf] * 3 + [-np.inf] * 3, np.float32), max(n, 1))
        for q, (p, st) in enumerate(add):
            params[q] = p
            vptrs[q] = ctypes.pointer(views[q])
            if st.domain is not None:
                dom[6 * q:6 * q + 3] = st.domain[0]
                dom[6 * q + 3:6 * q + 6] = st.domain[1]
        ids = np.zeros(max(n, 1), np.uint32)
        rc = self.lib.tsp_batch_update(self.h, len(rm), _capi.u32ptr(rm) if len(rm) else None, n, params, vptrs,
                                       _capi.fptr(dom), _capi.u32ptr(ids))
        self._check_update(rc, "tsp_batch_update")
        for r in rm:
            self.counts.pop(int(r), None)
        out = [int(v) for v in ids[:n]]
        for sid, (_, st) in zip(out, add):
            self.counts[sid] = st.n
        return out

    def _check_update(self, rc, what):
        """A refused update adds no scene, but the removals it was given may have taken effect
        (a refused admission keeps only the scenes that stay): the scene table follows the
        context's ids before the error is raised."""
        if rc == 0:
            return
        err = TspError(what, rc, self.lib.tsp_last_error(self.h).decode())
        live = set(self.ids())
        for sid in list(self.counts):
            if sid not in live:
                self.counts.pop(sid)
        raise err

    def add(self, params, state):
        return self.update(add=[(params, state)])[0]

    def update_specs(self, remove=(), add=()):
        """One relayout: scenes `remove` leave, scenes `add` = [(params, tsp_scene_spec)] are
        generated on the device and join; returns the new ids."""
        rm = np.array([int(v) for v in remove], np.uint32)
        n = len(add)
        params = (_capi.TspParams * max(n, 1))()
        specs = [sp for _, sp in add]
        sptrs = (ctypes.POINTER(_capi.TspSceneSpec) * max(n, 1))()
        for q, (p, sp) in enumerate(add):
            params[q] = p
            sptrs[q] = ctypes.pointer(sp)
        ids = np.zeros(max(n, 1), np.uint32)
        rc = self.lib.tsp_batch_update_specs(self.h, len(rm), _capi.u32ptr(rm) if len(rm) else None, n, params,
                                             sptrs, _capi.u32ptr(ids))
        self._check_update(rc, "tsp_batch_update_specs")
        for r in rm:
            self.counts.pop(int(r), None)
        out = [int(v) for v in ids[:n]]
        for sid, sp in zip(out, specs):
            self.counts[sid] = spec_counts(sp)
        return out

    def add_spec(self, params, spec):
        """a scene of a tsp_scene_spec generated on the device; returns its id"""
        sid = ctypes.c_uint32()
        self._check(self.lib.tsp_batch_add_scene_spec(self.h, ctypes.byref(params), ctypes.byref(spec), ctypes.byref(sid)),
                    "tsp_batch_add_scene_spec")
        c = (ctypes.c_uint32 * 4)()
        self._check(self.lib.tsp_batch_scene_counts(self.h, sid.value, c), "tsp_batch_scene_counts")
        self.counts[sid.value] = tuple(c)
        return sid.value

    def remove(self, scene_id):
        self.update(remove=[scene_id])

    def ids(self):
        n = ctypes.c_uint32()
        self._check(self.lib.tsp_batch_scene_ids(self.h, None, 0, ctypes.byref(n)), "tsp_batch_scene_ids")
        buf = np.zeros(max(n.value, 1), np.uint32)
        self._check(self.lib.tsp_batch_scene_ids(self.h, _capi.u32ptr(buf), n.value, ctypes.byref(n)),
                    "tsp_batch_scene_ids")
        return [int(v) for v in buf[:n.value]]

    def step(self, n_steps=1):
        self._check(self.lib.tsp_batch_step(self.h, n_steps), "tsp_batch_step")

    def push(self, scene_id):
        st = _capi.HostState(*self.counts[scene_id])
        v = st.view()
        self._check(self.lib.tsp_batch_push_scene(self.h, scene_id, ctypes.byref(v)), "tsp_batch_push_scene")
        st.step, st.time = v.step, v.time
        return st

    def info(self, scene_id):
        i = _capi.TspStepInfo()
        self._check(self.lib.tsp_batch_scene_info(self.h, scene_id, ctypes.byref(i)), "tsp_batch_scene_info")
        return i

    def synchronize(self):
        self._check(self.lib.tsp_synchronize(self.h), "tsp_synchronize")

    def export_snapshot(self, scene_id, surface_nbr_count_threshold=27, include_boundary=False, mesh_id_begin=None):
        """HipBackend.export_snapshot of one scene."""
        mesh = 0xFFFFFFFFFFFFFFFF if mesh_id_begin is None else int(mesh_id_begin)
        nF, nB, nR, _ = self.counts[scene_id]
        rc, snap = _capi.export_from(
            lambda v: self.lib.tsp_batch_export_snapshot(self.h, scene_id, surface_nbr_count_threshold, mesh,
                                                         ctypes.byref(v)),
            max(nF // 4, 1024), nB, nR, include_boundary)
        self._check(rc, "tsp_batch_export_snapshot")
        return snap

    def statistics(self, scene_id):
        st = _capi.TspStatistics()
        self._check(self.lib.tsp_batch_scene_statistics(self.h, scene_id, ctypes.byref(st)), "tsp_batch_scene_statistics")
        return st

    def profile_enable(self, on=True):
        self._check(self.lib.tsp_profile_enable(self.h, 1 if on else 0), "tsp_profile_enable")

    # the per-launch-group times (HipBackend.profile_read)
    profile_read = HipBackend.profile_read

    @staticmethod
    def footprint(n):
        """device bytes a scene with counts n = (fluid, boundary, rigid, bodies) holds while resident"""
        return int(_capi.hip_lib().tsp_batch_footprint(*[int(v) for v in n]))


def spec_counts(spec):
    """(fluid, boundary, rigid, bodies) that a tsp_scene_spec builds"""
    c = (ctypes.c_uint32 * 4)()
    rc = _capi.host_lib().tsp_scene_spec_counts(ctypes.byref(spec), c)
    if rc != 0:
        raise RuntimeError(f"tsp_scene_spec_counts failed ({rc})")
    return tuple(int(v) for v in c)
