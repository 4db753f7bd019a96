"""ctypes binding of the oracle (oracle/_build/liboracle.so) for the tests.

The oracle is the CPU restatement of the reference TBB path (oracle/iisph_oracle.h); it is the
checker of the HIP path and is used only by tests/, the smoke check and bench.py's cpu_baseline.
"""
import ctypes
import os
import subprocess

import numpy as np

from turbospharticles_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "_build", "liboracle.so")

FF, FB, FR, RF, RB, RR = range(6)
STAGE_DENSITY, STAGE_NONPRESSURE, STAGE_SOURCE_DIAG, STAGE_PRESSURE_ACC, STAGE_JACOBI, STAGE_INTEGRATION = range(1, 7)

_fp = ctypes.POINTER(ctypes.c_float)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u8p = ctypes.POINTER(ctypes.c_uint8)


class OrcParams(ctypes.Structure):
    _fields_ = _capi.TspParams._fields_


class OrcState(ctypes.Structure):
    _fields_ = [f for f in _capi.TspStateView._fields_ if f[0] not in ("step", "time")]


class OrcInfo(ctypes.Structure):
    _fields_ = [
        ("solver_iters", ctypes.c_uint32),
        ("avg_err_raw", ctypes.c_float), ("avg_err_pos", ctypes.c_float), ("avg_err_abs", ctypes.c_float),
        ("err_sum_raw", ctypes.c_double), ("err_sum_pos", ctypes.c_double), ("err_sum_abs", ctypes.c_double),
        ("frame_orig", ctypes.c_int32 * 3), ("frame_dim", ctypes.c_uint32 * 3),
        ("frame_builds", ctypes.c_uint32),
        ("n_dict_chunks", ctypes.c_uint32), ("n_chunks", ctypes.c_uint32),
        ("n_ff", ctypes.c_uint64), ("n_fb", ctypes.c_uint64), ("n_fr", ctypes.c_uint64),
        ("n_rf", ctypes.c_uint64), ("n_rb", ctypes.c_uint64), ("n_rr", ctypes.c_uint64),
        ("contact_pairs", ctypes.c_uint32),
        ("contact_iters", ctypes.c_uint32), ("contact_avg_err", ctypes.c_float),
        ("contact_converged", ctypes.c_uint32), ("contact_boundary", ctypes.c_uint32),
        ("contacts", ctypes.c_uint64),
    ]


_libs = {}


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib(variant=None):
    """The checker build (variant None), or an optimized build of the same source for the CPU
    baseline: "v3" (-O3, AVX2) or "v4" (-O3, AVX-512)."""
    if variant not in _libs:
        path = ORACLE_LIB if variant is None else os.path.join(ORACLE_DIR, "_build", f"liboracle_{variant}.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        vp = ctypes.c_void_p
        for p in ("orc32", "orc64"):
            sig = {
                "create": (vp, []),
                "destroy": (None, [vp]),
                "set_params": (None, [vp, ctypes.POINTER(OrcParams)]),
                "set_domain": (None, [vp, _fp, _fp]),
                "set_chunk_width": (None, [vp, ctypes.c_uint32]),
                "set_rigid_sums_f64": (None, [vp, ctypes.c_int]),
                "set_threads": (None, [vp, ctypes.c_int]),
                "set_kernel_alpha": (None, [vp, ctypes.c_float]),
                "pull": (ctypes.c_int, [vp, ctypes.POINTER(OrcState)]),
                "set_fluid_positions": (None, [vp, _fp, _fp, _fp]),
                "push": (ctypes.c_int, [vp, ctypes.POINTER(OrcState)]),
                "neighbor_search": (ctypes.c_int, [vp]),
                "step": (ctypes.c_int, [vp]),
                "get_info": (None, [vp, ctypes.POINTER(OrcInfo)]),
                "get_list": (ctypes.c_uint64, [vp, ctypes.c_int, _u32p, _u32p, _u8p, ctypes.c_uint64]),
                "get_order": (None, [vp, _u32p, _u64p, _u32p, _u64p]),
                "get_rigid_world": (None, [vp, _fp, _fp]),
                "get_body_force": (None, [vp, _fp]),
                "kernel_weight": (None, [ctypes.POINTER(OrcParams), _fp, ctypes.c_uint32, _fp, _fp, _fp]),
                "kernels_from_lists": (ctypes.c_int, [vp, _u32p, _u32p, _u8p, _u32p, _u32p, _u32p, _u32p,
                                                      ctypes.c_int]),
                "export_surface": (None, [vp, ctypes.c_uint32, _u8p]),
                "contact_probe": (ctypes.c_int64, [vp, _fp, _fp, _fp, _fp, _fp, _fp, _fp, _fp, _fp, _fp,
                                                   _u32p, _fp]),
            }
            for name, (res, args) in sig.items():
                fn = getattr(L, f"{p}_{name}")
                fn.restype = res
                fn.argtypes = args
        L.orc_jitter3d.restype = ctypes.c_float
        L.orc_jitter3d.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_float]
        L.orc_try_build_cell_frame.restype = ctypes.c_int
        L.orc_try_build_cell_frame.argtypes = [ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                                               ctypes.POINTER(ctypes.c_int32), _u32p]
        L.orc_spread_bits_21to63.restype = ctypes.c_uint64
        L.orc_spread_bits_21to63.argtypes = [ctypes.c_uint32]
        _libs[variant] = L
    return _libs[variant]


def to_orc_params(p):
    o = OrcParams()
    ctypes.memmove(ctypes.byref(o), ctypes.byref(p), ctypes.sizeof(OrcParams))
    return o


class Oracle:
    """One oracle context (prefix 'orc32' for the FP32 restatement, 'orc64' for FP64)."""

    def __init__(self, params, *, precision=32, chunk_width=1, threads=None, rigid_f64=True, alpha=0.0,
                 variant=None):
        self.L = lib(variant)
        self.p = f"orc{precision}"
        self.h = self.fn("create")()
        self.params = to_orc_params(params)
        self.fn("set_params")(self.h, ctypes.byref(self.params))
        self.fn("set_chunk_width")(self.h, chunk_width)
        self.fn("set_rigid_sums_f64")(self.h, 1 if rigid_f64 else 0)
        self.fn("set_threads")(self.h, threads or min(8, os.cpu_count() or 1))
        self.fn("set_kernel_alpha")(self.h, alpha)
        self.n = None

    def fn(self, name):
        return getattr(self.L, f"{self.p}_{name}")

    def __del__(self):
        try:
            self.fn("destroy")(self.h)
        except Exception:
            pass

    def set_domain(self, mn, mx):
        a = (ctypes.c_float * 3)(*mn)
        b = (ctypes.c_float * 3)(*mx)
        self.fn("set_domain")(self.h, a, b)

    def _view(self, st):
        v = st.view()
        o = OrcState()
        for name, _ in OrcState._fields_:
            setattr(o, name, getattr(v, name))
        return o

    def pull(self, st):
        self.n = st.n
        if st.domain is not None:
            self.set_domain(st.domain[0], st.domain[1])
        o = self._view(st)
        rc = self.fn("pull")(self.h, ctypes.byref(o))
        assert rc == 0, rc

    def push(self):
        st = _capi.HostState(*self.n)
        o = self._view(st)
        rc = self.fn("push")(self.h, ctypes.byref(o))
        assert rc == 0, rc
        return st

    def set_fluid_positions(self, x, y, z):
        """Moves the fluid particles (current order) without invalidating the cell frame."""
        a = [np.ascontiguousarray(v, np.float32) for v in (x, y, z)]
        self.fn("set_fluid_positions")(self.h, *(_capi.fptr(v) for v in a))

    def neighbor_search(self):
        assert self.fn("neighbor_search")(self.h) == 0

    def step(self):
        rc = self.fn("step")(self.h)
        assert rc == 0, rc

    def info(self):
        i = OrcInfo()
        self.fn("get_info")(self.h, ctypes.byref(i))
        return i

    def lists(self, rel):
        """(counts, concatenated indices, dict flags or None) in iteration order."""
        n = self.n[0] if rel < 3 else self.n[2]
        counts = np.zeros(n, np.uint32)
        tot = self.fn("get_list")(self.h, rel, _capi.u32ptr(counts), None, None, 0)
        idx = np.zeros(max(tot, 1), np.uint32)
        dict_flags = np.zeros(n, np.uint8) if rel == FF else None
        self.fn("get_list")(self.h, rel, _capi.u32ptr(counts), _capi.u32ptr(idx),
                            None if dict_flags is None else dict_flags.ctypes.data_as(_u8p), tot)
        return counts, idx[:tot], dict_flags

    def order(self):
        nF, nB = self.n[0], self.n[1]
        of = np.zeros(nF, np.uint32)
        kf = np.zeros(nF, np.uint64)
        ob = np.zeros(nB, np.uint32)
        kb = np.zeros(nB, np.uint64)
        self.fn("get_order")(self.h, _capi.u32ptr(of), _capi.u64ptr(kf), _capi.u32ptr(ob), _capi.u64ptr(kb))
        return of, kf, ob, kb

    def rigid_world(self):
        n = self.n[2]
        x = np.zeros(3 * n, np.float32)
        v = np.zeros(3 * n, np.float32)
        self.fn("get_rigid_world")(self.h, _capi.fptr(x), _capi.fptr(v))
        return x.reshape(-1, 3), v.reshape(-1, 3)

    def body_force(self):
        ft = np.zeros(6 * self.n[3], np.float32)
        self.fn("get_body_force")(self.h, _capi.fptr(ft))
        return ft.reshape(-1, 6)

    def export_surface(self, threshold=27):
        """flags (uint8, current particle order) of the fluid particles an export carries"""
        flags = np.zeros(max(self.n[0], 1), np.uint8)
        self.fn("export_surface")(self.h, threshold, flags.ctypes.data_as(_u8p))
        return flags[:self.n[0]]

    def contact_probe(self, p_rigid=None, p_boundary=None):
        """The contact problem after the last neighbor search, evaluated once with the given
        contact pressures: a dict of per-particle arrays (rigid *_r, contact boundary *_b), the
        boundary particles' local indices and the contact loads per body (force, torque)."""
        f = self.fn("contact_probe")
        nCB = f(self.h, None, None, *([None] * 10))
        assert nCB >= 0, nCB
        nR, nRB = self.n[2], self.n[3]
        out = {k: np.zeros(nR, np.float32) for k in ("rho_r", "diag_r", "omega_r", "err_r")}
        out.update({k: np.zeros(max(nCB, 1), np.float32) for k in ("rho_b", "diag_b", "omega_b", "err_b")})
        out["local_b"] = np.zeros(max(nCB, 1), np.uint32)
        out["loads"] = np.zeros(6 * max(nRB, 1), np.float32)
        pr = None if p_rigid is None else np.ascontiguousarray(p_rigid, np.float32)
        pb = None if p_boundary is None else np.ascontiguousarray(p_boundary, np.float32)
        fp = _capi.fptr
        rc = f(self.h, fp(pr), fp(pb), fp(out["rho_r"]), fp(out["diag_r"]), fp(out["omega_r"]), fp(out["err_r"]),
               fp(out["rho_b"]), fp(out["diag_b"]), fp(out["omega_b"]), fp(out["err_b"]),
               _capi.u32ptr(out["local_b"]), fp(out["loads"]))
        assert rc == nCB, rc
        for k in ("rho_b", "diag_b", "omega_b", "err_b", "local_b"):
            out[k] = out[k][:nCB]
        out["loads"] = out["loads"][:6 * nRB].reshape(-1, 6)
        return out

    def kernels_from_lists(self, ff, fb, fr, ff_dict, stage):
        def csr(lists):
            off = np.zeros(len(lists) + 1, np.uint32)
            off[1:] = np.cumsum([len(x) for x in lists])
            flat = np.array([j for x in lists for j in x] or [0], np.uint32)
            return off, flat
        o1, l1 = csr(ff)
        o2, l2 = csr(fb)
        o3, l3 = csr(fr)
        d = np.asarray(ff_dict, np.uint8)
        rc = self.fn("kernels_from_lists")(self.h, _capi.u32ptr(o1), _capi.u32ptr(l1), d.ctypes.data_as(_u8p),
                                           _capi.u32ptr(o2), _capi.u32ptr(l2), _capi.u32ptr(o3),
                                           _capi.u32ptr(l3), stage)
        assert rc == 0


def kernel_weight(params, offsets, precision=32):
    L = lib()
    off = np.ascontiguousarray(offsets, np.float32).reshape(-1, 3)
    n = off.shape[0]
    w = np.zeros(n, np.float32)
    g = np.zeros(3 * n, np.float32)
    wu = np.zeros(n, np.float32)
    op = to_orc_params(params)
    getattr(L, f"orc{precision}_kernel_weight")(ctypes.byref(op), _capi.fptr(off), n, _capi.fptr(w),
                                                _capi.fptr(g), _capi.fptr(wu))
    return w, g.reshape(-1, 3), wu


def try_build_cell_frame(cell_min, cell_max, frame=(0, 0, 0, 0, 0, 0)):
    """orc_try_build_cell_frame: (built, (origX, origY, origZ, dimX, dimY, dimZ)); a rejected
    build returns the frame it was given."""
    L = lib()
    mn = (ctypes.c_int64 * 3)(*cell_min)
    mx = (ctypes.c_int64 * 3)(*cell_max)
    orig = (ctypes.c_int32 * 3)(*frame[:3])
    dim = (ctypes.c_uint32 * 3)(*frame[3:])
    ok = L.orc_try_build_cell_frame(mn, mx, orig, dim)
    return bool(ok), tuple(orig) + tuple(dim)
