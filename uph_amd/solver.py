"""Host-side mirror of the reference's solver seam over libuph_gpu.so (ctypes).

`search_neighbors` is the reference's SearchNeighbors (src/lib/Search.cc:5)
on numpy arrays; `GpuSolver` runs the leapfrog of src/bin/sph.cc:90-168 with
derivatives() (src/lib/derivatives.c:3137) on the GPU for one scene or a batch
of independent scenes.  There is no CPU fallback: if the HIP library is not
built, loading fails with an error.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

import numpy as np

from .scenes import Scene

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libuph_gpu.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class UphScene(C.Structure):
    _fields_ = [("n_fluid", C.c_int), ("n_fixed", C.c_int), ("n_moving", C.c_int),
                ("pos", _dp), ("vel", _dp), ("rho", _dp), ("mass", _dp), ("itype", _ip),
                ("lvel", _dp), ("dx", C.c_double), ("h", C.c_double), ("rho0", C.c_double),
                ("gamma", C.c_double), ("c0", C.c_double), ("B", C.c_double), ("g", C.c_double),
                ("gravity", C.c_int), ("alpha_av", C.c_double), ("k_spring", C.c_double),
                ("c_damp", C.c_double), ("dt", C.c_double), ("K", C.c_int), ("motion", C.c_int),
                ("m_amp", C.c_double), ("m_omega", C.c_double), ("box", C.c_double * 4),
                ("kernel", C.c_int), ("pres", _dp), ("mode", C.c_int), ("kappa", C.c_double),
                ("mu", C.c_double), ("eps", C.c_double), ("lid_u", C.c_double), ("depth", C.c_double),
                ("flat_bottom", C.c_double), ("gate_y", _dp), ("vtype", C.c_int),
                ("lbp", C.c_double), ("rbp", C.c_double), ("stress_type", C.c_int)]


class UphDiag(C.Structure):
    _fields_ = [("kinetic_energy", C.c_double), ("potential_energy", C.c_double),
                ("probe_pressure", C.c_double), ("probe_norm", C.c_double),
                ("probe_neighbors", C.c_int), ("consistency_one", C.c_double),
                ("consistency_two", C.c_double), ("consistency_three", C.c_double),
                ("density_uncertainty", C.c_double)]


class UphOpts(C.Structure):
    _fields_ = [("use_graph", C.c_int), ("keep_rates", C.c_int), ("pair_variant", C.c_int),
                ("pair_block", C.c_int), ("cell_factor", C.c_double), ("sort_every_step", C.c_int),
                ("skin", C.c_double)]


# every entry point of include/uph.h: name -> (restype, argtypes)
SIGNATURES = {
    "SearchNeighbors": (C.c_int, [C.POINTER(_dp), _ip, _ip, _dp, C.c_int, C.c_int, C.c_int,
                                  C.c_int, C.c_double]),
    "uph_search_neighbors_cap": (C.c_int, [_dp, _ip, _ip, _dp, C.c_int, C.c_int, C.c_int, C.c_int,
                                           C.c_double, C.c_long]),
    "uph_search_neighbors_noalloc": (C.c_int, [_dp, _ip, _ip, _dp, C.c_int, C.c_int, C.c_int, C.c_int,
                                               C.c_double, C.c_long]),
    "uph_search_last_error": (C.c_char_p, []),
    "uph_create": (C.c_int, [C.POINTER(UphScene), C.c_int, C.POINTER(UphOpts),
                             C.POINTER(C.c_void_p)]),
    "uph_destroy": (None, [C.c_void_p]),
    "uph_step": (C.c_int, [C.c_void_p, C.c_int]),
    "uph_step_persistent": (C.c_int, [C.c_void_p, C.c_int]),
    "uph_sync": (C.c_int, [C.c_void_p]),
    "uph_get_state": (C.c_int, [C.c_void_p, C.c_int, _dp, _dp, _dp]),
    "uph_get_rates": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_float),
                                C.POINTER(C.c_float), C.POINTER(C.c_float),
                                C.POINTER(C.c_float)]),
    "uph_get_ref_rates": (C.c_int, [C.c_void_p, C.c_int] + [C.POINTER(C.c_float)] * 5),
    "uph_get_pressure": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double)]),
    "uph_get_ghosts": (C.c_int, [C.c_void_p, C.c_int, C.c_int] + [C.POINTER(C.c_double)] * 8),
    "uph_ghost_capacity": (C.c_int, [C.c_void_p, C.c_int]),
    "uph_derivatives_scheme": (C.c_int, [C.c_int, C.c_int]),
    "uph_itime": (C.c_int, [C.c_void_p]),
    "uph_resorts": (C.c_int, [C.c_void_p]),
    "uph_skin": (C.c_double, [C.c_void_p]),
    "uph_time_steps": (C.c_int, [C.c_void_p, C.c_int, _dp, _dp]),
    "uph_time_phases": (C.c_int, [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp]),
    "uph_last_error": (C.c_char_p, [C.c_void_p]),
    # surrogate kernels (pointers are device addresses passed as integers)
    "uph_dev_alloc": (C.c_int, [C.POINTER(C.c_void_p), C.c_size_t]),
    "uph_dev_free": (C.c_int, [C.c_void_p]),
    "uph_dev_upload": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "uph_dev_download": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "uph_dev_copy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "uph_dev_zero": (C.c_int, [C.c_void_p, C.c_size_t]),
    "uph_dev_sync": (C.c_int, []),
    "uph_gnn_last_error": (C.c_char_p, []),
    "uph_gemm_bf16": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_void_p,
                                C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int,
                                C.c_void_p, C.c_int, C.c_void_p]),
    "uph_cast_f32_bf16": (C.c_int, [C.c_long, C.c_void_p, C.c_void_p, C.c_void_p]),
    "uph_cast_bf16_f32": (C.c_int, [C.c_long, C.c_void_p, C.c_void_p, C.c_void_p]),
    "uph_relu_bwd": (C.c_int, [C.c_long, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "uph_add_f32": (C.c_int, [C.c_long, C.c_void_p, C.c_void_p, C.c_void_p]),
    "uph_colsum_f32": (C.c_int, [C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                 C.c_void_p]),
    "uph_segsum": (C.c_int, [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int,
                             C.c_void_p, C.c_int, C.c_int, C.c_void_p]),
    "uph_gather_rows": (C.c_int, [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                  C.c_int, C.c_int, C.c_void_p]),
    "uph_adam": (C.c_int, [C.c_long, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                           C.c_float, C.c_float, C.c_float, C.c_float, C.c_int, C.c_void_p]),
    "uph_mse": (C.c_int, [C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                          C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    "uph_transpose_bf16": (C.c_int, [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    "uph_gather_add2": (C.c_int, [C.c_long, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_int, C.c_void_p, C.c_int, C.c_void_p]),
    "uph_relu_to_bf16": (C.c_int, [C.c_long, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                   C.c_void_p]),
    "uph_rows_to_bf16": (C.c_int, [C.c_long, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                   C.c_void_p]),
    "uph_fill_col_bf16": (C.c_int, [C.c_long, C.c_int, C.c_int, C.c_float, C.c_void_p, C.c_void_p]),
    "uph_relu_bwd_rows": (C.c_int, [C.c_long, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                    C.c_void_p, C.c_int, C.c_void_p]),
    "uph_add_rows_f32": (C.c_int, [C.c_long, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                   C.c_void_p]),
    "uph_gns_edge_step": (C.c_int, [C.c_long, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_void_p, C.c_void_p]),
    "uph_lincomb_f64": (C.c_int, [C.c_long, C.c_double, C.c_void_p, C.c_double, C.c_void_p,
                                  C.c_void_p, C.c_void_p]),
    "uph_scale_ratio_f64": (C.c_int, [C.c_long, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p]),
    "uph_gns_node_feats": (C.c_int, [C.c_int, C.c_int, C.c_void_p, C.c_float, C.c_void_p, C.c_void_p]),
    "uph_gns_edge_feats": (C.c_int, [C.c_long, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                                     C.c_void_p, C.c_void_p]),
    "uph_gns_integrate": (C.c_int, [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_double,
                                    C.c_double, C.c_void_p]),
    "uph_sph_density": (C.c_int, [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                                  C.c_void_p, C.c_void_p]),
    "uph_sn_edge_agg": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_double, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                                  C.c_int, C.c_void_p, C.c_int, C.c_void_p]),
    "uph_sn_edge_bwd": (C.c_int, [C.c_int, C.c_int, C.c_int] + [C.c_void_p] * 6 +
                        [C.c_double, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                         C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                         C.c_void_p, C.c_void_p]),
    "uph_sn_fused": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double] +
                     [C.c_void_p] * 19),
    "uph_sn_last_error": (C.c_char_p, []),
    "uph_state_to_device": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    "uph_state_from_device": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_int]),
    "uph_graph_build": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_void_p, C.c_void_p,
                                  C.c_void_p, C.c_void_p, C.c_void_p, C.c_long,
                                  C.POINTER(C.c_long)]),
    "uph_graph_last_error": (C.c_char_p, []),
    "uph_derivatives": (C.c_int, [C.c_void_p, C.c_double]),
    "uph_diagnostics": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_void_p]),
    "uph_probe_two": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_double, _dp,
                                C.POINTER(C.c_int)]),
    "uph_vorticity": (C.c_int, [C.c_void_p, C.c_int, _dp]),
    "derivatives": (None, [C.c_void_p, C.c_double]),
    "uph_derivatives_last_error": (C.c_char_p, []),
}

_lib = None


def lib():
    """The loaded HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libuph_gpu.so is not built ({LIB_PATH}); run "
                               "`python -c 'import __graft_entry__ as g; g.build()'`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _ptr(a: np.ndarray, ct=C.c_double):
    return a.ctypes.data_as(C.POINTER(ct))


def search_neighbors(pos: np.ndarray, n_fluid: int, n_particles: int, particle_boundary: int,
                     n_neighbors: int, dist: float, capacity: int | None = None):
    """SearchNeighbors (Search.cc:5-44) on the GPU: (i_pair, j_pair, rij2) arrays."""
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 2)
    cap = int(capacity) if capacity is not None else max(1, n_fluid * n_neighbors)
    ip = np.zeros(cap, np.int32)
    jp = np.zeros(cap, np.int32)
    r2 = np.zeros(cap)
    m = lib().uph_search_neighbors_cap(_ptr(pos), _ptr(ip, C.c_int), _ptr(jp, C.c_int), _ptr(r2),
                                       int(n_fluid), int(n_particles), int(particle_boundary),
                                       int(n_neighbors), float(dist), cap)
    if m < 0:
        raise RuntimeError("SearchNeighbors failed: " + lib().uph_search_last_error().decode())
    return ip[:m].copy(), jp[:m].copy(), r2[:m].copy()


def search_neighbors_rows(pos: np.ndarray, i_pair: np.ndarray, j_pair: np.ndarray,
                          rij2: np.ndarray, n_fluid: int, n_particles: int,
                          particle_boundary: int, n_neighbors: int, dist: float) -> int:
    """The exact reference signature: Position passed as a double** row table."""
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 2)
    rows = (_dp * len(pos))(*[C.cast(pos[i:].ctypes.data, _dp) for i in range(len(pos))])
    m = lib().SearchNeighbors(rows, _ptr(i_pair, C.c_int), _ptr(j_pair, C.c_int), _ptr(rij2),
                              int(n_fluid), int(n_particles), int(particle_boundary),
                              int(n_neighbors), float(dist))
    if m < 0:
        raise RuntimeError("SearchNeighbors failed: " + lib().uph_search_last_error().decode())
    return m


class GpuSolver:
    """Device-resident leapfrog over a batch of independent scenes.

    sort_every_step: None (default) re-sorts lazily for handles of at most 16 scenes
    and every step for larger batches; True re-sorts every step; False re-sorts lazily
    whatever the batch size.  skin: the lazy re-sort's cell growth (0: 0.02 for one
    scene, 0.05 for a batch)."""

    def __init__(self, scenes: Sequence[Scene] | Scene, use_graph: bool = False,
                 keep_rates: bool = False, pair_variant: int = 0, pair_block: int = 256,
                 cell_factor: float = 1.0, sort_every_step: Optional[bool] = None, skin: float = 0.0):
        if isinstance(scenes, Scene):
            scenes = [scenes]
        self.scenes = list(scenes)
        arr = (UphScene * len(self.scenes))()
        self._keep = []
        for k, sc in enumerate(self.scenes):
            pos = np.ascontiguousarray(sc.pos, dtype=np.float64).reshape(-1)
            vel = np.ascontiguousarray(sc.vel, dtype=np.float64).reshape(-1)
            rho = np.ascontiguousarray(sc.rho, dtype=np.float64)
            mass = np.ascontiguousarray(sc.mass, dtype=np.float64)
            ity = np.ascontiguousarray(sc.itype, dtype=np.int32)
            lvel = np.ascontiguousarray(sc.lvel, dtype=np.float64).reshape(-1)
            pres = np.ascontiguousarray(sc.pressure, dtype=np.float64)
            self._keep += [pos, vel, rho, mass, ity, lvel, pres]
            u = arr[k]
            u.n_fluid, u.n_fixed, u.n_moving = sc.n_fluid, sc.n_fixed, sc.n_moving
            u.pos, u.vel, u.rho, u.mass = _ptr(pos), _ptr(vel), _ptr(rho), _ptr(mass)
            u.itype, u.lvel = _ptr(ity, C.c_int), _ptr(lvel)
            u.dx, u.h, u.rho0, u.gamma, u.c0, u.B, u.g = (sc.dx, sc.h, sc.rho0, sc.gamma, sc.c0,
                                                          sc.B, sc.g)
            u.gravity, u.alpha_av, u.k_spring, u.c_damp = (int(sc.gravity), sc.alpha_av,
                                                           sc.k_spring, sc.c_damp)
            u.dt, u.K, u.motion, u.m_amp, u.m_omega = (sc.dt, int(sc.K), int(sc.motion),
                                                       sc.m_amp, sc.m_omega)
            for q in range(4):
                u.box[q] = float(sc.box[q])
            u.kernel = int(sc.kernel)
            u.pres = _ptr(pres)
            u.mode = int(sc.mode)
            u.kappa, u.mu, u.eps = float(sc.kappa), float(sc.mu), float(sc.eps)
            u.lid_u, u.depth, u.flat_bottom = float(sc.lid_u), float(sc.depth), float(sc.flat_bottom)
            if sc.gate_y is not None:
                gy = np.ascontiguousarray(sc.gate_y, dtype=np.float64)
                self._keep.append(gy)
                u.gate_y = _ptr(gy)
            u.vtype = int(getattr(sc, "vtype", 0))
            u.lbp, u.rbp = 0.0, 0.0
            u.stress_type = int(getattr(sc, "stress_type", 1))
        opts = UphOpts(int(bool(use_graph)), int(bool(keep_rates)), int(pair_variant),
                       int(pair_block), float(cell_factor),
                       0 if sort_every_step is None else (1 if sort_every_step else -1), float(skin))
        self.opts = dict(use_graph=bool(use_graph), pair_variant=int(pair_variant),
                         pair_block=int(pair_block), cell_factor=float(cell_factor),
                         sort_every_step=sort_every_step, skin=float(skin))
        h = C.c_void_p()
        rc = lib().uph_create(arr, len(self.scenes), C.byref(opts), C.byref(h))
        if rc != 0:
            raise RuntimeError("uph_create failed: " + lib().uph_last_error(None).decode())
        self.h = h
        self._keep = None

    def close(self):
        if getattr(self, "h", None):
            lib().uph_destroy(self.h)
            self.h = None

    __del__ = close

    def _check(self, rc):
        if rc < 0:
            raise RuntimeError(lib().uph_last_error(self.h).decode())
        return rc

    def step(self, n: int = 1):
        self._check(lib().uph_step(self.h, int(n)))

    def step_persistent(self, n: int = 1):
        """n steps in one persistent cooperative launch (uph_step_persistent)."""
        self._check(lib().uph_step_persistent(self.h, int(n)))

    def sync(self) -> int:
        """Waits; returns how many steps took the capped (more than K neighbours) path."""
        return self._check(lib().uph_sync(self.h))

    @property
    def itime(self) -> int:
        return lib().uph_itime(self.h)

    def skin(self) -> float:
        """The cells' growth of a lazy re-sort handle; 0.0 for one that re-sorts every step."""
        return lib().uph_skin(self.h)

    def resorts(self) -> int:
        """Re-sorts of a lazy handle so far (waits); -1 for a handle that sorts every step."""
        return lib().uph_resorts(self.h)

    def _scene(self, scene: int):
        if not 0 <= scene < len(self.scenes):
            raise IndexError(f"scene {scene} out of range (batch of {len(self.scenes)})")
        return self.scenes[scene]

    def state(self, scene: int = 0):
        sc = self._scene(scene)
        x = np.zeros((sc.n, 2)); v = np.zeros((sc.n, 2)); rho = np.zeros(sc.n)
        self._check(lib().uph_get_state(self.h, scene, _ptr(x), _ptr(v), _ptr(rho)))
        return dict(x=x, v=v, rho=rho)

    def rates(self, scene: int = 0):
        nf = self._scene(scene).n_fluid
        R = np.zeros(nf, np.float32); Rt = np.zeros(nf, np.float32)
        a = np.zeros((nf, 2), np.float32)
        at = np.zeros((nf, 2), np.float32)
        self._check(lib().uph_get_rates(self.h, scene, _ptr(R, C.c_float), _ptr(Rt, C.c_float),
                                        _ptr(a, C.c_float), _ptr(at, C.c_float)))
        return dict(R=R, Rt=Rt, a=a, at=at)

    def ref_rates(self, scene: int = 0):
        """Ref-mode rates of the last step: dpdt, dpdt_approx, the density, pressure and
        velocity fluctuations (fluid rows, FP32)."""
        nf = self._scene(scene).n_fluid
        out = {k: np.zeros(nf, np.float32) for k in ("RP", "RPt", "drho", "dP")}
        out["dv"] = np.zeros((nf, 2), np.float32)
        self._check(lib().uph_get_ref_rates(self.h, scene, *[_ptr(out[k], C.c_float) for k in
                                                             ("RP", "RPt", "drho", "dP", "dv")]))
        return out

    def pressure(self, scene: int = 0):
        """DynamicPressure of every particle in reference order (ref mode: the integrated
        state; config mode: the Tait pressure of the density)."""
        P = np.zeros(self._scene(scene).n)
        self._check(lib().uph_get_pressure(self.h, scene, _ptr(P)))
        return P

    def ghosts(self, scene: int = 0):
        """Mirror rows of the last step (derivatives.c:3193-4192): x, v, vloc (ng, 2) and
        rho, rholoc, P, Ploc, mass (ng), in the reference's order."""
        self._scene(scene)
        cap = max(int(lib().uph_ghost_capacity(self.h, scene)), 1)
        out = {k: np.zeros((cap, 2)) for k in ("x", "v", "vloc")}
        out.update({k: np.zeros(cap) for k in ("rho", "rholoc", "P", "Ploc", "mass")})
        ng = lib().uph_get_ghosts(self.h, scene, cap, *[_ptr(out[k]) for k in
                                                         ("x", "v", "vloc", "rho", "rholoc", "P", "Ploc", "mass")])
        self._check(0 if ng >= 0 else -1)
        return {k: a[:ng].copy() for k, a in out.items()}

    def diagnostics(self, probe=(0.0, 0.0), scene: int = 0) -> dict:
        """Kinetic/potential energy, the pressure probe and the density consistency values
        of the scene's current fluid state (derivatives.c:2254-2338, 3090-3117)."""
        self._scene(scene)
        d = UphDiag()
        self._check(lib().uph_diagnostics(self.h, scene, float(probe[0]), float(probe[1]),
                                          C.byref(d)))
        return {k: getattr(d, k) for k, _ in UphDiag._fields_}

    def probe_two(self, x: float, y: float, scene: int = 0):
        """(mean positive pressure, count) of the fluid rows within 2h of (x, y)
        (derivatives.c:2340-2393)."""
        self._scene(scene)
        p, n = C.c_double(), C.c_int()
        self._check(lib().uph_probe_two(self.h, scene, float(x), float(y), C.byref(p), C.byref(n)))
        return p.value, n.value

    def vorticity(self, scene: int = 0) -> np.ndarray:
        """approxVorticity of the fluid rows as the last step evaluated it (derivatives.c:2840-2885)."""
        out = np.zeros(self._scene(scene).n_fluid)
        self._check(lib().uph_vorticity(self.h, scene, _ptr(out)))
        return out

    def state_to_device(self, pos: int, vel: int, rho: int, scene: int = 0):
        """Copies the scene's state, in reference order, into device buffers (pointers)."""
        self._scene(scene)
        self._check(lib().uph_state_to_device(self.h, scene, pos, vel, rho))

    def state_from_device(self, pos: int, vel: int, rho: int, restart: bool = True,
                          scene: int = 0):
        """Loads the fluid rows of (pos, vel, rho) from device buffers (pointers)."""
        self._scene(scene)
        self._check(lib().uph_state_from_device(self.h, scene, pos, vel, rho, int(restart)))

    def time_steps(self, n: int):
        """(ms for the whole loop, mean ms of the fused pair kernel), HIP events."""
        tot = C.c_double(); pair = C.c_double()
        self._check(lib().uph_time_steps(self.h, int(n), C.byref(tot), C.byref(pair)))
        return tot.value, pair.value

    def time_phases(self, n: int):
        """Run n steps timed by HIP events: {"total": ms for the whole loop, then the mean
        ms per step of "sort" (binning, scan, reorder), "pair" (the compacted pair pass)
        and "exact" (the capped pass, empty unless a query had more than K in range)}."""
        v = [C.c_double() for _ in range(4)]
        self._check(lib().uph_time_phases(self.h, int(n), *[C.byref(x) for x in v]))
        return dict(zip(("total", "sort", "pair", "exact"), (x.value for x in v)))

    def time_loop(self, n: int) -> float:
        """ms for n steps timed by two HIP events around the loop, none between kernels."""
        tot = C.c_double()
        self._check(lib().uph_time_phases(self.h, int(n), C.byref(tot), None, None, None))
        return tot.value
