"""ctypes binding of the FP64 CPU oracle (oracle/uph_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker and the CPU baseline.  The product
package (uph_amd) never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class Params(C.Structure):
    _fields_ = [("dx", C.c_double), ("h", C.c_double), ("rho0", C.c_double),
                ("gamma", C.c_double), ("c0", C.c_double), ("B", C.c_double),
                ("g", C.c_double), ("gravity", C.c_int), ("alpha_av", C.c_double),
                ("k_spring", C.c_double), ("c_damp", C.c_double), ("dt", C.c_double),
                ("K", C.c_int), ("motion", C.c_int), ("m_amp", C.c_double),
                ("m_omega", C.c_double), ("kernel", C.c_int), ("mode", C.c_int),
                ("kappa", C.c_double), ("mu", C.c_double), ("eps", C.c_double),
                ("lid_u", C.c_double), ("depth", C.c_double), ("flat_bottom", C.c_double),
                ("vtype", C.c_int), ("stress_type", C.c_int)]


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_search.argtypes = [_dp, _ip, _ip, _dp, C.c_int, C.c_int, C.c_int, C.c_int,
                                 C.c_double, C.c_long, C.c_int]
        L.orc_search.restype = C.c_int
        L.orc_search_brute.argtypes = [_dp, _ip, _ip, _dp, C.c_int, C.c_int, C.c_int, C.c_int,
                                       C.c_double, C.c_long]
        L.orc_search_brute.restype = C.c_int
        L.orc_cubic.argtypes = [_dp, C.c_double, C.c_double, _dp, _dp]
        L.orc_cubic.restype = None
        L.orc_poly7.argtypes = [_dp, C.c_double, C.c_double, _dp, _dp]
        L.orc_poly7.restype = None
        L.orc_create.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(Params), _dp, _dp, _dp,
                                 _dp, _dp, _ip, _dp, C.c_int]
        L.orc_create.restype = C.c_void_p
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_step.argtypes = [C.c_void_p]
        L.orc_step.restype = C.c_int
        L.orc_run.argtypes = [C.c_void_p, C.c_int]
        L.orc_run.restype = C.c_int
        L.orc_derivatives.argtypes = [C.c_void_p, C.c_double]
        L.orc_derivatives.restype = C.c_int
        L.orc_get_state.argtypes = [C.c_void_p] + [C.c_void_p] * 5
        L.orc_get_rates.argtypes = [C.c_void_p] + [C.c_void_p] * 5
        L.orc_get_ref.argtypes = [C.c_void_p] + [C.c_void_p] * 7
        L.orc_get_stress.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_get_stress.restype = None
        L.orc_set_gate_y.argtypes = [C.c_void_p, _dp]
        L.orc_set_local.argtypes = [C.c_void_p] + [C.c_void_p] * 3
        L.orc_get_ghosts.argtypes = [C.c_void_p] + [C.c_void_p] * 8
        L.orc_get_ghosts.restype = C.c_int
        L.orc_ghost_capacity.argtypes = [C.c_void_p]
        L.orc_ghost_capacity.restype = C.c_int
        L.orc_ghost_sections.argtypes = [C.c_int, C.c_void_p]
        L.orc_ghost_sections.restype = C.c_int
        L.orc_niac.argtypes = [C.c_void_p]
        L.orc_niac.restype = C.c_int
        L.orc_itime.argtypes = [C.c_void_p]
        L.orc_itime.restype = C.c_int
        _lib = L
    return _lib


def cubic(dx, dy, r, h):
    """(W, dW/dx, dW/dy) of the cubic spline at separation (dx, dy), |.| = r."""
    w = np.zeros(1)
    dw = np.zeros(2)
    lib().orc_cubic(np.array([dx, dy], dtype=np.float64), float(r), float(h), w, dw)
    return w[0], dw[0], dw[1]


def poly7(dx, dy, r, h):
    """(W, dW/dx, dW/dy) of the PolyOrderSeven kernel pair the reference evaluates
    (derivatives.c:767-823) at separation (dx, dy), |.| = r."""
    w = np.zeros(1)
    dw = np.zeros(2)
    lib().orc_poly7(np.array([dx, dy], dtype=np.float64), float(r), float(h), w, dw)
    return w[0], dw[0], dw[1]


def search(pos, nf, n, pb, K, dist, brute=False, cap=None, nthreads=0):
    """The SearchNeighbors contract: returns (i_pair, j_pair, rij2) in emission order."""
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
    cap = cap if cap is not None else max(1, nf * K)
    ip = np.zeros(cap, np.int32)
    jp = np.zeros(cap, np.int32)
    r2 = np.zeros(cap)
    if brute:
        m = lib().orc_search_brute(pos, ip, jp, r2, nf, n, pb, K, dist, cap)
    else:
        m = lib().orc_search(pos, ip, jp, r2, nf, n, pb, K, dist, cap, nthreads)
    if m < 0:
        raise RuntimeError("oracle search: pair capacity exceeded")
    return ip[:m].copy(), jp[:m].copy(), r2[:m].copy()


def params_of(sc) -> Params:
    return Params(sc.dx, sc.h, sc.rho0, sc.gamma, sc.c0, sc.B, sc.g, int(sc.gravity),
                  sc.alpha_av, sc.k_spring, sc.c_damp, sc.dt, int(sc.K), int(sc.motion),
                  sc.m_amp, sc.m_omega, int(getattr(sc, "kernel", 0)), int(getattr(sc, "mode", 0)),
                  float(getattr(sc, "kappa", 0.0)), float(getattr(sc, "mu", 0.0)),
                  float(getattr(sc, "eps", 0.0)), float(getattr(sc, "lid_u", 0.0)),
                  float(getattr(sc, "depth", 0.0)), float(getattr(sc, "flat_bottom", 0.0)),
                  int(getattr(sc, "vtype", 0)), int(getattr(sc, "stress_type", 1)))


class Sim:
    """FP64 config-mode run of one scene, stepped with the reference's leapfrog."""

    def __init__(self, sc, nthreads=1):
        self.sc = sc
        self.n = sc.n
        self._p = params_of(sc)
        f = lambda a: np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        self._keep = [f(sc.pos), f(sc.vel), f(sc.rho), f(sc.pressure), f(sc.mass),
                      np.ascontiguousarray(sc.itype, dtype=np.int32), f(sc.lvel)]
        self.h = lib().orc_create(sc.n_fluid, sc.n_fixed, sc.n_moving, C.byref(self._p),
                                  *self._keep, int(nthreads))
        if getattr(sc, "gate_y", None) is not None:
            self._gate = f(sc.gate_y)
            lib().orc_set_gate_y(self.h, self._gate)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def step(self, n=1):
        if lib().orc_run(self.h, int(n)) != 0:
            raise RuntimeError("oracle step failed (pair capacity)")

    def derivatives(self, t):
        if lib().orc_derivatives(self.h, float(t)) != 0:
            raise RuntimeError("oracle derivatives failed")

    @property
    def itime(self):
        return lib().orc_itime(self.h)

    @property
    def niac(self):
        return lib().orc_niac(self.h)

    def state(self):
        n = self.n
        x = np.zeros((n, 2)); v = np.zeros((n, 2)); rho = np.zeros(n)
        vl = np.zeros((n, 2)); rl = np.zeros(n)
        lib().orc_get_state(self.h, x.ctypes.data, v.ctypes.data, rho.ctypes.data,
                            vl.ctypes.data, rl.ctypes.data)
        return dict(x=x, v=v, rho=rho, vloc=vl, rholoc=rl)

    def rates(self):
        n = self.n
        R = np.zeros(n); Rt = np.zeros(n); a = np.zeros((n, 2)); at = np.zeros((n, 2))
        P = np.zeros(n)
        lib().orc_get_rates(self.h, R.ctypes.data, Rt.ctypes.data, a.ctypes.data,
                            at.ctypes.data, P.ctypes.data)
        return dict(R=R, Rt=Rt, a=a, at=at, P=P)

    def ref(self):
        """Ref-mode fields: DynamicPressure P, localPressure Ploc, dpdt RP, dpdt_approx
        RPt, fluctuations drho, dP (n) and dv (n, 2)."""
        n = self.n
        out = {k: np.zeros(n) for k in ("P", "Ploc", "RP", "RPt", "drho", "dP")}
        out["dv"] = np.zeros((n, 2))
        lib().orc_get_ref(self.h, *[out[k].ctypes.data for k in
                                    ("P", "Ploc", "RP", "RPt", "drho", "dP", "dv")])
        return out

    def stress(self):
        """ViscousStress xx, xy, yx, yy of the last derivatives() call, (n, 4)
        (ClassicShearStress, ref mode)."""
        out = np.zeros((self.n, 4))
        lib().orc_get_stress(self.h, out.ctypes.data)
        return out

    def ghosts(self):
        """Mirror rows of the last derivatives() call (derivatives.c:3193-4192): x, v,
        vloc (ng, 2) and rho, rholoc, P, Ploc, mass (ng), in the reference's order."""
        cap = max(lib().orc_ghost_capacity(self.h), 1)
        out = {k: np.zeros((cap, 2)) for k in ("x", "v", "vloc")}
        out.update({k: np.zeros(cap) for k in ("rho", "rholoc", "P", "Ploc", "mass")})
        ng = lib().orc_get_ghosts(self.h, *[out[k].ctypes.data for k in
                                            ("x", "v", "vloc", "rho", "rholoc", "P", "Ploc", "mass")])
        return {k: a[:ng].copy() for k, a in out.items()}

    def set_local(self, rholoc=None, Ploc=None, vloc=None):
        """Sets localDensity / localPressure / localVelocity before a derivatives() call."""
        f = lambda a, m: None if a is None else np.ascontiguousarray(a, np.float64).reshape(-1)
        self._loc = [f(rholoc, 1), f(Ploc, 1), f(vloc, 2)]
        lib().orc_set_local(self.h, *[None if a is None else a.ctypes.data for a in self._loc])
