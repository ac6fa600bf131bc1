"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so); test-side only."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

import disco_amd as D

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
_L = None


def lib():
    global _L
    if _L is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} missing: run `make -C oracle oracle`")
        L = C.CDLL(LIB)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int)
        L.oracle_create.restype = C.c_void_p
        L.oracle_create.argtypes = [C.POINTER(D.Params)] + [C.c_int] * 6 + [ip, dp, dp, dp, dp, dp, dp, dp, dp,
                                                                           C.c_double]
        L.oracle_getmindt.restype = C.c_double
        L.oracle_getmindt.argtypes = [C.c_void_p]
        L.oracle_timestep.argtypes = [C.c_void_p, C.c_double]
        L.oracle_get_state.argtypes = [C.c_void_p, dp, dp, dp, dp, dp, dp, dp]
        L.oracle_destroy.argtypes = [C.c_void_p]
        L.oracle_track_faces.argtypes = [dp, dp, C.c_int, dp, dp, C.c_int, C.c_double, ip, ip, ip, dp, dp]
        L.oracle_track_faces.restype = C.c_int
        _L = L
    return _L


def _d(a):
    return None if a is None else np.ascontiguousarray(a, np.float64).ctypes.data_as(C.POINTER(C.c_double))


class Oracle:
    def __init__(self, p: D.Params, grid: dict, state: dict, planet_phi=None, t: float = 0.0):
        self.p = p.copy()
        self.grid = {k: np.ascontiguousarray(v) for k, v in grid.items()}
        g = self.grid
        self.ncell = int(g["Np"].sum())
        keep = [np.ascontiguousarray(state[k], np.float64).ravel() for k in ("prim", "cons", "piph", "dphi",
                                                                              "wiph")]
        pp = None if planet_phi is None else np.ascontiguousarray(planet_phi, np.float64)
        Np = np.ascontiguousarray(g["Np"], np.int32)
        self._h = lib().oracle_create(C.byref(self.p), *[int(np.asarray(g[k]).reshape(-1)[0]) for k in ("Nr", "Nz", "NgRa", "NgRb", "NgZa")],
                                      int(np.asarray(g["NgZb"]).reshape(-1)[0]), Np.ctypes.data_as(C.POINTER(C.c_int)), _d(g["r_jph"]),
                                      _d(g["z_kph"]), *[_d(a) for a in keep], _d(pp), float(t))

    def getmindt(self) -> float:
        return lib().oracle_getmindt(self._h)

    def timestep(self, dt: float) -> None:
        lib().oracle_timestep(self._h, float(dt))

    def step(self, n: int = 1) -> None:
        for _ in range(n):
            self.timestep(self.getmindt())

    def state(self) -> dict:
        n = self.ncell
        prim, cons = np.zeros(5 * n), np.zeros(5 * n)
        piph, dphi, wiph = np.zeros(n), np.zeros(n), np.zeros(n)
        pp, t = np.zeros(2), np.zeros(1)
        lib().oracle_get_state(self._h, _d(prim), _d(cons), _d(piph), _d(dphi), _d(wiph), _d(pp), _d(t))
        # the pointers above reference fresh contiguous arrays, so results land in them
        return dict(prim=prim.reshape(5, n), cons=cons.reshape(5, n), piph=piph, dphi=dphi, wiph=wiph,
                    planet_phi=pp, t=float(t[0]))

    def close(self):
        if self._h:
            lib().oracle_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def track_faces(piA, dpA, piB, dpB, phimax=2 * np.pi):
    NA, NB = len(piA), len(piB)
    n = NA + NB
    iL, iR, LR = (np.zeros(n, np.int32) for _ in range(3))
    pm, pp = np.zeros(n), np.zeros(n)
    ip = C.POINTER(C.c_int)
    a = [np.ascontiguousarray(x, np.float64) for x in (piA, dpA, piB, dpB)]
    lib().oracle_track_faces(_d(a[0]), _d(a[1]), NA, _d(a[2]), _d(a[3]), NB, phimax, iL.ctypes.data_as(ip),
                             iR.ctypes.data_as(ip), LR.ctypes.data_as(ip), pm.ctypes.data_as(C.POINTER(C.c_double)),
                             pp.ctypes.data_as(C.POINTER(C.c_double)))
    return iL, iR, LR, pm, pp
