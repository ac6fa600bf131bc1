"""MI355X (gfx950) implementation of Disco's per-timestep finite-volume update.

Python mirror of the reference's driver-level interface over the C ABI in
``include/disco_gpu.h`` (``disco_amd/_lib/libdisco_gpu.so``):

* :func:`read_par_file` / :func:`parse_par`  -- ``read_par_file`` (readpar.c:44-210)
* :func:`setup_run`                          -- setupGrid/setupDomain/setupCells on the host
* :class:`Domain`                            -- the device-resident ``struct domain``:
  ``getmindt`` (misc.c:39-84), ``timestep`` (timestep.c:7-54), ``step`` (the
  main loop of main.c:68-75 without output), ``state``/``upload``.

There is no CPU fallback: if the HIP library is missing, every entry point
raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libdisco_gpu.so")

HYDRO = {"euler": 0, "mhd": 1, "grmhd": 2, "greuler": 3, "grmhd2": 4}
NQ_OF = {0: 5, 1: 8, 2: 8}


class Params(C.Structure):
    """Mirror of ``disco_params`` (include/disco_params.h)."""

    _fields_ = [
        ("t_min", C.c_double), ("t_max", C.c_double),
        ("Num_R", C.c_int), ("Num_Z", C.c_int),
        ("aspect", C.c_double),
        ("rmin", C.c_double), ("rmax", C.c_double), ("zmin", C.c_double), ("zmax", C.c_double),
        ("phimax", C.c_double),
        ("NoBC_Rmin", C.c_int), ("NoBC_Rmax", C.c_int), ("NoBC_Zmin", C.c_int), ("NoBC_Zmax", C.c_int),
        ("LogZoning", C.c_int), ("R_Periodic", C.c_int), ("Z_Periodic", C.c_int),
        ("LogRadius", C.c_double),
        ("Mesh_Motion", C.c_int), ("Riemann_Solver", C.c_int), ("Timestep", C.c_int),
        ("CFL", C.c_double), ("PLM", C.c_double), ("maxDT", C.c_double),
        ("Density_Floor", C.c_double), ("Pressure_Floor", C.c_double),
        ("Exact_Mesh_Omega", C.c_int),
        ("Exact_Mesh_Omega_Par", C.c_double),
        ("Energy_Omega", C.c_int),
        ("Energy_Omega_Par", C.c_double),
        ("RotFrame", C.c_int),
        ("RotOmega", C.c_double), ("RotD", C.c_double),
        ("Adiabatic_Index", C.c_double),
        ("isothermal_flag", C.c_int),
        ("Cs2_Profile", C.c_int),
        ("Cs2_Par", C.c_double),
        ("Disk_Mach", C.c_double),
        ("Mass_Ratio", C.c_double),
        ("grav2D", C.c_int),
        ("grav_eps", C.c_double),
        ("visc_flag", C.c_int),
        ("CT", C.c_int),
        ("metricPar0", C.c_int),
        ("metricPar1", C.c_double), ("metricPar2", C.c_double), ("metricPar3", C.c_double),
        ("metricPar4", C.c_double),
        ("initPar0", C.c_int),
        ("initPar1", C.c_double), ("initPar2", C.c_double), ("initPar3", C.c_double),
        ("initPar4", C.c_double), ("initPar5", C.c_double), ("initPar6", C.c_double),
        ("initPar7", C.c_double), ("initPar8", C.c_double),
        ("noiseType", C.c_int),
        ("noiseAbs", C.c_double), ("noiseRel", C.c_double),
        ("hydro", C.c_int), ("boundary", C.c_int), ("planets", C.c_int), ("initial", C.c_int),
        ("num_c", C.c_int), ("num_n", C.c_int),
        ("sinkType", C.c_int), ("sinkNumber", C.c_int),
        ("sinkPar1", C.c_double), ("sinkPar2", C.c_double), ("sinkPar3", C.c_double),
        ("sinkPar4", C.c_double), ("sinkPar5", C.c_double),
        ("coolType", C.c_int), ("nozzleType", C.c_int), ("dampInnerType", C.c_int),
        ("dampOuterType", C.c_int), ("dampUpperType", C.c_int), ("dampLowerType", C.c_int),
        ("coolPar1", C.c_double), ("coolPar2", C.c_double), ("coolPar3", C.c_double), ("coolPar4", C.c_double),
        ("dampTimeInner", C.c_double), ("dampLenInner", C.c_double), ("dampTimeOuter", C.c_double),
        ("dampLenOuter", C.c_double), ("dampTimeUpper", C.c_double), ("dampLenUpper", C.c_double),
        ("dampTimeLower", C.c_double), ("dampLenLower", C.c_double),
        ("nozzlePar1", C.c_double), ("nozzlePar2", C.c_double), ("nozzlePar3", C.c_double),
        ("nozzlePar4", C.c_double),
        ("viscosity", C.c_double), ("alpha_flag", C.c_int), ("include_atmos", C.c_int),
    ]

    def copy(self) -> "Params":
        p = Params()
        C.pointer(p)[0] = self
        return p


_lib = None


def lib() -> C.CDLL:
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m disco_amd.build` "
            "(there is no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int)
    vp = C.c_void_p
    L.disco_parse_par.argtypes = [C.c_char_p, C.POINTER(Params), C.c_char_p, C.c_int]
    L.disco_read_par.argtypes = [C.c_char_p, C.POINTER(Params), C.c_char_p, C.c_int]
    L.disco_gpu_create.argtypes = [C.POINTER(Params), C.c_int, C.POINTER(vp)]
    L.disco_setup_run.argtypes = [C.POINTER(Params), ip, ip, dp, dp, dp, dp, dp, dp, dp, dp]
    L.disco_gpu_info.argtypes = [vp, C.c_int, ip]
    L.disco_gpu_grid.argtypes = [vp, C.c_int, ip, dp, dp]
    L.disco_gpu_mindt.argtypes = [vp, dp]
    L.disco_gpu_step.argtypes = [vp, C.c_int]
    L.disco_gpu_advance.argtypes = [vp, C.c_int, C.c_int, ip, ip]
    L.disco_gpu_final.argtypes = [vp, ip]
    L.disco_gpu_report.argtypes = [vp, C.c_int, dp, C.c_int, ip]
    L.disco_gpu_write_checkpoint.argtypes = [vp, C.c_int, C.c_char_p]
    L.disco_gpu_restart.argtypes = [vp, C.c_int, C.c_char_p]
    if hasattr(L, "disco_gpu_report_noalloc"):
        L.disco_gpu_report_noalloc.argtypes = [vp, C.c_int, dp, C.c_int, ip]
    L.disco_gpu_timestep.argtypes = [vp, dp]
    L.disco_gpu_sync.argtypes = [vp]
    L.disco_gpu_download.argtypes = [vp, C.c_int, dp, dp, dp, dp, dp]
    L.disco_gpu_upload.argtypes = [vp, C.c_int, dp, dp, dp, dp, dp]
    L.disco_gpu_download_faces.argtypes = [vp, C.c_int, dp]
    L.disco_gpu_upload_faces.argtypes = [vp, C.c_int, dp]
    L.disco_gpu_time.argtypes = [vp, dp, dp, dp]
    L.disco_gpu_kernel_times.argtypes = [vp, C.c_int, C.c_int, C.c_char_p, dp, ip]
    L.disco_gpu_time_steps.argtypes = [vp, C.c_int, dp]
    fp = C.POINTER(C.c_float)
    L.disco_gpu_set_c2p_net.argtypes = [vp, fp, fp]
    L.disco_gpu_diagnostics.argtypes = [vp, C.c_int, dp, dp, C.c_int]
    L.disco_gpu_autotune.argtypes = [vp, C.c_int, ip, dp]
    L.disco_gpu_c2p_count.argtypes = [vp, C.c_int]
    L.disco_gpu_c2p_stats.argtypes = [vp, C.POINTER(C.c_longlong), C.c_int]
    L.disco_gpu_destroy.argtypes = [vp]
    L.disco_gpu_destroy.restype = None
    L.disco_gpu_last_error.restype = C.c_char_p
    _lib = L
    return L


def _err(L) -> str:
    return L.disco_gpu_last_error().decode(errors="replace")


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what} failed: {_err(lib())}")


def _dptr(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


def _iptr(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_int))


def parse_par(text: str) -> Params:
    """``read_par_file`` on in.par text; raises on missing keys (readpar.c:199-206)."""
    L = lib()
    p = Params()
    buf = C.create_string_buffer(2048)
    rc = L.disco_parse_par(text.encode(), C.byref(p), buf, len(buf))
    if rc != 0:
        raise ValueError(buf.value.decode())
    return p


def read_par_file(path: str) -> Params:
    with open(path) as f:
        return parse_par(f.read())


@dataclass
class RunInfo:
    Nr: int
    Nz: int
    NgRa: int
    NgRb: int
    NgZa: int
    NgZb: int
    ncell: int
    nreal: int
    nq: int
    npl: int
    nruns: int
    ncell_total: int

    @staticmethod
    def from_array(a) -> "RunInfo":
        return RunInfo(*[int(v) for v in a[:12]])


def setup_run(p: Params) -> dict:
    """Host-only set-up (gridsetup.c + domain.c), returned as numpy arrays."""
    L = lib()
    info = np.zeros(12, np.int32)
    _check(L.disco_setup_run(C.byref(p), _iptr(info), None, None, None, None, None, None, None, None, None),
           "disco_setup_run")
    I = RunInfo.from_array(info)
    out = dict(
        info=I,
        Np=np.zeros(I.Nr * I.Nz, np.int32),
        r_jph=np.zeros(I.Nr + 1),
        z_kph=np.zeros(I.Nz + 1),
        prim=np.zeros(I.nq * I.ncell),
        cons=np.zeros(I.nq * I.ncell),
        piph=np.zeros(I.ncell),
        dphi=np.zeros(I.ncell),
        wiph=np.zeros(I.ncell),
        planet_phi=np.zeros(2),
    )
    _check(L.disco_setup_run(C.byref(p), _iptr(info), _iptr(out["Np"]), _dptr(out["r_jph"]),
                             _dptr(out["z_kph"]), _dptr(out["prim"]), _dptr(out["cons"]), _dptr(out["piph"]),
                             _dptr(out["dphi"]), _dptr(out["wiph"]), _dptr(out["planet_phi"])),
           "disco_setup_run")
    out["prim"] = out["prim"].reshape(I.nq, I.ncell)
    out["cons"] = out["cons"].reshape(I.nq, I.ncell)
    return out


class Domain:
    """Device-resident batch of runs (the reference's ``struct domain``)."""

    def __init__(self, params):
        if isinstance(params, Params):
            params = [params]
        self.params = [p.copy() for p in params]
        L = lib()
        arr = (Params * len(self.params))(*self.params)
        h = C.c_void_p()
        _check(L.disco_gpu_create(arr, len(self.params), C.byref(h)), "disco_gpu_create")
        self._h = h
        self.nruns = len(self.params)
        self.info = [self._info(r) for r in range(self.nruns)]

    def _info(self, run: int) -> RunInfo:
        a = np.zeros(12, np.int32)
        _check(lib().disco_gpu_info(self._h, run, _iptr(a)), "disco_gpu_info")
        return RunInfo.from_array(a)

    def grid(self, run: int = 0) -> dict:
        I = self.info[run]
        Np = np.zeros(I.Nr * I.Nz, np.int32)
        r = np.zeros(I.Nr + 1)
        z = np.zeros(I.Nz + 1)
        _check(lib().disco_gpu_grid(self._h, run, _iptr(Np), _dptr(r), _dptr(z)), "disco_gpu_grid")
        return dict(Np=Np, r_jph=r, z_kph=z)

    def getmindt(self) -> np.ndarray:
        dt = np.zeros(self.nruns)
        _check(lib().disco_gpu_mindt(self._h, _dptr(dt)), "disco_gpu_mindt")
        return dt

    def timestep(self, dt) -> None:
        dt = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, float), (self.nruns,)))
        _check(lib().disco_gpu_timestep(self._h, _dptr(dt)), "disco_gpu_timestep")

    def step(self, nsteps: int = 1) -> None:
        _check(lib().disco_gpu_step(self._h, int(nsteps)), "disco_gpu_step")

    def advance(self, max_steps: int = 1 << 30, poll_files: bool = False) -> tuple:
        """main.c:68-75 without output: steps until every run has taken its
        step to T_End (check_dt, domain.c:283-310) or max_steps are done.
        Returns (steps taken, status) with status bits 1 every run final,
        2 an "abort" file ended the loop, 4 a "latest" file was seen (removed;
        the loop stopped before the step so the caller can write a
        checkpoint); the files are read only with poll_files."""
        n = C.c_int(0)
        st = C.c_int(0)
        _check(lib().disco_gpu_advance(self._h, int(max_steps), int(bool(poll_files)), C.byref(n), C.byref(st)),
               "disco_gpu_advance")
        return n.value, st.value

    def final(self) -> np.ndarray:
        """check_dt's final flag of every run."""
        f = np.zeros(self.nruns, np.int32)
        _check(lib().disco_gpu_final(self._h, f.ctypes.data_as(C.POINTER(C.c_int))), "disco_gpu_final")
        return f.astype(bool)

    def report(self, run: int = 0) -> np.ndarray:
        """The values of report.c's report.dat line for run `run` (see
        include/disco_gpu.h: t, q, 8 x Npl planet accumulators, Mass, Torque,
        Torque2, 5 x (Torque1_cut, Torque2_cut))."""
        out = np.zeros(15 + 8 * self.info[run].npl)
        n = C.c_int(0)
        _check(lib().disco_gpu_report(self._h, int(run), _dptr(out), out.size, C.byref(n)), "disco_gpu_report")
        return out[:n.value]

    def write_checkpoint(self, path: str, run: int = 0) -> None:
        """output() of the HDF5 module: the reference's checkpoint datasets
        (plus the conserved state and mesh speeds) in this library's
        container; read it with disco_amd.checkpoint.read.  Zeroes the
        accumulated diagnostics, as output() does."""
        _check(lib().disco_gpu_write_checkpoint(self._h, int(run), os.fsencode(path)), "disco_gpu_write_checkpoint")

    def restart(self, path: str, run: int = 0) -> None:
        """restart() from a checkpoint of a run with the same parameters."""
        _check(lib().disco_gpu_restart(self._h, int(run), os.fsencode(path)), "disco_gpu_restart")

    def report_noalloc(self, run: int = 0) -> np.ndarray:
        """The same line as report(), from the entry point that allocates
        nothing (a per-GPU static scratch; at most 65536 annuli per run;
        calls that overlap on one GPU race)."""
        out = np.zeros(15 + 8 * self.info[run].npl)
        n = C.c_int(0)
        _check(lib().disco_gpu_report_noalloc(self._h, int(run), _dptr(out), out.size, C.byref(n)),
               "disco_gpu_report_noalloc")
        return out[:n.value]

    def sync(self) -> None:
        _check(lib().disco_gpu_sync(self._h), "disco_gpu_sync")

    def time(self) -> dict:
        t = np.zeros(self.nruns)
        dt = np.zeros(self.nruns)
        pp = np.zeros(2 * self.nruns)
        _check(lib().disco_gpu_time(self._h, _dptr(t), _dptr(dt), _dptr(pp)), "disco_gpu_time")
        return dict(t=t, dt=dt, planet_phi=pp.reshape(self.nruns, 2))

    def state(self, run: int = 0) -> dict:
        I = self.info[run]
        prim = np.zeros(I.nq * I.ncell)
        cons = np.zeros(I.nq * I.ncell)
        piph = np.zeros(I.ncell)
        dphi = np.zeros(I.ncell)
        wiph = np.zeros(I.ncell)
        _check(lib().disco_gpu_download(self._h, run, _dptr(prim), _dptr(cons), _dptr(piph), _dptr(dphi),
                                        _dptr(wiph)), "disco_gpu_download")
        out = dict(prim=prim.reshape(I.nq, I.ncell), cons=cons.reshape(I.nq, I.ncell), piph=piph, dphi=dphi,
                   wiph=wiph)
        if self.params[run].hydro in (HYDRO["mhd"], HYDRO["grmhd"], HYDRO["grmhd2"]):
            out["Phi"] = self.faces(run)
        return out

    def faces(self, run: int = 0) -> np.ndarray:
        """Magnetic face fluxes Phi[5, ncell] (MHD or GRMHD with constrained transport)."""
        I = self.info[run]
        Phi = np.zeros(5 * I.ncell)
        _check(lib().disco_gpu_download_faces(self._h, run, _dptr(Phi)), "disco_gpu_download_faces")
        return Phi.reshape(5, I.ncell)

    def upload_faces(self, run: int, Phi) -> None:
        Phi = np.ascontiguousarray(Phi, dtype=np.float64).ravel()
        _check(lib().disco_gpu_upload_faces(self._h, run, _dptr(Phi)), "disco_gpu_upload_faces")

    def upload(self, run: int, prim=None, cons=None, piph=None, dphi=None, wiph=None) -> None:
        def c(a):
            return None if a is None else np.ascontiguousarray(a, dtype=np.float64).ravel()
        prim, cons, piph, dphi, wiph = map(c, (prim, cons, piph, dphi, wiph))
        _check(lib().disco_gpu_upload(self._h, run, _dptr(prim), _dptr(cons), _dptr(piph), _dptr(dphi),
                                      _dptr(wiph)), "disco_gpu_upload")

    def kernel_times(self, nsteps: int) -> dict:
        maxk = 32
        names = C.create_string_buffer(32 * maxk)
        avg = np.zeros(maxk)
        cnt = np.zeros(maxk, np.int32)
        k = lib().disco_gpu_kernel_times(self._h, int(nsteps), maxk, names, _dptr(avg), _iptr(cnt))
        if k < 0:
            raise RuntimeError(f"disco_gpu_kernel_times failed: {_err(lib())}")
        out = {}
        for i in range(k):
            nm = names.raw[32 * i:32 * (i + 1)].split(b"\0", 1)[0].decode()
            out[nm] = (float(avg[i]), int(cnt[i]))
        return out

    def cell_positions(self, run: int = 0):
        """Real cells of a run (host order): their cons2prim position x[3, n]
        (centroid r, mid phi, centroid z; misc.c:314-331) and volume dV[n]."""
        I = self.info[run]
        g = self.grid(run)
        st = self.state(run)
        Np = g["Np"]
        jk = np.repeat(np.arange(I.Nr * I.Nz), Np)
        j, k = jk % I.Nr, jk // I.Nr
        rp, rm = g["r_jph"][j + 1], g["r_jph"][j]
        zp, zm = g["z_kph"][k + 1], g["z_kph"][k]
        real = (j >= I.NgRa) & (j < I.Nr - I.NgRb) & (k >= I.NgZa) & (k < I.Nz - I.NgZb)
        phip = st["piph"]
        dphi = st["dphi"]
        x = np.stack([2.0 * (rp * rp + rp * rm + rm * rm) / (3.0 * (rp + rm)), phip - 0.5 * dphi, 0.5 * (zp + zm)])
        dV = 0.5 * (rp + rm) * (rp - rm) * dphi * (zp - zm)
        return x[:, real], dV[real], real, st

    def diagnostics(self, run: int = 0, reset: bool = False) -> dict:
        """Accumulated step diagnostics (add_diagnostics): Qrz[Nz, Nr, 8] =
        sum over steps of dt x annulus average, and t_avg."""
        I = self.info[run]
        Q = np.zeros(I.Nr * I.Nz * 8)
        t = np.zeros(1)
        _check(lib().disco_gpu_diagnostics(self._h, run, _dptr(Q), _dptr(t), int(bool(reset))),
               "disco_gpu_diagnostics")
        return dict(Qrz=Q.reshape(I.Nz, I.Nr, 8), t_avg=float(t[0]))

    def set_c2p_net(self, params=None, norm=None) -> None:
        """GRMHD: start the in-loop cons2prim from a trained guess network
        (params, norm as C2P.get_mlp returns them); None returns to the
        reference's start (the cell's previous primitives)."""
        if params is None:
            _check(lib().disco_gpu_set_c2p_net(self._h, None, None), "disco_gpu_set_c2p_net")
            return
        w = np.ascontiguousarray(params, np.float32)
        n = np.ascontiguousarray(norm, np.float32)
        fp = C.POINTER(C.c_float)
        _check(lib().disco_gpu_set_c2p_net(self._h, w.ctypes.data_as(fp), n.ctypes.data_as(fp)),
               "disco_gpu_set_c2p_net")

    def c2p_count(self, enable: bool = True) -> None:
        """GRMHD: count the in-loop cons2prim solves (see c2p_stats)."""
        _check(lib().disco_gpu_c2p_count(self._h, int(bool(enable))), "disco_gpu_c2p_count")

    def c2p_stats(self, reset: bool = False) -> dict:
        out = (C.c_longlong * 4)()
        _check(lib().disco_gpu_c2p_stats(self._h, out, int(bool(reset))), "disco_gpu_c2p_stats")
        return dict(iterations=out[0], solves=out[1], fallbacks=out[2], unconverged=out[3])

    def autotune(self, nsteps: int = 3) -> dict:
        """Pick the annulus tile and threads per block by timing (state kept)."""
        ch = np.zeros(9, np.int32)
        tab = np.zeros(15)
        _check(lib().disco_gpu_autotune(self._h, int(nsteps), _iptr(ch), _dptr(tab)), "disco_gpu_autotune")
        tiles = [1, 2, 4, 8, 16] if any(i.Nz > 1 for i in self.info) else [int(ch[0])]
        blocks = [64, 128, 256]
        ms = {f"{t}x{b}": float(tab[i * 3 + j]) for i, t in enumerate(tiles) for j, b in enumerate(blocks)}
        classes = ["stage", "slopes", "flux", "emf", "ct_update", "flip", "update", "zcopy"]
        return dict(tile=int(ch[0]), block=dict(zip(classes, (int(v) for v in ch[1:]))), ms_per_step=ms)

    def time_steps(self, nsteps: int) -> float:
        ms = C.c_double()
        _check(lib().disco_gpu_time_steps(self._h, int(nsteps), C.byref(ms)), "disco_gpu_time_steps")
        return ms.value

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().disco_gpu_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def config_path(name: str) -> str:
    """Path of a config shipped in the repository's ``configs/``."""
    return os.path.join(os.path.dirname(_HERE), "configs", name)


# ---------------------------------------------------------------- GRMHD cons2prim
class C2PParams(C.Structure):
    """Mirror of ``disco_c2p_params`` (include/disco_gpu.h)."""

    _fields_ = [("gamma_law", C.c_double), ("M", C.c_double), ("a", C.c_double), ("prec", C.c_double),
                ("max_iter", C.c_int), ("nextra", C.c_int), ("rmin", C.c_double), ("rmax", C.c_double),
                ("theta_cut", C.c_double)]


PLAIN, LEARNED = 0, 1
MLP_NPARAM, MLP_NNORM = 13186, 20


def _c2p_lib():
    L = lib()
    if getattr(L, "_c2p_ready", False):
        return L
    dp, fp, ip, vp = C.POINTER(C.c_double), C.POINTER(C.c_float), C.POINTER(C.c_int), C.c_void_p
    L.disco_c2p_create.argtypes = [C.POINTER(C2PParams), C.c_long, C.POINTER(vp)]
    L.disco_c2p_destroy.argtypes = [vp]
    L.disco_c2p_destroy.restype = None
    L.disco_c2p_last_error.restype = C.c_char_p
    L.disco_c2p_generate.argtypes = [vp, C.c_ulonglong, C.c_long, C.c_long]
    L.disco_c2p_upload.argtypes = [vp, dp, dp, C.c_long]
    L.disco_c2p_solve.argtypes = [vp, C.c_int, C.c_int, dp]
    L.disco_c2p_download.argtypes = [vp, dp, ip, ip, C.c_long]
    L.disco_c2p_errors.argtypes = [vp, C.c_ulonglong, C.c_long, fp]
    L.disco_mlp_train.argtypes = [vp, C.c_ulonglong, C.c_long, C.c_int, C.c_int, C.c_float, fp]
    L.disco_mlp_train_resident.argtypes = [vp, C.c_ulonglong, C.c_int, C.c_int, C.c_float, fp]
    L.disco_mlp_get.argtypes = [vp, fp, fp]
    L.disco_mlp_set.argtypes = [vp, fp, fp]
    L.disco_mlp_grad.argtypes = [vp, fp, fp, C.c_int, fp, fp]
    L.disco_mlp_infer.argtypes = [vp, fp, C.c_long]
    L.disco_c2p_features.argtypes = [vp, fp, fp, C.c_long]
    L._c2p_ready = True
    return L


def _fptr(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))


class C2P:
    """Batched GRMHD cons2prim on the GPU (plain and learned starts)."""

    def __init__(self, capacity: int, gamma: float, M: float = 1.0, a: float = 0.9, prec: float = 1e-12,
                 max_iter: int = 30, nextra: int = 10, rmin: float | None = None, rmax: float = 50.0,
                 theta_cut: float = 0.1):
        if rmin is None:
            rmin = 1.5 * (M + (M * M - a * a) ** 0.5)
        self.p = C2PParams(gamma, M, a, prec, max_iter, nextra, rmin, rmax, theta_cut)
        self.capacity = int(capacity)
        L = _c2p_lib()
        h = C.c_void_p()
        self._chk(L.disco_c2p_create(C.byref(self.p), self.capacity, C.byref(h)), "disco_c2p_create")
        self._h = h
        self.n = 0

    @staticmethod
    def _chk(rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed: {_c2p_lib().disco_c2p_last_error().decode()}")

    def generate(self, seed: int, k0: int, n: int) -> None:
        self._chk(_c2p_lib().disco_c2p_generate(self._h, seed, k0, n), "disco_c2p_generate")
        self.n = n

    def upload(self, cons: np.ndarray, x: np.ndarray) -> None:
        cons = np.ascontiguousarray(cons, np.float64)
        x = np.ascontiguousarray(x, np.float64)
        n = cons.shape[-1]
        self._chk(_c2p_lib().disco_c2p_upload(self._h, _dptr(cons), _dptr(x), n), "disco_c2p_upload")
        self.n = n

    def solve(self, mode: int = PLAIN, nextra: int | None = None, split: bool = False):
        """Solve the resident states; returns ms (or (ms, ms_network) with split)."""
        ms = np.zeros(2)
        ne = self.p.nextra if nextra is None else nextra
        self._chk(_c2p_lib().disco_c2p_solve(self._h, mode, ne, _dptr(ms)), "disco_c2p_solve")
        return (float(ms[0]), float(ms[1])) if split else float(ms[0])

    def download(self, n: int | None = None) -> dict:
        n = self.n if n is None else n
        prim = np.zeros(8 * n)
        it = np.zeros(n, np.int32)
        fl = np.zeros(n, np.int32)
        self._chk(_c2p_lib().disco_c2p_download(self._h, _dptr(prim), _iptr(it), _iptr(fl), n), "download")
        return dict(prim=prim.reshape(8, n), iters=it, flags=fl)

    def errors(self, seed: int, k0: int) -> np.ndarray:
        e = np.zeros(self.n, np.float32)
        self._chk(_c2p_lib().disco_c2p_errors(self._h, seed, k0, _fptr(e)), "disco_c2p_errors")
        return e

    def train(self, seed: int, ntrain: int, steps: int, batch: int, lr: float) -> np.ndarray:
        loss = np.zeros(steps, np.float32)
        self._chk(_c2p_lib().disco_mlp_train(self._h, seed, ntrain, steps, batch, lr, _fptr(loss)), "disco_mlp_train")
        return loss

    def train_resident(self, seed: int, steps: int, batch: int, lr: float) -> np.ndarray:
        """Train on the uploaded states (e.g. a run's own cells)."""
        loss = np.zeros(steps, np.float32)
        self._chk(_c2p_lib().disco_mlp_train_resident(self._h, seed, steps, batch, lr, _fptr(loss)),
                  "disco_mlp_train_resident")
        return loss

    def get_mlp(self):
        w = np.zeros(MLP_NPARAM, np.float32)
        nrm = np.zeros(MLP_NNORM, np.float32)
        self._chk(_c2p_lib().disco_mlp_get(self._h, _fptr(w), _fptr(nrm)), "disco_mlp_get")
        return w, nrm

    def set_mlp(self, w, nrm) -> None:
        w = np.ascontiguousarray(w, np.float32)
        nrm = np.ascontiguousarray(nrm, np.float32)
        self._chk(_c2p_lib().disco_mlp_set(self._h, _fptr(w), _fptr(nrm)), "disco_mlp_set")

    def grad(self, feat: np.ndarray, tgt: np.ndarray):
        feat = np.ascontiguousarray(feat, np.float32)
        tgt = np.ascontiguousarray(tgt, np.float32)
        g = np.zeros(MLP_NPARAM, np.float32)
        loss = np.zeros(1, np.float32)
        n = feat.shape[-1]
        self._chk(_c2p_lib().disco_mlp_grad(self._h, _fptr(feat), _fptr(tgt), n, _fptr(g), _fptr(loss)),
                  "disco_mlp_grad")
        return g, float(loss[0])

    def infer(self, n: int | None = None) -> np.ndarray:
        n = self.n if n is None else n
        out = np.zeros(2 * n, np.float32)
        self._chk(_c2p_lib().disco_mlp_infer(self._h, _fptr(out), n), "disco_mlp_infer")
        return out.reshape(n, 2)

    def features(self, n: int | None = None):
        n = self.n if n is None else n
        f = np.zeros(8 * n, np.float32)
        t = np.zeros(2 * n, np.float32)
        self._chk(_c2p_lib().disco_c2p_features(self._h, _fptr(f), _fptr(t), n), "disco_c2p_features")
        return f.reshape(8, n), t.reshape(2, n)

    def close(self):
        if getattr(self, "_h", None):
            _c2p_lib().disco_c2p_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
