"""ctypes binding of the CPU oracle (oracle/liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product (flowlb_amd) never
imports this module.
"""
import ctypes as C  # noqa: N812
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liborc.so"

FLUID, GAS, INTERFACE, WALL, MWALL, INLET, OUTLET, GAS_INTERFACE = (
    0x01, 0x02, 0x04, 0x08, 0x10, 0x20, 0x40, 0x80)


class OrcParams(C.Structure):
    _fields_ = [("nx", C.c_uint32), ("ny", C.c_uint32), ("inv_tau", C.c_float),
                ("fx", C.c_float), ("fy", C.c_float), ("use_force", C.c_int32),
                ("collision", C.c_int32), ("smagorinsky", C.c_int32),
                ("free_surface", C.c_int32), ("gas_check_corrected", C.c_int32),
                ("mass_mode", C.c_int32), ("threads", C.c_int32)]


class OrcRect(C.Structure):
    _fields_ = [("x0", C.c_double), ("y0", C.c_double), ("height", C.c_double),
                ("width", C.c_double)]


class OrcChannel(C.Structure):
    _fields_ = [("nx", C.c_uint32), ("ny", C.c_uint32), ("dx", C.c_double),
                ("x_min", C.c_double), ("y_min", C.c_double), ("free_surface", C.c_int32),
                ("bc_left", C.c_uint8), ("bc_right", C.c_uint8), ("bc_up", C.c_uint8),
                ("bc_down", C.c_uint8), ("n_obstacles", C.c_uint32),
                ("n_initial", C.c_uint32), ("obstacles", C.POINTER(OrcRect)),
                ("initial", C.POINTER(OrcRect))]


def build(force=False):
    """Compiles liborc.so with the committed Makefile (g++, no GPU needed)."""
    if force or not LIB_PATH.exists():
        subprocess.run(["make", "-s", "-C", str(HERE)] + (["-B"] if force else []), check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        P = C.c_void_p
        dp = C.POINTER(C.c_double)
        u8p = C.POINTER(C.c_uint8)
        L.orc_build_channel_flags.argtypes = [C.POINTER(OrcChannel), u8p]
        L.orc_build_channel_flags.restype = C.c_int
        L.orc_seed_free_surface.argtypes = [u8p, C.c_uint32, C.c_uint32]
        L.orc_init_free_surface_fields.argtypes = [u8p, C.c_size_t, dp, dp, dp]
        for pre in ("orc", "orcf"):
            getattr(L, pre + "_create").argtypes = [C.POINTER(OrcParams), u8p, dp, dp, dp]
            getattr(L, pre + "_create").restype = P
            getattr(L, pre + "_set_fs_fields").argtypes = [P, dp, dp, dp]
            getattr(L, pre + "_step").argtypes = [P, C.c_uint64]
            getattr(L, pre + "_get").argtypes = [P, dp, dp, dp, u8p, dp, dp, dp]
            getattr(L, pre + "_total_mass").argtypes = [P]
            getattr(L, pre + "_total_mass").restype = C.c_double
            getattr(L, pre + "_destroy").argtypes = [P]
        L.orc_time.argtypes = [P]
        L.orc_time.restype = C.c_uint64
        L.orc_set_f.argtypes = [P, dp]
        L.orc_set_cells.argtypes = [P, u8p, dp, dp]
        L.orc_set_time.argtypes = [P, C.c_uint64]
        L.orc_run_pass.argtypes = [P, C.c_int]
        L.orc_run_pass.restype = C.c_int
        L.orc_get_reconstructed.argtypes = [P, dp, C.POINTER(C.c_uint16)]
        L.orc_discharge.argtypes = [P, C.c_uint32]
        L.orc_discharge.restype = C.c_double
        L.orc_fs_discharge.argtypes = [P, C.c_uint32]
        L.orc_fs_discharge.restype = C.c_double
        if hasattr(L, "orc_adjoint"):
            L.orc_adjoint.argtypes = [P, C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint64,
                                      C.c_uint32, dp, dp, dp]
            L.orc_adjoint.restype = C.c_int
            L.orc_objective.argtypes = [P, C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint64]
            L.orc_objective.restype = C.c_double
        _lib = L
    return _lib


def _dp(a):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _u8(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint8))


def build_channel_flags(nx, ny, dx, bc_left, bc_right, bc_up, bc_down, free_surface=False,
                        obstacles=(), initial=(), x_min=0.0, y_min=0.0):
    """Oracle restatement of initData for rectangle-only decks.  Rectangles are
    (x0, y0, height, width) in SI units."""
    obs = (OrcRect * max(1, len(obstacles)))(*[OrcRect(*r) for r in obstacles])
    ini = (OrcRect * max(1, len(initial)))(*[OrcRect(*r) for r in initial])
    g = OrcChannel(nx, ny, dx, x_min, y_min, int(free_surface), bc_left, bc_right, bc_up,
                   bc_down, len(obstacles), len(initial), obs, ini)
    out = np.zeros(nx * ny, dtype=np.uint8)
    rc = lib().orc_build_channel_flags(C.byref(g), _u8(out))
    if rc:
        raise ValueError("orc_build_channel_flags failed")
    if free_surface:
        lib().orc_seed_free_surface(_u8(out), nx, ny)
    return out


def init_free_surface_fields(flags):
    flags = np.ascontiguousarray(flags, dtype=np.uint8)
    n = flags.size
    m, p, e = np.zeros(n), np.zeros(n), np.zeros(n)
    lib().orc_init_free_surface_fields(_u8(flags), n, _dp(m), _dp(p), _dp(e))
    return m, p, e


class Oracle:
    """Serial restatement of the reference step (FP64 by default, FP32 with fp32=True)."""

    def __init__(self, flags, nx, ny, tau, fx=0.0, fy=0.0, use_force=False, collision=0,
                 smagorinsky=False, free_surface=False, gas_check_corrected=False,
                 mass_mode=0, rho0=None, u0=None, alpha=None, threads=0, fp32=False,
                 inv_tau=None):
        self.nx, self.ny, self.N = nx, ny, nx * ny
        self.free_surface = free_surface
        self.pre = "orcf" if fp32 else "orc"
        if threads <= 0:
            threads = int(os.environ.get("FLB_ORACLE_THREADS", "0")) or min(8, os.cpu_count() or 1)
        if inv_tau is None:
            inv_tau = float(np.float32(1.0) / np.float32(tau))
        self.p = OrcParams(nx, ny, inv_tau, fx, fy, int(use_force), collision, int(smagorinsky),
                           int(free_surface), int(gas_check_corrected), mass_mode, threads)
        self.flags0 = np.ascontiguousarray(flags, dtype=np.uint8)
        rho0 = None if rho0 is None else np.ascontiguousarray(rho0, dtype=np.float64)
        u0 = None if u0 is None else np.ascontiguousarray(u0, dtype=np.float64)
        alpha = None if alpha is None else np.ascontiguousarray(alpha, dtype=np.float64)
        self._keep = (rho0, u0, alpha)
        self.h = getattr(lib(), self.pre + "_create")(C.byref(self.p), _u8(self.flags0), _dp(rho0),
                                                       _dp(u0), _dp(alpha))
        if free_surface:
            m, p, e = init_free_surface_fields(self.flags0)
            self.set_fs_fields(m, p, e)

    def set_fs_fields(self, mass, phi, excess=None):
        mass = np.ascontiguousarray(mass, dtype=np.float64)
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        excess = None if excess is None else np.ascontiguousarray(excess, dtype=np.float64)
        getattr(lib(), self.pre + "_set_fs_fields")(self.h, _dp(mass), _dp(phi), _dp(excess))

    def step(self, n=1):
        getattr(lib(), self.pre + "_step")(self.h, n)

    def get(self):
        N = self.N
        f = np.empty(9 * N)
        rho = np.empty(N)
        u = np.empty(2 * N)
        flags = np.empty(N, dtype=np.uint8)
        mass = np.zeros(N)
        excess = np.zeros(N)
        phi = np.zeros(N)
        getattr(lib(), self.pre + "_get")(self.h, _dp(f), _dp(rho), _dp(u), _u8(flags), _dp(mass),
                                          _dp(excess), _dp(phi))
        return dict(f=f, rho=rho, u=u, flags=flags, mass=mass, excess=excess, phi=phi)

    def set_f(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        lib().orc_set_f(self.h, _dp(f))

    def total_mass(self):
        return getattr(lib(), self.pre + "_total_mass")(self.h)

    # Single passes on a hand-built state (FP64 oracle only).
    PASSES = {"fs1": 1, "stream_collide": 2, "fs2": 3, "fs3": 4, "fs4": 5}

    def set_cells(self, flags=None, rho=None, u=None):
        flags = None if flags is None else np.ascontiguousarray(flags, dtype=np.uint8)
        rho = None if rho is None else np.ascontiguousarray(rho, dtype=np.float64)
        u = None if u is None else np.ascontiguousarray(u, dtype=np.float64)
        lib().orc_set_cells(self.h, None if flags is None else _u8(flags), _dp(rho), _dp(u))

    def set_time(self, t):
        lib().orc_set_time(self.h, t)

    def run_pass(self, name):
        if self.pre != "orc" or lib().orc_run_pass(self.h, self.PASSES[name]):
            raise ValueError(f"pass {name} not available on this oracle")

    def reconstructed(self):
        rec = np.zeros(9 * self.N)
        mask = np.zeros(self.N, dtype=np.uint16)
        lib().orc_get_reconstructed(self.h, _dp(rec), mask.ctypes.data_as(C.POINTER(C.c_uint16)))
        return rec.reshape(9, self.N), mask

    def time(self):
        return lib().orc_time(self.h)

    def discharge(self, column):
        return lib().orc_discharge(self.h, column)

    def fs_discharge(self, column):
        """sum of phi rho u_x over the wet cells of a column (FP64 oracle)."""
        return lib().orc_fs_discharge(self.h, column)

    def adjoint(self, nsteps, column, t0, t1, ckpt=100):
        g = np.zeros(self.N)
        duin = np.zeros(1)
        j = np.zeros(1)
        rc = lib().orc_adjoint(self.h, nsteps, column, t0, t1, ckpt, _dp(g), _dp(duin), _dp(j))
        if rc:
            raise RuntimeError("orc_adjoint failed")
        return j[0], g, duin[0]

    def objective(self, nsteps, column, t0, t1):
        return lib().orc_objective(self.h, nsteps, column, t0, t1)

    def close(self):
        if getattr(self, "h", None):
            getattr(lib(), self.pre + "_destroy")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
