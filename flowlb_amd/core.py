"""ctypes binding of libflbcore.so (the C ABI in include/flb_core.h).

The library must be built (python -m flowlb_amd.build); there is no fallback
implementation: loading fails loudly when the shared object is missing.
"""
import ctypes as C  # noqa: N812
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libflbcore.so"

FLUID, GAS, INTERFACE, WALL, MOVING_WALL, INLET, OUTLET, GAS_INTERFACE = (
    0x01, 0x02, 0x04, 0x08, 0x10, 0x20, 0x40, 0x80)
STORAGE = {"fp32": 0, "fp16s": 1, "bf16s": 2, "fp64": 3}
MONOFLUID, FREE_SURFACE = 0, 1     # flb_options.problem
SRT, TRT = 0, 1                    # flb_options.collision
K_STREAM_COLLIDE, K_FS1, K_FS_UPDATE = range(3)
KERNEL_NAMES = ["stream_collide", "fs1", "fs_update"]


class FlbError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{msg} (status {status})")
        self.status = status


class Options(C.Structure):
    _fields_ = [("tau", C.c_float), ("fx", C.c_float), ("fy", C.c_float),
                ("use_force", C.c_int32), ("collision", C.c_int32), ("smagorinsky", C.c_int32),
                ("problem", C.c_int32), ("storage", C.c_int32),
                ("gas_check_corrected", C.c_int32), ("mass_mode", C.c_int32),
                ("block_x", C.c_int32), ("device", C.c_int32), ("row_pitch", C.c_int32),
                ("whole_grid", C.c_int32)]


class IsoView(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("x_min", C.c_uint32),
                ("x_max", C.c_uint32), ("y_min", C.c_uint32), ("y_max", C.c_uint32),
                ("iso_value", C.c_float), ("field", C.c_int32), ("clear", C.c_int32)]


ISO_VELOCITY, ISO_PHI = 0, 1


class Rect(C.Structure):
    _fields_ = [("x0", C.c_double), ("y0", C.c_double), ("height", C.c_double),
                ("width", C.c_double)]


class Channel(C.Structure):
    _fields_ = [("nx", C.c_uint32), ("ny", C.c_uint32), ("dx", C.c_double),
                ("x_min", C.c_double), ("y_min", C.c_double), ("free_surface", C.c_int32),
                ("bc_left", C.c_uint8), ("bc_right", C.c_uint8), ("bc_up", C.c_uint8),
                ("bc_down", C.c_uint8), ("n_obstacles", C.c_uint32), ("n_initial", C.c_uint32),
                ("obstacles", C.POINTER(Rect)), ("initial", C.POINTER(Rect))]


class Culvert(C.Structure):
    _fields_ = [("nx", C.c_uint32), ("ny", C.c_uint32), ("free_surface", C.c_int32),
                ("bc_left", C.c_uint8), ("bc_right", C.c_uint8), ("bc_up", C.c_uint8),
                ("bc_down", C.c_uint8), ("emb_x0", C.c_uint32), ("emb_x1", C.c_uint32),
                ("opening", C.c_uint32), ("water_x1", C.c_uint32), ("water_depth", C.c_uint32),
                ("u_inlet", C.c_float), ("rho_noise", C.c_double), ("seed", C.c_uint64)]


class AdjointSpec(C.Structure):
    _fields_ = [("nsteps", C.c_uint64), ("t0", C.c_uint64), ("t1", C.c_uint64),
                ("column", C.c_uint32), ("ckpt", C.c_uint32),
                ("trace_budget_bytes", C.c_uint64)]


class Instance(C.Structure):
    _fields_ = [("nx", C.c_uint32), ("ny", C.c_uint32), ("tau", C.c_float), ("fx", C.c_float),
                ("fy", C.c_float), ("use_force", C.c_int32), ("flags", C.POINTER(C.c_uint8)),
                ("rho0", C.POINTER(C.c_float)), ("u0", C.POINTER(C.c_float))]


class VtiField(C.Structure):
    _fields_ = [("name", C.c_char_p), ("type", C.c_int32), ("components", C.c_uint32),
                ("data", C.c_void_p), ("scale", C.c_double), ("flip_y", C.c_int32)]


class UnitsIn(C.Structure):
    _fields_ = [("length_si", C.c_double), ("nx", C.c_uint32), ("ref_velocity_si", C.c_double),
                ("ref_velocity_lb", C.c_double), ("density_si", C.c_double),
                ("relaxation_time", C.c_double), ("kinematic_viscosity", C.c_double),
                ("use_force", C.c_int32), ("acceleration_x", C.c_double),
                ("acceleration_y", C.c_double), ("volumetric_force_x", C.c_double),
                ("volumetric_force_y", C.c_double), ("velocity_si_x", C.c_double),
                ("velocity_si_y", C.c_double)]


class UnitsOut(C.Structure):
    _fields_ = [("tau", C.c_float), ("inv_tau", C.c_float), ("fx", C.c_float), ("fy", C.c_float),
                ("velocity_lb_x", C.c_double), ("velocity_lb_y", C.c_double),
                ("length_scale", C.c_double), ("time_scale", C.c_double),
                ("mass_scale", C.c_double)]


class DeckInfo(C.Structure):
    _fields_ = [("nx", C.c_uint32), ("ny", C.c_uint32), ("x_min", C.c_double),
                ("x_max", C.c_double), ("y_min", C.c_double), ("y_max", C.c_double),
                ("size", C.c_double), ("n_points", C.c_uint32), ("n_cdws", C.c_uint32),
                ("n_obstacles", C.c_uint32), ("n_initial", C.c_uint32), ("problem", C.c_int32),
                ("bc_left", C.c_uint8), ("bc_right", C.c_uint8), ("bc_up", C.c_uint8),
                ("bc_down", C.c_uint8), ("collision", C.c_int32), ("smagorinsky", C.c_int32),
                ("surface_tension", C.c_int32), ("use_force", C.c_int32),
                ("precision", C.c_int32), ("block_x", C.c_int32), ("block_y", C.c_int32),
                ("block_z", C.c_int32), ("graphics", C.c_int32),
                ("ref_velocity_lb", C.c_double), ("ref_velocity_si", C.c_double),
                ("flow", C.c_double), ("relaxation_time", C.c_double),
                ("kinematic_viscosity", C.c_double), ("density", C.c_double),
                ("velocity_si", C.c_double * 3), ("acceleration", C.c_double * 3),
                ("volumetric_force", C.c_double * 3), ("time_simulation", C.c_float),
                ("time_save", C.c_float)]


DECK_DOMAIN, DECK_CALCULATION = 0, 1

# Exported symbols and their prototypes: (name, restype, argtypes).
_H = C.c_void_p
_F = C.POINTER(C.c_float)
_D = C.POINTER(C.c_double)
_U8 = C.POINTER(C.c_uint8)
_U64 = C.POINTER(C.c_uint64)
PROTOTYPES = [
    ("flb_default_options", None, [C.POINTER(Options)]),
    ("flb_create", C.c_int, [C.POINTER(Options), _U8, C.c_uint32, C.c_uint32, _F, _F,
                             C.POINTER(_H)]),
    ("flb_set_fs_fields", C.c_int, [_H, _F, _F, _F]),
    ("flb_set_fs_fields_f64", C.c_int, [_H, _D, _D, _D]),
    ("flb_step", C.c_int, [_H, C.c_uint64]),
    ("flb_sync", C.c_int, [_H]),
    ("flb_read", C.c_int, [_H, _F, _F, _U8, _F]),
    ("flb_read_fs", C.c_int, [_H, _F, _F]),
    ("flb_read_populations", C.c_int, [_H, _F]),
    ("flb_read_f64", C.c_int, [_H, _D, _D, _U8, _D]),
    ("flb_read_fs_f64", C.c_int, [_H, _D, _D]),
    ("flb_read_populations_f64", C.c_int, [_H, _D]),
    ("flb_write_populations", C.c_int, [_H, _F]),
    ("flb_write_populations_f64", C.c_int, [_H, _D]),
    ("flb_device_fields", C.c_int, [_H, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                     C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]),
    ("flb_mass", C.c_int, [_H, _D]),
    ("flb_export_velocity", C.c_int, [_H, C.c_void_p, C.c_size_t]),
    ("flb_bandwidth_probe", C.c_int, [C.c_int32, C.c_uint32, C.c_uint32, C.c_int32, C.c_uint32,
                                      _F, _D]),
    ("flb_isosurface", C.c_int, [_H, C.POINTER(IsoView), C.c_void_p, C.c_size_t]),
    ("flb_time", C.c_uint64, [_H]),
    ("flb_step_timed", C.c_int, [_H, C.c_uint64, _F, _F]),
    ("flb_flag_histogram", C.c_int, [_H, _U64]),
    ("flb_flag_histogram_noalloc", C.c_int, [_H, _U64]),
    ("flb_last_error", C.c_char_p, [_H]),
    ("flb_status_string", C.c_char_p, [C.c_int]),
    ("flb_destroy", None, [_H]),
    ("flb_set_brinkman", C.c_int, [_H, _F]),
    ("flb_objective", C.c_int, [_H, C.POINTER(AdjointSpec), _D]),
    ("flb_adjoint", C.c_int, [_H, C.POINTER(AdjointSpec), _D, _F, _D, _F]),
    ("flb_batch_create", C.c_int, [C.POINTER(Instance), C.c_uint32, C.c_int32, C.c_int32,
                                   C.POINTER(_H)]),
    ("flb_batch_configure", C.c_int, [_H, C.c_int32, C.c_int32, C.c_int32]),
    ("flb_batch_configure_persistent", C.c_int, [_H, C.c_int32, C.c_int32, C.c_int32]),
    ("flb_batch_set_order", C.c_int, [_H, C.c_int32]),
    ("flb_batch_order", C.c_int32, [_H]),
    ("flb_batch_autotune", C.c_int, [_H, C.c_uint32, C.POINTER(C.c_int32),
                                     C.POINTER(C.c_int32), _F]),
    ("flb_batch_step", C.c_int, [_H, C.c_uint64]),
    ("flb_batch_sync", C.c_int, [_H]),
    ("flb_batch_step_timed", C.c_int, [_H, C.c_uint64, _F]),
    ("flb_batch_read", C.c_int, [_H, C.c_uint32, _F, _F]),
    ("flb_batch_read_populations", C.c_int, [_H, C.c_uint32, _F]),
    ("flb_batch_time", C.c_uint64, [_H]),
    ("flb_batch_info", C.c_int, [_H, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                 C.POINTER(C.c_int32), C.POINTER(C.c_uint32)]),
    ("flb_batch_last_error", C.c_char_p, [_H]),
    ("flb_batch_destroy", None, [_H]),
    ("flb_write_vti", C.c_int, [C.c_char_p, C.c_uint32, C.c_uint32, C.c_double, C.c_double,
                                C.c_double, C.POINTER(VtiField), C.c_uint32, C.c_int32]),
    ("flb_write_vti_series", C.c_int, [C.c_char_p, C.c_uint32, C.c_double]),
    ("flb_save_vti", C.c_int, [_H, C.c_char_p, C.c_double, C.c_double, C.c_double, C.c_double,
                               C.c_double, C.c_int32]),
    ("flb_get_shape", C.c_int, [_H, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                C.POINTER(C.c_int32)]),
    ("flb_get_layout", C.c_int, [_H, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64)]),
    ("flb_build_channel_flags", C.c_int, [C.POINTER(Channel), _U8]),
    ("flb_seed_free_surface", C.c_int, [_U8, C.c_uint32, C.c_uint32]),
    ("flb_init_free_surface_fields", C.c_int, [_U8, C.c_size_t, _F, _F, _F]),
    ("flb_build_culvert", C.c_int, [C.POINTER(Culvert), _U8, _F, _F]),
    ("flb_units", C.c_int, [C.POINTER(UnitsIn), C.POINTER(UnitsOut)]),
    ("flb_deck_load", C.c_int, [C.c_char_p, C.c_int, C.POINTER(_H), C.c_char_p, C.c_size_t]),
    ("flb_deck_get_info", C.c_int, [_H, C.POINTER(DeckInfo)]),
    ("flb_deck_build_flags", C.c_int, [_H, _U8]),
    ("flb_deck_units_in", C.c_int, [_H, C.POINTER(UnitsIn)]),
    ("flb_deck_free", None, [_H]),
]

_lib = None
_hip = None


def hip():
    """The HIP runtime the core library itself links (libamdhip64), for the
    device buffers a caller hands to the viewer entry points."""
    global _hip
    if _hip is None:
        lib()
        H = C.CDLL("libamdhip64.so.7")
        H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        H.hipFree.argtypes = [C.c_void_p]
        H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        H.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        H.hipDeviceSynchronize.argtypes = []
        _hip = H
    return _hip


PROBE_KINDS = {"esoteric_inplace": 0, "copy9": 1, "inplace9": 2, "esoteric_inplace16": 3}


def bandwidth_probe(nx, ny, kind, launches=20, device=0):
    """Device time per launch (ms) and bytes moved per launch of the
    stream-collide memory pattern without arithmetic (flb_bandwidth_probe)."""
    ms, nbytes = C.c_float(), C.c_double()
    check(lib().flb_bandwidth_probe(device, nx, ny, PROBE_KINDS[kind], launches, C.byref(ms),
                                    C.byref(nbytes)), None, "flb_bandwidth_probe")
    return ms.value, nbytes.value


class DeviceBuffer:
    """A plain device allocation (hipMalloc) with host copies, standing in for
    the GL pixel buffer a viewer maps through HIP-GL interop."""
    H2D, D2H = 1, 2

    def __init__(self, nbytes, fill=None):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        if hip().hipMalloc(C.byref(p), self.nbytes) != 0:
            raise FlbError(-1, f"hipMalloc of {self.nbytes} bytes failed")
        self.ptr = p.value
        # The solver's stream does not wait for the null stream, so the fill
        # (and from_host's copy) completes before a solver call can use the buffer.
        if fill is not None and (hip().hipMemset(C.c_void_p(self.ptr), int(fill), self.nbytes)
                                 or hip().hipDeviceSynchronize()):
            raise FlbError(-1, "hipMemset failed")

    def to_host(self, dtype=np.uint8):
        out = np.empty(self.nbytes, np.uint8)
        if hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), self.nbytes,
                           self.D2H):
            raise FlbError(-1, "hipMemcpy to host failed")
        return out.view(dtype)

    def from_host(self, a):
        a = np.ascontiguousarray(a)
        if a.nbytes > self.nbytes:
            raise ValueError("array larger than the buffer")
        if (hip().hipMemcpy(C.c_void_p(self.ptr), a.ctypes.data_as(C.c_void_p), a.nbytes, self.H2D)
                or hip().hipDeviceSynchronize()):
            raise FlbError(-1, "hipMemcpy to device failed")

    def close(self):
        if getattr(self, "ptr", None):
            hip().hipFree(C.c_void_p(self.ptr))
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def lib():
    """Loads the HIP library; raises if it was not built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m flowlb_amd.build`")
        L = C.CDLL(str(LIB_PATH))
        for name, res, args in PROTOTYPES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _register_optional(L)
        _lib = L
    return _lib


_optional_hooks = []


def _register_optional(L):
    for hook in _optional_hooks:
        hook(L)


def fptr(a):
    return None if a is None else a.ctypes.data_as(_F)


def u8ptr(a):
    return None if a is None else a.ctypes.data_as(_U8)


def _f32(a, n=None):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float32)
    if n is not None and a.size != n:
        raise ValueError(f"expected {n} values, got {a.size}")
    return a


def check(status, h=None, what=""):
    if status != 0:
        msg = lib().flb_last_error(h).decode() if h else ""
        name = lib().flb_status_string(status).decode()
        raise FlbError(status, f"{what}: {name}{': ' + msg if msg else ''}")


def default_options(**kw):
    o = Options()
    lib().flb_default_options(C.byref(o))
    for k, v in kw.items():
        if k == "storage" and isinstance(v, str):
            v = STORAGE[v]
        setattr(o, k, v)
    return o


class Solver:
    """One simulation on the GPU: the role h_launchCudaCalculations2D plays in
    the reference (src/cuda/runCuda.cu:31-128), split into calls."""

    def __init__(self, flags, nx, ny, options=None, rho0=None, u0=None, **kw):
        N = nx * ny
        self.nx, self.ny, self.N = nx, ny, N
        self.options = options if options is not None else default_options(**kw)
        self.flags0 = np.ascontiguousarray(flags, dtype=np.uint8)
        if self.flags0.size != N:
            raise ValueError("flags must have nx*ny entries")
        rho0 = _f32(rho0, N)
        u0 = _f32(u0, 2 * N)
        h = C.c_void_p()
        st = lib().flb_create(C.byref(self.options), u8ptr(self.flags0), nx, ny, fptr(rho0),
                              fptr(u0), C.byref(h))
        check(st, None, "flb_create")
        self.h = h
        self.free_surface = self.options.problem == 1

    # -- set-up
    def set_fs_fields(self, mass, phi, excess=None):
        mass, phi, excess = _f32(mass, self.N), _f32(phi, self.N), _f32(excess, self.N)
        check(lib().flb_set_fs_fields(self.h, fptr(mass), fptr(phi), fptr(excess)), self.h,
              "flb_set_fs_fields")

    def set_fs_fields64(self, mass, phi, excess=None):
        """set_fs_fields from doubles (exact for FP64 storage)."""
        f64 = lambda a: None if a is None else np.ascontiguousarray(a, np.float64).reshape(-1)  # noqa: E731
        mass, phi, excess = f64(mass), f64(phi), f64(excess)
        for a in (mass, phi, excess):
            if a is not None and a.size != self.N:
                raise ValueError(f"expected {self.N} values, got {a.size}")
        P = lambda a: None if a is None else a.ctypes.data_as(_D)  # noqa: E731
        check(lib().flb_set_fs_fields_f64(self.h, P(mass), P(phi), P(excess)), self.h,
              "flb_set_fs_fields_f64")

    # -- advance
    def step(self, n=1):
        check(lib().flb_step(self.h, n), self.h, "flb_step")

    def sync(self):
        check(lib().flb_sync(self.h), self.h, "flb_sync")

    def step_timed(self, n, per_kernel=False):
        tot = C.c_float()
        if per_kernel:
            arr = (C.c_float * len(KERNEL_NAMES))()
            check(lib().flb_step_timed(self.h, n, C.byref(tot), arr), self.h, "flb_step_timed")
            return tot.value, list(arr)
        check(lib().flb_step_timed(self.h, n, C.byref(tot), None), self.h, "flb_step_timed")
        return tot.value

    @property
    def time(self):
        return lib().flb_time(self.h)

    @property
    def layout(self):
        """(row pitch, plane size) of the device fields (flb_get_layout)."""
        pitch, plane = C.c_uint32(), C.c_uint64()
        check(lib().flb_get_layout(self.h, C.byref(pitch), C.byref(plane)), self.h, "flb_get_layout")
        return pitch.value, plane.value

    # -- readback
    def read(self):
        N = self.N
        u = np.empty(2 * N, np.float32)
        rho = np.empty(N, np.float32)
        fl = np.empty(N, np.uint8)
        phi = np.empty(N, np.float32) if self.free_surface else None
        check(lib().flb_read(self.h, fptr(u), fptr(rho), u8ptr(fl), fptr(phi)), self.h, "flb_read")
        out = dict(u=u, rho=rho, flags=fl)
        if phi is not None:
            out["phi"] = phi
            m = np.empty(N, np.float32)
            e = np.empty(N, np.float32)
            check(lib().flb_read_fs(self.h, fptr(m), fptr(e)), self.h, "flb_read_fs")
            out["mass"], out["excess"] = m, e
        return out

    def populations(self):
        f = np.empty(9 * self.N, np.float32)
        check(lib().flb_read_populations(self.h, fptr(f)), self.h, "flb_read_populations")
        return f

    def read64(self):
        """read() in double precision (exact for FP64 storage)."""
        N = self.N
        d = lambda n: np.empty(n, np.float64)
        u, rho, fl = d(2 * N), d(N), np.empty(N, np.uint8)
        phi = d(N) if self.free_surface else None
        P = lambda a: None if a is None else a.ctypes.data_as(_D)
        check(lib().flb_read_f64(self.h, P(u), P(rho), u8ptr(fl), P(phi)), self.h, "flb_read_f64")
        out = dict(u=u, rho=rho, flags=fl)
        if phi is not None:
            out["phi"] = phi
            m, e = d(N), d(N)
            check(lib().flb_read_fs_f64(self.h, P(m), P(e)), self.h, "flb_read_fs_f64")
            out["mass"], out["excess"] = m, e
        return out

    def populations64(self):
        f = np.empty(9 * self.N, np.float64)
        check(lib().flb_read_populations_f64(self.h, f.ctypes.data_as(_D)), self.h,
              "flb_read_populations_f64")
        return f

    def write_populations(self, f):
        f = _f32(f, 9 * self.N)
        check(lib().flb_write_populations(self.h, fptr(f)), self.h, "flb_write_populations")

    def write_populations64(self, f):
        f = np.ascontiguousarray(f, np.float64).reshape(-1)
        if f.size != 9 * self.N:
            raise ValueError(f"expected {9 * self.N} values, got {f.size}")
        check(lib().flb_write_populations_f64(self.h, f.ctypes.data_as(_D)), self.h,
              "flb_write_populations_f64")

    def mass(self):
        m = C.c_double()
        check(lib().flb_mass(self.h, C.byref(m)), self.h, "flb_mass")
        return m.value

    def flag_histogram(self):
        c = np.zeros(256, np.uint64)
        check(lib().flb_flag_histogram(self.h, c.ctypes.data_as(_U64)), self.h, "flb_flag_histogram")
        return c

    # -- viewer output (device buffers owned by the caller)
    def export_velocity(self, dst, pitch_bytes=None):
        """Writes (ux, uy) float2 texels into the DeviceBuffer dst (rows of
        pitch_bytes, default 8 * nx) and waits for it."""
        pitch = 8 * self.nx if pitch_bytes is None else pitch_bytes
        if dst.nbytes < pitch * self.ny:
            raise ValueError("destination too small")
        check(lib().flb_export_velocity(self.h, C.c_void_p(dst.ptr), pitch), self.h,
              "flb_export_velocity")
        self.sync()

    def isosurface(self, tex, width, height, bounds, iso, field=ISO_VELOCITY, clear=True,
                   pitch_bytes=None):
        """Marching-squares isoline raster into the DeviceBuffer tex (rows of
        pitch_bytes, default width); bounds = (x_min, x_max, y_min, y_max)."""
        pitch = width if pitch_bytes is None else pitch_bytes
        if tex.nbytes < pitch * height:
            raise ValueError("texture buffer too small")
        v = IsoView(width, height, bounds[0], bounds[1], bounds[2], bounds[3], iso, field,
                    int(clear))
        check(lib().flb_isosurface(self.h, C.byref(v), C.c_void_p(tex.ptr), pitch), self.h,
              "flb_isosurface")
        self.sync()

    def flag_histogram_noalloc(self):
        """flag_histogram through flb_flag_histogram_noalloc (no device allocation)."""
        c = np.zeros(256, np.uint64)
        check(lib().flb_flag_histogram_noalloc(self.h, c.ctypes.data_as(_U64)), self.h,
              "flb_flag_histogram_noalloc")
        return c

    def device_fields(self):
        ptrs = [C.c_void_p() for _ in range(4)]
        check(lib().flb_device_fields(self.h, *[C.byref(p) for p in ptrs]), self.h,
              "flb_device_fields")
        return [p.value for p in ptrs]

    # -- adjoint
    def set_brinkman(self, alpha):
        alpha = _f32(alpha, self.N)
        check(lib().flb_set_brinkman(self.h, fptr(alpha)), self.h, "flb_set_brinkman")

    def objective(self, nsteps, column, t0, t1, ckpt=100):
        spec = AdjointSpec(nsteps, t0, t1, column, ckpt)
        j = C.c_double()
        check(lib().flb_objective(self.h, C.byref(spec), C.byref(j)), self.h, "flb_objective")
        return j.value

    def adjoint(self, nsteps, column, t0, t1, ckpt=100, trace_budget_bytes=0):
        """Returns J, dJ/dalpha (N floats), dJ/du_inlet and the phase times (ms):
        forward sweep, recompute of the segments without a resident trace,
        reverse sweep.  trace_budget_bytes: 0 = half the free device memory."""
        spec = AdjointSpec(nsteps, t0, t1, column, ckpt, trace_budget_bytes)
        j = C.c_double()
        duin = C.c_double()
        g = np.zeros(self.N, np.float32)
        ph = (C.c_float * 3)()
        check(lib().flb_adjoint(self.h, C.byref(spec), C.byref(j), fptr(g), C.byref(duin), ph),
              self.h, "flb_adjoint")
        return j.value, g, duin.value, list(ph)

    def save_vti(self, path, spacing=1.0, origin=(0.0, 0.0), velocity_scale=1.0,
                 density_scale=1.0, binary=False):
        check(lib().flb_save_vti(self.h, str(path).encode(), origin[0], origin[1], spacing,
                                 velocity_scale, density_scale, int(binary)), self.h, "flb_save_vti")

    def close(self):
        if getattr(self, "h", None):
            lib().flb_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many single-phase culverts advanced by one launch per step (flb_batch_*)."""

    def __init__(self, instances, storage="fp32", device=0):
        """instances: list of dicts with nx, ny, tau, flags, rho0, u0 (and fx,
        fy, use_force)."""
        self._keep = []
        arr = (Instance * len(instances))()
        self.sizes = []
        for k, d in enumerate(instances):
            fl = np.ascontiguousarray(d["flags"], dtype=np.uint8)
            r0 = _f32(d.get("rho0"))
            u0 = _f32(d.get("u0"))
            self._keep += [fl, r0, u0]
            arr[k] = Instance(d["nx"], d["ny"], d["tau"], d.get("fx", 0.0), d.get("fy", 0.0),
                              int(d.get("use_force", 0)), u8ptr(fl), fptr(r0), fptr(u0))
            self.sizes.append((d["nx"], d["ny"]))
        h = C.c_void_p()
        st = lib().flb_batch_create(arr, len(instances), STORAGE[storage] if isinstance(storage, str)
                                    else storage, device, C.byref(h))
        if st:
            raise FlbError(st, "flb_batch_create: " + lib().flb_status_string(st).decode())
        self.h = h

    def _check(self, st, what):
        if st:
            raise FlbError(st, f"{what}: {lib().flb_status_string(st).decode()}: "
                               f"{lib().flb_batch_last_error(self.h).decode()}")

    def configure(self, block_x=256, grid=2048, graph_chunk=100):
        self._check(lib().flb_batch_configure(self.h, block_x, grid, graph_chunk), "configure")

    def configure_persistent(self, block_x=256, blocks_per_cu=2, chunk=100):
        """One cooperative launch per `chunk` steps (grid barrier between steps)."""
        self._check(lib().flb_batch_configure_persistent(self.h, block_x, blocks_per_cu, chunk),
                    "configure_persistent")

    def autotune(self, steps=20):
        b, g, ms = C.c_int32(), C.c_int32(), C.c_float()
        self._check(lib().flb_batch_autotune(self.h, steps, C.byref(b), C.byref(g), C.byref(ms)),
                    "autotune")
        return b.value, g.value, ms.value

    def step(self, n=1):
        self._check(lib().flb_batch_step(self.h, n), "step")

    def sync(self):
        self._check(lib().flb_batch_sync(self.h), "sync")

    def step_timed(self, n):
        ms = C.c_float()
        self._check(lib().flb_batch_step_timed(self.h, n, C.byref(ms)), "step_timed")
        return ms.value

    def read(self, i):
        nx, ny = self.sizes[i]
        u = np.empty(2 * nx * ny, np.float32)
        rho = np.empty(nx * ny, np.float32)
        self._check(lib().flb_batch_read(self.h, i, fptr(u), fptr(rho)), "read")
        return dict(u=u, rho=rho)

    def populations(self, i):
        nx, ny = self.sizes[i]
        f = np.empty(9 * nx * ny, np.float32)
        self._check(lib().flb_batch_read_populations(self.h, i, fptr(f)), "read_populations")
        return f

    def info(self):
        b, g, c, t = C.c_int32(), C.c_int32(), C.c_int32(), C.c_uint32()
        self._check(lib().flb_batch_info(self.h, C.byref(b), C.byref(g), C.byref(c), C.byref(t)),
                    "info")
        return dict(block_x=b.value, grid=g.value, graph_chunk=c.value, tiles=t.value,
                    order=["rows", "instance", "bands"][lib().flb_batch_order(self.h)])

    def set_order(self, order):
        """Tile order of the packed list: "rows", "instance" or "bands"."""
        self._check(lib().flb_batch_set_order(self.h, ["rows", "instance", "bands"].index(order)),
                    "set_order")

    @property
    def time(self):
        return lib().flb_batch_time(self.h)

    def close(self):
        if getattr(self, "h", None):
            lib().flb_batch_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- host set-up

def build_channel_flags(nx, ny, dx, bc_left, bc_right, bc_up, bc_down, free_surface=False,
                        obstacles=(), initial=(), x_min=0.0, y_min=0.0):
    obs = (Rect * max(1, len(obstacles)))(*[Rect(*r) for r in obstacles])
    ini = (Rect * max(1, len(initial)))(*[Rect(*r) for r in initial])
    g = Channel(nx, ny, dx, x_min, y_min, int(free_surface), bc_left, bc_right, bc_up, bc_down,
                len(obstacles), len(initial), obs, ini)
    out = np.zeros(nx * ny, np.uint8)
    check(lib().flb_build_channel_flags(C.byref(g), u8ptr(out)), None, "flb_build_channel_flags")
    return out


def seed_free_surface(flags, nx, ny):
    flags = np.array(flags, dtype=np.uint8, copy=True)
    check(lib().flb_seed_free_surface(u8ptr(flags), nx, ny), None, "flb_seed_free_surface")
    return flags


def init_free_surface_fields(flags):
    flags = np.ascontiguousarray(flags, dtype=np.uint8)
    n = flags.size
    m, p, e = (np.empty(n, np.float32) for _ in range(3))
    check(lib().flb_init_free_surface_fields(u8ptr(flags), n, fptr(m), fptr(p), fptr(e)), None,
          "flb_init_free_surface_fields")
    return m, p, e


def build_culvert(nx, ny, free_surface=False, bc=(INLET, OUTLET, WALL, WALL), emb=(0, 0),
                  opening=0, water=(0, 0), u_inlet=0.0, rho_noise=0.0, seed=0):
    """Returns flags, rho0, u0 of a synthetic culvert (flb_build_culvert)."""
    c = Culvert(nx, ny, int(free_surface), bc[0], bc[1], bc[2], bc[3], emb[0], emb[1], opening,
                water[0], water[1], u_inlet, rho_noise, seed)
    N = nx * ny
    fl = np.zeros(N, np.uint8)
    rho0 = np.empty(N, np.float32)
    u0 = np.empty(2 * N, np.float32)
    check(lib().flb_build_culvert(C.byref(c), u8ptr(fl), fptr(rho0), fptr(u0)), None,
          "flb_build_culvert")
    return fl, rho0, u0


def write_vti(path, nx, ny, fields, origin=(0.0, 0.0), spacing=1.0, binary=False):
    """fields: list of (name, array, scale, flip_y); uint8 arrays are written as
    UInt8, float arrays as Float32, arrays of 2N values as 2-component vectors."""
    keep = []
    arr = (VtiField * max(1, len(fields)))()
    N = nx * ny
    for k, (name, data, scale, flip) in enumerate(fields):
        if data.dtype == np.uint8:
            d = np.ascontiguousarray(data)
            t = 0
        else:
            d = np.ascontiguousarray(data, dtype=np.float32)
            t = 1
        comps = d.size // N
        keep.append(d)
        arr[k] = VtiField(name.encode(), t, comps, d.ctypes.data, scale, int(flip))
    check(lib().flb_write_vti(str(path).encode(), nx, ny, origin[0], origin[1], spacing, arr,
                              len(fields), int(binary)), None, "flb_write_vti")


def write_vti_series(path, count, time_interval):
    check(lib().flb_write_vti_series(str(path).encode(), count, time_interval), None,
          "flb_write_vti_series")


def units(**kw):
    i = UnitsIn(**kw)
    o = UnitsOut()
    check(lib().flb_units(C.byref(i), C.byref(o)), None, "flb_units")
    return o


class Deck:
    """A FlowLB case directory (Geometry.txt, OptionsCalc.txt, CSV outlines)
    read by the native reader; `flags()` gives the domain's cell flags."""

    def __init__(self, directory, mode=DECK_DOMAIN):
        h = _H()
        err = C.create_string_buffer(512)
        st = lib().flb_deck_load(str(directory).encode(), mode, C.byref(h), err, len(err))
        if st != 0:
            raise FlbError(st, f"flb_deck_load: {lib().flb_status_string(st).decode()}: "
                               f"{err.value.decode()}")
        self.h = h
        self.info = DeckInfo()
        check(lib().flb_deck_get_info(self.h, C.byref(self.info)), None, "flb_deck_get_info")

    @property
    def nx(self):
        return self.info.nx

    @property
    def ny(self):
        return self.info.ny

    def flags(self):
        out = np.empty(self.info.nx * self.info.ny, np.uint8)
        check(lib().flb_deck_build_flags(self.h, u8ptr(out)), None, "flb_deck_build_flags")
        return out

    def units_in(self):
        i = UnitsIn()
        check(lib().flb_deck_units_in(self.h, C.byref(i)), None, "flb_deck_units_in")
        return i

    def close(self):
        if self.h:
            lib().flb_deck_free(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
