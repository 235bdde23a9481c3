"""CPU checks of the C ABI library: it loads, exports every function the public
headers declare, and its host set-up (flags, free-surface fields, culvert
builder, unit conversion) agrees with the oracle and the reference fixtures.
No GPU call is made here."""
import ctypes
import json
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import flowlb_amd as F
from flowlb_amd import cases, core

ROOT = Path(__file__).resolve().parent.parent
GOLDEN = ROOT / "tests" / "golden"


def declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"\b(flb_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    core.lib()
    out = subprocess.run(["nm", "-D", "--defined-only", str(core.LIB_PATH)], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing
    assert len(declared_functions()) >= 20


def test_prototypes_cover_core_entry_points():
    bound = {p[0] for p in core.PROTOTYPES}
    for n in ("flb_create", "flb_step", "flb_read", "flb_mass", "flb_destroy", "flb_last_error"):
        assert n in bound


def test_no_device_error_is_reported_not_fatal():
    """Without a GPU flb_create returns a status (the reference would exit)."""
    if (Path("/dev/kfd")).exists():
        pytest.skip("a GPU is visible")
    fl, r, u = cases.scaled(cases.SINGLE_PHASE, 8).build()
    with pytest.raises(core.FlbError) as e:
        core.Solver(fl, 64, 16, cases.SINGLE_PHASE.options(), rho0=r, u0=u)
    assert e.value.status == 2


def test_invalid_options_rejected():
    fl = np.full(16, F.FLUID, np.uint8)
    for kw in (dict(tau=0.5), dict(storage=7), dict(block_x=100), dict(collision=3)):
        with pytest.raises(core.FlbError) as e:
            core.Solver(fl, 4, 4, core.default_options(**kw))
        assert e.value.status == 1


def test_inlet_at_last_cell_rejected():
    fl = np.full(16, F.FLUID, np.uint8)
    fl[-1] = F.INLET
    with pytest.raises(core.FlbError):
        core.Solver(fl, 4, 4, core.default_options())


@pytest.mark.parametrize("name", ["flags_channel_41x9", "flags_dambreak_101x101"])
def test_channel_flags_match_fixtures(name):
    meta = json.loads((GOLDEN / "flags_fixtures.json").read_text())[name]
    want = np.load(GOLDEN / f"{name}.npy", allow_pickle=False)
    if name == "flags_channel_41x9":
        got = F.build_channel_flags(meta["nx"], meta["ny"], meta["dx"], F.INLET, F.OUTLET, F.WALL,
                                    F.WALL, free_surface=True)
    else:
        got = F.build_channel_flags(meta["nx"], meta["ny"], meta["dx"], F.WALL, F.WALL, F.WALL,
                                    F.WALL, free_surface=True, obstacles=[(0.45, 0.05, 0.1, 0.05)],
                                    initial=[(0.2, 0.35, 0.5, 0.2)])
    np.testing.assert_array_equal(got, want)


def test_culvert_flags_match_oracle(orc):
    for case in (cases.SINGLE_PHASE, cases.scaled(cases.FS_4096, 8), cases.FS_1024):
        fl, _, _ = case.build()
        emb = case.emb
        obst = [(0.5 * (emb[0] + emb[1]), 0.5 * (case.opening + 1 + case.ny - 1),
                 case.ny - 1 - (case.opening + 1), emb[1] - emb[0])]
        ini = []
        if case.water[0]:
            ini = [(0.5 * (1 + case.water[0]), 0.5 * (1 + case.water[1]), case.water[1] - 1,
                    case.water[0] - 1)]
        want = orc.build_channel_flags(case.nx, case.ny, 1.0, *case.bc,
                                       free_surface=case.free_surface, obstacles=obst, initial=ini)
        np.testing.assert_array_equal(fl, want)


def test_single_phase_case_counts():
    fl, rho0, u0 = cases.SINGLE_PHASE.build()
    assert (fl == F.WALL).sum() == 7134           # 2 x 512 rows + 65 x 94 embankment
    assert (fl == F.FLUID).sum() + (fl == F.INLET).sum() + (fl == F.OUTLET).sum() == 58402
    assert np.abs(rho0 - 1).max() <= 1e-4 + 1e-7 and np.abs(rho0 - 1).max() > 9e-5  # float32 ulp
    N = fl.size
    np.testing.assert_array_equal(u0[:N][fl == F.INLET], np.float32(0.05))
    assert (u0[:N][fl != F.INLET] == 0).all() and (u0[N:] == 0).all()
    # mt19937_64 draws are reproducible
    _, rho1, _ = cases.SINGLE_PHASE.build()
    np.testing.assert_array_equal(rho0, rho1)


def test_free_surface_fields_match_oracle(orc):
    fl, _, _ = cases.scaled(cases.FS_4096, 16).build()
    m, p, e = F.init_free_surface_fields(fl)
    mo, po, eo = orc.init_free_surface_fields(fl)
    np.testing.assert_array_equal(m, mo)
    np.testing.assert_array_equal(p, po)
    np.testing.assert_array_equal(e, eo)


def test_units_closed_forms():
    """Units (src/utils.cpp): L = len/(nx-1), T = L u_LB/u_SI, tau = 3 nu T/L^2 + 1/2,
    F = rho g L^2 T^2 / M with M = rho L^3."""
    o = F.units(length_si=1.0, nx=101, ref_velocity_si=10.0, ref_velocity_lb=0.1, density_si=1000.0,
                relaxation_time=0.0, kinematic_viscosity=1e-6, use_force=1, acceleration_x=0.0,
                acceleration_y=-9.80665, volumetric_force_x=0.0, volumetric_force_y=0.0,
                velocity_si_x=1.0, velocity_si_y=0.0)
    L = 0.01
    T = L * 0.1 / 10.0
    assert abs(o.length_scale - L) < 1e-15 and abs(o.time_scale - T) < 1e-18
    nu = np.float32(1e-6 * T / (L * L))
    assert o.tau == np.float32(np.float32(3.0) * nu + np.float32(0.5))
    assert abs(o.fy - (-9.80665 * T * T / L)) < 1e-7 * 9.80665 * T * T / L
    assert abs(o.velocity_lb_x - 1.0 / (L / T)) < 1e-12
    t2 = F.units(length_si=1.0, nx=101, ref_velocity_si=10.0, ref_velocity_lb=0.1, density_si=1.0,
                 relaxation_time=0.56, kinematic_viscosity=0.0, use_force=0, acceleration_x=0.0,
                 acceleration_y=0.0, volumetric_force_x=0.0, volumetric_force_y=0.0,
                 velocity_si_x=0.0, velocity_si_y=0.0)
    assert t2.tau == np.float32(0.56) and t2.fx == 0 and t2.fy == 0


def test_batch_instances_reproducible():
    a = cases.batch_instances(8, seed=1)
    b = cases.batch_instances(8, seed=1)
    assert [(c.nx, c.ny, c.emb, c.opening) for c in a] == [(c.nx, c.ny, c.emb, c.opening) for c in b]
    for c in a:
        assert 256 <= c.nx <= 2048 and 64 <= c.ny <= 512
        assert 0.52 <= c.tau <= 0.8 and 0.02 <= c.u_inlet <= 0.08
