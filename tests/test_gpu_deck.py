"""A deck drives the GPU path end to end: the damBreak example deck (read by the
native deck reader, units from its OptionsCalc.txt, Smagorinsky on, gravity from
ACCELERATION) steps on the GPU and agrees with the FP64 oracle run from the
same flags and constants; its VTI output holds the fields the solver reads
back.  Bars: flags equal in >= 99.9 % of cells, velocity within 1e-5
absolute (lattice units), mass within 1e-4."""
import re
from pathlib import Path

import numpy as np
import pytest

import flowlb_amd as F
from flowlb_amd import core

pytestmark = pytest.mark.gpu

DECK = Path(__file__).resolve().parent / "golden" / "decks" / "dambreak"


def deck_solver():
    with core.Deck(DECK, core.DECK_CALCULATION) as d:
        flags, info, ui = d.flags(), d.info, d.units_in()
    uo = F.units(**{k: getattr(ui, k) for k, _ in core.UnitsIn._fields_})
    opt = core.default_options(tau=uo.tau, fx=uo.fx, fy=uo.fy, use_force=info.use_force,
                               collision=info.collision, smagorinsky=info.smagorinsky,
                               problem=info.problem)
    s = core.Solver(flags, info.nx, info.ny, opt)
    s.set_fs_fields(*core.init_free_surface_fields(flags))
    return s, flags, info, uo


def test_dambreak_deck_against_oracle(orc):
    s, flags, info, uo = deck_solver()
    with s:
        s.step(300)
        g = s.read()
    o = orc.Oracle(flags, info.nx, info.ny, float(uo.tau), use_force=True, fx=uo.fx, fy=uo.fy,
                   smagorinsky=True, free_surface=True, mass_mode=1, inv_tau=float(uo.inv_tau))
    o.step(300)
    r = o.get()
    same = (g["flags"] == r["flags"]).mean()
    du = np.abs(g["u"] - r["u"]).max()
    dm = np.abs(g["mass"] - r["mass"]).max()
    print(f"\ndamBreak deck 101x101 x 300: tau {uo.tau:.7f} fy {uo.fy:.3e}; flags equal {same:.5f}, "
          f"max|du| {du:.2e}, max|dm| {dm:.2e}")
    assert same >= 0.999
    assert du < 1e-5 and dm < 1e-4


def appended(path):
    raw = path.read_bytes()
    head, data = raw.split(b'<AppendedData encoding="raw">\n_', 1)
    out = {}
    for m in re.finditer(rb'type="(\w+)" Name="(\w+)" NumberOfComponents="(\d)" format="appended" '
                         rb'offset="(\d+)"', head):
        t, name, comps, off = m.group(1), m.group(2).decode(), int(m.group(3)), int(m.group(4))
        n = int(np.frombuffer(data[off:off + 8], np.uint64)[0])
        dt = np.uint8 if t == b"UInt8" else np.float32
        out[name] = np.frombuffer(data[off + 8:off + 8 + n], dt).reshape(-1, comps)
    return out


def test_save_vti_holds_readback(tmp_path):
    s, flags, info, uo = deck_solver()
    with s:
        s.step(20)
        g = s.read()
        s.save_vti(tmp_path / "result_0.vti", spacing=info.size, velocity_scale=2.0, binary=True)
    a = appended(tmp_path / "result_0.vti")
    N = info.nx * info.ny
    np.testing.assert_array_equal(a["Flags"][:, 0], g["flags"])
    np.testing.assert_array_equal(a["U"][:, 0], (g["u"][:N] * 2.0).astype(np.float32))
    np.testing.assert_array_equal(a["rho"][:, 0], g["rho"])
    np.testing.assert_array_equal(a["phi"][:, 0], g["phi"])
