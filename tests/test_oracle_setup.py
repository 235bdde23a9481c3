"""Pins the oracle's set-up (flag generation, free-surface seeding and fields)
against the flag files held by the reference's test decks."""
import json
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden"
FLUID, GAS, INTERFACE, WALL, INLET, OUTLET = 0x01, 0x02, 0x04, 0x08, 0x20, 0x40


def load(name):
    meta = json.loads((GOLDEN / "flags_fixtures.json").read_text())[name]
    return np.load(GOLDEN / f"{name}.npy", allow_pickle=False), meta


def test_channel_fixture(orc):
    """test/mesh.vti: 2.0 x 0.4 m channel, spacing 0.05, INPUT left, OUTPUT right,
    WALL up/down, free surface, no initial fluid."""
    want, meta = load("flags_channel_41x9")
    got = orc.build_channel_flags(meta["nx"], meta["ny"], meta["dx"], INLET, OUTLET, WALL, WALL,
                                  free_surface=True)
    np.testing.assert_array_equal(got, want)


def test_dambreak_fixture(orc):
    """test/ejemplos/damBreak/mesh.vti: 1 x 1 m box of walls, spacing 0.01, WALL
    obstacle centred (0.45, 0.05) of 0.05 x 0.1, initial fluid block centred
    (0.2, 0.35) of 0.2 x 0.5 (the file predates the deck's current height 0.8)."""
    want, meta = load("flags_dambreak_101x101")
    got = orc.build_channel_flags(meta["nx"], meta["ny"], meta["dx"], WALL, WALL, WALL, WALL,
                                  free_surface=True, obstacles=[(0.45, 0.05, 0.1, 0.05)],
                                  initial=[(0.2, 0.35, 0.5, 0.2)])
    np.testing.assert_array_equal(got, want)


def test_dambreak_with_deck_height_differs(orc):
    """The deck's current initial height 0.8 does not reproduce the stored file."""
    want, meta = load("flags_dambreak_101x101")
    got = orc.build_channel_flags(meta["nx"], meta["ny"], meta["dx"], WALL, WALL, WALL, WALL,
                                  free_surface=True, obstacles=[(0.45, 0.05, 0.1, 0.05)],
                                  initial=[(0.2, 0.35, 0.8, 0.2)])
    assert (got != want).any()


def test_seeding_visit_order(orc):
    """initFreeSurfaceFlags visits cells in index order and only FLUID/INLET/OUTLET
    cells seed; an OUTLET becomes GAS|OUTLET and never turns INTERFACE."""
    nx, ny = 6, 4
    fl = np.full(nx * ny, GAS, np.uint8)
    fl[:nx] = WALL
    fl[-nx:] = WALL
    fl[1 * nx + 0] = INLET
    fl[2 * nx + 5] = OUTLET
    fl[2 * nx + 4] = FLUID
    orc.lib().orc_seed_free_surface(fl.ctypes.data_as(orc.C.POINTER(orc.C.c_uint8)), nx, ny)
    g = fl.reshape(ny, nx)
    assert g[1, 0] == INLET | FLUID
    assert g[2, 5] == OUTLET | GAS
    assert g[1, 1] == INTERFACE and g[2, 1] == INTERFACE   # seeded by the inlet
    assert g[1, 3] == INTERFACE and g[1, 4] == INTERFACE and g[1, 5] == INTERFACE
    assert g[2, 3] == INTERFACE
    assert g[2, 2] == GAS and g[1, 2] == GAS   # not adjacent to a seeding cell


def test_free_surface_fields(orc):
    flags = np.array([FLUID, FLUID | INLET, FLUID | OUTLET, GAS, GAS | OUTLET, INTERFACE, WALL],
                     np.uint8)
    m, p, e = orc.init_free_surface_fields(flags)
    np.testing.assert_array_equal(m, [1, 1, 1, 0, 0, 0.5, 0])
    np.testing.assert_array_equal(p, [1, 1, 1, 0, 0, 0.5, 0])
    np.testing.assert_array_equal(e, 0)
