"""Native deck reader (flb_deck_*): FlowLB's Geometry.txt / OptionsCalc.txt /
CSV outlines -> cell flags, checked against the mesh.vti files the reference's
example decks hold (written by its "Create Domain" action, src/ui/app.cpp:203-239).

The decks in the reference tree were edited after their mesh.vti was written.
Each case below restores the values the mesh file itself records -- its Spacing
(SIZE), its extent (end point x, CDW radius = half the height) -- and, where the
stored flags show it, the boundary types or the initial fluid block.  With
those, five meshes are reproduced bit for bit.  Three more (CSV-outline
obstacles) were written by an older FlowLB whose flag values differ; they agree
cell for cell through a fixed relabelling, which pins the geometry (outline
centroid, crossing test, walls)."""
import json
import math
import re
import shutil
from pathlib import Path

import numpy as np
import pytest

import flowlb_amd as F
from flowlb_amd import core

GOLDEN = Path(__file__).resolve().parent / "golden"
META = json.loads((GOLDEN / "flags_fixtures.json").read_text())["decks"]

# case -> {file: [(pattern, replacement), ...]}
EXACT = {
    "dambreak": {"OptionsCalc.txt": [(r"^INITIAL_CONDITION_FLAG\(1\)\.height = 0\.8",
                                      "INITIAL_CONDITION_FLAG(1).height = 0.5")]},
    "cavity": {"Geometry.txt": [(r"^SIZE = 0\.01", "SIZE = 0.001")],
               # the stored mesh has its inlet on the lid (top row), not on the left
               "OptionsCalc.txt": [(r"^LEFT_BOUNDARY = INPUT", "LEFT_BOUNDARY = WALL"),
                                   (r"^UP_BOUNDARY = WALL", "UP_BOUNDARY = INPUT")]},
    "modulos": {"Geometry.txt": [(r"^POINT\(2\) = \(15, 0, 0\)", "POINT(2) = (20, 0, 0)"),
                                 (r"^CDW\(1\)\.radius = 2\.5", "CDW(1).radius = 3.0")]},
    "poiseuille": {"Geometry.txt": [(r"^POINT\(2\) = \(10, 0, 0\)", "POINT(2) = (20, 0, 0)"),
                                    (r"^SIZE = 0\.0005", "SIZE = 0.005")],
                   # the stored mesh holds inlet and outlet columns
                   "OptionsCalc.txt": [(r"^LEFT_BOUNDARY = PERIODIC", "LEFT_BOUNDARY = INPUT"),
                                       (r"^RIGHT_BOUNDARY = PERIODIC", "RIGHT_BOUNDARY = OUTPUT")]},
    "channel": {"Geometry.txt": [(r"^POINT\(2\) = \(20, 0, 0\)", "POINT(2) = (2, 0, 0)"),
                                 (r"^SIZE = 0\.025", "SIZE = 0.05"),
                                 (r"^CDW\(1\)\.radius = 1\.0", "CDW(1).radius = 0.2")]},
}
# (our flag, stored flag) pairs of the older numbering: fluid 0, wall 6, inlet 8,
# outlet 8 or 9.
RELABELLED = {
    "pantalon": ({"Geometry.txt": [(r"^SIZE = 0\.005", "SIZE = 0.01")]},
                 {(1, 0), (8, 6), (32, 8), (64, 9)}),
    "plenum": ({"Geometry.txt": [(r"^SIZE = 0\.005", "SIZE = 0.02")]},
               {(1, 0), (8, 6), (32, 8), (64, 8)}),
    "coche": ({"Geometry.txt": [(r"^SIZE = 0\.005", "SIZE = 0.01"),
                                (r"^CDW\(1\)\.radius = 2\.5", "CDW(1).radius = 3.0")]},
              {(1, 0), (8, 6), (32, 8), (64, 8)}),
}


def deck_copy(tmp_path, case, subs):
    dst = tmp_path / case
    shutil.copytree(GOLDEN / "decks" / case, dst)
    for name, rules in subs.items():
        text = (dst / name).read_text()
        for pat, rep in rules:
            new = re.sub(pat, rep, text, flags=re.M)
            assert new != text, (case, pat)
            text = new
        (dst / name).write_text(text)
    return dst


def stored(case):
    return np.load(GOLDEN / "decks" / case / "mesh_flags.npz", allow_pickle=False)["flags"]


@pytest.mark.parametrize("case", sorted(EXACT))
def test_deck_reproduces_mesh_file(tmp_path, case):
    with core.Deck(deck_copy(tmp_path, case, EXACT[case])) as d:
        assert (d.nx, d.ny) == (META[case]["nx"], META[case]["ny"])
        assert d.info.size == pytest.approx(META[case]["dx"])
        np.testing.assert_array_equal(d.flags(), stored(case))


@pytest.mark.parametrize("case", sorted(RELABELLED))
def test_outline_decks_match_older_mesh_files(tmp_path, case):
    subs, pairs = RELABELLED[case]
    with core.Deck(deck_copy(tmp_path, case, subs)) as d:
        assert (d.nx, d.ny) == (META[case]["nx"], META[case]["ny"])
        got = d.flags()
    want = stored(case)
    assert set(zip(got.tolist(), want.tolist())) == pairs
    assert d.info.n_obstacles >= 1


def test_dambreak_deck_equals_channel_builder(tmp_path):
    with core.Deck(deck_copy(tmp_path, "dambreak", EXACT["dambreak"])) as d:
        got = d.flags()
    ref = F.build_channel_flags(101, 101, 0.01, F.WALL, F.WALL, F.WALL, F.WALL, free_surface=True,
                                obstacles=[(0.45, 0.05, 0.1, 0.05)], initial=[(0.2, 0.35, 0.5, 0.2)])
    np.testing.assert_array_equal(got, ref)


def test_calculation_mode_reads_every_option(tmp_path):
    with core.Deck(GOLDEN / "decks" / "dambreak", core.DECK_CALCULATION) as d:
        i = d.info
        assert i.problem == F.FREE_SURFACE and i.collision == F.SRT and i.smagorinsky == 1
        assert i.use_force == 1 and tuple(i.acceleration) == (0.0, -9.80665, 0.0)
        assert (i.block_x, i.block_y, i.block_z) == (16, 16, 0)
        assert i.ref_velocity_lb == 0.1 and i.ref_velocity_si == 10.0
        assert i.kinematic_viscosity == 1e-6 and i.density == 1000.0
        assert (i.time_simulation, i.time_save) == (np.float32(15), np.float32(0.1))
        assert (i.bc_left, i.bc_right, i.bc_up, i.bc_down) == (F.WALL,) * 4
        ui = d.units_in()
    # h_initConstantDataCuda2D on these inputs: L = 1/100, T = L * 0.1 / 10,
    # nu = 1e-6 T / L^2, tau = 3 nu + 0.5, F_y = a_y T^2 / L (lattice density 1).
    o = F.units(**{k: getattr(ui, k) for k, _ in core.UnitsIn._fields_})
    L, T = 0.01, 0.01 * 0.1 / 10.0
    assert o.tau == np.float32(3 * np.float32(1e-6 * T / L ** 2) + 0.5)
    assert o.fy == pytest.approx(-9.80665 * T * T / L, rel=1e-6)
    assert o.fx == 0.0


@pytest.mark.parametrize("case,line", [("coche", 17), ("channel", 17)])
def test_calculation_mode_rejects_unknown_options(case, line):
    """GRAVITY / REFERENCE_VELOCITY are not options of the current reader: the
    reference's readOptionsCalculation fails on them, and so does this one
    (the domain mode, which reads five keys, accepts the same decks)."""
    with pytest.raises(core.FlbError, match=f"line {line}"):
        core.Deck(GOLDEN / "decks" / case, core.DECK_CALCULATION)
    core.Deck(GOLDEN / "decks" / case, core.DECK_DOMAIN).close()


def write_deck(tmp, geometry, options, csv=None):
    tmp.mkdir(parents=True, exist_ok=True)
    (tmp / "Geometry.txt").write_text(geometry)
    (tmp / "OptionsCalc.txt").write_text(options)
    for name, rows in (csv or {}).items():
        (tmp / name).write_text("".join(f"{x};{y}\n" for x, y in rows))
    return tmp


GEOM = """# box 0..{L} x 0..{H}
TYPE_DOMAIN = 2D
POINT(1) = (0, 0, 0)
POINT(2) = ({L}, 0, 0)
SIZE = {dx}
INITIAL_POSITION_SHAPES = 1
END_POSITION_SHAPES = 2
CDW(1).type = CIRCLE
CDW(1).radius = {R}
{extra}
"""
OPTS = """TYPE_PROBLEM = MONOFLUID
LEFT_BOUNDARY = INPUT
RIGHT_BOUNDARY = OUTPUT
UP_BOUNDARY = WALL
DOWN_BOUNDARY = WALL
"""


def cell_centres(d):
    x = d.info.x_min + np.arange(d.nx) * d.info.size
    y = d.info.y_min + np.arange(d.ny) * d.info.size
    return np.meshgrid(x, y)


def test_rotated_wall_and_circle_against_geometry(tmp_path):
    """A wall rotated by 30 degrees and a circle obstacle: the WALL cells are
    those inside the rotated box / disc, computed here with half-planes and
    distances (cells within 1e-9 of an edge excluded)."""
    extra = """OBSTACLE(1).type = WALL
OBSTACLE(1).position = (1.3, 0.9, 0)
OBSTACLE(1).height = 0.8
OBSTACLE(1).xWidth = 0.3
OBSTACLE(1).rotation = (0, 0, 30)
OBSTACLE(2).type = CIRCLE
OBSTACLE(2).position = (2.6, 1.0, 0)
OBSTACLE(2).radius = 0.37"""
    deck = write_deck(tmp_path, GEOM.format(L=4, H=2, dx=0.02, R=1.0, extra=extra), OPTS)
    with core.Deck(deck) as d:
        fl = d.flags().reshape(d.ny, d.nx)
        X, Y = cell_centres(d)
    th = math.radians(30)
    # box axes: the rectangle turned anticlockwise by theta about its centre
    u = (X - 1.3) * math.cos(th) + (Y - 0.9) * math.sin(th)
    v = -(X - 1.3) * math.sin(th) + (Y - 0.9) * math.cos(th)
    box = (np.abs(u) <= 0.15) & (np.abs(v) <= 0.4)
    edge = (np.abs(np.abs(u) - 0.15) < 1e-9) | (np.abs(np.abs(v) - 0.4) < 1e-9)
    r2 = (X - 2.6) ** 2 + (Y - 1.0) ** 2
    disc = r2 <= 0.37 ** 2
    edge |= np.abs(r2 - 0.37 ** 2) < 1e-9
    interior = np.zeros_like(box)
    interior[1:-1, 1:-1] = True
    want = np.where(box | disc, F.WALL, F.FLUID)
    m = interior & ~edge
    np.testing.assert_array_equal(fl[m], want[m])
    assert box[m].sum() > 400 and disc[m].sum() > 800


def test_csv_outline_against_winding_number(tmp_path):
    """A concave CSV outline placed by its centroid and turned by 20 degrees:
    inside cells agree with a winding-number test on the same transformed
    polygon (centroid from the shoelace formula)."""
    pts = np.array([(0.0, 0.0), (1.0, 0.0), (1.0, 0.6), (0.5, 0.25), (0.0, 0.6)]) + 3.0
    extra = """OBSTACLE(1).type = CSV_2D
OBSTACLE(1).name = shape
OBSTACLE(1).position = (2.0, 1.0, 0)
OBSTACLE(1).rotation = (0, 0, 20)"""
    deck = write_deck(tmp_path, GEOM.format(L=4, H=2, dx=0.01, R=1.0, extra=extra), OPTS,
                      {"shape.csv": [(f"{x:.3f}", f"{y:.3f}") for x, y in pts]})
    with core.Deck(deck) as d:
        fl = d.flags().reshape(d.ny, d.nx)
        X, Y = cell_centres(d)
    x, y = pts[:, 0], pts[:, 1]
    cross = x * np.roll(y, -1) - np.roll(x, -1) * y
    A = cross.sum() / 2
    cx = ((x + np.roll(x, -1)) * cross).sum() / (6 * A)
    cy = ((y + np.roll(y, -1)) * cross).sum() / (6 * A)
    th = math.radians(20)
    px = (x - cx) * math.cos(th) - (y - cy) * math.sin(th) + 2.0
    py = (x - cx) * math.sin(th) + (y - cy) * math.cos(th) + 1.0
    ang = np.zeros_like(X)
    near = np.zeros(X.shape, bool)
    for i in range(len(px)):
        ax, ay = px[i] - X, py[i] - Y
        bx, by = px[(i + 1) % len(px)] - X, py[(i + 1) % len(px)] - Y
        ang += np.arctan2(ax * by - ay * bx, ax * bx + ay * by)
        ex, ey = bx - ax, by - ay
        t = np.clip(-(ax * ex + ay * ey) / (ex * ex + ey * ey), 0, 1)
        near |= np.hypot(ax + t * ex, ay + t * ey) < 1e-7
    inside = np.abs(ang) > np.pi
    interior = np.zeros_like(inside)
    interior[1:-1, 1:-1] = True
    m = interior & ~near
    np.testing.assert_array_equal(fl[m] == F.WALL, inside[m])
    assert inside[m].sum() > 2000


def test_sloped_bed_follows_the_point_line(tmp_path):
    """Three points: a flat reach then a 1:4 slope.  The circular CDW is the
    band within R of the point line raised vertically by R (the reference's
    vector-product distance), so on the slope it reaches R/cos - R below the
    line; fluid cells on or below the line become the DOWN boundary (WALL
    here), everything outside the band is WALL."""
    geom = """POINT(1) = (0, 0, 0)
POINT(2) = (1, 0, 0)
POINT(3) = (3, 0.5, 0)
SIZE = 0.01
INITIAL_POSITION_SHAPES = 1
END_POSITION_SHAPES = 3
CDW(1).type = CIRCLE
CDW(1).radius = 0.25
"""
    opts = OPTS.replace("INPUT", "WALL").replace("OUTPUT", "WALL")
    with core.Deck(write_deck(tmp_path, geom, opts)) as d:
        fl = d.flags().reshape(d.ny, d.nx)
        X, Y = cell_centres(d)
        assert d.ny == 101  # y_max = 0.5 + 2 * 0.25
    base = np.where(X <= 1.0, 0.0, (X - 1.0) * 0.25)
    # distance to the raised line, perpendicular (the reference's vector product)
    seg = X > 1.0
    dist = np.where(seg, np.abs((Y - base - 0.25)) * math.cos(math.atan(0.25)), np.abs(Y - 0.25))
    inside = dist <= 0.25
    m = np.abs(dist - 0.25) > 1e-9
    m[:, [0, -1]] = False
    m[[0, -1], :] = False
    m &= np.abs(Y - base) > 1e-9
    want = np.where(inside, F.FLUID, F.WALL)
    want[inside & (Y <= base)] = F.WALL   # DOWN boundary
    np.testing.assert_array_equal(fl[m], want[m])
    assert (inside & (Y < base - 1e-9) & m).sum() > 0   # the band dips below the line


@pytest.mark.parametrize("geometry,options,match", [
    ("POINT(1) = (0, 0, 0)\nPOINT(1) = (1, 0, 0)\n", OPTS, "Geometry.txt line 2"),
    ("POINT(1) = (0, 0, 0)\nPOINT(2) = (1, 0, 0)\nCDW(1).radius = 0.5\n", OPTS,
     "Geometry.txt line 3"),
    ("POINT(1) = (0, 0, 0)\nPOINT(2) = (1, 0, 0)\nSIZE = 0.1\nINITIAL_POSITION_SHAPES = 2\n"
     "END_POSITION_SHAPES = 1\nCDW(1).type = CIRCLE\nCDW(1).radius = 0.5\n", OPTS,
     "initial point must come before"),
    (GEOM.format(L=1, H=1, dx=0.1, R=0.5, extra=""),
     OPTS.replace("LEFT_BOUNDARY = INPUT", "LEFT_BOUNDARY = OUTPUT")
         .replace("RIGHT_BOUNDARY = OUTPUT", "RIGHT_BOUNDARY = WALL"), "left needs the same"),
    (GEOM.format(L=1, H=1, dx=0.1, R=0.5, extra=""),
     OPTS.replace("TYPE_PROBLEM = MONOFLUID", "TYPE_PROBLEM = VOF"), "OptionsCalc.txt line 1"),
    (GEOM.format(L=1, H=1, dx=0.1, R=0.5, extra="OBSTACLE(1).type = CSV_2D\n"
                                                  "OBSTACLE(1).name = missing"), OPTS,
     "cannot open"),
    (GEOM.format(L=1, H=1, dx=0.1, R=0.5, extra="TYPE_DOMAIN = 3D"), OPTS, "3D"),
])
def test_deck_errors_are_reported(tmp_path, geometry, options, match):
    with pytest.raises(core.FlbError, match=match):
        core.Deck(write_deck(tmp_path, geometry, options))


def test_missing_directory():
    with pytest.raises(core.FlbError, match="cannot open"):
        core.Deck("/nonexistent/deck")
