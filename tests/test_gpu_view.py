"""The viewer entry points (flb_export_velocity, flb_isosurface) on the GPU
against the CPU restatement oracle/view_oracle.py, through the C ABI, into
plain device buffers.  Bars: the float2 export equals the velocity readback
bit for bit; the isoline rasters equal the restatement pixel for pixel
(integer output, float32 arithmetic kept in the reference's order on both
sides)."""
import dataclasses

import numpy as np
import pytest

import view_oracle as vo
from flowlb_amd import cases, core

pytestmark = pytest.mark.gpu

SP = cases.scaled(cases.SINGLE_PHASE, 4)          # 128 x 32 culvert
FAST = dataclasses.replace(cases.scaled(cases.FS_4096, 16), name="dam_break_256x64_fast",
                           force=(0.0, -2e-4), tau=0.6)


@pytest.mark.parametrize("storage", ["fp32", "fp16s", "fp64"])
@pytest.mark.parametrize("pad", [0, 72])
def test_export_velocity_matches_readback(storage, pad):
    with SP.solver(storage) as s:
        s.step(120)
        g = s.read()
        pitch = 8 * SP.nx + pad
        with core.DeviceBuffer(pitch * SP.ny, fill=0xAB) as buf:
            s.export_velocity(buf, pitch)
            raw = buf.to_host()
    rows = raw.reshape(SP.ny, pitch)
    img = rows[:, :8 * SP.nx].copy().view(np.float32).reshape(SP.ny, SP.nx, 2)
    want = vo.export_velocity(g["u"][:SP.N].reshape(SP.ny, SP.nx),
                              g["u"][SP.N:].reshape(SP.ny, SP.nx))
    np.testing.assert_array_equal(img, want)
    assert (rows[:, 8 * SP.nx:] == 0xAB).all()     # row padding untouched
    assert np.abs(want).max() > 0.01


VIEWS = [
    dict(width=384, height=96, bounds=(0, 128, 0, 32)),       # whole domain, 3 px per node
    dict(width=200, height=150, bounds=(20, 70, 4, 29)),      # zoomed window
    dict(width=90, height=40, bounds=(100, 160, 10, 50)),     # window past the domain edge
]


@pytest.mark.parametrize("view", VIEWS)
@pytest.mark.parametrize("iso", [0.02, 0.06])
def test_velocity_isolines(view, iso):
    with SP.solver() as s:
        s.step(300)
        g = s.read()
        with core.DeviceBuffer(view["width"] * view["height"]) as tex:
            s.isosurface(tex, view["width"], view["height"], view["bounds"], iso)
            img = tex.to_host().reshape(view["height"], view["width"])
    mag = vo.magnitude_field(g["u"][:SP.N], g["u"][SP.N:]).reshape(SP.ny, SP.nx)
    want = vo.isosurface(mag, view["width"], view["height"], view["bounds"], iso)
    lit = int((want == 255).sum())
    print(f"\nvelocity isolines iso {iso} window {view['bounds']}: {lit} pixels")
    assert lit > 20
    np.testing.assert_array_equal(img, want)


@pytest.mark.parametrize("storage", ["fp32", "fp64"])
def test_fill_level_isolines(storage):
    """Fill-level isoline (iso 0.5) of a collapsing dam break, with a texture
    pitch wider than its rows; pixels not on a line keep what the buffer held
    when clear = 0."""
    W, H, pitch = 512, 128, 544
    bounds = (0, FAST.nx, 0, FAST.ny)
    with FAST.solver(storage, mass_mode=1, gas_check_corrected=1) as s:
        s.step(200)
        g = s.read()
        with core.DeviceBuffer(pitch * H, fill=7) as tex:
            s.isosurface(tex, W, H, bounds, 0.5, field=core.ISO_PHI, clear=False, pitch_bytes=pitch)
            kept = tex.to_host().reshape(H, pitch)
            s.isosurface(tex, W, H, bounds, 0.5, field=core.ISO_PHI, clear=True, pitch_bytes=pitch)
            img = tex.to_host().reshape(H, pitch)
    want = vo.isosurface(g["phi"].reshape(FAST.ny, FAST.nx), W, H, bounds, 0.5)
    assert (want == 255).sum() > 200
    np.testing.assert_array_equal(img[:, :W], want)
    assert (img[:, W:] == 7).all()
    np.testing.assert_array_equal(kept[:, :W], np.where(want == 255, 255, 7))


def test_fill_level_needs_free_surface():
    with SP.solver() as s, core.DeviceBuffer(64 * 64) as tex:
        with pytest.raises(core.FlbError):
            s.isosurface(tex, 64, 64, (0, 10, 0, 10), 0.5, field=core.ISO_PHI)
        with pytest.raises(core.FlbError):
            s.isosurface(tex, 64, 64, (10, 10, 0, 10), 0.5)
