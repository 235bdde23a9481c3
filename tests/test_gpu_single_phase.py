"""GPU parity of the single-phase stream-collide kernel against the FP64 oracle,
through the C ABI.  Tolerances: FP32 storage density and velocity within 1e-5
relative L2 of the oracle through t = 1000 of the 512x128 culvert (the FP64
reference algorithm itself overflows on this configuration between t = 1750
and t = 2000, so the 10^4-step run has no finite reference); at t = 1500 the
FP32 error has grown to ~1.6e-5 with the flow (|u| up to 0.32) and is bounded
by 1e-4; flags bit-exact; 16-bit storage formats within the bounds stated per
test."""
import numpy as np
import pytest

from flowlb_amd import cases, core

pytestmark = pytest.mark.gpu

SP = cases.SINGLE_PHASE


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def oracle_for(orc, case, flags, rho0, u0, **kw):
    o = dict(use_force=case.force != (0.0, 0.0), fx=case.force[0], fy=case.force[1])
    o.update(kw)
    return orc.Oracle(flags, case.nx, case.ny, case.tau, rho0=rho0.astype(np.float64),
                      u0=u0.astype(np.float64), free_surface=case.free_surface, **o)


def test_populations_after_init_match(orc):
    case = cases.scaled(SP, 8)
    flags, rho0, u0 = case.build()
    with case.solver() as s:
        f = s.populations()
    o = oracle_for(orc, case, flags, rho0, u0)
    np.testing.assert_allclose(f, o.get()["f"], rtol=0, atol=2e-7)


@pytest.mark.parametrize("steps", [1, 2, 3, 7])
def test_slots_match_after_few_steps(orc, steps):
    """Raw esoteric slot array equal to FP32 rounding: pins the slot algebra,
    walls, inlet and outlet on the device."""
    case = cases.scaled(SP, 8)
    flags, rho0, u0 = case.build()
    with case.solver() as s:
        s.step(steps)
        f = s.populations()
        g = s.read()
    o = oracle_for(orc, case, flags, rho0, u0)
    o.step(steps)
    r = o.get()
    np.testing.assert_allclose(f, r["f"], rtol=0, atol=2e-6)
    np.testing.assert_allclose(g["rho"], r["rho"], rtol=0, atol=2e-6)
    np.testing.assert_allclose(g["u"], r["u"], rtol=0, atol=2e-6)
    np.testing.assert_array_equal(g["flags"], flags)


def test_parity_window_fp32(orc):
    """512 x 128 culvert: GPU FP32 (DDF-shifted) against the FP64 oracle over the
    window in which the reference algorithm stays bounded (it overflows between
    t = 1750 and 2000 on this configuration, see
    test_reference_diverges_on_512x128).  Bar: 1e-5 relative L2 in u and rho
    through t = 1000; the t = 1500 value is printed and bounded by 1e-4 (the
    rounding difference grows with the accelerating jet).  The error is also
    compared with the reference's own FP32 arithmetic (FP32 oracle)."""
    flags, rho0, u0 = SP.build()
    s = SP.solver("fp32")
    o = oracle_for(orc, SP, flags, rho0, u0, threads=16)
    o32 = oracle_for(orc, SP, flags, rho0, u0, threads=16, fp32=True)
    t = 0
    for target in (100, 500, 1000, 1500):
        s.step(target - t)
        o.step(target - t)
        o32.step(target - t)
        t = target
        g, r, r32 = s.read(), o.get(), o32.get()
        e_u, e_rho = rel_l2(g["u"], r["u"]), rel_l2(g["rho"], r["rho"])
        e32 = rel_l2(r32["u"], r["u"])
        print(f"\nfp32 512x128 t={t}: GPU relL2 u {e_u:.3e} rho {e_rho:.3e} | "
              f"reference-FP32 relL2 u {e32:.3e}")
        if t <= 1000:
            assert e_u < 1e-5 and e_rho < 1e-5
        else:
            assert e_u < 1e-4 and e_rho < 1e-5
        np.testing.assert_array_equal(g["flags"], flags)
    s.close()


def test_reference_diverges_on_512x128(orc):
    """The reference algorithm (FP64) on the 512 x 128 culvert is unstable: at
    t = 1750 the jet through the culvert reaches |u| = 0.38 with rho in
    [0.85, 1.28]; by t = 2000 the density is of order 1e37 and the run
    overflows before t = 2500."""
    flags, rho0, u0 = SP.build()
    o = oracle_for(orc, SP, flags, rho0, u0, threads=16)
    o.step(1750)
    g = o.get()
    act = (flags & 0x08) == 0
    assert np.isfinite(g["rho"]).all() and g["rho"][act].max() < 1.3
    o.step(250)
    assert np.abs(o.get()["rho"][act]).max() > 1e6
    o.step(500)
    assert not np.isfinite(o.get()["u"]).all()


@pytest.mark.parametrize("storage,tol_u", [("fp16s", 5e-3), ("bf16s", 3e-2)])
def test_sixteen_bit_storage(orc, storage, tol_u):
    """DDF-shifted 16-bit storage, FP32 arithmetic: velocity error vs FP64 after
    1000 steps.  Measured relative L2 of u: FP16S 5.3e-4, BF16S 5.5e-3; the
    bars are about 10x and 5x those."""
    steps = 1000
    flags, rho0, u0 = SP.build()
    with SP.solver(storage) as s:
        s.step(steps)
        g = s.read()
    with SP.solver("fp32") as s32:
        s32.step(steps)
        g32 = s32.read()
    o = oracle_for(orc, SP, flags, rho0, u0, threads=16)
    o.step(steps)
    r = o.get()
    e = rel_l2(g["u"], r["u"])
    e32 = rel_l2(g32["u"], r["u"])
    print(f"\n{storage} 512x128 x {steps}: relL2 u {e:.3e} (fp32 {e32:.3e}) rho {rel_l2(g['rho'], r['rho']):.3e}")
    assert np.isfinite(g["u"]).all()
    assert e < tol_u
    assert e32 < 1e-5


@pytest.mark.parametrize("variant", [
    dict(use_force=1, fx=2e-6, fy=-1e-6),
    dict(smagorinsky=1),
    dict(collision=1),
    dict(collision=1, smagorinsky=1),
])
def test_option_variants(orc, variant):
    """Forcing, Smagorinsky and TRT (lbmKernel2D.cu:372-381, 417-442, 475-510)."""
    case = cases.scaled(SP, 4)
    flags, rho0, u0 = case.build()
    opts = core.default_options(tau=case.tau, **variant)
    with core.Solver(flags, case.nx, case.ny, opts, rho0=rho0, u0=u0) as s:
        s.step(300)
        g = s.read()
        f = s.populations()
    kw = dict(use_force=bool(variant.get("use_force")), fx=variant.get("fx", 0.0),
              fy=variant.get("fy", 0.0), smagorinsky=bool(variant.get("smagorinsky")),
              collision=variant.get("collision", 0))
    o = orc.Oracle(flags, case.nx, case.ny, case.tau, rho0=rho0.astype(np.float64),
                   u0=u0.astype(np.float64), **kw)
    o.step(300)
    r = o.get()
    assert rel_l2(g["u"], r["u"]) < 1e-5
    np.testing.assert_allclose(f, r["f"], rtol=0, atol=1e-5)


def test_stepwise_equals_batched():
    """flb_step(1) x n and flb_step(n) give the same state (moments are written
    on the last step of each call)."""
    case = cases.scaled(SP, 8)
    with case.solver() as a, case.solver() as b:
        for _ in range(25):
            a.step(1)
        b.step(25)
        ga, gb = a.read(), b.read()
        np.testing.assert_array_equal(a.populations(), b.populations())
        np.testing.assert_array_equal(ga["u"], gb["u"])
        np.testing.assert_array_equal(ga["rho"], gb["rho"])
        assert a.time == b.time == 25


@pytest.mark.parametrize("block_x", [64, 128, 512])
def test_block_shapes_agree(block_x):
    case = cases.scaled(SP, 2)
    with case.solver() as a, case.solver(block_x=block_x) as b:
        a.step(40)
        b.step(40)
        np.testing.assert_array_equal(a.populations(), b.populations())


def test_mass_conserved_closed_box():
    nx, ny = 256, 96
    fl = core.build_channel_flags(nx, ny, 1.0, core.WALL, core.WALL, core.WALL, core.WALL)
    rng = np.random.default_rng(0)
    rho0 = (1 + rng.uniform(-1e-3, 1e-3, nx * ny)).astype(np.float32)
    with core.Solver(fl, nx, ny, core.default_options(tau=0.7, use_force=1, fx=1e-5), rho0=rho0) as s:
        m0 = s.mass()
        s.step(2000)
        m1 = s.mass()
    assert abs(m1 - m0) / m0 < 1e-6


def test_readback_helpers():
    case = cases.scaled(SP, 8)
    flags, _, _ = case.build()
    with case.solver() as s:
        h = s.flag_histogram()
        for v, cnt in zip(*np.unique(flags, return_counts=True)):
            assert h[v] == cnt
        ptrs = s.device_fields()
        assert ptrs[0] and ptrs[1] and ptrs[2]


@pytest.mark.parametrize("pitch", [0, 320, 259])
def test_row_pitch_does_not_change_results(pitch):
    """Device planes padded to other row strides (flb_options.row_pitch) give
    bit-identical populations and fields: the layout is storage only.  The
    256 x 64 culvert has inlet and outlet columns, whose linear neighbours
    n +- 1 are read through the pitch."""
    case = cases.scaled(SP, 2)
    with case.solver() as a, case.solver(row_pitch=pitch) as b:
        assert a.layout == (256, 256 * 64)          # nx is a multiple of 64
        assert b.layout == (pitch or 256, (pitch or 256) * 64)
        a.step(40)
        b.step(40)
        ga, gb = a.read(), b.read()
        np.testing.assert_array_equal(a.populations(), b.populations())
        np.testing.assert_array_equal(ga["u"], gb["u"])
        np.testing.assert_array_equal(ga["rho"], gb["rho"])
        np.testing.assert_array_equal(ga["flags"], gb["flags"])
        np.testing.assert_array_equal(a.flag_histogram(), b.flag_histogram())
        assert a.mass() == pytest.approx(b.mass(), rel=1e-14)


def test_row_pitch_rejected_below_nx():
    case = cases.scaled(SP, 8)
    with pytest.raises(core.FlbError):
        case.solver(row_pitch=case.nx - 1)


def test_inlet_on_last_column_reads_next_row(orc):
    """An inlet in the last column reads rho[n + 1], which in the reference's
    row-major numbering is the first cell of the next row, and an outlet in
    the first column reads u[n - 1], the last cell of the previous row
    (lbmKernel2D.cu:336,344).  With padded device rows those are not the
    adjacent elements; the populations must still match the oracle."""
    nx, ny = 64, 16
    flags, rho0, u0 = core.build_culvert(nx, ny, False, (core.OUTLET, core.INLET, core.WALL, core.WALL),
                                         (24, 32), 6, u_inlet=-0.03, rho_noise=1e-4, seed=3)
    opts = core.default_options(tau=0.6, row_pitch=nx + 64)
    with core.Solver(flags, nx, ny, opts, rho0=rho0, u0=u0) as s:
        s.step(9)
        f = s.populations()
        g = s.read()
    o = orc.Oracle(flags, nx, ny, 0.6, rho0=rho0.astype(np.float64), u0=u0.astype(np.float64))
    o.step(9)
    r = o.get()
    np.testing.assert_allclose(f, r["f"], rtol=0, atol=2e-6)
    np.testing.assert_allclose(g["u"], r["u"], rtol=0, atol=2e-6)
