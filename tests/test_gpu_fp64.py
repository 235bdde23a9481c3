"""FP64 storage (the reference's PRECISION = 64 path, h_launchCudaCalculations2D
<double>) on the GPU against the FP64 oracle, which restates the same mixed
float/double arithmetic.  With both sides in double the comparison is tight:
the two differ only by the shifted storage, fused multiply-adds and summation
order (~1e-16 per operation).  Bars: populations within 1e-13 after a few
steps; density and velocity within 1e-10 relative L2 over 1500 steps of the
512 x 128 culvert for SRT and Smagorinsky.  Where the reference keeps float
temporaries inside the double path -- TRT's pair sums, the free surface's
fill level (__saturatef) and neighbour averages -- a 1e-17 difference of their
double inputs can round to a different float, so those runs agree to float
precision only: TRT within 1e-7 relative L2, free-surface velocity within 1e-7
and mass within 1e-6 (flags identical) after 400 steps.  Conservative-mode
drift below 1e-12 per 10^4 steps."""
import dataclasses

import numpy as np
import pytest

from flowlb_amd import cases, core

pytestmark = pytest.mark.gpu

SP = cases.SINGLE_PHASE
FS_SMALL = cases.scaled(cases.FS_4096, 16)
FAST = dataclasses.replace(FS_SMALL, name="dam_break_256x64_fast", force=(0.0, -2e-4), tau=0.6)


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def oracle(orc, case, flags, rho0=None, u0=None, **kw):
    o = dict(use_force=case.force != (0.0, 0.0), fx=case.force[0], fy=case.force[1],
             free_surface=case.free_surface, threads=16)
    o.update(kw)
    r = None if rho0 is None else rho0.astype(np.float64)
    u = None if u0 is None else u0.astype(np.float64)
    return orc.Oracle(flags, case.nx, case.ny, case.tau, rho0=r, u0=u, **o)


@pytest.mark.parametrize("steps", [1, 2, 7])
def test_slots_after_few_steps(orc, steps):
    case = cases.scaled(SP, 8)
    flags, rho0, u0 = case.build()
    with case.solver("fp64") as s:
        s.step(steps)
        f = s.populations64()
        g = s.read64()
    o = oracle(orc, case, flags, rho0, u0)
    o.step(steps)
    r = o.get()
    np.testing.assert_allclose(f, r["f"], rtol=0, atol=1e-13)
    np.testing.assert_allclose(g["u"], r["u"], rtol=0, atol=1e-13)
    np.testing.assert_allclose(g["rho"], r["rho"], rtol=0, atol=1e-13)


@pytest.mark.parametrize("variant,tol", [(dict(), 1e-10), (dict(collision=1), 1e-7),
                                         (dict(smagorinsky=1), 1e-10)],
                         ids=["srt", "trt", "smagorinsky"])
def test_culvert_1500_steps(orc, variant, tol):
    flags, rho0, u0 = SP.build()
    s = core.Solver(flags, SP.nx, SP.ny, SP.options("fp64", **variant), rho0=rho0, u0=u0)
    o = oracle(orc, SP, flags, rho0, u0, collision=variant.get("collision", 0),
               smagorinsky=bool(variant.get("smagorinsky", 0)))
    with s:
        s.step(1500)
        g = s.read64()
    o.step(1500)
    r = o.get()
    eu, er = rel_l2(g["u"], r["u"]), rel_l2(g["rho"], r["rho"])
    print(f"\nfp64 512x128 x 1500 {variant}: relL2 u {eu:.2e} rho {er:.2e}")
    assert eu < tol and er < tol


def test_free_surface_400_steps(orc):
    flags, _, _ = FAST.build()
    with FAST.solver("fp64", mass_mode=1) as s:
        s.step(400)
        g = s.read64()
    o = oracle(orc, FAST, flags, mass_mode=1)
    o.step(400)
    r = o.get()
    moved = (r["flags"] != flags).sum()
    dm = np.abs(g["mass"] - r["mass"]).max()
    du = np.abs(g["u"] - r["u"]).max()
    print(f"\nfp64 fs 256x64 x 400: cells changed {moved}, flags equal "
          f"{(g['flags'] == r['flags']).mean():.6f}, max|dm| {dm:.2e} max|du| {du:.2e}")
    assert moved > 100
    np.testing.assert_array_equal(g["flags"], r["flags"])
    assert dm < 1e-6 and du < 1e-7


def test_conservation_1e4_steps():
    case = cases.scaled(cases.FS_4096, 8)
    with case.solver("fp64", mass_mode=1) as s:
        m0 = s.mass()
        s.step(10000)
        m1 = s.mass()
    drift = abs(m1 - m0) / m0
    print(f"\nfp64 fs 512x128 x 1e4 conservative: drift {drift:.2e}")
    assert drift <= 1e-12


def test_fp64_free_surface_fields_round_trip():
    """An FP64 free-surface checkpoint restored through the double setters is
    exact: a run that writes back its own mass, fill level and excess shares
    (flb_set_fs_fields_f64) every step stays bit-identical to one that does
    not; populations written back in double (flb_write_populations_f64) move
    by at most the rounding of adding and removing the lattice weight."""
    kw = dict(mass_mode=1, gas_check_corrected=1)
    with FAST.solver("fp64", **kw) as a, FAST.solver("fp64", **kw) as b:
        for _ in range(150):
            a.step(1)
            b.step(1)
            g = b.read64()
            b.set_fs_fields64(g["mass"], g["phi"], g["excess"])
        ga, gb = a.read64(), b.read64()
        for k in ("flags", "mass", "excess", "phi", "u", "rho"):
            np.testing.assert_array_equal(ga[k], gb[k])
        f = b.populations64()
        b.write_populations64(f)
        np.testing.assert_allclose(b.populations64(), f, rtol=0, atol=4e-16)
        assert (ga["flags"] != FAST.build()[0]).sum() > 20


def test_float_readback_of_fp64_handle():
    case = cases.scaled(SP, 8)
    with case.solver("fp64") as s:
        s.step(5)
        g32, g64 = s.read(), s.read64()
        f32, f64 = s.populations(), s.populations64()
    np.testing.assert_array_equal(g32["u"], g64["u"].astype(np.float32))
    np.testing.assert_array_equal(f32, f64.astype(np.float32))


def test_fp64_not_offered_for_batch():
    # (The adjoint takes FP64 handles: test_gpu_adjoint.py.)
    c = cases.scaled(cases.batch_instances(1, seed=1)[0], 4)
    fl, r0, v0 = c.build()
    with pytest.raises(core.FlbError):
        core.Batch([dict(nx=c.nx, ny=c.ny, tau=c.tau, flags=fl, rho0=r0, u0=v0)], storage="fp64")


@pytest.mark.parametrize("storage,tol", [("fp64", 1e-7), ("fp32", 1e-5)])
def test_forced_trt_branch(orc, storage, tol):
    """FLB_TRT_GUO (collision 2, the defined forced TRT branch) on the forced
    256 x 64 culvert against the oracle's restatement."""
    case = cases.scaled(cases.ADJOINT, 4)
    flags, rho0, u0 = case.build()
    s = core.Solver(flags, case.nx, case.ny, case.options(storage, collision=2), rho0=rho0, u0=u0)
    o = oracle(orc, case, flags, rho0, u0, collision=2)
    with s:
        s.step(1000)
        g = s.read64()
    o.step(1000)
    r = o.get()
    eu = rel_l2(g["u"], r["u"])
    print(f"\n{storage} forced TRT 256x64 x 1000: relL2 u {eu:.2e}")
    assert eu < tol
