"""Batched instances (flb_batch_*): every instance of a batch advanced by the
packed kernels must equal the same culvert run alone on a handle, bit for
bit (same per-cell arithmetic), with one launch per step (with and without
graph replay), with the persistent cooperative kernel (grid barrier between
steps), and after autotuning; one instance is also checked against the
oracle.  The instances' rows are padded to a 64-cell pitch on the device, so
the readbacks also check the strided copies."""
import numpy as np
import pytest

from flowlb_amd import cases, core

pytestmark = pytest.mark.gpu


def instances(count=7, seed=11):
    out = []
    for c in cases.batch_instances(count, seed=seed):
        c = cases.scaled(c, 4)     # 64..512 x 16..128
        flags, rho0, u0 = c.build()
        out.append((c, dict(nx=c.nx, ny=c.ny, tau=c.tau, flags=flags, rho0=rho0, u0=u0)))
    return out


@pytest.mark.parametrize("graph_chunk,grid,order", [(0, 512, "rows"), (10, 512, "rows"),
                                                    (10, 0, "rows"), (10, 0, "instance"),
                                                    (0, 512, "bands")])
def test_batch_equals_single_handles(graph_chunk, grid, order):
    inst = instances()
    steps = 37
    with core.Batch([d for _, d in inst]) as b:
        b.set_order(order)
        b.configure(block_x=128, grid=grid, graph_chunk=graph_chunk)  # grid 0: a block per tile
        b.step(steps)
        got = [(b.populations(i), b.read(i)) for i in range(len(inst))]
        assert b.time == steps
    for (c, d), (f, fields) in zip(inst, got):
        with core.Solver(d["flags"], c.nx, c.ny, c.options(), rho0=d["rho0"], u0=d["u0"]) as s:
            s.step(steps)
            np.testing.assert_array_equal(f, s.populations())
            g = s.read()
            np.testing.assert_array_equal(fields["u"], g["u"])
            np.testing.assert_array_equal(fields["rho"], g["rho"])


@pytest.mark.parametrize("block_x,per_cu,chunk", [(256, 1, 10), (512, 2, 7), (128, 1, 100)])
def test_persistent_equals_single_handles(block_x, per_cu, chunk):
    inst = instances(seed=13)
    steps = 41
    with core.Batch([d for _, d in inst]) as b:
        b.configure_persistent(block_x=block_x, blocks_per_cu=per_cu, chunk=chunk)
        b.step(steps - 4)
        b.step(4)          # a second call: parity carried across cooperative launches
        got = [(b.populations(i), b.read(i)) for i in range(len(inst))]
        assert b.time == steps
    for (c, d), (f, fields) in zip(inst, got):
        with core.Solver(d["flags"], c.nx, c.ny, c.options(), rho0=d["rho0"], u0=d["u0"]) as s:
            s.step(steps)
            np.testing.assert_array_equal(f, s.populations())
            g = s.read()
            np.testing.assert_array_equal(fields["u"], g["u"])
            np.testing.assert_array_equal(fields["rho"], g["rho"])


def test_persistent_rejects_non_resident_grid():
    inst = instances(2, seed=1)
    with core.Batch([d for _, d in inst]) as b:
        with pytest.raises(core.FlbError):
            b.configure_persistent(block_x=1024, blocks_per_cu=64, chunk=10)


def test_autotune_then_same_result():
    inst = instances(5, seed=3)
    with core.Batch([d for _, d in inst]) as a, core.Batch([d for _, d in inst]) as b:
        bx, grid, ms = a.autotune(steps=4)
        assert bx in (64, 128, 256, 512) and ms > 0
        t = a.time
        b.step(t)
        a.step(30)
        b.step(30)
        for i in range(len(inst)):
            np.testing.assert_array_equal(a.populations(i), b.populations(i))


def test_batch_instance_matches_oracle(orc):
    inst = instances(3, seed=5)
    c, d = inst[1]
    with core.Batch([x for _, x in inst]) as b:
        b.step(200)
        g = b.read(1)
    o = orc.Oracle(d["flags"], c.nx, c.ny, c.tau, rho0=d["rho0"].astype(np.float64),
                   u0=d["u0"].astype(np.float64))
    o.step(200)
    r = o.get()
    assert np.linalg.norm(g["u"] - r["u"]) / np.linalg.norm(r["u"]) < 1e-5


def test_batch_rejects_bad_instances():
    fl = np.full(16, core.FLUID, np.uint8)
    with pytest.raises(core.FlbError):
        core.Batch([dict(nx=4, ny=4, tau=0.5, flags=fl)])
