"""flb_bandwidth_probe, the attainable-bandwidth reference of the roofline: every
kind runs through the C ABI, reports the bytes it moves (18 values per cell:
72 in FP32, 36 in 16-bit) and a rate in the range an MI355X gives (the
esoteric FP32 pattern measured 5.4-5.5 TB/s at 8192 x 2048, the 16-bit one
6.3 TB/s; the bounds here leave room for a smaller grid and another box).
Invalid arguments are rejected without a launch."""
import pytest

from flowlb_amd import core

pytestmark = pytest.mark.gpu

NX, NY = 8192, 2048   # 604 MB of FP32 planes, 302 MB of 16-bit: beyond the 256 MB cache


@pytest.mark.parametrize("kind,per_cell", [("esoteric_inplace", 72), ("copy9", 72),
                                           ("inplace9", 72), ("esoteric_inplace16", 36)])
def test_probe_kinds(kind, per_cell):
    ms, nbytes = core.bandwidth_probe(NX, NY, kind, launches=10)
    rate = nbytes / (ms * 1e-3) / 1e12
    print(f"\n{kind}: {ms * 1e3:.1f} us per launch, {rate:.2f} TB/s")
    assert nbytes == per_cell * NX * NY
    assert 2.0 < rate < 9.0


def test_probe_rejects_invalid_arguments():
    for nx, ny, kind, launches in [(1, 64, 0, 1), (64, 64, 9, 1), (64, 64, 0, 0)]:
        with pytest.raises(core.FlbError):
            ms = core.C.c_float()
            core.check(core.lib().flb_bandwidth_probe(0, nx, ny, kind, launches, core.C.byref(ms),
                                                      None), None, "flb_bandwidth_probe")
