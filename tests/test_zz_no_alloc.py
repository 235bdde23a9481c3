"""flb_flag_histogram_noalloc (no device allocation; file-scope scratch per
GPU) against flb_flag_histogram, which allocates its counters per call: equal
counts on consecutive calls, the second on a smaller grid (the shared scratch
must start from zero each call), and again on the first handle."""
import dataclasses

import numpy as np
import pytest

from flowlb_amd import cases, core

pytestmark = pytest.mark.gpu

FAST = dataclasses.replace(cases.scaled(cases.FS_4096, 16), name="dam_break_256x64_fast",
                           force=(0.0, -2e-4), tau=0.6)
SMALL = cases.scaled(cases.SINGLE_PHASE, 8)   # 64 x 16


def test_noalloc_histogram_matches_allocating_one():
    with FAST.solver(mass_mode=1, gas_check_corrected=1) as a, SMALL.solver() as b:
        a.step(60)
        b.step(5)
        ha, hb = a.flag_histogram(), b.flag_histogram()
        print(f"\nnoalloc histogram: {FAST.nx}x{FAST.ny} then {SMALL.nx}x{SMALL.ny}", flush=True)
        na = a.flag_histogram_noalloc()
        nb = b.flag_histogram_noalloc()       # smaller input, same scratch
        na2 = a.flag_histogram_noalloc()
        np.testing.assert_array_equal(na, ha)
        np.testing.assert_array_equal(nb, hb)
        np.testing.assert_array_equal(na2, ha)
        assert int(na.sum()) == FAST.nx * FAST.ny and int(nb.sum()) == SMALL.nx * SMALL.ny
        assert (na > 0).sum() >= 3            # wall, gas, fluid, interface present
