"""Index arithmetic of the kernels, checked on the host: the scalar
magic-number division of block_tile / block_tile_window (flb_device.h) gives
exactly the tiles of the plain division form for every block index of 270 grid
shapes (36 M cases), div_by equals n / d, and plane_stride (flb_internal.h)
keeps population planes 256-byte aligned, at least a plane long and 4 KiB past
a multiple of 8 KiB apart.  Host build with hipcc; no GPU."""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not Path(HIPCC).exists(), reason="hipcc not installed")
def test_block_mapping_matches_division(tmp_path):
    exe = tmp_path / "block_mapping_check"
    subprocess.run([HIPCC, "-O2", "-std=c++17", f"-I{ROOT / 'include'}",
                    f"-I{ROOT / 'flowlb_amd' / 'csrc'}",
                    str(ROOT / "tests" / "host" / "block_mapping_check.cpp"), "-o", str(exe)],
                   check=True, capture_output=True, text=True, timeout=300)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "mismatches 0" in r.stdout
