"""Writes the flag fixtures held by the reference's test decks.

Reads the ASCII "Flags" DataArray of two VTK ImageData files that FlowLB's
"Create Domain" action wrote (src/ui/app.cpp:203-239 -> src/io/writer.cpp) and
stores them as .npy arrays (row-major, index x + y*Nx) next to this script:

* test/mesh.vti: 41 x 9 free-surface INPUT/OUTPUT channel, spacing 0.05;
* test/ejemplos/damBreak/mesh.vti: 101 x 101 dam break, spacing 0.01.

Run once in a container that holds the reference tree:
    python tests/golden/make_fixtures.py /root/reference
"""
import json
import re
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent

CASES = {
    "flags_channel_41x9": "test/mesh.vti",
    "flags_dambreak_101x101": "test/ejemplos/damBreak/mesh.vti",
}


def read_vti_flags(path: Path):
    text = path.read_text()
    ext = re.search(r'WholeExtent="(\d+) (\d+) (\d+) (\d+) (\d+) (\d+)"', text)
    spacing = re.search(r'Spacing="([0-9.eE+-]+) ([0-9.eE+-]+) ([0-9.eE+-]+)"', text)
    body = re.search(r'Name="Flags"[^>]*>(.*?)</DataArray>', text, re.S).group(1)
    vals = np.array([int(v) for v in body.split()], dtype=np.uint8)
    nx = int(ext.group(2)) - int(ext.group(1)) + 1
    ny = int(ext.group(4)) - int(ext.group(3)) + 1
    assert vals.size == nx * ny, (vals.size, nx, ny)
    return vals, nx, ny, float(spacing.group(1))


def main(root: str):
    meta = {}
    for name, rel in CASES.items():
        vals, nx, ny, dx = read_vti_flags(Path(root) / rel)
        np.save(HERE / f"{name}.npy", vals, allow_pickle=False)
        hist = {int(k): int(v) for k, v in zip(*np.unique(vals, return_counts=True))}
        meta[name] = {"source": rel, "nx": nx, "ny": ny, "dx": dx, "histogram": hist}
    (HERE / "flags_fixtures.json").write_text(json.dumps(meta, indent=2) + "\n")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
