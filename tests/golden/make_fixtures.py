"""Writes the flag fixtures held by the reference's test decks.

Reads the ASCII "Flags" DataArray of the VTK ImageData files that FlowLB's
"Create Domain" action wrote (src/ui/app.cpp:203-239 -> src/io/writer.cpp):

* flags_channel_41x9.npy / .vti   test/mesh.vti (41 x 9, spacing 0.05);
* flags_dambreak_101x101.npy / .vti   test/ejemplos/damBreak/mesh.vti;
* decks/<case>/mesh_flags.npz   the mesh.vti of every example deck, with the
  deck inputs beside it (Geometry.txt, OptionsCalc.txt, CSV outlines), copied
  unchanged;
* flags_fixtures.json   extent, spacing, origin and flag histogram of each.

Run once in a container that holds the reference tree:
    python tests/golden/make_fixtures.py /root/reference
"""
import json
import re
import shutil
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent

CASES = {
    "flags_channel_41x9": "test/mesh.vti",
    "flags_dambreak_101x101": "test/ejemplos/damBreak/mesh.vti",
}
DECKS = {
    "channel": "test",
    "poiseuille": "test/ejemplos/Poiseuille_plano",
    "cavity": "test/ejemplos/cavity",
    "coche": "test/ejemplos/coche",
    "dambreak": "test/ejemplos/damBreak",
    "modulos": "test/ejemplos/modulos",
    "pantalon": "test/ejemplos/pantalon",
    "plenum": "test/ejemplos/plenum",
}


def read_vti_flags(path: Path):
    text = path.read_text()
    ext = re.search(r'WholeExtent="(\d+) (\d+) (\d+) (\d+) (\d+) (\d+)"', text)
    spacing = re.search(r'Spacing="([0-9.eE+-]+) ([0-9.eE+-]+) ([0-9.eE+-]+)"', text)
    origin = re.search(r'Origin="([0-9.eE+-]+) ([0-9.eE+-]+) ([0-9.eE+-]+)"', text)
    body = re.search(r'Name="Flags"[^>]*>(.*?)</DataArray>', text, re.S).group(1)
    vals = np.array([int(v) for v in body.split()], dtype=np.uint8)
    nx = int(ext.group(2)) - int(ext.group(1)) + 1
    ny = int(ext.group(4)) - int(ext.group(3)) + 1
    assert vals.size == nx * ny, (vals.size, nx, ny)
    return vals, nx, ny, float(spacing.group(1)), (float(origin.group(1)), float(origin.group(2)))


def histogram(vals):
    return {int(k): int(v) for k, v in zip(*np.unique(vals, return_counts=True))}


def main(root: str):
    meta = {}
    for name, rel in CASES.items():
        vals, nx, ny, dx, _ = read_vti_flags(Path(root) / rel)
        np.save(HERE / f"{name}.npy", vals, allow_pickle=False)
        meta[name] = {"source": rel, "nx": nx, "ny": ny, "dx": dx, "histogram": histogram(vals)}
        # the file itself, to check the VTK writer's output byte for byte
        shutil.copyfile(Path(root) / rel, HERE / f"{name}.vti")
    decks = {}
    for name, rel in DECKS.items():
        src = Path(root) / rel
        dst = HERE / "decks" / name
        dst.mkdir(parents=True, exist_ok=True)
        names = {"Geometry.txt", "OptionsCalc.txt"}
        for deck in ("Geometry.txt", "OptionsCalc.txt"):   # CSV outlines the deck names
            for m in re.finditer(r"^\s*[A-Z_]+\(\d+\)\.name\s*=\s*(\w+)",
                                 (src / deck).read_text(), re.M):
                names.add(m.group(1) + ".csv")
        for n in sorted(names):
            shutil.copyfile(src / n, dst / n)
        vals, nx, ny, dx, org = read_vti_flags(src / "mesh.vti")
        np.savez_compressed(dst / "mesh_flags.npz", flags=vals)
        decks[name] = {"source": rel, "nx": nx, "ny": ny, "dx": dx, "origin": list(org),
                       "histogram": histogram(vals)}
    meta["decks"] = decks
    (HERE / "flags_fixtures.json").write_text(json.dumps(meta, indent=2) + "\n")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
