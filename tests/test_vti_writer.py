"""VTK ImageData output (VTIWriter, src/io/writer.cpp): the ASCII writer
reproduces the reference's files byte for byte, the binary variant holds the
same arrays, and the series file follows Writer::createTimeSeriesFile."""
import json
import re
from pathlib import Path

import numpy as np
import pytest

from flowlb_amd import core

GOLDEN = Path(__file__).resolve().parent / "golden"


@pytest.mark.parametrize("name", ["flags_channel_41x9", "flags_dambreak_101x101"])
def test_ascii_matches_reference_file(tmp_path, name):
    meta = json.loads((GOLDEN / "flags_fixtures.json").read_text())[name]
    flags = np.load(GOLDEN / f"{name}.npy", allow_pickle=False)
    out = tmp_path / "mesh.vti"
    core.write_vti(out, meta["nx"], meta["ny"], [("Flags", flags, 1.0, False)],
                   spacing=meta["dx"])
    assert out.read_bytes() == (GOLDEN / f"{name}.vti").read_bytes()


def read_appended(path):
    raw = path.read_bytes()
    head, data = raw.split(b'<AppendedData encoding="raw">\n_', 1)
    arrays = {}
    for m in re.finditer(rb'type="(\w+)" Name="(\w+)" NumberOfComponents="(\d)" format="appended" '
                         rb'offset="(\d+)"', head):
        t, name, comps, off = m.group(1), m.group(2).decode(), int(m.group(3)), int(m.group(4))
        n = int(np.frombuffer(data[off:off + 8], np.uint64)[0])
        dt = np.uint8 if t == b"UInt8" else np.float32
        arrays[name] = np.frombuffer(data[off + 8:off + 8 + n], dt).reshape(-1, comps)
    return arrays


def test_binary_holds_same_arrays(tmp_path):
    nx, ny = 13, 7
    N = nx * ny
    rng = np.random.default_rng(0)
    flags = rng.integers(0, 256, N).astype(np.uint8)
    u = rng.standard_normal(2 * N).astype(np.float32)
    rho = rng.standard_normal(N).astype(np.float32)
    fields = [("Flags", flags, 1.0, False), ("U", u, 2.5, True), ("rho", rho, 1.0, False)]
    core.write_vti(tmp_path / "b.vti", nx, ny, fields, binary=True)
    a = read_appended(tmp_path / "b.vti")
    np.testing.assert_array_equal(a["Flags"][:, 0], flags)
    np.testing.assert_allclose(a["U"][:, 0], u[:N] * 2.5, rtol=1e-7)
    np.testing.assert_allclose(a["U"][:, 1], -u[N:] * 2.5, rtol=1e-7)
    np.testing.assert_array_equal(a["rho"][:, 0], rho)
    # ascii holds the same values to 6 significant digits
    core.write_vti(tmp_path / "a.vti", nx, ny, fields)
    txt = (tmp_path / "a.vti").read_text()
    body = re.search(r'Name="U"[^>]*>(.*?)</DataArray>', txt, re.S).group(1).split()
    np.testing.assert_allclose(np.array(body, float).reshape(-1, 2)[:, 1], -u[N:] * 2.5,
                               rtol=1e-5, atol=1e-6)


def test_series_file(tmp_path):
    core.write_vti_series(tmp_path / "loaderFiles.vti.series", 3, 0.1)
    txt = (tmp_path / "loaderFiles.vti.series").read_text()
    assert txt == ('{\n"file-series-version" : "1.0",\n"files" : [\n'
                   '{"name" : "result_0.vti", "time" : 0.000000},\n'
                   '{"name" : "result_1.vti", "time" : 0.100000},\n'
                   '{"name" : "result_2.vti", "time" : 0.200000},\n]\n}')
