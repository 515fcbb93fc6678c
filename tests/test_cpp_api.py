"""C++ host API (namespace Lpm, the reference's class interface) exercised by
compiled drivers under tests/cpp, which assert and exit with their failure
count.  Built by __graft_entry__.build()."""
import subprocess
from pathlib import Path

import pytest

BIN = Path(__file__).resolve().parent / "cpp" / "bin"


def _run(name, *args, timeout=300):
    exe = BIN / name
    assert exe.exists(), f"{exe} not built"
    r = subprocess.run([str(exe), *args], capture_output=True, text=True, timeout=timeout)
    print(r.stdout[-4000:], r.stderr[-4000:])
    return r


def _vtk_blocks(path):
    """Legacy-VTK polydata as {block header: list of data lines}."""
    lines = path.read_text().splitlines()
    assert lines[0] == "# vtk DataFile Version 2.0" and lines[2] == "ASCII" and lines[3] == "DATASET POLYDATA"
    blocks, key = {}, None
    for ln in lines[4:]:
        head = ln.split()[0] if ln else ""
        if head in ("POINTS", "POLYGONS", "POINT_DATA", "CELL_DATA", "SCALARS"):
            key = ln
            blocks[key] = []
        elif head != "LOOKUP_TABLE":
            blocks[key].append(ln)
    return blocks


def test_unit_api(tmp_path):
    import numpy as np
    import lpmcxx_amd as L
    vtk, csv = tmp_path / "mesh.vtk", tmp_path / "mesh.csv"
    r = _run("lpm_unit_tests", str(vtk), str(csv))
    assert r.returncode == 0 and "0 failure(s)" in r.stdout
    # the VTK file: 98 vertices, 96 quads, a vertex field and two face fields
    b = _vtk_blocks(vtk)
    m = L.Mesh(L.LPM_SEED_CUBED, 2)
    pts = np.array([list(map(float, ln.split())) for ln in b["POINTS 98 double"]])
    assert np.abs(pts - np.stack([m.vx, m.vy, m.vz], 1)).max() < 1e-15
    polys = [list(map(int, ln.split())) for ln in b["POLYGONS 96 480"]]
    assert all(p[0] == 4 and len(p) == 5 and max(p[1:]) < 98 for p in polys)
    assert "POINT_DATA 98" in b and "CELL_DATA 96" in b
    psi = np.array(list(map(float, b["SCALARS streamFn_m^2/s double 1"])))
    close = lambda a, b: np.abs(a - b).max() <= 1e-13 * np.abs(b).max()  # C++ vs numpy analytic functions
    assert close(psi, L.sphere_harmonic54(m.vx, m.vy, m.vz))
    vor = np.array(list(map(float, b["SCALARS vorticity_1/time double 1"])))
    assert close(vor, L.sphere_harmonic54(m.cx, m.cy, m.cz))
    vel = np.array([list(map(float, ln.split())) for ln in b["SCALARS velocity_m/s double 3"]])
    assert close(vel, np.stack(L.rh54_velocity(m.cx, m.cy, m.cz), 1))
    # the CSV file: header and one row per vertex
    rows = csv.read_text().splitlines()
    assert rows[0] == "x,y,z,streamFn" and len(rows) == 99
    tab = np.array([list(map(float, ln.split(","))) for ln in rows[1:]])
    assert np.abs(tab[:, 3] - psi).max() == 0.0 and np.abs(tab[:, :3] - pts).max() == 0.0


@pytest.mark.gpu
def test_series_and_point_cloud_tree():
    r = _run("lpm_series_tree")
    assert r.returncode == 0 and "0 failure(s)" in r.stdout


@pytest.mark.gpu
def test_biot_savart_rh_wave():
    r = _run("lpm_biot_savart_rh")
    assert r.returncode == 0 and "0 failure(s)" in r.stdout


@pytest.mark.gpu
def test_bve_gaussian_vortex(tmp_path):
    """Lagrangian BVE driver: invariants, beta drift, treecode vs direct, and
    its VTK / CSV output."""
    import numpy as np
    vtk, fcsv, vcsv = tmp_path / "bve.vtk", tmp_path / "faces.csv", tmp_path / "verts.csv"
    r = _run("lpm_bve", str(vtk), str(fcsv), str(vcsv))
    assert r.returncode == 0 and "0 failure(s)" in r.stdout
    b = _vtk_blocks(vtk)
    assert "POINTS 6146 double" in b and "POLYGONS 6144 30720" in b
    assert "SCALARS lagrangian_coords double 3" in b and "SCALARS relVort_1/s double 1" in b
    pts = np.array([list(map(float, ln.split())) for ln in b["POINTS 6146 double"]])
    assert np.abs(np.linalg.norm(pts, axis=1) - 1).max() < 1e-14
    lag = np.array([list(map(float, ln.split())) for ln in b["SCALARS lagrangian_coords double 3"]])
    assert np.abs(pts - lag).max() > 1e-3                 # the vertices moved
    rows = fcsv.read_text().splitlines()
    assert rows[0] == "x,y,z,relVort,absVort,area" and len(rows) == 6145
    tab = np.array([list(map(float, ln.split(","))) for ln in rows[1:]])
    assert abs(tab[:, 5].sum() - 4 * np.pi) < 1e-12      # areas are conserved
    assert np.abs(tab[:, 4] - tab[:, 3] - 4 * np.pi * tab[:, 2]).max() < 1e-12  # zeta_abs = zeta_rel + 2 Omega z
    assert len(vcsv.read_text().splitlines()) == 6147


@pytest.mark.gpu
def test_poisson_and_octree(tmp_path):
    vtk = tmp_path / "octree.vtk"
    r = _run("lpm_poisson", str(vtk))
    assert r.returncode == 0 and "0 failure(s)" in r.stdout
    text = vtk.read_text().splitlines()
    assert text[0].startswith("# vtk DataFile") and text[4].startswith("POINTS")
