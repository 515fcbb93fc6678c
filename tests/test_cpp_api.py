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


def test_unit_api():
    r = _run("lpm_unit_tests")
    assert r.returncode == 0 and "0 failure(s)" in r.stdout


@pytest.mark.gpu
def test_biot_savart_rh_wave():
    r = _run("lpm_biot_savart_rh")
    assert r.returncode == 0 and "0 failure(s)" in r.stdout


@pytest.mark.gpu
def test_poisson_and_octree(tmp_path):
    vtk = tmp_path / "octree.vtk"
    r = _run("lpm_poisson", str(vtk))
    assert r.returncode == 0 and "0 failure(s)" in r.stdout
    text = vtk.read_text().splitlines()
    assert text[0].startswith("# vtk DataFile") and text[4].startswith("POINTS")
