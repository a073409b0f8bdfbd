"""The reference's own Fortran driver with the GPU callbacks (oracle/_ref/
tabipb_hip.exe: src/main.f90 with the integration hook -- GPU set-up and
right-hand side after treecode_initialization, MATVECgpu/MSOLVEgpu bound for
DGMRES, the energy loop replaced by the GPU's; readin, DGMRES and
output_potential unchanged) on the golden cases: the unmodified reference's
iteration count, solvation energy and surface_potential.dat."""
import os
import re
import subprocess

import numpy as np
import pytest

import cases
from conftest import ROOT, load_golden
from tabipb import meshio

pytestmark = pytest.mark.gpu
EXE = os.path.join(ROOT, "oracle", "_ref", "tabipb_hip.exe")


@pytest.mark.skipif(not os.path.exists(EXE), reason="drop-in build needs the reference "
                    "sources at build time (make -C oracle dropin)")
@pytest.mark.parametrize("name", ["sph3_l100_p3", "sph4_off_p6", "sph5_born", "sph3_degen",
                                  "sph3_nano", "sph3_eses", "mol1500"])
def test_reference_driver_with_gpu_callbacks(name, tmp_path):
    g = load_golden(name)
    cases.stage(name, str(tmp_path))
    print(f"[gpu] drop-in {name}", flush=True)
    r = subprocess.run(["bash", "-c", f"ulimit -s unlimited; exec {EXE}"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    out = r.stdout
    assert "GPU set-up done" in out
    m = re.search(r"err=\s*(\S+)\s+ierr=\s*(\S+)\s+iter=\s*(\S+)", out)
    assert m, out[-2000:]
    assert int(m.group(3)) == int(g["iter"]) and int(m.group(2)) == int(g["ierr"])
    sol = float(re.search(r"Solvation energy=:\s*(\S+)", out).group(1))
    assert sol == pytest.approx(float(g["soleng"]), rel=1e-9)
    vpos, vnrm, vpot, faces = meshio.read_surface_potential(str(tmp_path / "surface_potential.dat"))
    s = int(g["stride"])
    ref = g["vpot"]
    ok = ~np.isnan(ref[:, 0])
    np.testing.assert_array_equal(np.isnan(vpot[::s, 0]), ~ok)
    np.testing.assert_allclose(vpot[::s][ok], ref[ok], rtol=0,
                               atol=1e-9 * np.abs(ref[ok]).max() + 5.1e-11)
    assert len(faces) == int(g["pot_nface"])
