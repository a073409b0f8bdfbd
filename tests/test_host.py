"""Host-side logic (no GPU): the driver mirror's discretisation against the
oracle, surface I/O, and the C-ABI library's exported symbols."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import cases
from conftest import ROOT
from tabipb import driver, hip, meshio


def test_driver_geometry_bit_exact_with_oracle(oracle_lib):
    v, nv, f, atoms, usr = cases.build("sph4_off_p6")
    xyz, nrm, area = driver.face_geometry(v, nv, f)
    ox, on, oa = oracle_lib.face_geometry(v, nv, f)
    np.testing.assert_array_equal(xyz, ox)
    np.testing.assert_array_equal(nrm, on)
    np.testing.assert_array_equal(area, oa)


def test_culling_matches_oracle_on_degenerate_mesh(oracle_lib):
    v, nv, f, _, _ = cases.build("sph3_l100_p3")
    rng = np.random.default_rng(1)
    # sliver faces (area < 1e-5) and near-duplicate faces (centroids closer than
    # sqrt(1e-5) to one of the 10 preceding faces)
    extra_v = v[:20] + 1e-7 * rng.standard_normal((20, 3))
    vv = np.vstack([v, extra_v])
    nvv = np.vstack([nv, nv[:20]])
    ff = [list(t) for t in f[:200]]
    for k in range(20):
        ff.insert(7 * k + 3, [k, len(v) + k, (k + 1) % 20])          # sliver
        ff.insert(9 * k + 5, list(ff[9 * k + 4][::-1]))              # duplicate
    ff = np.array(ff, dtype=np.int64)
    got, sa = driver.cull_faces(vv, ff)
    ref, sa_ref = oracle_lib.cull_faces(vv, ff)
    assert len(got) < len(ff)
    np.testing.assert_array_equal(got, ref)
    assert sa == sa_ref


def test_msms_and_pqr_round_trip(tmp_path):
    v, nv, f, atoms, usr = cases.build("sph3_k0_p2")
    pref = str(tmp_path / "s")
    meshio.write_msms(pref, v, nv, f)
    v2, n2, f2 = meshio.read_msms(pref)
    np.testing.assert_array_equal(v, v2)
    np.testing.assert_array_equal(nv, n2)
    np.testing.assert_array_equal(f, f2)
    meshio.write_pqr(str(tmp_path / "m.pqr"), atoms[:, :3], atoms[:, 3], atoms[:, 4])
    qp, q, r = meshio.read_pqr(str(tmp_path / "m.pqr"))
    np.testing.assert_array_equal(q, atoms[:, 3].astype(np.float32).astype(np.float64))
    usr.write(str(tmp_path / "usrdata.in"))
    u2 = meshio.UsrData.read(str(tmp_path / "usrdata.in"))
    assert u2 == usr


def test_kappa_matches_oracle(oracle_lib):
    for i_s, ew in [(0.15, 80.0), (0.0, 80.0), (0.5, 78.5)]:
        assert driver.kappa_of(i_s, ew) == oracle_lib.kappa_of(i_s, ew)


def _header_functions():
    txt = open(os.path.join(ROOT, "include", "tabi_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(tabi_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    """Loads libtabi_hip.so and resolves each declared entry point (no compute
    call is made, so this runs without a GPU)."""
    declared = _header_functions()
    assert sorted(hip.EXPORTS) == declared
    lib = hip.load()
    for name in declared:
        assert isinstance(getattr(lib, name), ctypes._CFuncPtr), name
    out = subprocess.run(["nm", "-D", "--defined-only", hip.LIB_PATH], capture_output=True,
                         text=True).stdout
    exported = set(re.findall(r" T (tabi_\w+)", out))
    assert set(declared) <= exported


def test_library_missing_raises(monkeypatch):
    monkeypatch.setattr(hip, "_LIB", None)
    monkeypatch.setattr(hip, "LIB_PATH", "/nonexistent/libtabi_hip.so")
    with pytest.raises(hip.TabiError):
        hip.load()


def test_header_compiles_as_c_and_cxx(tmp_path):
    """include/tabi_hip.h is consumed by C, C++ and Fortran drivers: it must
    stand alone in both C and C++."""
    import shutil
    import subprocess
    hdr = os.path.join(ROOT, "include", "tabi_hip.h")
    for cc, lang in (("gcc", "c"), ("g++", "c++")):
        if shutil.which(cc) is None:
            pytest.skip(f"{cc} not available")
        src = tmp_path / f"t.{'c' if lang == 'c' else 'cpp'}"
        src.write_text(f'#include "{hdr}"\nint main(void) {{ return 0; }}\n')
        subprocess.run([cc, "-std=c11" if lang == "c" else "-std=c++17", "-Wall", "-Werror",
                        "-fsyntax-only", str(src)], check=True)
