"""CPU pins of the oracle and host code in the regimes the reference fixtures
cover beyond the icospheres (fixtures from the flang-built reference,
tests/golden/gen_golden.py):

* output_potential's vertex values (src/main.f90:226-326): the oracle from the
  reference's own solution prints exactly as the reference's
  surface_potential.dat, and the host writer reproduces the file byte for byte;
* surface_area culling (src/readin.f90:264-333) on a mesh with duplicate,
  zero-area and sliver faces: the culled face count and the tree permutation;
* the NanoShaper and ESES readers (src/readin.f90:90-197): the geometry the
  reference built from those files, bit for bit;
* MSOLVEprec (src/main.f90:532-542), exactly;
* GMRES through a restart and into the NRMAX cap (IERR = 1,
  src/dgmres_dep.f:785-850): iteration count and status of the oracle's
  restatement against the reference.
"""
import os
import tempfile

import numpy as np
import pytest

import cases
from conftest import load_golden
from tabipb import driver, meshio

# solved cases whose fixtures hold every entry (up to 8,000 faces)
SOLVED_SMALL = ["sph3_l100_p3", "sph4_off_p6", "sph3_k0_p2", "sph3_direct", "sph3_cap",
                "sph3_degen", "sph3_nano", "sph3_eses", "sph3_diag"]


def staged_problem(name):
    with tempfile.TemporaryDirectory() as d:
        cases.stage(name, d)
        return driver.load_run_dir(d)


def orig_order(g, v):
    order = g["orderarr"].astype(np.int64) - 1
    n = order.size
    out = np.empty(2 * n)
    out[order] = v[:n]
    out[n + order] = v[n:]
    return out


@pytest.mark.parametrize("name", SOLVED_SMALL)
def test_oracle_vertex_potential_prints_as_reference_file(oracle_lib, name):
    O = oracle_lib
    g = load_golden(name)
    P = staged_problem(name)
    vp = O.vertex_potential(P.vpos, P.faces, P.xyz, orig_order(g, g["xvct"]))
    assert [f"{a:20.10f}" for a in vp.ravel()] == [f"{a:20.10f}" for a in g["vpot"].ravel()]


@pytest.mark.parametrize("name", SOLVED_SMALL)
def test_surface_potential_writer_reproduces_reference_file(name):
    g = load_golden(name)
    text = str(g["pot_text"])
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "ref.dat")
        open(src, "w").write(text)
        vpos, vnrm, vpot, faces = meshio.read_surface_potential(src)
        out = os.path.join(d, "mine.dat")
        meshio.write_surface_potential(out, vpos, vnrm, vpot, faces)
        assert open(out).read() == text


def test_culling_matches_reference(oracle_lib):
    """Four degenerate faces culled (1,284 -> 1,280), by the driver and by the
    oracle, and the tree over the kept faces is the reference's."""
    O = oracle_lib
    g = load_golden("sph3_degen")
    v, nv, f, _, _ = cases.build("sph3_degen")
    assert len(f) == 1284
    kept, _ = driver.cull_faces(v, f)
    kept_o, _ = O.cull_faces(v, f)
    assert len(kept) == int(g["n"]) == 1280
    np.testing.assert_array_equal(kept, kept_o)
    xyz, _, _ = driver.face_geometry(v, nv, kept)
    _, order, _ = O.build_tree(xyz, 100)
    np.testing.assert_array_equal(order + 1, g["orderarr"])
    # the vertex only the culled sliver used is NaN in the reference's file
    assert list(g["vpot_nan"]) == [len(v) - 1]


@pytest.mark.parametrize("name", ["sph3_nano", "sph3_eses", "sph3_degen", "sph3_diag"])
def test_readers_give_reference_geometry(name):
    g = load_golden(name)
    P = staged_problem(name)
    o = g["orderarr"].astype(np.int64) - 1
    np.testing.assert_array_equal(P.xyz[o], g["txyz"])
    np.testing.assert_array_equal(P.nrm[o], g["tnrm"])
    np.testing.assert_array_equal(P.area[o], g["tarea"])


def test_list_directed_reader_semantics():
    """Records continue across lines, surplus items on a record are dropped, a
    slash ends the read, commas separate like blanks."""
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "f.txt")
        open(p, "w").write("1 2\n3 4 5\n6,7 / 8\n9\n")
        r = meshio.ListDirected(p)
        assert r.read(3) == ["1", "2", "3"]      # rest of "3 4 5" dropped
        assert r.read(3) == ["6", "7", None]     # slash ends the read
        assert r.read(1) == ["9"]


def test_diag_prec_matches_reference(oracle_lib):
    g = load_golden("sph3_diag")
    z = oracle_lib.diag_prec(float(g["eps"]), g["x"])
    assert np.array_equal(z, g["z"])


def test_oracle_gmres_restart_and_cap_match_reference(oracle_lib):
    """The oracle's DGMRES restatement through one restart into the NRMAX cap:
    20 steps, IERR = 1, the reference's solution and energy."""
    O = oracle_lib
    g = load_golden("sph3_cap")
    P = staged_problem("sph3_cap")
    txyz, order, nodes = O.build_tree(P.xyz, P.usr.maxparnode)
    prm = O.params(P.eps, P.kappa, P.usr.mac, P.usr.order, P.usr.maxparnode)
    tn, ta = P.nrm[order], P.area[order]
    x, it, err, ierr, hist = O.gmres(prm, txyz, tn, ta, nodes, g["bvct"], tol=float(g["tol"]),
                                     nrmax=int(g["nrmax"]))
    assert (it, ierr) == (int(g["iter"]), int(g["ierr"])) == (20, 1)
    assert np.linalg.norm(x - g["xvct"]) <= 1e-12 * np.linalg.norm(g["xvct"])
    sol, _ = O.energy(txyz, tn, ta, x, P.qpos, P.q, P.eps, P.kappa, P.usr.epsp)
    assert sol == pytest.approx(float(g["soleng"]), rel=1e-12)


@pytest.mark.slow
def test_oracle_gmres_restart_on_ses_molecule_matches_reference(oracle_lib):
    """The SES molecule at tol 1e-8: the reference restarts once (14 steps,
    DRLCAL residual); the oracle's restatement takes the same steps and ends
    with the same error estimate and solution (about two minutes on one core)."""
    O = oracle_lib
    g = load_golden("mol150_restart")
    P = staged_problem("mol150_restart")
    txyz, order, nodes = O.build_tree(P.xyz, P.usr.maxparnode)
    np.testing.assert_array_equal(order + 1, g["orderarr"])
    prm = O.params(P.eps, P.kappa, P.usr.mac, P.usr.order, P.usr.maxparnode)
    n = order.size
    b = O.rhs(P.xyz, P.nrm, P.qpos, P.q, P.usr.epsp)
    bt = np.concatenate([b[order], b[n + order]])
    s = int(g["stride"])
    np.testing.assert_array_equal(bt[::s], g["bvct"])
    x, it, err, ierr, _ = O.gmres(prm, txyz, P.nrm[order], P.area[order], nodes, bt,
                                  tol=float(g["tol"]), nrmax=int(g["nrmax"]))
    assert (it, ierr) == (int(g["iter"]), int(g["ierr"])) == (14, 0)
    assert err == pytest.approx(float(g["err"]), rel=1e-12)
    assert np.linalg.norm(x[::s] - g["xvct"]) <= 1e-12 * np.linalg.norm(g["xvct"])


def test_oracle_reports_singular_leaf_block(oracle_lib):
    """The oracle's partial-pivot LU flags a zero pivot and names the first
    singular leaf (the GPU library turns the same condition into a set-up
    error naming leaf 1)."""
    O = oracle_lib
    P = staged_problem("sph3_l100_p3")
    txyz, order, nodes = O.build_tree(P.xyz, 100)
    ib, ie = O.leaves(nodes)
    prm = O.params(-1.0, 0.1, 0.8, 3, 100)
    _, bad = O.precond(prm, txyz, P.nrm[order], np.zeros(len(order)), ib, ie,
                       np.ones(2 * len(order)))
    assert bad == 1
