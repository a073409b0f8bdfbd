"""Deterministic inputs of the golden cases (shared by gen_golden.py and tests).

Every surface is built with element-wise IEEE arithmetic only (icospheres) or by
the package's own seeded mesher (synthetic molecules, the same code and library
versions on every machine of this image), so the same bits come out everywhere;
gen_golden.py records a SHA-256 of the inputs and the tests check it before
comparing outputs.

A case is a surface (``surf``: icosphere subdivision, radius, centre; or
``mol``: atom count, seed, grid spacing of the SES mesher), atoms, usrdata.in
overrides, optionally a surface file format other than MSMS (``fmt``), a
degenerate-face edit that the reference's culling must remove (``degen``), and
the options the reference harness runs with (``harness``: GMRES tol and NRMAX,
``nosolve``/``nomsolve`` for the large surfaces).
"""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

_ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(_ROOT, "tabi-pb_amd"))

from tabipb import meshio  # noqa: E402


def _sphere(subdiv, radius, center):
    v, n, f = meshio.icosphere(subdiv, radius)
    return v + np.asarray(center, dtype=np.float64), n, f


_MOL = (150, 11, 0.5)

# name -> (surface spec, atoms [(x,y,z,q,r)], usrdata overrides, ...)
CASES = {
    # small sphere, three tree levels, p=3
    "sph3_l100_p3": dict(surf=(3, 2.0, (0.0, 0.0, 0.0)), atoms=[(0, 0, 0, 1.0, 2.0)],
                         usr=dict(maxparnode=100, order=3, mac=0.8)),
    # the Kirkwood/Born sphere workload (20,480 faces, leaf 500, p=3, theta=0.8)
    "sph5_born": dict(surf=(5, 2.0, (0.0, 0.0, 0.0)), atoms=[(0, 0, 0, 1.0, 2.0)],
                      usr=dict(maxparnode=500, order=3, mac=0.8)),
    # off-centre sphere, two charges, p=6, theta=0.5
    "sph4_off_p6": dict(surf=(4, 3.0, (0.3, -0.7, 1.1)),
                        atoms=[(0.8, -0.2, 1.4, 0.6, 1.5), (-0.4, -1.5, 0.6, -0.35, 1.7)],
                        usr=dict(maxparnode=200, order=6, mac=0.5)),
    # no salt (kappa = 0), p=2, small leaves
    "sph3_k0_p2": dict(surf=(3, 2.5, (0.1, 0.2, -0.3)), atoms=[(0.5, 0.1, -0.2, -1.0, 2.0)],
                       usr=dict(maxparnode=50, order=2, mac=0.8, bulk_strength=0.0)),
    # theta = 0: the treecode reduces to the direct sum
    "sph3_direct": dict(surf=(3, 2.0, (0.0, 0.0, 0.0)), atoms=[(0.3, 0, 0, 1.0, 2.0)],
                        usr=dict(maxparnode=100, order=3, mac=0.0)),
    # SES surface of a 150-atom synthetic molecule (own mesher, 35,955 faces),
    # 150 charges, leaf 500, p=3, theta=0.8; tol 1e-8 so GMRES restarts (DRLCAL)
    "mol150_restart": dict(mol=_MOL, usr=dict(maxparnode=500, order=3, mac=0.8),
                           harness=dict(tol=1e-8)),
    # the same molecule at p=6, theta=0.5
    "mol150_p6": dict(mol=_MOL, usr=dict(maxparnode=500, order=6, mac=0.5),
                      harness=dict(tol=1e-6)),
    # a tolerance GMRES cannot reach in (NRMAX+1)*MAXL = 20 steps (one restart): IERR = 1
    "sph3_cap": dict(surf=(3, 2.0, (0.2, -0.1, 0.05)),
                     atoms=[(0.4, 0.2, -0.3, 1.0, 1.5), (-0.5, 0.1, 0.2, -0.5, 1.5)],
                     usr=dict(maxparnode=100, order=3, mac=0.8),
                     harness=dict(tol=1e-30, nrmax=1)),
    # degenerate faces the reference's surface_area culling removes
    "sph3_degen": dict(surf=(3, 2.0, (0.0, 0.0, 0.0)), degen=True,
                       atoms=[(0.3, -0.2, 0.1, 1.0, 2.0)],
                       usr=dict(maxparnode=100, order=3, mac=0.8)),
    # NanoShaper-format surface files (isurf_type 2)
    "sph3_nano": dict(surf=(3, 2.2, (0.5, 0.25, -0.75)), fmt="nanoshaper",
                      atoms=[(0.5, 0.3, -0.6, -1.0, 2.0)],
                      usr=dict(maxparnode=100, order=3, mac=0.8, isurf_type=2)),
    # ESES / MS_Intersection-format surface file (isurf_type 4)
    "sph3_eses": dict(surf=(3, 1.8, (-0.25, 0.5, 0.125)), fmt="eses",
                      atoms=[(-0.2, 0.4, 0.2, 1.0, 1.8)],
                      usr=dict(maxparnode=100, order=3, mac=0.8, isurf_type=4)),
    # diagonal-scaling preconditioner MSOLVEprec (iprcd != 1)
    "sph3_diag": dict(surf=(3, 2.0, (0.0, 0.1, 0.0)), atoms=[(0.2, 0.1, 0.0, 1.0, 2.0)],
                      usr=dict(maxparnode=100, order=3, mac=0.8, iprcd=0)),
}

# Large cases: the reference needs minutes to hours of CPU for them, so they are
# generated on request (gen_golden.py NAME) and their tests are the slow ones.
LARGE = {
    # the 1,500-atom synthetic molecule (180,900 faces), full solve at tol 1e-4
    "mol1500": dict(mol=(1500, 1, 0.5), usr=dict(maxparnode=500, order=3, mac=0.8)),
    # the 8,000-atom benchmark molecule (483,572 faces): matvec and
    # preconditioner of the harness vector only
    "mol8000_mv": dict(mol=(8000, 1, 0.6), usr=dict(maxparnode=500, order=3, mac=0.8),
                       harness=dict(nosolve=True)),
}
CASES.update(LARGE)
SMALL = [k for k in CASES if k not in LARGE]


def _degenerate(v, n, f):
    """Edits of an icosphere that surface_area (src/readin.f90:264-333) must
    cull: an exact duplicate right after face 100 (centroid distance 0 to the
    preceding face), a zero-area face with a repeated vertex after face 200, a
    duplicate of face 300 exactly ten faces after it (the far edge of the
    ten-face window, max(1,i-10)..i-1), and after face 400 a sliver of area
    ~1e-8 along one of its edges through a new vertex that no kept face uses
    (output_potential then averages over no faces for it)."""
    f = [tuple(t) for t in f]
    out = []
    for i, t in enumerate(f):
        out.append(t)
        if i == 100:
            out.append(t)                               # exact duplicate (alert)
        if i == 200:
            out.append((t[0], t[0], t[1]))              # zero area
        if i == 300:
            out.append(("dup10", 300))                  # placeholder
        if i == 400:
            out.append(("sliver", 400))
    # resolve placeholders: the duplicate of face 300 lands exactly ten
    # positions after it; the sliver is a thin face along an edge of face 400
    res = []
    vv = [tuple(p) for p in v]
    nn = [tuple(p) for p in n]
    pending = []
    for t in out:
        if t[0] == "dup10":
            pending.append((len(res) + 9, f[t[1]]))
            continue
        if t[0] == "sliver":
            a, b, _ = f[t[1]]
            pa, pb = np.asarray(vv[a]), np.asarray(vv[b])
            vv.append(tuple(pa + (pb - pa) * 0.5 + 1e-7))
            nn.append(nn[a])
            res.append((a, b, len(vv) - 1))
            continue
        res.append(t)
        while pending and pending[0][0] == len(res):
            res.append(pending.pop(0)[1])
    return np.array(vv), np.array(nn), np.array(res, dtype=np.int64)


def _molecule(natom, seed, h):
    """Seeded synthetic molecule and its SES mesh; atom records rounded to the
    six decimals the PQR file carries."""
    from tabipb import surface
    xyz, q, r = surface.synth_molecule(natom, seed=seed)
    v, nv, f = surface.ses_mesh(xyz, r, probe=1.4, h=h)
    atoms = np.round(np.column_stack([xyz, q, r]), 6)
    return v, nv, f, atoms


def build(name):
    """Returns (verts, normals, faces0, atoms[n,5], UsrData)."""
    spec = CASES[name]
    if "mol" in spec:
        v, n, f, atoms = _molecule(*spec["mol"])
    else:
        v, n, f = _sphere(*spec["surf"])
        atoms = np.asarray(spec["atoms"], dtype=np.float64).reshape(-1, 5)
    if spec.get("degen"):
        v, n, f = _degenerate(v, n, f)
    usr = meshio.UsrData(fname="mol", den="5", **spec["usr"])
    return v, n, f, atoms, usr


def harness_opts(name):
    """Options of the reference harness for the case and the GMRES settings the
    tests repeat: dict(tol, nrmax, solve, msolve)."""
    h = CASES[name].get("harness", {})
    return dict(tol=h.get("tol", 1e-4), nrmax=h.get("nrmax", 10),
                solve=not h.get("nosolve", False), msolve=not h.get("nomsolve", False))


def input_hash(v, n, f, atoms):
    h = hashlib.sha256()
    for a in (np.ascontiguousarray(v, np.float64), np.ascontiguousarray(n, np.float64),
              np.ascontiguousarray(f, np.int64), np.ascontiguousarray(atoms, np.float64)):
        h.update(a.tobytes())
    return h.hexdigest()


def stage(name, dirpath):
    """Write the case as a reference run directory; returns the UsrData."""
    v, n, f, atoms, usr = build(name)
    os.makedirs(dirpath, exist_ok=True)
    pref = os.path.join(dirpath, "surf")
    meshio.write_msms(pref, v, n, f)
    meshio.write_pqr(os.path.join(dirpath, "mol.pqr"), atoms[:, :3], atoms[:, 3], atoms[:, 4])
    meshio.stage_reference_run(dirpath, usr, os.path.join(dirpath, "mol.pqr"), pref)
    fmt = CASES[name].get("fmt", "msms")
    if fmt != "msms":
        # the MSMS files stay where readin ignores them for this isurf_type
        os.remove(os.path.join(dirpath, "test_proteins", usr.fname + ".vert"))
        os.remove(os.path.join(dirpath, "test_proteins", usr.fname + ".face"))
    if fmt == "nanoshaper":
        meshio.write_nanoshaper(dirpath, v, n, f)
    elif fmt == "eses":
        meshio.write_eses(os.path.join(dirpath, "marching_cubes.obj"), v, n, f)
    return usr


def harness_x(n2):
    """The matvec test vector the harness uses (oracle/ref_harness.f90)."""
    i = np.arange(1, n2 + 1, dtype=np.float64)
    return np.sin(0.37 * i) + 0.5 * np.cos(1.3 * i) + 0.25
