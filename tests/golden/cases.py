"""Deterministic inputs of the golden cases (shared by gen_golden.py and tests).

Every surface is built with element-wise IEEE arithmetic only, so the same
bits come out on any machine; gen_golden.py records a SHA-256 of the inputs
and the tests check it before comparing outputs.
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


# name -> (surface spec, atoms [(x,y,z,q,r)], usrdata overrides)
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
}


def build(name):
    """Returns (verts, normals, faces0, atoms[n,5], UsrData)."""
    spec = CASES[name]
    v, n, f = _sphere(*spec["surf"])
    atoms = np.asarray(spec["atoms"], dtype=np.float64).reshape(-1, 5)
    usr = meshio.UsrData(fname="mol", den="5", **spec["usr"])
    return v, n, f, atoms, usr


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
    return usr


def harness_x(n2):
    """The matvec test vector the harness uses (oracle/ref_harness.f90)."""
    i = np.arange(1, n2 + 1, dtype=np.float64)
    return np.sin(0.37 * i) + 0.5 * np.cos(1.3 * i) + 0.25
