"""Host-side mirror of the reference driver (src/main.f90, src/readin.f90).

Reads the same inputs (usrdata.in, PQR, MSMS surface), reproduces the
driver's discretisation arithmetic bit for bit (IEEE float64 element-wise
operations in the reference's operand order; numpy never contracts to FMA), and
hands the solve to libtabi_hip.so.
"""
from __future__ import annotations

import os
from dataclasses import dataclass

import numpy as np

from . import hip, meshio

PARA = 332.0716  # src/main.f90:64


def kappa_of(bulk_strength: float, eps_w: float) -> float:
    """src/main.f90:62-63; the literal 8.430325455 is single precision."""
    return float(np.sqrt(np.float64(np.float32(8.430325455)) * bulk_strength / eps_w))


def _dot3(a, b):
    return (a[:, 0] * b[:, 0] + a[:, 1] * b[:, 1]) + a[:, 2] * b[:, 2]


def _centroid(v, faces):
    t = 1.0 / 3.0
    return ((0.0 + t * v[faces[:, 0]]) + t * v[faces[:, 1]]) + t * v[faces[:, 2]]


def _heron(p0, p1, p2):
    d01, d02, d12 = p0 - p1, p0 - p2, p1 - p2
    aa, bb, cc = np.sqrt(_dot3(d01, d01)), np.sqrt(_dot3(d02, d02)), np.sqrt(_dot3(d12, d12))
    s = 0.5 * ((aa + bb) + cc)
    return np.sqrt(((s * (s - aa)) * (s - bb)) * (s - cc))


def cull_faces(vpos, faces):
    """surface_area (src/readin.f90:264-333): drop faces with Heron area < 1e-5
    or whose centroid lies within squared distance 1e-5 of any of the 10
    preceding faces in the original numbering.  Returns (kept faces, total area
    of all faces)."""
    vpos = np.asarray(vpos, dtype=np.float64)
    faces = np.asarray(faces, dtype=np.int64)
    nf = faces.shape[0]
    xx = _centroid(vpos, faces)
    area = _heron(vpos[faces[:, 0]], vpos[faces[:, 1]], vpos[faces[:, 2]])
    alert = np.zeros(nf, dtype=bool)
    for k in range(1, 11):
        if k >= nf:
            break
        d = xx[k:] - xx[:-k]
        alert[k:] |= _dot3(d, d) < 1.0e-5
    keep = ~((area < 1.0e-5) | alert)
    s_area = 0.0
    for a in area:  # sequential sum, as the reference accumulates it
        s_area += a
    return faces[keep], s_area


def face_geometry(vpos, vnrm, faces):
    """form_matrix geometry (src/main.f90:569-594): centroid, normalised mean
    vertex normal, Heron area -- original face order."""
    vpos = np.asarray(vpos, dtype=np.float64)
    vnrm = np.asarray(vnrm, dtype=np.float64)
    faces = np.asarray(faces, dtype=np.int64)
    r0 = _centroid(vpos, faces)
    v0 = _centroid(vnrm, faces)
    v0 = v0 / np.sqrt(_dot3(v0, v0))[:, None]
    area = _heron(vpos[faces[:, 0]], vpos[faces[:, 1]], vpos[faces[:, 2]])
    return r0, v0, area


@dataclass
class Problem:
    """One PB problem as the reference driver sets it up."""
    usr: meshio.UsrData
    xyz: np.ndarray       # face centroids, original order (n,3)
    nrm: np.ndarray       # face normals (n,3)
    area: np.ndarray      # face areas (n,)
    qpos: np.ndarray      # charge positions (nq,3)
    q: np.ndarray         # charges (nq,)
    vpos: np.ndarray = None
    vnrm: np.ndarray = None
    faces: np.ndarray = None

    @property
    def eps(self):
        return self.usr.epsw / self.usr.epsp

    @property
    def kappa(self):
        return kappa_of(self.usr.bulk_strength, self.usr.epsw)

    def params(self, precision=0, flags=0, order=None, maxparnode=None, theta=None):
        u = self.usr
        if u.iprcd != 1:
            # MSOLVE => MSOLVEprec unless iprcd = 1 (src/main.f90:118-123)
            flags |= hip.FLAG_DIAG_PREC
        return hip.make_params(self.eps, self.kappa, u.mac if theta is None else theta,
                               u.order if order is None else order,
                               u.maxparnode if maxparnode is None else maxparnode,
                               precision, flags)


def problem_from_mesh(usr: meshio.UsrData, v, vn, f, qpos, q) -> Problem:
    """Culls the faces as readin does (surface_area) and forms the geometry."""
    f, _ = cull_faces(v, f)
    xyz, nrm, area = face_geometry(v, vn, f)
    return Problem(usr, xyz, nrm, area, qpos, q, v, vn, f)


def load_problem(pqr_path: str, msms_prefix: str, usr: meshio.UsrData) -> Problem:
    qpos, q, _ = meshio.read_pqr(pqr_path)
    v, vn, f = meshio.read_msms(msms_prefix)
    return problem_from_mesh(usr, v, vn, f, qpos, q)


def load_run_dir(path: str) -> Problem:
    """A directory laid out for the reference program: usrdata.in,
    test_proteins/<fname>.pqr and the surface files of its isurf_type -- MSMS
    test_proteins/<fname>.{vert,face} (1), NanoShaper triangulatedSurf.{vert,face}
    (2) or ESES marching_cubes.obj (4) in the directory itself
    (src/readin.f90:18-30,90-246)."""
    usr = meshio.UsrData.read(os.path.join(path, "usrdata.in"))
    base = os.path.join(path, "test_proteins", usr.fname)
    qpos, q, _ = meshio.read_pqr(base + ".pqr")
    if usr.isurf_type == 4:
        v, vn, f = meshio.read_eses(os.path.join(path, "marching_cubes.obj"))
    elif usr.isurf_type == 2:
        v, vn, f = meshio.read_nanoshaper(path)
    else:
        v, vn, f = meshio.read_msms(base)
    return problem_from_mesh(usr, v, vn, f, qpos, q)


def output_potential(S: "hip.Surface", prob: Problem, x, path: str = None):
    """output_potential (src/main.f90:226-326) for a solved surface: the vertex
    values, and surface_potential.dat written to path when given."""
    vpot = S.vertex_potential(prob.vpos, prob.faces, x)
    if path is not None:
        hip.write_surface_potential(path, prob.vpos, prob.vnrm, vpot, prob.faces)
    return vpot


def solve(prob: Problem, tol=1e-4, precision=0, flags=0, model=None, **over):
    """Full GPU solve (from a zero start, or from the learned guess of a
    hip.WsModel); returns the dict of hip.solve plus the original-order solution."""
    res = hip.solve(prob.params(precision, flags, **over), prob.xyz, prob.nrm, prob.area,
                    prob.qpos, prob.q, prob.usr.epsp, tol, model=model)
    perm = res["orderarr"].astype(np.int64) - 1
    n = prob.xyz.shape[0]
    xo = np.empty(2 * n)
    xo[perm] = res["x"][:n]
    xo[n + perm] = res["x"][n:]
    res["x_orig"] = xo
    return res
