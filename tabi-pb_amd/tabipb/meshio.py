"""Surface and molecule file formats used by the TABI-PB driver.

* MSMS ``.vert`` / ``.face``: three header lines, the third starting with the
  record count; then one record per line.  The driver reads vertex lines as
  ``x y z nx ny nz`` in real*8 (src/readin.f90:204-224) and face lines as three
  1-based vertex indices (src/readin.f90:228-246), both list-directed, so any
  numeric precision parses.  Files written here carry full double precision.
* PQR: optional ``REMARK`` lines, then records whose first five tokens are
  skipped and whose next five are x, y, z, charge, radius parsed as real*4
  (src/readin.f90:29-84).  :func:`read_pqr` reproduces the real*4 rounding.
* ``usrdata.in``: ten positional ``key value`` lines (src/main.f90:39-50).
"""
from __future__ import annotations

import os
from dataclasses import dataclass

import numpy as np


# ----------------------------------------------------------------------------
# icosphere (Born / Kirkwood sphere surface)
# ----------------------------------------------------------------------------
def icosphere(subdiv: int, radius: float = 1.0):
    """Regular icosahedron, ``subdiv`` rounds of 4:1 midpoint subdivision, every
    vertex projected to ``radius``.  Returns (verts[V,3], normals[V,3],
    faces[F,3] 0-based, outward counter-clockwise).  subdiv=5 gives 10,242
    vertices and 20,480 faces."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0),
         (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11),
         (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6), (7, 1, 8),
         (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9),
         (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    def unit(p):
        p = np.asarray(p, dtype=np.float64)
        # element-wise IEEE arithmetic only (no BLAS), so every machine builds
        # the same bits
        return p / np.sqrt((p[0] * p[0] + p[1] * p[1]) + p[2] * p[2])

    verts = [unit(p) for p in v]
    faces = list(f)
    for _ in range(subdiv):
        cache = {}
        newf = []

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                verts.append(unit(verts[a] + verts[b]))
                cache[key] = len(verts) - 1
            return cache[key]

        for (a, b, c) in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            newf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = newf
    unit = np.array(verts)
    return unit * radius, unit.copy(), np.array(faces, dtype=np.int64)


# ----------------------------------------------------------------------------
# MSMS files
# ----------------------------------------------------------------------------
def write_msms(prefix: str, verts, normals, faces, nsph: int = 1,
               density: float = 0.0, probe: float = 1.4) -> None:
    """Write ``prefix.vert`` / ``prefix.face`` (faces given 0-based)."""
    verts = np.asarray(verts, dtype=np.float64)
    normals = np.asarray(normals, dtype=np.float64)
    faces = np.asarray(faces, dtype=np.int64)
    with open(prefix + ".vert", "w") as fh:
        fh.write("# MSMS solvent excluded surface vertices\n")
        fh.write("#vertex #sphere density probe_r\n")
        fh.write(f"{len(verts)} {nsph} {density:.2f} {probe:.2f}\n")
        rows = np.hstack([verts, normals])
        np.savetxt(fh, rows, fmt="%.17g")
    with open(prefix + ".face", "w") as fh:
        fh.write("# MSMS solvent excluded surface triangles\n")
        fh.write("#faces #sphere density probe_r\n")
        fh.write(f"{len(faces)} {nsph} {density:.2f} {probe:.2f}\n")
        np.savetxt(fh, faces + 1, fmt="%d")


def read_msms(prefix: str):
    """Read ``prefix.vert`` / ``prefix.face``; faces returned 0-based."""
    with open(prefix + ".vert") as fh:
        fh.readline(); fh.readline()
        nv = int(fh.readline().split()[0])
        vv = np.loadtxt(fh, dtype=np.float64, max_rows=nv, usecols=range(6), ndmin=2)
    with open(prefix + ".face") as fh:
        fh.readline(); fh.readline()
        nf = int(fh.readline().split()[0])
        ff = np.loadtxt(fh, dtype=np.int64, max_rows=nf, usecols=range(3), ndmin=2)
    return vv[:, :3].copy(), vv[:, 3:6].copy(), ff - 1


# ----------------------------------------------------------------------------
# PQR
# ----------------------------------------------------------------------------
def write_pqr(path: str, xyz, charge, radius) -> None:
    xyz = np.asarray(xyz, dtype=np.float64)
    with open(path, "w") as fh:
        fh.write("REMARK   synthetic molecule\n")
        for i, (p, q, r) in enumerate(zip(xyz, charge, radius)):
            fh.write(f"ATOM  {i + 1:5d}  X   MOL  {i + 1:4d}    "
                     f"{p[0]:.6f} {p[1]:.6f} {p[2]:.6f} {q:.6f} {r:.6f}\n")
        fh.write("END\n")


def read_pqr(path: str):
    """Atoms as the driver sees them: real*4 values widened to real*8."""
    xyzqr = []
    with open(path) as fh:
        lines = fh.readlines()
    k = 0
    while k < len(lines) and lines[k].split() and lines[k].split()[0][:6] == "REMARK":
        k += 1
    for ln in lines[k:]:
        tok = ln.split()
        if not tok or tok[0][:3] in ("END", "TER") or len(tok) < 10:
            continue
        xyzqr.append([float(t) for t in tok[5:10]])
    a = np.asarray(xyzqr, dtype=np.float32).astype(np.float64).reshape(-1, 5)
    return a[:, :3].copy(), a[:, 3].copy(), a[:, 4].copy()


# ----------------------------------------------------------------------------
# usrdata.in
# ----------------------------------------------------------------------------
@dataclass
class UsrData:
    fname: str = "molecule"
    den: str = "10"
    epsp: float = 1.0
    epsw: float = 80.0
    bulk_strength: float = 0.15
    order: int = 3
    maxparnode: int = 500
    mac: float = 0.8
    iprcd: int = 1
    isurf_type: int = 1

    def write(self, path: str) -> None:
        with open(path, "w") as fh:
            fh.write(f"fname {self.fname}\nden {self.den}\nepsp {self.epsp}\n"
                     f"epsw {self.epsw}\nbulk_strength {self.bulk_strength}\n"
                     f"order {self.order}\nmaxparnode {self.maxparnode}\n"
                     f"mac {self.mac}\niprcd {self.iprcd}\nisurf_type {self.isurf_type}\n")

    @classmethod
    def read(cls, path: str) -> "UsrData":
        vals = []
        with open(path) as fh:
            for ln in fh:
                tok = ln.split()
                if len(tok) >= 2:
                    vals.append(tok[1])
        u = cls()
        u.fname, u.den = vals[0], vals[1]
        u.epsp, u.epsw, u.bulk_strength = float(vals[2]), float(vals[3]), float(vals[4])
        u.order, u.maxparnode = int(vals[5]), int(vals[6])
        u.mac = float(vals[7])
        u.iprcd, u.isurf_type = int(vals[8]), int(vals[9])
        return u


def stage_reference_run(dirpath: str, usr: UsrData, pqr_src: str, msms_prefix: str) -> None:
    """Lay out a run directory the way the reference program expects it:
    ``usrdata.in`` plus ``test_proteins/<fname>.{pqr,vert,face}``
    (src/readin.f90:18,30,204,228)."""
    import shutil
    os.makedirs(os.path.join(dirpath, "test_proteins"), exist_ok=True)
    usr.write(os.path.join(dirpath, "usrdata.in"))
    base = os.path.join(dirpath, "test_proteins", usr.fname)
    shutil.copyfile(pqr_src, base + ".pqr")
    shutil.copyfile(msms_prefix + ".vert", base + ".vert")
    shutil.copyfile(msms_prefix + ".face", base + ".face")
