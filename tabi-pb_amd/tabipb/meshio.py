"""Surface and molecule file formats used by the TABI-PB driver.

* MSMS ``.vert`` / ``.face``: three header lines, the third starting with the
  record count; then one record per line.  The driver reads vertex lines as
  ``x y z nx ny nz`` in real*8 (src/readin.f90:204-224) and face lines as three
  1-based vertex indices (src/readin.f90:228-246), both list-directed, so any
  numeric precision parses.  Files written here carry full double precision.
* PQR: optional ``REMARK`` lines, then records whose first five tokens are
  skipped and whose next five are x, y, z, charge, radius parsed as real*4
  (src/readin.f90:29-84).  :func:`read_pqr` reproduces the real*4 rounding.
* NanoShaper ``triangulatedSurf.vert``/``.face`` (isurf_type 2,
  src/readin.f90:124-193) and ESES ``marching_cubes.obj`` (isurf_type 4,
  src/readin.f90:90-123), read with Fortran list-directed semantics
  (:class:`ListDirected`).
* ``surface_potential.dat`` written by output_potential (src/main.f90:226-326).
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
# Fortran list-directed input (the reference reads every surface file this way)
# ----------------------------------------------------------------------------
class ListDirected:
    """Sequential list-directed READ semantics over a text file: every READ
    starts at the next record (line), takes blank/comma-separated values across
    as many records as it needs, and drops what is left of its last record; a
    ``/`` ends the READ early (the remaining items keep their values, here None)."""

    def __init__(self, path: str):
        with open(path) as fh:
            self._lines = fh.read().splitlines()
        self._k = 0

    def read(self, nitems: int):
        out = []
        while len(out) < nitems:
            if self._k >= len(self._lines):
                raise EOFError("end of file in list-directed read")
            line = self._lines[self._k]
            self._k += 1
            stop = False
            for tok in line.replace(",", " ").split():
                if tok.startswith("/"):
                    stop = True
                    break
                out.append(tok)
                if len(out) == nitems:
                    break
            if stop:
                out += [None] * (nitems - len(out))
        return out


def _fint(tok):
    return int(tok)


def _freal(tok):
    # Fortran accepts d/D exponents
    return float(tok.replace("d", "e").replace("D", "e"))


# ----------------------------------------------------------------------------
# NanoShaper (isurf_type 2) and ESES / MS_Intersection (isurf_type 4) files
# ----------------------------------------------------------------------------
def write_nanoshaper(dirpath: str, verts, normals, faces) -> None:
    """``triangulatedSurf.vert`` / ``.face`` as read at src/readin.f90:154-193:
    two header records, a record starting with the count, then per vertex
    ``x y z nx ny nz kk naff nafft`` and per face ``a b c mface x`` (1-based)."""
    verts = np.asarray(verts, dtype=np.float64)
    normals = np.asarray(normals, dtype=np.float64)
    faces = np.asarray(faces, dtype=np.int64)
    with open(os.path.join(dirpath, "triangulatedSurf.vert"), "w") as fh:
        fh.write("# NanoShaper triangulated surface vertices\n")
        fh.write("#vertex #sphere density probe_r\n")
        fh.write(f"{len(verts)} 1 0.00 1.40\n")
        for i, (p, nv) in enumerate(zip(verts, normals)):
            fh.write(f"{p[0]:.17g} {p[1]:.17g} {p[2]:.17g} {nv[0]:.17g} {nv[1]:.17g} "
                     f"{nv[2]:.17g} 0 {i % 7 + 1} 1\n")
    with open(os.path.join(dirpath, "triangulatedSurf.face"), "w") as fh:
        fh.write("# NanoShaper triangulated surface faces\n")
        fh.write("#faces #sphere density probe_r\n")
        fh.write(f"{len(faces)} 1 0.00 1.40\n")
        for i, f in enumerate(faces + 1):
            fh.write(f"{f[0]} {f[1]} {f[2]} {i % 3 + 1} {i + 1}\n")


def read_nanoshaper(dirpath: str):
    """Vertices, normals and 0-based faces as readin takes them from NanoShaper's
    files (src/readin.f90:154-193), with list-directed semantics."""
    rv = ListDirected(os.path.join(dirpath, "triangulatedSurf.vert"))
    rv.read(1); rv.read(1)
    nspt = _fint(rv.read(1)[0])
    rf = ListDirected(os.path.join(dirpath, "triangulatedSurf.face"))
    rf.read(1); rf.read(1)
    nface = _fint(rf.read(1)[0])
    v = np.empty((nspt, 6))
    for i in range(nspt):
        t = rv.read(9)
        v[i] = [_freal(a) for a in t[:6]]
    f = np.empty((nface, 3), dtype=np.int64)
    for i in range(nface):
        t = rf.read(5)
        f[i] = [_fint(a) for a in t[:3]]
    return v[:, :3].copy(), v[:, 3:].copy(), f - 1


def write_eses(path: str, verts, normals, faces) -> None:
    """``marching_cubes.obj`` as read at src/readin.f90:92-123: the vertex count,
    one ``tag x y z nx ny nz`` record per vertex, the face count, one
    ``tag a b c`` record per face (1-based)."""
    verts = np.asarray(verts, dtype=np.float64)
    normals = np.asarray(normals, dtype=np.float64)
    faces = np.asarray(faces, dtype=np.int64)
    with open(path, "w") as fh:
        fh.write(f"{len(verts)}\n")
        for p, nv in zip(verts, normals):
            fh.write(f"v {p[0]:.17g} {p[1]:.17g} {p[2]:.17g} {nv[0]:.17g} {nv[1]:.17g} "
                     f"{nv[2]:.17g}\n")
        fh.write(f"{len(faces)}\n")
        for f in faces + 1:
            fh.write(f"f {f[0]} {f[1]} {f[2]}\n")


def read_eses(path: str):
    """Vertices, normals and 0-based faces of an ESES ``marching_cubes.obj``
    (src/readin.f90:92-123)."""
    r = ListDirected(path)
    nspt = _fint(r.read(1)[0])
    v = np.empty((nspt, 6))
    for i in range(nspt):
        t = r.read(7)
        v[i] = [_freal(a) for a in t[1:7]]
    nface = _fint(r.read(1)[0])
    f = np.empty((nface, 3), dtype=np.int64)
    for i in range(nface):
        t = r.read(4)
        f[i] = [_fint(a) for a in t[1:4]]
    return v[:, :3].copy(), v[:, 3:].copy(), f - 1


# ----------------------------------------------------------------------------
# surface_potential.dat (output_potential, src/main.f90:226-326)
# ----------------------------------------------------------------------------
def _fortran_f(x: float, w: int, d: int) -> str:
    """Fortran Fw.d edit descriptor: fixed notation right-aligned in w columns,
    asterisks when it does not fit, NaN/Infinity spelled out."""
    if np.isnan(x):
        s = "NaN"
    elif np.isinf(x):
        s = "Infinity" if x > 0 else "-Infinity"
    else:
        s = f"{x:.{d}f}"
    if len(s) > w:
        return "*" * w
    return s.rjust(w)


def write_surface_potential(path: str, vpos, vnrm, vpot, faces) -> None:
    """The file output_potential writes: ``nspt nface``, then per vertex
    ``(i10,6f12.6,2f20.10)`` of index, position, normal and the two averaged
    values, then the three 1-based vertex indices of every face (the integer
    records list-directed, as the flang runtime spaces them)."""
    vpos = np.asarray(vpos, dtype=np.float64)
    vnrm = np.asarray(vnrm, dtype=np.float64)
    vpot = np.asarray(vpot, dtype=np.float64).reshape(-1, 2)
    faces = np.asarray(faces, dtype=np.int64)
    with open(path, "w") as fh:
        fh.write(f" {len(vpos)} {len(faces)}\n")
        for i in range(len(vpos)):
            fh.write(f"{i + 1:10d}" + "".join(_fortran_f(a, 12, 6) for a in vpos[i])
                     + "".join(_fortran_f(a, 12, 6) for a in vnrm[i])
                     + "".join(_fortran_f(a, 20, 10) for a in vpot[i]) + "\n")
        for f in faces + 1:
            fh.write(f" {f[0]} {f[1]} {f[2]}\n")


def read_surface_potential(path: str):
    """Parses a surface_potential.dat (fixed columns for the vertex records);
    returns (vpos[nspt,3], vnrm[nspt,3], vpot[nspt,2], faces[nface,3] 0-based)."""
    with open(path) as fh:
        lines = fh.read().splitlines()
    nspt, nface = (int(t) for t in lines[0].split()[:2])
    vpos = np.empty((nspt, 3)); vnrm = np.empty((nspt, 3)); vpot = np.empty((nspt, 2))

    def num(s):
        s = s.strip()
        return float("nan") if s.startswith("*") else float(s)

    for i in range(nspt):
        ln = lines[1 + i]
        f = [num(ln[10 + 12 * k:22 + 12 * k]) for k in range(6)]
        vpos[i], vnrm[i] = f[:3], f[3:]
        vpot[i] = [num(ln[82:102]), num(ln[102:122])]
    faces = np.array([[int(t) for t in lines[1 + nspt + j].split()[:3]] for j in range(nface)],
                     dtype=np.int64).reshape(-1, 3)
    return vpos, vnrm, vpot, faces - 1


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
