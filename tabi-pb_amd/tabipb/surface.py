"""Synthetic molecules and a solvent-excluded-surface mesher.

The reference obtains its triangulations from closed third-party executables
(MSMS / NanoShaper / MS_Intersection, src/readin.f90:90-246).  This module is the
project's own stand-in that writes the same MSMS ``.vert``/``.face`` format:

* :func:`synth_molecule` -- a compact, seeded, self-avoiding chain: backbone
  atoms 3.8 A apart with one side atom 1.5-3 A off each, radii U(1.4, 2.0),
  charges N(0, 0.35) shifted to a fixed net charge.
* :func:`ses_mesh` -- the solvent-excluded surface for a probe radius on a
  uniform grid: the solvent-accessible region is rasterised, a Euclidean
  distance transform gives each interior grid point its distance to the
  accessible probe-centre set, and the level set ``distance = probe`` is
  triangulated by marching tetrahedra (six tetrahedra per cube along a common
  diagonal, so neighbouring cubes share faces and the mesh is closed).
  Vertex normals are the area-weighted means of the adjacent face normals.
"""
from __future__ import annotations

import numpy as np
from scipy import ndimage


# ---------------------------------------------------------------------------
# synthetic molecule
# ---------------------------------------------------------------------------
class _CellGrid:
    """Points in cubic cells of side ``cell`` (capacity ``cap`` each) for fast
    minimum-distance queries of many candidates at once."""

    def __init__(self, half, cell=3.0, cap=12):
        self.cell = cell
        self.n = int(np.ceil(2 * half / cell)) + 3
        self.off = half + cell
        self.pts = np.full((self.n, self.n, self.n, cap, 3), 1e9)
        self.cnt = np.zeros((self.n, self.n, self.n), dtype=np.int64)
        self.cap = cap

    def _idx(self, p):
        return np.clip(np.floor((p + self.off) / self.cell).astype(np.int64), 1, self.n - 2)

    def add(self, p):
        i, j, k = self._idx(p)
        c = self.cnt[i, j, k]
        if c < self.cap:
            self.pts[i, j, k, c] = p
            self.cnt[i, j, k] = c + 1

    def min_dist2(self, cand):
        """Squared distance from each candidate (K,3) to its nearest stored point."""
        idx = self._idx(cand)
        o = np.array([(a, b, c) for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1)])
        nb = idx[:, None, :] + o[None, :, :]
        q = self.pts[nb[..., 0], nb[..., 1], nb[..., 2]]          # (K,27,cap,3)
        d = q - cand[:, None, None, :]
        return (d * d).sum(axis=-1).reshape(len(cand), -1).min(axis=1)


def synth_molecule(natom: int, seed: int, net_charge: float = -3.0, density: float = 1 / 20.0):
    """Returns (xyz[natom,3], charge[natom], radius[natom]).

    Backbone atoms follow a self-avoiding walk with 3.8 A steps, drawn towards
    the centre so the chain folds into a globule of roughly protein density
    (``density`` atoms per cubic Angstrom); each backbone atom carries one side
    atom 1.5-3 A away.  Radii are U(1.4, 2.0) A and charges N(0, 0.35) e, shifted
    so they sum to ``net_charge``."""
    rng = np.random.default_rng(seed)
    nb = (natom + 1) // 2
    r_glob = (3.0 * natom / (4.0 * np.pi * density)) ** (1.0 / 3.0)
    grid = _CellGrid(2.0 * r_glob + 20.0)
    K = 48
    pos = np.zeros(3)
    grid.add(pos)
    backbone = [pos]
    while len(backbone) < nb:
        d = rng.standard_normal((K, 3))
        d /= np.sqrt((d * d).sum(axis=1))[:, None]
        rr = np.sqrt(pos @ pos)
        if rr > 0.6 * r_glob:
            d -= 1.5 * (rr / r_glob) * pos / rr
            d /= np.sqrt((d * d).sum(axis=1))[:, None]
        cand = pos + 3.8 * d
        ok = np.nonzero(grid.min_dist2(cand) >= 3.0 * 3.0)[0]
        if len(ok):
            pos = cand[ok[0]]
            grid.add(pos)
            backbone.append(pos)
        else:  # trapped: let the globule grow a little and continue the chain
            # from one of the outermost atoms
            r_glob *= 1.002
            bb = np.array(backbone)
            rad = (bb * bb).sum(axis=1)
            top = np.argsort(rad)[-max(1, len(bb) // 20):]
            pos = bb[top[rng.integers(len(top))]]
    side = []
    for b in backbone:
        d = rng.standard_normal((K, 3))
        d /= np.sqrt((d * d).sum(axis=1))[:, None]
        cand = b + rng.uniform(1.5, 3.0, K)[:, None] * d
        ok = np.nonzero(grid.min_dist2(cand) >= 2.0 * 2.0)[0]
        s_ = cand[ok[0]] if len(ok) else cand[0]
        grid.add(s_)
        side.append(s_)
    xyz = np.empty((2 * nb, 3))
    xyz[0::2] = np.array(backbone)
    xyz[1::2] = np.array(side)
    xyz = xyz[:natom]
    xyz -= xyz.mean(axis=0)
    radius = rng.uniform(1.4, 2.0, natom)
    q = rng.normal(0.0, 0.35, natom)
    q += (net_charge - q.sum()) / natom
    return xyz, q, radius


# ---------------------------------------------------------------------------
# solvent-excluded surface
# ---------------------------------------------------------------------------
# Kuhn triangulation of the unit cube into 6 tetrahedra sharing the 0-7 diagonal;
# corners are numbered by bits (x=1, y=2, z=4)
_TETS = np.array([[0, 1, 3, 7], [0, 3, 2, 7], [0, 2, 6, 7],
                  [0, 6, 4, 7], [0, 4, 5, 7], [0, 5, 1, 7]], dtype=np.int64)
_CORNER = np.array([[(c >> 0) & 1, (c >> 1) & 1, (c >> 2) & 1] for c in range(8)], dtype=np.int64)


def ses_field(xyz, radius, probe=1.4, h=0.5, pad=None):
    """Signed field on a grid: distance to the probe-centre set minus the probe
    radius (positive inside the solvent-excluded region)."""
    xyz = np.asarray(xyz, dtype=np.float64)
    radius = np.asarray(radius, dtype=np.float64)
    rs = radius + probe
    if pad is None:
        pad = probe + 3 * h
    lo = xyz.min(axis=0) - rs.max() - pad
    hi = xyz.max(axis=0) + rs.max() + pad
    shape = np.ceil((hi - lo) / h).astype(np.int64) + 1
    sas = np.zeros(shape, dtype=bool)
    for c, r in zip(xyz, rs):
        i0 = np.maximum(np.floor((c - r - lo) / h).astype(np.int64), 0)
        i1 = np.minimum(np.ceil((c + r - lo) / h).astype(np.int64) + 1, shape)
        gx = lo[0] + h * np.arange(i0[0], i1[0])
        gy = lo[1] + h * np.arange(i0[1], i1[1])
        gz = lo[2] + h * np.arange(i0[2], i1[2])
        d2 = ((gx - c[0]) ** 2)[:, None, None] + ((gy - c[1]) ** 2)[None, :, None] + \
             ((gz - c[2]) ** 2)[None, None, :]
        sas[i0[0]:i1[0], i0[1]:i1[1], i0[2]:i1[2]] |= d2 <= r * r
    # grid points outside every enlarged ball are admissible probe centres;
    # the distance of interior points to that set (in Angstrom)
    d = ndimage.distance_transform_edt(sas, sampling=h)
    f = (d - probe).astype(np.float64)
    f[~sas] = -probe
    return f, lo, h


def marching_tets(f, lo, h, snap=0.02):
    """Triangulate the zero level set of a grid field (positive inside).
    Returns (verts[V,3], faces[F,3]) with faces oriented outward."""
    f = np.array(f, dtype=np.float64, copy=True)
    # keep crossings off the grid points so no triangle degenerates to a point
    small = np.abs(f) < snap * h
    f[small] = np.where(f[small] >= 0, snap * h, -snap * h)
    nx, ny, nz = f.shape
    inside = f > 0
    # cubes with a sign change among their 8 corners
    c = [inside[ox:nx - 1 + ox, oy:ny - 1 + oy, oz:nz - 1 + oz]
         for ox, oy, oz in _CORNER]
    anyin = np.zeros_like(c[0])
    allin = np.ones_like(c[0])
    for cc in c:
        anyin |= cc
        allin &= cc
    cubes = np.argwhere(anyin & ~allin)
    if cubes.size == 0:
        return np.zeros((0, 3)), np.zeros((0, 3), dtype=np.int64)
    # corners of every tetrahedron: (ncube, 6, 4, 3)
    corner = cubes[:, None, None, :] + _CORNER[_TETS][None, :, :, :]
    corner = corner.reshape(-1, 4, 3)
    val = f[corner[..., 0], corner[..., 1], corner[..., 2]]
    ids = (corner[..., 0] * ny + corner[..., 1]) * nz + corner[..., 2]
    ins = val > 0
    nin = ins.sum(axis=1)
    keep = (nin > 0) & (nin < 4)
    corner, val, ids, ins, nin = corner[keep], val[keep], ids[keep], ins[keep], nin[keep]

    tris_k = []
    ntot = nx * ny * nz
    order = np.argsort(~ins, axis=1, kind="stable")  # inside corners first
    # one inside corner: triangle on the 3 edges from it
    for cnt in (1, 3):
        sel = nin == cnt
        if not np.any(sel):
            continue
        o = order[sel]
        if cnt == 1:
            apex, others = o[:, 0], o[:, 1:]
        else:
            apex, others = o[:, 3], o[:, :3]
        ks, ps = [], []
        for m in range(3):
            k, p = _edge(ids[sel], val[sel], corner[sel], apex, others[:, m], lo, h, ntot)
            ks.append(k)
            ps.append(p)
        tri_k = np.stack(ks, axis=1)
        apex_pt = np.take_along_axis(corner[sel], apex[:, None, None], 1)[:, 0] * h + lo
        # outward = away from inside corners
        sign = 1.0 if cnt == 1 else -1.0
        tris_k.append((tri_k, np.stack(ps, axis=1), apex_pt, sign))
    sel = nin == 2
    if np.any(sel):
        o = order[sel]
        a0, a1, b0, b1 = o[:, 0], o[:, 1], o[:, 2], o[:, 3]
        C, V, I = corner[sel], val[sel], ids[sel]
        k00, p00 = _edge(I, V, C, a0, b0, lo, h, ntot)
        k01, p01 = _edge(I, V, C, a0, b1, lo, h, ntot)
        k11, p11 = _edge(I, V, C, a1, b1, lo, h, ntot)
        k10, p10 = _edge(I, V, C, a1, b0, lo, h, ntot)
        in_pt = 0.5 * (np.take_along_axis(C, a0[:, None, None], 1)[:, 0] +
                       np.take_along_axis(C, a1[:, None, None], 1)[:, 0]) * h + lo
        q = np.stack([k00, k01, k11, k10], axis=1)
        qp = np.stack([p00, p01, p11, p10], axis=1)
        tris_k.append((q[:, [0, 1, 2]], qp[:, [0, 1, 2]], in_pt, 1.0))
        tris_k.append((q[:, [0, 2, 3]], qp[:, [0, 2, 3]], in_pt, 1.0))
    all_k, all_f = [], []
    for tri_k, tri_p, ref_pt, sign in tris_k:
        nrm = np.cross(tri_p[:, 1] - tri_p[:, 0], tri_p[:, 2] - tri_p[:, 0])
        cen = tri_p.mean(axis=1)
        outward = ((cen - ref_pt) * nrm).sum(axis=1) * sign
        flip = outward < 0
        tri_k = tri_k.copy()
        tri_k[flip] = tri_k[flip][:, [0, 2, 1]]
        all_k.append(tri_k)
        all_f.append(tri_p.reshape(-1, 3))
    keys = np.concatenate([t.reshape(-1) for t in all_k])
    pts = np.concatenate(all_f)
    ukeys, first, inv = np.unique(keys, return_index=True, return_inverse=True)
    verts = pts[first]
    faces = inv.reshape(-1, 3)
    return verts, faces


def _edge(ids, val, corner, a, b, lo, h, ntot):
    ia = np.take_along_axis(ids, a[:, None], 1)[:, 0]
    ib = np.take_along_axis(ids, b[:, None], 1)[:, 0]
    va = np.take_along_axis(val, a[:, None], 1)[:, 0]
    vb = np.take_along_axis(val, b[:, None], 1)[:, 0]
    pa = np.take_along_axis(corner, a[:, None, None], 1)[:, 0] * h + lo
    pb = np.take_along_axis(corner, b[:, None, None], 1)[:, 0] * h + lo
    # parametrise from the lower grid id so both tets sharing the edge agree
    swap = ia > ib
    va2 = np.where(swap, vb, va)
    vb2 = np.where(swap, va, vb)
    pa2 = np.where(swap[:, None], pb, pa)
    pb2 = np.where(swap[:, None], pa, pb)
    t = va2 / (va2 - vb2)
    p = pa2 + t[:, None] * (pb2 - pa2)
    k = np.minimum(ia, ib) * ntot + np.maximum(ia, ib)
    return k, p


def _heron_ref(v, faces):
    """Face areas exactly as the driver computes them (Heron's formula,
    src/readin.f90:285-288,336-340); NaN for numerically collinear faces."""
    p0, p1, p2 = v[faces[:, 0]], v[faces[:, 1]], v[faces[:, 2]]

    def ln(a, b):
        d = a - b
        return np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])

    aa, bb, cc = ln(p0, p1), ln(p0, p2), ln(p1, p2)
    s = 0.5 * ((aa + bb) + cc)
    with np.errstate(invalid="ignore"):
        return np.sqrt(((s * (s - aa)) * (s - bb)) * (s - cc))


def _collapse(v, faces, pairs):
    """Merge each connected set of vertices joined by ``pairs`` into its mean
    position; drop faces that lose a vertex and unused vertices."""
    nv = len(v)
    parent = np.arange(nv)

    def find(a):
        while True:
            pa = parent[a]
            if np.array_equal(pa, a):
                return a
            parent[a] = parent[pa]
            a = parent[a]

    if len(pairs):
        a, b = pairs[:, 0], pairs[:, 1]
        for _ in range(64):  # vectorised pointer jumping to a fixed point
            ra, rb = find(a), find(b)
            lo_, hi_ = np.minimum(ra, rb), np.maximum(ra, rb)
            if np.all(lo_ == hi_):
                break
            parent[hi_] = lo_
        root = find(np.arange(nv))
    else:
        root = np.arange(nv)
    cnt = np.bincount(root, minlength=nv).astype(np.float64)
    acc = np.zeros_like(v)
    np.add.at(acc, root, v)
    newv = acc / np.maximum(cnt, 1)[:, None]
    f = root[faces]
    ok = (f[:, 0] != f[:, 1]) & (f[:, 1] != f[:, 2]) & (f[:, 0] != f[:, 2])
    f = f[ok]
    # drop faces duplicated by the collapse (same vertex set)
    _, first = np.unique(np.sort(f, axis=1), axis=0, return_index=True)
    f = f[np.sort(first)]
    used = np.unique(f)
    remap = -np.ones(nv, dtype=np.int64)
    remap[used] = np.arange(len(used))
    return newv[used], remap[f]


def _smooth(v, faces, iters=4, lam=0.5):
    """Tangential Laplacian smoothing: move each vertex towards the mean of its
    neighbours within its tangent plane (the surface shape is kept, triangle
    shapes improve)."""
    e = np.concatenate([faces[:, [0, 1]], faces[:, [1, 2]], faces[:, [2, 0]]])
    e = np.concatenate([e, e[:, ::-1]])
    deg = np.bincount(e[:, 0], minlength=len(v)).astype(np.float64)
    for _ in range(iters):
        nb = np.zeros_like(v)
        np.add.at(nb, e[:, 0], v[e[:, 1]])
        d = nb / np.maximum(deg, 1)[:, None] - v
        n = vertex_normals(v, faces)
        d -= (d * n).sum(axis=1)[:, None] * n
        v = v + lam * d
    return v


def clean_mesh(v, faces, h, min_area=1e-4):
    """Collapse edges shorter than h/5, smooth tangentially, then collapse the
    shortest edge of every face whose driver-side Heron area is NaN or below
    ``min_area`` until none is left."""
    def edges_of(f):
        return np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])

    e = edges_of(faces)
    ln = np.sqrt(((v[e[:, 0]] - v[e[:, 1]]) ** 2).sum(axis=1))
    v, faces = _collapse(v, faces, e[ln < 0.2 * h])
    v = _smooth(v, faces)
    for _ in range(10):
        a = _heron_ref(v, faces)
        bad = ~(a >= min_area)
        if not np.any(bad):
            break
        fb = faces[bad]
        cand = np.stack([fb[:, [0, 1]], fb[:, [1, 2]], fb[:, [2, 0]]], axis=1)
        ln = np.sqrt(((v[cand[..., 0]] - v[cand[..., 1]]) ** 2).sum(axis=2))
        pick = cand[np.arange(len(fb)), np.argmin(ln, axis=1)]
        v, faces = _collapse(v, faces, pick)
    a = _heron_ref(v, faces)
    faces = faces[a >= min_area]
    used = np.unique(faces)
    remap = -np.ones(len(v), dtype=np.int64)
    remap[used] = np.arange(len(used))
    return v[used], remap[faces]


def _half_edges(faces):
    e = np.concatenate([faces[:, [0, 1]], faces[:, [1, 2]], faces[:, [2, 0]]])
    fid = np.tile(np.arange(len(faces)), 3)
    return e, fid


def repair_manifold(v, faces, rounds=8):
    """Make the mesh a closed, consistently oriented 2-manifold.  Edge collapses
    can pinch the surface (an edge shared by three faces, a directed edge used
    twice); the faces on such edges are removed, vertices where the remaining
    surface meets itself lose their faces too, and every hole left is closed
    by a fan around a new vertex at its boundary's centroid, oriented with the
    surrounding faces.  Faces away from defects are untouched."""
    v = np.asarray(v, dtype=np.float64)
    faces = np.asarray(faces, dtype=np.int64)
    for _ in range(rounds):
        nv = len(v)
        e, fid = _half_edges(faces)
        key = e[:, 0] * nv + e[:, 1]
        rkey = e[:, 1] * nv + e[:, 0]
        uk, inv, cnt = np.unique(key, return_inverse=True, return_counts=True)
        rep = cnt[inv] > 1
        has_twin = np.isin(rkey, uk)
        ukey = np.minimum(key, rkey)
        _, uinv, ucnt = np.unique(ukey, return_inverse=True, return_counts=True)
        bad_edge = rep | ~has_twin | (ucnt[uinv] != 2)
        if not np.any(bad_edge):
            used = np.unique(faces)
            remap = -np.ones(nv, dtype=np.int64)
            remap[used] = np.arange(len(used))
            return v[used], remap[faces]
        bad_f = np.zeros(len(faces), dtype=bool)
        bad_f[fid[rep | (ucnt[uinv] > 2)]] = True
        f = faces[~bad_f]
        # boundary half-edges; a vertex with more than one outgoing boundary
        # half-edge joins two holes (non-manifold): drop its faces and retry
        e, fid = _half_edges(f)
        key = e[:, 0] * nv + e[:, 1]
        bnd = ~np.isin(e[:, 1] * nv + e[:, 0], key)
        out_deg = np.bincount(e[bnd, 0], minlength=nv)
        pinched = out_deg > 1
        if np.any(pinched):
            f = f[~pinched[f].any(axis=1)]
            faces = f
            continue
        nxt = -np.ones(nv, dtype=np.int64)
        nxt[e[bnd, 0]] = e[bnd, 1]
        seen = np.zeros(nv, dtype=bool)
        newv, newf = [], []
        for start in e[bnd, 0]:
            if seen[start]:
                continue
            loop = [start]
            seen[start] = True
            cur = nxt[start]
            while cur != start and cur >= 0 and not seen[cur]:
                loop.append(cur)
                seen[cur] = True
                cur = nxt[cur]
            if cur != start:
                continue                      # not a closed loop: next round
            if len(loop) == 3:
                newf.append((loop[0], loop[2], loop[1]))
                continue
            c = len(v) + len(newv)
            newv.append(v[loop].mean(axis=0))
            for a, b in zip(loop, loop[1:] + loop[:1]):
                newf.append((b, a, c))
        if newv:
            v = np.vstack([v, np.array(newv)])
        faces = np.vstack([f, np.array(newf, dtype=np.int64).reshape(-1, 3)])
    raise RuntimeError("mesh repair did not converge")


def vertex_normals(verts, faces):
    fn = np.cross(verts[faces[:, 1]] - verts[faces[:, 0]], verts[faces[:, 2]] - verts[faces[:, 0]])
    vn = np.zeros_like(verts)
    for k in range(3):
        np.add.at(vn, faces[:, k], fn)
    nn = np.sqrt((vn * vn).sum(axis=1))
    nn[nn == 0] = 1.0
    return vn / nn[:, None]


def outer_component(v, faces):
    """Keep the connected component of largest enclosed volume (the outer
    surface; internal cavity surfaces are dropped, as MSMS does by default)."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components
    e = np.concatenate([faces[:, [0, 1]], faces[:, [1, 2]], faces[:, [2, 0]]])
    g = coo_matrix((np.ones(len(e)), (e[:, 0], e[:, 1])), shape=(len(v), len(v)))
    nc, lab = connected_components(g, directed=False)
    if nc == 1:
        return v, faces
    fl = lab[faces[:, 0]]
    six_vol = np.einsum("ij,ij->i", v[faces[:, 0]], np.cross(v[faces[:, 1]], v[faces[:, 2]]))
    vol = np.bincount(fl, weights=six_vol, minlength=nc)
    keep = faces[fl == np.argmax(vol)]
    used = np.unique(keep)
    remap = -np.ones(len(v), dtype=np.int64)
    remap[used] = np.arange(len(used))
    return v[used], remap[keep]


def ses_mesh(xyz, radius, probe=1.4, h=0.5, cavities=False):
    """Solvent-excluded surface triangulation: (verts, normals, faces 0-based).
    With cavities=False only the outer surface is kept."""
    f, lo, h = ses_field(xyz, radius, probe, h)
    v, fc = marching_tets(f, lo, h)
    v, fc = clean_mesh(v, fc, h)
    v, fc = repair_manifold(v, fc)
    if not cavities:
        v, fc = outer_component(v, fc)
    return v, vertex_normals(v, fc), fc
