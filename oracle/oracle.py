"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the checker.  The product path never imports it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

NODE_DTYPE = np.dtype([
    ("ibeg", np.int32), ("iend", np.int32), ("numpar", np.int32), ("level", np.int32),
    ("nchild", np.int32), ("child", np.int32, 8), ("_pad", np.int32),
    ("xmin", np.float64), ("xmax", np.float64), ("ymin", np.float64), ("ymax", np.float64),
    ("zmin", np.float64), ("zmax", np.float64), ("xmid", np.float64), ("ymid", np.float64),
    ("zmid", np.float64), ("radius", np.float64)])
assert NODE_DTYPE.itemsize == 136


class Params(C.Structure):
    _fields_ = [("eps", C.c_double), ("kappa", C.c_double), ("theta", C.c_double),
                ("order", C.c_int32), ("maxparnode", C.c_int32)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        _LIB = C.CDLL(path)
    return _LIB


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def kappa_of(bulk_strength, eps_w):
    f = lib().or_kappa
    f.restype = C.c_double
    return f(C.c_double(bulk_strength), C.c_double(eps_w))


def face_geometry(vpos, vnrm, faces):
    vpos, vnrm, faces = _d(vpos), _d(vnrm), _i(faces)
    n = faces.shape[0]
    xyz = np.zeros((n, 3)); nrm = np.zeros((n, 3)); area = np.zeros(n)
    lib().or_face_geometry(C.c_int32(n), _p(vpos), _p(vnrm), _p(faces), _p(xyz), _p(nrm), _p(area))
    return xyz, nrm, area


def cull_faces(vpos, faces):
    vpos, faces = _d(vpos), _i(faces)
    out = np.zeros_like(faces)
    sa = C.c_double(0.0)
    f = lib().or_cull_faces
    f.restype = C.c_int32
    k = f(C.c_int32(faces.shape[0]), _p(vpos), _p(faces), _p(out), C.byref(sa))
    return out[:k].copy(), sa.value


def rhs(xyz, nrm, qpos, q, eps_p):
    xyz, nrm, qpos, q = _d(xyz), _d(nrm), _d(qpos), _d(q)
    n = xyz.shape[0]
    b = np.zeros(2 * n)
    lib().or_rhs(C.c_int32(n), _p(xyz), _p(nrm), C.c_int32(q.shape[0]), _p(qpos), _p(q),
                 C.c_double(eps_p), _p(b))
    return b


def build_tree(xyz, maxparnode):
    """Returns (xyz in tree order, orderarr 0-based, nodes structured array)."""
    xyz = _d(xyz).copy()
    n = xyz.shape[0]
    order = np.zeros(n, dtype=np.int32)
    nodes = np.zeros(2 * n + 1, dtype=NODE_DTYPE)
    f = lib().or_build_tree
    f.restype = C.c_int32
    nn = f(C.c_int32(n), _p(xyz), C.c_int32(maxparnode), _p(order), _p(nodes))
    return xyz, order, nodes[:nn].copy()


def leaves(nodes):
    f = lib().or_leaves
    f.restype = C.c_int32
    nl = f(_p(nodes), None, None)
    ib = np.zeros(nl, dtype=np.int32); ie = np.zeros(nl, dtype=np.int32)
    f(_p(nodes), _p(ib), _p(ie))
    return ib, ie


def params(eps, kappa, theta, order, maxparnode):
    return Params(eps, kappa, theta, order, maxparnode)


def treecode_matvec(prm, xyz, nrm, area, nodes, x):
    xyz, nrm, area, x = _d(xyz), _d(nrm), _d(area), _d(x)
    n = xyz.shape[0]
    y = np.zeros(2 * n)
    cnt = np.zeros(3, dtype=np.int64)
    lib().or_treecode_matvec(C.byref(prm), C.c_int32(n), _p(xyz), _p(nrm), _p(area),
                             _p(nodes), _p(x), _p(y), _p(cnt))
    return y, cnt


def treecode_matvec_targets(prm, xyz, nrm, area, nodes, x, targets):
    xyz, nrm, area, x = _d(xyz), _d(nrm), _d(area), _d(x)
    targets = _i(targets)
    n = xyz.shape[0]
    y = np.zeros(2 * len(targets))
    cnt = np.zeros(3, dtype=np.int64)
    lib().or_treecode_matvec_targets(C.byref(prm), C.c_int32(n), _p(xyz), _p(nrm), _p(area),
                                     _p(nodes), _p(x), C.c_int32(len(targets)), _p(targets),
                                     _p(y), _p(cnt))
    return y.reshape(-1, 2), cnt


def direct_matvec(prm, xyz, nrm, area, x):
    xyz, nrm, area, x = _d(xyz), _d(nrm), _d(area), _d(x)
    n = xyz.shape[0]
    y = np.zeros(2 * n)
    lib().or_direct_matvec(C.byref(prm), C.c_int32(n), _p(xyz), _p(nrm), _p(area), _p(x), _p(y))
    return y


def precond(prm, xyz, nrm, area, ib, ie, r):
    xyz, nrm, area, r = _d(xyz), _d(nrm), _d(area), _d(r)
    ib, ie = _i(ib), _i(ie)
    n = xyz.shape[0]
    z = np.zeros(2 * n)
    f = lib().or_precond
    f.restype = C.c_int32
    bad = f(C.byref(prm), C.c_int32(n), _p(xyz), _p(nrm), _p(area), C.c_int32(len(ib)),
            _p(ib), _p(ie), _p(r), _p(z))
    return z, bad


def leaf_block(prm, xyz, nrm, area, ibeg, iend):
    xyz, nrm, area = _d(xyz), _d(nrm), _d(area)
    m2 = 2 * (iend - ibeg + 1)
    blk = np.zeros(m2 * m2)
    lib().or_leaf_block(C.byref(prm), _p(xyz), _p(nrm), _p(area), C.c_int32(ibeg),
                        C.c_int32(iend), _p(blk))
    return blk.reshape(m2, m2).T.copy()  # row-major view of the column-major block


def gmres(prm, xyz, nrm, area, nodes, b, x0=None, tol=1e-4, maxl=10, nrmax=10):
    xyz, nrm, area, b = _d(xyz), _d(nrm), _d(area), _d(b)
    n = xyz.shape[0]
    x = np.zeros(2 * n) if x0 is None else _d(x0).copy()
    it = C.c_int32(0)
    err = C.c_double(0.0)
    hist = np.zeros((nrmax + 1) * maxl + 1)
    f = lib().or_gmres
    f.restype = C.c_int32
    ierr = f(C.byref(prm), C.c_int32(n), _p(xyz), _p(nrm), _p(area), _p(nodes), _p(b), _p(x),
             C.c_double(tol), C.c_int32(maxl), C.c_int32(nrmax), C.byref(it), C.byref(err),
             _p(hist))
    return x, it.value, err.value, ierr, hist[:max(it.value, 0)]


def energy(xyz, nrm, area, x, qpos, q, eps, kappa, eps_p):
    xyz, nrm, area, x, qpos, q = _d(xyz), _d(nrm), _d(area), _d(x), _d(qpos), _d(q)
    cou = C.c_double(0.0)
    f = lib().or_energy
    f.restype = C.c_double
    s = f(C.c_int32(xyz.shape[0]), _p(xyz), _p(nrm), _p(area), _p(x), C.c_int32(q.shape[0]),
          _p(qpos), _p(q), C.c_double(eps), C.c_double(kappa), C.c_double(eps_p), C.byref(cou))
    return s, cou.value


def taylor_coeffs(d, kappa, lim):
    out = np.zeros((lim + 1) ** 3)
    lib().or_taylor_coeffs(C.c_double(d[0]), C.c_double(d[1]), C.c_double(d[2]),
                           C.c_double(kappa), C.c_int32(lim), _p(out))
    return out.reshape(lim + 1, lim + 1, lim + 1)


def direct_rows(prm, xyz, nrm, area, x, targets):
    """Rows of the O(n^2) direct operator for the listed targets, shape (k, 2)."""
    xyz, nrm, area, x = _d(xyz), _d(nrm), _d(area), _d(x)
    targets = _i(targets)
    y = np.zeros(2 * len(targets))
    lib().or_direct_rows(C.byref(prm), C.c_int32(xyz.shape[0]), _p(xyz), _p(nrm), _p(area),
                         _p(x), C.c_int32(len(targets)), _p(targets), _p(y))
    return y.reshape(-1, 2)


def diag_prec(eps, r):
    r = _d(r)
    z = np.zeros_like(r)
    lib().or_diag_prec(C.c_int32(r.size), C.c_double(eps), _p(r), _p(z))
    return z


def vertex_potential(vpos, faces, xyz, x):
    """output_potential's vertex values; faces 0-based and xyz, x in the
    original face order."""
    vpos, xyz, x = _d(vpos), _d(xyz), _d(x)
    f1 = _i(np.asarray(faces).reshape(-1, 3) + 1)
    out = np.zeros((vpos.shape[0], 2))
    f = lib().or_vertex_potential
    f.restype = C.c_int32
    rc = f(C.c_int32(f1.shape[0]), C.c_int32(vpos.shape[0]), _p(vpos), _p(f1), _p(xyz), _p(x),
           _p(out))
    if rc:
        raise ValueError("vertex index out of range")
    return out
