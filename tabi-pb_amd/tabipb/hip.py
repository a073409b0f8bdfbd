"""ctypes binding of libtabi_hip.so (the C ABI declared in include/tabi_hip.h).

The library is the only compute path: if it is missing or fails to load, every
entry point raises; nothing falls back to a CPU implementation.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.normpath(os.path.join(_HERE, "..", "lib", "libtabi_hip.so"))
_LIB = None

# every symbol include/tabi_hip.h declares
EXPORTS = (
    "tabi_init", "tabi_matvec", "tabi_msolve", "tabi_rhs", "tabi_energy", "tabi_gmres",
    "tabi_finalize", "tabi_last_error", "tabi_stats", "tabi_ctx_create", "tabi_ctx_destroy",
    "tabi_ctx_stats", "tabi_ctx_matvec_dev", "tabi_ctx_msolve_dev", "tabi_ctx_matvec",
    "tabi_ctx_msolve", "tabi_ctx_rhs", "tabi_ctx_energy", "tabi_ctx_gmres", "tabi_solve",
    "tabi_last_solve_times", "tabi_ctx_profile", "tabi_ctx_kernel_times", "tabi_ctx_sync",
    "tabi_device_alloc", "tabi_device_free", "tabi_release_memory", "tabi_copy_to_device", "tabi_copy_to_host",
    "tabi_ws_param_count", "tabi_ws_create", "tabi_ws_destroy", "tabi_ws_get_params",
    "tabi_ctx_ws_features", "tabi_ws_predict", "tabi_ws_grad", "tabi_ws_train",
    "tabi_ctx_gmres_ws", "tabi_solve_ws", "tabi_ctx_matvec_block", "tabi_ctx_msolve_block",
    "tabi_ctx_gmres_block", "tabi_ctx_solve_block", "tabi_solve_many", "tabi_ctx_create_packed",
    "tabi_last_solve_many_setup",
    "tabi_last_solve_many_stats", "tabi_solve_dev", "tabi_vertex_potential",
    "tabi_ctx_vertex_potential", "tabi_write_surface_potential", "tabi_ctx_energy_noalloc",
)

# tabi_params.flags bits
FLAG_MATVEC_L23 = 1       # preconditioner with matvec-consistent L2/L3 placement
FLAG_HOST_TREE = 2        # sequential host tree build
FLAG_DIAG_PREC = 4        # diagonal scaling MSOLVEprec (usrdata.in iprcd != 1)
FLAG_GRAPH_GMRES = 16     # GMRES cycles as captured graphs with the control on the device

SOLVE_PHASES = ("tree", "lists", "precond", "rhs", "gmres", "energy", "total", "warmstart")

KERNEL_SLOTS = ("moments", "near", "far", "reduce", "msolve", "lists", "lu_factor", "rhs_energy")

STAT_NAMES = ("nface", "nnode", "nleaf", "p_near", "p_far", "p_mom", "n_far_items",
              "n_near_items", "n_tasks", "lu_bytes")


class TabiError(RuntimeError):
    pass


class Params(C.Structure):
    _fields_ = [("eps", C.c_double), ("kappa", C.c_double), ("theta", C.c_double),
                ("order", C.c_int32), ("maxparnode", C.c_int32), ("precision", C.c_int32),
                ("flags", C.c_int32)]


def load():
    """Load the library (once).  Raises TabiError if it is not built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise TabiError(f"{LIB_PATH} is not built; run __graft_entry__.build() "
                        "(make -C tabi-pb_amd/csrc)")
    lib = C.CDLL(LIB_PATH)
    lib.tabi_last_error.restype = C.c_char_p
    lib.tabi_finalize.restype = None
    lib.tabi_ctx_destroy.restype = None
    lib.tabi_ws_destroy.restype = None
    lib.tabi_ws_param_count.restype = C.c_int64
    _LIB = lib
    return lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _d(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


def _check(rc):
    if rc != 0:
        raise TabiError(load().tabi_last_error().decode())


class DeviceArray:
    """A float64 device buffer owned through the library (no other GPU runtime
    is involved).  ``ptr`` is the device address for the *_dev entry points."""

    def __init__(self, n):
        self.n = int(n)
        p = C.c_void_p()
        _check(load().tabi_device_alloc(C.c_int64(8 * self.n), C.byref(p)))
        self.ptr = p.value

    @classmethod
    def from_host(cls, a):
        a = _d(a).ravel()
        d = cls(a.size)
        d.upload(a)
        return d

    def upload(self, a):
        a = _d(a).ravel()
        assert a.size == self.n
        _check(load().tabi_copy_to_device(C.c_void_p(self.ptr), _p(a), C.c_int64(8 * self.n)))

    def get(self):
        out = np.empty(self.n)
        _check(load().tabi_copy_to_host(_p(out), C.c_void_p(self.ptr), C.c_int64(8 * self.n)))
        return out

    def free(self):
        if self.ptr:
            load().tabi_device_free(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def make_params(eps, kappa, theta, order, maxparnode, precision=0, flags=0):
    return Params(float(eps), float(kappa), float(theta), int(order), int(maxparnode),
                  int(precision), int(flags))


class Surface:
    """One surface on the GPU: tree, interaction lists and factored leaf blocks.

    Mirrors the reference's set-up + callback pair (treecode_initialization,
    MATVECmain, MSOLVEdric; src/main.f90:329-529).  Vectors are tree-ordered
    [phi; dphi/dn] of length 2n.
    """

    def __init__(self, prm: Params, xyz, nrm, area):
        lib = load()
        xyz = _d(xyz, (-1, 3))
        nrm = _d(nrm, (-1, 3))
        area = _d(area)
        n = xyz.shape[0]
        if nrm.shape[0] != n or area.shape[0] != n:
            raise TabiError("xyz, nrm and area disagree in length")
        self.n = n
        self.prm = prm
        self.order = np.zeros(n, dtype=np.int32)
        h = C.c_void_p()
        _check(lib.tabi_ctx_create(C.byref(prm), C.c_int32(n), _p(xyz), _p(nrm), _p(area),
                                   _p(self.order), C.byref(h)))
        self._h = h
        self.orderarr = self.order.copy()          # 1-based, as the reference's orderarr
        self.perm = self.order.astype(np.int64) - 1  # tree position -> original index

    def close(self):
        if getattr(self, "_h", None):
            load().tabi_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def stats(self) -> dict:
        out = np.zeros(10, dtype=np.int64)
        _check(load().tabi_ctx_stats(self._h, _p(out)))
        return dict(zip(STAT_NAMES, (int(v) for v in out)))

    def to_tree(self, v):
        """Permute an original-order length-n or length-2n vector into tree order."""
        v = np.asarray(v)
        if v.shape[0] == self.n:
            return v[self.perm]
        return np.concatenate([v[: self.n][self.perm], v[self.n:][self.perm]])

    def from_tree(self, v):
        v = np.asarray(v)
        out = np.empty_like(v)
        if v.shape[0] == self.n:
            out[self.perm] = v
            return out
        out[: self.n][self.perm] = v[: self.n]
        out[self.n:][self.perm] = v[self.n:]
        return out

    def profile(self, enable=True):
        _check(load().tabi_ctx_profile(self._h, C.c_int32(1 if enable else 0)))

    def kernel_times(self):
        ms = np.zeros(8)
        cnt = np.zeros(8, dtype=np.int64)
        _check(load().tabi_ctx_kernel_times(self._h, _p(ms), _p(cnt)))
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(KERNEL_SLOTS)}

    def _len(self, v):
        if v.shape != (2 * self.n,):
            raise TabiError(f"vector length {v.size} does not match 2*nface = {2 * self.n}")

    def matvec(self, x):
        x = _d(x)
        self._len(x)
        y = np.empty_like(x)
        _check(load().tabi_ctx_matvec(self._h, _p(x), _p(y)))
        return y

    def msolve(self, r):
        r = _d(r)
        self._len(r)
        z = np.empty_like(r)
        _check(load().tabi_ctx_msolve(self._h, _p(r), _p(z)))
        return z

    def matvec_dev(self, d_x: int, d_y: int):
        """Device-pointer matvec (addresses as integers, e.g. DeviceArray.ptr)."""
        _check(load().tabi_ctx_matvec_dev(self._h, C.c_void_p(d_x), C.c_void_p(d_y)))

    def msolve_dev(self, d_r: int, d_z: int):
        _check(load().tabi_ctx_msolve_dev(self._h, C.c_void_p(d_r), C.c_void_p(d_z)))

    def sync(self):
        """Wait for work enqueued by the device-pointer entry points."""
        _check(load().tabi_ctx_sync(self._h))

    def rhs(self, qpos, q, eps_p):
        qpos = _d(qpos, (-1, 3))
        q = _d(q)
        b = np.zeros(2 * self.n)
        _check(load().tabi_ctx_rhs(self._h, C.c_int32(q.shape[0]), _p(qpos), _p(q),
                                   C.c_double(eps_p), _p(b)))
        return b

    def energy(self, qpos, q, eps_p, x):
        qpos = _d(qpos, (-1, 3))
        q = _d(q)
        x = _d(x)
        s = C.c_double()
        cc = C.c_double()
        _check(load().tabi_ctx_energy(self._h, C.c_int32(q.shape[0]), _p(qpos), _p(q),
                                      C.c_double(eps_p), _p(x), C.byref(s), C.byref(cc)))
        return s.value, cc.value

    def energy_noalloc(self, d_qpos: int, d_q: int, nq: int, eps_p, d_x: int):
        """tabi_ctx_energy_noalloc on device arrays (DeviceArray.ptr values)."""
        s = C.c_double()
        cc = C.c_double()
        _check(load().tabi_ctx_energy_noalloc(self._h, C.c_int32(nq), C.c_void_p(d_qpos),
                                              C.c_void_p(d_q), C.c_double(eps_p),
                                              C.c_void_p(d_x), C.byref(s), C.byref(cc)))
        return s.value, cc.value

    def vertex_potential(self, vpos, faces, x):
        """output_potential's vertex values (src/main.f90:226-326): (nvert, 2)
        array of (phi, dphi/dn) scaled by para*4*pi; faces 0-based, original
        order (the culled face list); x tree order."""
        vpos = _d(vpos, (-1, 3))
        f1 = np.ascontiguousarray(np.asarray(faces).reshape(-1, 3) + 1, dtype=np.int32)
        if f1.shape[0] != self.n:
            raise TabiError(f"{f1.shape[0]} faces for a surface of {self.n}")
        x = _d(x)
        out = np.empty((vpos.shape[0], 2))
        _check(load().tabi_ctx_vertex_potential(self._h, C.c_int32(vpos.shape[0]), _p(vpos),
                                                _p(f1), _p(x), _p(out)))
        return out

    def gmres(self, b, x0=None, tol=1e-4, maxl=10, nrmax=10):
        b = _d(b)
        x = np.zeros_like(b) if x0 is None else _d(x0).copy()
        it = C.c_int32()
        err = C.c_double()
        ierr = C.c_int32()
        _check(load().tabi_ctx_gmres(self._h, _p(b), _p(x), C.c_double(tol), C.c_int32(maxl),
                                     C.c_int32(nrmax), C.byref(it), C.byref(err), C.byref(ierr)))
        return x, it.value, err.value, ierr.value

    def _blk(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        if x.ndim != 2 or x.shape[1] != 2 * self.n:
            raise TabiError(f"block must have shape (nrhs, {2 * self.n})")
        return x

    def matvec_block(self, x):
        """Y[k] = A X[k] for a block X of shape (nrhs, 2n) in one tree pass."""
        x = self._blk(x)
        y = np.empty_like(x)
        _check(load().tabi_ctx_matvec_block(self._h, C.c_int32(x.shape[0]), _p(x), _p(y)))
        return y

    def msolve_block(self, r):
        r = self._blk(r)
        z = np.empty_like(r)
        _check(load().tabi_ctx_msolve_block(self._h, C.c_int32(r.shape[0]), _p(r), _p(z)))
        return z

    def gmres_block(self, b, x0=None, tol=1e-4, maxl=10, nrmax=10):
        b = self._blk(b)
        x = np.zeros_like(b) if x0 is None else self._blk(x0).copy()
        k = b.shape[0]
        it, ierr = np.zeros(k, dtype=np.int32), np.zeros(k, dtype=np.int32)
        err = np.zeros(k)
        _check(load().tabi_ctx_gmres_block(self._h, C.c_int32(k), _p(b), _p(x), C.c_double(tol),
                                           C.c_int32(maxl), C.c_int32(nrmax), _p(it), _p(err),
                                           _p(ierr)))
        return x, it, err, ierr

    def solve_block(self, qpos, q, eps_p, tol=1e-4, want_x=True):
        """Solves for the charge sets q (nrhs, nq) at common positions qpos:
        returns dict(x, iter, err, ierr, soleng, couleng) with per-set arrays."""
        qpos = _d(qpos, (-1, 3))
        q = np.ascontiguousarray(q, dtype=np.float64)
        if q.ndim != 2 or q.shape[1] != qpos.shape[0]:
            raise TabiError("charges must have shape (nrhs, nq)")
        k, nq = q.shape
        x = np.zeros((k, 2 * self.n)) if want_x else None
        it, ierr = np.zeros(k, dtype=np.int32), np.zeros(k, dtype=np.int32)
        err, sol, cou = np.zeros(k), np.zeros(k), np.zeros(k)
        _check(load().tabi_ctx_solve_block(self._h, C.c_int32(k), C.c_int32(nq), _p(qpos), _p(q),
                                           C.c_double(eps_p), C.c_double(tol),
                                           _p(x) if want_x else None, _p(it), _p(err), _p(ierr),
                                           _p(sol), _p(cou)))
        return dict(x=x, iter=it, err=err, ierr=ierr, soleng=sol, couleng=cou)

    def ws_features(self, qpos, q, eps_p):
        """Warm-start features [n, 12] (tree order) for these charges and the
        charge scale s (tabi_ctx_ws_features)."""
        qpos = _d(qpos, (-1, 3))
        q = _d(q)
        feat = np.zeros((self.n, 12), dtype=np.float32)
        sc = C.c_double()
        _check(load().tabi_ctx_ws_features(self._h, C.c_int32(q.shape[0]), _p(qpos), _p(q),
                                           C.c_double(eps_p), _p(feat), C.byref(sc)))
        return feat, sc.value

    def gmres_ws(self, model: "WsModel", qpos, q, eps_p, b, tol=1e-4, maxl=10, nrmax=10):
        """GMRES from the model's predicted guess (tabi_ctx_gmres_ws); returns
        (x, iter, err, ierr, used) where used is 0 if the zero start was taken."""
        qpos = _d(qpos, (-1, 3))
        q = _d(q)
        b = _d(b)
        self._len(b)
        x = np.zeros_like(b)
        it, err, ierr, used = C.c_int32(), C.c_double(), C.c_int32(), C.c_int32()
        _check(load().tabi_ctx_gmres_ws(self._h, model._h, C.c_int32(qpos.shape[0]), _p(qpos),
                                        _p(q), C.c_double(eps_p),
                                        _p(b), _p(x), C.c_double(tol), C.c_int32(maxl),
                                        C.c_int32(nrmax), C.byref(it), C.byref(err),
                                        C.byref(ierr), C.byref(used)))
        return x, it.value, err.value, ierr.value, used.value


class WsModel:
    """The learned initial guess on the GPU (tabi_ws_*): packed FP32 master
    parameters, BF16 images for the MFMA kernels, Adam state."""

    def __init__(self, params):
        lib = load()
        params = np.ascontiguousarray(params, dtype=np.float32).ravel()
        h = C.c_void_p()
        _check(lib.tabi_ws_create(_p(params), C.c_int64(params.size), C.byref(h)))
        self._h = h

    @staticmethod
    def param_count() -> int:
        return int(load().tabi_ws_param_count())

    def close(self):
        if getattr(self, "_h", None):
            load().tabi_ws_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        out = np.zeros(self.param_count(), dtype=np.float32)
        _check(load().tabi_ws_get_params(self._h, _p(out), C.c_int64(out.size)))
        return out

    @staticmethod
    def _feat(feat):
        feat = np.ascontiguousarray(feat, dtype=np.float32)
        if feat.ndim != 2 or feat.shape[1] != 12:
            raise TabiError("features must have shape (n, 12)")
        return feat

    def predict(self, feat, scale):
        feat = self._feat(feat)
        x0 = np.zeros(2 * feat.shape[0])
        _check(load().tabi_ws_predict(self._h, C.c_int64(feat.shape[0]), _p(feat),
                                      C.c_double(scale), _p(x0)))
        return x0

    def grad(self, feat, target):
        feat = self._feat(feat)
        target = np.ascontiguousarray(target, dtype=np.float32).reshape(-1, 2)
        if target.shape[0] != feat.shape[0]:
            raise TabiError("features and targets disagree in length")
        g = np.zeros(self.param_count(), dtype=np.float32)
        loss = C.c_double()
        _check(load().tabi_ws_grad(self._h, C.c_int64(feat.shape[0]), _p(feat), _p(target),
                                   _p(g), C.byref(loss)))
        return g, loss.value

    def train(self, feat, target, epochs, batch=8192, lr=1e-3, seed=0):
        feat = self._feat(feat)
        target = np.ascontiguousarray(target, dtype=np.float32).reshape(-1, 2)
        if target.shape[0] != feat.shape[0]:
            raise TabiError("features and targets disagree in length")
        losses = np.zeros(int(epochs), dtype=np.float32)
        _check(load().tabi_ws_train(self._h, C.c_int64(feat.shape[0]), _p(feat), _p(target),
                                    C.c_int32(epochs), C.c_int32(batch), C.c_float(lr),
                                    C.c_uint64(seed), _p(losses)))
        return losses


# ---- default-context drop-in API (what a Fortran or C driver binds) --------
def write_surface_potential(path, vpos, vnrm, vpot, faces):
    """surface_potential.dat through the library's writer (faces 0-based)."""
    vpos, vnrm, vpot = _d(vpos, (-1, 3)), _d(vnrm, (-1, 3)), _d(vpot, (-1, 2))
    f1 = np.ascontiguousarray(np.asarray(faces).reshape(-1, 3) + 1, dtype=np.int32)
    _check(load().tabi_write_surface_potential(os.fsencode(path), C.c_int32(vpos.shape[0]),
                                               _p(vpos), _p(vnrm), _p(vpot),
                                               C.c_int32(f1.shape[0]), _p(f1)))


def dropin_init(prm: Params, xyz, nrm, area):
    """tabi_init: returns (orderarr 1-based, nleaf)."""
    xyz, nrm, area = _d(xyz, (-1, 3)), _d(nrm, (-1, 3)), _d(area)
    n = xyz.shape[0]
    order = np.zeros(n, dtype=np.int32)
    nl = C.c_int32()
    _check(load().tabi_init(C.byref(prm), C.c_int32(n), _p(xyz), _p(nrm), _p(area),
                            _p(order), C.byref(nl)))
    return order, nl.value


def dropin_matvec(x):
    """tabi_matvec(n2 = len(x), x, y): the MATVEC callback."""
    x = _d(x)
    y = np.empty_like(x)
    _check(load().tabi_matvec(C.c_int32(x.size), _p(x), _p(y)))
    return y


def dropin_msolve(r):
    """tabi_msolve(n2 = len(r), r, z): the MSOLVE callback."""
    r = _d(r)
    z = np.empty_like(r)
    _check(load().tabi_msolve(C.c_int32(r.size), _p(r), _p(z)))
    return z


def dropin_finalize():
    load().tabi_finalize()


def solve(prm: Params, xyz, nrm, area, qpos, q, eps_p, tol=1e-4, model: WsModel = None):
    """One full PB solve on the GPU (tabi_solve, or tabi_solve_ws with a learned
    initial guess).  Returns a dict with the tree-order solution, permutation,
    GMRES counters, energies and phase times."""
    lib = load()
    xyz = _d(xyz, (-1, 3))
    nrm = _d(nrm, (-1, 3))
    area = _d(area)
    qpos = _d(qpos, (-1, 3))
    q = _d(q)
    n = xyz.shape[0]
    x = np.zeros(2 * n)
    order = np.zeros(n, dtype=np.int32)
    it = C.c_int32()
    err = C.c_double()
    ierr = C.c_int32()
    sol = C.c_double()
    cou = C.c_double()
    used = C.c_int32()
    _check(lib.tabi_solve_ws(C.byref(prm), C.c_int32(n), _p(xyz), _p(nrm), _p(area),
                             C.c_int32(q.shape[0]), _p(qpos), _p(q), C.c_double(eps_p),
                             C.c_double(tol), model._h if model is not None else None, _p(x),
                             _p(order), C.byref(it), C.byref(err), C.byref(ierr),
                             C.byref(sol), C.byref(cou), C.byref(used)))
    t = np.zeros(8)
    lib.tabi_last_solve_times(_p(t))
    return dict(x=x, orderarr=order, iter=it.value, err=err.value, ierr=ierr.value,
                soleng=sol.value, couleng=cou.value, ws_used=used.value,
                times_ms=dict(zip(SOLVE_PHASES, t)))


def solve_many(prm: Params, surfaces, eps_p, tol=1e-4, graph=True, packed=True):
    """Solves several surfaces together (tabi_solve_many).  surfaces: sequence
    of (xyz, nrm, area, qpos, q) in original face order.  packed (default): one
    context over all surfaces; otherwise one context per surface with a graph
    branch each (graph=False: launched eagerly).  Returns a dict of per-surface
    arrays (iter, err, ierr, soleng, couleng) and stats."""
    lib = load()
    k = len(surfaces)
    keep = []
    ptr = {name: (C.c_void_p * k)() for name in ("xyz", "nrm", "area", "qpos", "q")}
    nface = np.zeros(k, dtype=np.int32)
    nq = np.zeros(k, dtype=np.int32)
    for i, (xyz, nrm, area, qpos, q) in enumerate(surfaces):
        arrs = dict(xyz=_d(xyz, (-1, 3)), nrm=_d(nrm, (-1, 3)), area=_d(area),
                    qpos=_d(qpos, (-1, 3)), q=_d(q))
        if arrs["nrm"].shape[0] != arrs["xyz"].shape[0] or arrs["area"].shape[0] != arrs["xyz"].shape[0]:
            raise TabiError(f"surface {i}: xyz, nrm and area disagree in length")
        if arrs["q"].shape[0] != arrs["qpos"].shape[0]:
            raise TabiError(f"surface {i}: qpos and q disagree in length")
        keep.append(arrs)
        nface[i] = arrs["xyz"].shape[0]
        nq[i] = arrs["q"].shape[0]
        for name, a in arrs.items():
            ptr[name][i] = a.ctypes.data
    it, ierr = np.zeros(k, dtype=np.int32), np.zeros(k, dtype=np.int32)
    err, sol, cou = np.zeros(k), np.zeros(k), np.zeros(k)
    _check(lib.tabi_solve_many(C.byref(prm), C.c_int32(k), _p(nface), ptr["xyz"], ptr["nrm"],
                               ptr["area"], _p(nq), ptr["qpos"], ptr["q"], C.c_double(eps_p),
                               C.c_double(tol),
                               C.c_int32((0 if graph else 1) | (0 if packed else 2)), _p(it), _p(err),
                               _p(ierr), _p(sol), _p(cou)))
    st = np.zeros(6)
    lib.tabi_last_solve_many_stats(_p(st))
    stats = dict(zip(("captures", "graph_launches", "setup_ms", "gmres_ms", "energy_ms", "total_ms"),
                     (float(v) for v in st)))
    if packed:
        su = np.zeros(4)
        lib.tabi_last_solve_many_setup(_p(su))
        stats.update(zip(("tree_ms", "lists_ms", "precond_ms", "rhs_ms"), (float(v) for v in su)))
    return dict(iter=it, err=err, ierr=ierr, soleng=sol, couleng=cou, stats=stats)


class PackedSurfaces(Surface):
    """Several surfaces in one context (tabi_ctx_create_packed): surface k's
    faces sit at face0[k] .. face0[k] + n_k of a common tree-order numbering of
    length n (padding between surfaces); matvec / msolve act on packed vectors
    [first halves; second halves] of length 2n.  pack() / unpack() convert
    from / to per-surface tree-order vectors [phi_k; dphi_k/dn]."""

    def __init__(self, prm: Params, surfaces):
        lib = load()
        k = len(surfaces)
        keep = []
        ptr = {name: (C.c_void_p * k)() for name in ("xyz", "nrm", "area")}
        nface = np.zeros(k, dtype=np.int32)
        for i, (xyz, nrm, area) in enumerate(surfaces):
            arrs = dict(xyz=_d(xyz, (-1, 3)), nrm=_d(nrm, (-1, 3)), area=_d(area))
            keep.append(arrs)
            nface[i] = arrs["xyz"].shape[0]
            for name, a in arrs.items():
                ptr[name][i] = a.ctypes.data
        self.face0 = np.zeros(k + 1, dtype=np.int32)
        self.nface = nface
        self.prm = prm
        h = C.c_void_p()
        _check(lib.tabi_ctx_create_packed(C.byref(prm), C.c_int32(k), _p(nface), ptr["xyz"],
                                          ptr["nrm"], ptr["area"], _p(self.face0), C.byref(h)))
        self._h = h
        self.n = int(self.face0[-1])

    def pack(self, vecs):
        out = np.zeros(2 * self.n)
        for k, v in enumerate(vecs):
            f0, nk = int(self.face0[k]), int(self.nface[k])
            out[f0:f0 + nk] = v[:nk]
            out[self.n + f0:self.n + f0 + nk] = v[nk:]
        return out

    def unpack(self, v):
        return [np.concatenate([v[f0:f0 + nk], v[self.n + f0:self.n + f0 + nk]])
                for f0, nk in zip(self.face0[:-1].tolist(), self.nface.tolist())]


class DeviceProblem:
    """A surface and its charges held in device memory (original face order),
    for solves whose inputs are already resident (tabi_solve_dev)."""

    def __init__(self, xyz, nrm, area, qpos, q):
        self.n = int(_d(xyz, (-1, 3)).shape[0])
        self.nq = int(_d(q).shape[0])
        self.xyz = DeviceArray.from_host(_d(xyz, (-1, 3)))
        self.nrm = DeviceArray.from_host(_d(nrm, (-1, 3)))
        self.area = DeviceArray.from_host(_d(area))
        self.qpos = DeviceArray.from_host(_d(qpos, (-1, 3)))
        self.q = DeviceArray.from_host(_d(q))
        self.x = DeviceArray(2 * self.n)

    def solve(self, prm: Params, eps_p, tol=1e-4, model: WsModel = None):
        """One solve; the tree-order solution stays on the device (self.x)."""
        lib = load()
        it, err, ierr = C.c_int32(), C.c_double(), C.c_int32()
        sol, cou, used = C.c_double(), C.c_double(), C.c_int32()
        _check(lib.tabi_solve_dev(C.byref(prm), C.c_int32(self.n), C.c_void_p(self.xyz.ptr),
                                  C.c_void_p(self.nrm.ptr), C.c_void_p(self.area.ptr),
                                  C.c_int32(self.nq), C.c_void_p(self.qpos.ptr),
                                  C.c_void_p(self.q.ptr), C.c_double(eps_p), C.c_double(tol),
                                  model._h if model is not None else None,
                                  C.c_void_p(self.x.ptr), None, C.byref(it), C.byref(err),
                                  C.byref(ierr), C.byref(sol), C.byref(cou), C.byref(used)))
        t = np.zeros(8)
        lib.tabi_last_solve_times(_p(t))
        return dict(iter=it.value, err=err.value, ierr=ierr.value, soleng=sol.value,
                    couleng=cou.value, ws_used=used.value, times_ms=dict(zip(SOLVE_PHASES, t)))

    def free(self):
        for a in (self.xyz, self.nrm, self.area, self.qpos, self.q, self.x):
            a.free()
