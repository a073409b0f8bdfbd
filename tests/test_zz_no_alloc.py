"""The allocation-free energy entry point (tabi_ctx_energy_noalloc) against
tabi_ctx_energy on the same context and solution: bit-identical energies (the
same kernels in the same order, the scratch in the library's static device
array instead of pool memory), on two calls in a row, the second smaller, and
a clear error for a problem whose scratch exceeds the static array."""
import numpy as np
import pytest

import cases
from tabipb import driver, hip

pytestmark = pytest.mark.gpu


def _solved(name):
    v, nv, f, atoms, usr = cases.build(name)
    f2, _ = driver.cull_faces(v, f)
    xyz, nrm, area = driver.face_geometry(v, nv, f2)
    a32 = atoms.astype(np.float32).astype(np.float64)
    P = driver.Problem(usr, xyz, nrm, area, a32[:, :3], a32[:, 3])
    S = hip.Surface(P.params(), xyz, nrm, area)
    x, it, err, ierr = S.gmres(S.rhs(P.qpos, P.q, P.usr.epsp), tol=1e-6)
    assert ierr == 0
    return P, S, x


def test_noalloc_energy_equals_pool_energy_on_two_calls():
    for name in ("mol150_restart", "sph3_l100_p3"):     # the second call is smaller
        print(f"[gpu] no-alloc energy {name}", flush=True)
        P, S, x = _solved(name)
        ref = S.energy(P.qpos, P.q, P.usr.epsp, x)
        dq, dqp, dx = (hip.DeviceArray.from_host(a) for a in (P.q, P.qpos.ravel(), x))
        got = S.energy_noalloc(dqp.ptr, dq.ptr, P.q.shape[0], P.usr.epsp, dx.ptr)
        assert got == ref
        # again on the same context: the static scratch keeps the last call's
        # contents, which nothing reads
        assert S.energy_noalloc(dqp.ptr, dq.ptr, P.q.shape[0], P.usr.epsp, dx.ptr) == ref
        for a in (dq, dqp, dx):
            a.free()
        S.close()


def test_noalloc_energy_rejects_oversized_problems():
    P, S, x = _solved("sph3_l100_p3")
    # 1,280 faces, 500,000 charges: 500,032 energy partials, 500,000 potentials
    # and 1,954^2 Coulomb partials = 4.82 M doubles, beyond the 4.19 M (32 MiB)
    nq = 500000
    dq = hip.DeviceArray(nq)
    dqp = hip.DeviceArray(3 * nq)
    dx = hip.DeviceArray.from_host(x)
    with pytest.raises(hip.TabiError, match="static"):
        S.energy_noalloc(dqp.ptr, dq.ptr, nq, 1.0, dx.ptr)
    for a in (dq, dqp, dx):
        a.free()
    S.close()
