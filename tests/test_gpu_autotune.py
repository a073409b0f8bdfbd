"""Autotuner: it returns the fastest candidate within the energy error target,
and the task granularity it sweeps changes only the grouping of partial sums
(FP64 energies agree to 1e-10, iterations are equal)."""
import numpy as np
import pytest

from tabipb import autotune, driver, hip, meshio
from tabipb import warmstart as WS

pytestmark = pytest.mark.gpu


def _problem():
    m = WS.molecule(55, 150, h=0.6)
    usr = meshio.UsrData(maxparnode=500, order=3, mac=0.8)
    return driver.Problem(usr, m["xyz"], m["nrm"], m["area"], m["qpos"], m["q"])


def test_autotune_picks_fastest_candidate_within_error():
    prob = _problem()
    print(f"[gpu] autotune faces={prob.xyz.shape[0]}", flush=True)
    r = autotune.autotune(prob, leaves=(100, 300, 1000), orders=(2, 3, 6), task_exponents=(0, 4),
                          repeats=1, log=lambda s: print(s, flush=True))
    ok = [t for t in r["table"] if t["ok"]]
    assert r["best"] in ok
    assert r["best"]["ms"] == min(t["ms"] for t in ok)
    assert all(t["rel_err"] <= 1e-3 for t in ok)
    assert len(r["table"]) == 18


def test_task_granularity_only_regroups_sums():
    prob = _problem()
    for leaf, p in ((100, 2), (300, 3)):
        res = []
        for t in (0, 2, 6):
            prm = prob.params(0, flags=t << 8, order=p, maxparnode=leaf)
            res.append(hip.solve(prm, prob.xyz, prob.nrm, prob.area, prob.qpos, prob.q,
                                 prob.usr.epsp, 1e-4))
        for r in res[1:]:
            assert r["iter"] == res[0]["iter"]
            assert r["soleng"] == pytest.approx(res[0]["soleng"], rel=1e-10)
