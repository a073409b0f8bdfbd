"""Times the reference program and its drop-in build on a golden case
(test infrastructure; nothing here is shipped).

    python oracle/time_dropin.py CASE [--exe hip|ref] [--dir DIR] [--stage-only]

Stages the case of tests/golden/cases.py as a run directory of the reference
(usrdata.in, PQR, surface files) and runs oracle/_ref/tabipb_hip.exe -- the
reference's src/main.f90 with the GPU hook (oracle/Makefile `dropin`) -- or
oracle/_ref/tabipb_ref.exe, the unmodified reference, with an unlimited stack,
then prints the wall time of the whole run and the program's own lines for the
iteration count, the energy and its cpu_time splits (src/main.f90:139,158,
164-167).  --stage-only writes the directory and exits (for profiling the
executable under rocprofv3 from inside it).
"""
import argparse
import os
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
sys.path.insert(0, os.path.join(ROOT, "tabi-pb_amd"))
import cases  # noqa: E402

EXES = {"hip": os.path.join(HERE, "_ref", "tabipb_hip.exe"),
        "ref": os.path.join(HERE, "_ref", "tabipb_ref.exe")}
KEYS = ("err=", "Solvation energy", "setup cpu", "solving cpu", "computing solvation",
        "Total cpu", "GPU set-up")


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("case")
    ap.add_argument("--exe", choices=sorted(EXES), default="hip")
    ap.add_argument("--dir", default=None)
    ap.add_argument("--stage-only", action="store_true")
    a = ap.parse_args(argv)
    d = a.dir or tempfile.mkdtemp(prefix="tabi_run_")
    t0 = time.perf_counter()
    cases.stage(a.case, d)
    print(f"[time_dropin] staged {a.case} in {d} ({time.perf_counter() - t0:.1f} s)", flush=True)
    if a.stage_only:
        return 0
    t0 = time.perf_counter()
    r = subprocess.run(["bash", "-c", f"ulimit -s unlimited; exec {EXES[a.exe]}"], cwd=d,
                       capture_output=True, text=True)
    wall = time.perf_counter() - t0
    for ln in r.stdout.splitlines():
        if any(k in ln for k in KEYS):
            print("   ", ln.strip(), flush=True)
    print(f"[time_dropin] {a.case} {a.exe}: exit {r.returncode}, wall {wall:.3f} s", flush=True)
    if r.returncode != 0:
        print(r.stdout[-3000:] + r.stderr[-3000:])
    return r.returncode


if __name__ == "__main__":
    sys.exit(main())
