"""Generate the golden fixtures by running the reference solver binary.

Run from the repository root, in the build container only (the reference
sources and the MPICH it links exist only there):

    make -f oracle/ref.mk
    python tests/golden/make_golden.py [case ...]

For each case of tests/golden/cases.py this writes the reference's input
files (deck, PLOT3D grid / base state / restart IC, BC file) into a scratch
directory, runs `oracle/_ref/camille ic.q` there, and stores the outputs the
parity tests need as tests/golden/<case>.npz:

  sol      final solution (nv, nj, ni) as written by the reference
  tau      its PLOT3D tau header {0, 1, step, time}
  met      metrics file camille.met.q (11, nj, ni), when requested
  dts      dt printed at each report step (text, 6 digits; informative only)

Only data (inputs are regenerated from cases.py, outputs stored) is kept.
"""
import glob
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
sys.path.insert(0, ROOT)

import cases  # noqa: E402
import refio  # noqa: E402

REF_BIN = os.path.join(ROOT, "oracle", "_ref", "camille")
MPIEXEC = "/opt/conda/bin/mpiexec"
KEEP_METRICS = {"curv", "filt", "wall"}


def run_case(case, nranks=1):
    with tempfile.TemporaryDirectory(prefix=f"golden_{case['name']}_") as d:
        x, y = case["x"], case["y"]
        z = np.zeros_like(x)
        refio.write_grid(os.path.join(d, "grid.xyz"), [(x[None], y[None], z[None])])
        refio.write_function(os.path.join(d, "mean.q"), [case["mean"][:, None]])
        refio.write_solution(os.path.join(d, "ic.q"), [case["ic"][:, None]], 0, 0.0)
        refio.write_bc(os.path.join(d, "bc.in"), case["bc"])
        refio.write_deck(os.path.join(d, "camille.in"), cases.deck_items(case))
        env = dict(os.environ)
        env["LD_LIBRARY_PATH"] = "/usr/lib/x86_64-linux-gnu:/opt/conda/lib"
        cmd = [MPIEXEC, "-n", str(nranks), REF_BIN, "ic.q"]
        p = subprocess.run(cmd, cwd=d, env=env, capture_output=True, text=True, timeout=3600)
        out = p.stdout
        if ">>> " in out and "graceful" in out.lower():
            raise RuntimeError(out[-2000:])
        files = sorted(glob.glob(os.path.join(d, "camille.[0-9]*.q")))
        if len(files) < 2:
            raise RuntimeError(f"reference wrote no final solution:\n{out[-3000:]}\n{p.stderr[-2000:]}")
        if case.get("num_scalar"):
            # more than 5 variables: the reference writes the solution in the
            # PLOT3D function format, which has no tau (io.cpp:740-761); the
            # time is kept as NaN and checked against the report line
            sol = refio.read_function(files[-1])[0][:, 0]
            res = dict(sol=sol, tau=np.array([0.0, 1.0, float(case["nsteps"]), np.nan]))
        else:
            sols, taus = refio.read_solution(files[-1])
            res = dict(sol=sols[0][:, 0], tau=taus[0])
        if case["name"] in KEEP_METRICS:
            res["met"] = refio.read_function(os.path.join(d, "camille.met.q"))[0][:, 0]
        dts = [ln for ln in out.splitlines() if ln.startswith("Time step:")]
        res["log"] = "\n".join(dts)
        return res


def run_overset_case(case, nranks=2, report_every=None):
    """Two-block overset problem: one MPI rank per block (decomp.cpp:29-30).
    With report_every, the reference's min/max report file (io.cpp:440-497,
    ./report/solution/solution) is returned as a fourth value."""
    import overset_cases
    with tempfile.TemporaryDirectory(prefix=f"golden_{case['name']}_") as d:
        blocks = case["blocks"]
        refio.write_grid(os.path.join(d, "grid.xyz"),
                         [(b["x"][None], b["y"][None], np.zeros_like(b["x"])[None]) for b in blocks])
        refio.write_function(os.path.join(d, "mean.q"), [b["mean"][:, None] for b in blocks])
        refio.write_solution(os.path.join(d, "ic.q"), [b["ic"][:, None] for b in blocks], 0, 0.0)
        refio.write_bc(os.path.join(d, "bc.in"), case["bc"])
        dims = [(b["x"].shape[1], b["x"].shape[0], 1) for b in blocks]
        if "geo" in case:
            from camille_amd import overset_gen
            overset_gen.write_overture(os.path.join(d, "overset.bin"), case["geo"])
        else:
            refio.write_overture(os.path.join(d, "overset.bin"), 2, dims, case["recv"], case["widths"],
                                 case["holes"])
        items = cases.deck_items(case)
        items = [(k, v) for k, v in items if k != "OVERSET"]
        if report_every:
            items = [(k, str(report_every) if k == "REPORT_TIMESTEP" else v) for k, v in items]
        items.insert(12, ("OVERSET", f"OVERTURE, TYPE = EXPLICIT, ORDER_ACCURACY = {case['order']}"))
        items.append(("OVERSET_FILE", "overset.bin"))
        refio.write_deck(os.path.join(d, "camille.in"), items)
        env = dict(os.environ)
        env["LD_LIBRARY_PATH"] = "/usr/lib/x86_64-linux-gnu:/opt/conda/lib"
        p = subprocess.run([MPIEXEC, "-n", str(nranks), REF_BIN, "ic.q"], cwd=d, env=env, capture_output=True,
                           text=True, timeout=3600)
        files = sorted(glob.glob(os.path.join(d, "camille.[0-9]*.q")))
        if len(files) < 2:
            raise RuntimeError(f"reference wrote no final solution:\n{p.stdout[-3000:]}\n{p.stderr[-2000:]}")
        sols, taus = refio.read_solution(files[-1])
        log = "\n".join(ln for ln in p.stdout.splitlines() if ln.startswith("Time step:"))
        if report_every:
            return sols, taus, log, open(os.path.join(d, "report", "solution", "solution")).read()
        return sols, taus, log


def pressure_series(report):
    """(step, min p', max p') per line of the reference's min/max report."""
    out = []
    for ln in report.splitlines():
        if not ln.startswith("Time step:"):
            continue
        step = int(ln.split(",")[0].split(":")[1])
        seg = ln.split("P': ")[1].split(",")[0]
        lo, hi = (float(t) for t in seg.split("~"))
        out.append((step, lo, hi))
    return out


def cylinder_long(nsteps=1500, every=100):
    """The reference's own long run of the small cylinder case: min/max p'
    every `every` steps (the wall mode's growth), stored as
    tests/golden/cylinder_long.json."""
    import overset_cases
    case = overset_cases.case_cylinder_small()
    case["nsteps"] = nsteps
    _, taus, log, report = run_overset_case(case, report_every=every)
    series = pressure_series(report)
    out = dict(case="cylinder_small", nsteps=nsteps, report_every=every, time=float(taus[0][3]), ranks=2,
               p_min_max=[[st, lo, hi] for st, lo, hi in series])
    json.dump(out, open(os.path.join(HERE, "cylinder_long.json"), "w"), indent=1)
    print("cylinder_long:", series[-1])


def main(names):
    import overset_cases
    if names and names[0] == "cylinder_long":
        cylinder_long()
        return
    overset = {"overset_small": overset_cases.case_overset_small,
               "cylinder_small": overset_cases.case_cylinder_small}
    if names and names[0] in overset:
        name = names[0]
        case = overset[name]()
        sols, taus, log = run_overset_case(case)
        arrs = {f"sol{b}": sols[b][:, 0] for b in range(len(sols))}
        arrs["tau"] = taus[0]
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **arrs)
        mpath = os.path.join(HERE, "manifest.json")
        old = json.load(open(mpath)) if os.path.exists(mpath) else {}
        if "geo" in case:
            nrecv, nholes = len(case["geo"]["recv"]["ri"]), len(case["geo"]["holes"]["i"])
        else:
            nrecv, nholes = len(case["recv"]), len(case["holes"])
        old[name] = dict(nsteps=case["nsteps"], blocks=[list(b["x"].shape) for b in case["blocks"]],
                         nrecv=nrecv, nholes=nholes, time=float(taus[0][3]), report=log, ranks=2)
        json.dump(old, open(mpath, "w"), indent=1, sort_keys=True)
        print(name, "time", taus[0][3], "->", os.path.getsize(path), "bytes")
        return
    names = names or list(cases.ALL)
    manifest = {}
    for name in names:
        case = cases.ALL[name]()
        res = run_case(case)
        path = os.path.join(HERE, f"{name}.npz")
        arrs = dict(tau=res["tau"])
        sol = res["sol"]
        # keep only the variables that are not identically zero
        nz = [v for v in range(sol.shape[0]) if np.any(sol[v] != 0.0)]
        arrs["vars"] = np.array(nz, dtype=np.int32)
        arrs["sol"] = sol[nz]
        if "met" in res:
            arrs["met"] = res["met"]
        np.savez_compressed(path, **arrs)
        manifest[name] = dict(nsteps=case["nsteps"], ni=case["ni"], nj=case["nj"], mode=case["mode"],
                              cfl=case["cfl"], dt=case["dt"], simulation=case["simulation"],
                              time=None if np.isnan(res["tau"][3]) else float(res["tau"][3]),
                              step=int(res["tau"][2]), nonzero_vars=nz, num_scalar=case.get("num_scalar", 0),
                              report=res["log"])
        print(name, "time", res["tau"][3], "vars", nz, "->", os.path.getsize(path), "bytes", flush=True)
    mpath = os.path.join(HERE, "manifest.json")
    old = json.load(open(mpath)) if os.path.exists(mpath) else {}
    old.update(manifest)
    json.dump(old, open(mpath, "w"), indent=1, sort_keys=True)


if __name__ == "__main__":
    main(sys.argv[1:])
