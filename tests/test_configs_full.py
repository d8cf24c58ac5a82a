"""The overset-cylinder and batched-ensemble workloads at their full sizes,
on the GPU, through the C ABI.

Overset cylinder (bench.c3_inputs: background 2049 x 1025 with holes and
sponges, O-grid 513 x 129 with its seam and a wall-slip face, 4,626
receivers, potential-flow mean), 100 steps from the initial pulse:
  * the fused FP64 path against the plain FP64 path (the reference's
    arithmetic, pinned bit for bit to the reference on the small cylinder
    and the other fixtures) to 1e-12 relative to each field's max, in both
    blocks;
  * FP32 against fused FP64 to 1e-5 (p', relative to its max);
  * holes hold 0 and every value is finite.

Batched ensemble (bench.c5_inputs: 64 solves, 12.4M points), 20 steps:
  * in FP64 the batch equals single handles of the same solves bit for bit
    (four solves checked: the smallest, the largest and two others), in
    the tuner's launch form and in the persistent + graph form.
"""
import numpy as np
import pytest

from camille_amd import FP32, FP64, GRID_RESIDENT, PATH_FUSED, PATH_PLAIN, Batch, Solver

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def test_overset_cylinder_full_size():
    import bench
    inp = bench.c3_inputs()
    blocks = [dict(x=b["x"], y=b["y"], mean=b["mean"]) for b in inp["blocks"]]
    nsteps = 100
    sols = {}
    for label, prec, path in (("plain", FP64, PATH_PLAIN), ("fused64", FP64, PATH_FUSED), ("fused32", FP32, PATH_FUSED)):
        s = Solver(blocks, boundaries=inp["bc"], overset=inp["tables"], precision=prec, path=path, mode="FIX_CFL",
                   value=0.5)
        for ib, b in enumerate(inp["blocks"]):
            s.set_solution(b["ic"], ib)
        s.solve(nsteps)
        sols[label] = [s.solution(ib) for ib in range(len(blocks))]
        holes = [s.iblank(ib) == 1 for ib in range(len(blocks))]
        s.close()
        for ib, q in enumerate(sols[label]):
            assert np.all(np.isfinite(q)), (label, ib)
            assert holes[ib].sum() == 0 or np.all(q[:, holes[ib]] == 0.0), (label, ib)
    for ib in range(len(blocks)):
        ref = sols["plain"][ib]
        e64 = [_rel(sols["fused64"][ib][v], ref[v]) for v in (0, 1, 2, 4)]
        e32 = _rel(sols["fused32"][ib][4], sols["fused64"][ib][4])
        print(f"overset cylinder block {ib}: fused64 vs plain {['%.2e' % e for e in e64]}, "
              f"fused32 vs fused64 p' {e32:.2e}", flush=True)
        # the pulse starts in the background; after 100 steps its edge has
        # reached the O-grid
        assert np.max(np.abs(ref[4])) > (1e-6 if ib == 0 else 0.0)
        assert max(e64) <= 1e-12
        assert e32 <= 1e-5


def test_batched_ensemble_full_size_equals_single_handles():
    import bench
    solves = bench.c5_inputs()
    sizes = [sv["x"].size for sv in solves]
    pick = sorted({int(np.argmin(sizes)), int(np.argmax(sizes)), 7, 40})
    nsteps = 20
    singles = {}
    for k in pick:
        sv = solves[k]
        s = Solver([dict(x=sv["x"], y=sv["y"], mean=sv["mean"])], boundaries=sv["boundaries"], mode="FIX_CFL",
                   value=0.5, simulation="CASE_MONOPOLE", source_params=sv["source_params"], precision=FP64,
                   path=PATH_FUSED)
        s.set_solution(np.zeros((5,) + sv["x"].shape))
        s.solve(nsteps)
        singles[k] = s.solution()
        s.close()
        assert np.max(np.abs(singles[k][4])) > 0
    for form in ("tuned", "persistent"):
        b = Batch(solves, precision=FP64)
        for k, sv in enumerate(solves):
            b.set_solution(np.zeros((5,) + sv["x"].shape), k)
        if form == "persistent":
            b.set_launch(True, GRID_RESIDENT, GRID_RESIDENT)
        b.solve(nsteps)
        for k in pick:
            assert np.array_equal(b.solution(k), singles[k]), (form, k)
        print(f"batched ensemble ({form} form {b.launch()}): solves {pick} bit-identical to single handles",
              flush=True)
        b.close()
