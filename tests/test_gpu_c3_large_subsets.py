"""The batched workload's largest 4096^2 items at full size, both precisions,
against the oracle on a uniform sample of their POIs.

Pair 49 of the uneven batch is 4096^2 with radius 30, first order, step 6
(the batch's largest subsets on its largest image); pair 36 is 4096^2 with
radius 19, second order, step 12 (its largest second-order subsets there).
On the sampled POIs:
  * FP64 has the oracle's statuses and, where the oracle succeeds, u and v
    within 1e-5 px of it (the drop-in bar);
  * mixed has the oracle's statuses and, where the oracle converges before
    its iteration limit, u and v within 1e-3 px of it.
Over all POIs, mixed and FP64 agree in status, and where FP64 converges
before the limit, mixed lies within 1e-3 px of it.
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_IT, THR = 30, 1e-3


def _solver(pkg, ctx, radius, order, precision):
    cfg = pkg.SubsetConfig(subset_radius=radius, max_iterations=MAX_IT, convergence_threshold=THR,
                           shape_function="second_order" if order == 2 else "first_order")
    return pkg.ICGNSolver(cfg, precision=precision, context=ctx)


@pytest.mark.parametrize("j, expect", [(49, ((4096, 4096), 30, 1, 6)), (36, ((4096, 4096), 19, 2, 12))])
def test_largest_subsets_at_full_size(pkg, ctx, orc, synth, j, expect):
    t0 = time.time()
    c = synth.batched_pair(j)
    assert (c.reference.shape, c.radius, c.shape_order, c.step) == expect
    xy, ini = np.ascontiguousarray(c.poi_xy), c.init
    n = xy.shape[0]
    g64 = _solver(pkg, ctx, c.radius, c.shape_order, "fp64").solve_batch(c.reference, c.deformed, xy, ini)
    gmx = _solver(pkg, ctx, c.radius, c.shape_order, "mixed").solve_batch(c.reference, c.deformed, xy, ini)
    print(f"pair {j}: {n} POIs, FP64 Success {int((g64['status'] == 0).sum())} "
          f"(render + solves {time.time() - t0:.1f} s)", flush=True)

    # all POIs: the two precisions
    assert np.array_equal(gmx["status"], g64["status"])
    conv64 = (g64["status"] == 0) & (g64["iterations"] < MAX_IT)
    dmx = np.abs(gmx["p"][conv64, :2] - g64["p"][conv64, :2]).max()
    print(f"  mixed vs FP64 over {int(conv64.sum())} POIs converged before the limit: {dmx:.1e} px", flush=True)
    assert conv64.sum() > 0.95 * n and dmx <= 1e-3

    # a uniform sample: the oracle
    t1 = time.time()
    sample = np.arange(0, n, max(1, n // 600))
    pre = orc.Precomputed(c.deformed, degree=5)
    gx, gy = orc.reference_gradient_cache(c.reference)
    cpu = orc.icgn_solve(c.reference, pre, gx, gy, xy[sample], ini[sample], radius=c.radius,
                         max_iterations=MAX_IT, threshold=THR, shape_order=c.shape_order)
    assert np.array_equal(g64["status"][sample], cpu["status"])
    assert np.array_equal(gmx["status"][sample], cpu["status"])
    ok = cpu["status"] == 0
    conv = ok & (cpu["iterations"] < MAX_IT)
    d64 = np.abs(g64["p"][sample][ok, :2] - cpu["p"][ok, :2]).max()
    dmx = np.abs(gmx["p"][sample][conv, :2] - cpu["p"][conv, :2]).max()
    print(f"  sample of {sample.size} (oracle {time.time() - t1:.1f} s): FP64 {d64:.1e} px, "
          f"mixed {dmx:.1e} px over {int(conv.sum())} converged", flush=True)
    assert conv.sum() > 0.95 * sample.size and d64 <= 1e-5 and dmx <= 1e-3
