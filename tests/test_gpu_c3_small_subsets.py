"""Mixed precision on the batched workload's small second-order subsets, at
full size, against the oracle.

Pair 13 of the uneven batch is 4096^2 with radius 11, second order, step 7
(338724 POIs): the batched workload's largest small-subset item.  Where the
mixed and FP64 results part by more than 1e-3 px, the cause is the
problem, not the precision: those subsets are nearly degenerate, the
Gauss-Newton iteration barely contracts, and the reference's own FP64
result at its settings (||dp|| < 1e-3, 30 iterations) stops far from the
stationary point it is heading for.  The test pins that with the oracle:
  * on a uniform sample of the other POIs, FP64 meets the drop-in bar and
    mixed lies within 1e-3 px of the oracle wherever the oracle converges
    before its iteration limit;
  * every POI where the two precisions part (or their statuses differ) is
    one the oracle runs to its iteration limit, one where the oracle's
    iteration diverges and fails (there the FP64 status may differ from the
    oracle's only if the oracle itself reaches it from a guess 1e-9 px
    away: such an iteration is chaotic, and the order of FP64 roundings
    decides where it ends), or one whose oracle
    result lies more than 1e-3 px from the stationary point (the oracle run
    to ||dp|| < 1e-12); mixed lies no farther from that point than twice the
    oracle's own distance plus 1e-3 px.
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _solver(pkg, ctx, radius, order, precision):
    cfg = pkg.SubsetConfig(subset_radius=radius, max_iterations=30, convergence_threshold=1e-3,
                           shape_function="second_order" if order == 2 else "first_order")
    return pkg.ICGNSolver(cfg, precision=precision, context=ctx)


def test_small_second_order_subsets_at_full_size(pkg, ctx, orc, synth):
    t0 = time.time()
    c = synth.batched_pair(13)
    assert (c.reference.shape, c.radius, c.shape_order, c.step) == ((4096, 4096), 11, 2, 7)
    xy, ini = np.ascontiguousarray(c.poi_xy), c.init
    n = xy.shape[0]
    g64 = _solver(pkg, ctx, c.radius, 2, "fp64").solve_batch(c.reference, c.deformed, xy, ini)
    gmx = _solver(pkg, ctx, c.radius, 2, "mixed").solve_batch(c.reference, c.deformed, xy, ini)
    ok64 = (g64["status"] == 0) & (g64["valid"] == 1)
    d = np.maximum(np.abs(gmx["p"][:, 0] - g64["p"][:, 0]), np.abs(gmx["p"][:, 1] - g64["p"][:, 1]))
    split = np.nonzero((ok64 & (d > 1e-3)) | (gmx["status"] != g64["status"]))[0]
    hit64 = int((ok64 & (g64["iterations"] == 30)).sum())
    hitmx = int(((gmx["status"] == 0) & (gmx["iterations"] == 30)).sum())
    print(f"pair 13: {n} POIs, FP64 Success {int(ok64.sum())}, hit 30 iterations FP64 {hit64} mixed {hitmx}, "
          f"split {split.size} (render + solve {time.time() - t0:.1f} s)", flush=True)
    assert split.size < 1e-4 * n

    pre = orc.Precomputed(c.deformed, degree=5)
    gx, gy = orc.reference_gradient_cache(c.reference)

    # uniform sample of the other POIs
    sample = np.setdiff1d(np.arange(0, n, 211), split)
    cpu = orc.icgn_solve(c.reference, pre, gx, gy, xy[sample], ini[sample], radius=c.radius, max_iterations=30,
                         threshold=1e-3, shape_order=2)
    conv = (cpu["status"] == 0) & (cpu["iterations"] < 30)
    assert np.array_equal(g64["status"][sample], cpu["status"])
    d64 = np.abs(g64["p"][sample][conv, :2] - cpu["p"][conv, :2]).max()
    dmx = np.abs(gmx["p"][sample][conv, :2] - cpu["p"][conv, :2]).max()
    print(f"sample of {sample.size}: oracle converges on {int(conv.sum())}; FP64 {d64:.1e} px, mixed {dmx:.1e} px",
          flush=True)
    assert conv.sum() > 0.95 * sample.size and d64 <= 1e-5 and dmx <= 1e-3

    # the POIs where the precisions part: the reference's own stopping error
    at = orc.icgn_solve(c.reference, pre, gx, gy, xy[split], ini[split], radius=c.radius, max_iterations=30,
                        threshold=1e-3, shape_order=2)
    tight = orc.icgn_solve(c.reference, pre, gx, gy, xy[split], ini[split], radius=c.radius, max_iterations=2000,
                           threshold=1e-12, shape_order=2)
    for k, i in enumerate(split):
        ran_out = at["iterations"][k] == 30
        if ran_out:
            continue  # no tolerance holds for an iteration that never met its own test
        if g64["status"][i] != at["status"][k]:
            # a chaotic POI: the oracle's own status changes when its guess
            # moves by 1e-9 px; FP64 must reach a status the oracle reaches
            flips = set()
            for du, dv in ((1e-9, 0.0), (-1e-9, 0.0), (0.0, 1e-9), (0.0, -1e-9)):
                pi = ini[[i]].copy()
                pi["u"] += du
                pi["v"] += dv
                flips.add(int(orc.icgn_solve(c.reference, pre, gx, gy, xy[[i]], pi, radius=c.radius,
                                             max_iterations=30, threshold=1e-3, shape_order=2)["status"][0]))
            print(f"  POI {i}: FP64 status {g64['status'][i]}, oracle {at['status'][k]} after {at['iterations'][k]} "
                  f"iterations; oracle statuses with the guess moved by 1e-9 px: {sorted(flips)}", flush=True)
            assert int(g64["status"][i]) in flips, i
            continue
        assert g64["status"][i] == at["status"][k], i  # FP64 is the oracle there too
        if at["status"][k] != 0:
            # a diverging iteration (it leaves the image or its deformed norm
            # vanishes): where and whether it stops depends on rounding, so
            # the mixed status is recorded, not held; such a POI counts as
            # converged in neither precision's favour
            print(f"  POI {i}: oracle fails after {at['iterations'][k]} iterations with status {at['status'][k]}; "
                  f"mixed status {gmx['status'][i]} after {gmx['iterations'][i]} (ZNSSD {gmx['correlation'][i]:.3g})",
                  flush=True)
            continue
        assert gmx["status"][i] == 0, i
        assert np.abs(g64["p"][i, :2] - at["p"][k, :2]).max() <= 1e-5, i
        fixed = tight["p"][k, :2]
        e_ref = np.abs(at["p"][k, :2] - fixed).max()
        e_mx = np.abs(gmx["p"][i, :2] - fixed).max()
        print(f"  POI {i}: oracle stops after {at['iterations'][k]} iterations {e_ref:.2e} px from the stationary "
              f"point ({tight['iterations'][k]} iterations); mixed {e_mx:.2e} px", flush=True)
        assert e_ref > 1e-3, i
        assert e_mx <= 2.0 * e_ref + 1e-3, i
