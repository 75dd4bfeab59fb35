"""Host-side input checks of the Python solver surface (no GPU needed: every
case returns or raises before a device call, which the stand-in context
below would turn into an error)."""
import numpy as np
import pytest


class _NoDevice:
    """Context stand-in: any device call fails the test."""

    def __getattr__(self, name):
        raise AssertionError(f"device call {name} reached")


def _solver(pkg, **kw):
    cfg = pkg.SubsetConfig(subset_radius=5, **kw)
    return pkg.ICGNSolver(cfg, context=_NoDevice())


def _case(pkg):
    ref = np.zeros((40, 50), np.float32)
    xy = np.array([[20.0, 20.0], [25.0, 18.0]])
    ini = pkg.initial_guesses([0.5, -1.0], [0.25, 2.0], confidence=0.125)
    return ref, xy, ini


def _assert_invalid(pkg, out, xy, ini):
    """icgn.cpp:455-470, :639-652: InvalidInput, initial u, v, correlation = confidence."""
    assert (out["status"] == pkg.INVALID_INPUT).all() and not out["valid"].any()
    assert np.array_equal(out["x"], xy[:, 0]) and np.array_equal(out["y"], xy[:, 1])
    assert np.array_equal(out["p"][:, 0], ini["u"]) and np.array_equal(out["p"][:, 1], ini["v"])
    assert np.array_equal(out["correlation"], ini["confidence"])


@pytest.mark.parametrize("roi_shape", [(40, 49), (50, 40), (39, 50), (40 * 50,)])
def test_masked_solve_with_roi_of_another_shape_is_invalid_input(pkg, roi_shape):
    ref, xy, ini = _case(pkg)
    s = _solver(pkg)
    out = s.solve_batch(ref, ref, xy, ini, roi=np.ones(roi_shape, np.uint8))
    _assert_invalid(pkg, out, xy, ini)
    one = s.solve_with_mask(ref, ref, np.ones(roi_shape, bool), xy[0], ini[0])
    assert one["status"] == pkg.INVALID_INPUT and one["p"][0] == ini["u"][0]


def test_pair_of_different_sizes_is_invalid_input(pkg):
    ref, xy, ini = _case(pkg)
    out = _solver(pkg).solve_batch(ref, np.zeros((40, 51), np.float32), xy, ini)
    _assert_invalid(pkg, out, xy, ini)


@pytest.mark.parametrize("case", ["roi_shape", "transposed_roi", "pair_sizes", "no_seeds", "radius", "spacing"])
def test_propagate_returns_the_empty_result_on_bad_input(pkg, case):
    """reliability_propagation.cpp:140-146: an empty result, no device work."""
    ref = np.zeros((40, 50), np.float32)
    dfm, roi = ref, np.ones((40, 50), np.uint8)
    seeds = np.zeros(1, pkg.SEED_DTYPE)
    kw = {}
    if case == "roi_shape":
        roi = np.ones((40, 49), np.uint8)
    elif case == "transposed_roi":
        roi = np.ones((50, 40), np.uint8)
    elif case == "pair_sizes":
        dfm = np.zeros((41, 50), np.float32)
    elif case == "no_seeds":
        seeds = seeds[:0]
    elif case == "radius":
        kw["subset_radius"] = 0
    else:
        kw["propagation_spacing"] = -1
    cfg = pkg.SubsetConfig(**{"subset_radius": 5, **kw})
    solver = pkg.ICGNSolver(cfg, context=_NoDevice())
    d, grid, n = pkg.propagate(ref, dfm, roi, seeds, solver=solver)
    assert n == 0 and grid.size == 0 and all(v.size == 0 for v in d.values())


def test_context_wrappers_refuse_an_roi_of_another_shape(pkg):
    """The C-ABI reads W * H bytes from the ROI pointer, so the ctypes layer
    checks the shape before any call."""
    from traditional_dic_amd import _lib

    class Img:
        width, height = 50, 40

    _lib._checked_roi(np.ones((40, 50), bool), Img())
    for shape in ((40, 49), (50, 40), (2000,), (40, 50, 1)):
        with pytest.raises(ValueError, match="roi has shape"):
            _lib._checked_roi(np.ones(shape, np.uint8), Img())


def test_solver_keeps_no_state_between_calls(pkg):
    """No implicit pair cache (ids of freed host arrays are reused by
    CPython): besides its config, precision and context the solver holds only
    the device buffers of host-image calls, keyed by image size and the
    precompute configuration and refilled on every call."""
    s = _solver(pkg)
    assert set(vars(s)) == {"config", "precision", "gradient_mode", "ctx", "_ws", "_ws_lock"}
    assert s._ws is None
