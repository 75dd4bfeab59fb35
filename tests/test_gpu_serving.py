"""Serving loop (dic_server_*): continuous batching of incremental frame
sequences with HBM-budget admission.

Each POI's result depends only on its own inputs, so a request's results
must be bit-identical served alone and served in a mix of requests that
arrive while it runs.  Frame f's guesses are frame f-K's results (K the
guess lag; frames below K start from the request's guesses) or, with
extrapolation, 2 r(f-K) - r(f-2K) from frame 2K on, so the served
FP64 results are checked against the same chain of direct
dic_icgn_solve_batch calls (identical statuses, |du|, |dv| <= 1e-9 px) and
frame 0 against the CPU oracle (|du|, |dv| <= 1e-5 px).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _report(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def seqs(synth, pkg):
    out = []
    for k, (w, h, nf) in enumerate([(256, 256, 5), (320, 256, 3), (256, 256, 4), (288, 320, 6)]):
        ref, frames = synth.strain_sequence(w, h, nf, seed=4100 + k)
        xy = synth.strain_pois(w, h)
        ini = np.zeros(xy.shape[0], pkg.INITIAL_DTYPE)
        out.append((ref, frames, xy, ini))
    return out


def _guesses_from(pkg, r):
    g = np.zeros(r.shape[0], pkg.INITIAL_DTYPE)
    for k, name in enumerate(("u", "v", "du_dx", "du_dy", "dv_dx", "dv_dy")):
        g[name] = r["p"][:, k]
    g["confidence"] = r["correlation"]
    return g


def _extrapolated(pkg, a, b):
    """Guesses 2 a - b (u, v, gradients) where a and b are both Success, a's
    values otherwise; confidence = a's correlation."""
    g = _guesses_from(pkg, a)
    lin = (a["status"] == 0) & (b["status"] == 0)
    for k, name in enumerate(("u", "v", "du_dx", "du_dy", "dv_dx", "dv_dy")):
        g[name] = np.where(lin, 2.0 * a["p"][:, k] - b["p"][:, k], a["p"][:, k])
    return g


def _direct_chain(pkg, ctx, seq, precision, optimizer="icgn", objective="znssd", order=1, lag=1,
                  extrapolate=False):
    """Frame f solved directly from frame f - lag's results (the request's
    guesses below lag), or from the extrapolation of frames f - lag and
    f - 2 lag where f >= 2 lag."""
    ref, frames, xy, ini = seq
    cfg = pkg.SubsetConfig(subset_radius=15, optimizer=optimizer, objective=objective,
                           shape_function="second_order" if order == 2 else "first_order")
    params = pkg.make_params(cfg, precision)
    dref = ctx.upload_image(ref)
    grad = ctx.ref_gradient(dref, 0, 5, 3, precision)
    res = []
    for f, frame in enumerate(frames):
        co = ctx.prefilter(ctx.upload_image(frame), 5, 3, True, precision)
        if f < lag:
            cur = ini
        elif extrapolate and f >= 2 * lag:
            cur = _extrapolated(pkg, res[f - lag], res[f - 2 * lag])
        else:
            cur = _guesses_from(pkg, res[f - lag])
        res.append(ctx.solve(params, dref, grad, co, xy, cur))
    return np.stack(res)


@pytest.mark.parametrize("lag,extrapolate", [(1, False), (4, False), (1, True), (2, True)])
def test_served_fp64_matches_direct_chain_and_oracle(pkg, ctx, orc, seqs, lag, extrapolate):
    srv = ctx.server(radius=15, shape_order=1, degree=5, precision=pkg.PRECISION_FP64, max_active=8,
                     max_width=512, max_height=512, guess_lag=lag, guess_extrapolate=extrapolate)
    req = srv.submit(*seqs[0])
    got = req.wait(60.0)
    st = srv.stats()
    srv.close()
    n_frames = len(seqs[0][1])
    assert st["ticks"] == -(-n_frames // lag)  # K frames of the one request share a tick
    want = _direct_chain(pkg, ctx, seqs[0], pkg.PRECISION_FP64, lag=lag, extrapolate=extrapolate)
    assert np.array_equal(got["status"], want["status"])
    ok = want["status"] == 0
    d = np.abs(got["p"][..., :2] - want["p"][..., :2])[ok].max(initial=0)
    _report(f"served (guess lag {lag}, extrapolate {extrapolate}) vs direct chain: {ok.mean():.4f} ok, "
            f"max |du| {d:.2e}, {st['ticks']} ticks, mean iterations {got['iterations'].mean():.2f}")
    assert d <= 1e-9 and ok.mean() > 0.98
    ref, frames, xy, ini = seqs[0]
    pre = orc.Precomputed(frames[0], degree=5)
    gx, gy = orc.reference_gradient_cache(ref)
    cpu = orc.icgn_solve(ref, pre, gx, gy, xy, ini, radius=15, max_iterations=30, threshold=1e-3, shape_order=1)
    assert np.array_equal(got["status"][0], cpu["status"])
    okc = cpu["status"] == 0
    assert np.abs(got["p"][0, okc, :2] - cpu["p"][okc, :2]).max(initial=0) <= 1e-5
    # strain truth: u = f * 2e-4 * (x - xc)
    xc = 0.5 * (ref.shape[1] - 1)
    for f in range(len(frames)):
        u = (f + 1) * 2e-4 * (xy[:, 0] - xc)
        assert np.abs(got["p"][f, ok[f], 0] - u[ok[f]]).max() < 0.05


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_served_fgn_matches_direct_chain(pkg, ctx, seqs, precision):
    """A server configured with the forward Gauss-Newton optimizer: each frame
    of a request equals the chain of direct FGN solves (identical statuses,
    |du|, |dv| <= 1e-9 px)."""
    prec = pkg.PRECISION_FP64 if precision == "fp64" else pkg.PRECISION_MIXED
    srv = ctx.server(radius=15, shape_order=1, degree=5, precision=prec, max_active=8, max_width=512,
                     max_height=512, optimizer=pkg._lib.OPTIMIZER_FGN)
    req = srv.submit(*seqs[1])
    got = req.wait(60.0)
    srv.close()
    want = _direct_chain(pkg, ctx, seqs[1], prec, optimizer="forward_gauss_newton")
    assert np.array_equal(got["status"], want["status"])
    ok = want["status"] == 0
    d = np.abs(got["p"][..., :2] - want["p"][..., :2])[ok].max(initial=0)
    _report(f"served FGN {precision} vs direct chain: {ok.mean():.4f} ok, max |du| {d:.2e}")
    assert d <= 1e-9 and ok.mean() > 0.98


@pytest.mark.parametrize("objective,order", [("ssd", 1), ("znssd", 2), ("ssd", 2)])
def test_served_other_configs_match_direct_chain(pkg, ctx, seqs, objective, order):
    """Servers configured for the SSD criterion and the second-order shape
    function (mixed precision): each frame equals the chain of direct solves
    of the same configuration bit for bit."""
    obj = pkg._lib.OBJECTIVE_SSD if objective == "ssd" else pkg._lib.OBJECTIVE_ZNSSD
    srv = ctx.server(radius=15, shape_order=order, degree=5, precision=pkg.PRECISION_MIXED, objective=obj,
                     max_active=8, max_width=512, max_height=512)
    got = srv.submit(*seqs[2]).wait(60.0)
    srv.close()
    want = _direct_chain(pkg, ctx, seqs[2], pkg.PRECISION_MIXED, objective=objective, order=order)
    ok = want["status"] == 0
    _report(f"served {objective} order {order}: {ok.mean():.4f} ok, same bits {got.tobytes() == want.tobytes()}")
    assert ok.mean() > 0.95
    assert got.tobytes() == want.tobytes()


@pytest.mark.parametrize("lag,extrapolate", [(1, False), (4, False), (1, True), (4, True)])
def test_alone_and_in_mix_are_identical(pkg, ctx, seqs, lag, extrapolate):
    srv = ctx.server(radius=15, shape_order=1, degree=5, precision=pkg.PRECISION_MIXED, max_active=8,
                     max_width=512, max_height=512, guess_lag=lag, guess_extrapolate=extrapolate)
    alone = srv.submit(*seqs[0]).wait(60.0).copy()
    # the same request again, in a mix: others are submitted before and after it
    r1 = srv.submit(*seqs[1])
    r2 = srv.submit(*seqs[2])
    again = srv.submit(*seqs[0])
    r3 = srv.submit(*seqs[3])
    for r in (r1, r2, again, r3):
        r.wait(60.0)
    st = srv.stats()
    _report(f"mix stats: {st}")
    assert st["max_active_seen"] >= 2 and st["completed"] == 5
    assert again.results.tobytes() == alone.tobytes()
    assert np.all(np.isfinite(again.frame_done_s)) and np.all(np.diff(again.frame_done_s) >= 0)
    srv.close()


def _slab_bytes(w, h, n, elem=4, lag=1):
    """Device bytes a request holds: reference, two gradient planes, POIs
    and the request's guesses, and 3K frame, coefficient, guess and result
    buffers (three ticks in flight, K = the guess lag); each array rounded up
    to 256 bytes."""
    def a(v):
        return (v + 255) // 256 * 256
    px, cpx = w * h, (w + 6) * (h + 6)
    return (a(4 * px) + 2 * a(elem * px) + a(16 * n) + a(56 * n)
            + 3 * lag * (a(4 * px) + a(elem * cpx) + a(56 * n) + a(136 * n)))


def test_budget_admission_and_refusal(pkg, ctx, seqs):
    # room for one 256 x 256 request at a time: requests are served one after another
    assert seqs[0][2].shape[0] == seqs[2][2].shape[0]
    one = _slab_bytes(256, 256, seqs[0][2].shape[0])
    srv = ctx.server(radius=15, precision=pkg.PRECISION_MIXED, max_active=8, max_width=512, max_height=512,
                     hbm_budget=one + 64 * 1024)
    reqs = [srv.submit(*seqs[0]), srv.submit(*seqs[2])]
    big = srv.submit(*seqs[3])  # 288 x 320: never fits
    for r in reqs:
        r.wait(60.0)
    with pytest.raises(pkg.DicError):
        big.wait(60.0)
    st = srv.stats()
    _report(f"budget stats: {st}")
    assert st["max_active_seen"] == 1 and st["peak_bytes_held"] <= one + 64 * 1024
    assert st["completed"] == 2
    srv.close()


def test_server_rejects_bad_config(pkg, ctx):
    with pytest.raises(pkg.DicError):
        ctx.server(shape_order=3)
    with pytest.raises(pkg.DicError):
        ctx.server(max_active=0)
    for lag in (-1, 17):
        with pytest.raises(pkg.DicError, match="guess_lag"):
            ctx.server(guess_lag=lag)
