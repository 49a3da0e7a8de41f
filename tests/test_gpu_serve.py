"""GPU tests of the serving loop (geye_serve_*): several camera streams of different sizes
and two resident models, frames overlapping on the device.  Every result must be identical
to the same frame detected alone (geye_detect on an idle context), whatever else is in
flight, including when a small HBM budget forces working memory to be evicted and rebuilt."""
import numpy as np
import pytest

import geye_amd
from geye_amd import synth

pytestmark = pytest.mark.gpu

STREAMS = [(320, 240), (640, 480), (800, 600), (1280, 720)]


def say(*a):
    print(*a, flush=True)


def _alone(ctx, models, frames, precision):
    out = []
    for mi, f in frames:
        out.append(models[mi].car_detection(f, thresh=-0.5, overlap=0.5, precision=precision))
    return out


def _frames(n_per_stream, seed0=100):
    frames = []
    for k in range(n_per_stream):
        for s, (W, H) in enumerate(STREAMS):
            frames.append(((k + s) % 2, synth.make_frame(W, H, seed=seed0 + 17 * k + s)))
    return frames


@pytest.mark.parametrize("precision", ["bf16", "fp32"])
def test_results_identical_alone_and_in_the_mix(ctx, precision):
    say("serve mix", precision)
    models = [ctx.model(synth.synthetic_car()), ctx.model(synth.synthetic_person())]
    frames = _frames(3)
    ref = _alone(ctx, models, frames, precision)
    with geye_amd.Server(slots=4, precision=precision) as srv:
        tickets = [srv.submit(models[mi], f) for mi, f in frames]
        got = [srv.wait(t)[0] for t in tickets]
        st = srv.stats()
    say(st)
    assert st["completed"] == len(frames)
    for k, (g, r) in enumerate(zip(got, ref)):
        assert g.shape == r.shape and np.array_equal(g.view(np.int32), r.view(np.int32)), k


def test_small_budget_evicts_and_stays_exact(ctx):
    """A budget that holds roughly one 1280x720 working set: workers must evict idle plans,
    requests wait for memory, and results stay exact."""
    say("serve budget")
    car = ctx.model(synth.synthetic_car())
    frames = [(0, synth.make_frame(W, H, seed=7 + i)) for i, (W, H) in enumerate(STREAMS * 2)]
    ref = _alone(ctx, [car], frames, "bf16")
    big = car.plan_bytes(1280, 720)
    budget = int(big * 1.3)
    say("budget", budget, "1280x720 working set", big)
    with geye_amd.Server(slots=3, hbm_budget=budget, precision="bf16") as srv:
        tickets = [srv.submit(car, f) for _, f in frames]
        got = [srv.wait(t)[0] for t in tickets]
        st = srv.stats()
    say(st)
    assert st["peak_bytes"] <= budget and st["evictions"] > 0
    for g, r in zip(got, ref):
        assert np.array_equal(g.view(np.int32), r.view(np.int32))


def test_frame_larger_than_budget_is_refused(ctx):
    say("serve refuse")
    car = ctx.model(synth.synthetic_car())
    small = synth.make_frame(320, 240, seed=1)
    big = synth.make_frame(1280, 720, seed=2)
    budget = int(car.plan_bytes(320, 240) * 1.5)
    with geye_amd.Server(slots=2, hbm_budget=budget, precision="bf16") as srv:
        t_big = srv.submit(car, big)
        t_small = srv.submit(car, small)
        with pytest.raises(geye_amd.GeyeError):
            srv.wait(t_big)
        rows, _ = srv.wait(t_small)
    ref = car.car_detection(small, thresh=-0.5, overlap=0.5, precision="bf16")
    assert np.array_equal(rows.view(np.int32), ref.view(np.int32))


def test_queue_cap_reports_busy(ctx):
    say("serve queue cap")
    car = ctx.model(synth.synthetic_car())
    f = synth.make_frame(640, 480, seed=3)
    with geye_amd.Server(slots=1, queue_cap=2, precision="bf16") as srv:
        tickets, busy = [], 0
        for _ in range(12):
            try:
                tickets.append(srv.submit(car, f))
            except BlockingIOError:
                busy += 1
        res = [srv.wait(t)[0] for t in tickets]
    assert busy > 0 and len(tickets) >= 2
    for r in res[1:]:
        assert np.array_equal(r.view(np.int32), res[0].view(np.int32))
