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


@pytest.mark.parametrize("precision", ["bf16", "bf16x3", "fp32"])
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


def _other_interval_model():
    """A small model with interval 5: its pyramid differs from the interval-10 models'."""
    m = synth.tiny_model(seed=9)
    m.interval = 5
    return m


@pytest.mark.parametrize("precision", ["bf16", "bf16x3", "fp32"])
@pytest.mark.parametrize("W,H", [(640, 480), (1920, 1080)])
def test_detect_models_equals_each_model_alone(ctx, precision, W, H):
    """geye_detect_models: car and person share one pyramid (person's levels are a prefix of
    car's), the interval-5 model builds its own; every model's survivors are its
    car_detection alone, bit for bit, in either model order."""
    say("detect models", precision, W, H)
    models = [ctx.model(synth.synthetic_car()), ctx.model(synth.synthetic_person()),
              ctx.model(_other_interval_model())]
    frame = synth.make_frame(W, H, seed=77)
    alone = [m.car_detection(frame, thresh=-0.5, overlap=0.5, precision=precision) for m in models]
    for order in ([0, 1, 2], [1, 0, 2], [2, 1, 0]):
        got = ctx.detect_models([models[i] for i in order], frame, thresh=-0.5, overlap=0.5,
                                precision=precision)
        for g, i in zip(got, order):
            r = alone[i]
            assert g.shape == r.shape and np.array_equal(g.view(np.int32), r.view(np.int32)), \
                (order, i)
    d = ctx.dev_upload(frame)
    try:
        got = ctx.detect_models(models, frame, thresh=-0.5, overlap=0.5, precision=precision,
                                d_frame=d)
    finally:
        ctx.dev_free(d)
    for g, r in zip(got, alone):
        assert np.array_equal(g.view(np.int32), r.view(np.int32))
    for m in models:
        m.free()


@pytest.mark.parametrize("precision", ["bf16", "bf16x3", "fp32"])
def test_served_model_pairs_identical_alone_and_in_the_mix(ctx, precision):
    """Frames submitted once for car and person (geye_serve_submit_models: one pyramid per
    frame), mixed with single-model requests, on the four camera sizes."""
    say("serve model pairs", precision)
    car, person = ctx.model(synth.synthetic_car()), ctx.model(synth.synthetic_person())
    sizes = [(640, 480), (1280, 720), (1920, 1080), (3840, 2160)]
    frames = [synth.make_frame(w, h, seed=600 + i) for i, (w, h) in enumerate(sizes)]
    ref = [(car.car_detection(f, -0.5, 0.5, precision),
            person.car_detection(f, -0.5, 0.5, precision)) for f in frames]
    with geye_amd.Server(slots=3, precision=precision, queue_cap=64) as srv:
        pairs = [srv.submit_models([car, person], f) for f in frames]
        singles = [srv.submit(person, f) for f in frames]
        got_pairs = [[srv.wait(t)[0] for t in p] for p in pairs]
        got_singles = [srv.wait(t)[0] for t in singles]
        st = srv.stats()
    say(st)
    assert st["completed"] == 3 * len(frames)
    for k in range(len(frames)):
        for g, r in zip(got_pairs[k], ref[k]):
            assert np.array_equal(g.view(np.int32), r.view(np.int32)), (sizes[k], "pair")
        assert np.array_equal(got_singles[k].view(np.int32), ref[k][1].view(np.int32)), sizes[k]
    car.free()
    person.free()


@pytest.mark.parametrize("precision,W,H,nf,thresh", [("bf16", 640, 480, 5, -0.5),
                                                     ("fp32", 640, 480, 5, -0.5),
                                                     ("fp16", 800, 600, 2, -0.5),
                                                     ("bf16", 1920, 1080, 3, -0.5),
                                                     ("bf16x3", 1280, 720, 3, -0.5),
                                                     ("fp16x3", 640, 480, 2, -0.5),
                                                     ("fp32", 640, 480, 3, 1.5)])
def test_detect_batch_equals_each_frame_alone(ctx, precision, W, H, nf, thresh):
    """geye_detect_batch: nf frames through car, person (sharing car's pyramid) and an
    interval-5 model (its own pyramid), every launch before the NMS covering all frames and the
    frames' NMS passes running together; each (frame, model) result is that frame's
    car_detection alone, bit for bit (thresh 1.5 leaves frames with few or no candidates)."""
    say("detect batch", precision, W, H, nf, thresh)
    models = [ctx.model(synth.synthetic_car()), ctx.model(synth.synthetic_person()),
              ctx.model(_other_interval_model())]
    frames = [synth.make_frame(W, H, seed=900 + k) for k in range(nf)]
    before = ctx.nms_fallbacks()
    alone = [[m.car_detection(f, thresh=thresh, overlap=0.5, precision=precision) for m in models]
             for f in frames]
    replays = ctx.nms_fallbacks() - before  # score ties of the small model's NMS, if any
    for rep in range(2):
        before = ctx.nms_fallbacks()
        got = ctx.detect_batch(models, frames, thresh=thresh, overlap=0.5, precision=precision)
        for f in range(nf):
            for i in range(len(models)):
                g, r = got[f][i], alone[f][i]
                assert g.shape == r.shape and np.array_equal(g.view(np.int32), r.view(np.int32)), \
                    (rep, f, i, g.shape, r.shape)
        # the batch takes the host replay exactly where the frames alone took it
        assert ctx.nms_fallbacks() - before == replays
    d = [ctx.dev_upload(f) for f in frames]
    try:
        got = ctx.detect_batch(models[:2], frames, thresh=thresh, overlap=0.5, precision=precision,
                               d_frames=d)
    finally:
        for p in d:
            ctx.dev_free(p)
    for f in range(nf):
        for i in range(2):
            assert np.array_equal(got[f][i].view(np.int32), alone[f][i].view(np.int32)), (f, i)
    for m in models:
        m.free()



def test_detect_batch_after_models_are_freed_and_made_again(ctx):
    """The batch front's captured graph belongs to the models' plans: freeing the models
    releases those plans, and a batch of new models (whose plans may sit at the same
    addresses) captures its own graph.  Every cycle's result is each frame's alone."""
    W, H, nf = 320, 240, 2
    frames = [synth.make_frame(W, H, seed=950 + k) for k in range(nf)]
    for cycle in range(3):
        say("detect batch cycle", cycle)
        models = [ctx.model(synth.synthetic_car()), ctx.model(synth.synthetic_person())]
        alone = [[m.car_detection(f, thresh=-0.5, overlap=0.5, precision="bf16x3")
                  for m in models] for f in frames]
        for rep in range(2):  # the first call captures, the second replays
            got = ctx.detect_batch(models, frames, thresh=-0.5, overlap=0.5, precision="bf16x3")
            for f in range(nf):
                for i in range(len(models)):
                    assert np.array_equal(got[f][i].view(np.int32), alone[f][i].view(np.int32)), \
                        (cycle, rep, f, i)
        for m in models:
            m.free()
