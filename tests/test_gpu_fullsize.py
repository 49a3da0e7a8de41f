"""Parity of the benchmark configurations at their full sizes.

C2: one 1920x1080 rectangle-noise frame (seed 2), the synthetic car model, FP32, threshold
    -0.5, NMS overlap 0.5 -- every candidate row (about a million) and every NMS survivor
    bit-identical to the oracle (CT/get_boxes.cpp:1497-1570, CT/nms.cpp:94-217).
C3: the same frame with the convolution in BF16 and FP16 on MFMA (FP32 accumulation), and in
    the split forms (hi/lo operand pairs, three MFMAs per tap): score maps against the
    oracle's FP64 scoring (max-abs and relative error printed and bounded), the windowed DT's
    rows equal to the divide and conquer's on the scores the form computed, and the post-NMS
    detection set against the FP32 form's at IoU >= 0.99 (all of it for the split forms).
C5: the four camera sizes 640x480, 1280x720, 1920x1080 and 3840x2160, each frame detected
    with the car and the person model: every result identical alone and in the serving mix.
"""
import os

import numpy as np
import pytest

import geye_amd
from geye_amd import synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu

W, H = 1920, 1080
THREADS = max(1, min(16, os.cpu_count() or 1))


def say(*a):
    print(*a, flush=True)


def rows_equal(g, o):
    assert g.shape == o.shape, (g.shape, o.shape)
    if g.size:
        bad = np.nonzero(~np.all(g.view(np.int32) == o.view(np.int32), axis=1))[0]
        assert bad.size == 0, f"{bad.size} rows differ, first {bad[:5]}: {g[bad[0]]} vs {o[bad[0]]}"


def iou_agreement(ref, got, thr=0.99):
    """Fractions of each set's root boxes matched by a box of the other at IoU >= thr."""
    if len(ref) == 0 or len(got) == 0:
        same = float(len(ref) == len(got))
        return same, same
    a, b = ref[:, :4].astype(np.float64), got[:, :4].astype(np.float64)
    y1 = np.maximum(a[:, None, 0], b[None, :, 0])
    x1 = np.maximum(a[:, None, 1], b[None, :, 1])
    y2 = np.minimum(a[:, None, 2], b[None, :, 2])
    x2 = np.minimum(a[:, None, 3], b[None, :, 3])
    inter = np.clip(y2 - y1 + 1, 0, None) * np.clip(x2 - x1 + 1, 0, None)
    aa = (a[:, 2] - a[:, 0] + 1) * (a[:, 3] - a[:, 1] + 1)
    ab = (b[:, 2] - b[:, 0] + 1) * (b[:, 3] - b[:, 1] + 1)
    m = inter / (aa[:, None] + ab[None, :] - inter)
    return float((m.max(axis=1) >= thr).mean()), float((m.max(axis=0) >= thr).mean())


@pytest.fixture(scope="module")
def c2():
    model = synth.synthetic_car()
    frame = synth.make_frame(W, H, seed=2)
    say("oracle 1080p synthetic car,", THREADS, "threads")
    rows = O.detect(model, frame, -0.5, 0, THREADS)
    kept = O.nms(rows, 0.5)
    say("oracle:", len(rows), "candidates,", len(kept), "boxes")
    return model, frame, rows, kept


@pytest.mark.parametrize("dt_form", ["auto", "window"])
def test_c2_fp32_1080p_bit_exact(ctx, c2, dt_form):
    """FP32 with the reference's divide-and-conquer DT (the plain form) and with the windowed
    DT of the 16-bit forms: both bit-identical to the oracle at 1080p."""
    model, frame, rows, kept = c2
    gm = ctx.model(model)
    before = ctx.nms_fallbacks()
    ctx.dt_form(dt_form)
    try:
        g = gm.get_boxes(frame, thresh=-0.5, precision="fp32")
        say("candidates", len(g))
        assert len(g) > 500_000
        rows_equal(g, rows)
        k = gm.car_detection(frame, thresh=-0.5, overlap=0.5, precision="fp32")
    finally:
        ctx.dt_form("auto")
    say("boxes", len(k))
    rows_equal(k, kept)
    assert ctx.nms_fallbacks() == before, "the NMS was replayed on the host"
    gm.free()


@pytest.fixture(scope="module")
def c3_fp64(c2):
    model, frame, _, _ = c2
    say("oracle FP64 scores 1080p")
    maps = O.score_maps(model, frame, precision=1, threads=THREADS)
    return np.concatenate([mp.ravel() for _, _, mp in maps])


# Bounds: the measured errors of each form with headroom.  BF16 keeps 8 significant bits
# per operand, FP16 11, the split forms 16 and 22; the scores span about -4..6.  The split
# forms must reproduce the FP32 detection set: every box matched at IoU >= 0.99.
@pytest.mark.parametrize("prec,max_abs,min_agree", [("bf16", 0.03, 0.97), ("fp16", 0.005, 0.98),
                                                    ("bf16x3", 2e-4, 1.0), ("fp16x3", 1e-4, 1.0)])
def test_c3_precision_1080p(ctx, c2, c3_fp64, prec, max_abs, min_agree):
    model, frame, _, kept32 = c2
    gm = ctx.model(model)
    rows = gm.get_boxes(frame, thresh=-0.5, precision=prec)
    g = gm.last_scores().astype(np.float64)
    # the DT of this form's part scores is dt()'s: the divide and conquer gives the same rows
    ctx.dt_form("divide")
    try:
        rows_dc = gm.get_boxes(frame, thresh=-0.5, precision=prec)
    finally:
        ctx.dt_form("auto")
    rows_equal(rows, rows_dc)
    ref = c3_fp64
    assert g.shape == ref.shape
    err = np.abs(g - ref)
    rel = err.max() / np.abs(ref).max()
    say(prec, f"score error vs FP64: max abs {err.max():.3e}, mean abs {err.mean():.3e}, "
              f"max abs / max |score| {rel:.3e}")
    assert err.max() < max_abs
    before = ctx.nms_fallbacks()
    k = gm.car_detection(frame, thresh=-0.5, overlap=0.5, precision=prec)
    assert ctx.nms_fallbacks() == before
    ref_m, got_m = iou_agreement(kept32, k)
    say(prec, f"post-NMS boxes at IoU >= 0.99: {ref_m:.4f} of the FP32 set ({len(kept32)}) "
              f"matched, {got_m:.4f} of this form's ({len(k)})")
    assert min(ref_m, got_m) >= min_agree
    gm.free()


C5_STREAMS = [(640, 480), (1280, 720), (1920, 1080), (3840, 2160)]


@pytest.mark.parametrize("precision", ["bf16", "bf16x3", "fp32"])
def test_c5_camera_sizes_alone_and_in_the_mix(ctx, precision):
    say("serve C5 sizes", precision)
    models = [ctx.model(synth.synthetic_car()), ctx.model(synth.synthetic_person())]
    frames = [synth.make_frame(w, h, seed=500 + i) for i, (w, h) in enumerate(C5_STREAMS)]
    ref = [[m.car_detection(f, thresh=-0.5, overlap=0.5, precision=precision) for m in models]
           for f in frames]
    with geye_amd.Server(slots=3, precision=precision, queue_cap=64) as srv:
        tickets = [[srv.submit(m, f) for m in models] for f in frames for _ in range(2)]
        got = [[srv.wait(t)[0] for t in pair] for pair in tickets]
        st = srv.stats()
    say(st)
    for k, pair in enumerate(got):
        for mi, g in enumerate(pair):
            r = ref[k // 2][mi]
            assert g.shape == r.shape and np.array_equal(g.view(np.int32), r.view(np.int32)), \
                (C5_STREAMS[k // 2], mi)
    for m in models:
        m.free()
