"""GPU parity of every stage of the FP32 path against the CPU oracle (bit-exact).

The contract (FP32 plain form): resize, HOG features, convolution, DT values and argmax
offsets, scores and candidate rows are bit-identical to the oracle on the same inputs,
and the post-NMS boxes are identical.  Inputs are seeded synthetic frames and models.
"""
import os

import numpy as np
import pytest

import geye_amd
from geye_amd import synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(autouse=True)
def nms_path(request, ctx):
    """Every test here runs its NMS on the device: the host replay (geye_nms_fallbacks) must
    not have run, except in the test built to force it (a tie on a zero score), whose one NMS
    call must have taken it.  The tests built to create ties must have set the order of the
    tied runs on the device (geye_nms_tie_repairs) instead."""
    before, rep0 = ctx.nms_fallbacks(), ctx.nms_tie_repairs()
    yield
    ran, rep = ctx.nms_fallbacks() - before, ctx.nms_tie_repairs() - rep0
    if request.node.get_closest_marker("nms_replay"):
        assert ran == 1, f"host NMS replays: {ran}"
    else:
        assert ran == 0, f"host NMS replays: {ran}"
    if request.node.get_closest_marker("nms_ties"):
        assert rep >= 1, f"tie repairs: {rep}"


def bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    if a.shape != b.shape:
        return False
    if a.dtype == np.float32:
        return np.array_equal(a.view(np.int32), b.view(np.int32))
    return np.array_equal(a, b)


@pytest.mark.parametrize("W,H,scale", [(67, 45, 0.933033), (160, 120, 0.5), (97, 131, 0.707107),
                                       (320, 240, 0.535887), (64, 48, 1.0)])
def test_resize_matches_oracle(ctx, W, H, scale):
    say("resize", W, H, scale)
    img = O.ipl_to_float(synth.make_frame(W, H, seed=W * 7 + H))
    g = ctx.resize(img, scale)
    o = O.resize(img, scale)
    assert bits_equal(g, o)


def test_resize_matches_reference_vectors(ctx):
    from conftest import GOLDEN

    z = np.load(os.path.join(GOLDEN, "resize_ref.npz"), allow_pickle=False)
    for k, s in enumerate(z["scales"]):
        say("resize golden", k, float(s))
        assert bits_equal(ctx.resize(z["img"], float(s)), z[f"out{k}"])


@pytest.mark.parametrize("W,H,sbin", [(64, 48, 8), (64, 48, 4), (133, 97, 8), (133, 97, 4),
                                      (320, 240, 4), (203, 161, 6), (37, 29, 8),
                                      # 65 and 129 cell rows: vote tiles of 64 cells with one
                                      # left over; 65 cell columns
                                      (40, 520, 8), (36, 516, 4), (520, 40, 8)])
def test_calc_feature_matches_oracle(ctx, W, H, sbin):
    say("calc_feature", W, H, sbin)
    img = O.ipl_to_float(synth.make_frame(W, H, seed=3 + W))
    g = ctx.calc_feature(img, sbin)
    o = O.calc_feature(img, sbin)
    assert g.shape == o.shape
    assert bits_equal(g, o)


def test_pyramid_matches_oracle(ctx):
    say("pyramid 320x240")
    frame = synth.make_frame(320, 240, seed=9)
    gf, gh, gw, gs = ctx.calc_f_pyramid(frame, 8, 10, 20)
    of, oh, ow, os_ = O.pyramid(frame, 8, 10, 20)
    assert np.array_equal(gh, oh) and np.array_equal(gw, ow) and bits_equal(gs, os_)
    for l, (a, b) in enumerate(zip(gf, of)):
        assert bits_equal(a, b), f"level {l}"


@pytest.mark.parametrize("h,w,n", [(6, 6, 8), (6, 11, 2), (7, 10, 2), (8, 9, 2), (9, 6, 3),
                                   (5, 6, 1), (6, 13, 1), (14, 5, 4), (16, 16, 1)])
def test_conv_matches_oracle(ctx, h, w, n):
    say("conv", h, w, n)
    rng = np.random.default_rng(h * 100 + w)
    feat = rng.random((31, 83, 71), dtype=np.float32) * 0.3
    filt = [rng.standard_normal(31 * w * h).astype(np.float32) * 0.05 for _ in range(n)]
    g = ctx.fconvsMT(feat, filt, h, w)
    for k in range(n):
        assert bits_equal(g[k], O.conv(feat, filt[k], h, w)), f"filter {k}"


@pytest.mark.parametrize("d0,d1,defs", [
    (23, 17, (0.01, 0.0, 0.01, 0.0)),
    (41, 64, (0.023062, 0.010890, 0.015328, -0.009441)),
    (263, 40, (0.01, 0.0, 0.01, 0.0)),
    (30, 473, (0.01, 0.0, 0.01, 0.0)),
    (1, 9, (0.1, 0.0, 0.1, 0.0)),
    (9, 1, (0.1, 0.0, 0.1, 0.0)),
    (130, 70, (0.05, 0.02, 0.03, -0.01)),
])
def test_dt_matches_oracle(ctx, d0, d1, defs):
    say("dt", d0, d1, defs)
    rng = np.random.default_rng(d0 * 1000 + d1)
    v = rng.standard_normal((d1, d0)).astype(np.float32)
    gM, gIx, gIy = ctx.dt(v, *defs)
    oM, oIx, oIy = O.dt(v, *defs)
    assert bits_equal(gM, oM)
    assert np.array_equal(gIx, oIx)
    assert np.array_equal(gIy, oIy)


def test_dt_matches_reference_vectors(ctx):
    from conftest import GOLDEN

    z = np.load(os.path.join(GOLDEN, "dt_ref.npz"), allow_pickle=False)
    for k in range(int(z["n"])):
        say("dt golden", k)
        d = [float(x) for x in z[f"def{k}"]]
        M, Ix, Iy = ctx.dt(z[f"in{k}"], *d)
        assert bits_equal(M, z[f"M{k}"]) and np.array_equal(Ix, z[f"Ix{k}"]) and np.array_equal(Iy, z[f"Iy{k}"])


@pytest.mark.parametrize("d0,d1,defs,scale,ties", [
    (23, 17, (0.01, 0.0, 0.01, 0.0), 0.15, False),
    (41, 64, (0.023062, 0.010890, 0.015328, -0.009441), 0.3, False),
    (263, 40, (0.01, 0.0, 0.01, 0.0), 0.1, False),
    (30, 473, (0.05, 0.02, 0.03, -0.01), 0.5, False),
    (1, 9, (0.1, 0.0, 0.1, 0.0), 0.2, False),
    (9, 1, (0.1, 0.0, 0.1, 0.0), 0.2, False),
    (130, 70, (0.2, 0.1, 0.3, -0.2), 0.4, False),
    (5, 6, (0.01, 0.0, 0.01, 0.0), 0.2, False),
    (200, 90, (0.01, 0.0, 0.01, 0.0), 0.3, False),    # value ranges giving half-widths
    (150, 33, (0.01, 0.0, 0.01, 0.0), 0.4, False),    # of about 14, 16, 20 and 24 (the
    (96, 80, (0.01, 0.0, 0.01, 0.0), 0.6, False),     # last two run the divide and conquer)
    (70, 75, (0.01, 0.0, 0.01, 0.0), 0.85, False),
    (64, 64, (0.01, 0.0, 0.01, 0.0), 100.0, False),  # argmin band wider than any window
    (50, 40, (0.0, 0.0, 0.0, 0.0), 1.0, False),      # a = 0: no band
    (37, 29, (-0.01, 0.0, -0.02, 0.0), 1.0, False),  # a < 0: no band
    (120, 90, (0.01, 0.0, 0.01, 0.0), 0.3, True),    # values on a 1/8 grid: many exact ties
    (90, 120, (0.25, 0.5, 0.25, -0.5), 1.0, True),
    (60, 70, (0.001, 0.0002, 0.002, -0.0001), 0.01, True),  # weak costs, tiny ranges
])
def test_dt_window_matches_oracle(ctx, d0, d1, defs, scale, ties):
    """DT of the 16-bit forms (geye_dt_window) is dt() bit for bit: M, Ix and Iy equal the
    oracle's on random maps whose value ranges give window half-widths from 8 to 16, on ranges
    too wide for any window and on a <= 0 (the divide and conquer runs there), and on maps
    full of exact ties, where the first minimum decides."""
    say("dt window", d0, d1, defs, scale, ties)
    rng = np.random.default_rng(d0 * 7919 + d1)
    v = rng.standard_normal((d1, d0)) * scale
    if ties:
        v = np.round(v * 8 / scale) * scale / 8
    v = v.astype(np.float32)
    M, Ix, Iy = ctx.dt(v, *defs, window=True)
    oM, oIx, oIy = O.dt(v, *defs)
    assert bits_equal(M, oM)
    assert np.array_equal(Ix, oIx) and np.array_equal(Iy, oIy)


@pytest.mark.parametrize("kind", ["ramp", "wells", "spike", "vertex", "checker"])
def test_dt_window_structured_maps(ctx, kind):
    """The windowed DT on structured maps: smooth ramps (argmins at the window's edge or
    centre), isolated deep wells (distant sources win across a flat map), a single spike far
    below a line, b large against a (the cost's vertex away from the centre), and a checker of
    two values (exact ties at every offset).  All bit for bit against the oracle."""
    say("dt window structured", kind)
    rng = np.random.default_rng(len(kind))
    d0, d1 = 97, 83
    defs = (0.01, 0.0, 0.01, 0.0)
    if kind == "ramp":
        v = (np.add.outer(np.arange(d1), np.arange(d0)) * 0.003).astype(np.float32)
    elif kind == "wells":
        v = np.zeros((d1, d0), np.float32)
        v[rng.integers(0, d1, 40), rng.integers(0, d0, 40)] = -0.9
    elif kind == "spike":
        v = (rng.standard_normal((d1, d0)) * 0.05).astype(np.float32)
        v[::7, 11] = -1.4
    elif kind == "vertex":
        v = (rng.standard_normal((d1, d0)) * 0.2).astype(np.float32)
        defs = (0.01, 0.09, 0.012, -0.07)
    else:
        v = np.where(np.add.outer(np.arange(d1), np.arange(d0)) % 2 == 0, 0.25, -0.25).astype(np.float32)
    M, Ix, Iy = ctx.dt(v, *defs, window=True)
    oM, oIx, oIy = O.dt(v, *defs)
    assert bits_equal(M, oM)
    assert np.array_equal(Ix, oIx) and np.array_equal(Iy, oIy)


def test_dt_window_matches_reference_vectors(ctx):
    """The windowed DT on the reference's own dt() vectors (random maps, plateaus, ties,
    all-zero maps with a = 0, large coefficients)."""
    from conftest import GOLDEN

    z = np.load(os.path.join(GOLDEN, "dt_ref.npz"), allow_pickle=False)
    for k in range(int(z["n"])):
        say("dt window golden", k)
        d = [float(x) for x in z[f"def{k}"]]
        M, Ix, Iy = ctx.dt(z[f"in{k}"], *d, window=True)
        assert bits_equal(M, z[f"M{k}"]) and np.array_equal(Ix, z[f"Ix{k}"]) and np.array_equal(Iy, z[f"Iy{k}"])


def _rows_equal(g, o):
    assert g.shape == o.shape, (g.shape, o.shape)
    if g.size:
        bad = np.nonzero(~np.all(g.view(np.int32) == o.view(np.int32), axis=1))[0]
        assert bad.size == 0, f"{bad.size} rows differ, first {bad[:5]}: {g[bad[0]]} vs {o[bad[0]]}"


def test_get_boxes_tiny_model(ctx):
    say("get_boxes tiny 200x150")
    m = synth.tiny_model()
    frame = synth.make_frame(200, 150, seed=21)
    gm = ctx.model(m)
    g = gm.get_boxes(frame, thresh=-0.5)
    o = O.detect(m, frame, -0.5)
    _rows_equal(g, o)


def test_get_boxes_real_car(ctx, car_model):
    say("get_boxes real car 640x480")
    frame = synth.make_frame(640, 480, seed=1)
    gm = ctx.model(car_model)
    for thresh in (-0.5, -1.2):
        g = gm.get_boxes(frame, thresh=thresh)
        o = O.detect(car_model, frame, thresh)
        _rows_equal(g, o)


def test_scores_synthetic_car(ctx):
    say("scores synthetic car 480x360")
    m = synth.synthetic_car()
    frame = synth.make_frame(480, 360, seed=4)
    gm = ctx.model(m)
    g = gm.get_boxes(frame, thresh=-0.5)
    o = O.detect(m, frame, -0.5)
    _rows_equal(g, o)
    maps = O.score_maps(m, frame)
    flat = np.concatenate([mp.ravel() for _, _, mp in maps])
    assert bits_equal(gm.last_scores(), flat)


def test_car_detection_matches_oracle(ctx, car_model):
    say("car_detection real car 640x480")
    frame = synth.make_frame(640, 480, seed=1)
    gm = ctx.model(car_model)
    g = gm.car_detection(frame, thresh=-1.2, overlap=0.5)
    o = O.nms(O.detect(car_model, frame, -1.2), 0.5)
    _rows_equal(g, o)


def test_car_detection_synthetic(ctx):
    say("car_detection synthetic car 320x240")
    m = synth.synthetic_car()
    frame = synth.make_frame(320, 240, seed=8)
    gm = ctx.model(m)
    g = gm.car_detection(frame, thresh=-0.5, overlap=0.5)
    o = O.nms(O.detect(m, frame, -0.5), 0.5)
    _rows_equal(g, o)


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
def test_padded_width_step_reads_the_same_pixels(ctx, precision):
    """IplImage rows are padded to 4 bytes (widthStep > 3 W for W = 165).  The first resize
    round and the frame's gradients read the BGR bytes in place: with padding bytes of 255
    between the rows, every candidate row equals the unpadded frame's, bit for bit."""
    import ctypes as C

    W, H = 165, 123
    ws = (3 * W + 3) & ~3
    say("get_boxes widthStep", ws, "for", 3 * W, precision)
    frame = synth.make_frame(W, H, seed=21)
    padded = np.full((H, ws), 255, np.uint8)
    padded[:, :3 * W] = frame.reshape(H, 3 * W)
    gm = ctx.model(synth.synthetic_car())
    ref = gm.get_boxes(frame, thresh=-0.8, precision=precision)
    got = gm._rows(geye_amd.lib().geye_get_boxes, ctx.h, gm.h, padded.ctypes.data, W, H, ws,
                   C.c_float(-0.8), geye_amd.PRECISIONS[precision])
    assert len(ref) > 0
    _rows_equal(got, ref)
    gm.free()


@pytest.mark.parametrize("overlap", [0.3, 0.5, 0.7, 0.95])
def test_car_detection_overlaps(ctx, overlap):
    """Device NMS in the detect path against the oracle at several overlaps, from loose (the
    every-pick suppression pass) to nearly disjoint-only (the tile-indexed pass, 0.5 and up)."""
    say("car_detection synthetic car 480x360 overlap", overlap)
    m = synth.synthetic_car()
    frame = synth.make_frame(480, 360, seed=13)
    gm = ctx.model(m)
    g = gm.car_detection(frame, thresh=-0.5, overlap=overlap)
    o = O.nms(O.detect(m, frame, -0.5), overlap)
    _rows_equal(g, o)


def _random_boxes(n, seed, ties=False, GL=11):
    rng = np.random.default_rng(seed)
    y1 = rng.integers(0, 400, n).astype(np.float32)
    x1 = rng.integers(0, 600, n).astype(np.float32)
    h = rng.integers(20, 120, n).astype(np.float32)
    w = rng.integers(20, 160, n).astype(np.float32)
    sc = rng.standard_normal(n).astype(np.float32)
    if ties:
        sc = np.round(sc * 4) / 4  # many equal scores
    rows = np.zeros((n, GL), np.float32)
    rows[:, 0], rows[:, 1], rows[:, 2], rows[:, 3] = y1, x1, y1 + h - 1, x1 + w - 1
    rows[:, GL - 3] = rng.integers(0, 6, n)
    rows[:, GL - 2] = sc
    rows[:, GL - 1] = 8.0
    return rows


@pytest.mark.parametrize("n,ties", [(1, False), (2, False), (50, False), (3000, False),
                                    pytest.param(3000, True, marks=pytest.mark.nms_ties),
                                    (20000, False)])
def test_nms_matches_oracle(ctx, n, ties):
    """Device NMS against the oracle.  The ties case (scores rounded to quarters) is the test
    of the tie repair: equal scores decide outcomes there, so the tied runs must take the
    reference's quicksort order on the device (checked by the nms_path fixture)."""
    say("nms", n, ties)
    rows = _random_boxes(n, seed=n + ties, ties=ties)
    g = ctx.nms(rows, 0.5)
    o = O.nms(rows, 0.5)
    _rows_equal(g, o)


@pytest.mark.nms_ties
@pytest.mark.parametrize("n,levels,seed", [(4000, 3, 1), (20000, 7, 2), (1500, 1, 3),
                                           (60000, 24, 4)])
def test_nms_tie_runs_match_oracle(ctx, n, levels, seed):
    """Scores drawn from a handful of values (one value: every box tied), so runs of
    thousands of equal scores decide outcomes: the host's restricted quicksort must put each
    run in the reference's order (bit for bit against the oracle's whole quicksort)."""
    say("nms tie runs", n, levels, seed)
    rows = _random_boxes(n, seed=seed)
    rng = np.random.default_rng(seed + 100)
    rows[:, -2] = (rng.integers(1, levels + 1, n) * 0.375 - 2.0).astype(np.float32)
    rows[rows[:, -2] == 0.0, -2] = 0.125  # no zero score: that run goes to the host replay
    g = ctx.nms(rows, 0.5)
    o = O.nms(rows, 0.5)
    _rows_equal(g, o)


@pytest.mark.nms_replay
def test_nms_zero_score_tie_takes_host_replay(ctx):
    """A tie on score zero (+0 and -0 are one run for the reference's comparisons but two
    for the sort keys) is left to the host replay, which must match the oracle."""
    say("nms zero-score ties")
    rows = _random_boxes(3000, seed=9)
    rows[:, -2] = 0.0
    rows[1::3, -2] = -0.0
    g = ctx.nms(rows, 0.5)
    o = O.nms(rows, 0.5)
    _rows_equal(g, o)


def _hard_geometry_boxes(seed, GL=11):
    """Boxes built to stress the tile index of the suppression pass: a dense cluster of small
    boxes in one tile (so windows hold hundreds of picks and every mask word is used), boxes
    reaching far outside the boxes' extent on every side (negative coordinates included),
    boxes as large as the whole extent, thin boxes, boxes nested in one another, and exact
    duplicates of other boxes with their own scores."""
    rng = np.random.default_rng(seed)
    parts = []
    # cluster: 1500 boxes of 3..9 px inside a 60 x 60 px area (mostly disjoint at 0.5)
    n = 1500
    y = rng.uniform(300, 360, n); x = rng.uniform(500, 560, n)
    h = rng.integers(3, 10, n); w = rng.integers(3, 10, n)
    parts.append(np.stack([y, x, y + h - 1, x + w - 1], 1))
    # reaching outside: centres near the border, sizes up to 900 px
    n = 400
    y = rng.uniform(-300, 1000, n); x = rng.uniform(-300, 1800, n)
    h = rng.integers(20, 900, n); w = rng.integers(20, 900, n)
    parts.append(np.stack([y, x, y + h - 1, x + w - 1], 1))
    # whole-extent and thin boxes
    parts.append(np.array([[-50.0, -50.0, 1100.0, 1900.0], [0.0, 0.0, 1079.0, 1919.0],
                           [10.0, 0.0, 12.0, 1919.0], [0.0, 700.0, 1079.0, 701.0]]))
    # nested: shrinking squares around one point
    k = np.arange(1, 60, dtype=np.float64)
    parts.append(np.stack([500 - 4 * k, 900 - 4 * k, 500 + 4 * k, 900 + 4 * k], 1))
    b = np.concatenate(parts).astype(np.float32)
    dup = b[rng.integers(0, len(b), 200)]
    b = np.concatenate([b, dup])
    rows = np.zeros((len(b), GL), np.float32)
    rows[:, :4] = b
    rows[:, GL - 3] = rng.integers(0, 6, len(b))
    rows[:, GL - 2] = rng.permutation(len(b)).astype(np.float32) * 0.001 - 0.5  # distinct
    rows[:, GL - 1] = 8.0
    return rows


@pytest.mark.parametrize("overlap", [0.5, 0.6, 0.9, 0.999])
def test_nms_indexed_pass_hard_geometry(ctx, overlap):
    """The tile-indexed suppression pass (overlap >= 0.5) on geometry built against it, bit
    for bit against the oracle, scores distinct (so the device path must decide everything)."""
    say("nms hard geometry overlap", overlap)
    rows = _hard_geometry_boxes(seed=int(overlap * 1000))
    g = ctx.nms(rows, overlap)
    o = O.nms(rows, overlap)
    _rows_equal(g, o)


@pytest.mark.parametrize("prec,ulp", [("bf16", 2.0 ** -8), ("fp16", 2.0 ** -10),
                                      ("bf16x3", 2.0 ** -14), ("fp16x3", 2.0 ** -17)])
@pytest.mark.parametrize("h,w,n", [(6, 6, 48), (6, 6, 8), (8, 11, 6), (6, 13, 2), (14, 5, 20),
                                   (9, 6, 1), (12, 12, 48), (16, 4, 20), (16, 16, 4)])
def test_conv_mfma_error_bound(ctx, prec, ulp, h, w, n):
    """MFMA convolution vs the FP64 oracle: every output within ulp * sum |a*b|, the bound
    that rounding both operands to the 16-bit type (and accumulating in FP32) allows.  The
    split forms keep 16 (BF16) or 22 (FP16) significant bits per operand: three products of
    relative error 2^-16 / 2^-22 each, plus the FP32 accumulation of 3 fh fw MFMAs.
    12x12x48 and 16x4x20 have tap columns too large for the kernel's filter registers (plain
    copy path; the split forms' 16x16x4 runs one filter buffer); 16x16x4 has the widest
    feature ring."""
    say("conv mfma", prec, h, w, n)
    rng = np.random.default_rng(h * 31 + w + n)
    feat = (rng.random((31, 77, 69), dtype=np.float32) * 0.4).astype(np.float32)
    filt = [(rng.standard_normal(31 * w * h) * 0.05).astype(np.float32) for _ in range(n)]
    nt = min(3, (n + 15) // 16)
    if prec.endswith("x3") and (4 + w - 1) * (64 + h - 1) * 128 + h * nt * 2048 > 160 * 1024:
        # a split-form tile (64 rows of 128-byte cells) and one tap column of filters exceed
        # the LDS: refused with a status code
        with pytest.raises(geye_amd.GeyeError, match="LDS"):
            ctx.fconvsMT(feat, filt, h, w, precision=prec)
        return
    g = ctx.fconvsMT(feat, filt, h, w, precision=prec)
    for k in range(n):
        exact = O.conv(feat, filt[k], h, w, f64=True)
        mag = O.conv(np.abs(feat), np.abs(filt[k]), h, w, f64=True)
        err = np.abs(g[k].astype(np.float64) - exact)
        assert np.all(err <= ulp * mag * 1.01 + 1e-6), f"filter {k}: max err {err.max()}"


@pytest.mark.parametrize("prec,tol", [("bf16", 0.02), ("fp16", 0.02), ("bf16x3", 2e-4),
                                      ("fp16x3", 2e-4)])
def test_detect_mfma_scores_vs_fp64(ctx, prec, tol):
    """Whole pipeline in the MFMA form: score maps against the FP64 oracle (FP32 features,
    FP64 convolution/DT/part sums).  Tolerance 0.02 absolute for the 16-bit forms, 2e-4 for
    the split forms (scores span about -3..8)."""
    say("detect", prec, "320x240")
    m = synth.synthetic_car()
    frame = synth.make_frame(320, 240, seed=8)
    gm = ctx.model(m)
    gm.get_boxes(frame, thresh=-0.5, precision=prec)
    g = gm.last_scores()
    maps = O.score_maps(m, frame, precision=1)
    ref = np.concatenate([mp.ravel() for _, _, mp in maps])
    err = np.abs(g.astype(np.float64) - ref)
    say(prec, "score max abs err", float(err.max()), "mean", float(err.mean()))
    assert err.max() < tol


@pytest.mark.parametrize("prec,frac", [("bf16", 0.97), ("fp16", 0.97), ("bf16x3", 0.999),
                                       ("fp16x3", 0.999)])
def test_mfma_part_boxes_match_fp32(ctx, prec, frac):
    """The 16-bit forms' part boxes come from the DT of their own part scores (the same DT
    as the FP32 form's, bit for bit, on those scores): candidate rows matched to the FP32 rows
    by root box, component and scale carry the same part boxes except where the 16-bit part
    scores moved an argmin (>= 97% of the part boxes; 99.9% for the split forms)."""
    say("part boxes", prec, "640x480 synthetic car")
    frame = synth.make_frame(640, 480, seed=5)
    gm = ctx.model(synth.synthetic_car())
    g16 = gm.get_boxes(frame, thresh=-0.5, precision=prec)
    g32 = gm.get_boxes(frame, thresh=-0.5, precision="fp32")
    GL = g32.shape[1]
    npart = (GL - 7) // 4
    key = lambda r: (tuple(r[0:4].tolist()), float(r[GL - 3]), float(r[GL - 1]))
    ref = {key(r): r for r in g32}
    same = total = 0
    for r in g16:
        o = ref.get(key(r))
        if o is None:
            continue
        pb = r[4:4 + 4 * npart].reshape(npart, 4)
        po = o[4:4 + 4 * npart].reshape(npart, 4)
        same += int(np.sum(np.all(pb == po, axis=1)))
        total += npart
    say(prec, "matched part boxes", same, "of", total)
    assert total > 1000
    assert same >= frac * total


@pytest.mark.parametrize("prec", ["bf16", "fp16", "bf16x3"])
def test_detect_mfma_weak_deformation(ctx, prec):
    """Quadratic deformation costs 20x weaker (a = 5e-4): a job whose part responses span
    more than a * 16^2 = 0.13 needs a window wider than the windowed DT's 16, so its jobs
    take the divide and conquer, which must still write the selected outputs where the
    scoring and the part boxes read them.  Scores against the FP64 oracle (0.02 absolute),
    part boxes against the FP32 form's."""
    say("weak deformation", prec, "320x240 synthetic car")
    m = synth.synthetic_car()
    m.deff = (m.deff * np.array([0.05, 1.0, 0.05, 1.0], dtype=np.float32)).astype(np.float32)
    frame = synth.make_frame(320, 240, seed=8)
    gm = ctx.model(m)
    g16 = gm.get_boxes(frame, thresh=-0.5, precision=prec)
    g = gm.last_scores()
    maps = O.score_maps(m, frame, precision=1)
    ref = np.concatenate([mp.ravel() for _, _, mp in maps])
    err = np.abs(g.astype(np.float64) - ref)
    say(prec, "score max abs err", float(err.max()))
    assert err.max() < 0.02
    g32 = gm.get_boxes(frame, thresh=-0.5, precision="fp32")
    GL = g32.shape[1]
    npart = (GL - 7) // 4
    key = lambda r: (tuple(r[0:4].tolist()), float(r[GL - 3]), float(r[GL - 1]))
    ref_rows = {key(r): r for r in g32}
    same = total = 0
    for r in g16:
        o = ref_rows.get(key(r))
        if o is None:
            continue
        same += int(np.sum(np.all(r[4:4 + 4 * npart].reshape(npart, 4) ==
                                  o[4:4 + 4 * npart].reshape(npart, 4), axis=1)))
        total += npart
    say(prec, "matched part boxes", same, "of", total)
    assert total > 100
    assert same >= 0.97 * total  # measured 99.5% (bf16) and 99.9% (fp16) at 320x240


@pytest.mark.parametrize("which,W,H,thresh", [("car", 640, 480, -0.5), ("car", 640, 480, -1.2),
                                              ("synthetic", 480, 360, -0.5)])
def test_accumulated_score_matches_oracle(ctx, car_model, which, W, H, thresh):
    """A_SCORE side output of get_boxes: bit-exact against the oracle's calc_a_score loop,
    from the -100 initial map and from a caller's prior map (updated in place)."""
    say("accumulated score", which, W, H, thresh)
    m = car_model if which == "car" else synth.synthetic_car()
    frame = synth.make_frame(W, H, seed=5)
    gm = ctx.model(m)
    gm.get_boxes(frame, thresh=thresh)
    g = gm.accumulated_score(W, H, thresh)
    o = O.ascore(m, frame, thresh)
    assert bits_equal(g, o)
    if thresh < -1.0 or which == "synthetic":   # these frames have scores above thresh
        assert o.max() == 1.0
    prior = np.random.default_rng(1).uniform(-2, 3, (W, H)).astype(np.float32)
    gm.get_boxes(frame, thresh=thresh)
    g2 = gm.accumulated_score(W, H, thresh, prior.copy())
    assert bits_equal(g2, O.ascore(m, frame, thresh, prior.copy()))


@pytest.mark.parametrize("prec", ["fp32", "bf16", "fp16", "bf16x3", "fp16x3"])
@pytest.mark.parametrize("which", ["synthetic", "car"])
def test_dt_forms_give_the_same_detections(ctx, car_model, prec, which):
    """The windowed DT inside the pipeline (half-grid outputs read by the scoring and the
    part boxes) against the divide and conquer, for every precision: candidate rows, score
    maps and NMS survivors bit for bit; in FP32 both also equal the oracle."""
    say("dt forms", prec, which)
    m = ctx.model(car_model if which == "car" else synth.synthetic_car())
    frame = synth.make_frame(640, 480, seed=6)
    thresh = -1.2 if which == "car" else -0.5   # the real car model scores lower on noise
    out = {}
    try:
        for form in ("divide", "window"):
            ctx.dt_form(form)
            rows = m.get_boxes(frame, thresh=thresh, precision=prec)
            scores = m.last_scores().copy()
            kept = m.car_detection(frame, thresh=thresh, overlap=0.5, precision=prec)
            out[form] = (rows, scores, kept)
    finally:
        ctx.dt_form("auto")
    (r1, s1, k1), (r2, s2, k2) = out["divide"], out["window"]
    say("candidates", len(r1), "boxes", len(k1))
    assert len(r1) > 20
    _rows_equal(r2, r1)
    assert bits_equal(s2, s1)
    _rows_equal(k2, k1)
    if prec == "fp32":
        o = O.detect(car_model if which == "car" else synth.synthetic_car(), frame, thresh)
        _rows_equal(r2, o)


@pytest.mark.parametrize("prec", ["fp32", "bf16", "bf16x3"])
def test_graph_replays_match_eager(ctx, prec):
    """The per-frame front replayed as a HIP graph equals the launch-by-launch run on every
    replay, with unrelated eager work and host synchronisation between the replays, for host
    and device-resident frames."""
    say("graph replays", prec)
    m = ctx.model(synth.synthetic_car())
    other = ctx.model(synth.tiny_model())
    f1 = synth.make_frame(480, 360, seed=31)
    f2 = synth.make_frame(200, 150, seed=32)
    ctx.graphs(False)
    ref = m.car_detection(f1, thresh=-0.5, overlap=0.5, precision=prec)
    ref_scores = m.last_scores().copy()
    ctx.graphs(True)
    d = ctx.dev_upload(f1)
    try:
        for k in range(4):
            got = (m.car_detection(f1, thresh=-0.5, overlap=0.5, precision=prec) if k % 2 == 0
                   else m.detect_dev(d, 480, 360, -0.5, 0.5, prec))
            assert bits_equal(got, ref), k
            assert bits_equal(m.last_scores(), ref_scores), k
            ctx.graphs(False)
            other.car_detection(f2, thresh=-0.5, overlap=0.5, precision=prec)  # eager in between
            ctx.synchronize()
            ctx.graphs(True)
    finally:
        ctx.dev_free(d)


def test_graphs_survive_precision_switches(ctx):
    """One plan replays its captured graphs of every form in turn: switching between the
    16-bit and the split forms (whose feature copy is twice as wide) must not move a buffer
    under another form's graph -- every replay equals that form's first result."""
    say("graphs across precision switches")
    m = ctx.model(synth.synthetic_car())
    frame = synth.make_frame(400, 300, seed=41)
    ctx.graphs(True)
    order = ["bf16", "fp32", "bf16x3", "fp16", "fp16x3", "bf16", "bf16x3", "fp32", "fp16x3"]
    first = {}
    for prec in order:
        for _ in range(2):  # warm-up + capture, then a replay
            got = m.car_detection(frame, thresh=-0.5, overlap=0.5, precision=prec)
        scores = m.last_scores().copy()
        if prec not in first:
            first[prec] = (got, scores)
        else:
            assert bits_equal(got, first[prec][0]), prec
            assert bits_equal(scores, first[prec][1]), prec
    m.free()


@pytest.mark.parametrize("prec", ["bf16", "fp16", "bf16x3", "fp16x3"])
def test_every_mfma_schedule_gives_the_same_scores(ctx, prec):
    """The autotuner's candidates (one launch per class; the persistent work-queue launch at
    two tile heights and several grid sizes) produce identical score maps."""
    say("mfma schedules", prec)
    m = ctx.model(synth.synthetic_car())
    frame = synth.make_frame(480, 360, seed=12)
    ctx.graphs(False)
    try:
        m.get_boxes(frame, thresh=-0.5, precision=prec)
        tuned = ctx.mfma_tuning(prec)
        say(tuned)
        assert tuned["choice"] >= 0 and len(tuned["ms"]) >= 2
        ref = m.last_scores().copy()
        for choice in [0, 1, 2, 9, 10]:
            try:
                ctx.mfma_schedule(prec, choice)
            except Exception:
                continue  # not admissible for this plan's LDS
            m.get_boxes(frame, thresh=-0.5, precision=prec)
            assert bits_equal(m.last_scores(), ref), choice
    finally:
        ctx.graphs(True)
