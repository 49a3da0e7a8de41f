"""GPU parity of the drop-in surface: the reference-named C++ layer (libgeye_dpm_compat.so,
driven by tests/compat_detect.cpp in the reference's own call sequence) and the C ABI entry
points it rests on (geye_new_rects, geye_add_part, geye_get_boxes_features), bit-exact
against the oracle."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import geye_amd
from geye_amd import synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "_bin", "compat_detect")


def say(*a):
    print(*a, flush=True)


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.int32)


@pytest.mark.parametrize("which,W,H,thresh", [("car", 640, 480, -1.2), ("synthetic", 320, 240, -0.5)])
def test_new_rects_matches_oracle(ctx, car_model, which, W, H, thresh):
    say("new_rects", which, W, H)
    m = car_model if which == "car" else synth.synthetic_car()
    gm = ctx.model(m)
    frame = synth.make_frame(W, H, seed=1)
    rows = gm.car_detection(frame, thresh=thresh, overlap=0.5)
    assert rows.shape[0] > 0
    for ratio in (1.0, 0.75):
        g = gm.new_rects(rows, W, H, ratio)
        o = O.new_rects(m, rows, W, H, ratio)
        for k in o:
            assert np.array_equal(g[k], o[k]), k


def test_add_part_matches_oracle(ctx):
    say("add_part")
    rng = np.random.default_rng(3)
    R1, R0, P1, P0, ax, ay = 17, 11, 40, 30, 3, 2
    S = rng.standard_normal((R1, R0)).astype(np.float32)
    M = rng.standard_normal((P1, P0)).astype(np.float32)
    g = ctx.add_part_calculation(S, M, ax, ay)
    o = S.copy()
    rs = np.array([R0, R1], np.int32)
    ps = np.array([P0, P1], np.int32)
    fp, ip = C.POINTER(C.c_float), C.POINTER(C.c_int)
    O.lib().ora_add_part_f32(o.ctypes.data_as(fp), M.ctypes.data_as(fp), rs.ctypes.data_as(ip),
                             ps.ctypes.data_as(ip), ax, ay)
    assert np.array_equal(bits(g), bits(o))


def test_get_boxes_on_host_features_matches_oracle(ctx, car_model):
    say("get_boxes on host features, real car 640x480")
    frame = synth.make_frame(640, 480, seed=1)
    plan = O.level_plan(car_model, 640, 480)
    feats, _, _, _ = O.pyramid(frame, car_model.sbin, car_model.interval, plan["max_scale"])
    gm = ctx.model(car_model)
    g = gm.get_boxes_features(feats, 640, 480, thresh=-1.2)
    o = O.detect(car_model, frame, -1.2)
    assert g.shape == o.shape and np.array_equal(bits(g), bits(o))


def _read(path, W, H):
    raw = open(path, "rb").read()
    off = 0

    def take(n, dt):
        nonlocal off
        a = np.frombuffer(raw, dtype=dt, count=n, offset=off)
        off += n * np.dtype(dt).itemsize
        return a

    ncand, GL = take(2, np.int32)
    boxes = take(ncand * GL, np.float32).reshape(ncand, GL)
    nkept = int(take(1, np.int32)[0])
    rects = take(nkept * GL, np.float32).reshape(nkept, GL)
    ascore = take(W * H, np.float32).reshape(W, H)
    res = {}
    if nkept:
        res = dict(point=take(nkept * 4, np.int32).reshape(nkept, 4),
                   OR_point=take(nkept * 4, np.int32).reshape(nkept, 4),
                   type=take(nkept, np.int32), scale=take(nkept, np.float32),
                   score=take(nkept, np.float32))
    LEN = int(take(1, np.int32)[0])
    return boxes, rects, ascore, res, LEN


@pytest.mark.parametrize("which,W,H,thresh", [("car", 640, 480, -1.2), ("synthetic", 320, 240, -0.5)])
def test_reference_call_sequence_end_to_end(tmp_path, car_model, which, W, H, thresh):
    """detect() + nms() + get_new_rects() through the reference's names and types: the
    candidate rows, the NMS survivors, A_SCORE and RESULT all bit-exact against the oracle."""
    assert os.path.exists(DRIVER), "tests/_bin/compat_detect is built by __graft_entry__.build()"
    say("compat driver", which, W, H, thresh)
    m = car_model if which == "car" else synth.synthetic_car()
    comp, root, part = O.write_csv_model(m, str(tmp_path))
    frame = synth.make_frame(W, H, seed=1)
    fpath = tmp_path / "frame.bgr"
    frame.tofile(fpath)
    out = tmp_path / "out.bin"
    env = dict(os.environ, GEYE_F_NAME_COM=comp, GEYE_F_NAME_ROOT=root, GEYE_F_NAME_PART=part)
    r = subprocess.run([DRIVER, str(fpath), str(W), str(H), str(thresh), "0.5", str(out)],
                       env=env, capture_output=True, text=True, timeout=120)
    say(r.stdout.strip(), r.stderr.strip())
    assert r.returncode == 0
    boxes, rects, ascore, res, LEN = _read(out, W, H)
    assert LEN == O.level_plan(m, W, H)["LEN"]
    o_boxes = O.detect(m, frame, thresh)
    assert boxes.shape == o_boxes.shape and np.array_equal(bits(boxes), bits(o_boxes))
    o_rects = O.nms(o_boxes, 0.5)
    assert rects.shape == o_rects.shape and np.array_equal(bits(rects), bits(o_rects))
    assert np.array_equal(bits(ascore), bits(O.ascore(m, frame, thresh)))
    if rects.shape[0]:
        o_res = O.new_rects(m, o_rects, W, H, 1.0)
        for k in o_res:
            assert np.array_equal(res[k], o_res[k]), k


def test_failures_are_status_codes_and_the_context_survives(ctx, tmp_path):
    """Failure paths on the device: a model file that is missing (GEYE_ERR_IO, as
    CT/load_model.cpp:67-71 reports with exit()), a device allocation beyond the GPU's memory
    (GEYE_ERR_NOMEM) and a frame too small for a plan (GEYE_ERR_ARG); after each the same
    context still detects, bit-exact with the oracle."""
    L = geye_amd.lib()
    h = C.c_void_p()
    rc = L.geye_model_load_csv(ctx.h, str(tmp_path / "none_comp.csv").encode(),
                               str(tmp_path / "none_root.csv").encode(),
                               str(tmp_path / "none_part.csv").encode(), C.byref(h))
    assert rc == -4 and b"none_comp.csv" in L.geye_last_error()
    p = C.c_void_p()
    assert L.geye_dev_alloc(ctx.h, C.c_size_t(1 << 60), C.byref(p)) == -3
    m = synth.tiny_model()
    gm = ctx.model(m)
    tiny = synth.make_frame(6, 6, seed=1)
    with pytest.raises(geye_amd.GeyeError):
        gm.get_boxes(tiny, thresh=-0.5)
    frame = synth.make_frame(200, 150, seed=21)
    rows = gm.get_boxes(frame, thresh=-0.5)
    ref = O.detect(m, frame, -0.5)
    assert rows.shape == ref.shape and np.array_equal(rows.view(np.int32), ref.view(np.int32))
    gm.free()

