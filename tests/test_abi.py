"""CPU tests of the C ABI library (no GPU needed).

The library must load, export every entry point include/geye_dpm.h declares, and its
host-only parts (the level plan, the host NMS the device path falls back to, argument
checks) must agree with the oracle.  No compute call here touches a device.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

import geye_amd
from geye_amd import synth
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "geye_dpm.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(geye_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = geye_amd.lib()
    syms = header_symbols()
    assert len(syms) >= 28
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    # the Python mirror binds the same set
    assert sorted(geye_amd.exported_symbols()) == syms


def test_error_path_without_device_work():
    L = geye_amd.lib()
    rc = L.geye_level_plan_dims(None, 0, 8, 10, 640, 480, None, None, None, None, None)
    assert rc == -1
    assert b"" != L.geye_last_error()


def test_model_file_missing_or_malformed_is_an_io_error(tmp_path):
    """The model CSV reader (CT/load_model.cpp:67-71 exits on a missing file; here a status):
    a missing file and a file holding a non-number are GEYE_ERR_IO with a message naming the
    file; a well-formed file gives its numbers in order (commas, spaces and newlines between)."""
    L = geye_amd.lib()
    vals, n = C.POINTER(C.c_float)(), C.c_longlong()
    missing = str(tmp_path / "no_such_comp.csv").encode()
    assert L.geye_csv_values(missing, C.byref(vals), C.byref(n)) == -4
    assert b"no_such_comp.csv" in L.geye_last_error()
    bad = tmp_path / "bad.csv"
    bad.write_text("2,8,10,\n6,13,abc\n")
    assert L.geye_csv_values(str(bad).encode(), C.byref(vals), C.byref(n)) == -4
    good = tmp_path / "good.csv"
    good.write_text("2, 8,10\n-4.058444 ,1e-3\r\n\t7\n")
    assert L.geye_csv_values(str(good).encode(), C.byref(vals), C.byref(n)) == 0
    got = [vals[i] for i in range(n.value)]
    L.geye_free(vals)
    assert got == [2.0, 8.0, 10.0, np.float32(-4.058444), np.float32(1e-3), 7.0]


def test_plan_bytes_from_a_description_needs_no_device(car_model):
    """geye_plan_bytes_desc packs the model and lays out one frame's plan on the host: the
    size is positive, grows with the frame, and a description naming a part filter that does
    not exist is GEYE_ERR_ARG."""
    L = geye_amd.lib()
    d, keep = geye_amd.desc_from_model(car_model)
    sizes = []
    for W, H in [(320, 240), (640, 480), (1920, 1080), (3840, 2160)]:
        b = C.c_longlong()
        assert L.geye_plan_bytes_desc(C.byref(d), W, H, C.byref(b)) == 0, L.geye_last_error()
        sizes.append(b.value)
    assert sizes[0] > 0 and all(a < b for a, b in zip(sizes, sizes[1:]))
    # 4K needs at least its own feature pyramid: 31 planes of the level-0 cells, in floats
    assert sizes[-1] > 31 * (2160 // 8) * (3840 // 8) * 4
    bad_pidx = np.array(car_model.pidx, np.int32)
    bad_pidx[0] = car_model.NoP + 5
    keep.append(bad_pidx)
    d.pidx = bad_pidx.ctypes.data_as(C.POINTER(C.c_int))
    b = C.c_longlong()
    assert L.geye_plan_bytes_desc(C.byref(d), 640, 480, C.byref(b)) == -1


@pytest.mark.parametrize("W,H", [(640, 480), (1280, 720), (1920, 1080), (3840, 2160), (200, 150),
                                 (90, 400)])
def test_level_plan_matches_oracle(car_model, W, H):
    for m in (car_model, synth.synthetic_person(), synth.tiny_model()):
        a = geye_amd.level_plan_dims(m.rsize, m.sbin, m.interval, W, H)
        b = O.level_plan(m, W, H)
        assert (a["max_scale"], a["LEN"]) == (b["max_scale"], b["LEN"])
        assert np.array_equal(a["scales"].view(np.int32), b["scales"].view(np.int32))
        assert np.array_equal(a["feat_h"], b["feat_h"]) and np.array_equal(a["feat_w"], b["feat_w"])


def _random_rows(rng, n, GL=11, tie=False):
    x1 = rng.integers(0, 600, n).astype(np.float32)
    y1 = rng.integers(0, 400, n).astype(np.float32)
    w = rng.integers(8, 120, n).astype(np.float32)
    h = rng.integers(8, 120, n).astype(np.float32)
    r = np.zeros((n, GL), np.float32)
    r[:, 0], r[:, 1], r[:, 2], r[:, 3] = x1, y1, x1 + w, y1 + h
    r[:, 4:GL - 2] = rng.random((n, GL - 6))
    s = rng.standard_normal(n).astype(np.float32)
    r[:, GL - 2] = np.round(s * 2) / 2 if tie else s
    r[:, GL - 1] = rng.random(n).astype(np.float32)
    return r


@pytest.mark.parametrize("n,tie,overlap", [(0, False, 0.5), (1, False, 0.5), (2, False, 0.5),
                                           (50, False, 0.5), (400, False, 0.3),
                                           (400, True, 0.5), (3000, True, 0.5)])
def test_host_nms_matches_oracle(n, tie, overlap):
    rng = np.random.default_rng(n + int(tie))
    rows = _random_rows(rng, n, tie=tie)
    got = geye_amd.nms_host(rows, overlap) if n else np.zeros((0, 11), np.float32)
    exp = O.nms(rows, overlap) if n else np.zeros((0, 11), np.float32)
    assert got.shape == exp.shape
    assert np.array_equal(got.view(np.int32), exp.view(np.int32))


# the reference's declarations (CT/detect_func.h:23-45, CT/get_boxes_func.h:28-35,
# CT/featurepyramid.cpp.new2:114,135, CT/load_model.cpp:42-45, CT/for_use_GPU.h:128-130),
# as the demangler prints them with FLOAT = float
REFERENCE_SIGNATURES = [
    "ini_scales(Model_info*, _IplImage*, int, int)",
    "ini_featsize(Model_info*)",
    "calc_f_pyramid(_IplImage*, Model_info*, int*, float*)",
    "get_boxes(float**, float*, int*, MODEL*, int*, float*, float)",
    "free_features(float**, Model_info*)",
    "nms(float*, float, int*, MODEL*)",
    "get_new_rects(_IplImage*, MODEL*, float*, int*)",
    "fconvsMT(float*, float*, float**, int*, int, int, int*, int**, int*)",
    "dt(float*, float, float, float, float, int*, int*, int*)",
    "add_part_calculation(float*, float*, int*, int*, int, int)",
    "resize(float*, int*, int*, float)",
    "calc_feature(float*, int*, int*, int)",
    "load_model(float)",
    "free_model(MODEL*)",
    "init_cuda",
    "clean_cuda",
]


def test_compat_library_exports_reference_signatures():
    import subprocess

    so = os.path.join(ROOT, "geye_amd", "libgeye_dpm_compat.so")
    out = subprocess.run(["nm", "-D", "--defined-only", "-C", so], capture_output=True, text=True,
                         check=True).stdout
    names = {line.split(" ", 2)[2] for line in out.splitlines() if line.count(" ") >= 2}
    missing = [s for s in REFERENCE_SIGNATURES if s not in names]
    assert not missing, missing


def test_compat_header_compiles_the_reference_call_sequence():
    import subprocess

    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "compat_detect.cpp")], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stderr
