"""geye_amd -- MI355X (gfx950) DPM detection path behind the reference's entry points.

Python mirror of the C ABI in include/geye_dpm.h (ctypes over libgeye_amd.so).  The names
follow the reference's C++ entry points (CT = CAR_TRACKING/ of the reference repository):

    resize(img, scale)                     CT/resize.cc:135
    calc_feature(img, sbin)                CT/featurepyramid.cpp.new2:263
    calc_f_pyramid(frame, sbin, ...)       CT/featurepyramid.cpp.new2:952
    fconvsMT(feat, filters, h, w)          CT/get_boxes_func.h:28 (CT/fconvsMT.cpp:310)
    dt(vals, ax, bx, ay, by)               CT/dt.cpp:101
    get_boxes(frame, thresh)               CT/get_boxes.cpp:583 (with detect, CT/detect.cpp:151)
    nms(boxes, overlap)                    CT/nms.cpp:94
    car_detection(frame, thresh, overlap)  CT/detect.cpp:232 (detect + nms)

Arrays use the reference's layouts: images and features are column-major planes
([c][x][y], y fastest), maps are [x][y].  Every call runs the HIP kernels; there is no CPU
fallback, and a missing library raises at import of the first call.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgeye_amd.so")

# convolution arithmetic: FP32 (the reference's), one MFMA per tap on BF16 / FP16 operands, or
# three MFMAs per tap on hi/lo operand pairs (bf16x3, fp16x3)
GEYE_FP32, GEYE_BF16, GEYE_FP16, GEYE_BF16X3, GEYE_FP16X3 = 0, 1, 2, 3, 4
PRECISIONS = {"fp32": GEYE_FP32, "bf16": GEYE_BF16, "fp16": GEYE_FP16, "bf16x3": GEYE_BF16X3,
              "fp16x3": GEYE_FP16X3}

_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)
_llp = C.POINTER(C.c_longlong)
_dp = C.POINTER(C.c_double)


class GeyeError(RuntimeError):
    pass


class ModelDesc(C.Structure):
    """geye_model_desc (include/geye_dpm.h)."""

    _fields_ = [
        ("numcomponent", C.c_int), ("sbin", C.c_int), ("interval", C.c_int),
        ("max_Y", C.c_int), ("max_X", C.c_int), ("padx", C.c_int), ("pady", C.c_int),
        ("maxpart", C.c_int),
        ("ridx", _ip), ("oidx", _ip), ("offw", _fp), ("rsize", _ip), ("numpart", _ip),
        ("pidx", _ip), ("didx", _ip), ("psize", _ip),
        ("DefL", C.c_int), ("deff", _fp), ("anchor", _ip),
        ("NoR", C.c_int), ("root_size", _ip), ("rootfilter", _fp), ("root_off", _llp),
        ("NoP", C.c_int), ("part_size", _ip), ("partfilter", _fp), ("part_off", _llp),
        ("lsq", _fp),
    ]


# (name, restype, argtypes) of every C ABI entry point
_SIGS = [
    ("geye_last_error", C.c_char_p, []),
    ("geye_device_count", C.c_int, [_ip]),
    ("geye_init", C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    ("geye_shutdown", None, [C.c_void_p]),
    ("geye_model_load_csv", C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p, C.c_char_p,
                                      C.POINTER(C.c_void_p)]),
    ("geye_model_create", C.c_int, [C.c_void_p, C.POINTER(ModelDesc), C.POINTER(C.c_void_p)]),
    ("geye_model_free", None, [C.c_void_p]),
    ("geye_model_info", C.c_int, [C.c_void_p, _ip, _ip, _ip, _ip, _ip, _ip, _ip]),
    ("geye_level_plan", C.c_int, [C.c_void_p, C.c_int, C.c_int, _ip, _ip, _fp, _ip, _ip]),
    ("geye_level_plan_dims", C.c_int, [_ip, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _ip, _ip,
                                       _fp, _ip, _ip]),
    ("geye_nms_host", C.c_int, [_fp, C.c_int, C.c_int, C.c_float, _fp, _ip]),
    ("geye_get_boxes", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                 C.c_float, C.c_int, C.POINTER(_fp), _ip, _ip]),
    ("geye_detect", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int,
                              C.c_float, C.c_float, C.c_int, C.POINTER(_fp), _ip, _ip]),
    ("geye_detect_dev", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                  C.c_float, C.c_float, C.c_int, C.POINTER(_fp), _ip, _ip]),
    ("geye_detect_models", C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                     C.c_int, C.c_int, C.c_float, C.c_float, C.c_int,
                                     C.c_void_p, _ip, _ip]),
    ("geye_detect_models_dev", C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                         C.c_int, C.c_int, C.c_float, C.c_float, C.c_int,
                                         C.c_void_p, _ip, _ip]),
    ("geye_detect_batch", C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int,
                                    C.c_int, C.c_int, C.c_float, C.c_float, C.c_int, C.c_void_p,
                                    _ip, _ip]),
    ("geye_detect_batch_dev", C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                        C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_int,
                                        C.c_void_p, _ip, _ip]),
    ("geye_serve_submit_models", C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                           C.c_int, C.c_int, _llp]),
    ("geye_serve_submit_models_dev", C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                               C.c_int, C.c_int, C.c_int, _llp]),
    ("geye_last_scores", C.c_int, [C.c_void_p, _fp, C.c_longlong, _llp]),
    ("geye_accumulated_score", C.c_int, [C.c_void_p, C.c_float, _fp]),
    ("geye_accumulated_score_dev", C.c_int, [C.c_void_p, C.c_float, C.c_void_p]),
    ("geye_new_rects", C.c_int, [C.c_void_p, C.c_void_p, _fp, C.c_int, C.c_int, C.c_int,
                                 C.c_int, C.c_float, _ip, _ip, _ip, _fp, _fp]),
    ("geye_get_boxes_features", C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                          _fp, C.c_longlong, C.c_float, C.POINTER(_fp), _ip,
                                          _ip]),
    ("geye_add_part", C.c_int, [C.c_void_p, _fp, _fp, _ip, _ip, C.c_int, C.c_int]),
    ("geye_graph_enable", C.c_int, [C.c_void_p, C.c_int]),
    ("geye_mfma_tuning", C.c_int, [C.c_void_p, C.c_int, _ip, C.c_char_p, C.c_int]),
    ("geye_mfma_schedule", C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    ("geye_serve_create", C.c_int, [C.c_int, C.c_void_p, C.POINTER(C.c_void_p)]),
    ("geye_serve_submit", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                    _llp]),
    ("geye_serve_submit_dev", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int,
                                        C.c_int, _llp]),
    ("geye_serve_wait", C.c_int, [C.c_void_p, C.c_longlong, C.POINTER(_fp), _ip, _ip, _dp]),
    ("geye_serve_stats", C.c_int, [C.c_void_p, _llp, _dp, _dp, _dp, _llp, _llp]),
    ("geye_serve_reset_stats", C.c_int, [C.c_void_p]),
    ("geye_serve_nms_fallbacks", C.c_int, [C.c_void_p, _llp]),
    ("geye_serve_nms_tie_repairs", C.c_int, [C.c_void_p, _llp]),
    ("geye_plan_bytes", C.c_int, [C.c_void_p, C.c_int, C.c_int, _llp]),
    ("geye_plan_bytes_desc", C.c_int, [C.c_void_p, C.c_int, C.c_int, _llp]),
    ("geye_serve_destroy", None, [C.c_void_p]),
    ("geye_free", None, [C.c_void_p]),
    ("geye_resize", C.c_int, [C.c_void_p, _fp, _ip, C.c_float, _fp, _ip]),
    ("geye_calc_feature", C.c_int, [C.c_void_p, _fp, C.c_int, C.c_int, C.c_int, _fp, _ip, _ip]),
    ("geye_pyramid", C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int,
                               C.c_int, C.c_int, _fp, C.c_longlong, _ip, _ip, _fp]),
    ("geye_fconv", C.c_int, [C.c_void_p, _fp, C.c_int, C.c_int, _fp, C.c_int, C.c_int, C.c_int,
                             C.c_int, _fp]),
    ("geye_dt", C.c_int, [C.c_void_p, _fp, C.c_float, C.c_float, C.c_float, C.c_float, _ip, _fp,
                          _ip, _ip]),
    ("geye_dt_window", C.c_int, [C.c_void_p, _fp, C.c_float, C.c_float, C.c_float, C.c_float, _ip,
                                 _fp, _ip, _ip]),
    ("geye_nms", C.c_int, [C.c_void_p, _fp, C.c_int, C.c_int, C.c_float, _fp, _ip]),
    ("geye_dt_noalloc", C.c_int, [C.c_void_p, C.c_void_p, C.c_float, C.c_float, C.c_float,
                                  C.c_float, _ip, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("geye_dev_alloc", C.c_int, [C.c_void_p, C.c_size_t, C.POINTER(C.c_void_p)]),
    ("geye_dev_free", C.c_int, [C.c_void_p, C.c_void_p]),
    ("geye_memcpy_h2d", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    ("geye_synchronize", C.c_int, [C.c_void_p]),
    ("geye_profile_enable", C.c_int, [C.c_void_p, C.c_int]),
    ("geye_kernel_stats", C.c_int, [C.c_void_p, C.c_char_p, C.c_int, _dp, _llp, C.c_int, _ip]),
    ("geye_nms_fallbacks", C.c_int, [C.c_void_p, _llp]),
    ("geye_nms_tie_repairs", C.c_int, [C.c_void_p, _llp]),
    ("geye_csv_values", C.c_int, [C.c_char_p, C.POINTER(_fp), _llp]),
    ("geye_dt_form", C.c_int, [C.c_void_p, C.c_int]),
]

_lib = None


def lib():
    """Load libgeye_amd.so (raises when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GeyeError(f"{LIB_PATH} is missing: run __graft_entry__.build() "
                            "(python -c 'import __graft_entry__ as g; g.build()')")
        L = C.CDLL(LIB_PATH)
        for name, res, args in _SIGS:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def exported_symbols() -> list[str]:
    return [s[0] for s in _SIGS]


def _check(rc: int) -> None:
    if rc != 0:
        msg = lib().geye_last_error()
        raise GeyeError(f"geye error {rc}: {msg.decode() if msg else ''}")


def _p(a, t):
    return a.ctypes.data_as(t)


def desc_from_model(model):
    """Build a ModelDesc (and the arrays it points into) from a synth.DPMModel."""
    keep = []

    def arr(a, dt):
        a = np.ascontiguousarray(a, dtype=dt)
        keep.append(a)
        return a

    m = model
    d = ModelDesc()
    d.numcomponent, d.sbin, d.interval = m.numcomponent, m.sbin, m.interval
    d.max_Y, d.max_X, d.padx, d.pady, d.maxpart = m.max_Y, m.max_X, m.padx, m.pady, m.maxpart
    d.ridx = _p(arr(m.ridx, np.int32), _ip)
    d.oidx = _p(arr(m.oidx, np.int32), _ip)
    d.offw = _p(arr(m.offw, np.float32), _fp)
    d.rsize = _p(arr(m.rsize, np.int32), _ip)
    d.numpart = _p(arr(m.numpart, np.int32), _ip)
    d.pidx = _p(arr(m.pidx, np.int32), _ip)
    d.didx = _p(arr(m.didx, np.int32), _ip)
    d.psize = _p(arr(m.psize, np.int32), _ip)
    d.DefL = int(m.deff.shape[0])
    d.deff = _p(arr(m.deff, np.float32), _fp)
    d.anchor = _p(arr(m.anchor, np.int32), _ip)
    d.NoR = m.NoR
    d.root_size = _p(arr(m.root_size, np.int32), _ip)
    d.rootfilter = _p(arr(np.concatenate(m.rootfilters), np.float32), _fp)
    d.root_off = _p(arr(np.cumsum([0] + [len(f) for f in m.rootfilters[:-1]]), np.int64), _llp)
    d.NoP = m.NoP
    if m.NoP:
        d.part_size = _p(arr(m.part_size, np.int32), _ip)
        d.partfilter = _p(arr(np.concatenate(m.partfilters), np.float32), _fp)
        d.part_off = _p(arr(np.cumsum([0] + [len(f) for f in m.partfilters[:-1]]), np.int64), _llp)
    if getattr(m, "lsq", None) is not None:
        d.lsq = _p(arr(lsq_flat(m), np.float32), _fp)
    return d, keep


def lsq_flat(m) -> np.ndarray:
    """Least-squares coefficients as the C ABI takes them: per component x1, y1, x2, y2."""
    if isinstance(m.lsq, (list, tuple)):
        return np.concatenate([np.asarray(a, np.float32).ravel() for a in m.lsq])
    return np.asarray(m.lsq, np.float32).ravel()


def level_plan_dims(rsize, sbin: int, interval: int, width: int, height: int) -> dict:
    """Host-only level plan (no device work)."""
    rs = np.ascontiguousarray(rsize, dtype=np.int32)
    nc = rs.size // 2
    ms, LEN = C.c_int(), C.c_int()
    _check(lib().geye_level_plan_dims(_p(rs, _ip), nc, sbin, interval, width, height,
                                      C.byref(ms), C.byref(LEN), None, None, None))
    n = LEN.value
    sc = np.zeros(n, np.float32)
    fh, fw = np.zeros(n, np.int32), np.zeros(n, np.int32)
    _check(lib().geye_level_plan_dims(_p(rs, _ip), nc, sbin, interval, width, height,
                                      C.byref(ms), C.byref(LEN), _p(sc, _fp), _p(fh, _ip),
                                      _p(fw, _ip)))
    return dict(max_scale=ms.value, LEN=n, scales=sc, feat_h=fh, feat_w=fw)


def nms_host(boxes: np.ndarray, overlap: float) -> np.ndarray:
    """The exact host NMS the device path falls back to on score ties."""
    boxes = np.ascontiguousarray(boxes, dtype=np.float32)
    n, GL = boxes.shape
    out = np.zeros_like(boxes)
    k = C.c_int()
    _check(lib().geye_nms_host(_p(boxes, _fp), n, GL, C.c_float(overlap), _p(out, _fp),
                               C.byref(k)))
    return out[:k.value].copy()


class ServeConfig(C.Structure):
    """geye_serve_config (include/geye_dpm.h)."""

    _fields_ = [("slots", C.c_int), ("hbm_budget", C.c_longlong), ("queue_cap", C.c_int),
                ("precision", C.c_int), ("thresh", C.c_float), ("overlap", C.c_float)]


class Server:
    """Serving loop over several camera streams (geye_serve_*)."""

    def __init__(self, device: int = 0, slots: int = 4, hbm_budget: int = 0, queue_cap: int = 256,
                 precision: str = "bf16", thresh: float = -0.5, overlap: float = 0.5):
        cfg = ServeConfig(slots, hbm_budget, queue_cap, PRECISIONS[precision], thresh, overlap)
        h = C.c_void_p()
        rc = lib().geye_serve_create(device, C.byref(cfg), C.byref(h))
        if rc != 0:
            raise GeyeError(f"geye_serve_create failed ({rc}): {lib().geye_last_error().decode()}")
        self.h = h

    def submit(self, model, frame: np.ndarray) -> int:
        H, W, _ = frame.shape
        frame = np.ascontiguousarray(frame, dtype=np.uint8)
        t = C.c_longlong()
        rc = lib().geye_serve_submit(self.h, model.h, frame.ctypes.data, W, H, 3 * W, C.byref(t))
        if rc == -6:
            raise BlockingIOError("serving queue full")
        _check(rc)
        return t.value

    def submit_dev(self, model, d_frame, W: int, H: int) -> int:
        """Device-resident frame (Context.dev_upload); keep it alive until waited for."""
        t = C.c_longlong()
        rc = lib().geye_serve_submit_dev(self.h, model.h, d_frame, W, H, 3 * W, C.byref(t))
        if rc == -6:
            raise BlockingIOError("serving queue full")
        _check(rc)
        return t.value

    def submit_models(self, models, frame: np.ndarray) -> list:
        """One frame for several models (one pyramid); one ticket per model, in order."""
        H, W, _ = frame.shape
        frame = np.ascontiguousarray(frame, dtype=np.uint8)
        hs = (C.c_void_p * len(models))(*[m.h for m in models])
        t = (C.c_longlong * len(models))()
        rc = lib().geye_serve_submit_models(self.h, hs, len(models), frame.ctypes.data, W, H,
                                            3 * W, t)
        if rc == -6:
            raise BlockingIOError("serving queue full")
        _check(rc)
        return list(t)

    def submit_models_dev(self, models, d_frame, W: int, H: int) -> list:
        hs = (C.c_void_p * len(models))(*[m.h for m in models])
        t = (C.c_longlong * len(models))()
        rc = lib().geye_serve_submit_models_dev(self.h, hs, len(models), d_frame, W, H, 3 * W, t)
        if rc == -6:
            raise BlockingIOError("serving queue full")
        _check(rc)
        return list(t)

    def wait(self, ticket: int):
        bp = _fp()
        n, GL = C.c_int(), C.c_int()
        lat = C.c_double()
        _check(lib().geye_serve_wait(self.h, ticket, C.byref(bp), C.byref(n), C.byref(GL),
                                     C.byref(lat)))
        rows = np.ctypeslib.as_array(bp, shape=(max(n.value, 1) * GL.value,))[: n.value * GL.value]
        rows = rows.copy().reshape(n.value, GL.value)
        lib().geye_free(bp)
        return rows, lat.value

    def stats(self) -> dict:
        c, ev, pk = C.c_longlong(), C.c_longlong(), C.c_longlong()
        p50, p99, fps = C.c_double(), C.c_double(), C.c_double()
        _check(lib().geye_serve_stats(self.h, C.byref(c), C.byref(p50), C.byref(p99), C.byref(fps),
                                      C.byref(pk), C.byref(ev)))
        return dict(completed=c.value, p50_ms=p50.value, p99_ms=p99.value, fps=fps.value,
                    peak_bytes=pk.value, evictions=ev.value)

    def reset_stats(self):
        _check(lib().geye_serve_reset_stats(self.h))

    def nms_fallbacks(self) -> int:
        n = C.c_longlong()
        _check(lib().geye_serve_nms_fallbacks(self.h, C.byref(n)))
        return int(n.value)

    def nms_tie_repairs(self) -> int:
        n = C.c_longlong()
        _check(lib().geye_serve_nms_tie_repairs(self.h, C.byref(n)))
        return int(n.value)

    def close(self):
        if self.h:
            lib().geye_serve_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Context:
    """One device context (geye_init / geye_shutdown)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        _check(lib().geye_init(device, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().geye_shutdown(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- models
    def model(self, model) -> "Model":
        d, keep = desc_from_model(model)
        h = C.c_void_p()
        _check(lib().geye_model_create(self.h, C.byref(d), C.byref(h)))
        return Model(self, h, getattr(model, "name", "model"))

    def load_model_csv(self, comp: str, root: str, part: str) -> "Model":
        h = C.c_void_p()
        _check(lib().geye_model_load_csv(self.h, comp.encode(), root.encode(), part.encode(),
                                         C.byref(h)))
        return Model(self, h, os.path.basename(comp))

    # ---- stage entry points (reference names)
    def resize(self, img: np.ndarray, scale: float) -> np.ndarray:
        """img float32 [3][w][h] (column-major planes) -> resized, same layout."""
        c, w, h = img.shape
        img = np.ascontiguousarray(img, dtype=np.float32)
        sd = np.array([h, w, c], np.int32)
        od = np.zeros(3, np.int32)
        _check(lib().geye_resize(self.h, _p(img, _fp), _p(sd, _ip), C.c_float(scale), None,
                                 _p(od, _ip)))
        out = np.zeros((od[2], od[1], od[0]), np.float32)
        _check(lib().geye_resize(self.h, _p(img, _fp), _p(sd, _ip), C.c_float(scale),
                                 _p(out, _fp), _p(od, _ip)))
        return out

    def calc_feature(self, img: np.ndarray, sbin: int) -> np.ndarray:
        c, w, h = img.shape
        img = np.ascontiguousarray(img, dtype=np.float32)
        F0, F1 = C.c_int(), C.c_int()
        _check(lib().geye_calc_feature(self.h, _p(img, _fp), h, w, sbin, None, C.byref(F0),
                                       C.byref(F1)))
        out = np.zeros((31, F1.value, F0.value), np.float32)
        if out.size:
            _check(lib().geye_calc_feature(self.h, _p(img, _fp), h, w, sbin, _p(out, _fp),
                                           C.byref(F0), C.byref(F1)))
        return out

    def calc_f_pyramid(self, frame: np.ndarray, sbin: int, interval: int, max_scale: int):
        H, W, _ = frame.shape
        frame = np.ascontiguousarray(frame, dtype=np.uint8)
        LEN = max_scale + interval
        fh, fw = np.zeros(LEN, np.int32), np.zeros(LEN, np.int32)
        sc = np.zeros(LEN, np.float32)
        _check(lib().geye_pyramid(self.h, frame.ctypes.data, W, H, 3 * W, sbin, interval,
                                  max_scale, None, 0, _p(fh, _ip), _p(fw, _ip), _p(sc, _fp)))
        total = int(sum(31 * int(a) * int(b) for a, b in zip(fh, fw)))
        buf = np.zeros(max(total, 1), np.float32)
        _check(lib().geye_pyramid(self.h, frame.ctypes.data, W, H, 3 * W, sbin, interval,
                                  max_scale, _p(buf, _fp), total, _p(fh, _ip), _p(fw, _ip),
                                  _p(sc, _fp)))
        feats, off = [], 0
        for l in range(LEN):
            n = 31 * int(fh[l]) * int(fw[l])
            feats.append(buf[off:off + n].reshape(31, int(fw[l]), int(fh[l])))
            off += n
        return feats, fh, fw, sc

    def fconvsMT(self, feat: np.ndarray, filters, h: int, w: int,
                 precision: str = "fp32") -> np.ndarray:
        """feat [31][F1][F0]; filters: list of flat [31][w][h] arrays -> [n][W][H]."""
        _, F1, F0 = feat.shape
        feat = np.ascontiguousarray(feat, dtype=np.float32)
        fl = np.ascontiguousarray(np.concatenate([np.asarray(f, np.float32).ravel() for f in filters]))
        n = len(filters)
        out = np.zeros((n, F1 - w + 1, F0 - h + 1), np.float32)
        _check(lib().geye_fconv(self.h, _p(feat, _fp), F0, F1, _p(fl, _fp), n, h, w,
                                PRECISIONS[precision], _p(out, _fp)))
        return out

    def dt(self, vals: np.ndarray, ax, bx, ay, by, window: bool = False):
        """vals [dims1][dims0] -> (M, Ix, Iy) in the same layout.  window=True runs the
        16-bit forms' windowed minimum (geye_dt_window) instead of the reference's order."""
        d1, d0 = vals.shape
        v = np.ascontiguousarray(vals, dtype=np.float32)
        dims = np.array([d0, d1], np.int32)
        M = np.zeros_like(v)
        Ix = np.zeros((d1, d0), np.int32)
        Iy = np.zeros((d1, d0), np.int32)
        fn = lib().geye_dt_window if window else lib().geye_dt
        _check(fn(self.h, _p(v, _fp), C.c_float(ax), C.c_float(bx), C.c_float(ay),
                  C.c_float(by), _p(dims, _ip), _p(M, _fp), _p(Ix, _ip), _p(Iy, _ip)))
        return M, Ix, Iy

    def add_part_calculation(self, score: np.ndarray, M: np.ndarray, ax: int, ay: int) -> np.ndarray:
        """score [R1][R0] -= M[(ay-1+2y) + (ax-1+2x)*P0] (M is [P1][P0]); returns a copy."""
        s = np.ascontiguousarray(score, dtype=np.float32).copy()
        Mv = np.ascontiguousarray(M, dtype=np.float32)
        rs = np.array([s.shape[1], s.shape[0]], np.int32)
        ps = np.array([Mv.shape[1], Mv.shape[0]], np.int32)
        _check(lib().geye_add_part(self.h, _p(s, _fp), _p(Mv, _fp), _p(rs, _ip), _p(ps, _ip),
                                   ax, ay))
        return s

    def nms(self, boxes: np.ndarray, overlap: float) -> np.ndarray:
        boxes = np.ascontiguousarray(boxes, dtype=np.float32)
        n, GL = boxes.shape
        out = np.zeros_like(boxes)
        k = C.c_int()
        _check(lib().geye_nms(self.h, _p(boxes, _fp), n, GL, C.c_float(overlap), _p(out, _fp),
                              C.byref(k)))
        return out[:k.value].copy()

    # ---- profiling
    def profile(self, on: bool = True):
        _check(lib().geye_profile_enable(self.h, 1 if on else 0))

    def kernel_stats(self) -> dict:
        n = C.c_int()
        cap = 64
        names = C.create_string_buffer(8192)
        ms = np.zeros(cap, np.float64)
        cnt = np.zeros(cap, np.int64)
        _check(lib().geye_kernel_stats(self.h, names, 8192, _p(ms, _dp), _p(cnt, _llp), cap,
                                       C.byref(n)))
        keys = names.value.decode().split("\n")[: n.value]
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(keys)}

    def dt_form(self, form: str = "auto") -> None:
        """DT form of whole-frame detection: "auto" (divide and conquer for fp32, windowed
        for the 16-bit forms), "divide" or "window" for every precision."""
        _check(lib().geye_dt_form(self.h, {"auto": 0, "divide": 1, "window": 2}[form]))

    def nms_fallbacks(self) -> int:
        """NMS calls on this context that were replayed on the host (score ties the device
        order could not settle)."""
        n = C.c_longlong()
        _check(lib().geye_nms_fallbacks(self.h, C.byref(n)))
        return int(n.value)

    def nms_tie_repairs(self) -> int:
        """NMS calls on this context that ran again with tied scores in the reference's
        quicksort order (set by the host's partial quicksort)."""
        n = C.c_longlong()
        _check(lib().geye_nms_tie_repairs(self.h, C.byref(n)))
        return int(n.value)

    def synchronize(self):
        _check(lib().geye_synchronize(self.h))

    # ---- device memory
    def mfma_tuning(self, precision: str = "bf16") -> dict:
        """Autotuned MFMA schedule of the last plan: the choice and each schedule's ms."""
        ch = C.c_int()
        buf = C.create_string_buffer(4096)
        _check(lib().geye_mfma_tuning(self.h, PRECISIONS[precision], C.byref(ch), buf, 4096))
        times = {}
        for item in buf.value.decode().split(";"):
            if "=" in item:
                k, v = item.split("=")
                times[k] = float(v)
        return {"choice": ch.value, "ms": times}

    def mfma_schedule(self, precision: str, choice: int) -> None:
        _check(lib().geye_mfma_schedule(self.h, PRECISIONS[precision], choice))

    def graphs(self, on: bool) -> None:
        """Replay the per-frame front as a HIP graph (default) or launch kernels one by one."""
        _check(lib().geye_graph_enable(self.h, int(bool(on))))

    def detect_models(self, models, frame: np.ndarray, thresh: float = -0.5,
                      overlap: float = 0.5, precision: str = "fp32", d_frame=None) -> list:
        """geye_detect_models: each model's NMS survivors on one frame, one shared pyramid
        (frame: host (H, W, 3) uint8, or its shape with d_frame a device copy)."""
        H, W, _ = frame.shape
        n = len(models)
        hs = (C.c_void_p * n)(*[m.h for m in models])
        bps = (C.c_void_p * n)()
        num = np.zeros(n, np.int32)
        GL = np.zeros(n, np.int32)
        if d_frame is None:
            frame = np.ascontiguousarray(frame, dtype=np.uint8)
            rc = lib().geye_detect_models(self.h, hs, n, frame.ctypes.data, W, H, 3 * W,
                                          C.c_float(thresh), C.c_float(overlap),
                                          PRECISIONS[precision], bps, _p(num, _ip), _p(GL, _ip))
        else:
            rc = lib().geye_detect_models_dev(self.h, hs, n, d_frame, W, H, 3 * W,
                                              C.c_float(thresh), C.c_float(overlap),
                                              PRECISIONS[precision], bps, _p(num, _ip),
                                              _p(GL, _ip))
        _check(rc)
        out = []
        for i in range(n):
            k = int(num[i]) * int(GL[i])
            bp = C.cast(bps[i], _fp)
            rows = np.ctypeslib.as_array(bp, shape=(max(k, 1),))[:k].copy() if bps[i] else \
                np.zeros(0, np.float32)
            lib().geye_free(bps[i])
            out.append(rows.reshape(int(num[i]), int(GL[i])))
        return out

    def detect_batch(self, models, frames, thresh: float = -0.5, overlap: float = 0.5,
                     precision: str = "fp32", d_frames=None) -> list:
        """geye_detect_batch: frames (list of (H, W, 3) uint8 host frames of one size, or their
        device copies in d_frames with frames giving the shape) through every model; returns
        result[frame][model], each frame's NMS survivors for that model."""
        nf, nm = len(frames), len(models)
        H, W, _ = frames[0].shape
        hs = (C.c_void_p * nm)(*[m.h for m in models])
        bps = (C.c_void_p * (nf * nm))()
        num = np.zeros(nf * nm, np.int32)
        GL = np.zeros(nf * nm, np.int32)
        if d_frames is None:
            keep = [np.ascontiguousarray(f, dtype=np.uint8) for f in frames]
            fp = (C.c_void_p * nf)(*[f.ctypes.data for f in keep])
            rc = lib().geye_detect_batch(self.h, hs, nm, fp, nf, W, H, 3 * W, C.c_float(thresh),
                                         C.c_float(overlap), PRECISIONS[precision], bps,
                                         _p(num, _ip), _p(GL, _ip))
        else:
            fp = (C.c_void_p * nf)(*[C.cast(d, C.c_void_p).value for d in d_frames])
            rc = lib().geye_detect_batch_dev(self.h, hs, nm, fp, nf, W, H, 3 * W,
                                             C.c_float(thresh), C.c_float(overlap),
                                             PRECISIONS[precision], bps, _p(num, _ip),
                                             _p(GL, _ip))
        _check(rc)
        out = []
        for f in range(nf):
            row = []
            for i in range(nm):
                k = f * nm + i
                cnt = int(num[k]) * int(GL[k])
                rows = (np.ctypeslib.as_array(C.cast(bps[k], _fp), shape=(max(cnt, 1),))[:cnt].copy()
                        if bps[k] else np.zeros(0, np.float32))
                lib().geye_free(bps[k])
                row.append(rows.reshape(int(num[k]), int(GL[k])))
            out.append(row)
        return out

    def dev_upload(self, arr: np.ndarray) -> C.c_void_p:
        arr = np.ascontiguousarray(arr)
        p = C.c_void_p()
        _check(lib().geye_dev_alloc(self.h, arr.nbytes, C.byref(p)))
        _check(lib().geye_memcpy_h2d(self.h, p, arr.ctypes.data, arr.nbytes))
        return p

    def dev_free(self, p):
        _check(lib().geye_dev_free(self.h, p))


class Model:
    def __init__(self, ctx: Context, h, name: str):
        self.ctx, self.h, self.name = ctx, h, name

    def free(self):
        if self.h:
            lib().geye_model_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def info(self) -> dict:
        v = [C.c_int() for _ in range(7)]
        _check(lib().geye_model_info(self.h, *[C.byref(x) for x in v]))
        keys = ["numcomponent", "sbin", "interval", "max_Y", "max_X", "NoR", "NoP"]
        return {k: x.value for k, x in zip(keys, v)}

    def level_plan(self, width: int, height: int) -> dict:
        ms, LEN = C.c_int(), C.c_int()
        _check(lib().geye_level_plan(self.h, width, height, C.byref(ms), C.byref(LEN), None,
                                     None, None))
        n = LEN.value
        sc = np.zeros(n, np.float32)
        fh, fw = np.zeros(n, np.int32), np.zeros(n, np.int32)
        _check(lib().geye_level_plan(self.h, width, height, C.byref(ms), C.byref(LEN),
                                     _p(sc, _fp), _p(fh, _ip), _p(fw, _ip)))
        return dict(max_scale=ms.value, LEN=n, scales=sc, feat_h=fh, feat_w=fw)

    def _rows(self, fn, *args) -> np.ndarray:
        bp = _fp()
        n, GL = C.c_int(), C.c_int()
        _check(fn(*args, C.byref(bp), C.byref(n), C.byref(GL)))
        try:
            k = n.value * GL.value
            out = np.ctypeslib.as_array(bp, shape=(max(k, 1),))[:k].copy()
        finally:
            lib().geye_free(bp)
        return out.reshape(n.value, GL.value)

    def get_boxes(self, frame: np.ndarray, thresh: float = -0.5,
                  precision: str = "fp32") -> np.ndarray:
        H, W, _ = frame.shape
        frame = np.ascontiguousarray(frame, dtype=np.uint8)
        return self._rows(lib().geye_get_boxes, self.ctx.h, self.h, frame.ctypes.data, W, H,
                          3 * W, C.c_float(thresh), PRECISIONS[precision])

    def car_detection(self, frame: np.ndarray, thresh: float = -0.5, overlap: float = 0.5,
                      precision: str = "fp32") -> np.ndarray:
        H, W, _ = frame.shape
        frame = np.ascontiguousarray(frame, dtype=np.uint8)
        return self._rows(lib().geye_detect, self.ctx.h, self.h, frame.ctypes.data, W, H, 3 * W,
                          C.c_float(thresh), C.c_float(overlap), PRECISIONS[precision])

    def detect_dev(self, d_frame, W: int, H: int, thresh: float = -0.5, overlap: float = 0.5,
                   precision: str = "fp32") -> np.ndarray:
        return self._rows(lib().geye_detect_dev, self.ctx.h, self.h, d_frame, W, H, 3 * W,
                          C.c_float(thresh), C.c_float(overlap), PRECISIONS[precision])

    def accumulated_score(self, W: int, H: int, thresh: float = -0.5,
                          a_score: np.ndarray | None = None) -> np.ndarray:
        """A_SCORE of the last detection on this context (get_boxes' side output).

        a_score: the caller's map as [W][H] float32 (column-major image), initialised to
        -100 like ini_ac_score (CT/detect.cpp:136-145) when not given; updated in place."""
        if a_score is None:
            a_score = np.full((W, H), -100.0, np.float32)
        assert a_score.dtype == np.float32 and a_score.shape == (W, H) and a_score.flags.c_contiguous
        _check(lib().geye_accumulated_score(self.ctx.h, C.c_float(thresh), _p(a_score, _fp)))
        return a_score

    def new_rects(self, rows: np.ndarray, W: int, H: int, ratio: float = 1.0) -> dict:
        """get_new_rects (CT/tracking.cpp:90-166) of NMS survivor rows."""
        rows = np.ascontiguousarray(rows, dtype=np.float32)
        n, GL = rows.shape if rows.ndim == 2 else (0, 0)
        out = dict(point=np.zeros((n, 4), np.int32), OR_point=np.zeros((n, 4), np.int32),
                   type=np.zeros(n, np.int32), scale=np.zeros(n, np.float32),
                   score=np.zeros(n, np.float32))
        if n:
            _check(lib().geye_new_rects(self.ctx.h, self.h, _p(rows, _fp), n, GL, W, H,
                                        C.c_float(ratio), _p(out["point"], _ip),
                                        _p(out["OR_point"], _ip), _p(out["type"], _ip),
                                        _p(out["scale"], _fp), _p(out["score"], _fp)))
        return out

    def get_boxes_features(self, feats: list, W: int, H: int, thresh: float = -0.5,
                           max_scale: int = 0) -> np.ndarray:
        """get_boxes on host features (one [31][F1][F0] array per pyramid level)."""
        flat = np.ascontiguousarray(np.concatenate([np.asarray(f, np.float32).ravel() for f in feats]))
        return self._rows(lib().geye_get_boxes_features, self.ctx.h, self.h, W, H, max_scale,
                          _p(flat, _fp), flat.size, C.c_float(thresh))

    def plan_bytes(self, W: int, H: int) -> int:
        """Working memory of one W x H frame, as a server's HBM budget counts it."""
        b = C.c_longlong()
        _check(lib().geye_plan_bytes(self.h, W, H, C.byref(b)))
        return b.value

    def last_scores(self) -> np.ndarray:
        total = C.c_longlong()
        _check(lib().geye_last_scores(self.ctx.h, None, 0, C.byref(total)))
        out = np.zeros(max(total.value, 1), np.float32)
        _check(lib().geye_last_scores(self.ctx.h, _p(out, _fp), total.value, C.byref(total)))
        return out[: total.value]
