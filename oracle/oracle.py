"""ctypes binding of the CPU oracle (oracle/libdpm_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, never by the product package.  The C++ restatement it wraps cites the
reference file:line of every stage (oracle/dpm_oracle.cpp).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdpm_oracle.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libdpm_ref.so")
REF_DIR = "/root/reference/CAR_TRACKING"

_fp = C.POINTER(C.c_float)
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_llp = C.POINTER(C.c_longlong)


class ModelDesc(C.Structure):
    """Mirror of ora_model_desc (oracle/dpm_oracle.h)."""

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


def build(quiet: bool = True) -> None:
    """Compile the oracle (and oracle/_ref when the reference tree is present)."""
    out = subprocess.DEVNULL if quiet else None
    subprocess.check_call(["make", "-C", HERE, "-s", "all"], stdout=out)
    if os.path.isdir(REF_DIR):
        subprocess.check_call(["make", "-C", HERE, "-s", "ref"], stdout=out)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.ora_ipl_to_float.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, _fp]
        L.ora_level_plan.argtypes = [_ip, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _ip, _ip,
                                     _fp, _ip, _ip, _ip, _ip, _ip]
        L.ora_resize.argtypes = [_fp, _ip, C.c_float, _fp, _ip]
        L.ora_resize.restype = C.c_longlong
        L.ora_resize_table.argtypes = [C.c_int, C.c_int, C.c_int, _ip, _ip, _fp, C.c_int]
        L.ora_calc_feature.argtypes = [_fp, C.c_int, C.c_int, C.c_int, _fp, _ip, _ip]
        L.ora_conv_f32.argtypes = [_fp, C.c_int, C.c_int, _fp, C.c_int, C.c_int, _fp]
        L.ora_conv_f64.argtypes = [_fp, C.c_int, C.c_int, _fp, C.c_int, C.c_int, _dp]
        L.ora_dt_f32.argtypes = [_fp, C.c_float, C.c_float, C.c_float, C.c_float, _ip, _fp,
                                 _ip, _ip]
        L.ora_dt_f64.argtypes = [_dp, C.c_double, C.c_double, C.c_double, C.c_double, _ip,
                                 _dp, _ip, _ip]
        L.ora_dt_passes_f32.argtypes = [_fp, C.c_float, C.c_float, C.c_float, C.c_float, _ip,
                                        _fp, _ip, _fp, _ip]
        L.ora_add_part_f32.argtypes = [_fp, _fp, _ip, _ip, C.c_int, C.c_int]
        L.ora_nms.argtypes = [_fp, C.c_int, C.c_int, C.c_float, _fp]
        L.ora_quicksort.argtypes = [_fp, _ip, C.c_int]
        L.ora_detect.argtypes = [C.POINTER(ModelDesc), C.c_void_p, C.c_int, C.c_int, C.c_int,
                                 C.c_float, C.c_int, C.c_int, C.POINTER(_fp), _ip]
        L.ora_detect.restype = C.c_longlong
        L.ora_detect_maps.argtypes = [C.POINTER(ModelDesc), C.c_void_p, C.c_int, C.c_int,
                                      C.c_int, C.c_int, C.c_int, C.POINTER(C.c_void_p),
                                      C.POINTER(_ip), _ip]
        L.ora_detect_maps.restype = C.c_longlong
        L.ora_rows_one_map.argtypes = [C.POINTER(ModelDesc), C.c_int, C.c_float, _fp, C.c_int,
                                       C.c_int, _ip, _ip, _ip, C.c_float, _fp, C.c_longlong,
                                       _ip]
        L.ora_rows_one_map.restype = C.c_longlong
        L.ora_pyramid.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                  C.c_int, C.c_int, C.POINTER(_fp), _ip, _ip, _fp]
        L.ora_pyramid.restype = C.c_longlong
        L.ora_free.argtypes = [C.c_void_p]
        L.ora_detect_ascore.argtypes = [C.POINTER(ModelDesc), C.c_void_p, C.c_int, C.c_int,
                                        C.c_int, C.c_float, C.c_int, _fp]
        L.ora_calc_a_score.argtypes = [_fp, C.c_int, C.c_int, _fp, C.c_int, C.c_int, _ip,
                                       C.c_int, C.c_int, C.c_float]
        L.ora_ascore_finish.argtypes = [_fp, C.c_longlong, C.c_float]
        L.ora_new_rects.argtypes = [C.POINTER(ModelDesc), _fp, C.c_int, C.c_int, C.c_int, C.c_int,
                                    C.c_float, _ip, _ip, _ip, _fp, _fp]
        L.ora_load_csv.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.POINTER(ModelDesc)]
        L.ora_load_csv.restype = C.c_void_p
        L.ora_model_free.argtypes = [C.c_void_p]
        _lib = L
    return _lib


def ref_lib():
    """The reference's own resize.cc/dt.cpp (oracle/_ref), or None when not built."""
    global _ref
    if _ref is None and os.path.exists(REF_LIB_PATH):
        R = C.CDLL(REF_LIB_PATH)
        R.ref_resize.argtypes = [_fp, _ip, C.c_float, _fp, _ip]
        R.ref_resize.restype = C.c_longlong
        R.ref_dt.argtypes = [_fp, C.c_float, C.c_float, C.c_float, C.c_float, _ip, _fp, _ip,
                             _ip]
        R.ref_add_part.argtypes = [_fp, _fp, _ip, _ip, C.c_int, C.c_int]
        _ref = R
    return _ref


def _p(a, t):
    return a.ctypes.data_as(t)


class _DescKeeper:
    """Holds the numpy arrays a ModelDesc points into."""

    def __init__(self, model):
        m = model
        self.arrs = []

        def keep(a, dt):
            a = np.ascontiguousarray(a, dtype=dt)
            self.arrs.append(a)
            return a

        d = ModelDesc()
        d.numcomponent, d.sbin, d.interval = m.numcomponent, m.sbin, m.interval
        d.max_Y, d.max_X, d.padx, d.pady = m.max_Y, m.max_X, m.padx, m.pady
        d.maxpart = m.maxpart
        d.ridx = _p(keep(m.ridx, np.int32), _ip)
        d.oidx = _p(keep(m.oidx, np.int32), _ip)
        d.offw = _p(keep(m.offw, np.float32), _fp)
        d.rsize = _p(keep(m.rsize, np.int32), _ip)
        d.numpart = _p(keep(m.numpart, np.int32), _ip)
        d.pidx = _p(keep(m.pidx, np.int32), _ip)
        d.didx = _p(keep(m.didx, np.int32), _ip)
        d.psize = _p(keep(m.psize, np.int32), _ip)
        d.DefL = int(m.deff.shape[0])
        d.deff = _p(keep(m.deff, np.float32), _fp)
        d.anchor = _p(keep(m.anchor, np.int32), _ip)
        d.NoR = m.NoR
        d.root_size = _p(keep(m.root_size, np.int32), _ip)
        roff = np.cumsum([0] + [len(f) for f in m.rootfilters[:-1]]).astype(np.int64)
        d.rootfilter = _p(keep(np.concatenate(m.rootfilters), np.float32), _fp)
        d.root_off = _p(keep(roff, np.int64), _llp)
        d.NoP = m.NoP
        if m.NoP:
            poff = np.cumsum([0] + [len(f) for f in m.partfilters[:-1]]).astype(np.int64)
            d.part_size = _p(keep(m.part_size, np.int32), _ip)
            d.partfilter = _p(keep(np.concatenate(m.partfilters), np.float32), _fp)
            d.part_off = _p(keep(poff, np.int64), _llp)
        if getattr(m, "lsq", None) is not None:
            d.lsq = _p(keep(np.concatenate([np.asarray(a, np.float32).ravel() for a in m.lsq]),
                            np.float32), _fp)
        self.desc = d


def level_plan(model, width: int, height: int) -> dict:
    L = lib()
    ms, LEN = C.c_int(), C.c_int()
    rs = np.ascontiguousarray(model.rsize, dtype=np.int32)
    L.ora_level_plan(_p(rs, _ip), model.numcomponent, model.sbin, model.interval, width,
                     height, C.byref(ms), C.byref(LEN), None, None, None, None, None, None)
    n = LEN.value
    sc = np.zeros(n, np.float32)
    ih, iw, sb, fh, fw = (np.zeros(n, np.int32) for _ in range(5))
    L.ora_level_plan(_p(rs, _ip), model.numcomponent, model.sbin, model.interval, width,
                     height, C.byref(ms), C.byref(LEN), _p(sc, _fp), _p(ih, _ip), _p(iw, _ip),
                     _p(sb, _ip), _p(fh, _ip), _p(fw, _ip))
    return dict(max_scale=ms.value, LEN=n, scales=sc, img_h=ih, img_w=iw, sbin=sb, feat_h=fh,
                feat_w=fw)


def ipl_to_float(frame: np.ndarray) -> np.ndarray:
    H, W, _ = frame.shape
    out = np.zeros((3, W, H), np.float32)
    lib().ora_ipl_to_float(frame.ctypes.data, W, H, 3 * W, _p(out, _fp))
    return out


def resize(img: np.ndarray, scale: float, use_ref: bool = False) -> np.ndarray:
    """img: float32 [c][x][y] (column-major planes), returns the same layout."""
    c, w, h = img.shape
    img = np.ascontiguousarray(img, dtype=np.float32)
    sd = np.array([h, w, c], np.int32)
    od = np.zeros(3, np.int32)
    fn = ref_lib().ref_resize if use_ref else lib().ora_resize
    n = fn(_p(img, _fp), _p(sd, _ip), C.c_float(scale), None, _p(od, _ip)) if not use_ref else None
    if use_ref:
        # the reference allocates its own output; size it from the formula first
        lib().ora_resize(_p(img, _fp), _p(sd, _ip), C.c_float(scale), None, _p(od, _ip))
        out = np.zeros((od[2], od[1], od[0]), np.float32)
        fn(_p(img, _fp), _p(sd, _ip), C.c_float(scale), _p(out, _fp), _p(od, _ip))
        return out
    out = np.zeros((od[2], od[1], od[0]), np.float32)
    fn(_p(img, _fp), _p(sd, _ip), C.c_float(scale), _p(out, _fp), _p(od, _ip))
    return out


def resize_table(sheight: int, dheight: int, width: int = 1):
    cap = sheight + 2 * dheight + 8
    si, di = np.zeros(cap, np.int32), np.zeros(cap, np.int32)
    al = np.zeros(cap, np.float32)
    n = lib().ora_resize_table(sheight, dheight, width, _p(si, _ip), _p(di, _ip), _p(al, _fp), cap)
    return si[:n], di[:n], al[:n]


def calc_feature(img: np.ndarray, sbin: int) -> np.ndarray:
    """img float32 [3][w][h] -> features float32 [31][F1][F0]."""
    c, w, h = img.shape
    img = np.ascontiguousarray(img, dtype=np.float32)
    F0, F1 = C.c_int(), C.c_int()
    lib().ora_calc_feature(_p(img, _fp), h, w, sbin, None, C.byref(F0), C.byref(F1))
    out = np.zeros((31, F1.value, F0.value), np.float32)
    lib().ora_calc_feature(_p(img, _fp), h, w, sbin, _p(out, _fp), C.byref(F0), C.byref(F1))
    return out


def conv(feat: np.ndarray, filt: np.ndarray, h: int, w: int, f64: bool = False) -> np.ndarray:
    """feat [31][F1][F0] (float32), filt flat [31][w][h] -> out [W][H] (column-major)."""
    _, F1, F0 = feat.shape
    feat = np.ascontiguousarray(feat, dtype=np.float32)
    filt = np.ascontiguousarray(filt, dtype=np.float32)
    H, W = F0 - h + 1, F1 - w + 1
    if f64:
        out = np.zeros((W, H), np.float64)
        lib().ora_conv_f64(_p(feat, _fp), F0, F1, _p(filt, _fp), h, w, _p(out, _dp))
    else:
        out = np.zeros((W, H), np.float32)
        lib().ora_conv_f32(_p(feat, _fp), F0, F1, _p(filt, _fp), h, w, _p(out, _fp))
    return out


def dt(vals: np.ndarray, ax, bx, ay, by, use_ref: bool = False, f64: bool = False):
    """vals [dims1][dims0] column-major map (dims0 fastest).  Returns M, Ix, Iy."""
    d1, d0 = vals.shape
    dims = np.array([d0, d1], np.int32)
    Ix = np.zeros((d1, d0), np.int32)
    Iy = np.zeros((d1, d0), np.int32)
    if f64:
        v = np.ascontiguousarray(vals, dtype=np.float64)
        M = np.zeros((d1, d0), np.float64)
        lib().ora_dt_f64(_p(v, _dp), ax, bx, ay, by, _p(dims, _ip), _p(M, _dp), _p(Ix, _ip),
                         _p(Iy, _ip))
        return M, Ix, Iy
    v = np.ascontiguousarray(vals, dtype=np.float32)
    M = np.zeros((d1, d0), np.float32)
    fn = ref_lib().ref_dt if use_ref else lib().ora_dt_f32
    fn(_p(v, _fp), C.c_float(ax), C.c_float(bx), C.c_float(ay), C.c_float(by), _p(dims, _ip),
       _p(M, _fp), _p(Ix, _ip), _p(Iy, _ip))
    return M, Ix, Iy


def dt_passes(vals: np.ndarray, ax, bx, ay, by):
    d1, d0 = vals.shape
    dims = np.array([d0, d1], np.int32)
    v = np.ascontiguousarray(vals, dtype=np.float32)
    tmpM = np.zeros_like(v)
    M = np.zeros_like(v)
    tIy = np.zeros((d1, d0), np.int32)
    tIx = np.zeros((d1, d0), np.int32)
    lib().ora_dt_passes_f32(_p(v, _fp), C.c_float(ax), C.c_float(bx), C.c_float(ay),
                            C.c_float(by), _p(dims, _ip), _p(tmpM, _fp), _p(tIy, _ip),
                            _p(M, _fp), _p(tIx, _ip))
    return tmpM, tIy, M, tIx


def add_part(score: np.ndarray, M: np.ndarray, ax: int, ay: int, use_ref: bool = False):
    """add_part_calculation (CT/dt.cpp:38-55) on a copy of score [R1][R0] with the part's DT
    map M [P1][P0] (both column-major); returns the updated copy."""
    R1, R0 = score.shape
    P1, P0 = M.shape
    if ay - 1 + 2 * (R0 - 1) >= P0 or ax - 1 + 2 * (R1 - 1) >= P1 or ax < 1 or ay < 1:
        raise ValueError("anchor reads outside the part map")
    out = np.ascontiguousarray(score, dtype=np.float32).copy()
    Mc = np.ascontiguousarray(M, dtype=np.float32)
    rs = np.array([R0, R1], np.int32)
    ps = np.array([P0, P1], np.int32)
    fn = ref_lib().ref_add_part if use_ref else lib().ora_add_part_f32
    fn(_p(out, _fp), _p(Mc, _fp), _p(rs, _ip), _p(ps, _ip), ax, ay)
    return out


def nms(boxes: np.ndarray, overlap: float) -> np.ndarray:
    boxes = np.ascontiguousarray(boxes, dtype=np.float32)
    n, GL = boxes.shape
    out = np.zeros_like(boxes)
    k = lib().ora_nms(_p(boxes, _fp), n, GL, C.c_float(overlap), _p(out, _fp))
    return out[:k].copy()


def quicksort(scores: np.ndarray):
    s = np.ascontiguousarray(scores, dtype=np.float32).copy()
    o = np.arange(len(s), dtype=np.int32)
    lib().ora_quicksort(_p(s, _fp), _p(o, _ip), len(s))
    return s, o


def detect(model, frame: np.ndarray, thresh: float = -0.5, precision: int = 0,
           threads: int = 1) -> np.ndarray:
    """Candidate boxes (before NMS), rows of GL floats, in the reference's order."""
    keep = _DescKeeper(model)
    H, W, _ = frame.shape
    frame = np.ascontiguousarray(frame)
    bp = _fp()
    GL = C.c_int()
    n = lib().ora_detect(C.byref(keep.desc), frame.ctypes.data, W, H, 3 * W, C.c_float(thresh),
                         precision, threads, C.byref(bp), C.byref(GL))
    if n < 0:
        raise RuntimeError("oracle detect failed")
    out = np.ctypeslib.as_array(bp, shape=(max(n, 1) * GL.value,))[: n * GL.value].copy()
    lib().ora_free(bp)
    return out.reshape(n, GL.value)


def rows_one_map(model, jj: int, scale: float, score: np.ndarray, Ix: list, Iy: list,
                 thresh: float) -> np.ndarray:
    """Candidate rows of component jj's score map score[R1][R0] (column-major) at one level
    of the given scale, with the part argmin maps Ix[k], Iy[k] ([P1][P0] each)."""
    keep = _DescKeeper(model)
    score = np.ascontiguousarray(score, dtype=np.float32)
    R1, R0 = score.shape
    pd = np.array([[a.shape[1], a.shape[0]] for a in Ix], np.int32).ravel()
    ix = np.ascontiguousarray(np.concatenate([np.asarray(a, np.int32).ravel() for a in Ix]))
    iy = np.ascontiguousarray(np.concatenate([np.asarray(a, np.int32).ravel() for a in Iy]))
    GL = C.c_int()
    n = lib().ora_rows_one_map(C.byref(keep.desc), jj, C.c_float(scale), _p(score, _fp), R0, R1,
                               _p(ix, _ip), _p(iy, _ip), _p(pd, _ip), C.c_float(thresh), None, 0,
                               C.byref(GL))
    out = np.zeros((max(n, 1), GL.value), np.float32)
    lib().ora_rows_one_map(C.byref(keep.desc), jj, C.c_float(scale), _p(score, _fp), R0, R1,
                           _p(ix, _ip), _p(iy, _ip), _p(pd, _ip), C.c_float(thresh),
                           _p(out, _fp), n, C.byref(GL))
    return out[:n]


def score_maps(model, frame: np.ndarray, precision: int = 0, threads: int = 1):
    """List of (level, comp, map[R1][R0]) root+bias-parts score maps."""
    keep = _DescKeeper(model)
    H, W, _ = frame.shape
    frame = np.ascontiguousarray(frame)
    mp = C.c_void_p()
    dp = _ip()
    nm = C.c_int()
    total = lib().ora_detect_maps(C.byref(keep.desc), frame.ctypes.data, W, H, 3 * W, precision,
                                  threads, C.byref(mp), C.byref(dp), C.byref(nm))
    dt_ = np.float64 if precision == 1 else np.float32
    buf = np.ctypeslib.as_array(C.cast(mp, C.POINTER(C.c_double if precision == 1 else C.c_float)),
                                shape=(max(total, 1),))[:total].astype(dt_)
    dims = np.ctypeslib.as_array(dp, shape=(max(nm.value, 1) * 4,))[: nm.value * 4].reshape(-1, 4).copy()
    lib().ora_free(mp)
    lib().ora_free(dp)
    out, off = [], 0
    for lv, jj, r0, r1 in dims:
        n = int(r0) * int(r1)
        out.append((int(lv), int(jj), buf[off:off + n].reshape(int(r1), int(r0))))
        off += n
    return out


def ascore(model, frame: np.ndarray, thresh: float = -0.5, ac: np.ndarray | None = None,
           threads: int = 8) -> np.ndarray:
    """A_SCORE of one FP32 detection; ac is [W][H] float32 (column-major image), -100 if None."""
    keep = _DescKeeper(model)
    H, W, _ = frame.shape
    frame = np.ascontiguousarray(frame)
    if ac is None:
        ac = np.full((W, H), -100.0, np.float32)
    ac = np.ascontiguousarray(ac, dtype=np.float32)
    lib().ora_detect_ascore(C.byref(keep.desc), frame.ctypes.data, W, H, 3 * W, C.c_float(thresh),
                            threads, _p(ac, _fp))
    return ac


def calc_a_score(ac: np.ndarray, score: np.ndarray, rsize, scale: float, padx: int = 0,
                 pady: int = 0) -> np.ndarray:
    """One calc_a_score call: ac [IWID][IHEI], score [R1][R0] (both column-major)."""
    ac = np.ascontiguousarray(ac, dtype=np.float32)
    score = np.ascontiguousarray(score, dtype=np.float32)
    IWID, IHEI = ac.shape
    R1, R0 = score.shape
    rs = np.ascontiguousarray(rsize, dtype=np.int32)
    lib().ora_calc_a_score(_p(ac, _fp), IHEI, IWID, _p(score, _fp), R0, R1, _p(rs, _ip), padx, pady,
                           C.c_float(scale))
    return ac


def ascore_finish(ac: np.ndarray, thresh: float) -> np.ndarray:
    ac = np.ascontiguousarray(ac, dtype=np.float32)
    lib().ora_ascore_finish(_p(ac, _fp), ac.size, C.c_float(thresh))
    return ac


def new_rects(model, rows: np.ndarray, W: int, H: int, ratio: float = 1.0) -> dict:
    """get_new_rects (CT/tracking.cpp:90-166) of NMS survivor rows."""
    keep = _DescKeeper(model)
    rows = np.ascontiguousarray(rows, dtype=np.float32)
    n = rows.shape[0]
    GL = rows.shape[1] if rows.ndim == 2 else 0
    out = dict(point=np.zeros((n, 4), np.int32), OR_point=np.zeros((n, 4), np.int32),
               type=np.zeros(n, np.int32), scale=np.zeros(n, np.float32),
               score=np.zeros(n, np.float32))
    if n:
        lib().ora_new_rects(C.byref(keep.desc), _p(rows, _fp), n, GL, W, H, C.c_float(ratio),
                            _p(out["point"], _ip), _p(out["OR_point"], _ip), _p(out["type"], _ip),
                            _p(out["scale"], _fp), _p(out["score"], _fp))
    return out


def write_csv_model(model, directory: str, prefix: str = "model") -> tuple:
    """Write a DPMModel in the reference's CSV model format (CT/load_model.cpp:57-530):
    returns the (comp, root, part) paths.  Floats are written with 9 significant digits, so
    they parse back to the same float32 values."""
    import os as _os

    m = model
    f9 = lambda v: "%.9g" % float(np.float32(v))
    comp = [m.numcomponent, m.sbin, m.interval, m.max_Y, m.max_X]
    for c in range(m.numcomponent):
        comp += [int(m.ridx[c]) + 1, int(m.oidx[c]) + 1, f9(m.offw[c]), int(m.rsize[c][0]),
                 int(m.rsize[c][1]), int(m.numpart[c])]
        for k in range(int(m.numpart[c])):
            comp += [int(m.pidx[c][k]) + 1, int(m.didx[c][k]) + 1, int(m.psize[c][k][0]),
                     int(m.psize[c][k][1])]
    comp.append(int(m.deff.shape[0]))
    for q in range(m.deff.shape[0]):
        comp += [f9(v) for v in m.deff[q]] + [int(m.anchor[q][0]), int(m.anchor[q][1])]
    for c in range(m.numcomponent):
        comp += [f9(v) for v in np.asarray(m.lsq[c], np.float32).ravel()]
    root = [m.NoR]
    for r in range(m.NoR):
        root += [int(v) for v in m.root_size[r]] + [f9(v) for v in m.rootfilters[r]]
    part = [m.NoP]
    for p in range(m.NoP):
        part += [int(v) for v in m.part_size[p]] + [f9(v) for v in m.partfilters[p]] + [0]
    paths = []
    for name, vals in (("comp", comp), ("root", root), ("part", part)):
        path = _os.path.join(directory, f"{prefix}_{name}.csv")
        with open(path, "w") as fh:
            fh.write(",\n".join(str(v) for v in vals) + ",\n")
        paths.append(path)
    return tuple(paths)


def pyramid(frame: np.ndarray, sbin: int, interval: int, max_scale: int, threads: int = 1):
    H, W, _ = frame.shape
    frame = np.ascontiguousarray(frame)
    LEN = max_scale + interval
    fh, fw = np.zeros(LEN, np.int32), np.zeros(LEN, np.int32)
    sc = np.zeros(LEN, np.float32)
    fp = _fp()
    total = lib().ora_pyramid(frame.ctypes.data, W, H, 3 * W, sbin, interval, max_scale,
                              threads, C.byref(fp), _p(fh, _ip), _p(fw, _ip), _p(sc, _fp))
    buf = np.ctypeslib.as_array(fp, shape=(max(total, 1),))[:total].copy()
    lib().ora_free(fp)
    feats, off = [], 0
    for l in range(LEN):
        n = 31 * int(fh[l]) * int(fw[l])
        feats.append(buf[off:off + n].reshape(31, int(fw[l]), int(fh[l])))
        off += n
    return feats, fh, fw, sc


def load_csv_model(comp: str, root: str, part: str):
    """Parse the reference CSV model format with the oracle's loader into a DPMModel."""
    from geye_amd.synth import DPMModel

    d = ModelDesc()
    h = lib().ora_load_csv(comp.encode(), root.encode(), part.encode(), C.byref(d))
    if not h:
        raise FileNotFoundError(comp)
    try:
        NC, MP = d.numcomponent, d.maxpart
        arr = lambda p, n, dt: np.ctypeslib.as_array(p, shape=(n,)).astype(dt).copy()
        root_size = arr(d.root_size, 3 * d.NoR, np.int32).reshape(-1, 3)
        part_size = arr(d.part_size, 3 * d.NoP, np.int32).reshape(-1, 3)
        roff = arr(d.root_off, d.NoR, np.int64)
        poff = arr(d.part_off, d.NoP, np.int64)
        rtot = int(sum(int(a) * int(b) * int(c) for a, b, c in root_size))
        ptot = int(sum(int(a) * int(b) * int(c) for a, b, c in part_size))
        rf = arr(d.rootfilter, rtot, np.float32)
        pf = arr(d.partfilter, ptot, np.float32)
        rootfilters = [rf[roff[i]:roff[i] + int(np.prod(root_size[i]))].copy() for i in range(d.NoR)]
        partfilters = [pf[poff[i]:poff[i] + int(np.prod(part_size[i]))].copy() for i in range(d.NoP)]
        nps = arr(d.numpart, NC, np.int32)
        mls = [1 + 2 * (1 + int(n)) for n in nps]
        lflat = arr(d.lsq, 4 * sum(mls), np.float32)
        lsq, o = [], 0
        for ml in mls:
            lsq.append(lflat[o:o + 4 * ml].reshape(4, ml).copy())
            o += 4 * ml
        return DPMModel(
            numcomponent=NC, sbin=d.sbin, interval=d.interval, max_Y=d.max_Y, max_X=d.max_X,
            ridx=arr(d.ridx, NC, np.int32), oidx=arr(d.oidx, NC, np.int32),
            offw=arr(d.offw, NC, np.float32), rsize=arr(d.rsize, 2 * NC, np.int32).reshape(NC, 2),
            numpart=arr(d.numpart, NC, np.int32), pidx=arr(d.pidx, NC * MP, np.int32).reshape(NC, MP),
            didx=arr(d.didx, NC * MP, np.int32).reshape(NC, MP),
            psize=arr(d.psize, NC * MP * 2, np.int32).reshape(NC, MP, 2),
            deff=arr(d.deff, 4 * d.DefL, np.float32).reshape(-1, 4),
            anchor=arr(d.anchor, 2 * d.DefL, np.int32).reshape(-1, 2),
            root_size=root_size, rootfilters=rootfilters, part_size=part_size,
            partfilters=partfilters, lsq=lsq, name=os.path.basename(comp).split("_")[0])
    finally:
        lib().ora_model_free(h)
