"""Seeded synthetic inputs for the DPM detection path: frames and models.

Frames: uniform uint8 noise per channel per pixel, then 20 axis-aligned rectangles of
constant random colour with sides in [20, 200] px, clipped to the frame, stored BGR
interleaved with widthStep = 3*W (the IplImage fields the reference reads,
CT/resize_GPU.cc:75-90).

Models: the synthetic car- and person-shaped DPM models of the benchmark configurations.
Filters are stored in the reference's layout [31][w][h] (CT/load_model.cpp:250-390); the
32nd HOG channel of the 32-channel description is a zero pad that the convolution never
reads (CT/fconvsMT.cpp:414), so it is not stored.  Deformation coefficients are
(ax, bx, ay, by) = (0.01, 0, 0.01, 0) and anchors form a 2x4 grid inside the 2x-resolution
root window, with ax, ay >= 2 so that the part-box probe (CT/get_boxes.cpp:246-248) stays
inside the part map.

All randomness is counter-based SplitMix64, so a (seed, stream) pair fixes every value.
"""
from __future__ import annotations

import dataclasses

import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int, stream: int = 0) -> np.ndarray:
    """n SplitMix64 outputs of (seed, stream), as uint64."""
    with np.errstate(over="ignore"):
        base = np.uint64((seed * 0x100000001B3 + stream * 0xD6E8FEB86659FD93) & 0xFFFFFFFFFFFFFFFF)
        z = base + (np.arange(1, n + 1, dtype=np.uint64) * _GAMMA)
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, n: int, stream: int = 0) -> np.ndarray:
    """Uniform doubles in (0, 1) from the top 53 bits."""
    z = splitmix64(seed, n, stream) >> np.uint64(11)
    return (z.astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)


def randint(seed: int, n: int, lo: int, hi: int, stream: int = 0) -> np.ndarray:
    """Integers uniform in [lo, hi] (inclusive)."""
    z = splitmix64(seed, n, stream)
    return (lo + (z % np.uint64(hi - lo + 1))).astype(np.int64)


def normal(seed: int, n: int, std: float, stream: int = 0) -> np.ndarray:
    """Box-Muller normals N(0, std^2), float64."""
    m = (n + 1) // 2
    u1 = uniform01(seed, m, stream * 2 + 1)
    u2 = uniform01(seed, m, stream * 2 + 2)
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.concatenate([r * np.cos(2.0 * np.pi * u2), r * np.sin(2.0 * np.pi * u2)])[:n]
    return z * std


def make_frame(width: int, height: int, seed: int = 1, nrect: int = 20) -> np.ndarray:
    """Seeded rectangle-noise frame, BGR interleaved uint8 of shape (H, W, 3)."""
    img = (splitmix64(seed, width * height * 3, stream=1) >> np.uint64(56)).astype(np.uint8)
    img = img.reshape(height, width, 3)
    sides = randint(seed, 2 * nrect, 20, 200, stream=2).reshape(nrect, 2)
    corner = uniform01(seed, 2 * nrect, stream=3).reshape(nrect, 2)
    colour = randint(seed, 3 * nrect, 0, 255, stream=4).reshape(nrect, 3)
    for k in range(nrect):
        rw, rh = int(sides[k, 0]), int(sides[k, 1])
        x0 = int(corner[k, 0] * width)
        y0 = int(corner[k, 1] * height)
        img[y0:min(y0 + rh, height), x0:min(x0 + rw, width), :] = colour[k].astype(np.uint8)
    return np.ascontiguousarray(img)


@dataclasses.dataclass
class DPMModel:
    """A DPM model in the reference's field layout (CT/MODEL_info.h:75-142), 0-based."""

    numcomponent: int
    sbin: int
    interval: int
    max_Y: int
    max_X: int
    ridx: np.ndarray  # [C] int32
    oidx: np.ndarray  # [C] int32
    offw: np.ndarray  # [C] float32
    rsize: np.ndarray  # [C, 2] int32 (Y, X)
    numpart: np.ndarray  # [C] int32
    pidx: np.ndarray  # [C, MP] int32
    didx: np.ndarray  # [C, MP] int32
    psize: np.ndarray  # [C, MP, 2] int32
    deff: np.ndarray  # [DefL, 4] float32 (ax, bx, ay, by)
    anchor: np.ndarray  # [DefL, 2] int32 (x, y)
    root_size: np.ndarray  # [NoR, 3] int32 (h, w, 31)
    rootfilters: list  # NoR arrays, float32, flat [31][w][h]
    part_size: np.ndarray  # [NoP, 3] int32
    partfilters: list  # NoP arrays, float32, flat [31][w][h]
    padx: int = 0
    pady: int = 0
    lsq: np.ndarray | None = None  # [C, 4, 1+2(1+np)] float32 (x1, y1, x2, y2)
    name: str = "model"

    @property
    def NoR(self) -> int:
        return len(self.rootfilters)

    @property
    def NoP(self) -> int:
        return len(self.partfilters)

    @property
    def maxpart(self) -> int:
        return int(self.pidx.shape[1])


def _synthetic(roots, nparts: int, seed: int, name: str) -> DPMModel:
    C = len(roots)
    NoP = C * nparts
    rsize = np.array(roots, dtype=np.int32)
    offw = normal(seed, C, 1.0, stream=10).astype(np.float32)
    rootfilters = []
    for c, (h, w) in enumerate(roots):
        rootfilters.append(normal(seed, 31 * w * h, 0.05, stream=100 + c).astype(np.float32))
    partfilters = [normal(seed, 31 * 6 * 6, 0.05, stream=1000 + p).astype(np.float32)
                   for p in range(NoP)]
    pidx = np.arange(NoP, dtype=np.int32).reshape(C, nparts)
    didx = pidx.copy()
    deff = np.tile(np.array([0.01, 0.0, 0.01, 0.0], dtype=np.float32), (NoP, 1))
    anchor = np.zeros((NoP, 2), dtype=np.int32)
    for c, (h, w) in enumerate(roots):
        for k in range(nparts):
            r, col = divmod(k, 4)
            anchor[c * nparts + k, 0] = 1 + (col * (w - 2)) // 3
            anchor[c * nparts + k, 1] = 1 + r * (h - 2)
    # Least-squares box coefficients (x1, y1, x2, y2) over [1 + 2(1 + np)] terms, used by
    # the box regression that follows NMS (CT/tracking.cpp:90-166): identity on the root box
    # (width term 0, root x1/y1 picked by x1/y1 and offset by the root size for x2/y2).
    GLs = 1 + 2 * (1 + nparts)
    lsq = np.zeros((C, 4, GLs), dtype=np.float32)
    lsq[:, 0, 1] = 1.0
    lsq[:, 1, 2] = 1.0
    lsq[:, 2, 0] = 1.0
    lsq[:, 2, 1] = 1.0
    lsq[:, 3, 2] = 1.0
    for c, (h, w) in enumerate(roots):
        lsq[c, 3, 0] = float(h) / float(w)
    return DPMModel(
        numcomponent=C, sbin=8, interval=10,
        max_Y=int(rsize[:, 0].max()), max_X=int(rsize[:, 1].max()),
        ridx=np.arange(C, dtype=np.int32), oidx=np.arange(C, dtype=np.int32), offw=offw,
        rsize=rsize, numpart=np.full(C, nparts, dtype=np.int32), pidx=pidx, didx=didx,
        psize=np.full((C, nparts, 2), 6, dtype=np.int32), deff=deff, anchor=anchor,
        root_size=np.array([[h, w, 31] for h, w in roots], dtype=np.int32),
        rootfilters=rootfilters,
        part_size=np.tile(np.array([6, 6, 31], dtype=np.int32), (NoP, 1)),
        partfilters=partfilters, lsq=lsq, name=name)


CAR_ROOTS = [(6, 11), (6, 11), (7, 10), (7, 10), (8, 9), (8, 9)]
PERSON_ROOTS = [(12, 5), (12, 5), (11, 6), (11, 6), (14, 5), (14, 5)]


def synthetic_car(seed: int = 7) -> DPMModel:
    """Synthetic car-shaped model: 6 components, 8 parts of 6x6 each."""
    return _synthetic(CAR_ROOTS, 8, seed, "synthetic_car")


def synthetic_person(seed: int = 11) -> DPMModel:
    """Synthetic person-shaped model: 6 components, 8 parts of 6x6 each."""
    return _synthetic(PERSON_ROOTS, 8, seed, "synthetic_person")


def tiny_model(seed: int = 3, roots=((3, 4), (4, 3)), nparts: int = 2,
               psize: int = 3) -> DPMModel:
    """A small model for fast parity tests (same structure, smaller filters)."""
    C = len(roots)
    NoP = C * nparts
    rsize = np.array(roots, dtype=np.int32)
    rootfilters = [normal(seed, 31 * w * h, 0.05, stream=100 + c).astype(np.float32)
                   for c, (h, w) in enumerate(roots)]
    partfilters = [normal(seed, 31 * psize * psize, 0.05, stream=1000 + p).astype(np.float32)
                   for p in range(NoP)]
    pidx = np.arange(NoP, dtype=np.int32).reshape(C, nparts)
    anchor = np.zeros((NoP, 2), dtype=np.int32)
    for c, (h, w) in enumerate(roots):
        for k in range(nparts):
            anchor[c * nparts + k] = (1 + k % max(w - 1, 1), 1 + (k // max(w - 1, 1)) % max(h - 1, 1))
    deff = np.tile(np.array([0.05, 0.01, 0.04, -0.01], dtype=np.float32), (NoP, 1))
    return DPMModel(
        numcomponent=C, sbin=8, interval=10,
        max_Y=int(rsize[:, 0].max()), max_X=int(rsize[:, 1].max()),
        ridx=np.arange(C, dtype=np.int32), oidx=np.arange(C, dtype=np.int32),
        offw=normal(seed, C, 1.0, stream=10).astype(np.float32), rsize=rsize,
        numpart=np.full(C, nparts, dtype=np.int32), pidx=pidx, didx=pidx.copy(),
        psize=np.full((C, nparts, 2), psize, dtype=np.int32), deff=deff, anchor=anchor,
        root_size=np.array([[h, w, 31] for h, w in roots], dtype=np.int32),
        rootfilters=rootfilters,
        part_size=np.tile(np.array([psize, psize, 31], dtype=np.int32), (NoP, 1)),
        partfilters=partfilters, name="tiny")
