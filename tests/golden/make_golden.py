"""Writes the data fixtures under tests/golden/ (run in a container holding the reference).

* car_model.npz -- the reference's real car model (CT/car_comp.csv, car_root.csv,
  car_part.csv) parsed by the oracle's CSV loader: model weights are input data.
* resize_ref.npz -- input/output vectors of the reference's own resize() (CT/resize.cc),
  compiled from the reference tree into oracle/_ref.
* dt_ref.npz -- input/output vectors of the reference's own dt() (CT/dt.cpp), same build,
  on random maps, plateaus, exact ties and large-coefficient maps.
* add_part_ref.npz -- input/output vectors of the reference's own add_part_calculation()
  (CT/dt.cpp:38-55), same build, at several anchors and map sizes.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from geye_amd import synth  # noqa: E402


def car_model():
    m = O.load_csv_model(os.path.join(O.REF_DIR, "car_comp.csv"),
                         os.path.join(O.REF_DIR, "car_root.csv"),
                         os.path.join(O.REF_DIR, "car_part.csv"))
    d = dict(numcomponent=m.numcomponent, sbin=m.sbin, interval=m.interval, max_Y=m.max_Y,
             max_X=m.max_X, ridx=m.ridx, oidx=m.oidx, offw=m.offw, rsize=m.rsize,
             numpart=m.numpart, pidx=m.pidx, didx=m.didx, psize=m.psize, deff=m.deff,
             anchor=m.anchor, root_size=m.root_size, part_size=m.part_size, NoR=m.NoR,
             NoP=m.NoP)
    for i, f in enumerate(m.rootfilters):
        d[f"root{i}"] = f
    for i, f in enumerate(m.partfilters):
        d[f"part{i}"] = f
    for i, a in enumerate(m.lsq):
        d[f"lsq{i}"] = a
    np.savez_compressed(os.path.join(HERE, "car_model.npz"), **d)


def resize_vectors():
    frame = synth.make_frame(67, 45, seed=5)
    img = O.ipl_to_float(frame)
    out = {"img": img}
    scales = [1.0, 0.933033, 0.870551, 0.5, 0.707107, 0.535887]
    for k, s in enumerate(scales):
        out[f"out{k}"] = O.resize(img, s, use_ref=True)
    out["scales"] = np.array(scales, np.float32)
    np.savez_compressed(os.path.join(HERE, "resize_ref.npz"), **out)


def dt_vectors():
    rng = np.random.default_rng(123)
    cases = []
    cases.append((rng.standard_normal((23, 17)).astype(np.float32), (0.01, 0.0, 0.01, 0.0)))
    cases.append((rng.standard_normal((41, 64)).astype(np.float32), (0.023062, 0.010890, 0.015328, -0.009441)))
    plateau = np.round(rng.standard_normal((19, 30)) * 2).astype(np.float32)
    cases.append((plateau, (0.05, 0.0, 0.05, 0.0)))
    cases.append((np.zeros((13, 11), np.float32), (0.0, 0.0, 0.0, 0.0)))
    cases.append((rng.standard_normal((9, 70)).astype(np.float32), (50.0, 1.0, 50.0, -1.0)))
    ties = np.tile(np.array([1.0, 0.0, 1.0, 0.0], np.float32), (12, 5))
    cases.append((ties, (0.25, 0.5, 0.25, -0.5)))
    out = {}
    for k, (v, d) in enumerate(cases):
        M, Ix, Iy = O.dt(v, *d, use_ref=True)
        out[f"in{k}"] = v
        out[f"def{k}"] = np.array(d, np.float32)
        out[f"M{k}"] = M
        out[f"Ix{k}"] = Ix
        out[f"Iy{k}"] = Iy
    out["n"] = np.array(len(cases))
    np.savez_compressed(os.path.join(HERE, "dt_ref.npz"), **out)


def add_part_vectors():
    rng = np.random.default_rng(77)
    # (R0, R1, P0, P1, ax, ay): anchors at the map's start, inside, and at the last cell the
    # root map can reach (ay - 1 + 2 (R0 - 1) = P0 - 1)
    cases = [(5, 4, 12, 10, 1, 1), (11, 17, 30, 40, 3, 2), (6, 9, 16, 23, 6, 5),
             (1, 1, 3, 3, 2, 3), (13, 7, 30, 16, 3, 5)]
    out = {}
    for k, (R0, R1, P0, P1, ax, ay) in enumerate(cases):
        S = rng.standard_normal((R1, R0)).astype(np.float32)
        M = rng.standard_normal((P1, P0)).astype(np.float32)
        out[f"score{k}"] = S
        out[f"M{k}"] = M
        out[f"anchor{k}"] = np.array([ax, ay], np.int32)
        out[f"out{k}"] = O.add_part(S, M, ax, ay, use_ref=True)
    out["n"] = np.array(len(cases))
    np.savez_compressed(os.path.join(HERE, "add_part_ref.npz"), **out)


if __name__ == "__main__":
    O.build()
    if O.ref_lib() is None:
        raise SystemExit("oracle/_ref is not built (the reference tree is needed)")
    car_model()
    resize_vectors()
    dt_vectors()
    add_part_vectors()
    print("fixtures written")
