"""Shared fixtures.  GPU tests are marked `gpu`; everything else runs on the CPU."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built library")
    config.addinivalue_line("markers", "nms_ties: the test's NMS call must reorder tied runs on "
                                       "the device order (no host replay)")
    config.addinivalue_line("markers", "nms_replay: the test's NMS call must take the host replay")


@pytest.fixture(scope="session")
def ctx():
    import geye_amd

    c = geye_amd.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def car_model():
    """The reference's real car model (CT/car_*.csv), stored as a data fixture."""
    from geye_amd.synth import DPMModel

    z = np.load(os.path.join(GOLDEN, "car_model.npz"), allow_pickle=False)
    nroot = int(z["NoR"])
    npart = int(z["NoP"])
    return DPMModel(
        numcomponent=int(z["numcomponent"]), sbin=int(z["sbin"]), interval=int(z["interval"]),
        max_Y=int(z["max_Y"]), max_X=int(z["max_X"]), ridx=z["ridx"], oidx=z["oidx"],
        offw=z["offw"], rsize=z["rsize"], numpart=z["numpart"], pidx=z["pidx"], didx=z["didx"],
        psize=z["psize"], deff=z["deff"], anchor=z["anchor"], root_size=z["root_size"],
        rootfilters=[z[f"root{i}"] for i in range(nroot)], part_size=z["part_size"],
        partfilters=[z[f"part{i}"] for i in range(npart)],
        lsq=[z[f"lsq{i}"] for i in range(int(z["numcomponent"]))], name="car")
