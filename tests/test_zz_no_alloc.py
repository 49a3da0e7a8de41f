"""geye_dt_noalloc (the DT on caller-owned device memory, working memory in a static device
array) against geye_dt, bit for bit, on consecutive calls that share the scratch array: a
larger map, then a smaller one with other coefficients, then the largest map the array
holds, then the first map again; and the size limit reported without device work."""
import ctypes as C

import numpy as np
import pytest
import torch

import geye_amd

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def dt_noalloc(ctx, v, defs):
    """v: [dims1][dims0] float32 (column-major map); returns M, Ix, Iy as numpy arrays."""
    d1, d0 = v.shape
    dv = torch.from_numpy(np.ascontiguousarray(v, dtype=np.float32)).to("cuda")
    M = torch.empty((d1, d0), dtype=torch.float32, device="cuda")
    Ix = torch.full((d1, d0), -7, dtype=torch.int32, device="cuda")
    Iy = torch.full((d1, d0), -7, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()  # inputs in place before the library's stream reads them
    dims = (C.c_int * 2)(d0, d1)
    rc = geye_amd.lib().geye_dt_noalloc(ctx.h, C.c_void_p(dv.data_ptr()),
                                        *[C.c_float(x) for x in defs], dims,
                                        C.c_void_p(M.data_ptr()), C.c_void_p(Ix.data_ptr()),
                                        C.c_void_p(Iy.data_ptr()))
    assert rc == 0, geye_amd.lib().geye_last_error()
    return M.cpu().numpy(), Ix.cpu().numpy(), Iy.cpu().numpy()


def bits(a):
    return np.ascontiguousarray(a).view(np.int32)


CASES = [
    (263, 190, (0.01, 0.0, 0.01, 0.0), 1.0, 1),
    (37, 29, (0.05, 0.02, 0.03, -0.01), 0.5, 2),        # smaller, other coefficients
    (1024, 1024, (0.023062, 0.010890, 0.015328, -0.009441), 1.0, 3),  # the array's limit
    (2048, 3, (0.1, 0.0, 0.2, 0.05), 2.0, 4),            # the longest line
    (263, 190, (0.01, 0.0, 0.01, 0.0), 1.0, 1),          # the first map again
]


def test_dt_noalloc_matches_dt_on_consecutive_calls(ctx):
    for d0, d1, defs, scale, seed in CASES:
        say("dt_noalloc", d0, d1, defs)
        rng = np.random.default_rng(seed)
        v = np.round(rng.standard_normal((d1, d0)) * 8 * scale) / 8  # exact ties included
        v = v.astype(np.float32)
        M, Ix, Iy = dt_noalloc(ctx, v, defs)
        rM, rIx, rIy = ctx.dt(v, *defs)
        assert np.array_equal(bits(M), bits(rM))
        assert np.array_equal(Ix, rIx) and np.array_equal(Iy, rIy)


def test_dt_noalloc_rejects_maps_beyond_its_array(ctx):
    L = geye_amd.lib()
    for d0, d1 in [(2049, 4), (4, 2049), (1025, 1024)]:
        dims = (C.c_int * 2)(d0, d1)
        # the size check precedes any device work, so no device memory is needed here
        rc = L.geye_dt_noalloc(ctx.h, C.c_void_p(16), C.c_float(0.01), C.c_float(0.0),
                               C.c_float(0.01), C.c_float(0.0), dims, C.c_void_p(16),
                               C.c_void_p(16), C.c_void_p(16))
        assert rc == -1
        assert b"scratch limits" in L.geye_last_error()
