"""maple_oc_update_noalloc (the allocation-free OC update) against maple_oc_update.

Called through the C ABI.  Two calls in a row on the device's one scratch array,
the second with a smaller input (fewer, smaller cells), each compared bit for bit
with maple_oc_update of a context holding the same cells; sizes beyond the
scratch limit are refused before anything is launched.
"""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda", 0) if torch.cuda.is_available() else None


def _noalloc(lib, dims, x, dJ, vf, mv, xm, nbisect=60):
    from maple_amd import _lib as L
    n = len(dims)
    flat = (C.c_int * (3 * n))(*[v for d in dims for v in d])
    arr = lambda v: np.ascontiguousarray(np.broadcast_to(np.asarray(v, dtype=np.float64), (n,)))
    vf, mv, xm = arr(vf), arr(mv), arr(xm)
    out = torch.empty_like(x)
    rc = lib.maple_oc_update_noalloc(0, n, flat, C.c_void_p(x.data_ptr()), C.c_void_p(dJ.data_ptr()), L.dptr(vf),
                                     L.dptr(mv), L.dptr(xm), int(nbisect), C.c_void_p(out.data_ptr()),
                                     C.c_void_p(torch.cuda.current_stream(DEV).cuda_stream))
    return rc, out


def _inputs(dims, seed):
    rng = np.random.default_rng(seed)
    n = sum(int(np.prod(d)) for d in dims)
    x = np.clip(rng.uniform(0.0, 0.6, n), 1e-3, 1.0)
    dJ = rng.lognormal(0.0, 2.0, n) * 1e-3
    dJ[rng.uniform(size=n) < 0.05] = -1.0          # clamped to 1e-12 by the update
    return torch.from_numpy(x).to(DEV), torch.from_numpy(dJ).to(DEV)


@pytest.mark.parametrize("calls", [[[16, 24, 8], [8]], [[32, 32], [16]]])
def test_noalloc_oc_matches_context_oc(calls):
    import maple_amd
    from maple_amd import _lib as L
    lib = L.load()
    for k, sizes in enumerate(calls):
        print(f"[noalloc] call {k}: cells {sizes}", flush=True)
        dims = [(r, r, r) for r in sizes]
        x, dJ = _inputs(dims, k)
        vf = [0.1 + 0.05 * i for i in range(len(sizes))]
        mv = [0.2 - 0.02 * i for i in range(len(sizes))]
        xm = [1e-3] * len(sizes)
        cb = maple_amd.CellBatch(sizes, nlevels=2, device=DEV)
        want = cb.oc_update(x, dJ, vf, mv, xm)
        rc, got = _noalloc(lib, dims, x, dJ, vf, mv, xm)
        assert rc == 0, lib.maple_last_error().decode()
        torch.cuda.synchronize()
        assert torch.equal(got, want)
        cb.close()


def test_noalloc_refuses_sizes_beyond_its_scratch():
    from maple_amd import _lib as L
    lib = L.load()
    x = torch.zeros(8, dtype=torch.float64, device=DEV)
    rc, _ = _noalloc(lib, [(2, 2, 2)] * 65, x, x, 0.3, 0.2, 1e-3)
    assert rc == L.MAPLE_EINVAL and b"ncells" in lib.maple_last_error()
    rc, _ = _noalloc(lib, [(256, 256, 256)], x, x, 0.3, 0.2, 1e-3)
    assert rc == L.MAPLE_EINVAL and b"elements" in lib.maple_last_error()
