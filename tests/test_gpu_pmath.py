"""The portable libm of the geodesy and the projection (pmath.hpp), and the
IEEE division and square root it is built from, return the same bits on the
device as on the host: the condition for the GPU's ECEF targets and UTM
anchors to equal the oracle's bit for bit."""
import numpy as np
import pytest

import oracle as O
from sardine_amd import _native

pytestmark = pytest.mark.gpu

RANGES = {"sin": (-40, 40), "cos": (-40, 40), "exp": (-30, 30), "log": (1e-6, 1e6), "sinh": (-6, 6),
          "cosh": (-6, 6), "atanh": (-0.9999, 0.9999), "atan": (-1e3, 1e3), "atan2": (-5, 5), "div": (-1e6, 1e6),
          "sqrt": (0, 1e14)}


@pytest.mark.parametrize("fn", O.PMATH_FUNCS)
def test_device_bits_equal_host(cuda, fn):
    import torch

    print(f"pmath {fn}", flush=True)
    rng = np.random.default_rng(O.PMATH_FUNCS.index(fn) + 101)
    lo, hi = RANGES[fn]
    n = 200_000
    if fn in ("log", "sqrt"):
        x = np.exp(rng.uniform(np.log(max(lo, 1e-12)), np.log(hi), n))
    else:
        x = rng.uniform(lo, hi, n)
    y = rng.uniform(lo, hi, n)
    if fn == "div":
        y[np.abs(y) < 1e-3] = 1.0
    xd, yd = torch.from_numpy(x).to(cuda), torch.from_numpy(y).to(cuda)
    out = torch.empty_like(xd)
    st = _native.lib().sardine_diag_math(O.PMATH_FUNCS.index(fn), xd.data_ptr(), yd.data_ptr(), out.data_ptr(), n,
                                         None)
    assert st == 0, _native.last_error()
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    want = O.pmath(fn, x, y)
    bad = got.view(np.int64) != want.view(np.int64)
    assert not bad.any(), (fn, int(bad.sum()), x[bad][:3], got[bad][:3], want[bad][:3])
