"""GPU parity of the pre-TC multilook (apply_multilook, scene_prep.rs:
1060-1161) through the C ABI: σ⁰, NESZ and geometry bit-exact against the
oracle, for unrolled and generic range factors, NaN blocks and truncated
edges; the 1 x 1 pass-through keeps the input bits; a multilooked scene
geocodes to the oracle's result on the same multilooked image.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from sardine_amd import _native, abi, api, synthetic
from sardine_amd.workload import DeviceScene
from test_gpu_parity import _tc_compare

pytestmark = pytest.mark.gpu


def _info(lines, samples):
    i = abi.MergedInfo()
    i.lines, i.samples = lines, samples
    i.near_slant_range_time_s = 0.0053
    i.range_pixel_spacing_m = 2.329562
    i.range_pixel_spacing_s = 2.329562 / 299_792_458.0 * 2.0
    i.azimuth_start_time_us = 1_548_221_628_000_000
    return i


def _run(cuda, d, n, r, a, ati):
    import torch

    lines, samples = d.shape
    src = _info(lines, samples)
    shape = abi.MergedInfo()
    L = _native.lib()
    assert L.sardine_apply_multilook(None, None, C.byref(src), r, a, ati, None, None, C.byref(shape), None) == 0
    td, tn = torch.from_numpy(d).to(cuda), torch.from_numpy(n).to(cuda)
    od = torch.full((shape.lines, shape.samples), 7.0, dtype=torch.float32, device=cuda)
    on = torch.full_like(od, 7.0)
    out = abi.MergedInfo()
    st = L.sardine_apply_multilook(td.data_ptr(), tn.data_ptr(), C.byref(src), r, a, ati, od.data_ptr(),
                                   on.data_ptr(), C.byref(out), torch.cuda.current_stream().cuda_stream)
    assert st == 0, _native.last_error()
    torch.cuda.synchronize()
    return od.cpu().numpy(), on.cpu().numpy(), out


@pytest.mark.parametrize("r,a,lines,samples", [(2, 1, 33, 4097), (4, 1, 100, 2001), (3, 2, 51, 999), (5, 4, 64, 1000),
                                               (7, 7, 50, 701), (10, 2, 41, 3333), (6, 3, 17, 5000),
                                               (1, 3, 30, 515), (8, 1, 9, 8192)])
def test_multilook_bit_exact(cuda, r, a, lines, samples):
    print(f"multilook r={r} a={a} {lines}x{samples}", flush=True)
    rng = np.random.default_rng(r * 100 + a)
    d = rng.exponential(0.2, (lines, samples)).astype(np.float32)
    n = rng.exponential(0.01, (lines, samples)).astype(np.float32)
    d[rng.random(d.shape) < 0.15] = np.nan
    n[rng.random(n.shape) < 0.01] = np.nan  # NaN NESZ under a valid σ⁰ propagates, as in the reference
    d[:a, :r] = np.nan
    d[a:2 * a, :r] = np.float32(1e-41)
    od, on, out = _run(cuda, d, n, r, a, 2.055e-3)
    st, rd, rn, ref = O.multilook(d, n, _info(lines, samples), r, a, 2.055e-3)
    assert st == 0
    assert np.array_equal(od.view(np.uint32), rd.view(np.uint32))
    assert np.array_equal(on.view(np.uint32), rn.view(np.uint32))
    assert bytes(out) == bytes(ref)


def test_multilook_pass_through_keeps_bits(cuda):
    d = np.arange(64, dtype=np.float32).reshape(8, 8)
    d.view(np.uint32)[3] = 0x7FC00123  # a NaN payload
    n = np.full((8, 8), 0.5, np.float32)
    od, on, out = _run(cuda, d, n, 1, 1, 0.0)
    assert np.array_equal(od.view(np.uint32), d.view(np.uint32)) and np.array_equal(on, n)
    assert bytes(out) == bytes(_info(8, 8))
    m = api.MergedSigma0(None, None, 0.0, 0.0, 0)
    assert api.apply_multilook(m, 1, 1, 0.0) is m  # the reference returns its input


def test_multilooked_scene_geocodes_like_oracle(cuda):
    """prepare_merged → apply_multilook 4 x 1 → terrain_correction on the
    device, against the oracle's multilook and TC on the same scene."""
    sc = synthetic.scene_small_plain(5, lines=512, samples=1024, out_rows=220, out_cols=260, dem_side=512)
    ds = DeviceScene.build(sc, cuda, host_raw=True)
    info, s0, nz, _ = O.prepare_merged(sc)
    m = ds.prepare_merged()
    ml = api.apply_multilook(m, 4, 1, 0.0)
    st, s0m, nzm, info_m = O.multilook(s0, nz, info, 4, 1, 0.0)
    assert st == 0
    assert np.array_equal(ml.data.cpu().numpy(), s0m, equal_nan=True)
    assert np.array_equal(ml.nesz.cpu().numpy(), nzm, equal_nan=True)
    assert (ml.near_slant_range_time_s, ml.range_pixel_spacing_m) == (info_m.near_slant_range_time_s,
                                                                     info_m.range_pixel_spacing_m)
    _tc_compare(sc, ds, info_m, s0m, nzm, ml)
