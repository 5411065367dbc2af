"""GPU parity of the SLC measurement TIFF reader (SlcReader::read_burst_raw,
slc_reader.rs:532-575) through the C ABI: rows read from the file into HBM
equal the file's CInt16 values (in host order) for little- and big-endian
files, SHORT and LONG strip tables, a single strip, padding before row 0,
destinations that are not 8-byte aligned, and files larger than one set of
pinned chunks; a synthetic scene's swath raster comes back bit for bit.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O
import slc_fixtures as F
from sardine_amd import _native, abi, api, synthetic

pytestmark = pytest.mark.gpu


def _data(h, w, seed):
    return np.random.default_rng(seed).integers(-32768, 32768, (h, w, 2)).astype(np.int16)


@pytest.mark.parametrize("kw", [{}, dict(le=False), dict(offsets_type=3), dict(data_gap=24),
                                dict(le=False, data_gap=8)])
def test_slc_read_matches_file(cuda, tmp_path, kw):
    print(f"slc read {kw}", flush=True)
    d = _data(37, 53, len(kw))
    F.write_slc_tiff(tmp_path / "s.tiff", d, **kw)
    r = api.SlcReader.open(tmp_path / "s.tiff")
    assert (r.width, r.height, r.little_endian) == (53, 37, kw.get("le", True))
    for first, count in ((0, 37), (3, 5), (36, 1), (10, 0)):
        got = r.read_burst_raw(first, count).cpu().numpy()
        st, ref = O.slc_read(tmp_path / "s.tiff", r.tiff, first, count)
        assert st == 0 and np.array_equal(got, d[first:first + count]) and np.array_equal(got, ref)


def test_slc_read_single_strip_and_unaligned_destination(cuda, tmp_path):
    import torch

    d = _data(1, 301, 7)
    F.write_slc_tiff(tmp_path / "one.tiff", d, le=False)
    r = api.SlcReader.open(tmp_path / "one.tiff")
    buf = torch.zeros(301 * 2 + 2, dtype=torch.int16, device=cuda)
    view = buf[2:].view(1, 301, 2)  # 4-byte aligned only: the per-value byte swap
    assert view.data_ptr() % 8 != 0
    r.read_burst_raw(0, 1, out=view)
    assert np.array_equal(view.cpu().numpy(), d) and int(buf[0]) == 0 and int(buf[1]) == 0


def test_slc_read_large_file_in_several_chunk_sets(cuda, tmp_path):
    """1100 rows x 17000 samples (75 MB): more than one set of pinned chunks,
    both byte orders."""
    d = _data(1100, 17000, 11)
    for le in (True, False):
        F.write_slc_tiff(tmp_path / "big.tiff", d, le=le)
        r = api.SlcReader.open(tmp_path / "big.tiff")
        got = r.read_burst_raw(0, 1100)
        assert np.array_equal(got.cpu().numpy(), d)
        got = r.read_burst_raw(517, 400)
        assert np.array_equal(got.cpu().numpy(), d[517:917])


def test_slc_read_feeds_the_scene(cuda, tmp_path):
    """A synthetic IW swath written as its measurement TIFF and read back per
    burst (as deburst reads it, deburst.rs:315-334) is the raster the scene's
    calibration reads."""
    sc = synthetic.scene_iw_full(4, n_swaths=2, n_bursts=3, lines_per_burst=300, raw_cols=900, overlap_lines=40,
                                 swath_overlap=80, utm=False, dem_max_side=256)
    for k, sw in enumerate(sc.swaths):
        raw = sw.raw_host()
        F.write_slc_tiff(tmp_path / f"iw{k}.tiff", raw)
        r = api.SlcReader.open(tmp_path / f"iw{k}.tiff")
        parts = [r.read_burst_raw(b * sw.lines_per_burst, sw.lines_per_burst) for b in range(sw.n_bursts)]
        import torch
        assert np.array_equal(torch.cat(parts).cpu().numpy(), raw)


def test_multi_slice_reads(cuda, tmp_path):
    """MultiSlcReader (multi_slc_reader.rs:151-193): a swath split in three
    slice files read per burst in its logical line space."""
    d = _data(90, 77, 13)
    for k, (a, b) in enumerate(((0, 30), (30, 60), (60, 90))):
        F.write_slc_tiff(tmp_path / f"s{k}.tiff", d[a:b], le=(k != 1))
    m = api.MultiSlcReader.from_readers([api.SlcReader.open(tmp_path / f"s{k}.tiff") for k in range(3)])
    assert (m.width, m.height) == (77, 90)
    for first, count in ((0, 30), (30, 30), (65, 25), (31, 7)):
        assert np.array_equal(m.read_burst_raw(first, count).cpu().numpy(), d[first:first + count])
