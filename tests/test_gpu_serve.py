"""The native serving loop (sardine_server_*): every request's result is
the same alone and in the mix — bit for bit equal to the per-call path on
the same request — with batches formed as requests arrive, several batches
in flight, and admission held back by a small HBM budget."""
import numpy as np
import pytest

from sardine_amd import api, synthetic
from sardine_amd.workload import DeviceScene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scenes(cuda):
    a = synthetic.scene_small_plain(7, lines=512, samples=1024, out_rows=300, out_cols=360, dem_side=512)
    b = synthetic.scene_iw_full(3, n_swaths=3, n_bursts=3, lines_per_burst=400, raw_cols=1600, overlap_lines=60,
                                swath_overlap=120, utm=True)
    out = []
    for sc in (a, b):
        ds = DeviceScene.build(sc, cuda)
        m = ds.prepare_merged()
        out.append((sc, ds, m, api.ResidentScene(m, ds.scene, ds.dem)))
    return out


def _requests(scenes, n, seed):
    rng = np.random.default_rng(seed)
    F = api.SpeckleFilter
    filters = [None, F.Lee(5, 4.0), F.Frost(5, 1.0), F.GammaMap(7, 4.0), F.RefinedLee(4.0)]
    out = []
    for _ in range(n):
        s = int(rng.integers(0, 2))
        ds = scenes[s][1]
        rows, cols, _ = api.output_grid(ds.scene, ds.tc_config())
        h = int(rng.integers(1, min(rows, 180) + 1))
        w = int(rng.integers(1, min(cols, 220) + 1))
        r0, c0 = int(rng.integers(0, rows - h + 1)), int(rng.integers(0, cols - w + 1))
        cfg = ds.tc_config(aoi=(r0, c0, h, w), resampling=int(rng.integers(0, 3)))
        out.append((s, cfg, filters[int(rng.integers(0, len(filters)))]))
    return out


def _alone(scenes, s, cfg, filt):
    _, ds, m, _ = scenes[s]
    img = api.terrain_correction(m, ds.scene, ds.dem, cfg)
    sp = filt.apply(img.data, img.cols, img.rows) if filt is not None else None
    return img, sp


def _equal(res, ref):
    img, sp = ref
    assert res.status == 0
    g = res.image
    assert (g.rows, g.cols, g.geotransform) == (img.rows, img.cols, img.geotransform)
    assert np.array_equal(g.data.cpu().numpy(), img.data.cpu().numpy(), equal_nan=True)
    assert g.valid_pixel_count == img.valid_pixel_count
    assert list(g.newton_iter_histogram) == list(img.newton_iter_histogram)
    if sp is not None:
        assert np.array_equal(res.speckled.cpu().numpy(), sp.reshape(g.rows, g.cols).cpu().numpy(), equal_nan=True)


@pytest.mark.parametrize("budget,slots", [(1 << 32, 2), (4 << 20, 1)])
def test_served_equals_alone(scenes, budget, slots):
    """budget 4 MiB holds about three requests (outputs until release plus
    work in flight), so admission waits for releases; the client releases
    each result once it has checked it."""
    res_scenes = [s[3] for s in scenes]
    reqs = _requests(scenes, 24, seed=slots)
    refs = [_alone(scenes, s, c, f) for s, c, f in reqs]
    srv = api.Server(res_scenes, hbm_budget_bytes=budget, max_batch=8, slots=slots)
    tickets = [srv.submit(s, c, f) for s, c, f in reqs]
    sizes = []
    for t, ref in zip(tickets, refs):
        r = srv.wait(t)
        _equal(r, ref)
        assert r.t_submit_us <= r.t_admit_us <= r.t_launch_us <= r.t_done_us
        sizes.append(r.batch_size)
        srv.release(t)
    if budget > (1 << 30):
        assert max(sizes) > 1  # requests shared batches
    # one request alone on the same server
    s, c, f = reqs[0]
    t = srv.submit(s, c, f)
    r = srv.wait(t)
    assert r.batch_size == 1
    _equal(r, refs[0])
    srv.release(t)
    srv.close()


def test_server_rejects_oversized_and_bad_scene(scenes):
    from sardine_amd import SardineError

    res_scenes = [s[3] for s in scenes]
    srv = api.Server(res_scenes, hbm_budget_bytes=1 << 20, max_batch=4, slots=1)
    ds = scenes[0][1]
    with pytest.raises(SardineError):
        srv.submit(0, ds.tc_config())  # the full 300x360 grid needs more than 1 MiB
    with pytest.raises(SardineError):
        srv.submit(5, ds.tc_config(aoi=(0, 0, 4, 4)))
    t = srv.submit(0, ds.tc_config(aoi=(0, 0, 4, 4)))
    assert srv.wait(t).status == 0
    srv.release(t)
    srv.close()


def test_release_waits_for_queued_reads(scenes):
    """A result released right after the client enqueued a read of it (behind
    long-running work on its stream) is read intact: the server returns the
    memory only after that stream's queued work, even though the next
    requests reuse it at once."""
    import torch

    res_scenes = [s[3] for s in scenes]
    reqs = _requests(scenes, 12, seed=11)
    srv = api.Server(res_scenes, hbm_budget_bytes=1 << 30, max_batch=4, slots=2)
    ref = _alone(scenes, *reqs[0])[0].data.cpu().numpy()
    t = srv.submit(*reqs[0])
    r = srv.wait(t)
    assert r.status == 0
    a = torch.randn(4096, 4096, device="cuda")
    for _ in range(40):  # tens of ms of queued work ahead of the read
        a = (a @ a) * (1.0 / 64.0)
    copy = r.image.data.clone()
    srv.release(t)
    others = [srv.submit(*q) for q in reqs[1:]]  # these outputs reuse the freed memory
    for o in others:
        assert srv.wait(o).status == 0
        srv.release(o)
    torch.cuda.synchronize()
    assert np.array_equal(copy.cpu().numpy(), ref, equal_nan=True)
    srv.close()
