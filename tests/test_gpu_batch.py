"""Batched geocoding (sardine_batch_*) against the per-call path.

Every job of a batch — AOI tiles of uneven size (down to one row and one
pixel) on two resident scenes (EPSG:4326 and UTM), all three resamplers,
all five speckle filters — must equal what sardine_terrain_correction and
sardine_speckle_apply give for it alone, bit for bit, eagerly, on HIP-graph
replays with other work in between, with the job order changed, and after
the autotuner has picked a launch shape.  The per-call path itself is held
to the oracle by test_gpu_parity.py.
"""
import numpy as np
import pytest

from sardine_amd import abi, api, synthetic
from sardine_amd.workload import DeviceScene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scenes(cuda):
    a = synthetic.scene_small_plain(5, lines=512, samples=1024, out_rows=400, out_cols=400, dem_side=512)
    b = synthetic.scene_iw_full(2, n_swaths=3, n_bursts=3, lines_per_burst=400, raw_cols=1600, overlap_lines=60,
                                swath_overlap=120, utm=True)
    out = []
    for sc in (a, b):
        ds = DeviceScene.build(sc, cuda)
        m = ds.prepare_merged()
        out.append((sc, ds, m, api.ResidentScene(m, ds.scene, ds.dem)))
    return out


def _jobs(scenes):
    (_, da, _, _), (_, db, _, _) = scenes
    F = api.SpeckleFilter
    spec = [
        (0, da.tc_config(compute_lia=True), F.RefinedLee(1.0)),
        (0, da.tc_config(aoi=(37, 11, 150, 211), resampling=api.ResamplingKernel.Bicubic), F.Lee(5, 4.0)),
        (0, da.tc_config(aoi=(0, 0, 1, 400), resampling=api.ResamplingKernel.Lanczos3), F.Frost(5, 1.0)),
        (0, da.tc_config(aoi=(399, 399, 1, 1)), None),
        (1, db.tc_config(compute_lia=True), F.GammaMap(7, 4.0)),
        (1, db.tc_config(aoi=(5, 17, 97, 301), resampling=api.ResamplingKernel.Lanczos3, flatten=True,
                         compute_lia=True), F.Boxcar(9)),
        (1, db.tc_config(aoi=(60, 3, 40, 40), resampling=api.ResamplingKernel.Bicubic), F.RefinedLee(2.0)),
    ]
    return [api.BatchJob(s, c, f) for s, c, f in spec]


def _alone(scenes, job):
    _, ds, m, _ = scenes[job.scene]
    img = api.terrain_correction(m, ds.scene, ds.dem, job.config)
    sp = None
    if job.speckle is not None:
        sp = job.speckle.apply(img.data, img.cols, img.rows).reshape(img.rows, img.cols)
    return img, sp


def _same(a, b):
    a, b = a.cpu().numpy(), b.cpu().numpy()
    return a.shape == b.shape and np.array_equal(a, b, equal_nan=True)


def _check(results, refs):
    for (res, (img, sp)) in zip(results, refs):
        assert res.status == 0
        g = res.image
        assert (g.rows, g.cols) == (img.rows, img.cols)
        assert g.geotransform == img.geotransform
        assert _same(g.data, img.data)
        if img.mask is not None:
            assert _same(g.mask, img.mask) and _same(g.cos_lia, img.cos_lia)
        for k in ("valid_pixel_count", "dem_missing_count", "outside_footprint_count", "non_converged_count",
                  "degenerate_geometry_count", "flat_masked_count", "noise_masked_count"):
            assert getattr(g, k) == getattr(img, k), k
        assert list(g.newton_iter_histogram) == list(img.newton_iter_histogram)
        if sp is not None:
            assert _same(res.speckled.reshape(g.rows, g.cols), sp)


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_equals_per_call(scenes, use_graph):
    import torch

    res_scenes = [s[3] for s in scenes]
    jobs = _jobs(scenes)
    refs = [_alone(scenes, j) for j in jobs]
    b = api.Batch(max_jobs=16, use_graph=use_graph)
    for rep in range(3):
        _check(b.run(res_scenes, jobs), refs)
        # unrelated eager work between replays
        x = torch.randn(1 << 20, device=scenes[0][2].data.device)
        torch.cuda.synchronize()
        assert torch.isfinite(x.sum())
    # the same executor with a different table: reversed order, a subset
    sub = list(reversed(jobs))[:5]
    for j in sub:
        j.out = None
        j.speckle_out = None
    _check(b.run(res_scenes, sub), [_alone(scenes, j) for j in sub])
    b.close()


def test_batch_scenes_with_longer_orbit(cuda, scenes):
    """A resident scene whose orbit has more state vectors (33 at 5 s instead
    of 17 at 10 s) joins a batch: the chain kernels' LDS holds the largest
    orbit of the batch, and the graph is captured again whenever that size
    changes (small, large, small).  Every job equals its per-call result."""
    import dataclasses

    sc = synthetic.scene_small_plain(6, lines=512, samples=1024, out_rows=300, out_cols=300, dem_side=512)
    t, pos, vel = synthetic.state_vectors(sc.orbit, n=33, spacing_s=5.0)
    sc = dataclasses.replace(sc, sv_t_us=t, sv_pos=pos, sv_vel=vel)
    ds = DeviceScene.build(sc, cuda)
    m = ds.prepare_merged()
    all3 = list(scenes) + [(sc, ds, m, api.ResidentScene(m, ds.scene, ds.dem))]
    res_scenes = [s[3] for s in all3]
    (_, da, _, _), (_, db, _, _) = scenes
    F = api.SpeckleFilter

    def small():
        return [api.BatchJob(0, da.tc_config(aoi=(10, 10, 120, 130)), None),
                api.BatchJob(1, db.tc_config(aoi=(5, 5, 60, 70), resampling=api.ResamplingKernel.Bicubic),
                             F.Lee(5, 4.0))]

    def large():
        return small() + [api.BatchJob(2, ds.tc_config(compute_lia=True), F.RefinedLee(1.0)),
                          api.BatchJob(2, ds.tc_config(aoi=(0, 0, 50, 300),
                                                       resampling=api.ResamplingKernel.Lanczos3), None)]

    b = api.Batch(max_jobs=8, use_graph=True)
    for make in (small, large, small):
        jobs = make()
        print(f"batch of {len(jobs)} jobs", flush=True)
        _check(b.run(res_scenes, jobs), [_alone(all3, j) for j in jobs])
    b.close()


def test_batch_invalid_job_is_skipped(scenes):
    res_scenes = [s[3] for s in scenes]
    jobs = _jobs(scenes)[:3]
    bad = api.BatchJob(0, scenes[0][1].tc_config(aoi=(390, 0, 20, 10)), None)  # outside the 400-row grid
    b = api.Batch(max_jobs=8, use_graph=True)
    res = b.run(res_scenes, jobs[:1] + [bad] + jobs[1:], check=False)
    assert res[1].status == abi.ERR_TC_CONFIG
    _check([res[0]] + res[2:], [_alone(scenes, j) for j in jobs])
    b.close()


def test_batch_autotune_then_equal(scenes):
    res_scenes = [s[3] for s in scenes]
    jobs = _jobs(scenes)
    b = api.Batch(max_jobs=16, use_graph=True)
    best, ms = b.autotune(res_scenes, jobs)
    assert ms > 0.0
    assert best.segment_columns in (16, 24, 32, 48) and best.tc_groups_per_cu in (1, 2)
    t = b.tuning
    assert (t.segment_columns, t.tc_groups_per_cu, t.speckle_groups_per_cu) == \
        (best.segment_columns, best.tc_groups_per_cu, best.speckle_groups_per_cu)
    _check(b.run(res_scenes, jobs), [_alone(scenes, j) for j in jobs])
    b.close()
