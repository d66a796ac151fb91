"""Batched and served solves match the same problem solved alone, bit for bit.

Every reduction, factorization and LM decision of the library is per window,
so a window's arithmetic must not depend on the other windows of its batch or
on when the serving loop admitted it.  Every served job's final RMSE is also
checked against the FP64 oracle's (1e-3 px).  Also checks the serving loop's
admission control and bookkeeping, that one failing job leaves the jobs
batched with it as if alone, and that a malformed job is refused at submit.
"""
import dataclasses

import numpy as np
import pytest

from hs_sfm_amd import abi, synthetic as syn
from hs_sfm_amd.hsba import Batch, Problem, Server
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def alone(scene, iters, precision=abi.PRECISION_MIXED):
    p = Problem(scene, precision=precision)
    rc, x, st = p.lm(max_iterations=iters)
    p.close()
    assert rc == 0
    return x, st


def test_batch_bitwise_matches_single(gpu):
    scenes = syn.windows_batch(count=6, scale_points=0.1)
    rc, xs, sts = Batch(scenes).lm(max_iterations=6)
    assert rc == 0
    for s, x, st in zip(scenes, xs, sts):
        xa, sta = alone(s, 6)
        assert np.array_equal(x, xa), s.name
        assert st.final_cost == sta.final_cost and st.iterations == sta.iterations


def test_large_windows_batch_bitwise_matches_single(gpu):
    """Windows whose reduced systems take the tiled factorization (more than
    four 64-row tiles) share each launch of the large path; tile counts
    differ (6, 8 and 5 tiles), so windows finish their steps at different
    launches.  Each result equals the window solved alone, bit for bit, and
    the FP64 oracle's RMSE within 1e-3 px."""
    scenes = [syn.mid_scene(seed=21, num_cameras=40, num_points=1500, kmax=20, name="L40"),
              syn.mid_scene(seed=22, num_cameras=53, num_points=1800, kmax=20, name="L53"),
              syn.mid_scene(seed=23, num_cameras=33, num_points=1200, kmax=20, name="L33")]
    assert [(9 * s.num_images + 63) // 64 for s in scenes] == [6, 8, 5]
    rc, xs, sts = Batch(scenes).lm(max_iterations=5)
    assert rc == 0
    for s, x, st in zip(scenes, xs, sts):
        xa, sta = alone(s, 5)
        assert np.array_equal(x, xa), s.name
        assert st.final_cost == sta.final_cost and st.iterations == sta.iterations
        _, xo, _, _ = Oracle(s).lm(max_iterations=5)
        assert abs(syn.reprojection_rmse(s, x) - syn.reprojection_rmse(s, xo)) < 1e-3, s.name


def mixed_jobs():
    rng = np.random.default_rng(7)
    jobs = []
    for j in range(10):
        c, p = syn.window_sizes(rng, 1)
        jobs.append(syn.window_scene(rng, int(c[0]), max(8, int(p[0]) // 10), name=f"w{j}"))
    # a second shape class: 9-parameter cameras
    jobs.insert(3, syn.mid_scene(seed=11, num_cameras=12, num_points=300, kmax=12, name="m0"))
    jobs.insert(8, syn.mid_scene(seed=12, num_cameras=14, num_points=400, kmax=12, name="m1"))
    return jobs


@pytest.mark.parametrize("streams,batch", [(1, 4), (3, 2)])
def test_server_results_match_alone(gpu, streams, batch):
    jobs = mixed_jobs()
    iters = 5
    srv = Server(streams=streams, max_batch_jobs=batch, max_iterations=iters)
    prepared = [srv.prepare(s) for s in jobs]
    ids = [srv.submit(j) for j in prepared]
    srv.drain()
    res = {r.job_id: r for r in srv.poll(64)}
    assert sorted(res) == sorted(ids)
    for jid, s in zip(ids, jobs):
        r = res[jid]
        assert r.status == 0, (s.name, r.status)
        assert r.submit_time_s <= r.admit_time_s <= r.finish_time_s
        xa, sta = alone(s, iters)
        assert np.array_equal(srv.x_out(jid), xa), s.name
        assert r.stats.final_cost == sta.final_cost
        _, xo, _, _ = Oracle(s).lm(max_iterations=iters)
        assert abs(syn.reprojection_rmse(s, srv.x_out(jid)) - syn.reprojection_rmse(s, xo)) < 1e-3, s.name
    assert srv.poll(8) == []
    srv.close()


def with_unobserved_image(s):
    """The scene plus one image that no key observes (its reduced-system
    block is missing, so the reference's SimplicialLDLT fails: -1)."""
    P, I = s.num_points, s.num_images
    x0 = np.concatenate([s.x0[:3 * P + 6 * I], s.x0[3 * P:3 * P + 6], s.x0[3 * P + 6 * I:]])
    return dataclasses.replace(s, num_images=I + 1, image_camera_map=np.append(s.image_camera_map, 0),
                               x0=x0, x_true=None, name=s.name + "+unobserved")


def test_failing_job_leaves_its_batch_alone(gpu):
    jobs = mixed_jobs()[:6]
    bad = with_unobserved_image(jobs[2])
    iters = 4
    srv = Server(streams=1, max_batch_jobs=8, max_iterations=iters)
    ids = [srv.submit(srv.prepare(s)) for s in jobs[:2] + [bad] + jobs[2:]]
    srv.drain()
    res = {r.job_id: r for r in srv.poll(64)}
    for jid, s in zip(ids, jobs[:2] + [bad] + jobs[2:]):
        r = res[jid]
        if s is bad:
            assert r.status == -1  # HSBA_ERR_GENERIC, the reference's -1
            continue
        assert r.status == 0, (s.name, r.status)
        xa, _ = alone(s, iters)
        assert np.array_equal(srv.x_out(jid), xa), s.name
    srv.close()


def test_malformed_job_refused_at_submit(gpu):
    s = mixed_jobs()[0]
    fm = s.feature_maps.copy()
    fm[1] = fm[0]  # a duplicate (image, point) key
    dup = dataclasses.replace(s, feature_maps=fm)
    srv = Server(streams=1, max_iterations=2)
    with pytest.raises(RuntimeError, match="HSBA_ERR_INVALID"):
        srv.submit(srv.prepare(dup))
    out = fm.copy()
    out[0, 1] = s.num_points  # point id out of range
    with pytest.raises(RuntimeError, match="HSBA_ERR_INVALID"):
        srv.submit(srv.prepare(dataclasses.replace(s, feature_maps=out)))
    jid = srv.submit(srv.prepare(s))  # the server still serves good jobs
    srv.drain()
    assert [r.status for r in srv.poll(4) if r.job_id == jid] == [0]
    srv.close()


def test_server_admission(gpu):
    jobs = mixed_jobs()[:4]
    # a budget below one mid-class job's footprint: that job is refused at submit
    srv = Server(hbm_budget_bytes=1 << 20, max_iterations=3)
    with pytest.raises(RuntimeError):
        srv.submit(srv.prepare(jobs[3]))
    srv.close()
    # room for about one job at a time: every job still completes, in order
    srv = Server(hbm_budget_bytes=96 << 20, max_active_jobs=1, streams=2, max_iterations=3)
    ids = [srv.submit(srv.prepare(s)) for s in jobs]
    srv.drain()
    res = sorted(srv.poll(16), key=lambda r: r.admit_time_s)
    assert [r.job_id for r in res] == ids
    for a, b in zip(res, res[1:]):
        assert b.admit_time_s >= a.finish_time_s - 1e-9  # one active job at a time
    assert all(r.status == 0 for r in res)
    srv.close()
