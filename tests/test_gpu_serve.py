"""Continuous-batching server: every job's result equals the same job run
alone, whatever ran beside it and however admission was throttled."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REQS = [dict(id=i, side=side, iterations=it, kind=kind, seed=500 + i, arrival_s=0.0)
        for i, (side, it, kind) in enumerate([
            (64, 23, "vpipes+thermal"), (128, 40, "thermal"), (192, 31, "vpipes+thermal"),
            (96, 10, "vpipes+thermal"), (256, 9, "thermal"), (64, 0, "thermal"),
            (160, 57, "vpipes+thermal"), (128, 3, "vpipes+thermal"), (320, 12, "thermal"),
            (64, 0, "vpipes+thermal")])]


def _params(hm):
    return hm.VpipesParams.make(water_height=1e-3, evap_rate=0.01, dt=0.5, k_capacity=0.1,
                                k_erode=0.05, k_depose=0.05)


def test_job_larger_than_budget_is_rejected(hm):
    from highmap_amd import serve

    sv = serve.Server(storage=hm.Storage.BF16, hbm_budget=1 << 20, params=_params(hm))
    small = sv.submit(64, 64, serve.JOB_THERMAL, 3, 0.01, seed=1)
    big = sv.submit(512, 512, serve.JOB_THERMAL, 3, 0.01, seed=2)
    sv.wait_all()
    assert sv.poll(small) == serve.DONE
    assert sv.poll(big) == -2  # HMG_ERR_OOM
    sv.close()


# budget 0: 80% of free HBM (everything admitted at once); 6 MiB: the
# largest job (320^2) fits alone, so admission throttles the mix; a cap of
# 200^2 active cells runs the larger jobs alone and the small ones in pairs
@pytest.mark.parametrize("budget,cap", [(0, 0), (6 << 20, 0), (0, 200 * 200)])
def test_served_jobs_equal_jobs_alone(hm, budget, cap):
    from highmap_amd import serve

    p = _params(hm)
    sv = serve.Server(storage=hm.Storage.BF16, hbm_budget=budget, params=p,
                      max_active_cells=cap)
    ids = []
    for r in REQS:
        kind = serve.JOB_VPIPES_THERMAL if r["kind"] == "vpipes+thermal" else serve.JOB_THERMAL
        ids.append(sv.submit(r["side"], r["side"], kind, r["iterations"],
                             serve.job_talus(kind, r["side"]), seed=r["seed"]))
    sv.wait_all()
    for r, j in zip(REQS, ids):
        got = sv.result(j)
        ref = serve.run_alone(hm, r, hm.Storage.BF16, p, mode="fused")
        assert np.array_equal(got, ref), r
        t_sub, t_adm, t_done = sv.times(j)
        assert t_sub <= t_adm <= t_done
    st = sv.stats()
    assert st["used"] == 0 and st["ticks"] >= 1
    if cap:
        # arrival order: no job is admitted before an earlier one
        adm = [sv.times(j)[1] for j in ids]
        assert adm == sorted(adm)
    sv.close()


def test_server_with_host_input_matches_session(hm, orc):
    from highmap_amd import serve

    p = _params(hm)
    z0 = orc.fbm_terrain(100, 100, 3)
    sv = serve.Server(storage=hm.Storage.F32, params=p)
    j = sv.submit(100, 100, serve.JOB_VPIPES_THERMAL, 25, 0.01, z=z0)
    sv.wait_all()
    got = sv.result(j)
    sv.close()
    s = hm.Session(z0, params=p)
    for k in (10, 10, 5):
        s.vpipes(k)
        if k == 10:
            s.thermal(0.01, 1)
    assert np.array_equal(got, s.export()[0])
    s.close()


def test_server_rejects_bad_jobs(hm):
    from highmap_amd import serve

    sv = serve.Server(storage=hm.Storage.BF16, params=_params(hm))
    with pytest.raises(hm.HmgError):
        sv.submit(2, 2, serve.JOB_THERMAL, 5, 0.1)
    with pytest.raises(hm.HmgError):
        sv.submit(64, 64, 7, 5, 0.1)
    sv.close()
