"""Non-finite detection: a NaN or an infinity in an input ends as
HMG_ERR_NONFINITE (-4) from the call that returns the result, not as an
ordinary result; in the server only the job that holds it fails."""
import numpy as np
import pytest

from conftest import VP_WORKLOAD

pytestmark = pytest.mark.gpu

NONFINITE = -4


def _with_nan(orc, n, seed, cell=(17, 23)):
    z = orc.fbm_terrain(n, n, seed)
    z[cell] = np.nan
    return z


def test_vpipes_drop_in_reports_nan_input(hm, orc):
    z = _with_nan(orc, 96, 3)
    with pytest.raises(hm.HmgError) as e:
        hm.hydraulic_vpipes(z, iterations=12, **VP_WORKLOAD)
    assert e.value.code == NONFINITE


def test_clean_input_is_not_reported(hm, orc):
    z = orc.fbm_terrain(96, 96, 3)
    hm.hydraulic_vpipes(z, iterations=12, **VP_WORKLOAD)
    assert np.isfinite(z).all()
    z = orc.fbm_terrain(96, 96, 3)
    hm.thermal(z, 1.0 / 96, 5)
    assert np.isfinite(z).all()


def test_thermal_drop_in_reports_inf_input(hm, orc):
    z = orc.fbm_terrain(80, 64, 4)
    z[30, 40] = np.inf
    with pytest.raises(hm.HmgError) as e:
        hm.thermal(z, 1.0 / 80, 5)
    assert e.value.code == NONFINITE


@pytest.mark.parametrize("storage", [0, 1, 2])
def test_session_reports_infinite_water(hm, orc, storage):
    """An infinite rain cell makes the water plane (stored as FP32, BF16 or
    FP16) non-finite: export and sums report it."""
    n = 64
    z = orc.fbm_terrain(n, n, 5)
    rain = np.ones((n, n), np.float32)
    rain[10, 12] = np.inf
    s = hm.Session(z, rain_map=rain, storage=storage, **dict(VP_WORKLOAD, maintain_water_volume=0))
    s.vpipes(3)
    with pytest.raises(hm.HmgError) as e:
        s.export()
    assert e.value.code == NONFINITE
    with pytest.raises(hm.HmgError) as e:
        s.sums()
    assert e.value.code == NONFINITE
    s.close()


def test_server_fails_only_the_nan_job(hm, orc):
    from highmap_amd import serve

    p = hm.VpipesParams.make(**VP_WORKLOAD)
    jobs = [(serve.JOB_VPIPES_THERMAL, 96, 25, 1), (serve.JOB_THERMAL, 80, 9, 2),
            (serve.JOB_VPIPES_THERMAL, 128, 31, 3), (serve.JOB_THERMAL, 64, 12, 4)]
    inputs = [orc.fbm_terrain(n, n, seed) for _, n, _, seed in jobs]
    bad = [_with_nan(orc, 96, 9), _with_nan(orc, 72, 10, cell=(5, 60))]
    sv = serve.Server(storage=hm.Storage.BF16, params=p)
    ids = []
    for (kind, n, it, _), z in zip(jobs, inputs):
        ids.append(sv.submit(n, n, kind, it, serve.job_talus(kind, n), z=z))
    bad_ids = [sv.submit(96, 96, serve.JOB_VPIPES_THERMAL, 25, 1.0 / 96, z=bad[0]),
               sv.submit(72, 72, serve.JOB_THERMAL, 7, 2.0 / 72, z=bad[1])]
    sv.wait_all()
    for j in bad_ids:
        assert sv.poll(j) == NONFINITE
    for (kind, n, it, _), z, j in zip(jobs, inputs, ids):
        assert sv.poll(j) == serve.DONE
        got = sv.result(j)
        if kind == serve.JOB_THERMAL:
            want = z.copy()
            hm.thermal(want, 2.0 / n, it)
        else:
            s = hm.Session(z, params=p, storage=hm.Storage.BF16)
            left = it
            while left > 0:
                k = min(10, left)
                s.vpipes(k)
                left -= k
                if k == 10:
                    s.thermal(1.0 / n, 1)
            want = s.export()[0]
            s.close()
        assert np.array_equal(got, want), (kind, n, it)
    sv.close()
