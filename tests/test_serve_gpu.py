"""The serving loop on the GPU: results are the same alone and in a mix of
concurrent requests of different sizes (bit for bit: every kernel is
deterministic and requests share no state), they equal the plan path and
the oracle, and admission refuses, queues and cancels as specified."""
import math

import numpy
import pytest

from guanaco_amd import _capi as C
from guanaco_amd.serve import Server
from oracle import oracle as O

pytestmark = pytest.mark.gpu

L300 = O.electron_wavelength(300000)


def case(X, T, ndf, Z, storage, seed, ps=1.5, df=25000.0):
    rng = numpy.random.default_rng(seed)
    Y = X // 2
    proj = (rng.standard_normal((T, Y, X)) + 4.0).astype(numpy.float32)
    angles = numpy.deg2rad(numpy.linspace(-60, 60, T)).astype(numpy.float32)
    centre = numpy.full(Y, X / 2.0, numpy.float32)
    mn, mx, st, n = O.get_defocus_steps(df, ndf, None, X, centre, ps)
    dfs = numpy.array([mn + d * st for d in range(n)], numpy.float64)
    kw = dict(num_defocus=ndf, pixel_size=ps, min_defocus=mn - df,
              max_defocus=mx - df, storage=storage, defocus=dfs,
              wavelength=L300, Cs=2.7e7, phi=math.asin(0.1))
    return proj, angles, Z, kw


MIX = [
    case(128, 21, 1, 32, "f16", 1),
    case(256, 41, 4, 64, "f32", 2),
    case(64, 9, 2, 16, "f16", 3),
    case(256, 31, 1, 64, "bf16", 4),
    case(128, 61, 8, 32, "f32", 5),
    case(64, 13, 1, 16, "f32", 6),
]


def plan_result(proj, angles, Z, kw):
    X, Y = proj.shape[2], proj.shape[1]
    plan = C.VolumePlan(X, Y, angles, Z, **kw)
    d_in = C.DeviceArray.from_host(proj)
    d_vol = C.DeviceArray((Y, Z, X))
    plan.run(d_in, d_vol)
    C.call("gx_device_synchronize")
    out = d_vol.to_host()
    plan.destroy()
    return out


def test_results_same_alone_and_in_the_mix(capi):
    alone = []
    with Server(budget_bytes=1 << 30, max_inflight=1) as srv:
        for proj, angles, Z, kw in MIX:
            r = srv.submit(proj, angles, Z, **kw)
            alone.append(r.result())
            r.release()
    with Server(budget_bytes=1 << 30, max_inflight=4, backfill=True) as srv:
        reqs = [srv.submit(p, a, Z, **kw) for p, a, Z, kw in MIX * 2]
        mixed = [r.result() for r in reqs]
        infos = [r.info() for r in reqs]
        for r in reqs:
            r.release()
        st = srv.stats()
    assert st["completed"] == len(MIX) * 2 and st["failed"] == 0
    assert st["used_bytes"] == 0 and st["p99_ms"] >= st["p50_ms"] > 0
    # the mix really overlapped: some request started before another ended
    spans = sorted((i["start_s"], i["end_s"]) for i in infos)
    assert any(spans[k + 1][0] < spans[k][1] for k in range(len(spans) - 1))
    for k, (p, a, Z, kw) in enumerate(MIX):
        numpy.testing.assert_array_equal(mixed[k], alone[k])
        numpy.testing.assert_array_equal(mixed[k + len(MIX)], alone[k])
        numpy.testing.assert_array_equal(alone[k], plan_result(p, a, Z, kw))


def test_served_f32_volume_matches_oracle(capi):
    proj, angles, Z, kw = case(128, 17, 3, 32, "f32", 11)
    with Server(budget_bytes=1 << 28) as srv:
        r = srv.submit(proj, angles, Z, **kw)
        got = r.result()
        r.release()
    centre = O.get_centre(proj.shape, None, False)
    c, mn, mx = O.correct_projections(proj, centre, kw["pixel_size"], 300,
                                      25000.0, 3, None, 2.7, 0, 0,
                                      math.degrees(kw["phi"]))
    expect = O.reconstruct(c, angles, None, kw["pixel_size"], mn, mx, depth=Z)
    err = float(numpy.linalg.norm(got - expect) / numpy.linalg.norm(expect))
    assert err < 1e-5, err


def test_device_input_equals_host_input(capi):
    proj, angles, Z, kw = MIX[1]
    with Server(budget_bytes=1 << 29) as srv:
        r1 = srv.submit(proj, angles, Z, **kw)
        d = C.DeviceArray.from_host(proj)
        r2 = srv.submit(d, angles, Z, **kw)
        a, b = r1.result(), r2.result()
        r1.release()
        r2.release()
    numpy.testing.assert_array_equal(a, b)


def test_admission_refusal_and_cancellation(capi):
    proj, angles, Z, kw = MIX[0]
    need = Server.request_bytes(proj.shape[2], proj.shape[1], angles, Z,
                                **kw)
    with Server(budget_bytes=need // 2) as srv:
        with pytest.raises(C.GxError, match="budget"):
            srv.submit(proj, angles, Z, **kw)
    # room for one running request beside one held volume: r1 and r2 run in
    # turn; with both volumes held, r3 cannot start until a release, and is
    # cancelled while queued
    keep = proj.shape[1] * Z * proj.shape[2] * 4
    with Server(budget_bytes=need + keep + 4096, max_inflight=4) as srv:
        r1 = srv.submit(proj, angles, Z, **kw)
        r2 = srv.submit(proj, angles, Z, **kw)
        r1.wait()
        r2.wait()
        r3 = srv.submit(proj, angles, Z, **kw)
        assert r3.info()["state"] == "queued"
        r3.cancel()
        assert r3.wait()["state"] == "cancelled"
        r1.release()
        r4 = srv.submit(proj, angles, Z, **kw)
        assert r4.wait()["state"] == "done"
        numpy.testing.assert_array_equal(r4.result(), r2.result())
        for r in (r2, r3, r4):
            r.release()
        assert srv.stats()["used_bytes"] == 0


def test_served_request_with_per_row_centres(capi):
    proj, angles, Z, kw = case(64, 15, 2, 16, "f32", 12)
    centres = (32.0 + numpy.linspace(-2.0, 2.0, proj.shape[1])).astype(
        numpy.float32)
    kw = dict(kw, centre=centres)
    with Server(budget_bytes=1 << 28) as srv:
        r = srv.submit(proj, angles, Z, **kw)
        got = r.result()
        r.release()
    numpy.testing.assert_array_equal(got, plan_result(proj, angles, Z, kw))
