"""GPU tests of the continuous-batching solve server (openlb_amd/serving.py).

Requests arrive while others run, are admitted as slots and arena memory free
up, run their own step counts, and finish with results that are bit-identical
to the same solve run alone."""
import numpy as np
import pytest

from openlb_amd import serving
from openlb_amd.ensemble import Ensemble
from openlb_amd.lattice import BlockLattice

pytestmark = pytest.mark.gpu


def alone(req):
    with BlockLattice.from_case(req.case, req.storage) as lat:
        lat.ini_equilibrium(serving.initial_moments(req.case))
        lat.initialize()
        lat.collide_and_stream(req.steps)
        return lat.download_populations("f64")


def small_stream(count=10, rate=400.0, seed=3):
    return serving.poisson_channel_stream(count=count, rate=rate, seed=seed, size=(6, 18), steps=(20, 150),
                                          u_mean=0.05, u_std=0.01)


def test_results_in_the_mix_equal_results_alone(gpu_available):
    reqs = small_stream()
    assert {r.storage for r in reqs} == {"f32", "f16"}
    with serving.SolveServer(budget_bytes=64 << 20, max_members=16, tick_steps=17,
                             keep_populations=True) as srv:
        results = srv.run(reqs, realtime=True)
        assert srv.max_active > 1          # solves did overlap on the device
    assert sorted(r.rid for r in results) == [r.rid for r in reqs]
    by_id = {r.rid: r for r in reqs}
    for res in results:
        req = by_id[res.rid]
        assert res.steps == req.steps and res.finished >= res.admitted >= res.arrival - 1e-9
        np.testing.assert_array_equal(res.populations, alone(req))


def test_admission_waits_for_memory(gpu_available):
    """An arena that holds two of these channels at a time: the server never runs more
    than fits, and still completes every request."""
    reqs = [serving.channel_request(k, 0.0, 16, 40 + 7 * k, 0.05, "f32") for k in range(6)]
    pop, _ = Ensemble.bytes_for(19, "f32", (16, 16, 16))
    with serving.SolveServer(budget_bytes=2 * pop + 256, storages=("f32",), max_members=8,
                             tick_steps=16, keep_moments=True) as srv:
        results = srv.run(reqs, realtime=False)
        assert srv.max_active == 2
    assert len(results) == 6
    for res in results:
        assert np.isfinite(res.moments).all()
        # driven along +x by the velocity inlet
        assert res.moments[1][res.moments[0] > 0].mean() > 0
    summary = serving.summarize(results)
    assert summary["requests"] == 6 and summary["p99_s"] >= summary["p50_s"] > 0


def test_oversize_request_is_rejected_not_queued(gpu_available):
    """A solve larger than its storage's whole arena comes back as an error result at
    once; the other requests still run."""
    pop, _ = Ensemble.bytes_for(19, "f32", (16, 16, 16))
    reqs = [serving.channel_request(0, 0.0, 40, 30, 0.05, "f32"),      # > the arena
            serving.channel_request(1, 0.0, 16, 30, 0.05, "f32")]
    with serving.SolveServer(budget_bytes=2 * pop + 256, storages=("f32",), max_members=4,
                             tick_steps=16) as srv:
        assert srv.never_fits(reqs[0]) is not None and srv.never_fits(reqs[1]) is None
        results = srv.run(reqs, realtime=False)
    by_id = {r.rid: r for r in results}
    assert set(by_id) == {0, 1}
    assert by_id[0].error and "population bytes" in by_id[0].error and by_id[0].steps == 0
    assert by_id[1].error is None and by_id[1].steps == 30
    summary = serving.summarize(results)
    assert summary["requests"] == 1 and summary["rejected"] == 1


def test_small_requests_pass_a_waiting_large_one(gpu_available):
    """Skip-ahead admission: while a large solve waits for memory, later small solves
    that fit are admitted; after max_bypass of them the large one is next."""
    big_pop, _ = Ensemble.bytes_for(19, "f32", (24, 24, 24))
    small_pop, _ = Ensemble.bytes_for(19, "f32", (8, 8, 8))
    first = serving.channel_request(0, 0.0, 8, 200, 0.05, "f32")      # holds memory for a while
    big = serving.channel_request(1, 0.0, 24, 20, 0.05, "f32")
    smalls = [serving.channel_request(2 + k, 0.0, 8, 20, 0.05, "f32") for k in range(5)]
    # the arena holds the large solve alone, or a few small ones, never both the large and the first
    budget = big_pop + small_pop // 2
    with serving.SolveServer(budget_bytes=budget, storages=("f32",), max_members=8, tick_steps=10,
                             max_bypass=2) as srv:
        results = srv.run([first, big] + smalls, realtime=False)
    by_id = {r.rid: r for r in results}
    assert all(r.error is None for r in results) and len(results) == 7
    # two small solves were admitted while the large one waited, the others only after it
    passed = [k for k in range(2, 7) if by_id[k].admitted < by_id[1].admitted]
    assert passed == [2, 3], (passed, {k: (r.admitted, r.finished) for k, r in by_id.items()})
