"""Admission logic of the solve server (openlb_amd/serving.py) on the CPU.

The ensembles are replaced by a host stand-in with the same arena accounting
(Ensemble.bytes_for) that "runs" a member by counting steps, so the tests check
only the scheduling: skip-ahead admission bounded by max_bypass, rejection of
requests that can never fit, and that run() always returns."""
import numpy as np
import pytest

from openlb_amd import abi, serving
from openlb_amd.ensemble import Ensemble


class HostEnsemble:
    def __init__(self, q, storage, kind, max_members, population_bytes, cell_capacity, device):
        self.storage, self.cap, self.max_members = storage, population_bytes, max_members
        self.used = 0
        self.members = {}                      # slot -> [bytes, steps done, steps wanted]

    bytes_for = staticmethod(Ensemble.bytes_for)

    def usage(self):
        return {"members": len(self.members), "bytes_used": self.used, "bytes_capacity": self.cap}

    def add(self, case, steps):
        pop, _ = Ensemble.bytes_for(19, self.storage, case.shape)
        free = [m for m in range(self.max_members) if m not in self.members]
        if not free or self.used + pop > self.cap:
            raise abi.LbmhError(abi.E_MEMORY, "arena full")
        self.members[free[0]] = [pop, 0, steps]
        self.used += pop
        return free[0]

    def ini_equilibrium(self, m, rho_u):
        pass

    def initialize(self, m):
        pass

    def remove(self, m):
        self.used -= self.members.pop(m)[0]

    def step(self, n, graph=False):
        for rec in self.members.values():
            rec[1] = min(rec[2], rec[1] + n)

    def status_all(self):
        done = np.zeros(self.max_members, dtype=np.int64)
        left = np.ones(self.max_members, dtype=np.int64)
        for m, (_, d, want) in self.members.items():
            done[m], left[m] = d, want - d
        return done, left

    def close(self):
        pass


@pytest.fixture
def host_server(monkeypatch):
    monkeypatch.setattr(serving, "Ensemble", HostEnsemble)
    return serving.SolveServer


def test_oversize_request_is_rejected(host_server):
    pop, _ = Ensemble.bytes_for(19, "f32", (16, 16, 16))
    reqs = [serving.channel_request(0, 0.0, 40, 30, 0.05, "f32"),
            serving.channel_request(1, 0.0, 16, 30, 0.05, "f32"),
            serving.channel_request(2, 0.0, 16, 30, 0.05, "f16")]     # storage not served
    srv = host_server(budget_bytes=2 * pop + 256, storages=("f32",), max_members=4, tick_steps=16)
    results = {r.rid: r for r in srv.run(reqs, realtime=False)}
    assert "population bytes" in results[0].error and results[0].steps == 0
    assert "not served" in results[2].error
    assert results[1].error is None and results[1].steps == 30
    s = serving.summarize(list(results.values()))
    assert s["requests"] == 1 and s["rejected"] == 2


def test_skip_ahead_is_bounded_by_max_bypass(host_server):
    big_pop, _ = Ensemble.bytes_for(19, "f32", (24, 24, 24))
    small_pop, _ = Ensemble.bytes_for(19, "f32", (8, 8, 8))
    first = serving.channel_request(0, 0.0, 8, 200, 0.05, "f32")
    big = serving.channel_request(1, 0.0, 24, 20, 0.05, "f32")
    smalls = [serving.channel_request(2 + k, 0.0, 8, 20, 0.05, "f32") for k in range(5)]
    for bypass, expect in ((2, [2, 3]), (0, []), (10, [2, 3, 4, 5, 6])):
        srv = host_server(budget_bytes=big_pop + small_pop // 2, storages=("f32",), max_members=8,
                          tick_steps=10, max_bypass=bypass)
        res = {r.rid: r for r in srv.run([first, big] + smalls, realtime=False)}
        assert len(res) == 7 and all(r.error is None for r in res.values())
        order = sorted(res, key=lambda k: (res[k].admitted, k))
        passed = [k for k in range(2, 7) if order.index(k) < order.index(1)]
        assert passed == expect, (bypass, order)


def test_fifo_when_everything_fits(host_server):
    reqs = [serving.channel_request(k, 0.0, 8, 10 + k, 0.05, "f32") for k in range(6)]
    srv = host_server(budget_bytes=1 << 30, storages=("f32",), max_members=16, tick_steps=4)
    res = srv.run(reqs, realtime=False)
    assert sorted(r.rid for r in res) == list(range(6)) and srv.max_active == 6


def test_precision_guard_runs_low_tau_fp16_requests_in_fp32(host_server):
    reqs = [serving.channel_request(0, 0.0, 8, 20, 0.05, "f16", tau=0.6),
            serving.channel_request(1, 0.0, 8, 20, 0.05, "f16", tau=0.51),
            serving.channel_request(2, 0.0, 8, 20, 0.05, "f32", tau=0.51)]
    srv = host_server(budget_bytes=1 << 26, storages=("f32", "f16"), max_members=8, tick_steps=8)
    res = {r.rid: r for r in srv.run(reqs, realtime=False)}
    assert [(res[k].storage, res[k].promoted) for k in range(3)] == [("f16", False), ("f32", True),
                                                                     ("f32", False)]
    assert serving.summarize(list(res.values()))["promoted_to_f32"] == 1
    off = host_server(budget_bytes=1 << 26, storages=("f32", "f16"), max_members=8, tick_steps=8,
                      precision_guard=None)
    assert {r.rid: r.storage for r in off.run(reqs, realtime=False)} == {0: "f16", 1: "f16", 2: "f32"}
    # a server without an FP32 ensemble keeps FP16 requests where they are
    only16 = host_server(budget_bytes=1 << 26, storages=("f16",), max_members=8, tick_steps=8)
    assert {r.rid: r.storage for r in only16.run(reqs[:2], realtime=False)} == {0: "f16", 1: "f16"}


def test_guarded_storage():
    from openlb_amd import workloads as W
    from openlb_amd.ensemble import FP16_TAU_GUARD, guarded_storage
    low = W.lid_driven_cavity((6, 6, 6), tau=0.52, lid_u=0.05)
    high = W.lid_driven_cavity((6, 6, 6), tau=FP16_TAU_GUARD, lid_u=0.05)
    assert guarded_storage(low, "f16") == "f32"
    assert guarded_storage(high, "f16") == "f16"
    assert guarded_storage(low, "f16", None) == "f16"
    assert guarded_storage(low, "f64") == "f64" and guarded_storage(low, "f32") == "f32"
