"""GPU tests of lattice ensembles (C ABI lbmh_ensemble_*).

Each member runs exactly the single-block cell body with the same derived
parameters, so a member must be bit-identical to the same lattice advanced
on its own by lbmh_step. The single-lattice path is pinned to the oracle and
to OpenLB by test_gpu_parity.py, so bit-identity extends that pin to every
member."""
import dataclasses

import numpy as np
import pytest

import golden_cases as G
from oracle import oracle as O
from openlb_amd import workloads as W
from openlb_amd.ensemble import Ensemble
from openlb_amd.lattice import BlockLattice

pytestmark = pytest.mark.gpu


def small_members(seed=3, count=7):
    rng = np.random.Generator(np.random.PCG64(seed))
    cases = []
    for k in range(count):
        shape = tuple(int(v) for v in rng.integers(4, 23, size=3))
        c = W.lid_driven_cavity(shape, tau=float(rng.uniform(0.52, 1.2)), lid_u=float(rng.uniform(0.02, 0.1)))
        cases.append(c)
    # a periodic member with NoDynamics holes and a second wall id
    p = W.random_periodic((9, 5, 13), seed=seed, tau=0.8)
    p.ids = p.ids.copy()
    p.ids[2, :, 3] = 0
    p.ids[4, 1, :] = 5
    p.walls = {5: (1.02, (0.01, -0.02, 0.005))}
    cases.append(p)
    return cases


def standalone(case, storage, f0, steps):
    with BlockLattice.from_case(case, storage) as lat:
        lat.upload_populations(f0, "f64")
        lat.collide_and_stream(steps)
        return lat.download_populations("f64")


@pytest.mark.parametrize("storage", ["f64", "f32", "f16"])
def test_members_match_standalone_lattices(gpu_available, storage):
    cases = small_members()
    with Ensemble(19, storage, "bgk", max_members=16, population_bytes=64 << 20, cell_capacity=1 << 20) as ens:
        members, f0s = [], []
        for c in cases:
            m = ens.add(c)
            f0 = O.initial_populations(c, storage)
            f0s.append(f0)
            ens.upload(m, f0)
            members.append(m)
        ens.step(17)
        for m, c, f0 in zip(members, cases, f0s):
            assert ens.status(m) == (17, -1)
            np.testing.assert_array_equal(ens.download(m), standalone(c, storage, f0, 17))


def test_step_budgets_stop_members_individually(gpu_available):
    cases = small_members(seed=5, count=3)[:3]
    budgets = [4, 11, -1]
    with Ensemble(19, "f32", "bgk", max_members=4, population_bytes=16 << 20, cell_capacity=1 << 18) as ens:
        ms = [ens.add(c, steps=b) for c, b in zip(cases, budgets)]
        f0s = []
        for m, c in zip(ms, cases):
            f0 = O.initial_populations(c, "f32")
            ens.upload(m, f0)
            f0s.append(f0)
        ens.step(20)
        assert [ens.status(m) for m in ms] == [(4, 0), (11, 0), (20, -1)]
        for m, c, f0, b in zip(ms, cases, f0s, [4, 11, 20]):
            np.testing.assert_array_equal(ens.download(m), standalone(c, "f32", f0, b))
        # a finished member can be given more steps
        ens.set_steps(ms[0], 3)
        ens.step(5)
        assert ens.status(ms[0]) == (7, 0)


def test_graph_replay_matches_eager_steps(gpu_available):
    """The captured step loop is replayed several times, with eager work and
    synchronisations between replays. Every replay must match eager stepping bit for bit."""
    cases = small_members(seed=7, count=5)
    with Ensemble(19, "f32", "bgk", max_members=8, population_bytes=32 << 20, cell_capacity=1 << 20) as a, \
            Ensemble(19, "f32", "bgk", max_members=8, population_bytes=32 << 20, cell_capacity=1 << 20) as b:
        for c in cases:
            f0 = O.initial_populations(c, "f32")
            a.upload(a.add(c), f0)
            b.upload(b.add(c), f0)
        for replay in range(3):
            a.step(70, graph=True)     # two graph replays of 32 steps plus 6 plain steps
            b.step(70, graph=False)
            for m in range(len(cases)):
                np.testing.assert_array_equal(a.download(m), b.download(m), err_msg=f"replay {replay}")
                assert a.status(m) == b.status(m) == (70 * (replay + 1), -1)
        # membership change invalidates the captured graph
        extra = W.lid_driven_cavity((6, 7, 8), tau=0.7, lid_u=0.05)
        f0 = O.initial_populations(extra, "f32")
        ma, mb = a.add(extra), b.add(extra)
        a.upload(ma, f0)
        b.upload(mb, f0)
        a.step(40, graph=True)
        b.step(40)
        for m in range(len(cases) + 1):
            np.testing.assert_array_equal(a.download(m), b.download(m))


def test_add_remove_reuses_arena(gpu_available):
    cases = small_members(seed=9, count=6)
    with Ensemble(19, "f64", "bgk", max_members=6, population_bytes=48 << 20, cell_capacity=1 << 18) as ens:
        ms = [ens.add(c) for c in cases[:5]]
        used = ens.usage()["bytes_used"]
        ens.remove(ms[1])
        ens.remove(ms[3])
        assert ens.usage()["members"] == 3
        assert ens.usage()["bytes_used"] < used
        # refill the freed slots with other extents
        new = [W.lid_driven_cavity((7, 9, 5), tau=0.9, lid_u=0.03), cases[5]]
        for c in new:
            m = ens.add(c)
            f0 = O.initial_populations(c, "f64")
            ens.upload(m, f0)
            ens.step(12)
            np.testing.assert_array_equal(ens.download(m), standalone(c, "f64", f0, 12))
        with pytest.raises(Exception):
            for _ in range(10):
                ens.add(W.lid_driven_cavity((60, 60, 60)))


def test_member_matches_oracle(gpu_available):
    c = W.lid_driven_cavity((14, 11, 17), tau=0.58, lid_u=0.09)
    f0 = O.initial_populations(c, "f64")
    with Ensemble(19, "f64", "bgk", max_members=2, population_bytes=8 << 20, cell_capacity=1 << 16) as ens:
        m = ens.add(c)
        ens.upload(m, f0)
        ens.step(60, graph=True)
        got = ens.download(m)
        mom = ens.moments(m)
    ora = O.OracleLattice(c, "f64")
    ora.upload(f0)
    ora.step(60, threads=4)
    assert G.rel_error(got, ora.download(), 19, c.ids == 1) <= 1e-13
    np.testing.assert_allclose(mom[:, c.ids == 1], ora.moments()[:, c.ids == 1], atol=1e-13)


def test_nan_in_one_member_leaves_the_others_untouched(gpu_available):
    """Members share one launch but no memory: a diverged member (NaN seeded in
    one cell) must not leak into its neighbours in the arena, which stay
    bit-identical to their standalone runs; the diverged member shows the NaN
    in its moments."""
    cases = small_members(seed=11, count=5)
    bad = 2
    with Ensemble(19, "f32", "bgk", max_members=16, population_bytes=64 << 20, cell_capacity=1 << 20) as ens:
        members, f0s = [], []
        for k, c in enumerate(cases):
            m = ens.add(c)
            f0 = O.initial_populations(c, "f32")
            f0s.append(f0)
            seeded = f0.copy()
            if k == bad:
                centre = tuple(n // 2 for n in c.shape)
                seeded[(slice(None),) + centre] = np.nan
            ens.upload(m, seeded)
            members.append(m)
        ens.step(32, graph=True)
        for k, (m, c, f0) in enumerate(zip(members, cases, f0s)):
            got = ens.download(m)
            if k == bad:
                assert np.isnan(got).any()
                assert np.isnan(ens.moments(m)).any()
            else:
                assert np.isfinite(got).all()
                np.testing.assert_array_equal(got, standalone(c, "f32", f0, 32))


def test_statistics_flag_a_diverged_lattice(gpu_available):
    """A NaN population shows up in the lattice statistics (average rho, max |u|)."""
    case = W.lid_driven_cavity((10, 9, 12), tau=0.6, lid_u=0.05)
    f0 = O.initial_populations(case, "f32")
    f0[:, 4, 4, 5] = np.nan
    with BlockLattice.from_case(case, "f32") as lat:
        lat.upload_populations(f0, "f64")
        lat.collide_and_stream(3)
        st = lat.statistics()
    assert np.isnan(st["average_rho"]) and np.isnan(st["average_energy"])


@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_local_boundary_members_match_standalone(gpu_available, storage):
    """Members with LocalVelocity / LocalPressure / InterpolatedVelocity /
    InterpolatedPressure boundaries (edge, corner and plane post-processors
    included) beside bounce-back members: each is bit-identical
    to the same lattice stepped alone, in eager and graph steps, and member 0's
    initialize() equals the lattice's."""
    cases = [W.local_velocity_cavity((9, 8, 11), tau=0.6, lid_u=0.07),
             W.lid_driven_cavity((7, 9, 6), tau=0.7, lid_u=0.05),
             W.velocity_channel((16, 8, 7), u_in=0.03, tau=0.65),
             W.local_velocity_cavity((6, 7, 5), tau=0.8, lid_u=0.1),
             W.local_velocity_cavity((8, 6, 7), tau=0.7, lid_u=0.06, boundary="interp_velocity"),
             W.velocity_channel((14, 7, 6), u_in=0.03, rho_out=1.0, tau=0.6, inlet="interp_velocity",
                                outlet="interp_pressure")]
    with Ensemble(19, storage, "bgk", max_members=8, population_bytes=64 << 20, cell_capacity=1 << 20) as ens:
        members, f0s = [], []
        for c in cases:
            m = ens.add(c)
            ens.ini_equilibrium(m, c.rho_u)
            ens.initialize(m)
            members.append(m)
            f0s.append(ens.download(m))
        with BlockLattice.from_case(cases[0], storage) as lat:
            lat.ini_equilibrium(cases[0].rho_u)
            lat.initialize()
            np.testing.assert_array_equal(lat.download_populations("f64"), f0s[0])
        ens.step(9)
        ens.step(32, graph=True)
        for m, c, f0 in zip(members, cases, f0s):
            assert ens.status(m) == (41, -1)
            np.testing.assert_array_equal(ens.download(m), standalone(c, storage, f0, 41), err_msg=f"{m} {c.name}")


@pytest.mark.parametrize("kind", ["smagorinsky", "forced_trt"])
def test_local_boundary_members_with_other_bulk_kinds_match_standalone(gpu_available, kind):
    """Local-boundary members of a Smagorinsky or forced-TRT ensemble (the
    boundary cells keep their BGK mixin, the post-processors read a forced
    neighbour's velocity with the member's F/2): each member equals the same
    lattice stepped alone.  Smagorinsky members are bit-identical; forced-TRT
    members agree to 1e-14 (about one ulp of the populations after 25 steps):
    the lattice and the ensemble run the forced collision in different
    kernels, which may contract the same arithmetic into different FMAs."""
    def make(c):
        extra = {"smago": 0.14} if kind == "smagorinsky" else {"magic": 0.25, "force": (1e-5, 0.0, -4e-6)}
        return dataclasses.replace(c, kind=kind, **extra)
    cases = [make(W.velocity_channel((16, 8, 7), u_in=0.03, tau=0.56, inlet="interp_velocity",
                                     outlet="interp_pressure")),
             make(W.local_velocity_cavity((9, 8, 11), tau=0.6, lid_u=0.07)),
             make(W.velocity_channel((14, 7, 6), u_in=0.03, rho_out=1.0, tau=0.6))]
    with Ensemble(19, "f64", kind, max_members=4, population_bytes=64 << 20, cell_capacity=1 << 20) as ens:
        members, f0s = [], []
        for c in cases:
            m = ens.add(c)
            ens.ini_equilibrium(m, c.rho_u)
            ens.initialize(m)
            members.append(m)
            f0s.append(ens.download(m))
        ens.step(5)
        ens.step(20, graph=True)
        for m, c, f0 in zip(members, cases, f0s):
            got, ref = ens.download(m), standalone(c, "f64", f0, 25)
            if kind == "smagorinsky":
                np.testing.assert_array_equal(got, ref, err_msg=f"{m} {c.name}")
            else:
                np.testing.assert_allclose(got, ref, rtol=0, atol=1e-14, err_msg=f"{m} {c.name}")


def test_autotune_leaves_the_members_as_they_were(gpu_available):
    """lbmh_ensemble_autotune runs real steps on the members: their populations
    and step state must come back exactly, bounded budgets included, and the
    ensemble then steps bit-identically to one that was never tuned."""
    cases = small_members(seed=11, count=4) + [W.local_velocity_cavity((9, 8, 7), tau=0.7, lid_u=0.06)]
    finals = []
    for tune in (False, True):
        with Ensemble(19, "f32", "bgk", max_members=8, population_bytes=64 << 20, cell_capacity=1 << 20) as ens:
            members = []
            for c in cases:
                m = ens.add(c, steps=25)
                ens.upload(m, O.initial_populations(c, "f32"))
                if c.local:
                    ens.initialize(m)
                members.append(m)
            ens.step(7)
            if tune:
                before = [ens.download(m) for m in members]
                chosen = ens.autotune(trial_steps=2)
                assert chosen["lds_bytes"] in (0, 16384, 24576, 32768, 40960, 49152)
                assert chosen["xcd_group"] in (4, 8)
                for m, f in zip(members, before):
                    assert ens.status(m) == (7, 18)
                    np.testing.assert_array_equal(ens.download(m), f)
            ens.step(30, graph=True)
            for m in members:
                assert ens.status(m) == (25, 0)
            finals.append([ens.download(m) for m in members])
    for a, b in zip(*finals):
        np.testing.assert_array_equal(a, b)


def test_guarded_ensemble_members_match_plain_ensembles(gpu_available):
    """The precision guard routes members with tau < FP16_TAU_GUARD to FP32
    storage; every member is bit-identical to a plain ensemble of the storage
    it runs in."""
    from openlb_amd.ensemble import FP16_TAU_GUARD, GuardedEnsemble
    cases = small_members(seed=5, count=6)
    cases[1] = W.lid_driven_cavity((7, 9, 6), tau=0.515, lid_u=0.08)
    cases[4] = W.lid_driven_cavity((5, 8, 11), tau=0.525, lid_u=0.05)
    with GuardedEnsemble.for_cases(cases) as ens:
        members = [ens.add(c, steps=19) for c in cases]
        for m, c in zip(members, cases):
            ens.ini_equilibrium(m, c.rho_u)
        ens.step(25, graph=True)
        got = {m: ens.download(m) for m in members}
        where = {m: ens.storage_of(m) for m in members}
        assert all(ens.status(m) == (19, 0) for m in members)
    for m, c in zip(members, cases):
        want_st = "f32" if 1.0 / c.omega < FP16_TAU_GUARD else "f16"
        assert where[m] == want_st
        with Ensemble(19, want_st, "bgk", max_members=1, population_bytes=16 << 20, cell_capacity=1 << 20) as one:
            k = one.add(c, steps=19)
            one.ini_equilibrium(k, c.rho_u)
            one.step(25)
            np.testing.assert_array_equal(got[m], one.download(k))
