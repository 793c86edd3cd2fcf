"""GPU tests of OpenLB checkpoints (lbmh_checkpoint_*, lbmh_ensemble_checkpoint_*).

The checkers are OpenLB's own checkpoint files and populations
(tests/golden/checkpoints/) and the oracle's restatement of the layout
(oracle/checkpoint.py), which test_checkpoint_cpu.py pins to those files byte
for byte.

* The library resumes from OpenLB's checkpoint after k steps and reaches
  OpenLB's populations after m steps, within the time-step bounds of the
  parity tests (FP64 1e-12, FP32 2e-5 of max |f_i|, populations of the cells
  that carry the flow, except those that stream in from beyond a
  non-periodic face: the overlap leftovers there depend on how each side
  fills its overlap and are discarded by every dynamics).
* A checkpoint the library loads and saves again is the same bytes.
* A checkpoint the library writes after further steps has OpenLB's shifts
  for that step count, the library's own core populations in rotated order,
  and every other field value OpenLB's set-up writes, overlap included.
* Stopping, saving, destroying and resuming a lattice or an ensemble member
  gives bit for bit the populations of the uninterrupted run, in every storage
  type (FP16 storage through FP32 checkpoints).
* A checkpoint whose size or field values do not fit the lattice is refused
  and leaves the lattice unchanged.
"""
import numpy as np
import pytest

import golden_cases as G
from oracle import checkpoint as K
from oracle import oracle as O
from openlb_amd import abi
from openlb_amd import checkpoint as CK
from openlb_amd import workloads as W
from openlb_amd.ensemble import Ensemble
from openlb_amd.lattice import BlockLattice

pytestmark = pytest.mark.gpu

TOL = {"f64": 1e-12, "f32": 2e-5}


def flow_populations(case):
    """The populations of the cells that carry the flow (bulk and boundary
    cells), without the slots that stream in from beyond a non-periodic face."""
    role, _ = K.cell_roles(case)
    flow = np.isin(role, ["bulk", "velocity", "pressure"])
    return flow[None] & ~G.streamed_from_outside(case)


@pytest.mark.parametrize("name", G.checkpoint_names())
def test_resume_from_openlb_checkpoint(gpu_available, name):
    case, meta, arrays, text = G.load_checkpoint(name)
    k, m = meta["steps"]
    real = meta["precision"]
    print(f"checkpoint {name}: restore at step {k}, run to {m}", flush=True)
    with BlockLattice.from_case(case, real) as lat:
        lat.restore(CK.decode_dat(text), real)
        n = int(np.prod(K.padded_shape(case)))
        np.testing.assert_array_equal(lat.shifts(), K.rotate_shifts(case, k) % n)
        np.testing.assert_array_equal(lat.download_populations("f64"), arrays[f"f_{k}"])
        lat.collide_and_stream(m - k)
        got = lat.download_populations("f64")
    err = G.rel_error(got, arrays[f"f_{m}"], case.q, flow_populations(case))
    assert err <= TOL[real], err


@pytest.mark.parametrize("name", G.checkpoint_names())
def test_load_then_save_reproduces_the_file(gpu_available, name):
    case, meta, _, text = G.load_checkpoint(name)
    real = meta["precision"]
    raw = CK.decode_dat(text)
    with BlockLattice.from_case(case, real) as lat:
        lat.restore(raw, real)
        assert lat.checkpoint(real) == raw


@pytest.mark.parametrize("name", G.checkpoint_names())
def test_save_after_steps_has_openlb_layout(gpu_available, name):
    case, meta, _, text = G.load_checkpoint(name)
    k, m = meta["steps"]
    real = meta["precision"]
    with BlockLattice.from_case(case, real) as lat:
        lat.restore(CK.decode_dat(text), real)
        lat.collide_and_stream(m - k)
        core = lat.download_populations(real)
        raw = lat.checkpoint(real)
    assert len(raw) == K.serializable_size(case, real)
    shifts, pops, fields = K.unpack(case, raw, real)
    np.testing.assert_array_equal(shifts, K.rotate_shifts(case, m))
    np.testing.assert_array_equal(K.core_of(case, pops), core)
    want = K.unpack(case, K.pack(case, K.padded_of(case, core), shifts, real), real)[2]
    assert fields.keys() == want.keys()
    for key in fields:
        np.testing.assert_array_equal(fields[key], want[key], err_msg=key)


def _resume_cases():
    return {
        "couette_forced": W.couette_channel((9, 7, 6), tau=0.7, wall_u=0.05, force=1e-5),
        "velocity_pressure": W.velocity_channel((12, 8, 6), u_in=0.03, rho_out=1.0, tau=0.6),
        "interp_channel": W.velocity_channel((12, 8, 6), u_in=0.03, rho_out=1.0, tau=0.6,
                                             inlet="interp_velocity", outlet="interp_pressure"),
        "tgv27": W.taylor_green((8, 6, 10), u0=0.05, tau=0.6, q=27),
    }


@pytest.mark.parametrize("storage", ["f64", "f32", "f16"])
@pytest.mark.parametrize("which", sorted(_resume_cases()))
def test_stop_and_resume_is_bit_identical(gpu_available, which, storage):
    case = _resume_cases()[which]
    real = "f64" if storage == "f64" else "f32"
    f0 = O.initial_populations(case, storage)
    print(f"resume {which} {storage}", flush=True)
    with BlockLattice.from_case(case, storage) as lat:
        lat.upload_populations(f0, "f64")
        lat.initialize()
        lat.collide_and_stream(23)
        straight = lat.download_populations(storage)
    with BlockLattice.from_case(case, storage) as lat:
        lat.upload_populations(f0, "f64")
        lat.initialize()
        lat.collide_and_stream(9)
        saved = lat.checkpoint(real)
    with BlockLattice.from_case(case, storage) as lat:
        lat.restore(saved, real)
        lat.collide_and_stream(14)
        resumed = lat.download_populations(storage)
    np.testing.assert_array_equal(resumed, straight)


def test_file_round_trip(gpu_available, tmp_path):
    case = W.couette_channel((9, 7, 6), tau=0.7, wall_u=0.05, force=1e-5)
    with BlockLattice.from_case(case, "f64") as lat:
        lat.upload_populations(O.initial_populations(case, "f64"), "f64")
        lat.collide_and_stream(5)
        path = lat.save(str(tmp_path / "couette"))
        assert path.endswith("couette_rank0_size1.dat")
        want = lat.download_populations("f64")
    with BlockLattice.from_case(case, "f64") as lat:
        lat.load(str(tmp_path / "couette"))
        np.testing.assert_array_equal(lat.download_populations("f64"), want)


def test_restore_refuses_a_foreign_checkpoint(gpu_available):
    case = W.couette_channel((9, 7, 6), tau=0.7, wall_u=0.05, force=1e-5)
    with BlockLattice.from_case(case, "f64") as lat:
        lat.upload_populations(O.initial_populations(case, "f64"), "f64")
        lat.collide_and_stream(3)
        good = lat.checkpoint("f64")
        before = lat.download_populations("f64")
        shifts = lat.shifts()
        with pytest.raises(abi.LbmhError, match="bytes"):
            lat.restore(good[:-8], "f64")
        with pytest.raises(abi.LbmhError, match="bytes"):
            lat.restore(good, "f32")          # a float lattice's checkpoint has other sizes
    other = W.couette_channel((9, 7, 6), tau=0.7, wall_u=0.05, force=2e-5)
    with BlockLattice.from_case(other, "f64") as lat2:
        lat2.upload_populations(before, "f64")
        with pytest.raises(abi.LbmhError, match="FORCE"):
            lat2.restore(good, "f64")
        np.testing.assert_array_equal(lat2.download_populations("f64"), before)
        np.testing.assert_array_equal(lat2.shifts(), np.zeros_like(shifts))
    moved = W.couette_channel((9, 7, 6), tau=0.7, wall_u=0.06, force=1e-5)
    with BlockLattice.from_case(moved, "f64") as lat3:
        with pytest.raises(abi.LbmhError, match="VELOCITY"):
            lat3.restore(good, "f64")


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_ensemble_member_preempt_and_resume(gpu_available, storage):
    """Serving preemption: a member is saved to host memory and removed; later
    it is added again (in whatever slot is free) and restored, and finishes
    bit-identical to a lattice that never stopped."""
    real = "f32"
    cases = [W.lid_driven_cavity((10, 9, 8), tau=0.6, lid_u=0.08),
             W.couette_channel((9, 7, 6), tau=0.7, wall_u=0.05, force=0.0, kind="bgk"),
             W.local_velocity_cavity((9, 8, 7), tau=0.7, lid_u=0.06)]
    total, cut = 31, 12
    with Ensemble(19, storage, "bgk", max_members=8, population_bytes=16 << 20, cell_capacity=1 << 18) as ens:
        members, f0s = [], []
        for c in cases:
            mbr = ens.add(c, steps=total)
            f0 = O.initial_populations(c, storage)
            ens.upload(mbr, f0)
            if c.local:
                ens.initialize(mbr)
            members.append(mbr)
            f0s.append(f0)
        for mbr in members:
            ens.set_steps(mbr, cut)
        ens.step(cut)
        saved = []
        for mbr, c in zip(members, cases):
            raw = ens.checkpoint(mbr, real)
            with BlockLattice.from_case(c, storage) as lat:
                lat.upload_populations(f0s[cases.index(c)], "f64")
                if c.local:
                    lat.initialize()
                lat.collide_and_stream(cut)
                assert lat.checkpoint(real) == raw   # a member's bytes are a lattice's bytes
            saved.append(raw)
        # preempt member 1, let the others go on, resume it in a new slot
        ens.remove(members[1])
        for mbr in (members[0], members[2]):
            ens.set_steps(mbr, total - cut)
        ens.step(5)
        again = ens.add(cases[1], steps=total - cut)
        ens.restore(again, saved[1], real)
        ens.step(total)
        for mbr, c, f0 in zip([members[0], again, members[2]], cases, f0s):
            with BlockLattice.from_case(c, storage) as lat:
                lat.upload_populations(f0, "f64")
                if c.local:
                    lat.initialize()
                lat.collide_and_stream(total)
                np.testing.assert_array_equal(ens.download(mbr, storage), lat.download_populations(storage))
