"""GPU test of the UQ driver (openlb_amd/uq.py ensemble_study): OpenLB's
samplers feeding the ensemble.

A UQ study of lid-driven cavities (lid velocity and relaxation time uncertain)
runs every sampling point as a member of one ensemble.  Each member's exported
fields must equal, bit for bit, the same lattice run alone (the single-lattice
path is pinned to OpenLB by the parity tests), and the statistics must be
uq.mean / uq.std of those fields point by point (pinned to OpenLB's UQ module
by test_uq_cpu.py)."""
import numpy as np
import pytest

from openlb_amd import uq as U
from openlb_amd import vtk as V
from openlb_amd import workloads as W
from openlb_amd.lattice import BlockLattice

pytestmark = pytest.mark.gpu


def cavity(sample):
    tau, lid = (float(sample[0]), float(sample[1])) if len(sample) > 1 else (0.6, float(sample[0]))
    return W.lid_driven_cavity((14, 12, 10), tau=tau, lid_u=lid)


def local_cavity(sample):
    return W.local_velocity_cavity((12, 10, 9), tau=0.7, lid_u=float(sample[0]))


def alone(case, storage, steps, field):
    with BlockLattice.from_case(case, storage) as lat:
        lat.ini_equilibrium(case.rho_u)
        if case.local:
            lat.initialize()
        lat.collide_and_stream(steps)
        f = {"velocity": V.SuperLatticeVelocity3D, "density": V.SuperLatticeDensity3D}[field](lat)
        return f.evaluate(0)


@pytest.mark.parametrize("setup", ["gpc_lid", "mc_tau_lid", "qmc_tau_lid", "gpc_local"])
def test_ensemble_study_matches_lattices_run_alone(gpu_available, setup):
    storage, steps = "f32", 40
    if setup == "gpc_lid":
        u = U.UncertaintyQuantification("gpc")
        u.initialize_gpc(2, 5, U.uniform(0.045, 0.055))
        make = cavity
    elif setup == "mc_tau_lid":
        u = U.UncertaintyQuantification("monte_carlo")
        u.initialize_monte_carlo(10, [U.uniform(0.52, 1.2), U.uniform(0.02, 0.1)], 123456)
        make = cavity
    elif setup == "qmc_tau_lid":
        u = U.UncertaintyQuantification("quasi_monte_carlo")
        u.initialize_quasi_monte_carlo(8, [U.uniform(0.52, 1.2), U.uniform(0.02, 0.1)], "sobol")
        make = cavity
    else:
        u = U.UncertaintyQuantification("gpc")
        u.initialize_gpc(2, 3, U.normal(0.05, 0.005))
        make = local_cavity
    fields, stats = U.ensemble_study(u, make, steps, storage, fields=("velocity", "density"))
    samples = u.sampling_points()
    assert fields["velocity"].shape[0] == len(samples)
    for k, s in enumerate(samples):
        np.testing.assert_array_equal(fields["velocity"][k], alone(make(s), storage, steps, "velocity"))
    for name in ("velocity", "density"):
        mean, std = stats[name]
        np.testing.assert_array_equal(mean, u.mean(fields[name]))
        np.testing.assert_array_equal(std, u.std(fields[name]))
    # the lid speed drives the flow: the velocity spread is non-zero
    assert np.abs(stats["velocity"][1]).max() > 0
