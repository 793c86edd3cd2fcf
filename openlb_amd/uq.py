"""Uncertainty quantification on the ensemble: OpenLB's samplers and statistics
(src/uq) feeding many lattices at once.

``UncertaintyQuantification`` mirrors ``olb::uq::UncertaintyQuantification<T>``
(src/uq/uncertaintyQuantification.h:49-120): the same methods, sampling points
and statistics, so a UQ study written against OpenLB gets the same samples
here.  What OpenLB then does one sample after the other (one simulation per
sampling point, examples/uncertaintyQuantification/cavity3d/cavity3d.cpp:
122-149) runs here as one ensemble: every sample is a member, all members
advance in the same launches, and the statistics of their fields are taken
from the device exports (``ensemble_study``).

Samplers (restated from the reference's headers):
  Monte Carlo      std::mt19937(seed) with libstdc++'s uniform_real_distribution
                   and normal_distribution (Marsaglia polar), one distribution
                   object per dimension, samples drawn point by point
                   (monteCarlo.h:60-110) -- bit-identical points
  GPC / stochastic collocation
                   Gauss quadrature of the Legendre (uniform) or probabilists'
                   Hermite (normal) measure from the eigen-decomposition of the
                   Jacobi matrix (polynomial.hh:80-190, quadrature.h), tensor
                   points with the last dimension fastest, affinely mapped
                   (generalizedPolynomialChaos.hh: findIndex,
                   getStochasticCollocationSample); mean and std from the
                   chaos coefficients (randomToChaos, mean, std)
  quasi Monte Carlo  Sobol (Joe-Kuo direction numbers, unscrambled, first point
                   0) and Halton (first primes, index from 1) mapped to the
                   uniform ranges (quasiMonteCarlo.h)
  Latin hypercube  one stratified uniform draw per interval and dimension,
                   shuffled (latinHypercubeSampling.h); OpenLB seeds it from the
                   clock, here from `seed`
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

UNIFORM, NORMAL = "uniform", "normal"


@dataclasses.dataclass(frozen=True)
class Distribution:
    """uq::Distribution<T> (distribution.h): uniform(min, max) or normal(mean, std)."""
    type: str
    param1: float = 0.0
    param2: float = 1.0


def uniform(lo: float, hi: float) -> Distribution:
    return Distribution(UNIFORM, float(lo), float(hi))


def normal(mean: float, std: float) -> Distribution:
    return Distribution(NORMAL, float(mean), float(std))


def affine(x: float, d: Distribution) -> float:
    """distribution.h affine(): a standard variable to the distribution's."""
    if d.type == UNIFORM:
        return 0.5 * (d.param2 - d.param1) * x + 0.5 * (d.param1 + d.param2)
    return d.param1 + d.param2 * x


# -- std::mt19937 and libstdc++'s real distributions ----------------------------

class MT19937:
    """std::mt19937 (the 32-bit Mersenne twister, seeded as std::mt19937(seed))."""

    def __init__(self, seed: int):
        mt = [0] * 624
        mt[0] = seed & 0xFFFFFFFF
        for i in range(1, 624):
            mt[i] = (1812433253 * (mt[i - 1] ^ (mt[i - 1] >> 30)) + i) & 0xFFFFFFFF
        self.mt, self.index = mt, 624

    def _twist(self):
        mt = self.mt
        for i in range(624):
            y = (mt[i] & 0x80000000) | (mt[(i + 1) % 624] & 0x7FFFFFFF)
            mt[i] = mt[(i + 397) % 624] ^ (y >> 1) ^ (0x9908B0DF if y & 1 else 0)
        self.index = 0

    def __call__(self) -> int:
        if self.index >= 624:
            self._twist()
        y = self.mt[self.index]
        self.index += 1
        y ^= y >> 11
        y ^= (y << 7) & 0x9D2C5680
        y ^= (y << 15) & 0xEFC60000
        y ^= y >> 18
        return y & 0xFFFFFFFF


def generate_canonical(rng: MT19937) -> float:
    """std::generate_canonical<double, 53> over a 32-bit engine: two draws."""
    s = float(rng()) + float(rng()) * 4294967296.0
    r = s / 18446744073709551616.0
    return r if r < 1.0 else math.nextafter(1.0, 0.0)


class _UniformReal:
    def __init__(self, a, b):
        self.a, self.b = a, b

    def __call__(self, rng):
        return generate_canonical(rng) * (self.b - self.a) + self.a


class _Normal:
    """libstdc++'s normal_distribution: Marsaglia's polar method, keeping the
    second variate of each pair for the next call."""

    def __init__(self, mean, std):
        self.mean, self.std, self.saved = mean, std, None

    def __call__(self, rng):
        if self.saved is not None:
            ret, self.saved = self.saved, None
        else:
            while True:
                x = 2.0 * generate_canonical(rng) - 1.0
                y = 2.0 * generate_canonical(rng) - 1.0
                r2 = x * x + y * y
                if r2 <= 1.0 and r2 != 0.0:
                    break
            mult = math.sqrt(-2.0 * math.log(r2) / r2)
            self.saved = x * mult
            ret = y * mult
        return ret * self.std + self.mean


# -- Gauss quadrature and orthogonal polynomials ---------------------------------

def _jacobi_matrix(kind: str, n: int) -> np.ndarray:
    J = np.zeros((n, n))
    for i in range(1, n):
        a = i / math.sqrt(4.0 * i * i - 1.0) if kind == UNIFORM else math.sqrt(i)
        J[i, i - 1] = J[i - 1, i] = a
    return J


def gauss_quadrature(kind: str, nq: int) -> tuple[np.ndarray, np.ndarray]:
    """Nodes (ascending) and probability weights of the nq-point Gauss rule
    of the uniform measure on [-1, 1] (Legendre) or the standard normal
    (Hermite): eigenvalues of the Jacobi matrix and the squared first
    components of its eigenvectors (Golub-Welsch)."""
    if nq == 1:
        return np.zeros(1), np.ones(1)
    w, v = np.linalg.eigh(_jacobi_matrix(kind, nq))
    order = np.argsort(w)
    return w[order], (v[0, order] ** 2)


def _poly_coefficients(kind: str, n: int) -> np.ndarray:
    """Monomial coefficients of P_n (Legendre) or He_n (probabilists' Hermite),
    lowest power first, by the recurrences of polynomial.hh."""
    if kind == UNIFORM:
        p0, p1 = np.array([1.0]), np.array([0.0, 1.0])
        if n == 0:
            return p0
        for k in range(2, n + 1):
            xp = np.concatenate(([0.0], p1))
            p2 = (2.0 * k - 1.0) / k * xp
            p2[: len(p0)] -= (k - 1.0) / k * p0
            p0, p1 = p1, p2
        return p1
    h0, h1 = np.array([1.0]), np.array([0.0, 1.0])
    if n == 0:
        return h0
    for k in range(2, n + 1):
        h2 = np.concatenate(([0.0], h1))
        h2[: len(h0)] -= (k - 1) * h0
        h0, h1 = h1, h2
    return h1


def _evaluate(kind: str, n: int, x: float) -> float:
    c = _poly_coefficients(kind, n)
    r = c[-1]
    for a in c[-2::-1]:
        r = r * x + a
    return r


def multi_indices(dim: int, order: int) -> list[tuple[int, ...]]:
    """calculateMultiIndices: total degree 0..order, first dimension slowest."""
    out = []

    def rec(pos, total, top, index):
        if pos == dim - 1:
            out.append(tuple(index + [top - total]))
            return
        for i in range(top - total + 1):
            rec(pos + 1, total + i, top, index + [i])

    for top in range(order + 1):
        rec(0, 0, top, [])
    return out


def sobol_points(n: int, dim: int) -> np.ndarray:
    """Unscrambled Sobol points with Joe-Kuo direction numbers, first point 0."""
    from scipy.stats import qmc
    return qmc.Sobol(dim, scramble=False).random(n)


def halton_points(n: int, dim: int) -> np.ndarray:
    primes = [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47]
    out = np.zeros((n, dim))
    for d in range(dim):
        base = primes[d]
        for i in range(n):
            index, f, r = i + 1, 1.0 / base, 0.0
            while index > 0:
                r += f * (index % base)
                index //= base
                f /= base
            out[i, d] = r
    return out


# -- the UQ object ----------------------------------------------------------------

class UncertaintyQuantification:
    """uq::UncertaintyQuantification<double>: methods "gpc", "monte_carlo",
    "quasi_monte_carlo", "latin_hypercube"."""

    def __init__(self, method: str):
        if method not in ("gpc", "monte_carlo", "quasi_monte_carlo", "latin_hypercube"):
            raise ValueError(f"unknown UQ method {method}")
        self.method = method
        self._points = None

    # initializeGPC(order, nq, distribution(s))
    def initialize_gpc(self, order: int, nq: int, distributions):
        self._require("gpc")
        dists = list(distributions) if isinstance(distributions, (list, tuple)) else [distributions]
        self.distributions, self.order, self.nq = dists, int(order), int(nq)
        rules = [gauss_quadrature(d.type, self.nq) for d in dists]
        dim = len(dists)
        total = self.nq ** dim
        idx = np.array([np.unravel_index(k, (self.nq,) * dim) for k in range(total)]).reshape(total, dim)
        self.weights = np.ones(total)
        xi = np.zeros((total, dim))
        for j in range(dim):
            xi[:, j] = rules[j][0][idx[:, j]]
            self.weights *= rules[j][1][idx[:, j]]
        self.indices = multi_indices(dim, self.order)
        self.phi = np.array([[np.prod([_evaluate(dists[j].type, ind[j], xi[k, j]) for j in range(dim)])
                              for ind in self.indices] for k in range(total)])   # [points][No]
        self.t2 = (self.phi ** 2 * self.weights[:, None]).sum(axis=0)
        self._points = np.array([[affine(xi[k, j], dists[j]) for j in range(dim)] for k in range(total)])

    # initializeMonteCarlo(numSamples, distribution(s), seed)
    def initialize_monte_carlo(self, num_samples: int, distributions, seed: int):
        self._require("monte_carlo")
        dists = list(distributions) if isinstance(distributions, (list, tuple)) else [distributions]
        rng = MT19937(int(seed))
        gens = [_UniformReal(d.param1, d.param2) if d.type == UNIFORM else _Normal(d.param1, d.param2) for d in dists]
        pts = np.zeros((int(num_samples), len(dists)))
        for i in range(int(num_samples)):
            for j, g in enumerate(gens):
                pts[i, j] = g(rng)
        self.distributions, self._points = dists, pts

    # initializeQuasiMonteCarlo(numSamples, distribution(s), dirFile, generator)
    def initialize_quasi_monte_carlo(self, num_samples: int, distributions, generator: str = "sobol"):
        self._require("quasi_monte_carlo")
        dists = list(distributions) if isinstance(distributions, (list, tuple)) else [distributions]
        if any(d.type != UNIFORM for d in dists):
            raise ValueError("quasi Monte Carlo supports uniform distributions only")
        raw = sobol_points(int(num_samples), len(dists)) if generator == "sobol" else \
            halton_points(int(num_samples), len(dists))
        lo = np.array([d.param1 for d in dists])
        hi = np.array([d.param2 for d in dists])
        self.distributions, self._points = dists, lo + (hi - lo) * raw

    # initializeLatinHypercubeSampling(numSamples, dimension)
    def initialize_latin_hypercube(self, num_samples: int, dimension: int, seed: int = 0):
        self._require("latin_hypercube")
        rng = np.random.Generator(np.random.PCG64(seed))
        n = int(num_samples)
        pts = np.zeros((n, int(dimension)))
        for d in range(int(dimension)):
            pts[:, d] = rng.permutation((np.arange(n) + rng.random(n)) / n)
        self._points = pts

    def sampling_points(self) -> np.ndarray:
        """[samples][dimension], getSamplingPoints()."""
        if self._points is None:
            raise RuntimeError("initialize the UQ method first")
        return self._points

    def samples_number(self) -> int:
        return len(self.sampling_points())

    def mean(self, values) -> np.ndarray:
        """Mean of a response sampled at the points; `values` is [samples, ...]
        and the statistic is taken over the first axis, point by point."""
        v = np.asarray(values, dtype=np.float64)
        if self.method == "gpc":
            return self._chaos(v)[0]
        if self.method == "latin_hypercube":
            raise RuntimeError("OpenLB defines no statistics for Latin hypercube sampling")
        acc = np.zeros(v.shape[1:])
        for row in v:                 # std::accumulate order
            acc = acc + row
        return acc / len(v)

    def std(self, values) -> np.ndarray:
        v = np.asarray(values, dtype=np.float64)
        if self.method == "gpc":
            c = self._chaos(v)
            var = np.zeros(v.shape[1:])
            for i in range(1, len(self.indices)):
                var = var + self.t2[i] * c[i] * c[i]
            return np.sqrt(var)
        if self.method == "latin_hypercube":
            raise RuntimeError("OpenLB defines no statistics for Latin hypercube sampling")
        avg = self.mean(v)
        acc = np.zeros(v.shape[1:])
        for row in v:
            acc = acc + (row - avg) * (row - avg)
        return np.sqrt(acc / (len(v) - 1))

    def _chaos(self, v):
        """randomToChaos: c_i = sum_k w_k v_k phi_i(xi_k) / <phi_i^2>."""
        wv = self.weights.reshape((-1,) + (1,) * (v.ndim - 1)) * v
        out = []
        for i in range(len(self.indices)):
            phi = self.phi[:, i].reshape((-1,) + (1,) * (v.ndim - 1))
            out.append((wv * phi).sum(axis=0) / self.t2[i])
        return out

    def _require(self, method):
        if self.method != method:
            raise RuntimeError(f"this UQ object uses {self.method}, not {method}")


# -- the driver: a UQ study as one ensemble ---------------------------------------

def ensemble_study(uq: UncertaintyQuantification, make_case, steps: int, storage: str = "f32",
                   fields=("velocity",), converter=None, overlap: int = 0, ensemble=None):
    """Run one lattice per sampling point as members of one ensemble and return
    ({field: [samples][points*dim] float32 export}, {field: (mean, std)}).

    make_case(sample) -> workloads.Case builds the lattice of one sampling
    point; every member runs `steps` steps (a member that finishes stops, the
    others go on).  Fields are the device exports of vtk.py (density,
    velocity, physVelocity, physPressure), the statistics OpenLB's
    computeMeanAndStdAndWriteVTI takes point by point (src/uq/postprocessing.h:
    167-260) with uq.mean / uq.std.  All members must share one extent."""
    from .ensemble import Ensemble
    from . import vtk

    samples = uq.sampling_points()
    cases = [make_case(s) for s in samples]
    shape = tuple(cases[0].shape)
    if any(tuple(c.shape) != shape for c in cases):
        raise ValueError("the members of a study must share one extent")
    own = ensemble is None
    if own:
        pop, ids = Ensemble.bytes_for(cases[0].q, storage, shape)
        ensemble = Ensemble(cases[0].q, storage, cases[0].kind, max_members=len(cases),
                            population_bytes=(pop + 4096) * len(cases),
                            cell_capacity=int(np.prod(shape)) * len(cases) + 4096)
    try:
        members = []
        for c in cases:
            m = ensemble.add(c, steps=int(steps))
            ensemble.ini_equilibrium(m, c.rho_u)
            if getattr(c, "local", None):
                ensemble.initialize(m)
            members.append(m)
        ensemble.step(int(steps))
        kinds = {"density": vtk.SuperLatticeDensity3D, "velocity": vtk.SuperLatticeVelocity3D,
                 "physVelocity": vtk.SuperLatticePhysVelocity3D, "physPressure": vtk.SuperLatticePhysPressure3D}
        out, stats = {}, {}
        for name in fields:
            cls = kinds[name]
            arrays = []
            for m in members:
                f = cls(ensemble, converter, member=m) if name.startswith("phys") else cls(ensemble, member=m)
                arrays.append(f.evaluate(overlap))
            arr = np.stack(arrays)
            out[name] = arr
            stats[name] = (uq.mean(arr), uq.std(arr))
        for m in members:
            ensemble.remove(m)
        return out, stats
    finally:
        if own:
            ensemble.close()
