"""Synthetic workloads of the hot path (seeded, reproducible).

Each builder returns a ``Workload``: parameters, positions/velocities in the
reference's (3,n) layout (numpy shape (n, 3), C order == Fortran (3,n)),
per-particle Z and A, and the timestep.  Random numbers come from numpy's
PCG64 generator seeded with the stated seed; species are assigned by a seeded
shuffle; velocities are Maxwell-Boltzmann per mass, the centre-of-mass
momentum removed and then rescaled to exactly 1.5 kT per particle (the
reference semantics of cancel_vcom + tnorm, src/gen_utils.f:64-110).

Workloads (sizes as named):
  pure_ion_3456      3,456 Z=8 A=16 ions, n=7.18e-5 fm^-3, kappa=1, Gamma=175
  ion_mixture_27648  27,648 ions, 0.6 Z8A16 / 0.3 Z26A56 / 0.1 Z34A80,
                     kappa=0.7, Gamma(Z=8)=100
  nucleon_51200      51,200 nucleons, x_p=0.3, rho=0.05 fm^-3, T=1 MeV, HPP +
                     screened p-p Coulomb (lambda_C=10 fm), cutoffs 20 fm
Plasma frequency of a mixture uses the mean ion <Z^2>/<A>.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .params import PI, VC, XMASS_DEFAULT, SimParams


@dataclass
class Workload:
    name: str
    params: SimParams
    x: np.ndarray      # (n, 3) float64, positions (fm)
    v: np.ndarray      # (n, 3) float64, velocities (c)
    Z: np.ndarray      # (n,) charge numbers
    A: np.ndarray      # (n,) mass numbers
    dt: float          # fm/c
    steps: int         # steps of the named workload

    @property
    def n(self) -> int:
        return self.x.shape[0]

    @property
    def pairs(self) -> int:
        return self.n * (self.n - 1) // 2


def wigner_seitz(rho: float) -> float:
    return (3.0 / (4.0 * PI * rho)) ** (1.0 / 3.0)


def maxwell_velocities(rng, A, kT, xmass):
    """Maxwell-Boltzmann velocities, COM momentum removed, KE = 1.5 kT/particle."""
    n = A.shape[0]
    sig = np.sqrt(kT / (A * xmass))
    v = rng.standard_normal((n, 3)) * sig[:, None]
    vcom = (A[:, None] * v).sum(axis=0) / A.sum()
    v -= vcom
    eka = 0.5 * xmass * (A * (v * v).sum(axis=1)).sum() / n
    v *= math.sqrt(1.5 * kT / eka)
    return v


def uniform_positions(rng, n, L):
    return rng.random((n, 3)) * L


def bcc_positions(rng, n, L, jitter=0.02):
    """Body-centred cubic lattice (n = 2 m^3) with a small seeded jitter."""
    m = round((n / 2) ** (1.0 / 3.0))
    if 2 * m ** 3 != n:
        raise ValueError(f"n={n} is not 2*m^3")
    a = L / m
    g = np.arange(m) * a
    cx, cy, cz = np.meshgrid(g, g, g, indexing="ij")
    base = np.stack([cx.ravel(), cy.ravel(), cz.ravel()], axis=1)
    pos = np.concatenate([base, base + 0.5 * a], axis=0)
    pos += rng.standard_normal(pos.shape) * (jitter * a)
    return np.mod(pos, L)


def ion_system(n, species, fractions, rho, kappa, gamma, gamma_z, seed,
               precision="fp64", start="random", steps=0, cutoff_lambdas=12.0):
    """Generic ion plasma (pure if one species).  species: [(Z, A), ...]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    counts = [int(round(f * n)) for f in fractions]
    counts[0] += n - sum(counts)
    Z = np.concatenate([np.full(c, z, dtype=np.float64) for c, (z, _) in zip(counts, species)])
    A = np.concatenate([np.full(c, a, dtype=np.float64) for c, (_, a) in zip(counts, species)])
    if len(species) > 1:
        perm = rng.permutation(n)
        Z, A = Z[perm], A[perm]
    L = (n / rho) ** (1.0 / 3.0)
    a_ws = wigner_seitz(rho)
    lam = a_ws / kappa
    rccut = min(L / 2.0, cutoff_lambdas * lam)
    kT = gamma_z ** 2 * VC / (a_ws * gamma)
    x = uniform_positions(rng, n, L) if start == "random" else bcc_positions(rng, n, L)
    v = maxwell_velocities(rng, A, kT, XMASS_DEFAULT)
    z2 = float((Z * Z).mean())
    abar = float(A.mean())
    omega_p = math.sqrt(4.0 * PI * rho * z2 * VC / (abar * XMASS_DEFAULT))
    dt = 0.05 / omega_p
    pure = len(species) == 1
    params = SimParams(
        sim_type="pure-ion" if pure else "ion-mixture",
        xl=(L, L, L), xmuc=1.0 / lam, rccut=rccut, rho=rho, kT=kT,
        zi=float(species[0][0]) if pure else 0.0,
        ai=float(species[0][1]) if pure else 0.0,
        precision=precision,
    )
    return params, x, v, Z, A, dt


def pure_ion(n=3456, seed=1000, precision="fp64", start="random", kappa=1.0, gamma=175.0):
    p, x, v, Z, A, dt = ion_system(n, [(8.0, 16.0)], [1.0], 7.18e-5, kappa, gamma, 8.0,
                                   seed, precision, start)
    return Workload(f"pure_ion_{n}", p, x, v, Z, A, dt, 1000)


MIX_SPECIES = [(8.0, 16.0), (26.0, 56.0), (34.0, 80.0)]
MIX_FRACTIONS = [0.6, 0.3, 0.1]


def ion_mixture(n=27648, seed=1001, precision="fp64", start="random", kappa=0.7,
                gamma8=100.0):
    p, x, v, Z, A, dt = ion_system(n, MIX_SPECIES, MIX_FRACTIONS, 7.18e-5, kappa, gamma8, 8.0,
                                   seed, precision, start)
    return Workload(f"ion_mixture_{n}", p, x, v, Z, A, dt, 2000)


def nucleon(n=51200, seed=1002, precision="fp64", proton_fraction=0.3, rho=0.05, kT=1.0,
            cutoff=20.0, lambda_c=10.0):
    rng = np.random.Generator(np.random.PCG64(seed))
    npro = int(round(proton_fraction * n))
    Z = np.zeros(n)
    Z[:npro] = 1.0
    Z = Z[rng.permutation(n)]
    A = np.ones(n)
    L = (n / rho) ** (1.0 / 3.0)
    x = uniform_positions(rng, n, L)
    xmass = 939.0
    v = maxwell_velocities(rng, A, kT, xmass)
    params = SimParams(
        sim_type="nucleon", xl=(L, L, L), xmuc=1.0 / lambda_c, rccut=cutoff, rncut=cutoff,
        rho=rho, kT=kT, xmass=xmass, precision=precision,
    )
    return Workload(f"nucleon_{n}", params, x, v, Z, A, 2.0, 5000)


def replica(seed: int, n: int, mixture: bool, gamma: float, kappa: float,
            precision="mixed") -> Workload:
    """One replica of the batched ensemble / one serving request."""
    if mixture:
        p, x, v, Z, A, dt = ion_system(n, MIX_SPECIES, MIX_FRACTIONS, 7.18e-5, kappa, gamma,
                                       8.0, seed, precision)
    else:
        p, x, v, Z, A, dt = ion_system(n, [(8.0, 16.0)], [1.0], 7.18e-5, kappa, gamma, 8.0,
                                       seed, precision)
    return Workload(f"replica_{seed}", p, x, v, Z, A, dt, 0)


def ensemble(replicas: int = 256, seed: int = 1003, precision: str = "mixed",
             n_min: int = 1024, n_max: int = 16384) -> list:
    """The batched ensemble: per replica (seed sequence [seed, r]) N log-uniform
    in [n_min, n_max], pure Z=8 or the three-species mixture with probability
    1/2, Gamma uniform in [150, 220], kappa uniform in [0.5, 2]; same density,
    cutoff min(L/2, 12 lambda) and dt = 0.05/omega_p rule."""
    out = []
    for r in range(replicas):
        rng = np.random.Generator(np.random.PCG64(np.random.SeedSequence([seed, r])))
        n = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
        mixture = bool(rng.random() < 0.5)
        gamma = float(rng.uniform(150.0, 220.0))
        kappa = float(rng.uniform(0.5, 2.0))
        sub_seed = int(rng.integers(0, 2 ** 31 - 1))
        out.append(replica(sub_seed, n, mixture, gamma, kappa, precision))
    return out


def serving_requests(count: int = 2000, seed: int = 1004, n_min: int = 2048,
                     n_max: int = 262144, steps_min: int = 100, steps_max: int = 2000,
                     rate_per_s: float = 10.0, mixed_fraction: float = 0.8):
    """Serving stream: Poisson arrivals at rate_per_s; per request the
    three-species mixture with N log-uniform in [n_min, n_max], steps uniform
    in [steps_min, steps_max], Gamma U[150, 220], kappa U[0.5, 2]; a fraction
    mixed_fraction asks for mixed precision, the rest FP64.  Returns request
    descriptions (the particle state is generated when a request is admitted)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    t = 0.0
    reqs = []
    for k in range(count):
        t += float(rng.exponential(1.0 / rate_per_s))
        reqs.append({
            "id": k, "arrival_s": t,
            "n": int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max))))),
            "steps": int(rng.integers(steps_min, steps_max + 1)),
            "gamma": float(rng.uniform(150.0, 220.0)),
            "kappa": float(rng.uniform(0.5, 2.0)),
            "precision": "mixed" if rng.random() < mixed_fraction else "fp64",
            "seed": int(rng.integers(0, 2 ** 31 - 1)),
        })
    return reqs
