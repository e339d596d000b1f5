"""Synthetic workloads of the visibility hot path at its headline size,
shared by the GPU tests and bench.py.

  - stations: log-radial layout from 100 m to 35 km, w within +-20 m;
  - sources: uniform over the sky above the horizon (n uniform in (0, 1],
    azimuth uniform), Stokes I from a power law (index -1.8, 0.01-10 Jy);
  - one aperture-array station design: elements uniform in a 38 m disc.
"""
import numpy as np


def power_law(rng, n, index=-1.8, lo=0.01, hi=10.0):
    """Inverse CDF of p(S) ~ S^index on [lo, hi]
    (oskar/math/src/oskar_random_power_law.c:15-36)."""
    u = rng.uniform(size=n)
    a = index + 1.0
    return (lo ** a + u * (hi ** a - lo ** a)) ** (1.0 / a)


def stations(rng, num_stations):
    r = np.exp(rng.uniform(np.log(100.0), np.log(35000.0), num_stations))
    t = rng.uniform(0, 2 * np.pi, num_stations)
    return dict(u=r * np.cos(t), v=r * np.sin(t),
                w=rng.uniform(-20.0, 20.0, num_stations))


def sky(rng, num_sources):
    z = rng.uniform(0.0, 1.0, num_sources)
    z[z == 0.0] = 1e-7
    ph = rng.uniform(0, 2 * np.pi, num_sources)
    s = np.sqrt(1.0 - z * z)
    return dict(l=s * np.sin(ph), m=s * np.cos(ph), n=z,
                I=power_law(rng, num_sources))


def elements(rng, num_elements, radius=19.0):
    re = radius * np.sqrt(rng.uniform(size=num_elements))
    te = rng.uniform(0, 2 * np.pi, num_elements)
    return np.stack([re * np.cos(te), re * np.sin(te),
                     np.zeros(num_elements)])


def make_workload(rng, num_stations, num_sources, num_elements):
    st = stations(rng, num_stations)
    sk = sky(rng, num_sources)
    el = elements(rng, num_elements)
    return st, sk, el
