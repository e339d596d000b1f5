"""Writes tests/golden/beam_direction.npz: great-circle distances from the
reference's own oskar_angular_distance (oskar/math/src/
oskar_angular_distance.c, compiled into oracle/_ref/liboskar_sky_ref.so by
``make -C oracle ref``), the function that snaps a station beam to the
nearest permitted direction.

Run in a container that has the reference tree (the GPU box does not):
    make -C oracle ref && python tests/golden/make_beam_direction_golden.py

Inputs: 2000 pairs of (lon, lat) over the sphere, 200 of them within
1e-6 rad of each other and 200 near antipodal; output: the distances.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.abspath(os.path.join(HERE, "..", "..")))
import oracle  # noqa: E402


def main():
    ref = oracle.ref_sky()
    if ref is None:
        raise SystemExit("oracle/_ref/liboskar_sky_ref.so missing: "
                         "run `make -C oracle ref` first")
    rng = np.random.default_rng(13)
    n = 2000
    lon1 = rng.uniform(-np.pi, np.pi, n)
    lat1 = np.arcsin(rng.uniform(-1, 1, n))
    lon2 = rng.uniform(-np.pi, np.pi, n)
    lat2 = np.arcsin(rng.uniform(-1, 1, n))
    lon2[:200] = lon1[:200] + rng.uniform(-1e-6, 1e-6, 200)
    lat2[:200] = np.clip(lat1[:200] + rng.uniform(-1e-6, 1e-6, 200),
                         -np.pi / 2, np.pi / 2)
    lon2[200:400] = lon1[200:400] + np.pi + rng.uniform(-1e-4, 1e-4, 200)
    lat2[200:400] = -lat1[200:400] + rng.uniform(-1e-4, 1e-4, 200)
    dist = np.array([ref.oskar_angular_distance(a, b, c, d) for a, b, c, d
                     in zip(lon1, lon2, lat1, lat2)])
    path = os.path.join(HERE, "beam_direction.npz")
    np.savez_compressed(path, lon1=lon1, lon2=lon2, lat1=lat1, lat2=lat2,
                        dist=dist)
    print(path, os.path.getsize(path))


if __name__ == "__main__":
    main()
