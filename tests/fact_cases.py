"""Synthetic skies, stations and the FP64 oracle of the reference's direct
algorithm for the factorised correlator tests.

The oracle forms J_p(s) = K_p(s) E(s) (oskar_evaluate_jones_K semantics,
oskar/interferometer/define_evaluate_jones_K.h:56-75) and runs the
reference's double-precision cross-correlation on it
(oskar/correlate/src/oskar_cross_correlate_omp.cpp:33-156).
"""
import numpy as np

import oracle

C0 = 299792458.0


def make_case(seed, ns, nsrc, spread=20000.0, pol=True, beam=True):
    rng = np.random.default_rng(seed)
    r = spread / 2 * np.sqrt(rng.uniform(size=ns))
    t = rng.uniform(0, 2 * np.pi, ns)
    u, v = r * np.cos(t), r * np.sin(t)
    w = rng.uniform(-30, 30, ns)
    th = np.arccos(rng.uniform(0.05, 1.0, nsrc))
    ph = rng.uniform(0, 2 * np.pi, nsrc)
    l, m, n = np.sin(th) * np.sin(ph), np.sin(th) * np.cos(ph), np.cos(th)
    I_ = 0.01 * (10.0 / 0.01) ** rng.uniform(size=nsrc)
    if pol:
        p = rng.uniform(0, 0.2, nsrc) * I_
        d = rng.normal(size=(3, nsrc))
        d /= np.linalg.norm(d, axis=0)
        Q, U, V = p * d[0], p * d[1], p * d[2]
    else:
        Q = U = V = np.zeros(nsrc)
    E = None
    if beam:
        E = rng.normal(size=(nsrc, 8)) * 0.3
        E[:, 0] += 1.0
        E[:, 6] += 1.0
    return dict(u=u, v=v, w=w, l=l, m=m, n=n, I=I_, Q=Q, U=U, V=V, E=E)


def rounded(c, prec):
    """The inputs as the GPU receives them: FP32 runs see FP32-rounded
    positions and directions, whose rounding alone moves phases by ~1e-3 rad
    on 10 km baselines, so each precision is checked against the oracle on
    the same rounded inputs."""
    if prec == "d":
        return c
    r = {}
    for k, v in c.items():
        r[k] = None if v is None else np.asarray(v, np.float32).astype(
            np.float64)
    return r


def oracle_direct(c, freq, casa=0, **kw):
    ns, nsrc = c["u"].size, c["l"].size
    k = 2 * np.pi * freq / C0
    K = oracle.jones_K(c["l"], c["m"], c["n"], c["u"], c["v"], c["w"], k,
                       np.ones(nsrc)).reshape(ns, nsrc, 2)
    Kc = K[..., 0] + 1j * K[..., 1]
    if c["E"] is None:
        Ec = np.broadcast_to(np.eye(2), (nsrc, 2, 2))
    else:
        Ec = (c["E"][:, 0::2] + 1j * c["E"][:, 1::2]).reshape(nsrc, 2, 2)
    J = Kc[:, :, None, None] * Ec[None]
    jr = np.stack([J.real, J.imag], -1).ravel()
    z = np.zeros(ns)
    return oracle.cross_correlate(
        "double", casa, 0, 1, jr, c["I"], c["Q"], c["U"], c["V"], c["l"],
        c["m"], c["n"], None, None, None, c["u"], c["v"], c["w"], z, z,
        freq=freq, **kw)


def station_beams(seed, ns, nsrc, spread=0.3):
    """One 2x2 complex beam per (station, source): identity plus complex
    Gaussian noise, (ns, nsrc, 8) reals."""
    rng = np.random.default_rng(seed)
    E = rng.normal(size=(ns, nsrc, 8)) * spread
    E[..., 0] += 1.0
    E[..., 6] += 1.0
    return E


def oracle_direct_stations(c, E, freq, casa=0, **kw):
    """The FP64 oracle of the reference's direct algorithm on
    J_p(s) = K_p(s) E_p(s) with one beam per station."""
    ns, nsrc = c["u"].size, c["l"].size
    k = 2 * np.pi * freq / C0
    K = oracle.jones_K(c["l"], c["m"], c["n"], c["u"], c["v"], c["w"], k,
                       np.ones(nsrc)).reshape(ns, nsrc, 2)
    Kc = K[..., 0] + 1j * K[..., 1]
    Ec = (E[..., 0::2] + 1j * E[..., 1::2]).reshape(ns, nsrc, 2, 2)
    J = Kc[:, :, None, None] * Ec
    jr = np.stack([J.real, J.imag], -1).ravel()
    z = np.zeros(ns)
    return oracle.cross_correlate(
        "double", casa, 0, 1, jr, c["I"], c["Q"], c["U"], c["V"], c["l"],
        c["m"], c["n"], None, None, None, c["u"], c["v"], c["w"], z, z,
        freq=freq, **kw)
