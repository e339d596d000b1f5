"""Pins the CPU oracle (oracle/oskar_oracle.c) before it is trusted.

  - against golden vectors that the reference's own compiled CPU kernels
    produced (tests/golden/make_golden.py);
  - against closed forms and invariants that a dropped term, a wrong sign,
    a wrong index or a transposed operand would break.
"""
import numpy as np
import pytest

import oracle
from helpers import (XCORR_CASES, load_golden, max_rel, rel_rms, xcorr_key)

C0 = 299792458.0


def run_oracle_xcorr(g, prec, matrix, gauss, casa, bs, ts, **kw):
    return oracle.cross_correlate(
        "double" if prec == "d" else "single", casa, gauss, matrix,
        g["jones_m"] if matrix else g["jones_s"], g["I"], g["Q"], g["U"],
        g["V"], g["l"], g["m"], g["n"], g["a"], g["b"], g["c"], g["u"],
        g["v"], g["w"], g["x"], g["y"],
        channel_bw=float(g["channel_bw"]) if bs else 0.0,
        time_avg=float(g["time_avg"]) if ts else 0.0,
        ra0=float(g["ra0"]), dec0=float(g["dec0"]), gast=float(g["gast"]),
        freq=float(g["freq"]), **kw)


@pytest.mark.parametrize("fixture", ["xcorr_ref.npz", "xcorr_long.npz"])
@pytest.mark.parametrize("case", XCORR_CASES,
                         ids=[xcorr_key(*c) for c in XCORR_CASES])
def test_xcorr_oracle_matches_reference_kernels(fixture, case):
    g = load_golden(fixture)
    prec = case[0]
    got = run_oracle_xcorr(g, *case)
    ref = g[xcorr_key(*case)]
    # Same operations in the same order as the reference's OpenMP kernels.
    tol = 1e-13 if prec == "d" else 2e-6
    assert max_rel(got, ref) < tol


def brute_force_vis(J, I_, Q, U, V, casa, smear=None):
    """V_pq = sum_s smear J_first B J_second^H with numpy matrix algebra."""
    ns, nsrc = J.shape[0], J.shape[1]
    B = np.zeros((nsrc, 2, 2), complex)
    B[:, 0, 0] = I_ + Q
    B[:, 0, 1] = U + 1j * V
    B[:, 1, 0] = U - 1j * V
    B[:, 1, 1] = I_ - Q
    out = []
    for q in range(ns):
        for p in range(q + 1, ns):
            a, b = (q, p) if casa else (p, q)
            t = J[a] @ B @ np.conj(np.transpose(J[b], (0, 2, 1)))
            if smear is not None:
                t = t * smear[len(out)][:, None, None]
            out.append(t.sum(0))
    return np.array(out)


def to_complex_matrix(v):
    v = v.reshape(-1, 4, 2)
    return (v[..., 0] + 1j * v[..., 1]).reshape(-1, 2, 2)


@pytest.mark.parametrize("casa", [0, 1])
def test_xcorr_oracle_matrix_algebra_brute_force(casa):
    rng = np.random.default_rng(11)
    ns, nsrc = 6, 40
    J = rng.normal(size=(ns, nsrc, 2, 2)) + 1j * rng.normal(size=(ns, nsrc,
                                                                   2, 2))
    I_ = rng.uniform(1, 2, nsrc)
    Q, U, V = (rng.uniform(-0.2, 0.2, nsrc) for _ in range(3))
    jr = np.stack([J.real, J.imag], -1).ravel()
    z = np.zeros(ns)
    got = oracle.cross_correlate("double", casa, 0, 1, jr, I_, Q, U, V,
                                 np.zeros(nsrc), np.zeros(nsrc),
                                 np.ones(nsrc), None, None, None,
                                 rng.normal(size=ns), rng.normal(size=ns),
                                 rng.normal(size=ns), z, z)
    want = brute_force_vis(J, I_, Q, U, V, casa)
    assert np.max(np.abs(to_complex_matrix(got) - want)) < 1e-12 * np.max(
        np.abs(want))


def test_xcorr_oracle_casa_is_conjugate_transpose():
    g = load_golden("xcorr_long.npz")
    a = to_complex_matrix(run_oracle_xcorr(g, "d", 1, 1, 0, 1, 1))
    b = to_complex_matrix(run_oracle_xcorr(g, "d", 1, 1, 1, 1, 1))
    assert np.max(np.abs(a - np.conj(np.transpose(b, (0, 2, 1))))) < \
        1e-12 * np.max(np.abs(a))


def test_xcorr_oracle_plain_dft_special_case():
    """J = K (scalar phases), point sources, no smearing:
    V_pq = sum_s I_s exp(i k ((u_p - u_q) l + (v_p - v_q) m + (w_p - w_q)(n-1)))."""
    rng = np.random.default_rng(12)
    ns, nsrc = 9, 64
    freq = 120e6
    k = 2 * np.pi * freq / C0
    u, v, w = (rng.uniform(-2000, 2000, ns) for _ in range(3))
    l, m = rng.uniform(-0.1, 0.1, nsrc), rng.uniform(-0.1, 0.1, nsrc)
    n = np.sqrt(1 - l * l - m * m)
    I_ = rng.uniform(0.5, 3, nsrc)
    K = oracle.jones_K(l, m, n, u, v, w, k, I_)
    got = oracle.cross_correlate("double", 0, 0, 0, K, I_, None, None, None,
                                 l, m, n, None, None, None, u, v, w,
                                 np.zeros(ns), np.zeros(ns), freq=freq)
    got = got[0::2] + 1j * got[1::2]
    want = []
    for q in range(ns):
        for p in range(q + 1, ns):
            ph = k * ((u[p] - u[q]) * l + (v[p] - v[q]) * m +
                      (w[p] - w[q]) * (n - 1))
            want.append(np.sum(I_ * np.exp(1j * ph)))
    assert np.max(np.abs(got - np.array(want))) < 1e-11 * np.max(np.abs(want))


def test_xcorr_oracle_smearing_closed_form_single_source():
    """One Gaussian source with bandwidth and time smearing: the visibility
    is B scaled by exp(-(a uu^2 + b 2uu vv + c vv^2)) sinc(.) sinc(.)."""
    ns = 4
    rng = np.random.default_rng(13)
    u, v, w = (rng.uniform(-9000, 9000, ns) for _ in range(3))
    x, y = rng.uniform(-9000, 9000, ns), rng.uniform(-9000, 9000, ns)
    l, m = np.array([0.05]), np.array([-0.03])
    n = np.sqrt(1 - l * l - m * m)
    a, b, c = np.array([2e-8]), np.array([-5e-9]), np.array([3e-8])
    freq, bw, tav, gast, ra0, dec0 = 150e6, 200e3, 8.0, 0.7, 0.2, -0.4
    J = np.zeros((ns, 1, 8))
    J[:, :, 0] = J[:, :, 6] = 1.0
    got = oracle.cross_correlate("double", 0, 1, 1, J.ravel(), np.ones(1),
                                 np.zeros(1), np.zeros(1), np.zeros(1), l, m,
                                 n, a, b, c, u, v, w, x, y, channel_bw=bw,
                                 time_avg=tav, ra0=ra0, dec0=dec0, gast=gast,
                                 freq=freq)
    got = got.reshape(-1, 8)
    inv_wl = freq / C0
    i = 0
    for q in range(ns):
        for p in range(q + 1, ns):
            uu, vv, ww = ((u[p] - u[q]) * inv_wl, (v[p] - v[q]) * inv_wl,
                          (w[p] - w[q]) * inv_wl)
            g = np.exp(-(a * uu ** 2 + b * 2 * uu * vv + c * vv ** 2))
            f = np.pi * bw / freq
            t1 = f * (uu * l + vv * m + ww * (n - 1))
            ha = gast - ra0
            xx = (x[p] - x[q]) * np.pi * inv_wl
            yy = (y[p] - y[q]) * np.pi * inv_wl
            rot = 7.272205217e-5 * tav
            du = (xx * np.cos(ha) - yy * np.sin(ha)) * rot
            tmp = (xx * np.sin(ha) + yy * np.cos(ha)) * rot
            dv, dw = tmp * np.sin(dec0), -tmp * np.cos(dec0)
            t2 = du * l + dv * m + dw * (n - 1)
            s = g * np.sin(t1) / t1 * np.sin(t2) / t2
            assert abs(got[i, 0] - s[0]) < 1e-14
            assert abs(got[i, 6] - s[0]) < 1e-14
            assert np.all(got[i, [1, 2, 3, 4, 5, 7]] == 0)
            i += 1


def test_xcorr_oracle_linearity_and_uv_filter():
    g = load_golden("xcorr_long.npz")
    full = run_oracle_xcorr(g, "d", 1, 1, 0, 1, 1)
    # Two half-sky calls accumulate into the same vis.
    h = dict(g)
    ns = g["l"].size // 2
    first = {k: g[k][:ns] for k in ("I", "Q", "U", "V", "l", "m", "n", "a",
                                    "b", "c")}
    second = {k: g[k][ns:] for k in first}
    nst = g["u"].size
    jm = g["jones_m"].reshape(nst, -1, 8)
    h.update(first)
    h["jones_m"] = jm[:, :ns].ravel()
    part = run_oracle_xcorr(h, "d", 1, 1, 0, 1, 1)
    h.update(second)
    h["jones_m"] = jm[:, ns:].ravel()
    both = run_oracle_xcorr(h, "d", 1, 1, 0, 1, 1, vis=part)
    assert max_rel(both, full) < 1e-12
    # uv filter in metres: filtered baselines keep their previous contents.
    sentinel = np.full_like(full, 7.0)
    filt = run_oracle_xcorr(g, "d", 1, 1, 0, 1, 1, vis=sentinel.copy(),
                            uv_min=5000.0, uv_max=15000.0)
    u, v = g["u"], g["v"]
    i = 0
    for q in range(nst):
        for p in range(q + 1, nst):
            L = np.hypot(u[p] - u[q], v[p] - v[q])
            row = filt[8 * i:8 * i + 8]
            if L < 5000.0 or L > 15000.0:
                assert np.all(row == 7.0)
            else:
                assert np.allclose(row - 7.0, full[8 * i:8 * i + 8],
                                   rtol=1e-12, atol=1e-12)
            i += 1


@pytest.mark.parametrize("prec", ["f", "d"])
@pytest.mark.parametrize("kind", ["c2c", "m2m"])
@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("typed", ["full", "typed"])
def test_dftw_oracle_matches_reference_kernels(prec, kind, dim, typed):
    g = load_golden("dftw.npz")
    matrix = kind == "m2m"
    data = g["data_%s_%s" % ("m" if matrix else "c", typed)]
    got = oracle.dftw("double" if prec == "d" else "single", matrix, 1,
                      float(g["wavenumber"]), g["weights"], g["x_in"],
                      g["y_in"], g["z_in"] if dim == 3 else None, g["x_out"],
                      g["y_out"], g["z_out"] if dim == 3 else None,
                      g["idx"] if typed == "typed" else None, data)
    ref = g["out_%s_%s_%dd_%s" % (prec, kind, dim, typed)]
    assert max_rel(got, ref) < (1e-13 if prec == "d" else 2e-6)


def test_dftw_oracle_dirichlet_closed_form():
    """Uniform linear array, weights steering to l0: the array factor is a
    Dirichlet kernel N at l = l0 and sin(N x/2)/sin(x/2) elsewhere."""
    N, d, freq = 16, 1.5, 100e6
    k = 2 * np.pi * freq / C0
    xe = d * np.arange(N)
    l0 = 0.2
    w = np.exp(-1j * k * xe * l0)
    wr = np.stack([w.real, w.imag], 1).ravel()
    lo = np.linspace(-0.9, 0.9, 181)
    data = np.zeros((N, lo.size, 2))
    data[..., 0] = 1.0
    got = oracle.dftw("double", 0, 0, k, wr, xe, np.zeros(N), None, lo,
                      np.zeros(lo.size), None, None, data.ravel())
    got = got[0::2] + 1j * got[1::2]
    x = k * d * (lo - l0)
    with np.errstate(invalid="ignore", divide="ignore"):
        mag = np.where(np.abs(np.sin(x / 2)) < 1e-12, N,
                       np.abs(np.sin(N * x / 2) / np.sin(x / 2)))
    assert np.max(np.abs(np.abs(got) - mag)) < 1e-11
    i0 = np.argmin(np.abs(lo - l0))
    assert abs(got[i0] - N) < 1e-10


def test_beam_parts_match_reference_kernels():
    g = load_golden("beam_parts.npz")
    w = np.zeros(2 * g["ew_x"].size)
    oracle.lib().oracle_element_weights_dft_d(
        g["ew_x"].size, *[oracle._p(np.ascontiguousarray(g[k])) for k in
                          ("ew_x", "ew_y", "ew_z", "ew_cable")],
        float(g["ew_k"]), *[float(v) for v in g["ew_beam"]], oracle._p(w))
    assert max_rel(w, g["ew_weights"]) < 1e-15
    m = g["tp_x"].size
    th, p1, p2 = np.zeros(m + 1), np.zeros(m + 1), np.zeros(m + 1)
    oracle.lib().oracle_theta_phi_d(
        0, m, *[oracle._p(np.ascontiguousarray(g[k])) for k in
                ("tp_x", "tp_y", "tp_z")], 1, float(g["tp_d1"]),
        float(g["tp_d2"]), oracle._p(th), oracle._p(p1), oracle._p(p2))
    assert np.array_equal(th, g["tp_theta"])
    assert np.array_equal(p1, g["tp_phi1"])
    assert np.array_equal(p2, g["tp_phi2"])
    kL = float(g["dip_kL"])
    for swap in (0, 1):
        pat = np.zeros(8 * (m + 1))
        oracle.lib().oracle_dipole_pattern_d(
            1, m + 1, oracle._p(th), oracle._p(p1), oracle._p(p2), kL,
            np.cos(kL), swap, 0, oracle._p(pat))
        assert max_rel(pat, g["dip_m_swap%d" % swap]) < 1e-15
    pat = np.zeros(2 * (m + 1))
    oracle.lib().oracle_dipole_pattern_d(
        0, m + 1, oracle._p(th), oracle._p(p1), oracle._p(p2), kL,
        np.cos(kL), 0, 0, oracle._p(pat))
    assert max_rel(pat, g["dip_s"]) < 1e-15
    jk = oracle.jones_K(g["k_l"], g["k_m"], g["k_n"], g["k_u"], g["k_v"],
                        g["k_w"], float(g["k_k"]), g["k_I"],
                        float(g["k_fmin"]), float(g["k_fmax"]))
    assert max_rel(jk, g["k_jones"]) < 1e-15


def test_dipole_oracle_closed_form():
    """Half-wave X dipole along phi = 0: in the plane perpendicular to the
    dipole (phi = 90 deg) the pattern is -cos(theta) * (cos(kL cos phi
    sin theta) - cos kL)/(...) = E_theta 0, E_phi = (1 - cos kL)."""
    kL = np.pi / 2
    th = np.linspace(0.0, 1.4, 8)
    phx = np.full(8, np.pi / 2)
    phy = np.zeros(8)
    pat = np.zeros(8 * 8)
    oracle.lib().oracle_dipole_pattern_d(
        1, 8, oracle._p(th), oracle._p(phx), oracle._p(phy), kL, np.cos(kL),
        0, 0, oracle._p(pat))
    pat = pat.reshape(8, 8)
    assert np.allclose(pat[:, 0], 0.0, atol=1e-15)      # X E_theta
    assert np.allclose(pat[:, 2], 1.0 - np.cos(kL))     # X E_phi
    # Y dipole seen along its own axis direction in the (y,z) plane:
    # phi = 0 for the Y feed is broadside, E_theta = -cos(theta)(1-cos kL)/cos^2
    want = -np.cos(th) * (np.cos(kL * np.sin(th)) - np.cos(kL)) / (
        1 + (np.cos(th) ** 2 - 1))
    assert np.allclose(pat[:, 4], want)


def test_station_beam_oracle_isotropic_closed_form():
    """Isotropic elements, unnormalised array: the beam at the pointing
    direction equals N_el; with normalise_array it equals 1."""
    rng = np.random.default_rng(21)
    ne = 32
    r = 19 * np.sqrt(rng.uniform(size=ne))
    t = rng.uniform(0, 2 * np.pi, ne)
    x, y, z = r * np.cos(t), r * np.sin(t), np.zeros(ne)
    bd = np.array([0.2, 0.1, np.sqrt(1 - 0.05)])
    dirs = np.array([bd, [0.0, 0.0, 1.0], [0.5, -0.2, -0.3]]).T
    for norm, want in ((0, ne), (1, 1.0)):
        beam = oracle.station_beam_aa(
            1, x, y, z, x, y, z, np.zeros(ne), np.zeros(ne, np.int32),
            np.array([1], np.int32), np.array([0.0]), bd, dirs[0], dirs[1],
            dirs[2], 100e6, norm_array=norm)
        b = beam.reshape(-1, 8)
        assert abs(b[0, 0] - want) < 1e-12 and abs(b[0, 1]) < 1e-12
        assert abs(b[0, 6] - want) < 1e-12
        assert np.all(b[2] == 0.0)          # below the horizon


def test_factorised_reference_matches_direct_oracle():
    """The long-double factorised reference equals the FP64 direct oracle
    for J_p = K_p E, unpolarised point sources, no smearing."""
    rng = np.random.default_rng(31)
    ns, nsrc = 8, 50
    freq = 90e6
    k = 2 * np.pi * freq / C0
    u, v, w = (rng.uniform(-500, 500, ns) for _ in range(3))
    l, m = rng.uniform(-0.3, 0.3, nsrc), rng.uniform(-0.3, 0.3, nsrc)
    n = np.sqrt(1 - l * l - m * m)
    I_ = rng.uniform(0.1, 5, nsrc)
    E = rng.normal(size=(nsrc, 8))
    K = oracle.jones_K(l, m, n, u, v, w, k, I_).reshape(ns, nsrc, 2)
    Kc = K[..., 0] + 1j * K[..., 1]
    Ec = (E[:, 0::2] + 1j * E[:, 1::2]).reshape(nsrc, 2, 2)
    J = Kc[:, :, None, None] * Ec[None]
    jr = np.stack([J.real, J.imag], -1).ravel()
    z = np.zeros(ns)
    zs = np.zeros(nsrc)
    direct = oracle.cross_correlate("double", 0, 0, 1, jr, I_, zs, zs, zs, l,
                                    m, n, None, None, None, u, v, w, z, z,
                                    freq=freq)
    p, q = [], []
    for qq in range(ns):
        for pp in range(qq + 1, ns):
            p.append(pp)
            q.append(qq)
    fact = oracle.factorised_subset(l, m, n, I_, E.ravel(), u, v, w, k,
                                    np.array(p), np.array(q))
    assert rel_rms(fact, direct) < 1e-13


def _single_dipole_beam(ex, ey, x, y, z, freq, swap_xy=0):
    """The oracle station beam of one dipole at the origin (array factor 1,
    no normalisation): the element pattern alone, (n, 2, 2) complex."""
    zero = np.zeros(1)
    beam = oracle.station_beam_aa(
        1, zero, zero, zero, zero, zero, zero, zero, np.zeros(1, np.int32),
        np.array([0], np.int32), np.array([0.5 * C0 / freq]),
        np.array([0.0, 0.0, 1.0]), x, y, z, freq, euler_x0=ex, euler_y0=ey,
        swap_xy=swap_xy)
    b = beam.reshape(-1, 4, 2)
    return (b[..., 0] + 1j * b[..., 1]).reshape(-1, 2, 2)


@pytest.mark.parametrize("ex_deg,ey_deg,swap_xy", [
    (0.0, 0.0, 0), (90.0, 90.0, 0), (30.0, 30.0, 0), (-50.0, -50.0, 1),
    (20.0, 75.0, 0)])
def test_station_beam_oracle_dipole_axes_closed_form(ex_deg, ey_deg, swap_xy):
    """The dipole branch of the AA-beam oracle (feed offsets from the Euler
    angles, theta/phi, dipole pattern, row order) equals the textbook dipole
    field in vector form: X along (cos ex, sin ex, 0), Y along
    (sin ey, -cos ey, 0).  The X feed is null towards its own axis on the
    horizon and maximal across it; the Y feed likewise."""
    from dipole_closed_form import dipole_matrix, feed_axes
    freq = 120e6
    kh = np.pi / 2                      # half-wave dipole: h = lambda/4
    ex, ey = np.radians(ex_deg), np.radians(ey_deg)
    rng = np.random.default_rng(77)
    n = 400
    th = np.arccos(rng.uniform(0.0, 1.0, n))
    ph = rng.uniform(0, 2 * np.pi, n)
    dx, dy = feed_axes(ex, ey)
    horizon = np.stack([dx, -dx, dy, -dy])
    x = np.concatenate([np.sin(th) * np.cos(ph), horizon[:, 0]])
    y = np.concatenate([np.sin(th) * np.sin(ph), horizon[:, 1]])
    z = np.concatenate([np.cos(th), horizon[:, 2]])
    got = _single_dipole_beam(ex, ey, x, y, z, freq, swap_xy)
    want = dipole_matrix(ex, ey, x, y, z, kh, swap_xy)
    assert np.max(np.abs(got - want)) < 1e-12
    rx, ry = (1, 0) if swap_xy else (0, 1)
    # Nulls along each feed's own axis; the other feed is broadside there
    # (|E| = 1 - cos(kh) = 1 at kh = pi/2).
    hx, hy = got[n:n + 2], got[n + 2:n + 4]
    assert np.all(np.abs(hx[:, rx, :]) < 1e-7)
    assert np.all(np.abs(hy[:, ry, :]) < 1e-7)
    if abs(ex - ey) < 1e-12:
        assert np.allclose(np.linalg.norm(hx[:, ry, :], axis=1), 1.0)
        assert np.allclose(np.linalg.norm(hy[:, rx, :], axis=1), 1.0)
    # The check separates the plausible mistakes: X and Y axes swapped, the
    # Euler angle's sign flipped, the field's sign flipped.
    wrong = [dipole_matrix(ey, ex, x, y, z, kh, 1 - swap_xy),
             -want]
    if ex != 0.0 or ey != 0.0:
        wrong.append(dipole_matrix(-ex, -ey, x, y, z, kh, swap_xy))
    for w in wrong:
        assert np.max(np.abs(got - w)) > 0.1


def test_station_beam_oracle_dipole_array_closed_form():
    """Dipoles at random positions with a common orientation: the beam is
    sum_a exp(i k x_a.(s - b)) g(s) with g the textbook dipole field."""
    from dipole_closed_form import dipole_matrix
    freq = 150e6
    k = 2 * np.pi * freq / C0
    rng = np.random.default_rng(78)
    ne = 24
    r = 12 * np.sqrt(rng.uniform(size=ne))
    t = rng.uniform(0, 2 * np.pi, ne)
    xe, ye, ze = r * np.cos(t), r * np.sin(t), np.zeros(ne)
    b = np.array([0.3, -0.2, np.sqrt(1 - 0.13)])
    n = 300
    th = np.arccos(rng.uniform(0.0, 1.0, n))
    ph = rng.uniform(0, 2 * np.pi, n)
    x, y, z = np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)
    ex = np.radians(40.0)
    beam = oracle.station_beam_aa(
        1, xe, ye, ze, xe, ye, ze, np.zeros(ne), np.zeros(ne, np.int32),
        np.array([0], np.int32), np.array([0.5 * C0 / freq]), b, x, y, z,
        freq, euler_x0=ex, euler_y0=ex)
    bb = beam.reshape(-1, 4, 2)
    got = (bb[..., 0] + 1j * bb[..., 1]).reshape(-1, 2, 2)
    af = np.exp(1j * k * (np.outer(x, xe) + np.outer(y, ye)
                          - (xe * b[0] + ye * b[1]))).sum(1)
    want = af[:, None, None] * dipole_matrix(ex, ex, x, y, z, np.pi / 2)
    assert np.max(np.abs(got - want)) < 1e-10 * ne
