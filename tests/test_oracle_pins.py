"""Pins of the oracle's functions by what the mathematics and physics fix,
independently of the oracle's own formulas.

- dustmat (dustmat.f90:42-67): at g = 0, pl = 1 the matrix is Rayleigh
  scattering (Chandrasekhar): p1 = 3/4 (1 + mu^2), p2/p1 = -(1 - mu^2)/(1 + mu^2),
  p3/p1 = 2 mu/(1 + mu^2), p4 = 0, at every mu.
- stokes / stokespeel (stokes.f90:117-336, stokespeel.f90:77-322) with
  polarized input, both ri1 branches: against S' = R(psi1) M(Theta) R(psi0) S
  built from the geometry alone (scattering plane from the two directions,
  meridian planes, frame rotations of the Stokes vector).  Of the four sign
  conventions for (Q, U) the construction admits, exactly one reproduces
  every case; the rotation angle ri1 equals the angle from the incident
  meridian plane to the scattering plane (mod pi); findangle3 (bit-exact to
  the compiled reference) supplies ri1 for stokespeel.
- tfinal (tfinal_5_test1.f90:206-281) with asymmetric upper/lower tallies and
  densities: T(upper) = (T(na_up, rho_up V_up) + T(na_low, rho_low V_up)) / 2
  (the mirror cell's mass uses the upper cell's volume), the lower half a
  mirror, phi copies of ip = 1; the inversion through np.interp.
- dusttemp normalization (dusttemp_5.f90:60-61) end to end: the Lucy
  temperatures of an optically thin model equal radiative equilibrium in the
  diluted stellar field, kP(T) T^4 = <W> kP(T*) T*^4, with Planck means of
  kappa_abs by quadrature and the dilution W = (1 - sqrt(1 - (R*/r)^2))/2
  averaged over the cell volume.
- lucydustfreq (lucydustfreq_2_test.f90:19-160) at temperatures between
  table nodes: sampled frequencies against the inverse CDF of
  kappa_abs B_nu(T) computed by quadrature.
- diskflux (flux_test.f90:19-80): in an empty medium the normalized escaped
  and peeled SEDs sum to L*/Lsun at every inclination; the error columns
  equal sqrt(N) std(w) / sum(w) (+ 1/N for si2) from numpy's sample standard
  deviation of the weights of a bin.
"""
import ctypes as C
import math

import numpy as np
import pytest

from src3_mpi_cfitsio_amd import model as M


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


# ---------------------------------------------------------------------------
# dustmat: Rayleigh limit
# ---------------------------------------------------------------------------
def test_dustmat_rayleigh_limit(oracle_lib):
    L = oracle_lib.oracle()
    p = np.zeros(4)
    for mu in np.linspace(-1.0, 1.0, 401):
        L.orc_dustmat(0.0, 1.0, mu, mu * mu, _d(p))
        assert abs(p[0] - 0.75 * (1.0 + mu * mu)) < 1e-14
        assert abs(p[1] / p[0] + (1.0 - mu * mu) / (1.0 + mu * mu)) < 1e-14
        assert abs(p[2] / p[0] - 2.0 * mu / (1.0 + mu * mu)) < 1e-14
        assert p[3] == 0.0
    # pl scales the linear polarization only: p2(pl) = pl p2(1), p1 and p3 unchanged
    for g in (0.0, 0.3, 0.6):
        for mu in (-0.7, -0.1, 0.2, 0.9):
            a, b = np.zeros(4), np.zeros(4)
            L.orc_dustmat(g, 1.0, mu, mu * mu, _d(a))
            L.orc_dustmat(g, 0.37, mu, mu * mu, _d(b))
            assert abs(b[1] - 0.37 * a[1]) < 1e-15 * abs(a[0]) + 1e-300
            assert b[0] == a[0] and b[2] == a[2]


# ---------------------------------------------------------------------------
# stokes / stokespeel against a rotation-matrix construction
# ---------------------------------------------------------------------------
def _u(c, s, ph):
    return np.array([s * math.cos(ph), s * math.sin(ph), c])


def _e_theta(u):
    s = math.hypot(u[0], u[1])
    ph = math.atan2(u[1], u[0])
    return np.array([u[2] * math.cos(ph), u[2] * math.sin(ph), -s])


def _rot(S, psi):
    """Stokes vector in the frame turned by psi (a' = cos psi a + sin psi b)."""
    c2, s2 = math.cos(2.0 * psi), math.sin(2.0 * psi)
    return np.array([S[0], S[1] * c2 + S[2] * s2, -S[1] * s2 + S[2] * c2, S[3]])


def _angle(a, b, v):
    return math.atan2(float(np.dot(v, b)), float(np.dot(v, a)))


def _construct(L, g, pl, u0, u1, S, sq, su):
    """S' = R M R S for incident u0, scattered u1; Stokes (Q, U) of the
    reference taken as (sq Q, su U) of the meridian frame (e_theta, u x e_theta)."""
    a0, a1 = _e_theta(u0), _e_theta(u1)
    b0, b1 = np.cross(u0, a0), np.cross(u1, a1)
    n = np.cross(u0, u1)
    n /= np.linalg.norm(n)
    s0, s1 = np.cross(n, u0), np.cross(n, u1)          # in-plane axes of the scattering frame
    mu = float(np.dot(u0, u1))
    p = np.zeros(4)
    L.orc_dustmat(g, pl, mu, mu * mu, _d(p))
    Mm = np.array([[p[0], p[1], 0, 0], [p[1], p[0], 0, 0], [0, 0, p[2], p[3]], [0, 0, -p[3], p[2]]])
    Sm = np.array([S[0], sq * S[1], su * S[2], S[3]])
    So = _rot(Mm @ _rot(Sm, _angle(a0, b0, s0)), _angle(s1, np.cross(u1, s1), a1))
    return np.array([So[0], sq * So[1], su * So[2], So[3]]), _angle(a0, b0, s0)


def _cases(n, seed):
    rng = np.random.default_rng(seed)
    for _ in range(n):
        g, pl = rng.uniform(0.0, 0.8), rng.uniform(0.1, 0.9)
        ct = math.cos(rng.uniform(0.1, math.pi - 0.1))
        st, ph = math.sqrt(1.0 - ct * ct), rng.uniform(0.0, 2.0 * math.pi)
        q, u, v = rng.uniform(-0.5, 0.5), rng.uniform(-0.45, 0.45), rng.uniform(-0.3, 0.3)
        bmu, ri1 = rng.uniform(-0.95, 0.95), rng.uniform(0.0, 2.0 * math.pi)
        yield g, pl, ct, st, ph, np.array([1.0, q, u, v]), bmu, ri1


def test_stokes_polarized_equals_rotation_construction(oracle_lib):
    L = oracle_lib.oracle()
    convs = [(sq, su) for sq in (1, -1) for su in (1, -1)]
    hits = {c: 0 for c in convs}
    n = branch_hi = 0
    for g, pl, ct, st, ph, S, bmu, ri1 in _cases(600, 21):
        d = np.array([ct, st, ph, math.cos(ph), math.sin(ph)])
        s = S.copy()
        L.orc_stokes_apply(g, pl, bmu, ri1, _d(d), _d(s))
        if d[1] < 1e-4:
            continue
        u0, u1 = _u(ct, st, ph), _u(d[0], d[1], d[2])
        n += 1
        branch_hi += ri1 > math.pi
        for cv in convs:
            So, psi0 = _construct(L, g, pl, u0, u1, S, *cv)
            # stokes keeps the weight: the construction scaled to the input I
            if np.allclose(s, So * (S[0] / So[0]), rtol=0, atol=1e-10):
                hits[cv] += 1
            if cv == (1, -1):
                assert abs(((ri1 - psi0 + 0.5 * math.pi) % math.pi) - 0.5 * math.pi) < 1e-7
    assert n > 550 and 150 < branch_hi < n - 150
    # Q against the meridian plane, U of the opposite handedness: every case,
    # and no other convention fits any case with U != 0
    assert hits[(1, -1)] == n
    assert max(hits[c] for c in convs if c != (1, -1)) < n // 10


def test_stokespeel_polarized_equals_rotation_construction(oracle_lib):
    L = oracle_lib.oracle()
    n = 0
    for g, pl, ct, st, ph, S, _, _ in _cases(600, 22):
        # an observer direction; findangle3 gives (cos Theta, ri1) of the peel
        ce = math.cos(np.random.default_rng(n).uniform(0.1, math.pi - 0.1))
        se, pe = math.sqrt(1.0 - ce * ce), (ph + 1.0 + n * 0.37) % (2.0 * math.pi)
        q4 = [C.c_double() for _ in range(4)]
        L.orc_findangle3(ct, st, ph, math.cos(ph), math.sin(ph), ce, se, pe, math.cos(pe), math.sin(pe),
                         *[C.byref(x) for x in q4])
        bmu, ri1 = q4[0].value, q4[3].value
        u0, u1 = _u(ct, st, ph), _u(ce, se, pe)
        if np.linalg.norm(np.cross(u0, u1)) < 1e-3:
            continue
        sp = S.copy()
        L.orc_stokespeel(g, pl, _d(sp), ce, se, ri1, bmu, ct, st)
        So, _ = _construct(L, g, pl, u0, u1, S, 1, -1)
        np.testing.assert_allclose(sp, So, rtol=0, atol=1e-9 * max(1.0, abs(So[0])))
        n += 1
    assert n > 550


# ---------------------------------------------------------------------------
# tfinal with asymmetric halves
# ---------------------------------------------------------------------------
def _T_interp(m, na, nphot, mcell):
    """dusttemp_5.f90:60-61 normalization, inverted through np.interp on
    T^4 kP(T) (the inversion pinned by test_dusttemp_inverts_kappaP_T4)."""
    kT = m.Tint ** 4 * m.kappap[0]
    if na <= 0.0:
        return 0.1 / 11605.0
    a = m.tstar ** 4 * na * math.pi * m.rstar ** 2 / (nphot * m.kappav[0]) * M.RSUN ** 2 / mcell ** 4
    return float(np.interp(a, kT, m.Tint))


def test_tfinal_asymmetric_halves(oracle_lib):
    m = M.small("asym", nr=8, nt=6, nphi=3)
    rng = np.random.default_rng(31)
    nt, ntg = m.ntg - 1, m.ntg
    rho = m.rho.copy()
    rho[rho > 0] *= rng.uniform(0.3, 3.0, size=int((rho > 0).sum()))       # lower and upper differ
    m.rho[...] = rho
    # lower-half volumes differ from their mirrors (on a symmetric grid they
    # are equal, which would hide whose volume the mirror mass takes)
    m.vcell[:, nt // 2:, :] *= rng.uniform(0.5, 2.0, size=m.vcell[:, nt // 2:, :].shape)
    dt = rng.uniform(1e-6, 1e-3, m.cell_shape)
    dt[:, nt // 2:, :] *= rng.uniform(0.1, 10.0, size=dt[:, nt // 2:, :].shape)
    nphot = 2e5
    T = oracle_lib.tfinal(m, dt, nphot)
    checked = 0
    for ir in range(m.nrg - 1):
        for it in range(nt // 2):                # 0-based upper index; mirror ntg - it - 2
            lo = ntg - (it + 1) - 1
            if m.species[0, it, ir] == 0 or m.rho[0, it, ir] <= 0:
                assert T[0, it, ir] == 0.1 / 11605.0
                continue
            v_up = m.vcell[0, it, ir]
            t1 = _T_interp(m, dt[0, it, ir], nphot, (m.rho[0, it, ir] * v_up) ** 0.25)
            t2 = _T_interp(m, dt[0, lo, ir], nphot, (m.rho[0, lo, ir] * v_up) ** 0.25)
            want = min((t1 + t2) / 2.0, 199999.0 / 11605.0)
            assert abs(T[0, it, ir] - want) <= 1e-9 * want, (ir, it)
            assert T[0, lo, ir] == T[0, it, ir]
            checked += 1
    for ip in range(1, m.npg - 1):
        np.testing.assert_array_equal(T[ip], T[0])
    assert checked > 10


# ---------------------------------------------------------------------------
# dusttemp normalization: optically thin radiative equilibrium
# ---------------------------------------------------------------------------
def _kabs(m, nu):
    nud, kap, alb = m.nudust[0][::-1], m.kappad[0][::-1], m.adust[0][::-1]   # ascending in nu
    return np.interp(nu, nud, kap * (1.0 - alb))


_XG = np.linspace(1e-5, 60.0, 120001)


def _planck_mean(m, T):
    """Planck mean of kappa_abs at T by the trapezoid rule on a fine grid in x = nu/T."""
    f = _kabs(m, _XG * T) * _XG ** 3 / np.expm1(_XG)
    return float(np.trapezoid(f, _XG)) / (math.pi ** 4 / 15.0)


def test_dusttemp_thin_radiative_equilibrium(oracle_lib):
    m = M.small("thin", nr=8, nt=4, nphi=1, nfreq=10)
    # optical depth along a radius below 1e-2 at every wavelength
    m.rho[...] = np.where(m.rho > 0, 1e-21, 0.0)
    n = 200_000
    o = oracle_lib.run_packets(m, n, seed=41, nthreads=4)
    T = oracle_lib.tfinal(m, o["dtauabs"], n)
    kps = _planck_mean(m, m.tstar)
    checked = 0
    for ir in range(m.nrg - 1):
        r1, r2 = m.rarr[ir], m.rarr[ir + 1]
        rr = np.linspace(r1, r2, 2001)
        W = 0.5 * (1.0 - np.sqrt(1.0 - (1.0 / rr) ** 2))
        Wv = np.trapezoid(W * rr ** 2, rr) / np.trapezoid(rr ** 2, rr)
        target = Wv * kps * m.tstar ** 4
        lo, hi = m.Tint[1], m.tstar
        for _ in range(80):
            mid = math.sqrt(lo * hi)
            if _planck_mean(m, mid) * mid ** 4 > target:
                hi = mid
            else:
                lo = mid
        want = math.sqrt(lo * hi)
        for it in range(m.ntg - 1):
            if m.species[0, it, ir] == 0:
                continue
            got = T[0, it, ir]
            assert abs(got / want - 1.0) < 0.02, (ir, it, got * 11605, want * 11605)
            checked += 1
    assert checked >= 8


# ---------------------------------------------------------------------------
# lucydustfreq between table nodes
# ---------------------------------------------------------------------------
def test_lucydustfreq_off_node_inverse_cdf(oracle_lib):
    m = M.small()
    rng = np.random.default_rng(51)
    for frac in (0.25, 0.5, 0.8):
        i = int(rng.integers(int(0.3 * m.nT), int(0.8 * m.nT)))
        T0 = math.exp(m.lTint[i] + frac * (m.lTint[i + 1] - m.lTint[i]))
        # CDF of kappa_abs B_nu(T0) in x = nu/T0 on [nuint[0], nuint[-1]] (the table's range)
        xs = np.concatenate([np.linspace(1e-9, m.nuint[0], 200), np.linspace(m.nuint[0], m.nuint[-1], 20001)[1:]])
        dens = _kabs(m, xs * T0) * xs ** 3 / np.expm1(xs)
        cdf = np.concatenate([[0.0], np.cumsum(0.5 * (dens[1:] + dens[:-1]) * np.diff(xs))])
        cdf /= cdf[-1]
        xi = np.sort(rng.uniform(0.0, 1.0, 4000))
        x_true = np.interp(xi, cdf, xs)
        x_got = np.array([oracle_lib.lucydustfreq(m, T0, 0, float(v)) / T0 for v in xi])
        inner = (xi > 0.01) & (xi < 0.99)
        rel = np.abs(x_got[inner] / x_true[inner] - 1.0)
        assert np.max(rel) < 0.02, (T0 * 11605, np.max(rel))
        assert np.median(rel) < 0.003


# ---------------------------------------------------------------------------
# diskflux
# ---------------------------------------------------------------------------
def test_diskflux_empty_medium_apparent_luminosity(oracle_lib):
    from src3_mpi_cfitsio_amd.engine import diskflux
    m = M.small("empty", nr=12, nt=8, nphi=1, nfreq=30)
    m.rho[...] = 0.0
    m.species[...] = 0
    n = 40_000
    o = oracle_lib.run_packets(m, n, seed=61, nthreads=4)
    f = diskflux(o, m, n)
    lsol = m.lstar / M.LSUN
    si = f["si"][0, 0, 0].sum(axis=-1)            # total component, per inclination bin
    ti = f["ti"][0, 0].sum(axis=-1)               # peeled, per direction
    # escaped: np/nmu packets per equal-solid-angle bin on average
    assert np.all(np.abs(si / lsol - 1.0) < 6 * math.sqrt(m.nmu / n))
    se = math.sqrt((16.0 / 6.0 - 1.0) / n)
    assert np.all(np.abs(ti / lsol - 1.0) < 6 * se)


def test_diskflux_error_columns_are_sample_statistics():
    from src3_mpi_cfitsio_amd.engine import SED_NAMES, PSED_NAMES, IMAGE_NAMES, diskflux
    m = M.small("df", nr=6, nt=4, nphi=1, nfreq=5, npeel=2)
    rng = np.random.default_rng(71)
    shp = (8, m.nap, m.nph, m.nmu, m.nfreq)
    pshp = (8, m.nap, m.npeel, m.nfreq)
    tal = {k: np.zeros(shp, np.float32) for k in SED_NAMES}
    tal.update({k: np.zeros(pshp, np.float32) for k in PSED_NAMES})
    tal.update({k: np.zeros((m.nband, m.nx, m.nx, m.npeel), np.float32) for k in IMAGE_NAMES})
    want_si2, want_ti2 = {}, {}
    for idx in [(0, 0, 0, 1, 2), (1, 0, 0, 0, 4), (0, 0, 0, 2, 0)]:
        w = rng.gamma(2.0, 0.5, size=int(rng.integers(5, 60)))
        for k, mult in (("si", 1.0), ("sq", 0.3), ("su", -0.2), ("sv", 0.05)):
            tal[k][idx] = np.float32((mult * w).sum())
            tal[k + "2"][idx] = np.float32(((mult * w) ** 2).sum())
        tal["nums"][idx] = len(w)
        tal["aveinc"][idx] = np.float32(0.4 * len(w))
        want_si2[idx] = math.sqrt(len(w)) * np.std(w, ddof=1) / w.sum() + 1.0 / len(w)
    for idx in [(0, 0, 1, 3), (2, 0, 0, 1)]:
        w = rng.gamma(3.0, 0.2, size=int(rng.integers(5, 60)))
        tal["ti"][idx] = np.float32(w.sum())
        tal["ti2"][idx] = np.float32((w ** 2).sum())
        tal["numt"][idx] = len(w)
        want_ti2[idx] = math.sqrt(len(w)) * np.std(w, ddof=1) / w.sum()
    npk = 12345
    f = diskflux(tal, m, npk)
    for idx, want in want_si2.items():
        assert abs(f["si2"][idx] - want) < 2e-4 * want + 1e-5
        assert abs(f["aveinc"][idx] - 0.4) < 1e-6
    for idx, want in want_ti2.items():
        assert abs(f["ti2"][idx] - want) < 2e-4 * want + 1e-5
    # single-packet and empty bins carry no error
    assert f["si2"][(0, 0, 0, 0, 0)] == 0.0 and f["ti2"][(0, 0, 0, 0)] == 0.0
    # normalization: the sums scale by nmu nph (L/Lsun) / np and (L/Lsun) / np
    idx = (0, 0, 0, 1, 2)
    lsol = m.lstar / M.LSUN
    assert abs(f["si"][idx] / (tal["si"][idx] * m.nmu * m.nph * lsol / npk) - 1.0) < 1e-6
    idx = (0, 0, 1, 3)
    assert abs(f["ti"][idx] / (tal["ti"][idx] * lsol / npk) - 1.0) < 1e-6


# ---------------------------------------------------------------------------
# tfinal's convergence statistic
# ---------------------------------------------------------------------------
def test_lucy_sigma_is_population_std_of_abs_change(oracle_lib):
    """tfinal_5_test1.f90:334-368 takes sqrt(<x^2> - <|x|>^2) of
    x = (tnew - told)/tnew over upper-hemisphere ip = 1 cells where both
    temperatures exceed 0.1 K and differ: numpy's population standard
    deviation of |x| over the same cells."""
    m = M.small("sig", nr=14, nt=10, nphi=3)
    rng = np.random.default_rng(81)
    floor_t = 0.1 / 11605.0
    told = rng.uniform(5, 900, m.cell_shape) / 11605.0
    tnew = told * rng.uniform(0.7, 1.4, m.cell_shape)
    told[0, :, :3] = floor_t                                    # floored before
    tnew[0, 1, 3:6] = floor_t                                   # floored after
    tnew[0, 2, 6:9] = told[0, 2, 6:9]                           # unchanged
    sig, n = oracle_lib.lucy_sigma(m, told, tnew)
    nh = (m.ntg - 1) // 2
    to, tn = told[0, :nh, :], tnew[0, :nh, :]
    keep = (tn != to) & (tn > floor_t) & (to > floor_t)
    x = (tn[keep] - to[keep]) / tn[keep]
    assert n == keep.sum() and 0 < n < keep.size
    assert abs(sig - np.std(np.abs(x))) < 1e-13
    # lower hemisphere and other azimuths do not enter
    told2 = told.copy()
    told2[1:] *= 2.0
    told2[0, nh:] *= 3.0
    assert oracle_lib.lucy_sigma(m, told2, tnew) == (sig, n)
