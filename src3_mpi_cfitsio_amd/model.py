"""Synthetic young-stellar-object models for the packet engine.

Builds, on the host, exactly the module state the reference's packet loop
reads (grid walls and face tables, per-cell density and species, dust and
Planck-emission tables, star, observer set-up), following the reference's
conventions (paths relative to src3_mpi_cfitsio/):

* walls: 1-based rarr(1..nrg) with rarr(1) = stellar surface, r2arr = rarr**2
  (gridset_5_model13new.f90:516); theta walls with the equator flagged by
  tan2arr = -1 and the poles by tan2arr = 0, ntg odd (:706-731); phi walls
  as half-planes aarr = sin(phi), barr = -cos(phi) (:766-772); cell volumes
  -(r2^3-r1^3)(cos t2-cos t1)(p2-p1)/3 (R* rsol)^3 (:1108-1152);
* dust tables as dustinit builds them (dust_5.f90:384-455): nudust = 12398/A
  in eV (descending), kappa normalized by its 5500 A value kappav;
* emission tables as setemiss builds them (dust_5.f90:507-699): nT = 500
  log-spaced temperatures 0.1 K..2e5 K (in K/11605), nnu = 513 linear
  x = h nu/kT in [1/32, 32], kapint = normalized cumulative integral of
  kappa_abs(nu) B_x(x) with kappa linear across each x interval
  (kappaB, dust_5.f90:784-796), kappap = the unnormalized total;
* SED bins (setup.f90:36-41, initarr_2_new.f90:21-25) and peel-off
  directions (tts_mod_5.f90, gridset check :695-701).

The densities themselves (flared disk, Ulrich infall envelope, log-normal
fractal clumping) are this project's synthetic workloads: the reference
builds its models from hydrodynamic tables that are not part of the path.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Any, Dict, List, Optional

import numpy as np

RSUN = 6.9598e10          # constants.f90
MSUN = 1.989e33
AU = 1.49598e13
G = 6.67428e-8
SIGT = 5.670400e-5
LSUN = 3.839e33
YEAR = 3.15576e7
TK = 11605.0              # K per eV (the reference's temperature unit)


@dataclasses.dataclass
class Model:
    name: str
    # grid
    nrg: int
    ntg: int
    npg: int
    rarr: np.ndarray
    r2arr: np.ndarray
    thetarr: np.ndarray
    costarr: np.ndarray
    tan2arr: np.ndarray
    phiarr: np.ndarray
    aarr: np.ndarray
    barr: np.ndarray
    vcell: np.ndarray            # (np-1, nt-1, nr-1) C order == reference cell order
    # medium
    rho: np.ndarray              # same shape, g/cm^3
    species: np.ndarray          # uint8, 0 = empty, 1 disk, 3 envelope
    species_opac: np.ndarray     # int32[11]
    # dust
    nopac: int
    nlam: int
    nudust: np.ndarray           # (nopac, nlam) eV, descending
    kappad: np.ndarray
    adust: np.ndarray
    gdust: np.ndarray
    pldust: np.ndarray
    kappav: np.ndarray           # (nopac,) cm^2/g
    nT: int
    nnu: int
    lTint: np.ndarray
    Tint: np.ndarray
    nuint: np.ndarray
    kapint: np.ndarray           # (nopac, nnu, nT)
    kappap: np.ndarray           # (nopac, nT)
    # source
    rstar: float                 # R_sun
    tstar: float                 # K/11605
    rmin: float
    # observer
    ipeel: int
    npeel: int
    thete: np.ndarray
    phie: np.ndarray
    nx: int
    rmaxi: float
    nband: int
    nwmax: int
    band_nw: np.ndarray
    band_freq: np.ndarray        # (nband, nwmax) Hz
    band_phi: np.ndarray
    nfreq: int
    nmu: int
    nph: int
    nap: int
    aperture2: np.ndarray
    u_mu: np.ndarray
    numin: float
    lnurat: float
    # run set-up
    npackets: int = 0
    niter: int = 1
    tdust0: Optional[np.ndarray] = None
    tables: Optional[tuple] = dataclasses.field(default=None, repr=False, compare=False)
    # external illumination (initpout.f90, rep_isrf.f90): the fraction of
    # the packets that enter from outside, and the ISRF sampling tables
    isrf_frac: float = 0.0
    isrf: Optional[Dict[str, np.ndarray]] = dataclasses.field(default=None, repr=False, compare=False)
    # stellar hot spots (initpspot.f90, spotset.f90): nspot, fspot,
    # spotlat/spotlon (deg), tshock (K/11605), accfrac2 (hot-spot photons)
    spots: Optional[Dict[str, float]] = None
    # disk-accretion photons (initpacc_4.f90): accfrac (the fraction of the
    # packets), ldint and hdisk (grid_mod's Ldint and hdisk, nrg - 1 each)
    accretion: Optional[Dict[str, Any]] = dataclasses.field(default=None, repr=False, compare=False)
    # Rosseland means of each opacity table (nopac, nT), as setemiss leaves kappar
    kappar: Optional[np.ndarray] = dataclasses.field(default=None, repr=False, compare=False)
    # the diffusion approximation (setdiffus_5.f90): kappar, hdisk (nrg - 1,
    # R*), rmaxd (R*), nr0
    diffusion: Optional[Dict[str, Any]] = dataclasses.field(default=None, repr=False, compare=False)

    @property
    def ncell(self) -> int:
        return (self.nrg - 1) * (self.ntg - 1) * (self.npg - 1)

    @property
    def cell_shape(self):
        return (self.npg - 1, self.ntg - 1, self.nrg - 1)

    @property
    def sini_mu(self) -> np.ndarray:
        return np.sqrt(1.0 - self.u_mu ** 2)

    @property
    def lstar(self) -> float:
        """4 pi R^2 sigma T^4 (newdisktherm_5_fuv.f90:195), erg/s."""
        return 4.0 * math.pi * SIGT * (self.tstar * TK) ** 4 * (self.rstar * RSUN) ** 2

    def tdust_initial(self) -> np.ndarray:
        if self.tdust0 is not None:
            return self.tdust0
        return np.full(self.cell_shape, 0.1 / TK)

    def with_packets(self, n: int) -> "Model":
        return dataclasses.replace(self, npackets=int(n))


# ---------------------------------------------------------------------------
# walls
# ---------------------------------------------------------------------------
def radial_walls(ncell_r: int, rin_au: float, rout_au: float, rstar_rsun: float) -> np.ndarray:
    """Wall 1 = stellar surface, wall 2 = Rin, log-spaced walls to Rout (R*)."""
    autors = AU / (rstar_rsun * RSUN)
    r = np.empty(ncell_r + 1)
    r[0] = 1.0
    r[1:] = np.logspace(math.log10(rin_au * autors), math.log10(rout_au * autors), ncell_r)
    return r


def theta_walls(ncell_t: int, midplane_refine: float = 0.9) -> np.ndarray:
    """ncell_t even; walls 0..pi with an equator wall, finer toward the midplane."""
    if ncell_t % 2:
        raise ValueError("the theta cell count must be even (ntg odd, equator wall)")
    nh = ncell_t // 2
    s = np.linspace(0.0, 1.0, nh + 1)
    w = midplane_refine
    upper = 0.5 * math.pi * ((1.0 - w) * s + w * (1.0 - (1.0 - s) ** 3))
    th = np.concatenate([upper, math.pi - upper[::-1][1:]])
    th[0], th[nh], th[-1] = 0.0, 0.5 * math.pi, math.pi
    return th


def theta_tables(thetarr: np.ndarray):
    """costarr / tan2arr with the reference's pole and equator flags."""
    ntg = len(thetarr)
    nttmp = (ntg + 1) // 2
    cost = np.cos(thetarr)
    tan2 = np.tan(thetarr) ** 2
    cost[0], tan2[0] = 1.0, 0.0
    cost[-1], tan2[-1] = -1.0, 0.0
    if ntg > 2:
        cost[nttmp - 1], tan2[nttmp - 1] = 0.0, -1.0
    return cost, tan2


def phi_walls(ncell_p: int):
    dp = 2.0 * math.pi / ncell_p
    phi = dp * np.arange(ncell_p + 1, dtype=np.float64)
    return phi, np.sin(phi), -np.cos(phi)


def cell_volumes(rarr, thetarr, phiarr, rstar_rsun):
    dr3 = rarr[1:] ** 3 - rarr[:-1] ** 3
    dcost = np.cos(thetarr[1:]) - np.cos(thetarr[:-1])
    dp = phiarr[1:] - phiarr[:-1]
    rmincgs = rstar_rsun * RSUN
    return (-(dr3[None, None, :] * dcost[None, :, None] * dp[:, None, None]) / 3.0 * rmincgs ** 3)


# ---------------------------------------------------------------------------
# densities
# ---------------------------------------------------------------------------
def zerod(xmu, z):
    """zerod.f90:3-69 elementwise: the root in (0,1] of
    xmu0^3 + (z-1) xmu0 - z xmu = 0 by the trigonometric form (three real
    roots; the last positive one) or Cardano's (one).  Returns (xmu0, flag):
    flag 1 where the reference reports two positive roots, 2 where none is
    positive; xmu0 is NaN there (the reference leaves it unset)."""
    xmu = np.asarray(xmu, dtype=np.float64)
    z = np.asarray(z, dtype=np.float64)
    a1 = 0.0
    a2 = z - 1.0
    a3 = -z * xmu
    q = (a1 * a1 - 3.0 * a2) / 9.0
    r = (2.0 * (a1 * a1 * a1) - 9.0 * a1 * a2 + 27.0 * a3) / 54.0
    q3 = q * q * q
    three = (q3 - r * r) >= 0.0
    with np.errstate(invalid="ignore", divide="ignore"):
        qh = np.sqrt(q)
        theta = np.arccos(r / np.sqrt(q3))
        x1 = -2.0 * qh * np.cos(theta / 3.0) - a1 / 3.0
        x2 = -2.0 * qh * np.cos((theta + 2.0 * math.pi) / 3.0) - a1 / 3.0
        x3 = -2.0 * qh * np.cos((theta + 4.0 * math.pi) / 3.0) - a1 / 3.0
        factor = np.sqrt(r * r - q3) + np.abs(r)
        factor3 = factor ** (1.0 / 3.0)
        xc = -1.0 * np.copysign(1.0, r) * (factor3 + q / factor3) - a1 / 3.0
    p1, p2, p3 = x1 > 0.0, x2 > 0.0, x3 > 0.0
    two = (p1 & p2) | (p1 & p3) | (p2 & p3)
    none = ~(p1 | p2 | p3)
    xt = np.where(p3, x3, np.where(p2, x2, x1))
    flag = np.where(three & two, 1, np.where(three & none, 2, 0)).astype(np.int32)
    xmu0 = np.where(three, xt, xc)
    return np.where(flag == 0, xmu0, np.nan), flag


def ulrich_mu0(mu: np.ndarray, rx: np.ndarray) -> np.ndarray:
    """Streamline origin mu0 of the Ulrich infall at |cos theta| = mu and
    r/Rc = rx, as zerod.f90 solves it."""
    return zerod(np.abs(mu), rx)[0]


def densenv(rad, costin, rc, rhoe0, rmine, rmax, rhoamb=0.0):
    """densenv.f90:30-45, 69-73, 111-113 elementwise (no holes, gaps or
    bubbles): Ulrich (1976) infall density at radius rad (R*) and polar
    cosine costin, centrifugal radius rc (R*), density scale rhoe0, zero
    inside rmine and beyond rmax, floored at rhoamb.  Returns (dens, flag)
    with zerod's flag per point."""
    rad = np.asarray(rad, dtype=np.float64)
    cosa = np.abs(costin)
    zp = rad * cosa
    xmu = zp / rad
    rx = rad / rc
    xmu0, flag = zerod(xmu, rx)
    with np.errstate(invalid="ignore", divide="ignore"):
        factor = 1.0 / ((xmu / xmu0) + 2.0 * (xmu0 * xmu0) / rx)
        dens = rhoe0 * rx ** -1.5 * factor / np.sqrt(1.0 + xmu / xmu0)
    out = rad > rmine
    dens = np.where(out, dens, 0.0)
    dens = np.where(out & (dens < rhoamb), rhoamb, dens)
    dens = np.where((rad >= rmax) | (rad <= rmine), 0.0, dens)
    return dens, np.where(out, flag, 0)


def fractal_field(shape, sigma_ln: float, dimension: float, seed: int) -> np.ndarray:
    """Log-normal multiplicative clumping exp(sigma g - sigma^2/2), where g is
    a unit-variance Gaussian random field with power spectrum k^-(2(3-D)+3)
    (fractional Brownian field of isosurface dimension D) on the cell index grid."""
    rng = np.random.default_rng(seed)
    beta = 2.0 * (3.0 - dimension) + 3.0
    kz = np.fft.fftfreq(shape[0])[:, None, None]
    ky = np.fft.fftfreq(shape[1])[None, :, None]
    kx = np.fft.rfftfreq(shape[2])[None, None, :]
    k = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
    k[0, 0, 0] = 1.0
    amp = k ** (-beta / 2.0)
    amp[0, 0, 0] = 0.0
    noise = rng.standard_normal(shape)
    g = np.fft.irfftn(np.fft.rfftn(noise) * amp, s=shape, axes=(0, 1, 2))
    g = (g - g.mean()) / g.std()
    return np.exp(sigma_ln * g - 0.5 * sigma_ln ** 2)


def ulrich_rhoe0(mdot_msun_yr: float, mstar_msun: float, rc_au: float) -> float:
    """Envelope density scale Mdot / (4 pi sqrt(G M Rc^3)), g/cm^3."""
    return (mdot_msun_yr * MSUN / YEAR) / (4.0 * math.pi * math.sqrt(G * mstar_msun * MSUN * (rc_au * AU) ** 3))


def density_params(rstar_rsun, mdisk_msun, rin_au, rout_au, h0_au=0.03, flare=1.25, alpha=2.25,
                   mdot_msun_yr=1e-6, rc_au=100.0, mstar_msun=0.5, rhoamb=0.0) -> np.ndarray:
    """The 12 per-model parameters of the GPU cell builder
    (mcrt_build_models), from the arguments yso_density takes."""
    autors = AU / (rstar_rsun * RSUN)
    return np.array([rstar_rsun * RSUN, autors, rin_au, rout_au, h0_au, flare, alpha, rc_au * autors,
                     ulrich_rhoe0(mdot_msun_yr, mstar_msun, rc_au), rin_au * autors, mdisk_msun * MSUN, rhoamb])


def yso_density(rarr, thetarr, phiarr, vcell, rstar_rsun, mdisk_msun, rin_au, rout_au,
                h0_au=0.03, flare=1.25, alpha=2.25, mdot_msun_yr=1e-6, rc_au=100.0,
                mstar_msun=0.5, clump_sigma=0.0, clump_dim=2.6, clump_seed=0):
    """Flared disk + Ulrich envelope at cell centres; one species per cell
    (1 = disk where the disk dominates, 3 = envelope), as the reference
    zeroes envelope species inside the disk (gridset_5_model13new.f90:1233-1239)."""
    autors = AU / (rstar_rsun * RSUN)
    rc = 0.5 * (rarr[1:] + rarr[:-1])[None, None, :]
    tc = 0.5 * (thetarr[1:] + thetarr[:-1])[None, :, None]
    r_au = rc / autors
    varpi = r_au * np.sin(tc)
    z = r_au * np.cos(tc)
    h = h0_au * np.maximum(varpi, 1e-30) ** flare
    inside = (r_au >= rin_au) & (varpi >= rin_au) & (r_au <= rout_au)
    rho_d = np.where(inside, np.maximum(varpi, 1e-30) ** (-alpha) * np.exp(-0.5 * (z / h) ** 2), 0.0)
    rho_d = np.broadcast_to(rho_d, vcell.shape).copy()
    mass = float(np.sum(rho_d * vcell))
    if mass > 0:
        rho_d *= mdisk_msun * MSUN / mass
    # Ulrich (1976) infall as densenv evaluates it at gridset's cell centres
    # (gridset_5_model13new.f90:1107,1123-1125): radii in R*
    rho_e, flag = densenv(rc, np.cos(tc), rc_au * autors, ulrich_rhoe0(mdot_msun_yr, mstar_msun, rc_au),
                          rin_au * autors, float(rarr[-1]))
    if np.any(flag):
        raise ValueError("zerod found no single streamline root at a cell centre")
    rho_e = np.broadcast_to(rho_e, vcell.shape).copy()
    if clump_sigma > 0:
        f = fractal_field(vcell.shape, clump_sigma, clump_dim, clump_seed)
        rho_d *= f
        rho_e *= f
    species = np.where(rho_d + rho_e > 0, np.where(rho_d >= rho_e, 1, 3), 0).astype(np.uint8)
    return rho_d + rho_e, species


# ---------------------------------------------------------------------------
# dust and emission tables
# ---------------------------------------------------------------------------
def synthetic_dust(nlam=100, lam_min=0.1, lam_max=1000.0, kappa0=200.0, slope=1.5,
                   albedo=(0.5, 0.01), g=(0.6, 0.0), pl=0.0):
    """kappa = kappa0 (lambda/um)^-slope; albedo and g linear in log lambda."""
    lam = np.logspace(math.log10(lam_min), math.log10(lam_max), nlam)     # um, ascending
    s = np.linspace(0.0, 1.0, nlam)
    kap = kappa0 * lam ** (-slope)
    alb = albedo[0] + (albedo[1] - albedo[0]) * s
    gg = g[0] + (g[1] - g[0]) * s
    pp = np.full(nlam, pl)
    lamA = 10000.0 * lam
    nu = 12398.0 / lamA                                                    # eV, descending
    iv = int(np.nonzero(lam <= 0.55)[0][-1])                               # dust_5.f90:426,459-460
    f = (5500.0 - lamA[iv]) / (lamA[iv + 1] - lamA[iv])
    kappav1 = (1.0 - f) * kap[iv] + f * kap[iv + 1]
    return dict(nudust=nu, kappad=kap / kappav1, adust=alb, gdust=gg, pldust=pp, kappav=kappav1)


_AS_EXPM1 = (0.9999999995, 0.4999999206, 0.1666653019, 0.0416573475,
             0.0083013598, 0.0013298820, 0.0001413161)


def expm1_as(x):
    """exp(x)-1 as mathsubs.f90:79-113 computes it: the Abramowitz & Stegun
    4.2.45 polynomial for |x| <= 0.6931472, exp(x)-1 beyond."""
    x = np.asarray(x, dtype=np.float64)
    s = np.zeros_like(x)
    for c in _AS_EXPM1[::-1]:
        s = x * (c + s)
    return np.where(np.abs(x) > 0.6931472, np.exp(x) - 1.0, s)


def planck_bx(x):
    """Bnu of mathsubs.f90:4-22: (15/pi^4) x^3 / expm1(x), 0 at x = 0."""
    x = np.asarray(x, dtype=np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where(x == 0.0, 0.0, 0.1539897 * x * x * x / expm1_as(x))


def planck_dbx(x):
    """dBnu of mathsubs.f90:29-47: (15/pi^4) x^4 e^x / (4 (e^x - 1)^2), the
    temperature derivative of planck_bx normalized to unit integral."""
    x = np.asarray(x, dtype=np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where(x == 0.0, 0.0, -0.1539897 * x ** 4 / (4.0 * expm1_as(x) * expm1_as(-x)))


def _opacset_vec(nu, nudust, tab):
    """dust_5.f90:21-55 linear interpolation in nu, clamped at the table ends."""
    return np.interp(nu, nudust[::-1], tab[::-1])


def emission_tables(dust, nT=500, nnu=513, Tmin=0.1, Tmax=200000.0, nodes=8):
    """setemiss (dust_5.f90:507-699) by Gauss-Legendre quadrature per x interval."""
    lTint = math.log(Tmin / TK) + np.arange(nT) * math.log(Tmax / Tmin) / (nT - 1)
    Tint = np.exp(lTint)
    numax, numin = 32.0, 32.0 / 1024.0
    nuint = numin + np.arange(nnu) * (numax - numin) / (nnu - 1)
    xg, wg = np.polynomial.legendre.leggauss(nodes)
    fg = 0.5 * (xg + 1.0)
    wg = 0.5 * wg
    x0 = np.concatenate([[0.0], nuint[:-1]])
    x1 = nuint
    dx = x1 - x0
    kabs = lambda nu: ((1.0 - _opacset_vec(nu, dust["nudust"], dust["adust"]))
                       * _opacset_vec(nu, dust["nudust"], dust["kappad"]))
    kap0 = kabs(x0[None, :] * Tint[:, None])                               # (nT, nnu)
    kap1 = kabs(x1[None, :] * Tint[:, None])
    xs = x0[:, None] + fg[None, :] * dx[:, None]                           # (nnu, nodes)
    B = planck_bx(xs)                                                      # mathsubs.f90:4-22
    kf = (1.0 - fg)[None, None, :] * kap0[:, :, None] + fg[None, None, :] * kap1[:, :, None]
    seg = dx[None, :] * np.sum(kf * B[None, :, :] * wg[None, None, :], axis=2)
    cum = np.cumsum(seg, axis=1)
    kappap = cum[:, -1].copy()
    kapint = cum / kappap[:, None]
    kapint[:, -1] = 1.0
    # kappar: 1 / sum over the intervals of (dB/dT)/kappa_ext (chidB,
    # dust_5.f90:814-826), the Rosseland mean of kappa/kappav
    kext = lambda nu: _opacset_vec(nu, dust["nudust"], dust["kappad"])    # noqa: E731
    e0, e1 = kext(x0[None, :] * Tint[:, None]), kext(x1[None, :] * Tint[:, None])
    ef = (1.0 - fg)[None, None, :] * e0[:, :, None] + fg[None, None, :] * e1[:, :, None]
    dB = planck_dbx(xs)
    kappar = 1.0 / np.sum(dx[None, :] * np.sum(dB[None, :, :] / ef * wg[None, None, :], axis=2), axis=1)
    return dict(nT=nT, nnu=nnu, lTint=lTint, Tint=Tint, nuint=nuint,
                kapint=np.ascontiguousarray(kapint.T), kappap=kappap, kappar=kappar)


# ---------------------------------------------------------------------------
# external illumination
# ---------------------------------------------------------------------------
# Black (1987) interstellar radiation field as CLOUDY tabulates it
# (isrf_cdf.f90:35-41): log10 nu [Hz] and log10 nu F_nu
_ISRF_LNU = (6.00, 10.72, 11.00, 11.23, 11.47, 11.55, 11.85, 12.26, 12.54, 12.71, 13.10, 13.64, 14.14, 14.38,
             14.63, 14.93, 15.08, 15.36, 15.43, 16.25, 17.09, 18.00, 23.00)
_ISRF_NUFNU = (-16.708, -2.96, -2.47, -2.09, -2.11, -2.34, -3.66, -2.72, -2.45, -2.57, -3.85, -3.34, -2.30, -1.79,
               -1.79, -2.34, -2.72, -2.55, -2.62, -5.68, -6.45, -6.30, -11.3)


def _locate2(xx, n, x):
    """nr_2.f:181-206 (1-based result, xx ascending or descending, first n used)."""
    jl, ju = 0, n + 1
    asc = xx[n - 1] > xx[0]
    while ju - jl > 1:
        jm = (ju + jl) // 2
        if asc == (x > xx[jm - 1]):
            jl = jm
        else:
            ju = jm
    if jl == 0 and x == xx[0]:
        jl = 1
    return jl


def isrf_tables(nisrf: int = 200) -> Dict[str, object]:
    """The ISRF sampling tables of isrf_cdf.f90 without extinction (isrf_Av
    = 0; the extinction law file the reference reads is not part of it):
    the spectrum log-interpolated onto nisrf log-spaced wavelengths in
    0.05..5000 um, a trapezoid CDF in frequency, then cdf (normalized),
    logcdf = log10 cdf (logcdf(1) = -99) and logfreq = log10 of h nu in eV
    (logfreq(1) = 0, as the reference leaves it).  isrf_int is the
    unnormalized integral (erg/s/cm^2/sr)."""
    c = 2.99792458e14
    freqin = 10.0 ** np.array(_ISRF_LNU)
    pdfin = 10.0 ** np.array(_ISRF_NUFNU) / freqin
    logfreqin, logpdfin = np.log10(freqin), np.log10(pdfin)
    nin = len(freqin)
    lmin = c / (freqin[nin - 1] * 0.99999999)
    lmax = c / (freqin[0] * 1.0000001)
    if lmin < 0.05:
        lmin = 0.05
    if lmax > 5000.0:
        lmax = 5000.0
    freq = np.zeros(nisrf)
    for i in range(1, nisrf + 1):
        lam = lmin * (lmax / lmin) ** (float(i) / float(nisrf))
        freq[nisrf - i] = c / lam
    pdf = np.zeros(nisrf)
    for i in range(nisrf):
        j = _locate2(freqin, nin, freq[i])
        logpdf = logpdfin[j - 1] + (math.log10(freq[i]) - logfreqin[j - 1]) / (logfreqin[j] - logfreqin[j - 1]) * \
            (logpdfin[j] - logpdfin[j - 1])
        pdf[i] = 10.0 ** logpdf * 1.0
    cdf = np.zeros(nisrf)
    for i in range(1, nisrf):
        cdf[i] = cdf[i - 1] + 0.5 * (pdf[i] + pdf[i - 1]) * (freq[i] - freq[i - 1])
    isrf_int = float(cdf[nisrf - 1])
    logcdf = np.zeros(nisrf)
    logfreq = np.zeros(nisrf)
    logcdf[0] = -99.0
    for i in range(1, nisrf):
        cdf[i] = cdf[i] / cdf[nisrf - 1]
        logcdf[i] = math.log10(cdf[i])
        freq[i] = freq[i] * 1.2398 / c
        logfreq[i] = math.log10(freq[i])
    return dict(cdf=cdf, logcdf=logcdf, logfreq=logfreq, isrf_int=isrf_int, pdf=pdf)


def isrf_luminosity(isrf_int: float, rmax_rstar: float, rstar_rsun: float, isrf_scl: float = 1.0) -> float:
    """l_isrf = isrf_int 4 pi (rmax R*)^2 / 2 * isrf_scl (newtts.f90:155), erg/s."""
    return isrf_int * 4.0 * math.pi * (rmax_rstar * RSUN * rstar_rsun) ** 2 / 2.0 * isrf_scl


def isrf_split(npackets: int, isrf_frac: float) -> int:
    """Packets of a run that are stellar; the rest (the last ones) enter from
    outside.  newdisktherm_5_fuv.f90:158-187 with no accretion: npout =
    nint-style rounding of isrf_frac * np, ns = np - npout."""
    if isrf_frac <= 0.0:
        return npackets
    npout = int(isrf_frac * float(npackets))
    if isrf_frac * float(npackets) - npout > 0.5:
        npout += 1
    return npackets - npout


def with_spots(m: Model, nspot: int = 1, fspot: float = 0.01, spotlat: float = 60.0, spotlon: float = 0.0,
               tshock_k: float = 8000.0, accfrac2: float = 0.1) -> Model:
    """The model with nspot accretion hot spots covering fspot of the star
    at tshock_k, and accfrac2 of every run's packets emitted from them."""
    return dataclasses.replace(m, spots=dict(nspot=int(nspot), fspot=float(fspot), spotlat=float(spotlat),
                                             spotlon=float(spotlon), tshock=float(tshock_k) / TK,
                                             accfrac2=float(accfrac2)))


def accretion_tables(m: Model, rin_au: float = 0.1, rout_au: float = 1000.0, h0_au: float = 0.03,
                     flare: float = 1.25) -> Dict[str, np.ndarray]:
    """Ldint and hdisk for the synthetic flared disk of build(): per radial
    cell j (mid radius r_j = ravearr(j), gridset_5_model13new.f90:525) the
    scale height h0 r^flare (AU, in R*) and the luminosity of a steady
    accretion disk, dL_j ~ r_j^-2 (1 - sqrt(R*/r_j)) dr_j over the cells
    with rin <= r_j <= rout.  initpacc_4.f90 places a photon in the cell j
    with Ldint(j) <= xi < Ldint(j+1), so Ldint(j) here is the sum over the
    cells before j (normalized to 1 in the last cell, whose own luminosity
    is left out): each cell emits its own share.  diskprofile_model11.f90
    sums through cell j instead, which moves every share one cell inward,
    the innermost into the dust-free cell inside the disk, where no photon
    can be placed.  The tables stand in for diskprofile's alpha-disk
    vertical structure (Bell-Lin opacities), which is part of the model
    set-up, not of the packet path: any Ldint (0 first, 1 last,
    non-decreasing) and hdisk >= 0 may be given instead."""
    rarr = np.asarray(m.rarr, dtype=np.float64)
    autors = AU / (m.rstar * RSUN)
    rave = 0.5 * (rarr[:-1] + rarr[1:])
    r_au = rave / autors
    hdisk = h0_au * r_au ** flare * autors
    dr = rarr[1:] - rarr[:-1]
    inside = (r_au >= rin_au) & (r_au <= rout_au)
    dl = np.where(inside, rave ** -2 * (1.0 - np.sqrt(1.0 / rave)) * dr, 0.0)
    before = np.concatenate(([0.0], np.cumsum(dl[:-1])))
    if not before[-1] > 0.0:
        raise ValueError("accretion_tables: no radial cell of the grid but the last lies in the disk")
    ldint = before / before[-1]
    return dict(ldint=ldint, hdisk=hdisk)


def with_accretion(m: Model, accfrac: float, tables: Optional[Dict[str, np.ndarray]] = None, **kw) -> Model:
    """The model with accfrac of every run's packets emitted by disk
    accretion, from tables (ldint, hdisk) or accretion_tables(m, **kw)."""
    t = tables if tables is not None else accretion_tables(m, **kw)
    return dataclasses.replace(m, accretion=dict(accfrac=float(accfrac), ldint=np.asarray(t["ldint"], np.float64),
                                                 hdisk=np.asarray(t["hdisk"], np.float64)))


def with_dusts(m: Model, dusts, r_split_au: float = 10.0) -> Model:
    """The model with several dust tables, as findopac assigns opacity
    tables to the species of densarr (grid_mod_5.f90, findopacid_5.f90):
    dusts is a list of 2 to 4 synthetic_dust tables of one length.  With two
    tables the disk (species 1) takes the first and the envelope (species 3)
    the second.  With three or four, cells beyond r_split_au move to
    species 2 (disk) and 4 (envelope), which take tables 2 and 3 (the
    fourth shared with the envelope when there are three).  Each table gets
    its own emission tables (kapint, kappap) on the model's nT x nnu grid."""
    nd = len(dusts)
    if not 2 <= nd <= 4:
        raise ValueError("with_dusts: 2 to 4 dust tables")
    nlam = len(dusts[0]["nudust"])
    if any(len(d["nudust"]) != nlam for d in dusts):
        raise ValueError("with_dusts: every table needs the same number of wavelengths")
    ems = [emission_tables(d, nT=m.nT, nnu=m.nnu) for d in dusts]
    species = m.species.copy()
    so = np.zeros(11, dtype=np.int32)
    so[1], so[3] = 0, 1
    if nd >= 3:
        autors = AU / (m.rstar * RSUN)
        rc = (0.5 * (m.rarr[1:] + m.rarr[:-1]) / autors)[None, None, :]
        far = np.broadcast_to(rc > r_split_au, species.shape)
        species = np.where(far & (species == 1), 2, np.where(far & (species == 3), 4, species)).astype(np.uint8)
        so[2], so[4] = 2, (3 if nd == 4 else 1)
    st = lambda k: np.stack([np.asarray(d[k], np.float64) for d in dusts])   # noqa: E731
    return dataclasses.replace(
        m, species=species, species_opac=so, nopac=nd, nlam=nlam, nudust=st("nudust"), kappad=st("kappad"),
        adust=st("adust"), gdust=st("gdust"), pldust=st("pldust"),
        kappav=np.array([d["kappav"] for d in dusts], np.float64),
        kapint=np.stack([e["kapint"] for e in ems]), kappap=np.stack([e["kappap"] for e in ems]),
        kappar=np.stack([e["kappar"] for e in ems]))


def with_diffusion(m: Model, rin_au: float = 0.1, rout_au: float = 1000.0, h0_au: float = 0.03,
                   flare: float = 1.25) -> Model:
    """The model with the diffusion approximation on (DIFFUSION = YES,
    setdiffus_5.f90): its Rosseland tables, the disk scale height of each
    radial cell (the flared disk of build(): h0 r^flare, at the cell's mid
    radius), the disk's outer radius and the first radial cell that may
    diffuse (1 when the disk reaches the inner wall, else 3)."""
    if m.kappar is None:
        raise ValueError("with_diffusion: the model has no Rosseland tables (kappar)")
    autors = AU / (m.rstar * RSUN)
    hdisk = accretion_tables(m, rin_au=rin_au, rout_au=rout_au, h0_au=h0_au, flare=flare)["hdisk"]
    nr0 = 1 if rin_au * autors <= float(m.rarr[0]) * (1.0 + 1e-12) else 3
    return dataclasses.replace(m, diffusion=dict(kappar=np.ascontiguousarray(m.kappar, np.float64), hdisk=hdisk,
                                                 rmaxd=float(min(rout_au * autors, m.rarr[-1])), nr0=nr0))


def with_isrf(m: Model, isrf_frac: float, nisrf: int = 200) -> Model:
    """The model with external illumination carrying isrf_frac of the packets."""
    return dataclasses.replace(m, isrf_frac=float(isrf_frac), isrf=isrf_tables(nisrf))


# ---------------------------------------------------------------------------
# configurations
# ---------------------------------------------------------------------------
def tophat_bands(lam_ranges):
    """Filter tables for top-hat bands (lambda ranges in um): three lines per
    band so that filt_nu's lookup (peeloff_3d_2_test.f90:250-263) returns
    weight 1 strictly inside (f_lo, f_hi)."""
    nb = len(lam_ranges)
    freq = np.zeros((max(nb, 1), 3))
    phi = np.zeros((max(nb, 1), 3))
    for b, (l0, l1) in enumerate(lam_ranges):
        f_lo, f_hi = 2.9979e14 / l1, 2.9979e14 / l0
        freq[b] = [f_lo, f_hi, f_hi]
        phi[b] = [1.0, 1.0, 1.0]
    return np.full(nb, 3, dtype=np.int32), freq, phi


def build(name: str, *, nr: int, nt: int, nphi: int, rin_au=0.1, rout_au=1000.0,
          rstar=2.0, tstar_k=4000.0, mdisk=0.01, mdot=1e-6, rc_au=100.0,
          clump_sigma=0.0, clump_dim=2.6, clump_seed=1001, pl=0.0,
          npeel=10, nx=101, bands=((0.01, 5000.0),), nfreq=100, nmu=10,
          npackets=1_000_000, niter=5, nlam=100, lam_max=1000.0, g=(0.6, 0.0), nT=500, nnu=513,
          thete_deg=None, phie_deg=0.0, flare=1.25, _tables=None, device=False, _cells=None) -> Model:
    """One model; device=True builds its cell arrays (volumes, densities,
    species) on the GPU (engine.build_models) instead of in numpy."""
    rarr = radial_walls(nr, rin_au, rout_au, rstar)
    thetarr = theta_walls(nt) if nt > 1 else np.array([0.0, math.pi])
    costarr, tan2arr = theta_tables(thetarr)
    phiarr, aarr, barr = phi_walls(nphi)
    if _cells is not None:
        rho, species, vcell = _cells
    elif device:
        from .engine import build_models
        clump = None
        if clump_sigma > 0:
            clump = fractal_field((len(phiarr) - 1, len(thetarr) - 1, len(rarr) - 1), clump_sigma, clump_dim,
                                  clump_seed)
        par = density_params(rstar, mdisk, rin_au, rout_au, flare=flare, mdot_msun_yr=mdot, rc_au=rc_au)
        cells, _ = build_models([(rarr, thetarr, phiarr)], par[None, :], None if clump is None else [clump])
        rho, species, vcell = cells[0]
    else:
        vcell = cell_volumes(rarr, thetarr, phiarr, rstar)
        rho, species = yso_density(rarr, thetarr, phiarr, vcell, rstar, mdisk, rin_au, rout_au,
                                   flare=flare, mdot_msun_yr=mdot, rc_au=rc_au, clump_sigma=clump_sigma,
                                   clump_dim=clump_dim, clump_seed=clump_seed)
    if _tables is None:          # (dust, emission tables) reused across models of a batch
        dust = synthetic_dust(nlam=nlam, lam_max=lam_max, g=g, pl=pl)
        em = emission_tables(dust, nT=nT, nnu=nnu)
    else:
        dust, em = _tables
    if thete_deg is None:
        thete_deg = (np.arange(npeel) + 0.5) * (90.0 / max(npeel, 1))
    thete = np.radians(np.asarray(thete_deg, dtype=np.float64))
    if np.any(np.min(np.abs(thete[:, None] - thetarr[None, :]), axis=1) == 0.0):
        raise ValueError("a peel-off angle equals a wall angle (gridset_5_model13new.f90:695-701)")
    band_nw, band_freq, band_phi = tophat_bands(bands)
    rmax = float(rarr[-1])
    numin, numax = 1.2398 / 5000.0, 1.2398 / 0.01                 # setup.f90:36-41
    species_opac = np.zeros(11, dtype=np.int32)
    return Model(
        name=name, nrg=len(rarr), ntg=len(thetarr), npg=len(phiarr),
        rarr=rarr, r2arr=rarr ** 2, thetarr=thetarr, costarr=costarr, tan2arr=tan2arr,
        phiarr=phiarr, aarr=aarr, barr=barr, vcell=vcell, rho=rho, species=species,
        species_opac=species_opac, nopac=1, nlam=nlam,
        nudust=dust["nudust"][None, :].copy(), kappad=dust["kappad"][None, :].copy(),
        adust=dust["adust"][None, :].copy(), gdust=dust["gdust"][None, :].copy(),
        pldust=dust["pldust"][None, :].copy(), kappav=np.array([dust["kappav"]]),
        nT=em["nT"], nnu=em["nnu"], lTint=em["lTint"], Tint=em["Tint"], nuint=em["nuint"],
        kapint=em["kapint"][None, :, :].copy(), kappap=em["kappap"][None, :].copy(),
        kappar=em["kappar"][None, :].copy(),
        rstar=rstar, tstar=tstar_k / TK, rmin=1.0,
        ipeel=1 if npeel > 0 else 0, npeel=npeel, thete=thete,
        phie=np.full(npeel, math.radians(phie_deg)), nx=nx, rmaxi=1.1 * rmax,
        nband=len(bands), nwmax=3, band_nw=band_nw, band_freq=band_freq, band_phi=band_phi,
        nfreq=nfreq, nmu=nmu, nph=1, nap=1, aperture2=np.array([(2.0 * rmax) ** 2]),
        u_mu=1.0 - (2.0 / nmu) * (np.arange(nmu) + 0.5), numin=numin,
        lnurat=math.log(numax / numin), npackets=npackets, niter=niter, tables=(dust, em))


CONFIGS: Dict[str, dict] = {
    # 2-D axisymmetric, FP64 plain method (CPU-runnable case)
    "c0": dict(nr=100, nt=90, nphi=1, npeel=10, nx=101, bands=((0.01, 5000.0),),
               npackets=1_000_000, niter=5),
    # 3-D polarized clumpy; FP32 transport with FP64 face roots and tallies.
    # Scattering matrix of the dustmat form (dustmat.f90:42-48) with g = 0.5
    # and peak linear polarization 0.5 at every wavelength
    "c1": dict(nr=200, nt=180, nphi=64, clump_sigma=0.5, clump_seed=1001, pl=0.5, g=(0.5, 0.5),
               npeel=10, nx=256, bands=((1.1, 1.4), (2.0, 2.4), (3.2, 4.0)),
               npackets=100_000_000, niter=5),
    # adjoint model; the observer sits 0.05 deg off 60 deg, which is a theta
    # wall of this grid (a peel angle on a wall is refused, gridset:695-701)
    "c3": dict(nr=150, nt=120, nphi=32, clump_sigma=0.3, clump_seed=1003, pl=0.5, g=(0.5, 0.5),
               npeel=1, thete_deg=[60.05], nx=128, bands=((2.0, 2.4), (3.2, 4.0)),
               npackets=10_000_000, niter=1),
    # large optically thick Lucy run, 250 wavelengths over 0.1-3000 um
    "c4": dict(nr=400, nt=360, nphi=128, mdisk=0.1, pl=0.5, nlam=250, lam_max=3000.0, npeel=10, nx=512,
               bands=((2.0, 2.4),), npackets=1_000_000_000, niter=10),
}


# ---------------------------------------------------------------------------
# batched model grid: many independent axisymmetric models
# ---------------------------------------------------------------------------
@dataclasses.dataclass
class Batch:
    """Independent axisymmetric models with one shared star, dust and SED
    binning (`shared`, a Model whose grid is the first model's), each with
    its own walls and medium.  Arrays are per model, concatenated for the
    device in the order of `models`."""
    shared: Model
    models: List[Model]
    params: List[dict]

    @property
    def nmodels(self) -> int:
        return len(self.models)

    def concat(self):
        ms = self.models
        cat = lambda f: np.ascontiguousarray(np.concatenate([np.asarray(f(x)).ravel() for x in ms]))
        return dict(
            nrg=np.array([x.nrg for x in ms], dtype=np.int32), ntg=np.array([x.ntg for x in ms], dtype=np.int32),
            rarr=cat(lambda x: x.rarr), r2arr=cat(lambda x: x.r2arr), thetarr=cat(lambda x: x.thetarr),
            costarr=cat(lambda x: x.costarr), tan2arr=cat(lambda x: x.tan2arr),
            vcell=cat(lambda x: x.vcell), rho=cat(lambda x: x.rho),
            species=np.ascontiguousarray(np.concatenate([x.species.ravel() for x in ms]).astype(np.uint8)),
            tdust=cat(lambda x: x.tdust_initial()))

    def cell_offsets(self) -> np.ndarray:
        n = np.array([x.ncell for x in self.models], dtype=np.int64)
        return np.concatenate([[0], np.cumsum(n)])


def batch_models(n: int = 512, seed: int = 1002, nr_range=(50, 400), nt_range=(40, 200), npeel_sed: int = 10,
                 nfreq: int = 100, npackets: int = 1_000_000, nlam: int = 100, nT: int = 500, nnu: int = 513,
                 device: bool = False) -> Batch:
    """Seeded grid of axisymmetric disk + envelope models with uneven grids:
    nr uniform in nr_range, n_theta uniform even in nt_range, Mdisk
    log-uniform 1e-4..1e-1 Msun, flaring exponent 1.0..1.3, Rin 0.1..1 AU,
    envelope infall rate log-uniform 1e-7..1e-5 Msun/yr.  Same star and dust
    as the 100x90 disk; SEDs binned at npeel_sed inclinations (cos theta
    bins of the escaping packets), no peel-off images.  device=True builds
    every model's cell arrays in one GPU launch."""
    rng = np.random.default_rng(seed)
    params = []
    for k in range(n):
        nr = int(rng.integers(nr_range[0], nr_range[1] + 1))
        nt = int(rng.integers(nt_range[0] // 2, nt_range[1] // 2 + 1)) * 2
        params.append(dict(nr=nr, nt=nt, mdisk=float(10 ** rng.uniform(-4.0, -1.0)),
                           flare=float(rng.uniform(1.0, 1.3)), rin_au=float(rng.uniform(0.1, 1.0)),
                           mdot=float(10 ** rng.uniform(-7.0, -5.0))))
    rstar, rout_au = 2.0, 1000.0                 # build()'s defaults, shared by every model of the grid
    cells = [None] * n
    if device:                                   # every model's cell arrays in one GPU launch
        from .engine import build_models
        walls = [(radial_walls(p["nr"], p["rin_au"], rout_au, rstar), theta_walls(p["nt"]), phi_walls(1)[0])
                 for p in params]
        par = np.array([density_params(rstar, p["mdisk"], p["rin_au"], rout_au, flare=p["flare"],
                                       mdot_msun_yr=p["mdot"]) for p in params])
        cells, _ = build_models(walls, par)
    tables = None
    models = []
    for k, p in enumerate(params):
        m = build(f"batch{k}", nr=p["nr"], nt=p["nt"], nphi=1, rin_au=p["rin_au"], rout_au=rout_au, rstar=rstar,
                  mdisk=p["mdisk"], mdot=p["mdot"], flare=p["flare"], npeel=0, nx=1, bands=(), nfreq=nfreq,
                  nmu=npeel_sed, npackets=npackets, niter=1, nlam=nlam, nT=nT, nnu=nnu, _tables=tables,
                  _cells=cells[k])
        tables = m.tables
        models.append(m)
    return Batch(shared=models[0], models=models, params=params)


def config(name: str, **over) -> Model:
    kw = dict(CONFIGS[name])
    kw.update(over)
    return build(name, **kw)


def small(name="small", nr=24, nt=16, nphi=1, **over) -> Model:
    """A few-thousand-cell model with the same physics, for parity tests."""
    kw = dict(nr=nr, nt=nt, nphi=nphi, npeel=4, nx=33, nfreq=20, nmu=4, npackets=2000, niter=1,
              nlam=40, nT=100, nnu=65)
    kw.update(over)
    return build(name, **kw)
