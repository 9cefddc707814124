"""ctypes wrapper of the CPU oracle (liboracle.so) and of the reference leaf
build (oracle/_ref/libref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libref.so")

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int32)
_u8p = C.POINTER(C.c_uint8)


class OrcModel(C.Structure):
    _fields_ = [
        ("nrg", C.c_int32), ("ntg", C.c_int32), ("npg", C.c_int32),
        ("rarr", _dp), ("r2arr", _dp), ("thetarr", _dp), ("costarr", _dp), ("tan2arr", _dp),
        ("phiarr", _dp), ("aarr", _dp), ("barr", _dp), ("vcell", _dp),
        ("rho", _dp), ("species", _u8p), ("species_opac", C.c_int32 * 11),
        ("nopac", C.c_int32), ("nlam", C.c_int32),
        ("nudust", _dp), ("kappad", _dp), ("adust", _dp), ("gdust", _dp), ("pldust", _dp), ("kappav", _dp),
        ("nT", C.c_int32), ("nnu", C.c_int32),
        ("lTint", _dp), ("Tint", _dp), ("nuint", _dp), ("kapint", _dp), ("kappap", _dp),
        ("rstar", C.c_double), ("tstar", C.c_double), ("rmin", C.c_double),
        ("ipeel", C.c_int32), ("npeel", C.c_int32),
        ("coste", _dp), ("sinte", _dp), ("cospe", _dp), ("sinpe", _dp), ("phie", _dp),
        ("nx", C.c_int32), ("rmaxi", C.c_double),
        ("nband", C.c_int32), ("nwmax", C.c_int32), ("band_nw", _ip), ("band_freq", _dp), ("band_phi", _dp),
        ("nfreq", C.c_int32), ("nmu", C.c_int32), ("nph", C.c_int32), ("nap", C.c_int32),
        ("aperture2", _dp), ("u_mu", _dp), ("sini_mu", _dp),
        ("numin", C.c_double), ("lnurat", C.c_double),
        ("iintmax", C.c_int32), ("taumax", C.c_double),
        ("nisrf", C.c_int32), ("isrf_cdf", _dp), ("isrf_logcdf", _dp), ("isrf_logfreq", _dp),
        ("isrf_from", C.c_int64),
        ("ispot", C.c_int32), ("nspot", C.c_int32), ("fspot", C.c_double), ("csprad", C.c_double),
        ("tshock", C.c_double), ("scspot1", C.c_double), ("ssspot1", C.c_double), ("cspot1", C.c_double),
        ("sspot1", C.c_double), ("scspot2", C.c_double), ("ssspot2", C.c_double), ("cspot2", C.c_double),
        ("sspot2", C.c_double), ("hot_to", C.c_int64),
        ("acc_to", C.c_int64), ("acc_ldint", C.POINTER(C.c_double)), ("acc_hdisk", C.POINTER(C.c_double)),
        ("diff_itmin", _ip), ("diff_delta", C.c_double), ("diff_count", C.POINTER(C.c_uint32)),
    ]


class OrcTallies(C.Structure):
    _fields_ = [("dtauabs", _dp), ("nabs", _ip), ("dbig", _ip), ("killed", _ip),
                ("sed", _fp), ("psed", _fp), ("image", _fp), ("counters", C.c_double * 16)]


class OrcWalls(C.Structure):
    _fields_ = [("tn1", C.c_double), ("tp1", C.c_double), ("tp2", C.c_double),
                ("in1", C.c_int32), ("ip1", C.c_int32), ("ip2", C.c_int32)]


def build_oracle():
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


def build_ref():
    """Compile the reference's MPI-free leaves (needs /root/reference)."""
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "ref")], check=True)


_orc = None
_ref = None


def oracle():
    global _orc
    if _orc is None:
        if not os.path.exists(ORACLE_SO):
            build_oracle()
        L = C.CDLL(ORACLE_SO)
        L.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.orc_packet_word.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32]
        L.orc_packet_word.restype = C.c_uint32
        L.orc_lecuyer_set_all_seeds.argtypes = [C.c_int32, C.c_int32]
        L.orc_lecuyer_set_generator.argtypes = [C.c_int32]
        L.orc_lecuyer_ran.restype = C.c_double
        L.orc_bbfreq_lecuyer.argtypes = [C.c_double]
        L.orc_bbfreq_lecuyer.restype = C.c_double
        W = C.POINTER(OrcWalls)
        L.orc_reset_t.argtypes = [W]
        L.orc_insert_t.argtypes = [W, C.c_double, C.c_int32]
        L.orc_find_next_wall.argtypes = [W, C.c_int, _dp, _ip]
        L.orc_find_next_wall.restype = C.c_int
        L.orc_radface.argtypes = [W] + [C.c_double] * 8
        L.orc_thetaface.argtypes = [W] + [C.c_double] * 8
        L.orc_phiface.argtypes = [W] + [C.c_double] * 10
        L.orc_locate.argtypes = [_dp, C.c_int32, C.c_double]
        L.orc_locate.restype = C.c_int32
        L.orc_findangle3.argtypes = [C.c_double] * 10 + [_dp] * 4
        L.orc_dustmat.argtypes = [C.c_double] * 4 + [_dp]
        L.orc_stokes_apply.argtypes = [C.c_double] * 4 + [_dp, _dp]
        L.orc_stokespeel.argtypes = [C.c_double, C.c_double, _dp] + [C.c_double] * 6
        L.orc_lucydustfreq.argtypes = [C.POINTER(OrcModel), C.c_double, C.c_int32, C.c_double]
        L.orc_lucydustfreq.restype = C.c_double
        L.orc_dusttemp.argtypes = [C.POINTER(OrcModel), C.c_double, C.c_double, C.c_int32, C.c_double]
        L.orc_dusttemp.restype = C.c_double
        L.orc_tfinal.argtypes = [C.POINTER(OrcModel), _dp, C.c_double, _dp]
        L.orc_trace_ray.argtypes = [C.POINTER(OrcModel), _dp, _dp, _ip, C.c_int, C.c_double, C.c_int32,
                                    _ip, _ip, _dp, _ip, C.c_uint64]
        L.orc_trace_ray.restype = C.c_int32
        L.orc_run_packets.argtypes = [C.POINTER(OrcModel), _dp, C.POINTER(OrcTallies), C.c_uint64,
                                      C.c_int64, C.c_uint64, C.c_int32, C.c_int32, _dp]
        L.orc_run_packets.restype = C.c_int32
        for f in (L.orc_expm1_as, ):
            f.argtypes = [C.c_double]
            f.restype = C.c_double
        for f in (L.orc_bnu, L.orc_dbnu):
            f.argtypes = [C.c_double, C.c_double]
            f.restype = C.c_double
        L.orc_qsimp_seg.argtypes = [C.c_int32] + [C.c_double] * 5 + [_dp]
        L.orc_qsimp_seg.restype = C.c_int32
        L.orc_setemiss.argtypes = [C.c_int32, C.c_int32, _ip, _dp, _dp, _dp, C.c_int32, _dp, C.c_int32, _dp, _ip,
                                   _dp, _dp, _dp]
        L.orc_setemiss.restype = C.c_int64
        L.orc_zerod.argtypes = [C.c_double, C.c_double, _dp]
        L.orc_zerod.restype = C.c_int32
        L.orc_densenv.argtypes = [C.c_double] * 8 + [_ip]
        L.orc_densenv.restype = C.c_double
        L.orc_cell_volume.argtypes = [C.c_double] * 7
        L.orc_cell_volume.restype = C.c_double
        _orc = L
    return _orc


def ref():
    """The reference's own compiled leaf routines, or None when unbuilt."""
    global _ref
    if _ref is None:
        if not os.path.exists(REF_SO):
            return None
        L = C.CDLL(REF_SO)
        L.ref_walls_state.argtypes = [_dp, _ip]
        L.ref_radface.argtypes = [_dp, _dp, C.c_double, C.c_double]
        L.ref_thetaface.argtypes = [_dp, _dp, C.c_double, C.c_double]
        L.ref_phiface.argtypes = [C.c_double] * 4 + [_dp, _dp]
        L.ref_find_next_wall.argtypes = [C.c_int, _dp, _ip]
        L.ref_findangle3.argtypes = [_dp, _dp]
        L.ref_check.argtypes = [C.c_double]
        L.ref_check.restype = C.c_double
        L.ref_locate.argtypes = [_dp, C.c_int32, C.c_double]
        L.ref_locate.restype = C.c_int32
        L.ref_user_set_all.argtypes = [C.c_int32, C.c_int32, C.c_int32]
        L.ref_ran.restype = C.c_double
        L.ref_bbfreq.argtypes = [C.c_double]
        L.ref_bbfreq.restype = C.c_double
        L.ref_bnu.argtypes = [C.c_double, C.c_double]
        L.ref_bnu.restype = C.c_double
        if hasattr(L, "ref_qsimp_seg"):
            L.ref_dbnu.argtypes = [C.c_double, C.c_double]
            L.ref_dbnu.restype = C.c_double
            L.ref_expm1.argtypes = [C.c_double]
            L.ref_expm1.restype = C.c_double
            L.ref_qsimp_seg.argtypes = [C.c_int] + [C.c_double] * 5 + [_dp, C.POINTER(C.c_int)]
        if hasattr(L, "ref_zerod"):
            L.ref_zerod.argtypes = [C.c_double, C.c_double, _dp, C.POINTER(C.c_int)]
        L.ref_open_log()
        _ref = L
    return _ref


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def spot_geometry(nspot, fspot, spotlat, spotlon):
    """gridset_5_model13new.f90:159-165 (spot radius in degrees from the
    filling factor) and spotset.f90:9-27 (cap centres and cos radius)."""
    import math
    pi = 3.1415926535897932384626433
    deg2rad, rad2deg = pi / 180.0, 180.0 / pi
    thspot = (math.acos(1.0 - 2.0 * fspot) if nspot == 1 else math.acos(1.0 - fspot)) * rad2deg
    th1, ph1 = (90.0 - spotlat) * deg2rad, spotlon * deg2rad
    th2, ph2 = pi - th1, ph1 + pi
    return dict(scspot1=math.sin(th1) * math.cos(ph1), ssspot1=math.sin(th1) * math.sin(ph1), cspot1=math.cos(th1),
                sspot1=math.sin(th1), scspot2=math.sin(th2) * math.cos(ph2), ssspot2=math.sin(th2) * math.sin(ph2),
                cspot2=math.cos(th2), sspot2=math.sin(th2), csprad=math.cos(thspot * deg2rad))


class OracleModel:
    """Holds the numpy arrays alive and exposes the C struct."""

    def __init__(self, m):
        f = lambda x: np.ascontiguousarray(x, dtype=np.float64)
        self.arrs = dict(
            rarr=f(m.rarr), r2arr=f(m.r2arr), thetarr=f(m.thetarr), costarr=f(m.costarr),
            tan2arr=f(m.tan2arr), phiarr=f(m.phiarr), aarr=f(m.aarr), barr=f(m.barr), vcell=f(m.vcell),
            rho=f(m.rho), species=np.ascontiguousarray(m.species, dtype=np.uint8),
            nudust=f(m.nudust), kappad=f(m.kappad), adust=f(m.adust), gdust=f(m.gdust),
            pldust=f(m.pldust), kappav=f(m.kappav), lTint=f(m.lTint), Tint=f(m.Tint),
            nuint=f(m.nuint), kapint=f(m.kapint), kappap=f(m.kappap),
            coste=f(np.cos(m.thete)), sinte=f(np.sin(m.thete)), cospe=f(np.cos(m.phie)),
            sinpe=f(np.sin(m.phie)), phie=f(m.phie),
            band_nw=np.ascontiguousarray(m.band_nw, dtype=np.int32), band_freq=f(m.band_freq),
            band_phi=f(m.band_phi), aperture2=f(m.aperture2), u_mu=f(m.u_mu), sini_mu=f(m.sini_mu))
        s = OrcModel()
        for k, v in self.arrs.items():
            t = C.c_uint8 if v.dtype == np.uint8 else (C.c_int32 if v.dtype == np.int32 else C.c_double)
            setattr(s, k, _p(v, t))
        s.nrg, s.ntg, s.npg = m.nrg, m.ntg, m.npg
        for i in range(11):
            s.species_opac[i] = int(m.species_opac[i])
        s.nopac, s.nlam, s.nT, s.nnu = m.nopac, m.nlam, m.nT, m.nnu
        s.rstar, s.tstar, s.rmin = m.rstar, m.tstar, m.rmin
        s.ipeel, s.npeel, s.nx, s.rmaxi = m.ipeel, m.npeel, m.nx, m.rmaxi
        s.nband, s.nwmax = m.nband, m.nwmax
        s.nfreq, s.nmu, s.nph, s.nap = m.nfreq, m.nmu, m.nph, m.nap
        s.numin, s.lnurat = m.numin, m.lnurat
        s.iintmax, s.taumax = 10000, 25.0
        s.isrf_from = -1
        if getattr(m, "isrf", None) is not None and m.isrf_frac > 0.0:
            for k in ("cdf", "logcdf", "logfreq"):
                self.arrs["isrf_" + k] = f(m.isrf[k])
                setattr(s, "isrf_" + k, _p(self.arrs["isrf_" + k], C.c_double))
            s.nisrf = len(self.arrs["isrf_cdf"])
        sp = getattr(m, "spots", None)
        if sp:
            g = spot_geometry(sp["nspot"], sp["fspot"], sp["spotlat"], sp["spotlon"])
            s.ispot, s.nspot, s.fspot, s.tshock = 1, sp["nspot"], sp["fspot"], sp["tshock"]
            for k, v in g.items():
                setattr(s, k, v)
        acc = getattr(m, "accretion", None)
        if acc and acc["accfrac"] > 0.0:
            for k in ("ldint", "hdisk"):
                self.arrs["acc_" + k] = f(acc[k])
                setattr(s, "acc_" + k, _p(self.arrs["acc_" + k], C.c_double))
        self.s = s
        self.m = m

    def set_run(self, npackets):
        """The split of a run of npackets (newdisktherm_5_fuv.f90:158-187):
        accretion photons first, then hot-spot photons, the external-field
        packets last."""
        m = self.m
        if getattr(m, "isrf", None) is not None and m.isrf_frac > 0.0:
            npout = int(m.isrf_frac * float(npackets))
            if m.isrf_frac * float(npackets) - npout > 0.5:
                npout += 1
            self.s.isrf_from = int(npackets) - npout
        else:
            self.s.isrf_from = -1
        # newdisktherm_5_fuv.f90:169-170, :242: iphot < npacc = int(accfrac np)
        # (1-based) are accretion photons
        acc = getattr(m, "accretion", None)
        npacc = int(acc["accfrac"] * float(npackets)) if acc and acc["accfrac"] > 0.0 else 0
        self.s.acc_to = max(npacc - 1, 0)
        sp = getattr(m, "spots", None)
        if sp:
            # :172-177, :296: then hot-spot photons while iphot < npacc + npaccstar
            npaccstar = int(sp["accfrac2"] * float(npackets))
            self.s.hot_to = max(npacc + npaccstar - 1, 0)
        else:
            self.s.hot_to = 0


def run_packets(m, npackets, seed, first_packet=0, tdust=None, nthreads=1, rng_mode=0, trace=False, diffusion=None):
    """Oracle packet loop; returns tallies named like Engine.tallies().
    diffusion: dict(itmin, count) from setdiffus (count, a uint32 array of
    cell_shape, is updated in place) to run with the diffusion approximation;
    its surface increment is 4 pi Tstar^4 / ns of this run."""
    L = oracle()
    om = OracleModel(m)
    if diffusion is not None:
        om.arrs["diff_itmin"] = np.ascontiguousarray(diffusion["itmin"], dtype=np.int32)
        om.s.diff_itmin = _p(om.arrs["diff_itmin"], C.c_int32)
        cnt = diffusion["count"]
        assert cnt.dtype == np.uint32 and cnt.flags["C_CONTIGUOUS"]
        om.s.diff_count = cnt.ctypes.data_as(C.POINTER(C.c_uint32))
        om.s.diff_delta = diffusion_delta(m, npackets)
    cs = m.cell_shape
    out = dict(dtauabs=np.zeros(cs), nabs=np.zeros(cs, np.int32), dbig=np.zeros(cs, np.int32),
               killed=np.zeros(cs, np.int32))
    sed = np.zeros((10, 8, m.nap, m.nph, m.nmu, m.nfreq), np.float32)
    psed = np.zeros((9, 8, m.nap, max(m.npeel, 1), m.nfreq), np.float32)
    img = np.zeros((8, max(m.nband, 1), m.nx, m.nx, max(m.npeel, 1)), np.float32)
    t = OrcTallies()
    t.dtauabs = _p(out["dtauabs"], C.c_double)
    t.nabs = _p(out["nabs"], C.c_int32)
    t.dbig = _p(out["dbig"], C.c_int32)
    t.killed = _p(out["killed"], C.c_int32)
    t.sed, t.psed, t.image = _p(sed, C.c_float), _p(psed, C.c_float), _p(img, C.c_float)
    td = np.ascontiguousarray(m.tdust_initial() if tdust is None else tdust, dtype=np.float64)
    tr = np.zeros((npackets, 8)) if trace else None
    om.set_run(npackets)
    rc = L.orc_run_packets(C.byref(om.s), _p(td, C.c_double), C.byref(t), int(first_packet), int(npackets),
                           C.c_uint64(int(seed)), int(rng_mode), int(nthreads),
                           _p(tr, C.c_double) if trace else None)
    if rc != 0:
        raise RuntimeError(f"orc_run_packets failed: {rc}")
    from src3_mpi_cfitsio_amd.engine import COUNTERS, IMAGE_NAMES, PSED_NAMES, SED_NAMES
    for i, n in enumerate(SED_NAMES):
        out[n] = sed[i]
    for i, n in enumerate(PSED_NAMES):
        out[n] = psed[i]
    for i, n in enumerate(IMAGE_NAMES):
        out[n] = img[i]
    out["counters"] = {k: t.counters[i] for i, k in enumerate(COUNTERS)}
    if trace:
        out["trace"] = tr
    return out


def diffusion_delta(m, npackets):
    """4 pi Tstar^4 / ns, ns = the run's stellar photons (np - npacc - npout)."""
    nout = 0
    if getattr(m, "isrf", None) is not None and m.isrf_frac > 0.0:
        nout = int(m.isrf_frac * float(npackets))
        if m.isrf_frac * float(npackets) - nout > 0.5:
            nout += 1
    acc = getattr(m, "accretion", None)
    nacc = int(acc["accfrac"] * float(npackets)) if acc and acc["accfrac"] > 0.0 else 0
    return 4.0 * np.pi * m.tstar ** 4 / float(npackets - nacc - nout)


def setdiffus(m, tdave, iteration):
    """setdiffus_5.f90 (orc_setdiffus) with the model's diffusion set-up:
    returns diffus (bool, cell_shape), diffdir (int8) and itmin (nrg - 1)."""
    L = oracle()
    om = OracleModel(m)
    d = m.diffusion
    kr, hd = np.ascontiguousarray(d["kappar"], np.float64), np.ascontiguousarray(d["hdisk"], np.float64)
    td = np.ascontiguousarray(tdave, np.float64)
    dif = np.zeros(m.cell_shape, np.uint8)
    ddir = np.zeros(m.cell_shape, np.int8)
    itmin = np.zeros(m.nrg - 1, np.int32)
    L.orc_setdiffus(C.byref(om.s), _p(kr, C.c_double), _p(hd, C.c_double), _p(td, C.c_double), C.c_double(d["rmaxd"]),
                    C.c_int32(int(d["nr0"])), C.c_int32(int(iteration)), _p(dif, C.c_uint8),
                    ddir.ctypes.data_as(C.POINTER(C.c_int8)),
                    _p(itmin, C.c_int32))
    return dif.astype(bool), ddir, itmin


def diff_columns(m, itmin, count, nabs, npackets, tdust):
    """orc_diff_columns: tdust with the diffusion columns' temperatures after a run."""
    L = oracle()
    om = OracleModel(m)
    om.arrs["diff_itmin"] = np.ascontiguousarray(itmin, dtype=np.int32)
    om.s.diff_itmin = _p(om.arrs["diff_itmin"], C.c_int32)
    kr = np.ascontiguousarray(m.diffusion["kappar"], np.float64)
    cnt = np.ascontiguousarray(count, np.uint32)
    na = np.ascontiguousarray(nabs, np.int32)
    out = np.array(tdust, dtype=np.float64, order="C", copy=True)
    L.orc_diff_columns(C.byref(om.s), _p(kr, C.c_double), cnt.ctypes.data_as(C.POINTER(C.c_uint32)),
                       _p(na, C.c_int32), C.c_double(diffusion_delta(m, npackets)), _p(out, C.c_double))
    return out


def tfinal2(m, dtauabs, nphot, tdust, itmin):
    """tfinal with the diffusion region itmin keeping its run temperatures."""
    L = oracle()
    om = OracleModel(m)
    om.arrs["diff_itmin"] = np.ascontiguousarray(itmin, dtype=np.int32)
    om.s.diff_itmin = _p(om.arrs["diff_itmin"], C.c_int32)
    d = np.ascontiguousarray(dtauabs, dtype=np.float64)
    td = np.ascontiguousarray(tdust, dtype=np.float64)
    t = np.zeros(m.cell_shape)
    L.orc_tfinal2(C.byref(om.s), _p(d, C.c_double), C.c_double(nphot), _p(td, C.c_double), _p(t, C.c_double))
    return t


def tfinal(m, dtauabs, nphot):
    L = oracle()
    om = OracleModel(m)
    d = np.ascontiguousarray(dtauabs, dtype=np.float64)
    out = np.zeros(m.cell_shape)
    L.orc_tfinal(C.byref(om.s), _p(d, C.c_double), float(nphot), _p(out, C.c_double))
    return out


def lucy_sigma(m, told, tnew):
    """tfinal's convergence statistic (orc_lucy_sigma): (sigma, ncells)."""
    L = oracle()
    om = OracleModel(m)
    a = np.ascontiguousarray(told, dtype=np.float64)
    b = np.ascontiguousarray(tnew, dtype=np.float64)
    out = np.zeros(3)
    L.orc_lucy_sigma.restype = C.c_double
    sig = L.orc_lucy_sigma(C.byref(om.s), _p(a, C.c_double), _p(b, C.c_double), _p(out, C.c_double))
    return float(sig), int(out[2])


def dusttemp(m, na, nphot, iop, mcell):
    om = OracleModel(m)
    return oracle().orc_dusttemp(C.byref(om.s), float(na), float(nphot), int(iop), float(mcell))


def lucydustfreq(m, T0, iop, xi):
    om = OracleModel(m)
    return oracle().orc_lucydustfreq(C.byref(om.s), float(T0), int(iop), float(xi))


def trace_ray(m, pos, u, ijk, on_wall, tau_stop, max_steps=4096, om=None, stream=0):
    L = oracle()
    om = om or OracleModel(m)
    p = np.array(pos, dtype=np.float64)
    uu = np.array(u, dtype=np.float64)
    ij = np.array(ijk, dtype=np.int32)
    faces = np.zeros(max_steps, np.int32)
    cells = np.zeros((max_steps, 3), np.int32)
    tau = C.c_double()
    st = C.c_int32()
    n = L.orc_trace_ray(C.byref(om.s), _p(p, C.c_double), _p(uu, C.c_double), _p(ij, C.c_int32), int(on_wall),
                        float(tau_stop), int(max_steps), _p(faces, C.c_int32), _p(cells, C.c_int32),
                        C.byref(tau), C.byref(st), C.c_uint64(int(stream)))
    return dict(n=n, faces=faces[:max(n, 0)], cells=cells[:max(n, 0)], pos=p, ijk=ij, tau=tau.value,
                status=st.value)


def setemiss(nudust, kappad, adust, Tint, nuint, smooth=None, nlambda=None):
    """setemiss (dust_5.f90:507-699) for tables of shape (nopac, nlam):
    returns kapint (nopac, nnu, nT), kappap (nopac, nT), kappar (nopac, nT)
    and the number of qsimp calls that did not converge."""
    L = oracle()
    nud, kap, alb = (np.ascontiguousarray(np.atleast_2d(a), dtype=np.float64) for a in (nudust, kappad, adust))
    nopac, nlam = nud.shape
    T = np.ascontiguousarray(Tint, dtype=np.float64)
    nu = np.ascontiguousarray(nuint, dtype=np.float64)
    nT, nnu = len(T), len(nu)
    kapint = np.zeros((nopac, nnu, nT))
    kappap = np.zeros((nopac, nT))
    kappar = np.zeros((nopac, nT))
    sm = None if smooth is None else np.ascontiguousarray(smooth, dtype=np.int32)
    nl = None if nlambda is None else np.ascontiguousarray(nlambda, dtype=np.int32)
    nfail = L.orc_setemiss(nopac, nlam, None if nl is None else _p(nl, C.c_int32), _p(nud, C.c_double),
                           _p(kap, C.c_double), _p(alb, C.c_double), nT, _p(T, C.c_double), nnu, _p(nu, C.c_double),
                           None if sm is None else _p(sm, C.c_int32), _p(kapint, C.c_double),
                           _p(kappap, C.c_double), _p(kappar, C.c_double))
    return kapint, kappap, kappar, int(nfail)
