"""Write tests/golden/ref_vectors.npz: input/output vectors of the reference's
own compiled routines (oracle/_ref/libref.so, built from the reference
sources by oracle/ref/Makefile).  Run in a container that has the reference:

    make -C oracle/ref && python tests/golden/make_ref_vectors.py

The fixture holds only data (inputs and the reference's outputs); the tests
compare the C oracle against it when libref.so itself is not available.
"""
import ctypes as C
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle_py  # noqa: E402

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


def _d(a):
    return a.ctypes.data_as(_dp)


def wall_cases(n, seed):
    """Random rays plus the degenerate ones: on-wall starts, equator plane,
    pole walls, tangent and axis-parallel directions."""
    rng = np.random.default_rng(seed)
    u = rng.normal(size=(n, 3))
    x = rng.normal(size=(n, 3)) * rng.uniform(0.1, 300.0, size=(n, 1))
    u[0::13, 2] = 0.0                       # horizontal rays (uz = 0): equator plane parallel
    u[1::17, :2] = 0.0                      # rays along the z axis (aa ~ 0 on cones)
    u /= np.linalg.norm(u, axis=1)[:, None]
    r2 = np.sort(rng.uniform(0.5, 4e4, size=(n, 2)), axis=1)
    # start exactly on the inner shell for a quarter of the rays
    on = rng.uniform(size=n) < 0.25
    rr = np.sqrt(np.sum(x * x, axis=1))
    x[on] *= (np.sqrt(r2[on, 0]) / rr[on])[:, None]
    t2 = rng.uniform(0.0, 6.0, size=(n, 2))
    t2[0::7, 0] = -1.0                      # equator flag
    t2[3::11, 1] = 0.0                      # pole wall
    t2[5::19, 0] = 0.0
    # tangent rays: x perpendicular to u
    tang = np.arange(n) % 23 == 4
    x[tang] -= (np.sum(x[tang] * u[tang], axis=1))[:, None] * u[tang]
    ab = rng.normal(size=(n, 4))
    ow = (rng.uniform(size=n) < 0.5).astype(np.int32)
    return u, x, r2, t2, ab, ow


def main():
    R = oracle_py.ref()
    if R is None:
        raise SystemExit("build oracle/_ref/libref.so first (make -C oracle/ref)")
    out = {}
    # --- L'Ecuyer-Cote streams (ecuyer_cote_mod.f90 via user_set_all + ran)
    seeds = np.array([[12345678, 87654321, 1], [12345678 - 7, 87654321, 3], [1234567890, 123456789, 32],
                      [12345678 - 1, 87654321, 31]], dtype=np.int32)
    lec = np.zeros((len(seeds), 256))
    for k, (s1, s2, g) in enumerate(seeds):
        R.ref_user_set_all(int(s1), int(s2), int(g))
        lec[k] = [R.ref_ran() for _ in range(256)]
    out["lec_seeds"], out["lec_draws"] = seeds, lec
    # --- bbfreq (random.f90:44-72) on a seeded stream
    R.ref_user_set_all(12345678, 87654321, 2)
    teffs = np.array([4000.0 / 11605.0, 0.01, 1.0])
    bb = np.zeros((len(teffs), 400))
    for k, t in enumerate(teffs):
        bb[k] = [R.ref_bbfreq(float(t)) for _ in range(400)]
    out["bb_seed"] = np.array([12345678, 87654321, 2], np.int32)
    out["bb_teff"], out["bb_draws"] = teffs, bb
    # --- face roots and closest-wall selection
    n = 3000
    u, x, r2, t2, ab, ow = wall_cases(n, 7)
    st_t = np.zeros((n, 3))
    st_i = np.zeros((n, 3), np.int32)
    nw_t = np.zeros(n)
    nw_i = np.zeros(n, np.int32)
    for k in range(n):
        R.ref_reset_t()
        uk, xk = np.ascontiguousarray(u[k]), np.ascontiguousarray(x[k])
        R.ref_radface(_d(uk), _d(xk), float(r2[k, 0]), float(r2[k, 1]))
        R.ref_thetaface(_d(uk), _d(xk), float(t2[k, 0]), float(t2[k, 1]))
        R.ref_phiface(float(ab[k, 0]), float(ab[k, 1]), float(ab[k, 2]), float(ab[k, 3]), _d(uk), _d(xk))
        tt = np.zeros(3)
        ii = np.zeros(3, np.int32)
        R.ref_walls_state(_d(tt), ii.ctypes.data_as(_ip))
        st_t[k], st_i[k] = tt, ii
        # find_next_wall stops the program without a candidate: only call it when one exists
        cand = ii[1] if not ow[k] else (ii[1] if abs(tt[0]) < abs(tt[1]) else ii[2])
        if cand != 0:
            t = C.c_double()
            i = C.c_int32()
            R.ref_find_next_wall(int(ow[k]), C.byref(t), C.byref(i))
            nw_t[k], nw_i[k] = t.value, i.value
    out.update(w_u=u, w_x=x, w_r2=r2, w_t2=t2, w_ab=ab, w_onwall=ow, w_state_t=st_t, w_state_i=st_i,
               w_next_t=nw_t, w_next_i=nw_i)
    # --- findangle3 (both ri1 branches, poles, equal azimuths)
    rng = np.random.default_rng(3)
    m = 3000
    t1, t2a = rng.uniform(0, math.pi, m), rng.uniform(0, math.pi, m)
    p1, p2 = rng.uniform(0, 2 * math.pi, m), rng.uniform(0, 2 * math.pi, m)
    t1[0::29] = 1e-9
    p2[1::31] = p1[1::31]
    t2a[2::37] = t1[2::37]
    p2[2::37] = p1[2::37]
    fa_in = np.stack([np.cos(t1), np.sin(t1), p1, np.cos(p1), np.sin(p1),
                      np.cos(t2a), np.sin(t2a), p2, np.cos(p2), np.sin(p2)], axis=1)
    fa_out = np.zeros((m, 4))
    for k in range(m):
        a = np.ascontiguousarray(fa_in[k])
        o = np.zeros(4)
        R.ref_findangle3(_d(a), _d(o))
        fa_out[k] = o
    out["fa_in"], out["fa_out"] = fa_in, fa_out
    # --- locate (nr_2.f:131-154) on ascending and descending tables
    asc = np.cumsum(rng.uniform(0.1, 1.0, 50))
    desc = np.cos(np.linspace(0, math.pi, 41))
    q_asc = np.concatenate([rng.uniform(asc[0] - 1, asc[-1] + 1, 300), asc])
    q_desc = np.concatenate([rng.uniform(-1.1, 1.1, 300), desc])
    out["loc_asc"], out["loc_desc"], out["loc_qa"], out["loc_qd"] = asc, desc, q_asc, q_desc
    out["loc_ra"] = np.array([R.ref_locate(_d(asc), len(asc), float(q)) for q in q_asc], np.int32)
    out["loc_rd"] = np.array([R.ref_locate(_d(desc), len(desc), float(q)) for q in q_desc], np.int32)
    # --- check.f90 clamp and mathsubs.f90 Bnu
    cv = np.array([-2.0, -1.0000001, -1.0, -0.3, 0.0, 0.7, 1.0, 1.0000001, 1.02, 5.0])
    out["chk_in"], out["chk_out"] = cv, np.array([R.ref_check(float(v)) for v in cv])
    xs = np.linspace(0.0, 40.0, 161)
    out["bnu_x"], out["bnu"] = xs, np.array([R.ref_bnu(float(v), 1.0) for v in xs])
    # --- setemiss pieces: mathsubs.f90 expm1/dBnu and nr_2.f qsimp over
    # one interval of kappaB / chidB (dust_5.f90:560-575), including the
    # first interval (nu0 = 0) and intervals near the x = 32 end
    ex = np.concatenate([np.linspace(-3.0, 3.0, 241), [0.6931472, -0.6931472, 0.69314721, 1e-12, -1e-9]])
    out["em1_x"], out["em1"] = ex, np.array([R.ref_expm1(float(v)) for v in ex])
    dx_ = np.concatenate([np.linspace(0.0, 40.0, 161)[1:], [-0.5, -2.0]])
    out["dbnu_x"], out["dbnu"] = dx_, np.array([R.ref_dbnu(float(v), 1.0) for v in dx_])
    nseg = 400
    seg = np.zeros((nseg, 6))
    seg[:, 0] = np.arange(nseg) % 2                                   # kind
    seg[:, 1:3] = rng.uniform(0.01, 5.0, (nseg, 2))                   # k0, k1
    T = 10.0 ** rng.uniform(-5.0, 1.5, nseg)
    x0 = rng.uniform(0.0, 31.9, nseg)
    x0[:20] = 0.0
    x0[20:40] = 31.9375
    dxs = rng.uniform(1e-3, 0.07, nseg)
    seg[:, 3], seg[:, 4], seg[:, 5] = x0 * T, dxs * T, T              # nu0, dnu, teff
    res, flag = np.zeros(nseg), np.zeros(nseg, np.int32)
    for k in range(nseg):
        r, cv = C.c_double(), C.c_int()
        R.ref_qsimp_seg(int(seg[k, 0]), *[float(v) for v in seg[k, 1:]], C.byref(r), C.byref(cv))
        res[k], flag[k] = r.value, cv.value
    out["qs_in"], out["qs_out"], out["qs_flag"] = seg, res, flag
    # --- zerod.f90: the Ulrich streamline cubic, both branches, the rx = 1
    # boundary and the equator (mu = 0); xmu0 starts at -7 so that a case
    # that leaves it unset shows
    zmu = np.concatenate([rng.uniform(0.0, 1.0, 400), [0.0, 0.0, 1.0, 1e-12, 0.5, 0.999999]])
    zz = np.concatenate([10.0 ** rng.uniform(-4.0, 4.0, 400), [0.5, 2.0, 1.0, 1.0, 1.0 - 1e-9, 1e-6]])
    zout, zfl = np.zeros(len(zz)), np.zeros(len(zz), np.int32)
    for k in range(len(zz)):
        r, cv = C.c_double(-7.0), C.c_int()
        R.ref_zerod(float(zmu[k]), float(zz[k]), C.byref(r), C.byref(cv))
        zout[k], zfl[k] = r.value, cv.value
    out["zd_mu"], out["zd_z"], out["zd_out"], out["zd_flag"] = zmu, zz, zout, zfl
    path = os.path.join(HERE, "ref_vectors.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
