"""mcrt_setemiss_noalloc (the allocation-free setemiss) against mcrt_setemiss.

Both run the same kernels on the same inputs; the variant takes device
tables and keeps its interval integrals in the library's fixed scratch
array, species after species in chunks.  The kernels are deterministic (no
atomics in the tables), so the tables must be bit-identical, and the
qsimp counters equal.  Two calls in a row, the second smaller, check that
the scratch kept by the first call does not leak into the second.  Sizes
beyond the scratch are refused.
"""
import ctypes as C

import numpy as np
import pytest

from src3_mpi_cfitsio_amd import model as M

pytestmark = pytest.mark.gpu


def _dust_set(nopac, nlam, seed):
    rng = np.random.default_rng(seed)
    nud, kap, alb, nl = [], [], [], []
    for k in range(nopac):
        d = M.synthetic_dust(nlam=nlam, kappa0=rng.uniform(50, 400), slope=rng.uniform(0.5, 2.0),
                             albedo=(rng.uniform(0.3, 0.7), rng.uniform(0.0, 0.1)))
        nud.append(d["nudust"]), kap.append(d["kappad"]), alb.append(d["adust"])
        nl.append(nlam - (k % 3) * 5)
    return np.array(nud), np.array(kap), np.array(alb), np.array(nl, np.int32)


class _Hip:
    """Device buffers for the test through the HIP runtime libmcrt.so links
    (the test allocates; the entry point under test must not)."""

    def __init__(self):
        self.h = C.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.live = []

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        self.live.append(p)
        assert self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0        # host to device
        return p

    def get(self, p, shape):
        out = np.zeros(shape)
        assert self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0    # device to host
        return out

    def free(self):
        for p in self.live:
            self.h.hipFree(p)
        self.live = []


def _noalloc(lib, h, nud, kap, alb, T, nu, smooth, nl):
    hip = _Hip()
    try:
        nopac, nlam = nud.shape
        nT, nnu = len(T), len(nu)
        d = {k: hip.put(v) for k, v in dict(nud=nud, kap=kap, alb=alb, T=T, nu=nu).items()}
        kapint = hip.put(np.full((nopac, nnu, nT), -1.0))
        kappap = hip.put(np.full((nopac, nT), -1.0))
        kappar = hip.put(np.full((nopac, nT), -1.0))
        nfail, nevals = C.c_int64(), C.c_int64()
        sm = np.ascontiguousarray(smooth, np.int32)
        nla = np.ascontiguousarray(nl, np.int32)
        ip = C.POINTER(C.c_int32)
        rc = lib.mcrt_setemiss_noalloc(h, nopac, nlam, nla.ctypes.data_as(ip), d["nud"], d["kap"], d["alb"],
                                       nT, d["T"], nnu, d["nu"], sm.ctypes.data_as(ip), kapint, kappap,
                                       kappar, C.byref(nfail), C.byref(nevals))
        if rc != 0:
            raise RuntimeError(lib.mcrt_last_error(h).decode())
        return dict(kapint=hip.get(kapint, (nopac, nnu, nT)), kappap=hip.get(kappap, (nopac, nT)),
                    kappar=hip.get(kappar, (nopac, nT)), nfail=nfail.value, nevals=nevals.value)
    finally:
        hip.free()


def test_noalloc_equals_setemiss_two_calls():
    from src3_mpi_cfitsio_amd import _lib
    from src3_mpi_cfitsio_amd.engine import setemiss
    lib = _lib.load()
    h = C.c_void_p()
    assert lib.mcrt_create(0, C.byref(h)) == 0
    try:
        # first call: 6 species at 500 x 513 (4.1 MB of integrals each: two
        # chunks of 4 and 2 species through the 16 MiB scratch); second
        # call smaller: 3 species at 60 x 65
        for (nT, nnu, nopac, seed) in [(500, 513, 6, 11), (60, 65, 3, 12)]:
            em = M.emission_tables(M.synthetic_dust(nlam=20), nT=nT, nnu=nnu)
            nud, kap, alb, nl = _dust_set(nopac, 40, seed)
            smooth = np.array([(k + 1) % 2 for k in range(nopac)], np.int32)
            print(f"noalloc setemiss nT={nT} nnu={nnu} nopac={nopac}", flush=True)
            ref = setemiss(nud, kap, alb, em["Tint"], em["nuint"], smooth=smooth, nlambda=nl)
            got = _noalloc(lib, h, nud, kap, alb, em["Tint"], em["nuint"], smooth, nl)
            for k in ("kapint", "kappap", "kappar"):
                assert got[k].shape == ref[k].shape
                assert np.array_equal(got[k], ref[k]), f"{k} differs (nT={nT}, nopac={nopac})"
            assert got["nfail"] == ref["nfail"]
            assert got["nevals"] == ref["nevals"]
        # beyond the scratch: nnu * nT > 2^20 is refused
        nud, kap, alb, nl = _dust_set(1, 40, 13)
        T_big = np.geomspace(1e-5, 17.0, 2100)
        nu_big = np.linspace(0.03125, 32.0, 513)
        with pytest.raises(RuntimeError, match="16 MiB scratch"):
            _noalloc(lib, h, nud, kap, alb, T_big, nu_big, np.zeros(1, np.int32), nl)
        # host arrays where device tables are expected are refused
        nT, nnu = 60, 65
        em = M.emission_tables(M.synthetic_dust(nlam=20), nT=nT, nnu=nnu)
        nud, kap, alb, nl = _dust_set(1, 40, 14)
        out = np.zeros((1, nnu, nT))
        d = C.c_double
        dp = lambda a: a.ctypes.data_as(C.POINTER(d))   # noqa: E731
        T = np.ascontiguousarray(em["Tint"])
        nu = np.ascontiguousarray(em["nuint"])
        rc = lib.mcrt_setemiss_noalloc(h, 1, 40, None, dp(nud), dp(kap), dp(alb), nT, dp(T), nnu, dp(nu), None,
                                       dp(out), dp(out), dp(out), None, None)
        assert rc != 0 and b"device pointers" in lib.mcrt_last_error(h)
    finally:
        lib.mcrt_destroy(h)
