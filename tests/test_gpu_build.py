"""GPU model building (mcrt_build_models) against the host builder.

Both evaluate the same expressions in the same order (gridset's cell loop,
densenv and zerod, the synthetic flared disk); the GPU's libm and numpy's
may differ in the last bit of pow/cos/acos/exp, and Cardano's form in zerod
cancels for r >> Rc near the equator, so densities agree to 1e-10 relative
and volumes to 1e-13.  Species must be identical, and the disk must carry
its mass.  Two builds of the same input are bit-identical (fixed-order
reductions, no atomics).
"""
import numpy as np
import pytest

from src3_mpi_cfitsio_amd import model as M

pytestmark = pytest.mark.gpu


def _same(a, b, what):
    np.testing.assert_allclose(a.vcell, b.vcell, rtol=1e-13, err_msg=f"{what} vcell")
    np.testing.assert_allclose(a.rho, b.rho, rtol=1e-10, atol=0, err_msg=f"{what} rho")
    assert np.array_equal(a.species, b.species), f"{what} species"


@pytest.mark.parametrize("nphi,clump", [(1, 0.0), (8, 0.7)])
def test_device_model_equals_host_model(nphi, clump):
    kw = dict(nr=60, nt=40, nphi=nphi, npeel=2, nx=9, nfreq=10, nmu=4, nlam=20, nT=50, nnu=33,
              clump_sigma=clump, mdisk=0.02, mdot=3e-6)
    host = M.build("h", **kw)
    dev = M.build("d", device=True, **kw)
    print(f"nphi={nphi} clump={clump}: {host.ncell} cells, disk cells {np.sum(host.species == 1)}", flush=True)
    _same(dev, host, f"nphi={nphi}")
    again = M.build("d2", device=True, **kw)
    assert np.array_equal(again.rho, dev.rho) and np.array_equal(again.vcell, dev.vcell)
    disk = np.sum(np.where(dev.species == 1, dev.rho, 0.0) * dev.vcell)
    assert disk > 0.5 * 0.02 * M.MSUN


def test_device_batch_equals_host_batch():
    from src3_mpi_cfitsio_amd.engine import build_models
    host = M.batch_models(24, seed=1002, nfreq=20, nlam=40, nT=100, nnu=65)
    dev = M.batch_models(24, seed=1002, nfreq=20, nlam=40, nT=100, nnu=65, device=True)
    for k, (a, b) in enumerate(zip(dev.models, host.models)):
        _same(a, b, f"model {k}")
    walls = [(m.rarr, m.thetarr, m.phiarr) for m in host.models]
    par = np.array([M.density_params(2.0, p["mdisk"], p["rin_au"], 1000.0, flare=p["flare"], mdot_msun_yr=p["mdot"])
                    for p in host.params])
    _, ms = build_models(walls, par)
    print(f"24 models, {sum(m.ncell for m in host.models)} cells: GPU build {ms:.3f} ms", flush=True)


def test_build_rejects_bad_grids():
    from src3_mpi_cfitsio_amd.engine import MCRTError, build_models
    w = (np.array([1.0]), M.theta_walls(4), M.phi_walls(1)[0])
    with pytest.raises(MCRTError, match="nrg"):
        build_models([w], np.zeros((1, 12)))
