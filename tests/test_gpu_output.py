"""GPU: one Lucy iteration's products written in the reference's formats,
and a run resumed from the tdave checkpoint (initarr_2_new.f90:55-75)
giving the run that continues in the same process."""
import os

import numpy as np
import pytest

from src3_mpi_cfitsio_amd import model as M
from src3_mpi_cfitsio_amd import output as O

pytestmark = pytest.mark.gpu


def test_iteration_outputs_and_resume(tmp_path):
    from src3_mpi_cfitsio_amd.engine import Engine, diskflux
    m = M.build("out", nr=20, nt=16, nphi=1, npeel=2, nx=9, nfreq=10, nmu=3, nlam=20, nT=50, nnu=33,
                bands=((1.0, 2.0),), thete_deg=[30.0, 80.0])
    n = 20000
    d = str(tmp_path)
    with Engine(m) as e:
        tal = e.disk(n, seed=5)
        flux = diskflux(tal, m, n)
        files = O.output(d, flux, m, band_names=["2K"], nabs=tal["nabs"], tdust=tal["tdave"], iteration=1)
        print("wrote", len(files), "files", flush=True)
        e.zero_tallies()
        e.set_tdust((tal["tdave"] * M.TK) / M.TK)      # the value a checkpoint round trip gives
        e.run_packets(n, seed=6)
        cont = e.tallies()

    cube = O.read_fits(os.path.join(d, "flux_hypercube.fits.gz"))[0].data
    si = flux["si"]
    pos = si > 0
    assert pos.any()
    assert np.array_equal(cube[0][pos], (si / O._nunorm(m))[pos]) and np.all(cube[0][~pos] == 0)
    assert np.array_equal(O.read_grid(os.path.join(d, "nabs.fits.gz"))[0], tal["nabs"])
    img = O.read_fits(os.path.join(d, "e_2K_80.0_0.0_I_img.fits.gz"))[0].data
    assert np.array_equal(img, flux["image_b_i"][0, :, :, 1])
    t_fits = O.read_grid(os.path.join(d, "tdave_001.fits.gz"))[0]
    np.testing.assert_allclose(t_fits, tal["tdave"] * M.TK, rtol=1e-7)

    with Engine(m) as e2:
        e2.set_tdust(O.read_tdave(os.path.join(d, "tdave_001.unf"), m))
        e2.zero_tallies()
        e2.run_packets(n, seed=6)
        res = e2.tallies()
    for k in ("nabs", "killed", "nums", "numt"):
        assert np.array_equal(res[k], cont[k]), k
    np.testing.assert_allclose(res["dtauabs"], cont["dtauabs"], rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(res["si"], cont["si"], rtol=1e-5, atol=1e-6 * float(cont["si"].max()))
