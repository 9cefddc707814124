"""GPU tests of the path-replay adjoint (mcrt_adjoint_image).

The adjoint's gradient of L = sum(W * image_I) with respect to the cell
densities is checked against
  - the loss of an ordinary run on the same RNG counters (the forward pass
    must reproduce the peeled image the transport writes),
  - central finite differences of the loss with common random numbers, along
    +rho over the cells of largest gradient,
  - the FP64 adjoint (mixed precision against FP64, same counters),
  - torch.autograd through tests/adjoint_autograd.py.
"""
import numpy as np
import pytest

from src3_mpi_cfitsio_amd import model as M

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng_adj():
    from src3_mpi_cfitsio_amd.engine import Engine
    m = M.small("adj3d", nr=20, nt=12, nphi=8, clump_sigma=0.3, clump_seed=1003, pl=0.5,
                npeel=1, thete_deg=[57.0], nx=33, bands=((0.3, 3.0), (3.0, 300.0)))
    e = Engine(m)
    # a warm medium: re-emission from heated dust is part of the paths
    for k in range(2):
        e.zero_tallies()
        e.run_packets(20000, seed=900 + k, precision=0)
        e.dusttemp(20000, update_tdust=True)
    yield e
    e.close()


def _weights(m, seed=5):
    rng = np.random.default_rng(seed)
    return rng.uniform(0.0, 1.0, size=(m.nband, m.nx, m.nx, m.npeel)).astype(np.float32)


def test_adjoint_forward_equals_image_loss(eng_adj):
    e = eng_adj
    m = e.model
    W = _weights(m)
    n = 50000
    for prec in (0, 1):
        e.zero_tallies()
        e.run_packets(n, seed=11, precision=prec)
        img = e.tallies()["image_b_i"].astype(np.float64)       # [nband][nx][nx][npeel]
        l_img = float(np.sum(W.astype(np.float64) * img))
        l_adj, _ = e.adjoint_image(W, n, seed=11, precision=prec, grad=False)
        print(f"precision {prec}: image loss {l_img:.6e}, adjoint forward loss {l_adj:.6e}", flush=True)
        assert l_img > 0.0
        # the image is a sum of FP32 atomics, the adjoint loss FP64
        assert abs(l_adj - l_img) <= 3e-4 * l_img


def test_adjoint_gradient_matches_central_fd(eng_adj):
    """Central differences with common random numbers along +rho in the 64
    cells of largest |dL/drho * rho|.

    dL/drho has two parts: the attenuation of the peel-off walks (smooth in
    rho: FD reproduces it at any step) and the change of the flight sampling
    (the score term).  In optically thin cells the latter is mostly flights
    that switch between escaping and interacting; a 1e-3 step almost never
    samples such a switch, so common-random-number differences need a step
    of order 0.1 to see it.  The score term is unbiased but heavy-tailed: a
    packet that interacts K times in a perturbed cell adds R (1 - tau_k) per
    flight, a zero-mean sum whose spread grows as sqrt(K) (per-packet
    std/mean of about 90 for this model).  The comparison is therefore
    statistical: the cells come from an independent run (choosing them on
    the estimate under test biases it), and the adjoint and FD are paired on
    the same packets in batches; the mean paired difference must lie within
    4 standard errors plus 1%, and the standard error must itself be below 2%
    of the derivative.  Uniform weights keep the loss free of pixel-edge
    jumps."""
    e = eng_adj
    m = e.model
    W = np.ones((m.nband, m.nx, m.nx, m.npeel), dtype=np.float32)
    rho0 = m.rho.copy()
    _, gsel = e.adjoint_image(W, 200000, seed=7, precision=0, grad=True)
    cells = np.argsort(np.abs(gsel * rho0).ravel())[::-1][:64]
    d = np.zeros(rho0.size)
    d[cells] = rho0.ravel()[cells]
    d = d.reshape(rho0.shape)

    h = 0.05
    nb, nbatch = 2500000, 8

    def central(h, first):
        e.set_density(rho0 + h * d)
        lp, _ = e.adjoint_image(W, nb, seed=21, precision=0, first_packet=first, grad=False)
        e.set_density(rho0 - h * d)
        lm, _ = e.adjoint_image(W, nb, seed=21, precision=0, first_packet=first, grad=False)
        return (lp - lm) / 2.0

    preds, fds = [], []
    for b in range(nbatch):
        first = b * nb
        e.set_density(rho0)
        _, g = e.adjoint_image(W, nb, seed=21, precision=0, first_packet=first, grad=True)
        assert np.all(np.isfinite(g))
        # Richardson: (8 FD(h) - FD(2h)) / 6 cancels the h^3 term of the
        # central difference (optically thick cells curve exp(-tau))
        f1, f2 = central(h, first), central(2 * h, first)
        fds.append((8.0 * f1 - f2) / 6.0)
        preds.append(h * float(np.sum(g * d)))
    e.set_density(rho0)
    preds, fds = np.array(preds), np.array(fds)
    diff = preds - fds
    fd, se = fds.mean(), diff.std(ddof=1) / np.sqrt(nbatch)
    print(f"{nbatch} x {nb} packets: FD {fd:.6e}, adjoint {preds.mean():.6e}, ratio {preds.mean() / fd:.4f}, "
          f"paired difference {diff.mean():.4e} +- {se:.4e}", flush=True)
    assert se <= 0.02 * abs(fd)
    assert abs(diff.mean()) <= 4.0 * se + 0.01 * abs(fd)


def test_adjoint_mixed_matches_fp64(eng_adj):
    e = eng_adj
    m = e.model
    W = _weights(m)
    n = 100000
    l64, g64 = e.adjoint_image(W, n, seed=31, precision=0, grad=True)
    l32, g32 = e.adjoint_image(W, n, seed=31, precision=1, grad=True)
    rel_l = abs(l32 - l64) / abs(l64)
    rel_g = np.linalg.norm(g32 - g64) / np.linalg.norm(g64)
    print(f"mixed vs FP64: loss {rel_l:.2e}, gradient (relative L2) {rel_g:.2e}", flush=True)
    assert rel_l <= 1e-3
    assert rel_g <= 2e-2


def test_adjoint_through_torch_autograd(eng_adj):
    torch = pytest.importorskip("torch")
    from adjoint_autograd import ImageLoss
    e = eng_adj
    m = e.model
    W = _weights(m)
    n = 20000
    rho = torch.tensor(m.rho, dtype=torch.float64, requires_grad=True)
    loss = ImageLoss.apply(rho, e, W, n, 41, 1)
    (2.0 * loss).backward()
    e.set_density(m.rho)
    l_ref, g_ref = e.adjoint_image(W, n, seed=41, precision=1, grad=True)
    # the two adjoint runs differ only in the order of their FP64 atomics
    assert abs(float(loss.detach()) - l_ref) <= 1e-9 * abs(l_ref)
    np.testing.assert_allclose(rho.grad.numpy(), 2.0 * g_ref, rtol=0, atol=1e-9 * np.abs(g_ref).max())


@pytest.mark.parametrize("extra", ["accretion", "spots", "isrf"])
def test_adjoint_refuses_sources_it_does_not_replay(extra):
    """The replay runs stellar photons only: a context that would also emit
    accretion, hot-spot or external photons is refused, not silently
    differentiated as another model."""
    from src3_mpi_cfitsio_amd.engine import Engine, MCRTError
    m = M.small("adjx", nr=16, nt=10, nphi=1)
    m = {"accretion": lambda x: M.with_accretion(x, 0.2), "spots": lambda x: M.with_spots(x),
         "isrf": lambda x: M.with_isrf(x, 0.1)}[extra](m)
    W = np.ones((m.nband, m.nx, m.nx, m.npeel), np.float32)
    with Engine(m) as e:
        with pytest.raises(MCRTError, match="not part of the replayed model"):
            e.adjoint_image(W, 100, seed=1)
        {"accretion": lambda: e.set_accretion(None), "spots": lambda: e.set_spots(None),
         "isrf": lambda: e.set_isrf(0.0)}[extra]()
        loss, _ = e.adjoint_image(W, 100, seed=1, grad=False)
        assert loss > 0.0
