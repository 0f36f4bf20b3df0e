"""Request generation of the serving loop and the snapshot series (CPU)."""
import numpy as np

from fastqsl_amd import serve, synth


def test_poisson_requests_are_reproducible_and_inside_the_field():
    a = serve.poisson_requests(rate=50, duration=20, nsnap=16, shape=(320, 320, 256), seed=3)
    b = serve.poisson_requests(rate=50, duration=20, nsnap=16, shape=(320, 320, 256), seed=3)
    assert [(r.t_arrival, r.n, r.x0) for r in a] == [(r.t_arrival, r.n, r.x0) for r in b]
    assert 800 < len(a) < 1200                       # ~ rate x duration
    t = np.array([r.t_arrival for r in a])
    assert np.all(np.diff(t) > 0)
    gaps = np.diff(np.concatenate([[0], t]))
    assert abs(gaps.mean() - 1 / 50) < 0.003          # exponential gaps of mean 1/rate
    n = np.array([r.n for r in a])
    assert n.min() >= 128 and n.max() <= 2048
    assert abs(np.mean(np.log(n)) - (np.log(128) + np.log(2048)) / 2) < 0.1   # log-uniform
    for r in a:
        assert 0 <= r.x0 and r.x0 + r.side <= 319 and 0 <= r.y0 and r.y0 + r.side <= 319
        assert 0 <= r.snapshot < 16
    assert len(serve.poisson_requests(rate=1, duration=0.001, nsnap=1, shape=(64, 64, 64),
                                      min_count=5)) == 5


def test_request_seeds_span_the_window():
    import torch
    r = serve.QRequest(0, 0.0, 0, x0=10.0, y0=20.0, side=30.0, n=7)
    x, y, z = r.seeds(device="cpu")
    assert x.numel() == 49 and torch.all(z == 0)
    X = x.numpy().reshape(7, 7)
    Y = y.numpy().reshape(7, 7)
    np.testing.assert_allclose(X[0], np.linspace(10, 40, 7), rtol=1e-6)
    np.testing.assert_allclose(Y[:, 0], np.linspace(20, 50, 7), rtol=1e-6)


def test_snapshot_perturbation_amplitude():
    shape = (48, 40, 32)
    base = synth.multipole(shape, seed=0, dtype=np.float64, device="cpu")
    for i in (0, 5):
        s = synth.snapshot(i, shape=shape, dtype=np.float64, device="cpu")
        eps = np.random.default_rng(1000 + i).uniform(0.01, 0.05)
        d = [a - b for a, b in zip(s, base)]
        r = np.sqrt(sum(x[:, :, 0] ** 2 for x in d).mean() / sum(x[:, :, 0] ** 2 for x in base).mean())
        np.testing.assert_allclose(r, eps, rtol=1e-6)
    assert not np.allclose(synth.snapshot(1, shape=shape, device="cpu")[2],
                           synth.snapshot(2, shape=shape, device="cpu")[2])


def test_fourier_perturbation_is_divergence_free():
    """Second-order convergence of the centred-difference divergence to 0."""
    errs = []
    for n in (33, 65):
        dbx, dby, dbz = synth.fourier_perturbation((n, n, n), 3, device="cpu", modes=2, kmax=2)
        h = 2 / (n - 1)
        div = ((dbx[:, :, 2:] - dbx[:, :, :-2])[1:-1, 1:-1, :] + (dby[:, 2:, :] - dby[:, :-2, :])[1:-1, :, 1:-1]
               + (dbz[2:] - dbz[:-2])[:, 1:-1, 1:-1]) / (2 * h)
        g = (dbx[:, :, 2:] - dbx[:, :, :-2]).abs().max() / (2 * h)
        errs.append(float(div.abs().max() / g))
    assert errs[1] < errs[0] / 3
