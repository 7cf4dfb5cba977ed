"""The CPU reference of the con2prim-guess network (oracle/net_ref.py) is
pinned by central finite differences of its own loss, so that it can check
the GPU training kernel's gradients."""
import numpy as np


def test_reference_gradient_matches_finite_differences():
    import net_ref
    rng = np.random.default_rng(3)
    x = (rng.normal(size=(64, 8)) * np.array([2.0, 0.5, 3.0, 0.3, 1.0, 0.7, 2.5, 0.4])
         + np.array([1.0, 0.2, -1.0, 0.5, 0.0, -0.3, 2.0, 1.0]))
    y = rng.normal(size=64) * 0.05
    w = net_ref.random_weights(x, y, seed=5).astype(np.float64)
    _, g = net_ref.loss_and_grad(w, x, y)
    idx = list(range(net_ref.W1, net_ref.B3 + 1, 37)) + [net_ref.B3, net_ref.W3 + 7, net_ref.B1 + 3]
    for q in idx:
        h = 1e-6 * max(1.0, abs(w[q]))
        wp, wm = w.copy(), w.copy()
        wp[q] += h
        wm[q] -= h
        fd = (net_ref.loss_and_grad(wp, x, y)[0] - net_ref.loss_and_grad(wm, x, y)[0]) / (2 * h)
        assert abs(fd - g[q]) <= 1e-6 + 1e-5 * abs(fd), (q, fd, g[q])
    # the fixed entries (normalisation, target scale) carry no gradient
    assert np.all(g[:net_ref.W1] == 0) and g[net_ref.YSC] == 0
