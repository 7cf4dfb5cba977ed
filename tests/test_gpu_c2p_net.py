"""The learned con2prim guess (RMHD): GPU training gradients against the CPU
reference (oracle/net_ref.py, pinned by finite differences), BF16 matrix-core
inference against the float64 forward pass, and what the trained network
gains over the reference's Eq. (A27) guess -- same converged primitives (the
Newton tolerance), fewer Newton passes -- trained on sampled states or on the
states of an earlier run."""
import numpy as np
import pytest

from gpluto_dev_amd import c2p_net

pytestmark = pytest.mark.gpu


def test_pairs_are_the_methods_own():
    """y = log(W/W_A27) of the generated pairs equals what the reference solve
    recovers from the conserved states."""
    u, x, y = c2p_net.make_data(4096, seed=11)
    v, passes, err, _ = c2p_net.solve(u)
    assert np.all(err == 0)
    g2 = 1.0 / (1.0 - np.sum(v[:, 1:4] ** 2, axis=1))
    h = 1.0 + 4.0 * v[:, 7] / v[:, 0]           # gamma/(gamma-1) = 4 for gamma = 4/3
    W = v[:, 0] * h * g2
    D, E = u[:, 0], u[:, 7]
    m2 = np.sum(u[:, 1:4] ** 2, 1)
    B2 = np.sum(u[:, 4:7] ** 2, 1)
    Y1 = -4.0 * (E + D - B2)
    Y2 = m2 - 2.0 * (E + D) * B2 + B2 * B2
    WA = np.maximum(D, (-Y1 + np.sqrt(np.maximum(0.0, Y1 * Y1 - 12.0 * Y2))) / 6.0)
    assert np.max(np.abs(np.log(W / WA) - y)) < 1e-5
    assert np.all(np.isfinite(x))


def test_training_gradient_matches_cpu_reference():
    import net_ref
    _, x, y = c2p_net.make_data(1000, seed=12)
    w = net_ref.random_weights(x, y, seed=1)
    loss_g, g = c2p_net.grad(w, x, y)
    loss_c, gc = net_ref.loss_and_grad(w, x, y)
    assert abs(loss_g - loss_c) <= 1e-4 * loss_c
    sl = net_ref.TRAIN
    rel = np.linalg.norm(g[sl] - gc[sl]) / np.linalg.norm(gc[sl])
    assert rel < 1e-4, rel
    assert np.max(np.abs(g[sl] - gc[sl])) <= 1e-3 * np.max(np.abs(gc[sl]))


def test_matrix_core_inference_matches_fp32_forward():
    import net_ref
    _, x, y = c2p_net.make_data(5000, seed=13)
    w = net_ref.random_weights(x, y, seed=2)
    out = c2p_net.infer(w, x)
    ref = net_ref.forward(w, x) * w[net_ref.YSC]
    # BF16 weights, inputs and activations (8 significant bits) with FP32
    # accumulation: errors of a few 2^-8 of the output's scale
    scale = np.max(np.abs(ref))
    print(f"max |out - ref| / max |ref| = {np.max(np.abs(out - ref)) / scale:.2e}", flush=True)
    assert np.max(np.abs(out - ref)) < 3e-2 * scale
    assert np.corrcoef(out, ref)[0, 1] > 0.999


def test_trained_guess_saves_newton_passes_at_the_same_answer():
    w, loss = c2p_net.train(nsamples=1 << 18, steps=3000, batch=4096, lr=3e-3, seed=21)
    u, x, y = c2p_net.make_data(1 << 17, seed=22)          # held-out states
    va, pa, ea, ms_a = c2p_net.solve(u)
    vn, pn, en, ms_n = c2p_net.solve(u, w)
    print(f"loss {loss:.4f}; Newton passes A27 {pa.mean():.3f} learned {pn.mean():.3f}; "
          f"kernel ms A27 {ms_a:.3f} learned {ms_n:.3f}", flush=True)
    assert np.all(ea == 0) and np.all(en == 0)
    assert loss < 0.2                     # explained >80% of the A27 guess's error variance
    assert pn.mean() < pa.mean()
    for q in (0, 1, 2, 3, 7):
        scale = np.maximum(np.abs(va[:, q]), 1e-3 * np.max(np.abs(va[:, q])))
        assert np.max(np.abs(vn[:, q] - va[:, q]) / scale) < 1e-8, q


def test_rmhd_step_with_learned_guess_agrees_with_drop_in():
    """The RMHD stage update with the learned guess in its con2prim matches the
    drop-in (A27 start) run to the Newton tolerance after several steps."""
    from gpluto_dev_amd import ic
    from gpluto_dev_amd.stage import Definitions, RuntimeSetup, StageUpdate
    w, _ = c2p_net.train(nsamples=1 << 16, steps=1500, batch=4096, lr=3e-3, seed=31)
    n = (20, 18, 24)
    vc, vs = ic.rmhd_turbulence(n, 14, lorentz_max=3.0, sigma=1.0)
    out = []
    for use in (False, True):
        g = StageUpdate(Definitions(physics="RMHD"), RuntimeSetup(n=n, solver="hll", gamma=4 / 3))
        if use:
            g.set_c2p_network(w)
        g.upload(vc, vs)
        err, t, dt = g.run(6, 1e-3)
        out.append((err, t, dt, g.download()))
        g.close()
    (e0, t0, d0, (v0, s0, _)), (e1, t1, d1, (v1, s1, _)) = out
    assert e0 == e1
    assert abs(t1 - t0) <= 1e-10 * t0 and abs(d1 - d0) <= 1e-9 * d0
    for q in range(8):
        l1 = np.sum(np.abs(v1[q] - v0[q])) / np.sum(np.abs(v0[q]))
        assert l1 < 1e-9, (q, l1)


def test_guess_trained_on_an_earlier_run():
    """Pairs from one RMHD run (its final conserved and primitive states), the
    network trained on them, and the guess used on another run's states:
    fewer Newton passes, the same answers."""
    from gpluto_dev_amd import ic
    from gpluto_dev_amd.stage import Definitions, RuntimeSetup, StageUpdate

    def run_states(seed):
        n = (32, 32, 32)
        vc, vs = ic.rmhd_turbulence(n, seed, lorentz_max=4.0, sigma=1.0)
        g = StageUpdate(Definitions(physics="RMHD"), RuntimeSetup(n=n, solver="hll", gamma=4 / 3))
        g.upload(vc, vs)
        err, _, _ = g.run(4, 1e-3)
        v, _, u = g.download()
        g.close()
        assert err == 0
        prim = np.stack([v[q].ravel() for q in range(8)], 1)
        cons = np.stack([u[q].ravel() for q in (0, 1, 2, 3)] + [v[q].ravel() for q in (4, 5, 6)]
                        + [u[7].ravel()], 1)
        return cons, prim

    u1, v1 = run_states(41)
    x1, y1 = c2p_net.pairs(u1, v1)
    assert np.all(np.isfinite(x1)) and np.all(np.isfinite(y1))
    w, loss = c2p_net.fit(x1, y1, steps=2000, batch=4096, lr=3e-3, seed=3)
    u2, _ = run_states(42)                       # another run: held out
    va, pa, ea, _ = c2p_net.solve(u2)
    vn, pn, en, _, rs = c2p_net.solve(u2, w, restarts=True)
    print(f"loss {loss:.4f}; Newton passes A27 {pa.mean():.3f} learned {pn.mean():.3f}; "
          f"restarts {rs.mean():.2e}", flush=True)
    assert np.all(ea == 0) and np.all(en == 0)
    assert pn.mean() < pa.mean()
    for q in (0, 1, 2, 3, 7):
        scale = np.maximum(np.abs(va[:, q]), 1e-3 * np.max(np.abs(va[:, q])))
        assert np.max(np.abs(vn[:, q] - va[:, q]) / scale) < 1e-8, q
