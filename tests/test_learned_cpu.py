"""CPU checks of the warm-start classifier's oracle and of the learned pipeline's host logic."""
import numpy as np

import mlp_oracle as M
import synth


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    w = rng.normal(0, 0.3, M.NPAR)
    X = rng.normal(0, 1, (64, M.NIN))
    y = (rng.random(64) < 0.5).astype(np.float64)
    g, l0 = M.gradient(w, X, y)
    assert abs(l0 - M.loss(w, X, y)) < 1e-14
    h = 1e-6
    # every bias and output weight, and a sample of the hidden weights
    idx = np.concatenate([np.arange(512, 576), np.arange(4672, 4801), rng.choice(M.NPAR, 200, replace=False)])
    for i in idx:
        e = np.zeros(M.NPAR)
        e[i] = h
        fd = (M.loss(w + e, X, y) - M.loss(w - e, X, y)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd), (i, fd, g[i])


def test_gradient_of_one_sample_closed_form():
    # one sample, all hidden units active: dL/db3 = sigmoid(z) - y
    rng = np.random.default_rng(1)
    w = np.abs(rng.normal(0, 0.2, M.NPAR))
    X = np.abs(rng.normal(0, 1, (1, M.NIN)))
    z, _ = M.forward(w, X)
    for yv in (0.0, 1.0):
        g, _ = M.gradient(w, X, np.array([yv]))
        assert np.isclose(g[-1], 1.0 / (1.0 + np.exp(-z[0])) - yv, rtol=1e-14)


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -2.5, 1e-30, 3.0e38], np.float32)
    r = M.bf16_round(v)
    # ties to even: 1 + 2^-8 -> 1; 1 + 3*2^-8 -> 1 + 2^-6
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == np.float32(1.0 + 2.0 ** -6)
    assert r[3] == -2.5
    assert (r.view(np.uint32) & 0xFFFF == 0).all()
    assert np.all(np.abs(r - v) <= np.abs(v) * 2.0 ** -8)


def test_previous_state_layout():
    import learned
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 2, 0], [5, 5, 5], [6, 5, 5]], np.float32)
    pop = synth.Population(pos, np.zeros_like(pos), np.ones(5, np.float32), np.array([0, 3, 5]), np.arange(5))
    res = np.zeros(2, dtype=[("nbound", np.int64), ("mbound", np.float32), ("com_pos", np.float32, 3),
                             ("com_vel", np.float32, 3)])
    res["nbound"] = [2, 0]
    res["mbound"] = [2.0, 0.0]
    res["com_pos"][0] = [0.5, 0, 0]
    res["com_pos"][1] = [5.5, 5, 5]
    out = dict(order=np.array([1, 0, 2, 1, 0]), energy=np.array([-3, -2, 4, 1, 2], np.float32), result=res)
    st = learned.previous_state(pop, out)
    assert st["rank"].tolist() == [1, 0, 2, 1, 0]
    assert st["bound"].tolist() == [1, 1, 0, 0, 0]
    assert st["energy"].tolist() == [-2, -3, 4, 2, 1]
    assert np.isclose(st["rscale"][0], 0.5)
    assert np.isclose(st["rscale"][1], 0.5)  # nothing bound: rms over all particles
    assert st["mbound"].tolist() == [2.0, 2.0]
