"""Pins the MLP oracle's gradients by central finite differences."""

from oracle import mlp_oracle as M


def test_mlp_oracle_gradients_match_finite_differences(rng):
    n_in, H, B = 5, 16, 7
    n = n_in * H + H + H * H + H + H + 1
    p = rng.normal(0, 0.4, n)
    X = rng.normal(0, 1, (B, n_in))
    t = rng.normal(0, 1, B)
    loss, g = M.loss_and_grad(p, X, t, n_in, H)
    for i in rng.choice(n, 60, replace=False):
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (M.loss_and_grad(pp, X, t, n_in, H)[0] - M.loss_and_grad(pm, X, t, n_in, H)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 * (1 + abs(fd)), i
