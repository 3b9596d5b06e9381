"""The FP64 MLP oracle's gradient against central finite differences (CPU)."""
import numpy as np

import mlp


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    params = rng.normal(0, 0.3, mlp.N_PARAMS)
    x = rng.normal(0, 1, (40, mlp.NI))
    t = rng.normal(0, 1, (40, mlp.NO))
    _, grad = mlp.loss_and_grad(params, x, t)
    idx = rng.choice(mlp.N_PARAMS, 60, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros(mlp.N_PARAMS)
        e[i] = h
        fd = (mlp.loss_only(params + e, x, t) - mlp.loss_only(params - e, x, t)) / (2 * h)
        assert abs(fd - grad[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, grad[i])


def test_oracle_layout_matches_binding():
    import pfh

    assert mlp.N_PARAMS == pfh.MLP_PARAMS
    assert mlp.NO == pfh.MLP_OUTPUTS
