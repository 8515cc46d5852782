"""Pins the FP64 numpy MLP oracle (oracle/mlp.py) by central finite
differences of its own loss — the check the GPU's BF16 training kernels are
then compared against.  CPU only."""
from __future__ import annotations

import numpy as np

from oracle import mlp as OM


def random_problem(rng, B=64):
    P = np.zeros(OM.OFF["N"])
    for k, (a, n) in ((k, v) for k, v in OM.OFF.items() if k != "N"):
        fan = OM.NIN if k == "W1" else OM.HID
        P[a:a + n] = rng.uniform(-1, 1, n) * np.sqrt(6.0 / fan) * (0.3 if k.startswith("b") else 1.0)
    a5, n5 = OM.OFF["W5"]
    W5 = P[a5:a5 + n5].reshape(OM.NOUTP, OM.HID)
    W5[OM.NOUT:] = 0.0
    ab, nb = OM.OFF["b5"]
    P[ab + OM.NOUT:ab + nb] = 0.0
    X = rng.normal(size=(B, OM.NIN))
    d = rng.normal(size=(B, 3)); d /= np.linalg.norm(d, axis=1, keepdims=True)
    n = rng.normal(size=(B, 3)); n /= np.linalg.norm(n, axis=1, keepdims=True)
    fate = (rng.random(B) < 0.3).astype(float)
    valid = (rng.random(B) < 0.95).astype(float)
    T = np.concatenate([d, fate[:, None], valid[:, None], n], axis=1)
    return P, X, T


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    P, X, T = random_problem(rng)
    w_dir = 50.0
    L, g, _ = OM.loss_grad(P, X, T, w_dir)
    # sample coordinates from every parameter block
    picks = []
    for k, (a, n) in ((k, v) for k, v in OM.OFF.items() if k != "N"):
        size = n if k not in ("W5", "b5") else (OM.NOUT * OM.HID if k == "W5" else OM.NOUT)
        picks += list(a + rng.choice(size, size=min(6, size), replace=False))
    for i in picks:
        h = 1e-6
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += h; Pm[i] -= h
        fd = (OM.loss_only(Pp, X, T, w_dir) - OM.loss_only(Pm, X, T, w_dir)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])
    # padded output rows carry no gradient
    a5, _ = OM.OFF["W5"]
    assert np.all(g[a5 + OM.NOUT * OM.HID:a5 + OM.NOUTP * OM.HID] == 0)
