"""Pins the FP64 reference gradient (oracle/diff_oracle.cpp) on the CPU.

splat-tui has no backward, so the reference gradient is checked against
what the mathematics fixes: central finite differences of the same FP64
loss (step 1e-6, truncation error ~1e-12 relative) for every parameter
group, at every SH degree; its forward image against the independent FP64
renderer of the cpu-naive semantics; and a planted sign error in the
analytic gradient must fail the finite-difference check.
"""
import numpy as np
import pytest

import splat_amd as S
from splat_amd.diff import params_from_scene

GROUPS = {"mean": range(0, 3), "log_scale": range(3, 6), "quat": range(6, 10), "logit": [10],
          "colour": range(11, 14)}


def session(oracle, deg, n=300, seed=7, size=(96, 64), nviews=2):
    s = S.random_scene(n, seed=seed, sh_degree=deg)
    prm = params_from_scene(s)
    views = S.orbit_views(s, 40, 20.0)[:nviews]
    sizes = [size] * nviews
    rng = np.random.default_rng(seed)
    tprm = {k: (v + rng.normal(0, 0.01, v.shape)).astype(np.float32) for k, v in prm.items()}
    tgt = oracle.DiffF64(tprm, deg, views, sizes, (0.1, 0.2, 0.3))
    targets = [tgt.image(v).astype(np.float32) for v in range(nviews)]
    return s, prm, views, sizes, oracle.DiffF64(prm, deg, views, sizes, (0.1, 0.2, 0.3), targets)


@pytest.mark.parametrize("deg", [0, 1, 2, 3])
def test_fp64_gradient_matches_finite_differences(oracle, deg):
    _, _, _, _, D = session(oracle, deg)
    g = D.grad_flat()
    rows = np.nonzero(np.abs(g).sum(1))[0]
    assert len(rows) > 50
    rng = np.random.default_rng(deg)
    groups = dict(GROUPS)
    if deg > 0:
        groups["sh"] = range(14, D.np)
    for name, ks in groups.items():
        rel = []
        for _ in range(12):
            i = int(rng.choice(rows))
            k = int(rng.choice(list(ks)))
            fd = D.fd(i, k, 1e-6)
            an = g[i, k]
            rel.append(abs(fd - an) / max(abs(fd), abs(an), 1e-9))
        # Threshold crossings (alpha = 1/255, T = 1e-4) inside +-1e-6 would show as
        # isolated outliers; none are expected at this step size.
        assert np.median(rel) < 1e-6, (name, rel)
        assert np.max(rel) < 1e-4, (name, rel)


def test_fp64_forward_matches_fp64_renderer(oracle):
    s, _, views, sizes, D = session(oracle, 3)
    for v in range(len(views)):
        w, h = sizes[v]
        rgb, _ = oracle.render_f64(s, views[v], w, h, (0.1, 0.2, 0.3))
        # The reference scene stores activated f32 values; the session re-activates the
        # f32 raw parameters in double, a 1e-7 relative difference per parameter.
        assert np.abs(D.image(v) - rgb).max() < 1e-4


def test_planted_sign_error_fails_finite_differences(oracle):
    """A flipped sign in one gradient row is caught by the finite-difference check."""
    _, _, _, _, D = session(oracle, 1)
    g = D.grad_flat()
    rows = np.nonzero(np.abs(g[:, 0]))[0]
    i = int(rows[np.argmax(np.abs(g[rows, 0]))])
    planted = -g[i, 0]
    fd = D.fd(i, 0, 1e-6)
    assert abs(fd - g[i, 0]) < 1e-6 * max(abs(fd), 1e-9) + 1e-12
    assert abs(fd - planted) > 0.5 * abs(fd)


def test_loss_is_sum_of_squares(oracle):
    """With zero targets the loss is the sum of squared pixel values."""
    s = S.random_scene(300, seed=7, sh_degree=0)
    views = S.orbit_views(s, 40, 20.0)[:2]
    sizes = [(96, 64)] * 2
    Z = oracle.DiffF64(params_from_scene(s), 0, views, sizes, (0.1, 0.2, 0.3))
    zimg = sum(float((Z.image(v) ** 2).sum()) for v in range(len(views)))
    assert Z.loss() == pytest.approx(zimg, rel=1e-12)
