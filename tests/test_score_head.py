"""ScoreHead (pirate_amd.scorer): the logistic-regression head over the network's logit and
the peak finder's statistic, checked against what the mathematics fixes:
  * on labels drawn from a known logistic model in (1, z, p, z p), the fitted log-odds match
    the true log-odds (maximum-likelihood consistency, n = 400k);
  * at the fit, the gradient of the log-likelihood vanishes (first-order optimality of the
    Newton iterations, up to the ridge term);
  * flipping every label negates the fitted log-odds (the model's symmetry);
  * the mean log-loss falls monotonically over the Newton iterations."""

import numpy as np

from pirate_amd.scorer import ScoreHead


def _data(n, seed):
    rng = np.random.default_rng(seed)
    z = rng.standard_normal(n) * 2.0 - 1.0
    p = rng.gamma(4.0, 1.5, n)
    L = -3.0 + 0.8 * z + 0.5 * p + 0.15 * z * p
    y = (rng.random(n) < 1.0 / (1.0 + np.exp(-L))).astype(np.float64)
    return z, p, y, L


def test_recovers_known_log_odds():
    z, p, y, L = _data(400_000, 1)
    h = ScoreHead()
    hist = h.fit(z, p, y)
    s = h(z, p)
    # the error of the fitted log-odds in units of its standard error (from the inverse Fisher
    # information at the fit): for a correct maximum-likelihood fit the largest such ratio over
    # all inputs is sqrt(chi^2 with 4 degrees of freedom), below 6 with probability 1 - 1e-6
    X = h._features(z, p)
    q = 1.0 / (1.0 + np.exp(-s))
    cov = np.linalg.inv((X * (q * (1.0 - q))[:, None]).T @ X)
    se = np.sqrt(np.einsum('ij,jk,ik->i', X, cov, X))
    ratio = np.max(np.abs(s - L) / se)
    assert ratio < 6.0, ratio
    assert all(b <= a + 1e-12 for a, b in zip(hist, hist[1:])), hist


def test_gradient_vanishes_at_the_fit():
    z, p, y, _ = _data(50_000, 2)
    h = ScoreHead()
    h.fit(z, p, y, ridge=0.0)
    X = h._features(z, p)
    q = 1.0 / (1.0 + np.exp(-(X @ h.w)))
    g = X.T @ (y - q) / len(y)
    assert np.max(np.abs(g)) < 1e-9, g


def test_flipped_labels_negate_the_score():
    z, p, y, _ = _data(50_000, 3)
    a, b = ScoreHead(), ScoreHead()
    a.fit(z, p, y)
    b.fit(z, p, 1.0 - y)
    np.testing.assert_allclose(b(z, p), -a(z, p), rtol=0, atol=1e-8)
