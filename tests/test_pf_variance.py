"""Analytic peak-finder variances (PfAvarApproximation, fill_analytic_weights' input).

Reference: src_lib/PfVariance.cpp:238-366 (sparse-tile evaluation), pirate_frb/slow_avar/
PfVariance.py:465-503 (definition of the approximation). The reference's Python cannot be
imported here and its C++ needs ksgpu, so the library's closed form is checked against:

  * oracle/avar.py, a dense restatement (impulse of every channel gridded, pushed through
    k passes of the tree recursion on whole arrays, filtered by every peak-finding kernel
    built from the boxcar recursion, squared and summed) -- to 1e-12 relative;
  * the dense restatement itself is pinned here: its tree levels against the reference's
    closed-form delays (src_lib/utils.cpp:287-301) and the oracle's whole-array tree, its
    kernels against the bare-kernel variances and the kernel widths, its energies against
    direct convolution;
  * Monte Carlo through the oracle's full float64 chain (lagged downsampling, two-stage
    trees, subbands, the reference peak finder's sum-of-squares output): with weights
    1/sqrt(variance), the weighted outputs of white noise with the given per-channel
    variances have unit variance up to the approximation's error (the reference's purpose:
    out_max comes out as an SNR).
"""

import numpy as np
import pytest

from gpu_cases import CASES, LARGE


def bitrev(i, n):
    j = 0
    for _ in range(n):
        j = (j << 1) | (i & 1)
        i >>= 1
    return j


def _plan(d):
    import pirate_amd as pa
    return pa.DedispersionPlan(pa.DedispersionConfig.from_dict(d))


@pytest.mark.parametrize('name', sorted(set(CASES) - LARGE))
def test_closed_form_matches_dense_oracle(name, orc):
    import pirate_amd as pa
    from oracle import avar
    d = CASES[name]
    plan = _plan(d)
    fv = np.random.default_rng(7).uniform(0.3, 3.0, plan.nfreq)
    tv_ref, ptf_ref = avar.pf_variances(orc.OracleConfig.from_yaml_dict(d), fv)
    av = pa.PfAvarApproximation(plan, fv)
    assert av.ntrees == plan.ntrees == len(tv_ref)
    for it in range(plan.ntrees):
        t = plan.trees[it]
        assert av.tree_variance[it].shape == (t.nsubbands_N, t.ndm_wt, t.nprofiles) == tv_ref[it].shape
        np.testing.assert_allclose(av.tree_variance[it], tv_ref[it], rtol=1e-12, atol=0)
        np.testing.assert_allclose(av.per_tf[it], ptf_ref[it], rtol=1e-12, atol=1e-300)
        assert np.all(av.tree_variance[it] > 0)


def test_linear_in_freq_variances():
    import pirate_amd as pa
    plan = _plan(CASES['zones4'])
    rng = np.random.default_rng(3)
    a, b = rng.uniform(0.1, 1, plan.nfreq), rng.uniform(0.1, 1, plan.nfreq)
    va, vb = pa.PfAvarApproximation(plan, a), pa.PfAvarApproximation(plan, b)
    vab = pa.PfAvarApproximation(plan, 2.0 * a + 3.0 * b)
    for it in range(plan.ntrees):
        np.testing.assert_allclose(vab.tree_variance[it], 2.0 * va.tree_variance[it] + 3.0 * vb.tree_variance[it], rtol=1e-13)


def test_dense_tree_levels_pinned(orc):
    from oracle import avar
    R, nt = 5, 48
    x = np.random.default_rng(1).standard_normal((2 ** R, nt))
    # full depth: the reference's closed-form delays (natural row d <-> dm_brev = bitrev(d))
    full = avar.tree_levels(x, R)[0]
    for d in range(2 ** R):
        ref = np.zeros(nt)
        for n in range(2 ** R):
            s = orc.dedispersion_delay(R, n, bitrev(d, R))
            ref[s:] += x[n, :nt - s]
        np.testing.assert_allclose(full[d], ref / 2 ** (R / 2), rtol=1e-12, atol=1e-12)
    # full depth: the oracle's whole-array tree (rows in bit-reversed order)
    y = orc.dedisperse_non_incremental(x.copy())
    np.testing.assert_allclose(full, y[[bitrev(d, R) for d in range(2 ** R)]], rtol=1e-12, atol=1e-12)
    # partial depth k: each band is a rank-k tree over its 2^k channels
    k = 3
    part = avar.tree_levels(x, k)
    assert part.shape == (2 ** (R - k), 2 ** k, nt)
    for F in range(2 ** (R - k)):
        for d in range(2 ** k):
            ref = np.zeros(nt)
            for j in range(2 ** k):
                s = orc.dedispersion_delay(k, j, bitrev(d, k))
                ref[s:] += x[F * 2 ** k + j, :nt - s]
            np.testing.assert_allclose(part[F, d], ref / 2 ** (k / 2), rtol=1e-12, atol=1e-12)


def test_dense_kernels_pinned(orc):
    from oracle import avar
    P = 16
    ks = avar.pf_kernels(P)
    assert len(ks) == P
    np.testing.assert_allclose([np.sum(h * h) for h in ks], orc.bare_profile_variances(P), rtol=1e-15)
    assert len(ks[0]) == 1 and ks[0][0] == 1.0
    for lev in range(5):
        w = 2 ** lev
        q1, q2, q3 = ks[1 + 3 * lev:4 + 3 * lev]
        assert (len(q1), len(q2), len(q3)) == (2 * w, 3 * w, 4 * w)
        assert (q1.sum(), q2.sum(), q3.sum()) == (2 * w, 2 * w, 3 * w)
        assert q1.min() == 1.0 and q2.max() == 1.0 and q3[0] == q3[-1] == 0.5
    rows = np.random.default_rng(2).standard_normal((3, 40))
    e = avar._filtered_sumsq(rows, ks)
    for p, h in enumerate(ks):
        np.testing.assert_allclose(e[:, p], [np.sum(np.convolve(r, h) ** 2) for r in rows], rtol=1e-12)


def test_bad_arguments():
    import pirate_amd as pa
    from pirate_amd._lib import PddError
    plan = _plan(CASES['one_tree'])
    with pytest.raises(ValueError):
        pa.PfAvarApproximation(plan, np.ones(plan.nfreq - 1))
    bad = np.ones(plan.nfreq)
    bad[5] = 0.0
    with pytest.raises(PddError, match='positive'):
        pa.PfAvarApproximation(plan, bad)
    bad[5] = np.nan
    with pytest.raises(PddError, match='positive'):
        pa.PfAvarApproximation(plan, bad)


@pytest.mark.parametrize('name,nchunks', [('toy', 96), ('zones4', 96)])
def test_monte_carlo_unit_variance(name, nchunks, orc):
    """Weighted outputs of the oracle's float64 chain have variance ~1 with analytic weights."""
    import pirate_amd as pa
    d = CASES[name]
    plan = _plan(d)
    rng = np.random.default_rng(11)
    fv = rng.uniform(0.5, 2.0, plan.nfreq)
    av = pa.PfAvarApproximation(plan, fv)
    dd = orc.Dedisperser(orc.OracleConfig.from_yaml_dict(d), 1, dbl=True, enable_variances=True)
    for it in range(plan.ntrees):
        B, D, T, P, N = dd.wt_shape(it)
        w = (1.0 / np.sqrt(av.tree_variance[it])).transpose(1, 2, 0)         # (D, P, N)
        dd.set_weights(it, np.broadcast_to(w[None, :, None], (B, D, T, P, N)).astype(np.float32))
    acc = [0.0] * plan.ntrees
    nuse = 0
    for c in range(nchunks):
        for b in range(dd.nbatches):
            x = rng.standard_normal((dd.B, plan.nfreq, plan.nt_in)) * np.sqrt(fv)[None, :, None]
            dd.dedisperse(x, c, b)
            if c >= 4:                        # past the start-up transient (zero history)
                nuse += 1
                for it in range(plan.ntrees):
                    acc[it] = acc[it] + dd.variances(it).mean(axis=0)
    for it, t in enumerate(dd.trees):
        m = acc[it] / nuse                                                    # (ndm_out, M, P)
        g = m.reshape(t['ndm_wt'], -1, *m.shape[1:]).mean(axis=1)              # per weight row
        print(f'{name} tree {it}: weighted variance {g.min():.3f}..{g.max():.3f}, mean {g.mean():.3f}')
        assert 0.97 < g.mean() < 1.05
        assert 0.88 < g.min() and g.max() < 1.15        # approximation error + Monte Carlo noise
