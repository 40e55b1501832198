"""Pins the CPU oracle (oracle/oracle.hpp) against closed forms, invariants and brute force.

The reference holds no golden vectors for numerical outputs (its tests are randomized
self-consistency checks), so each oracle component is pinned here by something
independent of its own code path, chosen so that a plausible mistake fails:

  * tree delays          -> closed-form endpoints (delay(0,dm)=dm, delay(N-1,dm)=0),
                            monotonicity, near-linearity; one-hot impulse responses.
  * incremental tree     -> brute-force delay sums and the whole-array tree over the
                            concatenated chunks (reference ReferenceTree::test_basics).
  * subband outputs      -> independent sub-trees per subband (reference test_subbands).
  * lag buffer           -> whole-array shift of the concatenated stream.
  * tree gridding        -> overlap-weight closed form and intensity conservation.
  * peak finder          -> pure-Python brute force of the boxcar/profile definition on
                            exact (integer) data, including the tie-break order.
  * whole dedisperser    -> soph 0 (one-stage tree, plain downsampling) == soph 1
                            (two-stage tree, lagged downsampling, ring lags), bit-exact.
"""

import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def bitrev(i, n):
    j = 0
    for _ in range(n):
        j = (j << 1) | (i & 1)
        i >>= 1
    return j


# ----------------------------------------------------------------------------------------------
# Tree delays


@pytest.mark.parametrize('rank', [1, 2, 3, 5, 7])
def test_delay_closed_forms(orc, rank):
    N = 2**rank
    for dmb in range(N):
        dm = bitrev(dmb, rank)
        d = [orc.dedispersion_delay(rank, f, dmb) for f in range(N)]
        assert d[0] == dm                      # highest frequency: full-band delay = dm
        assert d[N - 1] == 0                   # lowest frequency: zero delay
        assert all(d[i] >= d[i + 1] for i in range(N - 1))   # monotone in frequency
        # Tree delays approximate the linear sweep dm*(N-1-f)/(N-1) to within rank/2 samples.
        lin = dm * (N - 1 - np.arange(N)) / max(N - 1, 1)
        assert np.max(np.abs(np.array(d) - lin)) <= rank / 2 + 1e-9


@pytest.mark.parametrize('rank', [2, 4, 6])
def test_tree_impulse_response(orc, rank):
    # One-hot input at (f, t0): output row dm_brev is c^rank at t0 + delay(f, dm_brev), else 0.
    N, T, t0 = 2**rank, 3 * 2**rank + 8, 2**rank + 3
    c = np.float32(0.7071067811865476)
    for f in [0, 1, N // 2, N - 1]:
        arr = np.zeros((N, T), dtype=np.float32)
        arr[f, t0] = 1.0
        orc.dedisperse_non_incremental(arr)
        for dmb in range(N):
            expect = np.zeros(T, dtype=np.float32)
            expect[t0 + orc.dedispersion_delay(rank, f, dmb)] = 1.0
            np.testing.assert_allclose(arr[dmb], expect * c**rank, rtol=2e-6, atol=0)


@pytest.mark.parametrize('seed', [0, 1, 2])
def test_incremental_tree_vs_brute_force_and_whole_array(orc, seed):
    rng = np.random.default_rng(seed)
    B, amb, dd, T, nchunks = 2, 1, 5, 24, 5
    A, D = 2**amb, 2**dd
    x = rng.standard_normal((B, A, D, nchunks * T)).astype(np.float32)

    # Whole-array (non-incremental) tree.
    whole = x.copy()
    for b in range(B):
        for a in range(A):
            tmp = np.ascontiguousarray(whole[b, a])
            orc.dedisperse_non_incremental(tmp)
            whole[b, a] = tmp

    # Brute-force delay sums, in float64.
    w = 1.0 / np.sqrt(D)
    for (b, a, dmb, t) in [(0, 0, 0, 7), (1, 1, 31, 100), (0, 1, 13, 60), (1, 0, 22, nchunks * T - 1)]:
        s = 0.0
        for f in range(D):
            t0 = t - orc.dedispersion_delay(dd, f, dmb)
            if t0 >= 0:
                s += w * float(x[b, a, f, t0])
        assert abs(s - whole[b, a, dmb, t]) < 1e-4

    # Incremental tree over chunks == whole-array tree (bit-exact: same operations, same order).
    tree = orc.Tree(B, amb, dd, T)
    out = np.zeros_like(x)
    for c in range(nchunks):
        buf = np.ascontiguousarray(x[..., c * T:(c + 1) * T])
        tree.dedisperse(buf)
        out[..., c * T:(c + 1) * T] = buf
    assert np.array_equal(out, whole)


def _subband_lags(fs_M, m_to_fhi, pf_rank, dd_rank, Dpf, B):
    lags = np.zeros((B, Dpf, fs_M), dtype=np.int64)
    for dpf in range(Dpf):
        for m in range(fs_M):
            lags[:, dpf, m] = (2**pf_rank - m_to_fhi[m]) * (dpf & (2**(dd_rank - pf_rank) - 1))
    return lags


def _subbands(sc):
    """Independent restatement of the subband tables (FrequencySubbands.cpp:32-60)."""
    n_flo, n_fhi, n_mbase, m_to_n = [], [], [], []
    M = 0
    for lev, cnt in enumerate(sc):
        sp = 2**max(lev - 1, 0)
        for b in range(cnt):
            n_flo.append(b * sp)
            n_fhi.append(b * sp + 2**lev)
            n_mbase.append(M)
            m_to_n += [len(n_flo) - 1] * 2**lev
            M += 2**lev
    return n_flo, n_fhi, n_mbase, m_to_n, M


@pytest.mark.parametrize('sc,amb,dd', [([1], 1, 3), ([2, 1], 1, 3), ([2, 2, 1], 0, 4), ([4, 3, 1], 2, 3),
                                      ([0, 0, 1], 1, 2), ([3, 5, 2, 1], 1, 4), ([5, 9, 7, 3, 1], 0, 5)])
def test_subbands_vs_independent_subtrees(orc, sc, amb, dd):
    rng = np.random.default_rng(17)
    B, T, nchunks = 2, 16, 4
    pf_rank = len(sc) - 1
    A, Din = 2**amb, 2**dd
    Dpf = 2**(amb + dd - pf_rank)
    n_flo, n_fhi, n_mbase, m_to_n, M = _subbands(sc)
    m_to_fhi = [n_fhi[m_to_n[m]] for m in range(M)]

    tree = orc.Tree(B, amb, dd, T, subband_counts=sc)
    assert tree.M == M
    subtrees = []
    for n in range(len(n_flo)):
        size = (n_fhi[n] - n_flo[n]) << (dd - pf_rank)
        subtrees.append(orc.Tree(B, amb, int(np.log2(size)), T))
    lagbuf = orc.Lagbuf(_subband_lags(M, m_to_fhi, pf_rank, dd, Dpf, B), T)

    for c in range(nchunks):
        x = rng.standard_normal((B, A, Din, T)).astype(np.float32)
        buf = x.copy()
        out = np.zeros((B, Dpf, M, T), dtype=np.float32)
        tree.dedisperse(buf, out)

        out2 = np.zeros_like(out)
        for n, st in enumerate(subtrees):
            flo, fhi = n_flo[n] << (dd - pf_rank), n_fhi[n] << (dd - pf_rank)
            size = fhi - flo
            srank = int(np.log2(size))
            lev = int(np.log2(n_fhi[n] - n_flo[n]))
            sub = np.ascontiguousarray(x[:, :, flo:fhi, :])
            st.dedisperse(sub)
            for a in range(A):
                dm_c = bitrev(a, amb)
                for d in range(size):
                    dm_f = bitrev(d, srank)
                    dpf = (dm_c << (dd - pf_rank)) + (dm_f >> lev)
                    m = n_mbase[n] + (dm_f & (2**lev - 1))
                    out2[:, dpf, m, :] = sub[:, a, d, :]
        lagbuf.apply(out2)
        assert np.array_equal(out, out2), f'chunk {c}'
        # The full band is the last subband (its sub-tree is the main tree).
        assert n_flo[-1] == 0 and n_fhi[-1] == 2**pf_rank


def test_lagbuf_whole_array_shift(orc):
    rng = np.random.default_rng(5)
    nrows, T, nchunks = 23, 7, 9
    lags = rng.integers(0, 3 * T, size=nrows)
    lags[0], lags[1] = 0, T
    lb = orc.Lagbuf(lags, T)
    x = rng.standard_normal((nrows, nchunks * T)).astype(np.float32)
    y = np.zeros_like(x)
    for c in range(nchunks):
        chunk = np.ascontiguousarray(x[:, c * T:(c + 1) * T])
        lb.apply(chunk)
        y[:, c * T:(c + 1) * T] = chunk
    for r in range(nrows):
        expect = np.zeros(nchunks * T, dtype=np.float32)
        L = int(lags[r])
        expect[L:] = x[r, :nchunks * T - L]
        assert np.array_equal(y[r], expect)


# ----------------------------------------------------------------------------------------------
# Tree gridding and channel map


def _cfg(orc, zone_nfreq, edges, R, T=64, B=1, sc=(1,), pts=None):
    if pts is None:
        pts = [dict(num_early_triggers=0, max_width=4, wt_dm_downsampling=4, wt_time_downsampling=4)]
    return orc.OracleConfig(zone_nfreq, edges, R, T, B, B, list(sc), pts)


@pytest.mark.parametrize('zones,edges,R', [([640], [400, 800], 8), ([1024], [400, 800], 10),
                                           ([300, 100, 50], [300, 500, 800, 1500], 7)])
def test_channel_map_closed_form(orc, zones, edges, R):
    cfg = _cfg(orc, zones, edges, R)
    cm = cfg.channel_map()
    flo, fhi = edges[0], edges[-1]
    n = np.arange(2**R + 1, dtype=np.float64)
    freq = 1.0 / np.sqrt(n / (2**R / (flo**-2 - fhi**-2)) + fhi**-2)
    # Piecewise-linear zone mapping freq -> fractional channel.
    idx = np.interp(freq, edges, np.concatenate([[0], np.cumsum(zones)]))
    np.testing.assert_allclose(cm, idx, rtol=0, atol=1e-8 * sum(zones))
    assert cm[0] == pytest.approx(sum(zones)) and cm[-1] == pytest.approx(0.0, abs=1e-9)
    assert np.all(np.diff(cm) < 0)


@pytest.mark.parametrize('zones,edges,R', [([640], [400, 800], 8), ([300, 100, 50], [300, 500, 800, 1500], 7)])
def test_gridding_weights_and_conservation(orc, zones, edges, R):
    cfg = _cfg(orc, zones, edges, R)
    cm = cfg.channel_map()
    F, N, T = sum(zones), 2**R, 5
    # Constant input: each tree channel receives its width in input channels.
    out = orc.tree_gridding(np.ones((1, F, T), dtype=np.float64), cm)
    np.testing.assert_allclose(out[0, :, 0], cm[:-1] - cm[1:], rtol=1e-12, atol=1e-12)
    # One-hot input channel: total weight over tree channels is 1 (intensity conservation).
    for f in [0, 1, F // 3, F - 1]:
        x = np.zeros((1, F, T), dtype=np.float64)
        x[0, f, :] = 1.0
        out = orc.tree_gridding(x, cm)
        assert out[0, :, 0].sum() == pytest.approx(1.0, abs=1e-12)
        nz = np.nonzero(out[0, :, 0])[0]
        # Overlap weights of one input channel go to a contiguous run of tree channels.
        assert np.all(np.diff(nz) == 1)


# ----------------------------------------------------------------------------------------------
# Peak finder: brute force on exact integer data


def pf_brute_force(stream, wt, M, N, m_to_n, Wmax, Dout, Dcore, nt_out, nt_wt, tout0=0):
    """Pure-Python restatement of the peak finder's definition (PeakFindingKernel.hpp:20-120,
    notes/tree_dedispersion.tex kernels). 'stream' is (M, ntot) holding all past samples
    (zeros before time 0); outputs are for the last chunk: tout in [tout0, tout0+nt_out).

    Candidate (level l, multiplet m, isamp, profile p) ends at
        e = tout*Dout + (isamp+1)*dt,    dt = min(Dcore, 2^l),
    with boxcar B(e') = sum(stream[m, e'-2^l : e']) and profiles
        p=0 (l=0 only): B(e);  q=1: B(e-w)+B(e);  q=2: B(e-2w)/2+B(e-w)+B(e)/2;
        q=3: B(e-3w)/2+B(e-2w)+B(e-w)+B(e)/2,  w = 2^l, p = 3l+q.
    The max uses strict '>' in loop order (l, m, isamp, p): earliest wins ties.
    Token = t | (p << 8) | (m << 16), t = isamp*dt.
    """
    P = 1 + 3 * int(np.log2(Wmax))
    nlev = max(int(np.log2(Wmax)), 1)
    ntot = stream.shape[1]

    def box(m, e, w):
        lo, hi = e - w, e
        return float(stream[m, max(lo, 0):max(hi, 0)].sum())

    out_max = np.zeros(nt_out)
    out_arg = np.zeros(nt_out, dtype=np.uint32)
    Tds = nt_out // nt_wt
    for j in range(nt_out):
        tout = tout0 + j
        wrow = wt[j // Tds]
        best, arg = -1e30, 0xffffffff
        for l in range(nlev):
            w = 2**l
            dt = min(Dcore, w)
            for m in range(M):
                n = m_to_n[m]
                for isamp in range(Dout // dt):
                    e = tout * Dout + (isamp + 1) * dt
                    assert e <= ntot
                    b = [box(m, e - k * w, w) for k in (3, 2, 1, 0)]
                    cands = []
                    if l == 0:
                        cands.append((0, b[3]))
                    if P > 1:
                        cands.append((3 * l + 1, b[2] + b[3]))
                        cands.append((3 * l + 2, 0.5 * b[1] + b[2] + 0.5 * b[3]))
                        cands.append((3 * l + 3, 0.5 * b[0] + b[1] + b[2] + 0.5 * b[3]))
                    for p, y in cands:
                        v = wrow[p, n] * y
                        if v > best:
                            best, arg = v, (isamp * dt) | (p << 8) | (m << 16)
        out_max[j], out_arg[j] = best, arg
    return out_max, out_arg


@pytest.mark.parametrize('sc,Wmax,Dout,Dcore', [([1], 4, 4, 4), ([2, 1], 8, 8, 4), ([0, 2, 1], 16, 4, 4),
                                                ([2, 1], 1, 2, 2), ([1], 32, 16, 8), ([1, 1], 2, 2, 1)])
def test_peak_finder_brute_force_exact(orc, sc, Wmax, Dout, Dcore):
    # Integer-valued inputs and power-of-two weights: every sum and product is exact, so
    # the comparison is association-free and exercises ties (integer data ties often).
    rng = np.random.default_rng(Wmax * 100 + Dout)
    n_flo, n_fhi, n_mbase, m_to_n, M = _subbands(sc)
    N = len(n_flo)
    P = 1 + 3 * int(np.log2(Wmax))
    B, D, nt_out, nchunks = 1, 2, 6, 4
    nt_in = nt_out * Dout
    nt_wt = 3
    pf = orc.PeakFinder(sc, Wmax, B, B, D, D, nt_out, nt_in, nt_wt, Dcore)
    hist = np.zeros((D, M, 0), dtype=np.float32)
    for c in range(nchunks):
        x = rng.integers(-4, 5, size=(B, D, M, nt_in)).astype(np.float32)
        wt = (2.0 ** rng.integers(-2, 3, size=(B, D, nt_wt, P, N))).astype(np.float32)
        wt *= (rng.random(wt.shape) < 0.8)     # some zero weights -> +-0 ties
        mx, am = pf.apply(x, wt, 0)
        hist = np.concatenate([hist, x[0]], axis=2)
        for d in range(D):
            emax, earg = pf_brute_force(hist[d], wt[0, d], M, N, m_to_n, Wmax, Dout, Dcore, nt_out, nt_wt,
                                        tout0=c * nt_out)
            np.testing.assert_array_equal(mx[0, d], emax.astype(np.float32))
            np.testing.assert_array_equal(am[0, d], earg)
        ev = pf.eval_tokens(am, wt)
        np.testing.assert_array_equal(ev, mx)


def test_bare_profile_variances(orc):
    # Zero-lag autocorrelation of each profile kernel: sum of squared taps.
    v = orc.bare_profile_variances(16)
    expect = [1.0]
    for l in range(5):
        w = 2**l
        expect += [2.0 * w, w * (0.25 + 1 + 0.25), w * (0.25 + 1 + 1 + 0.25)]
    np.testing.assert_allclose(v, expect[:16])


# ----------------------------------------------------------------------------------------------
# Whole dedisperser


CASES = {
    'toy_like': dict(zones=[640], edges=[400, 800], R=8, T=256, beams=4, bpb=2, sc=[2, 2, 1],
                     pts=[(0, 16, 16, 16), (1, 8, 16, 16), (1, 4, 16, 16)]),
    'one_tree': dict(zones=[1024], edges=[400, 800], R=10, T=256, beams=1, bpb=1, sc=[1],
                     pts=[(0, 16, 64, 64)]),
    'four_trees_zones': dict(zones=[200, 300, 500], edges=[300, 500, 800, 1500], R=9, T=512, beams=2, bpb=1,
                             sc=[3, 3, 1], pts=[(0, 8, 32, 32), (1, 4, 32, 32), (0, 4, 32, 32), (2, 2, 16, 16)]),
}


def make_case(orc, name):
    c = CASES[name]
    pts = [dict(num_early_triggers=a, max_width=b, wt_dm_downsampling=d, wt_time_downsampling=e)
           for (a, b, d, e) in c['pts']]
    return orc.OracleConfig(c['zones'], c['edges'], c['R'], c['T'], c['beams'], c['bpb'], c['sc'], pts)


@pytest.mark.parametrize('name', sorted(CASES))
def test_soph0_equals_soph1_bit_exact(orc, name):
    cfg = make_case(orc, name)
    rng = np.random.default_rng(3)
    d0, d1 = orc.Dedisperser(cfg, 0), orc.Dedisperser(cfg, 1)
    nb = d0.nbatches
    for ichunk in range(3):
        for ib in range(nb):
            x = rng.uniform(-1, 1, size=(cfg.beams_per_batch, cfg.nfreq, cfg.nt_in)).astype(np.float32)
            for it in range(len(d0.trees)):
                w = orc.sparse_random_weights(rng, *d0.wt_shape(it))
                d0.set_weights(it, w)
                d1.set_weights(it, w)
            d0.dedisperse(x, ichunk, ib)
            d1.dedisperse(x, ichunk, ib)
            for it in range(len(d0.trees)):
                m0, a0 = d0.output(it)
                m1, a1 = d1.output(it)
                assert np.array_equal(m0, m1), (ichunk, ib, it)
                assert np.array_equal(a0, a1), (ichunk, ib, it)
                assert np.array_equal(d0.eval_tokens(it, a1), m0)


def test_double_oracle_close_to_float(orc):
    cfg = make_case(orc, 'toy_like')
    rng = np.random.default_rng(9)
    df, dd = orc.Dedisperser(cfg, 1), orc.Dedisperser(cfg, 1, dbl=True)
    for ichunk in range(2):
        for ib in range(df.nbatches):
            x = rng.standard_normal((cfg.beams_per_batch, cfg.nfreq, cfg.nt_in))
            for it in range(len(df.trees)):
                w = orc.bare_weights(*df.wt_shape(it))
                df.set_weights(it, w)
                dd.set_weights(it, w)
            df.dedisperse(x.astype(np.float32), ichunk, ib)
            dd.dedisperse(x, ichunk, ib)
            for it in range(len(df.trees)):
                mf, _ = df.output(it)
                md, _ = dd.output(it)
                np.testing.assert_allclose(mf, md, rtol=2e-5, atol=2e-5)


# ----------------------------------------------------------------------------------------------
# Trees derived by the oracle == the reference's own example plan (golden fixture)


def test_oracle_trees_match_reference_plan(orc):
    with open(os.path.join(HERE, 'golden', 'reference_plans.json')) as f:
        G = json.load(f)
    c = G['configs']['chord_sb2_et']
    cfg = orc.OracleConfig.from_yaml_dict(c, nelts_per_segment=64)
    trees = cfg.trees()
    gtrees = G['plans']['chord_sb2_et']['trees']
    assert len(trees) == len(gtrees) == G['plans']['chord_sb2_et']['ntrees']
    for t, g in zip(trees, gtrees):
        assert t['ipri'] == g['primary_tree_index']
        assert t['et'] == g['early_trigger_level']
        assert t['amb_rank'] == g['amb_rank'] and t['dd_rank'] == g['dd_rank']
        assert t['nt_ds'] == g['nt_ds'] and t['Wmax'] == g['max_width']
        assert t['dm_ds'] == g['dm_downsampling'] and t['time_ds'] == g['time_downsampling']
        assert t['Dcore'] == g['Dcore'] and t['nprofiles'] == g['nprofiles']
        assert t['ndm_out'] == g['ndm_out'] and t['nt_out'] == g['nt_out']
        assert t['ndm_wt'] == g['ndm_wt'] and t['nt_wt'] == g['nt_wt']
        assert t['subband_counts'] == g['frequency_subband_counts']


# ----------------------------------------------------------------------------------------------
# 16-bit storage rounding (the oracle's model of the GPU chain's fp16/bf16 stores)

@pytest.mark.parametrize('storage', ['float16', 'bfloat16'])
def test_storage_rounding_matches_library_casts(storage, orc):
    """Round-to-nearest-even to binary16 / bfloat16 equals numpy's float16 cast and torch's
    bfloat16 cast on normal, subnormal, overflowing, tie and special values."""
    import torch
    rng = np.random.default_rng(0)
    parts = [rng.standard_normal(20000).astype(np.float32) * s for s in (1e-9, 6e-8, 1e-5, 1.0, 1e3, 6e4, 7e4, 1e30)]
    special = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 65504, 65519.99, 65520, 65536, 2.0**-24, 2.0**-25,
                        3 * 2.0**-26, 2.0**-14, 2.0**-14 * (1 - 2.0**-11), 1 + 2.0**-11, 1 + 3 * 2.0**-11,
                        1 + 2.0**-8, 1 + 3 * 2.0**-8], dtype=np.float32)
    x = np.concatenate(parts + [special])
    got = orc.round_to_storage(x, storage)
    with np.errstate(over='ignore'):
        want = (x.astype(np.float16).astype(np.float32) if storage == 'float16'
                else torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy())
    same = (got == want) | (np.isnan(got) & np.isnan(want))
    num = ~np.isnan(want)
    assert np.all(same) and np.array_equal(np.signbit(got[num]), np.signbit(want[num]))   # a nan's sign is free


def test_storage_rounding_oracle_is_close_to_float(orc):
    """The 16-bit-storage oracle's outputs are representable in the storage type and within
    the storage precision of the float oracle's."""
    from gpu_cases import CASES
    d = dict(CASES['toy'])
    cfg = orc.OracleConfig.from_yaml_dict(d, nelts_per_segment=64)
    a = orc.Dedisperser(cfg, 1)
    b = orc.Dedisperser(cfg, 1, storage='float16')
    rng = np.random.default_rng(3)
    w = [orc.bare_weights(*a.wt_shape(it)) for it in range(len(a.trees))]
    for it in range(len(a.trees)):
        a.set_weights(it, w[it])
        b.set_weights(it, w[it])
    for c in range(3):
        for ib in range(a.nbatches):
            x = orc.round_to_storage(rng.uniform(-1, 1, size=(a.B, cfg.nfreq, cfg.nt_in)), 'float16')
            x = x.reshape(a.B, cfg.nfreq, cfg.nt_in)
            a.dedisperse(x, c, ib)
            b.dedisperse(x, c, ib)
            for it in range(len(a.trees)):
                ma, _ = a.output(it)
                mb, _ = b.output(it)
                assert np.array_equal(mb, orc.round_to_storage(mb, 'float16').reshape(mb.shape))
                assert np.max(np.abs(ma - mb)) <= 0.02 * max(1.0, float(np.max(np.abs(ma))))
