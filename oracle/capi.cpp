// ORACLE — test infrastructure only (see oracle.hpp). Plain-C entry points over the oracle
// templates, loaded by tests/ and bench.py through ctypes (oracle/oracle.py).

#include "oracle.hpp"

#include <cstdio>
#include <string>

using namespace oracle;

static thread_local std::string g_err;

extern "C" const char *orc_last_error() { return g_err.c_str(); }

#define ORC_TRY(expr)                                 \
    try { expr; return 0; }                           \
    catch (std::exception &e) { g_err = e.what(); return -1; }

struct orc_cfg {
    long nzones;
    const long *zone_nfreq;
    const double *zone_freq_edges;
    long R, nt_in, beams_per_gpu, beams_per_batch, nelts_per_segment;
    long n_sbc;
    const long *subband_counts;
    long npri;
    const long *pt;          // npri x 6: num_et, max_width, dm_ds, time_ds, wt_dm_ds, wt_time_ds
    long reference_dcore;
};

static OrcConfig to_cfg(const orc_cfg *c)
{
    OrcConfig r;
    r.zone_nfreq.assign(c->zone_nfreq, c->zone_nfreq + c->nzones);
    r.zone_freq_edges.assign(c->zone_freq_edges, c->zone_freq_edges + c->nzones + 1);
    r.R = c->R;
    r.nt_in = c->nt_in;
    r.beams_per_gpu = c->beams_per_gpu;
    r.beams_per_batch = c->beams_per_batch;
    r.nelts_per_segment = c->nelts_per_segment;
    r.subband_counts.assign(c->subband_counts, c->subband_counts + c->n_sbc);
    for (long i = 0; i < c->npri; i++) {
        const long *p = c->pt + 6 * i;
        r.pts.push_back({p[0], p[1], p[2], p[3], p[4], p[5]});
    }
    r.reference_dcore = (c->reference_dcore != 0);
    return r;
}


// ---- config-derived helpers ----

extern "C" int orc_channel_map(const orc_cfg *c, double *out)
{
    ORC_TRY({
        auto m = to_cfg(c).channel_map();
        std::copy(m.begin(), m.end(), out);
    });
}

// Per tree, 17 longs: ipri, et, amb_rank, dd_rank, nt_ds, M, N, Wmax, dm_ds, time_ds,
// wt_dm_ds, wt_time_ds, Dcore, nprofiles, ndm_out, ndm_wt, nt_out, nt_wt (18 total).
extern "C" long orc_trees(const orc_cfg *c, long *out, long maxtrees)
{
    try {
        auto trees = to_cfg(c).make_trees();
        long n = 0;
        for (auto &t : trees) {
            if (n >= maxtrees) break;
            long v[18] = {t.ipri, t.et, t.amb_rank, t.dd_rank, t.nt_ds, t.M, t.N, t.Wmax, t.dm_ds, t.time_ds,
                          t.wt_dm_ds, t.wt_time_ds, t.Dcore, t.nprofiles, t.ndm_out, t.ndm_wt, t.nt_out, t.nt_wt};
            std::copy(v, v + 18, out + 18 * n);
            n++;
        }
        return long(trees.size());
    }
    catch (std::exception &e) { g_err = e.what(); return -1; }
}

extern "C" long orc_tree_subband_counts(const orc_cfg *c, long itree, long *out)
{
    try {
        auto trees = to_cfg(c).make_trees();
        auto &sc = trees.at(itree).subband_counts;
        std::copy(sc.begin(), sc.end(), out);
        return long(sc.size());
    }
    catch (std::exception &e) { g_err = e.what(); return -1; }
}

extern "C" long orc_dedispersion_delay(int rank, long f, long dmb) { return dedispersion_delay(rank, f, dmb); }
extern "C" long orc_rb_lag(long fc, long dmb, int s1, int s2, int u) { return rb_lag(fc, dmb, s1, s2, u != 0); }

extern "C" int orc_bare_profile_variances(long P, double *out)
{
    ORC_TRY({ auto v = bare_profile_variances(P); std::copy(v.begin(), v.end(), out); });
}


// ---- stateless kernels ----

extern "C" int orc_dedisperse_non_incremental(void *arr, long nfreq, long ninner, long nspec, int dbl)
{
    ORC_TRY({
        if (dbl) dedisperse_non_incremental((double *) arr, nfreq, ninner, nspec);
        else dedisperse_non_incremental((float *) arr, nfreq, ninner, nspec);
    });
}

extern "C" int orc_tree_gridding(void *out, const void *in, const double *cmap, long B, long nchan, long nfreq, long ntime, int dbl)
{
    ORC_TRY({
        if (dbl) tree_gridding((double *) out, (const double *) in, cmap, B, nchan, nfreq, ntime);
        else tree_gridding((float *) out, (const float *) in, cmap, B, nchan, nfreq, ntime);
    });
}


// ---- stateful objects behind an opaque handle ----

struct Handle {
    int kind = 0;   // 1 = tree, 2 = lagbuf, 3 = lds, 4 = pf, 5 = dedisperser
    int dbl = 0;
    std::shared_ptr<void> f, d;   // float / double instance
};

template<typename T> static T *get(Handle *h) { return (T *) (h->dbl ? h->d.get() : h->f.get()); }

extern "C" void orc_destroy(void *h) { delete (Handle *) h; }

extern "C" void *orc_tree_create(long B, long amb, long dd, long nt, long ns, const long *sbc, long nsbc, int dbl)
{
    try {
        std::vector<long> sc(sbc, sbc + nsbc);
        auto *h = new Handle;
        h->kind = 1; h->dbl = dbl;
        if (dbl) h->d = std::make_shared<RefTree<double>>(B, amb, dd, nt, ns, sc);
        else h->f = std::make_shared<RefTree<float>>(B, amb, dd, nt, ns, sc);
        return h;
    }
    catch (std::exception &e) { g_err = e.what(); return nullptr; }
}

extern "C" int orc_tree_dedisperse(void *hv, void *buf, long bsb, long bsa, long bsd, void *out, long osb, long osd, long osm)
{
    Handle *h = (Handle *) hv;
    ORC_TRY({
        ORC_CHECK(h && h->kind == 1, "bad handle");
        if (h->dbl) get<RefTree<double>>(h)->dedisperse((double *) buf, bsb, bsa, bsd, (double *) out, osb, osd, osm);
        else get<RefTree<float>>(h)->dedisperse((float *) buf, bsb, bsa, bsd, (float *) out, osb, osd, osm);
    });
}

extern "C" void *orc_lagbuf_create(const long *lags, long n, long ntime, int dbl)
{
    try {
        std::vector<long> l(lags, lags + n);
        auto *h = new Handle;
        h->kind = 2; h->dbl = dbl;
        if (dbl) h->d = std::make_shared<RefLagbuf<double>>(l, ntime);
        else h->f = std::make_shared<RefLagbuf<float>>(l, ntime);
        return h;
    }
    catch (std::exception &e) { g_err = e.what(); return nullptr; }
}

extern "C" int orc_lagbuf_apply(void *hv, void *data)
{
    Handle *h = (Handle *) hv;
    ORC_TRY({
        ORC_CHECK(h && h->kind == 2, "bad handle");
        if (h->dbl) { auto *lb = get<RefLagbuf<double>>(h); long nt = lb->ntime; lb->apply((double *) data, [&](long r) { return r * nt; }); }
        else { auto *lb = get<RefLagbuf<float>>(h); long nt = lb->ntime; lb->apply((float *) data, [&](long r) { return r * nt; }); }
    });
}

extern "C" void *orc_lds_create(long B, long R, long odr, long npri, long ntime, int dbl)
{
    try {
        auto *h = new Handle;
        h->kind = 3; h->dbl = dbl;
        if (dbl) h->d = std::make_shared<RefLDS<double>>(B, R, odr, npri, ntime);
        else h->f = std::make_shared<RefLDS<float>>(B, R, odr, npri, ntime);
        return h;
    }
    catch (std::exception &e) { g_err = e.what(); return nullptr; }
}

extern "C" int orc_lds_apply(void *hv, const void *in, void *const *outs)
{
    Handle *h = (Handle *) hv;
    ORC_TRY({
        ORC_CHECK(h && h->kind == 3, "bad handle");
        if (h->dbl) get<RefLDS<double>>(h)->apply((const double *) in, (double *const *) outs);
        else get<RefLDS<float>>(h)->apply((const float *) in, (float *const *) outs);
    });
}

extern "C" void *orc_pf_create(const long *sbc, long nsbc, long Wmax, long B, long total_beams, long ndm_out, long ndm_wt,
                               long nt_out, long nt_in, long nt_wt, long Dcore, int dbl)
{
    try {
        std::vector<long> sc(sbc, sbc + nsbc);
        auto *h = new Handle;
        h->kind = 4; h->dbl = dbl;
        if (dbl) h->d = std::make_shared<RefPF<double>>(sc, Wmax, B, total_beams, ndm_out, ndm_wt, nt_out, nt_in, nt_wt, Dcore);
        else h->f = std::make_shared<RefPF<float>>(sc, Wmax, B, total_beams, ndm_out, ndm_wt, nt_out, nt_in, nt_wt, Dcore);
        return h;
    }
    catch (std::exception &e) { g_err = e.what(); return nullptr; }
}

extern "C" int orc_pf_apply(void *hv, void *out_max, uint32_t *out_argmax, double *out_var, const void *in, const float *wt, long ibatch)
{
    Handle *h = (Handle *) hv;
    ORC_TRY({
        ORC_CHECK(h && h->kind == 4, "bad handle");
        if (h->dbl) get<RefPF<double>>(h)->apply((double *) out_max, out_argmax, out_var, (const double *) in, wt, ibatch);
        else get<RefPF<float>>(h)->apply((float *) out_max, out_argmax, out_var, (const float *) in, wt, ibatch);
    });
}

extern "C" int orc_pf_eval_tokens(void *hv, void *out, const uint32_t *tokens, const float *wt)
{
    Handle *h = (Handle *) hv;
    ORC_TRY({
        ORC_CHECK(h && h->kind == 4, "bad handle");
        if (h->dbl) get<RefPF<double>>(h)->eval_tokens((double *) out, tokens, wt);
        else get<RefPF<float>>(h)->eval_tokens((float *) out, tokens, wt);
    });
}

extern "C" void *orc_dd_create(const orc_cfg *c, int soph, int dbl, int enable_var, int tree_domain_input, int storage)
{
    try {
        OrcConfig cfg = to_cfg(c);
        if (storage != 0 && dbl)
            throw std::runtime_error("orc_dd_create: 16-bit storage rounding needs the float oracle");
        auto *h = new Handle;
        h->kind = 5; h->dbl = dbl;
        if (dbl) h->d = std::make_shared<RefDedisperser<double>>(cfg, soph, enable_var != 0, tree_domain_input != 0);
        else h->f = std::make_shared<RefDedisperser<float>>(cfg, soph, enable_var != 0, tree_domain_input != 0, storage);
        return h;
    }
    catch (std::exception &e) { g_err = e.what(); return nullptr; }
}

template<typename T>
static void dd_set_input(RefDedisperser<T> *d, const void *data)
{
    std::memcpy(d->input.data(), data, d->input.size() * sizeof(T));
}

extern "C" int orc_dd_set_input(void *hv, const void *data)
{
    Handle *h = (Handle *) hv;
    ORC_TRY({
        ORC_CHECK(h && h->kind == 5, "bad handle");
        if (h->dbl) dd_set_input(get<RefDedisperser<double>>(h), data);
        else dd_set_input(get<RefDedisperser<float>>(h), data);
    });
}

template<typename T>
static void dd_set_weights(RefDedisperser<T> *d, long itree, const float *w)
{
    auto &v = d->wt.at(itree);
    std::memcpy(v.data(), w, v.size() * sizeof(float));
}

extern "C" int orc_dd_set_weights(void *hv, long itree, const float *w)
{
    Handle *h = (Handle *) hv;
    ORC_TRY({
        ORC_CHECK(h && h->kind == 5, "bad handle");
        if (h->dbl) dd_set_weights(get<RefDedisperser<double>>(h), itree, w);
        else dd_set_weights(get<RefDedisperser<float>>(h), itree, w);
    });
}

extern "C" int orc_dd_dedisperse(void *hv, long ichunk, long ibatch)
{
    Handle *h = (Handle *) hv;
    ORC_TRY({
        ORC_CHECK(h && h->kind == 5, "bad handle");
        if (h->dbl) get<RefDedisperser<double>>(h)->dedisperse(ichunk, ibatch);
        else get<RefDedisperser<float>>(h)->dedisperse(ichunk, ibatch);
    });
}

template<typename T>
static void dd_get(RefDedisperser<T> *d, long it, void *mx, uint32_t *am, double *var)
{
    auto &m = d->out_max.at(it);
    if (mx) std::memcpy(mx, m.data(), m.size() * sizeof(T));
    if (am) std::memcpy(am, d->out_argmax.at(it).data(), m.size() * sizeof(uint32_t));
    if (var) {
        auto &v = d->out_var.at(it);
        ORC_CHECK(!v.empty(), "variances not enabled");
        std::memcpy(var, v.data(), v.size() * sizeof(double));
    }
}

extern "C" int orc_dd_get_output(void *hv, long itree, void *out_max, uint32_t *out_argmax, double *out_var)
{
    Handle *h = (Handle *) hv;
    ORC_TRY({
        ORC_CHECK(h && h->kind == 5, "bad handle");
        if (h->dbl) dd_get(get<RefDedisperser<double>>(h), itree, out_max, out_argmax, out_var);
        else dd_get(get<RefDedisperser<float>>(h), itree, out_max, out_argmax, out_var);
    });
}

extern "C" int orc_dd_eval_tokens(void *hv, long itree, const uint32_t *tokens, void *out)
{
    Handle *h = (Handle *) hv;
    ORC_TRY({
        ORC_CHECK(h && h->kind == 5, "bad handle");
        if (h->dbl) { auto *d = get<RefDedisperser<double>>(h); d->pf.at(itree)->eval_tokens((double *) out, tokens, d->wt.at(itree).data()); }
        else { auto *d = get<RefDedisperser<float>>(h); d->pf.at(itree)->eval_tokens((float *) out, tokens, d->wt.at(itree).data()); }
    });
}

extern "C" int orc_dequantize(const uint8_t *data, const uint16_t *scoff, long nbeams, long nfreq, long ntime, float *out)
{
    if (ntime % 256) {
        g_err = "orc_dequantize: ntime must be a multiple of 256";
        return 1;
    }
    ref_dequantize(data, scoff, nbeams, nfreq, ntime, out);
    return 0;
}

extern "C" int orc_ringbuf_copy(uint8_t *ringbuf, long nbytes, const uint32_t *oct, long noct, long seg_bytes,
                                long nbeams, unsigned long long iframe)
{
    for (long i = 0; i < noct; i++)
        for (int side = 0; side < 2; side++) {
            const uint32_t *q = oct + 8 * i + 4 * side;
            if (q[2] == 0 || (long(q[0]) + long(q[2] - 1) * long(q[3]) + 1) * seg_bytes > nbytes) {
                g_err = "orc_ringbuf_copy: octuple addresses a segment outside the buffer";
                return 1;
            }
        }
    ref_ringbuf_copy(ringbuf, oct, noct, seg_bytes, nbeams, iframe);
    return 0;
}

extern "C" int orc_round_storage(float *x, long n, int storage)
{
    round_to_storage(x, n, storage);
    return 0;
}
