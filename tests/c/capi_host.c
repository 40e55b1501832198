/* A plain-C caller of the host side of include/pirate_amd.h (no GPU calls): parse a YAML
 * config, build a plan, read its trees, decode a token, compute analytic variances.
 * Prints one line per check; exit status 0 iff all checks pass. */

#include <stdio.h>
#include <stdlib.h>
#include "pirate_amd.h"

static const char *yaml =
    "zone_nfreq: [1024]\n"
    "zone_freq_edges: [400, 800]\n"
    "time_sample_ms: 1.0\n"
    "toplevel_tree_rank: 10\n"
    "time_samples_per_chunk: 2048\n"
    "dtype: float32\n"
    "frequency_subband_counts: [1]\n"
    "beams_per_gpu: 1\n"
    "beams_per_batch: 1\n"
    "num_active_batches: 1\n"
    "future_write_max_samples: 0\n"
    "primary_trees:\n"
    "  - {num_early_triggers: 0, max_width: 16, dm_downsampling: 0, time_downsampling: 0,\n"
    "     wt_dm_downsampling: 64, wt_time_downsampling: 64}\n";

#define CHECK(cond, what)                                                         \
    do {                                                                          \
        if (!(cond)) {                                                            \
            printf("FAIL %s: %s\n", what, pdd_last_error());                      \
            return 1;                                                             \
        }                                                                         \
        printf("ok %s\n", what);                                                  \
    } while (0)

int main(void)
{
    pdd_config *cfg = NULL;
    pdd_plan *plan = NULL;
    int64_t ntrees = 0, out5[5];
    pdd_tree_info info;
    double *fv, *tv;
    long i, n;

    CHECK(pdd_abi_version() == 4, "abi version");
    CHECK(pdd_config_from_yaml_string(yaml, &cfg) == 0, "config from yaml");
    CHECK(pdd_plan_create(cfg, 1, &plan) == 0, "plan create");
    CHECK(pdd_plan_ntrees(plan, &ntrees) == 0 && ntrees == 1, "one tree");
    CHECK(pdd_plan_tree_info(plan, 0, &info) == 0 && info.ndm_out * info.nt_out > 0, "tree info");
    CHECK(pdd_plan_decode_argmax(plan, 0u, 0, 0, 1, out5) == 0 && out5[4] == 0, "decode token 0");
    CHECK(pdd_plan_decode_argmax(plan, 0xffffffffu, 0, 0, 0, out5) != 0, "bad token rejected");

    fv = (double *) malloc(1024 * sizeof(double));
    for (i = 0; i < 1024; i++)
        fv[i] = 1.0;
    n = info.nsubbands_N * info.ndm_wt * info.nprofiles;
    tv = (double *) malloc(n * sizeof(double));
    CHECK(pdd_plan_pf_variances(plan, fv, 1024, 0, tv, n, NULL, 0) == 0, "analytic variances");
    for (i = 0; i < n; i++)
        if (!(tv[i] > 0.0))
            break;
    CHECK(i == n, "variances positive");

    free(fv);
    free(tv);
    pdd_plan_destroy(plan);
    pdd_config_destroy(cfg);
    return 0;
}
