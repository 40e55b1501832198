"""Extracts golden plan data from the reference's example plan and config files.

Reads (as data, with yaml.safe_load) the reference's
  configs/dedispersion/chord_sb2_et.yml      (a dedispersion config)
  configs/dedispersion/toy.yml               (a dedispersion config)
  configs/example_dedispersion_plan.yml      (the plan the reference derived from chord_sb2_et.yml)
and writes tests/golden/reference_plans.json: the config fields, and per output tree the
fields the plan derives (ranks, shapes, downsampling factors, subband counts, DM range).

Run from the repository root:  python tests/golden/make_plan_fixture.py /root/reference
"""

import json
import os
import sys

import yaml


def main(refroot):
    def load(rel):
        with open(os.path.join(refroot, rel)) as f:
            return yaml.safe_load(f)

    out = {'configs': {}, 'plans': {}}
    for name in ['chord_sb2_et', 'toy']:
        out['configs'][name] = load(f'configs/dedispersion/{name}.yml')

    plan = load('configs/example_dedispersion_plan.yml')
    keep = ['ndm_out', 'nt_out', 'dm_min', 'dm_max', 'trigger_frequency', 'primary_tree_index',
            'early_trigger_level', 'amb_rank', 'dd_rank', 'nt_ds', 'max_width', 'dm_downsampling',
            'time_downsampling', 'Dcore', 'nprofiles', 'wt_dm_downsampling', 'wt_time_downsampling',
            'ndm_wt', 'nt_wt', 'frequency_subband_counts']
    top = {k: plan[k] for k in ['dtype', 'nfreq', 'nt_in', 'toplevel_tree_rank', 'num_primary_trees',
                                'beams_per_gpu', 'beams_per_batch', 'num_active_batches',
                                'stage1_dd_rank', 'stage1_amb_rank', 'ntrees']}
    top['trees'] = [{k: t[k] for k in keep} for t in plan['trees']]
    out['plans']['chord_sb2_et'] = top

    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'reference_plans.json')
    with open(dst, 'w') as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print('wrote', dst)


if __name__ == '__main__':
    main(sys.argv[1] if len(sys.argv) > 1 else '/root/reference')
