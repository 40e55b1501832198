"""The rocprof summary names each dispatch by its place in the chain's host launch order
(dispatch ids), not by start time: primary trees run on streams of their own, so their
kernels start out of launch order."""

import csv
import os

from pirate_amd.profiling import chain_names, kernel_times


def test_chain_names_follow_launch_order():
    assert chain_names(2, 2) == ['grid_lds', 'stage1_p0', 'stage2_t0', 'pf_t0', 'stage1_p1', 'stage2_t1', 'pf_t1']
    # early-trigger trees: several trees per primary tree, in tree order within it
    assert chain_names(2, 3, [0, 0, 1]) == ['grid_lds', 'stage1_p0', 'stage2_t0', 'pf_t0', 'stage2_t1', 'pf_t1',
                                             'stage1_p1', 'stage2_t2', 'pf_t2']


def test_dispatches_labelled_by_dispatch_id(tmp_path):
    names = chain_names(2, 2)
    # one chain: dispatch ids in launch order; start times put primary tree 1's stage 1
    # before tree 0's stage 2 (overlap on another stream), plus a non-library copy kernel
    kernels = ['k_grid_lds', 'k_stage1_fast_a', 'k_stage2_fast_a', 'k_pf_a', 'k_stage1_fast_b', 'k_stage2_fast_b', 'k_pf_b']
    starts = [0, 10, 30, 50, 20, 40, 60]
    rows = [dict(Dispatch_Id=1, Kernel_Name='copyBuffer', Start_Timestamp=-5, End_Timestamp=-1)]
    for i, (k, t0) in enumerate(zip(kernels, starts)):
        rows.append(dict(Dispatch_Id=i + 2, Kernel_Name='void pirate_amd::%s(args)' % k, Start_Timestamp=t0,
                         End_Timestamp=t0 + 1000000 * (i + 1)))
    d = tmp_path / 'stats'
    os.makedirs(d)
    with open(d / 'run_kernel_trace.csv', 'w', newline='') as f:
        w = csv.DictWriter(f, fieldnames=list(rows[0]))
        w.writeheader()
        w.writerows(rows)
    t = kernel_times(str(d), names)
    for i, name in enumerate(names):
        assert t[name]['kernel'].startswith(kernels[i]), (name, t[name])
        assert abs(t[name]['avg_ms'] - (i + 1)) < 1e-9
