"""Per-kernel SQ counter summary of a rocprofv3 --pmc run of bench.py (chain names as in
pirate_amd.profiling). Usage: python -m pirate_amd.profiling_sq DIR [--npri 4 --ntrees 4]

Prints the average of every counter per chain kernel and, when the counters are present,
instructions per wave (SQ_INSTS_* / SQ_WAVES) and the share of wave cycles spent waiting
on any instruction (SQ_WAIT_INST_ANY / SQ_WAVE_CYCLES)."""

import argparse
import collections
import csv
import glob
import os
import sys

from .profiling import chain_names, label_dispatches


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('dir')
    ap.add_argument('--npri', type=int, default=4)
    ap.add_argument('--ntrees', type=int, default=4)
    a = ap.parse_args(argv)
    paths = glob.glob(os.path.join(a.dir, '**', '*counter_collection.csv'), recursive=True)
    if not paths:
        raise SystemExit(f'no *counter_collection.csv under {a.dir}')
    rows = list(csv.DictReader(open(paths[0])))
    per = collections.defaultdict(dict)
    kn = {}
    for r in rows:
        d = int(r['Dispatch_Id'])
        per[d][r['Counter_Name']] = per[d].get(r['Counter_Name'], 0.0) + float(r['Counter_Value'])
        kn[d] = r['Kernel_Name']
    order = sorted(per)
    labels = label_dispatches([kn[d] for d in order], chain_names(a.npri, a.ntrees))
    acc = collections.defaultdict(lambda: collections.defaultdict(list))
    for d, lab in zip(order, labels):
        if lab:
            for c, v in per[d].items():
                acc[lab][c].append(v)
    names = sorted({c for d in per.values() for c in d})
    derived = []
    if 'SQ_WAVES' in names:
        derived += [(c.replace('SQ_INSTS_', '') + '/wave', c) for c in names if c.startswith('SQ_INSTS_')]
    if 'SQ_WAIT_INST_ANY' in names and 'SQ_WAVE_CYCLES' in names:
        derived.append(('wait %', None))
    print('| kernel | ' + ' | '.join(names + [d for d, _ in derived]) + ' |')
    print('|---|' + '---|' * (len(names) + len(derived)))
    for lab in chain_names(a.npri, a.ntrees):
        if lab not in acc:
            continue
        avg = {c: sum(acc[lab][c]) / len(acc[lab][c]) for c in names}
        cells = ['%.4g' % avg[c] for c in names]
        for d, c in derived:
            if c is not None:
                cells.append('%.1f' % (avg[c] / avg['SQ_WAVES']))
            else:
                cells.append('%.0f' % (100.0 * avg['SQ_WAIT_INST_ANY'] / avg['SQ_WAVE_CYCLES']))
        print('| %s | ' % lab + ' | '.join(cells) + ' |')


if __name__ == '__main__':
    sys.exit(main())
