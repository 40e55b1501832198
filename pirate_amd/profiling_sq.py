"""Per-kernel SQ counter summary of a rocprofv3 --pmc run of bench.py (chain names as in
pirate_amd.profiling). Usage: python -m pirate_amd.profiling_sq DIR [--npri 4 --ntrees 4]"""

import argparse
import collections
import csv
import os
import sys

from .profiling import chain_names, label_dispatches


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('dir')
    ap.add_argument('--npri', type=int, default=4)
    ap.add_argument('--ntrees', type=int, default=4)
    a = ap.parse_args(argv)
    rows = list(csv.DictReader(open(os.path.join(a.dir, 'run_counter_collection.csv'))))
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
    print('| kernel | ' + ' | '.join(names) + ' |')
    print('|---|' + '---|' * len(names))
    for lab in chain_names(a.npri, a.ntrees):
        if lab in acc:
            print('| %s | ' % lab + ' | '.join('%.4g' % (sum(acc[lab][c]) / len(acc[lab][c])) for c in names) + ' |')


if __name__ == '__main__':
    sys.exit(main())
