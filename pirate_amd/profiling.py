"""Summaries of rocprofv3 runs of the dedispersion chain.

Usage (after profiling bench.py, see --help):

    python -m pirate_amd.profiling --stats DIR [--fetch DIR --write DIR] --md OUT.md --json OUT.json

DIR is a rocprofv3 output directory written with ``--output-format csv -o run``. Kernel
dispatches of this library are named by their place in the chain, in host launch order
(dispatch ids): grid_lds, then per primary tree p its stage1_p<p> followed by stage2_t<i>,
pf_t<i> of each tree i of p. One chain per step, so the i-th dispatch of the library within
a step has a fixed name, whichever stream it ran on. HBM counters are scaled by what the
same counter reads on a reference copy of known size (the 1 GiB copy that bench.py makes
with --calib-copy): bytes = counter * (2^30 / counter_of_copy).
"""

import argparse
import collections
import csv
import json
import os
import sys

LIB_PREFIX = 'pirate_amd::'


def _rows(path):
    with open(path) as f:
        return list(csv.DictReader(f))


# Optional substring that a library dispatch's kernel name must also contain (e.g. '__half'
# to keep the float16 chain of a run that holds float32 and float16 chains), set by --only.
_ONLY = ''


def _is_lib(name):
    return LIB_PREFIX in name and _ONLY in name


def _short(name):
    s = name.replace('void ', '').replace(LIB_PREFIX, '')
    return s.split('(')[0]


def chain_names(npri, ntrees, tree_primary=None):
    """Chain kernel names in launch order; tree_primary[i] = primary tree of tree i (default:
    tree i of primary tree i, which needs ntrees == npri)."""
    if tree_primary is None:
        if ntrees != npri:
            raise ValueError('give tree_primary when trees and primary trees differ in number')
        tree_primary = list(range(ntrees))
    names = ['grid_lds']
    for p in range(npri):
        names.append('stage1_p%d' % p)
        for i in range(ntrees):
            if tree_primary[i] == p:
                names += ['stage2_t%d' % i, 'pf_t%d' % i]
    return names


def _launch_key(r):
    for k in ('Dispatch_Id', 'Correlation_Id'):
        if r.get(k) not in (None, ''):
            return int(r[k])
    return int(r['Start_Timestamp'])


def label_dispatches(kernel_names, names):
    """Labels library dispatches (in launch order) with their chain names."""
    out = []
    k = 0
    for kn in kernel_names:
        if _is_lib(kn):
            out.append(names[k % len(names)])
            k += 1
        else:
            out.append(None)
    if k % len(names):
        raise RuntimeError('dispatch count %d is not a whole number of chains of %d' % (k, len(names)))
    return out


def kernel_times(stats_dir, names):
    tr = _rows(os.path.join(stats_dir, 'run_kernel_trace.csv'))
    tr.sort(key=_launch_key)
    labels = label_dispatches([r['Kernel_Name'] for r in tr], names)
    acc = collections.defaultdict(list)
    kname = {}
    for r, lab in zip(tr, labels):
        if lab is None:
            continue
        acc[lab].append((int(r['End_Timestamp']) - int(r['Start_Timestamp'])) * 1e-6)   # ms
        grid = 'x'.join(r.get('Grid_Size_' + c, '?') for c in 'XY')
        kname[lab] = '%s grid=%s wg=%s vgpr=%s' % (_short(r['Kernel_Name']), grid, r.get('Workgroup_Size_X', '?'),
                                                   r.get('VGPR_Count', '?'))
    return {lab: dict(n=len(v), avg_ms=sum(v) / len(v), kernel=kname[lab]) for lab, v in acc.items()}


def counter_bytes(pmc_dir, counter, names):
    rows = _rows(os.path.join(pmc_dir, 'run_counter_collection.csv'))
    per = collections.defaultdict(float)       # dispatch id -> counter sum
    kn = {}
    for r in rows:
        if r['Counter_Name'] != counter:
            continue
        d = int(r['Dispatch_Id'])
        per[d] += float(r['Counter_Value'])
        kn[d] = r['Kernel_Name']
    order = sorted(per)
    calib = [per[d] for d in order if 'copyBuffer' in kn[d] or 'copy_kernel' in kn[d]]
    if not calib:
        raise RuntimeError('%s: no reference copy dispatch found (run bench.py with --calib-copy)' % pmc_dir)
    scale = float(1 << 30) / max(calib)
    labels = label_dispatches([kn[d] for d in order], names)
    acc = collections.defaultdict(list)
    for d, lab in zip(order, labels):
        if lab is not None:
            acc[lab].append(per[d] * scale)
    return scale, max(calib), {lab: sum(v) / len(v) for lab, v in acc.items()}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument('--stats', required=True)
    ap.add_argument('--fetch')
    ap.add_argument('--write')
    ap.add_argument('--npri', type=int, default=4)
    ap.add_argument('--ntrees', type=int, default=4)
    ap.add_argument('--tree-primary', help='comma-separated primary tree of each tree (default: i -> i)')
    ap.add_argument('--note', action='append', default=[], help='paragraph to add under the commands')
    ap.add_argument('--alg-json', help='bench.py JSON line (per-kernel algorithmic bytes)')
    ap.add_argument('--title', default='rocprofv3 summary')
    ap.add_argument('--command', action='append', default=[], help='command lines to record')
    ap.add_argument('--md', required=True)
    ap.add_argument('--json', required=True)
    ap.add_argument('--only', default='', help='keep the library dispatches whose kernel name contains this '
                                                '(e.g. __half: the float16 chain of a run with both storage types)')
    ap.add_argument('--alg-key', default=None, help='JSON path of the per-kernel table in --alg-json (e.g. fp16)')
    a = ap.parse_args(argv)
    global _ONLY
    _ONLY = a.only

    tp = [int(x) for x in a.tree_primary.split(',')] if a.tree_primary else None
    names = chain_names(a.npri, a.ntrees, tp)
    times = kernel_times(a.stats, names)
    alg = {}
    if a.alg_json:
        with open(a.alg_json) as f:
            for line in f:
                line = line.strip()
                if line.startswith('{'):
                    rec = json.loads(line)
                    cfg = rec[a.alg_key] if a.alg_key else rec['config']
                    pk = cfg.get('per_kernel_serialized', cfg.get('per_kernel', {}))
                    alg = {k: v['alg_bytes'] for k, v in pk.items()}
    fetch = write = None
    if a.fetch:
        fscale, fcal, fetch = counter_bytes(a.fetch, 'FETCH_SIZE', names)
    if a.write:
        wscale, wcal, write = counter_bytes(a.write, 'WRITE_SIZE', names)

    out = {'kernels': {}}
    lines = ['# ' + a.title, '', 'Commands (run from the repository root on the GPU host):', '']
    lines += ['    ' + c for c in a.command]
    lines += ['']
    for n in a.note:
        lines += [n, '']
    if fetch is not None:
        out['fetch_scale'] = fscale
        out['write_scale'] = wscale
        lines += ['HBM counters are scaled by a 1 GiB device copy profiled in the same process:',
                  'FETCH_SIZE of the copy = %.0f KiB (scale %.4f bytes/KiB-unit), '
                  'WRITE_SIZE of the copy = %.0f KiB (scale %.4f).' % (fcal, fscale, wcal, wscale), '']
    hdr = '| kernel | dispatch (name, grid, workgroup) | n | avg ms | alg GB | alg GB/s |'
    sep = '|---|---|---|---|---|---|'
    if fetch is not None:
        hdr += ' HBM read GB | HBM write GB | HBM GB/s |'
        sep += '---|---|---|'
    lines += [hdr, sep]
    for lab in names:
        if lab not in times:
            continue
        t = times[lab]
        rec = dict(avg_ms=t['avg_ms'], n=t['n'], kernel=t['kernel'])
        row = '| %s | `%s` | %d | %.4f |' % (lab, t['kernel'], t['n'], t['avg_ms'])
        if lab in alg:
            rec['alg_bytes'] = alg[lab]
            row += ' %.3f | %.0f |' % (alg[lab] / 1e9, alg[lab] / (t['avg_ms'] * 1e-3) / 1e9)
        else:
            row += ' | |'
        if fetch is not None:
            rb, wb = fetch.get(lab, 0.0), write.get(lab, 0.0)
            rec['hbm_read_bytes'], rec['hbm_write_bytes'] = rb, wb
            row += ' %.3f | %.3f | %.0f |' % (rb / 1e9, wb / 1e9, (rb + wb) / (t['avg_ms'] * 1e-3) / 1e9)
        out['kernels'][lab] = rec
        lines.append(row)
    with open(a.md, 'w') as f:
        f.write('\n'.join(lines) + '\n')
    with open(a.json, 'w') as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print('\n'.join(lines))


if __name__ == '__main__':
    sys.exit(main())
