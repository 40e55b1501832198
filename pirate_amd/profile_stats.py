"""Markdown summary of a rocprofv3 --kernel-trace --stats run (its kernel_stats.csv).

    python -m pirate_amd.profile_stats DIR --md OUT.md --title TITLE --command CMD [--command CMD ...]
                                      [--note TEXT ...] [--flop NAME_SUBSTRING=FLOP_PER_CALL ...] [--lib-only]

DIR is a rocprofv3 output directory written with ``--output-format csv -o run``. The table
lists every kernel (name, calls, average / total time, share of the total); with --flop, the
kernels whose name contains the substring also get an achieved TFLOP/s column (flop per call /
average duration). --lib-only keeps this library's kernels (pirate_amd::) and drops those of
torch and its libraries (test-data generation), whose share is then left out of the percentages.
"""

import argparse
import csv
import os


def read_stats(d):
    path = os.path.join(d, 'run_kernel_stats.csv')
    with open(path) as f:
        rows = list(csv.DictReader(f))
    out = []
    for r in rows:
        out.append(dict(name=r['Name'], calls=int(r['Calls']), total_ns=float(r['TotalDurationNs']),
                        avg_ns=float(r['AverageNs']), pct=float(r['Percentage'])))
    return sorted(out, key=lambda r: -r['total_ns'])


def short(name, n=110):
    s = name.replace('void ', '').replace('pirate_amd::', '')
    return s if len(s) <= n else s[:n - 3] + '...'


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split('\n')[0])
    ap.add_argument('dir')
    ap.add_argument('--md', required=True)
    ap.add_argument('--title', default='rocprofv3 kernel statistics')
    ap.add_argument('--command', action='append', default=[])
    ap.add_argument('--note', action='append', default=[])
    ap.add_argument('--flop', action='append', default=[])
    ap.add_argument('--lib-only', action='store_true')
    a = ap.parse_args(argv)
    flop = {}
    for f in a.flop:
        k, v = f.split('=')
        flop[k] = float(v)
    rows = read_stats(a.dir)
    if a.lib_only:
        rows = [r for r in rows if 'pirate_amd::' in r['name']]
        tot = sum(r['total_ns'] for r in rows) or 1.0
        for r in rows:
            r['pct'] = 100.0 * r['total_ns'] / tot
    lines = [f'# {a.title}', '', 'Commands (run from the repository root on the GPU host):', '']
    lines += ['    ' + c for c in a.command] + ['']
    for n in a.note:
        lines += [n, '']
    hdr = '| kernel | calls | avg us | total ms | % |' + (' TFLOP/s |' if flop else '')
    lines += [hdr, '|---|---|---|---|---|' + ('---|' if flop else '')]
    for r in rows:
        extra = ''
        if flop:
            fl = [v for k, v in flop.items() if k in r['name']]
            extra = f' {fl[0] / (r["avg_ns"] * 1e-9) / 1e12:.1f} |' if fl else ' |'
        lines.append(f'| `{short(r["name"])}` | {r["calls"]} | {r["avg_ns"] / 1e3:.1f} | {r["total_ns"] / 1e6:.3f} | '
                     f'{r["pct"]:.1f} |{extra}')
    with open(a.md, 'w') as f:
        f.write('\n'.join(lines) + '\n')


if __name__ == '__main__':
    main()
