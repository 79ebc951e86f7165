"""Summaries of rocprofv3 results for the kernels of this library.

``python -m opensfm_amd.profiling DIR`` turns every ``*_results.db`` under
DIR (the SQLite output of ``rocprofv3 -d DIR -o NAME ...``) into a text
table next to it (``*_summary.txt``): per kernel, the dispatch count and
the average duration, and with ``--pmc`` the average of each counter per
dispatch (the maximum for the runtime's copy kernel, where the 1 GiB
calibration copy of ``bench.py --copy-probe-mb 1024`` is the largest
dispatch); without counters, the total time and its share.  A kernel
launched with several grid sizes (batches of different reference counts,
the autotuner's trial launches) gets one more row per grid size below the
table, so that the launches of one configuration can be compared on their
own.  ``--delete`` removes the databases afterwards.
"""
from __future__ import annotations

import collections
import glob
import os
import re
import sqlite3
import sys


_TYPES = {"f": "float", "d": "double", "i": "int", "j": "unsigned int", "DF16_": "_Float16"}


def _demangle(name: str) -> str:
    """Readable name of a mangled kernel symbol (rocprofv3 leaves kernels with
    _Float16 template arguments mangled): the last component of the nested
    name with its template arguments (int literals and the scalar types
    above), or the symbol unchanged when it has another form."""
    if not name.startswith("_ZN"):
        return name
    pos, last = 3, None
    while pos < len(name) and name[pos].isdigit():
        n = re.match(r"\d+", name[pos:]).group(0)
        pos += len(n)
        last = name[pos:pos + int(n)]
        pos += int(n)
    if last is None:
        return name
    args = []
    if name.startswith("I", pos):
        pos += 1
        while pos < len(name) and name[pos] != "E":
            lit = re.match(r"Li(-?\d+)E", name[pos:])
            if lit:
                args.append(lit.group(1))
                pos += len(lit.group(0))
                continue
            for code, tname in _TYPES.items():
                if name.startswith(code, pos):
                    args.append(tname)
                    pos += len(code)
                    break
            else:
                return name
    targs = f"<{', '.join(args)}>" if args else ""
    return f"pmhip::{last}{targs}("


def kernel_name(full: str) -> str:
    """Template-qualified name without the argument list."""
    full = _demangle(full)
    m = re.search(r"::((?:\w+)(?:<[^()]*>)?)\(", full)
    name = m.group(1) if m else full
    return name[:80]


def _grid_expr(c, table: str) -> str:
    """SQL expression of the grid size (work-items) of a row of ``table``, or
    '0' when the table has no grid columns."""
    cols = [r[1] for r in c.execute(f"pragma table_info({table})")]
    gcols = [x for x in cols if re.fullmatch(r"grid(_size)?_?[xyz]", x, re.I)]
    if not gcols:
        gcols = [x for x in cols if x.lower() in ("grid_size", "grid")]
    return " * ".join(gcols) if gcols else "0"


def _row(k: str, n: int, dur_ms: float, kind: str, vals: dict) -> str:
    return (f"{k:80s} disp {n:6d} avg_ms {dur_ms:10.3f} {kind} " +
            " ".join(f"{cn} {vals[cn]:.6g}" for cn in sorted(vals)))


def summarize(db: str) -> str:
    c = sqlite3.connect(db)
    lines = [f"# {os.path.basename(db)}"]
    have_pmc = c.execute("select count(*) from counters_collection").fetchone()[0] > 0
    by_grid = collections.OrderedDict()
    if have_pmc:
        per = collections.OrderedDict()
        grid = _grid_expr(c, "counters_collection")
        for did, kn, cn, v, du, g in c.execute(
                "select dispatch_id, kernel_name, counter_name, value, duration, "
                f"{grid} from counters_collection"):
            k = kernel_name(kn)
            per.setdefault(k, {}).setdefault(did, {"dur": du})[cn] = v
            by_grid.setdefault(k, {}).setdefault(g, set()).add(did)
        for k, d in per.items():
            names = sorted({cn for e in d.values() for cn in e if cn != "dur"})
            n = len(d)
            if "copyBuffer" in k:
                vals = {cn: max(e.get(cn, 0) for e in d.values()) for cn in names}
                kind = "max"
            else:
                vals = {cn: sum(e.get(cn, 0) for e in d.values()) / n for cn in names}
                kind = "avg"
            lines.append(_row(k, n, sum(e["dur"] for e in d.values()) / n / 1e6, kind, vals))
        extra = []
        for k, grids in by_grid.items():
            if len(grids) < 2 or "copyBuffer" in k:
                continue
            for g, dids in sorted(grids.items(), key=lambda kv: -kv[0]):
                es = [per[k][d] for d in dids]
                names = sorted({cn for e in es for cn in e if cn != "dur"})
                vals = {cn: sum(e.get(cn, 0) for e in es) / len(es) for cn in names}
                extra.append(_row(f"{k} grid {int(g)}", len(es), sum(e["dur"] for e in es) / len(es) / 1e6, "avg",
                                  vals))
    else:
        per = collections.OrderedDict()
        grid = _grid_expr(c, "kernels")
        for n, du, g in c.execute(f"select name, duration, {grid} from kernels"):
            k = kernel_name(n)
            per.setdefault(k, []).append(du)
            by_grid.setdefault(k, {}).setdefault(g, []).append(du)
        tot = sum(sum(v) for v in per.values())
        for k, v in sorted(per.items(), key=lambda kv: -sum(kv[1])):
            lines.append(f"{k:80s} disp {len(v):6d} total_ms {sum(v) / 1e6:11.3f} avg_ms {sum(v) / len(v) / 1e6:10.4f} "
                         f"share {100 * sum(v) / tot:5.1f}%")
        extra = []
        for k, grids in sorted(by_grid.items(), key=lambda kv: -sum(sum(v) for v in kv[1].values())):
            if len(grids) < 2 or "copyBuffer" in k or sum(sum(v) for v in grids.values()) < 0.01 * tot:
                continue
            for g, v in sorted(grids.items(), key=lambda kv: -kv[0]):
                extra.append(f"{k + ' grid ' + str(int(g)):80s} disp {len(v):6d} total_ms {sum(v) / 1e6:11.3f} "
                             f"avg_ms {sum(v) / len(v) / 1e6:10.4f}")
    if extra:
        lines.append("# by grid size (work-items), kernels launched with more than one")
        lines += extra
    return "\n".join(lines) + "\n"


def main(argv) -> int:
    delete = "--delete" in argv
    roots = [a for a in argv if a != "--delete"]
    for root in roots:
        for db in glob.glob(os.path.join(root, "**", "*_results.db"), recursive=True):
            with open(db.replace("_results.db", "_summary.txt"), "w") as f:
                f.write(summarize(db))
            if delete:
                os.remove(db)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
