"""Summaries of rocprofv3 results for the kernels of this library.

``python -m opensfm_amd.profiling DIR`` turns every ``*_results.db`` under
DIR (the SQLite output of ``rocprofv3 -d DIR -o NAME ...``) into a text
table next to it (``*_summary.txt``): per kernel, the dispatch count and
the average duration, and with ``--pmc`` the average of each counter per
dispatch (the maximum for the runtime's copy kernel, where the 1 GiB
calibration copy of ``bench.py --copy-probe-mb 1024`` is the largest
dispatch); without counters, the total time and its share.  ``--delete``
removes the databases afterwards.
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


def summarize(db: str) -> str:
    c = sqlite3.connect(db)
    lines = [f"# {os.path.basename(db)}"]
    have_pmc = c.execute("select count(*) from counters_collection").fetchone()[0] > 0
    if have_pmc:
        per = collections.OrderedDict()
        for did, kn, cn, v, du in c.execute(
                "select dispatch_id, kernel_name, counter_name, value, duration from counters_collection"):
            per.setdefault(kernel_name(kn), {}).setdefault(did, {"dur": du})[cn] = v
        for k, d in per.items():
            names = sorted({cn for e in d.values() for cn in e if cn != "dur"})
            n = len(d)
            if "copyBuffer" in k:
                vals = {cn: max(e.get(cn, 0) for e in d.values()) for cn in names}
                kind = "max"
            else:
                vals = {cn: sum(e.get(cn, 0) for e in d.values()) / n for cn in names}
                kind = "avg"
            dur = sum(e["dur"] for e in d.values()) / n / 1e6
            lines.append(f"{k:80s} disp {n:6d} avg_ms {dur:10.3f} {kind} " +
                         " ".join(f"{cn} {vals[cn]:.6g}" for cn in names))
    else:
        per = collections.OrderedDict()
        for n, du in c.execute("select name, duration from kernels"):
            per.setdefault(kernel_name(n), []).append(du)
        tot = sum(sum(v) for v in per.values())
        for k, v in sorted(per.items(), key=lambda kv: -sum(kv[1])):
            lines.append(f"{k:80s} disp {len(v):6d} total_ms {sum(v) / 1e6:11.3f} avg_ms {sum(v) / len(v) / 1e6:10.4f} "
                         f"share {100 * sum(v) / tot:5.1f}%")
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
