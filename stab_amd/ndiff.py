"""Numerical text comparison for eigenfunction files (time.N / space.N).

The reference's test scripts compare their output with the goldens by
`ndiff -abserr 1e-8` (thesis/TStest/README.md; the tool is not part of the
reference).  This is an equivalent: lines starting with '#' are skipped, the
remaining lines are split into whitespace-separated fields, numeric fields
are compared with |a - b| <= max(abserr, relerr * max(|a|, |b|)), other
fields must be equal, and the files must have the same shape.
"""
from __future__ import annotations

import argparse
import sys


def _rows(text):
    return [line.split() for line in text.splitlines() if line.strip() and not line.lstrip().startswith("#")]


def compare(a: str, b: str, abserr: float = 0.0, relerr: float = 0.0):
    """Returns (ok, max_abs_diff, first_bad) with first_bad = (line, field) or None."""
    ra, rb = _rows(a), _rows(b)
    if len(ra) != len(rb):
        return False, float("inf"), (min(len(ra), len(rb)), 0)
    worst, bad = 0.0, None
    for i, (x, y) in enumerate(zip(ra, rb)):
        if len(x) != len(y):
            return False, float("inf"), (i, min(len(x), len(y)))
        for j, (u, v) in enumerate(zip(x, y)):
            try:
                fu, fv = float(u), float(v)
            except ValueError:
                if u != v and bad is None:
                    bad = (i, j)
                continue
            d = abs(fu - fv)
            worst = max(worst, d)
            if d > max(abserr, relerr * max(abs(fu), abs(fv))) and bad is None:
                bad = (i, j)
    return bad is None, worst, bad


def main(argv=None):
    ap = argparse.ArgumentParser(description="numerical diff of two text files")
    ap.add_argument("-abserr", type=float, default=0.0)
    ap.add_argument("-relerr", type=float, default=0.0)
    ap.add_argument("file1")
    ap.add_argument("file2")
    a = ap.parse_args(argv)
    ok, worst, bad = compare(open(a.file1).read(), open(a.file2).read(), a.abserr, a.relerr)
    if not ok:
        print(f"files differ: first at data line {bad[0] + 1}, field {bad[1] + 1}; max abs diff {worst:.3e}")
        return 1
    print(f"files agree (max abs diff {worst:.3e})")
    return 0


if __name__ == "__main__":
    sys.exit(main())
