"""Compares two sets of golden fixtures array by array and prints a report:
for each fixture and array, whether the two are bitwise equal and, if not,
the largest absolute and relative difference.

    python tests/golden/compare_fixtures.py <dir_a> <dir_b>

Used to check a regeneration of the fixtures (make_golden.py <dir>) against
the committed ones.
"""
import os
import sys

import numpy as np


def compare(a_dir, b_dir):
    lines, all_equal = [], True
    names = sorted(f for f in os.listdir(a_dir) if f.endswith(".npz"))
    for name in names:
        a = np.load(os.path.join(a_dir, name), allow_pickle=False)
        b = np.load(os.path.join(b_dir, name), allow_pickle=False)
        keys = sorted(set(a.files) | set(b.files))
        diff = []
        for k in keys:
            if k not in a.files or k not in b.files:
                diff.append("%s: only in %s" % (k, "a" if k in a.files else "b"))
                continue
            x, y = a[k], b[k]
            if x.shape != y.shape or x.dtype != y.dtype:
                diff.append("%s: shape/dtype %s %s vs %s %s" % (k, x.shape, x.dtype, y.shape, y.dtype))
            elif x.tobytes() != y.tobytes():
                xd, yd = x.astype(np.float64), y.astype(np.float64)
                m = np.abs(xd - yd).max()
                s = max(np.abs(yd).max(), 1e-300)
                diff.append("%s: differs, max |a-b| %.3e (%.3e of max |b|)" % (k, m, m / s))
        all_equal &= not diff
        lines.append("%-18s %d arrays: %s" % (name, len(keys), "bitwise equal" if not diff else "DIFFERS"))
        lines += ["    " + d for d in diff]
    return all_equal, lines


if __name__ == "__main__":
    eq, lines = compare(sys.argv[1], sys.argv[2])
    print("\n".join(lines))
    print("all bitwise equal" if eq else "differences found")
