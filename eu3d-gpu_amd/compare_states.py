#!/usr/bin/env python3
"""Compare EU3D_STATE_V1_LE state dumps.

Same file format and acceptance rule as the reference's comparer
(tools/compare_states.py): per field |error| <= atol + rtol*|reference| with
atol 1e-12, rtol 1e-10; Nchem must match exactly.  Single-block dumps only
(the build runs the whole domain on one GPU).

    python compare_states.py reference.bin candidate.bin
"""
import argparse
import math
import sys

import numpy as np

MAGIC = b"EU3D_STATE_V1_LE\n"
BASE_FIELDS = "X Y Z D T U V W E Wav Rgas Cp P H Gamma C Ma".split()


def field_names(ns):
    names = list(BASE_FIELDS)
    names += [f"{n}{s}" for n in ("Di", "Yi", "Mc") for s in range(ns)]
    names += [f"CS{s}" for s in range(ns + 4)] + [f"CMS{s}" for s in range(ns)] + ["Nchem"]
    return names


def read(path):
    """Return (meta ints[17], (final_time, dt), field names, data[n_nodes, n_fields])."""
    with open(path, "rb") as f:
        if f.readline() != MAGIC:
            raise ValueError(f"{path}: wrong dump format")
        tok = f.readline().decode().split()
        if len(tok) != 19:
            raise ValueError(f"{path}: wrong metadata length")
        meta = [int(t) for t in tok[:17]]
        times = tuple(float(t) for t in tok[17:])
        fields = f.readline().decode().split()
        ns = meta[0]
        if ns < 1 or fields != field_names(ns):
            raise ValueError(f"{path}: wrong field schema")
        n = meta[4:7]
        if any(x < 2 for x in meta[1:7]):
            raise ValueError(f"{path}: invalid grid dimensions")
        if meta[15] < 1 or any(not math.isfinite(t) or t <= 0 for t in times):
            raise ValueError(f"{path}: invalid iteration/time metadata")
        data = np.frombuffer(f.read(), dtype="<f8")
        nodes = n[0] * n[1] * n[2]
        if data.size != nodes * len(fields):
            raise ValueError(f"{path}: truncated or oversized dump")
    return meta, times, fields, data.reshape(nodes, len(fields))


def compare(reference, parts, atol=1e-12, rtol=1e-10, out=sys.stdout):
    """Number of mismatching values of the candidate dump(s) against the reference."""
    rm, rt, names, ref = read(reference)
    if not np.all(np.isfinite(ref)):
        raise ValueError("reference contains non-finite data")
    mismatches = 0
    for path in parts:
        m, t, fields, val = read(path)
        if (m[:7], m[15:17], t, fields) != (rm[:7], rm[15:17], rt, names):
            raise ValueError(f"{path}: different grid/species/iteration/output/time/fields")
        err = np.where(np.isfinite(val), np.abs(ref - val), np.inf)
        tol = atol + rtol * np.abs(ref)
        tol[:, names.index("Nchem")] = 0.0
        bad = err > tol
        mismatches += int(bad.sum())
        for q, name in enumerate(names):
            out.write(f"{name}: max_abs={err[:, q].max():.6e}, mismatches={int(bad[:, q].sum())}\n")
    out.write(f"State validation: {'Passed' if not mismatches else 'Failed'}, mismatches={mismatches}\n")
    return mismatches


def main():
    p = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    p.add_argument("reference")
    p.add_argument("parts", nargs="+")
    a = p.parse_args()
    try:
        bad = compare(a.reference, a.parts)
    except (OSError, ValueError) as e:
        p.exit(2, f"Validation input error: {e}\n")
    sys.exit(1 if bad else 0)


if __name__ == "__main__":
    main()
