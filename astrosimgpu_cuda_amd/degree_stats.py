"""Degree statistics of a connectivity dump.

    python -m astrosimgpu_cuda_amd.degree_stats edges.csv

The reference's ``--dump-edges`` CSV (src/astrosim_fused.cu:873-902; written
here by ``as_network_dump_edges`` and the CLI's ``--dump-edges``) lists one
``type,source,target`` row per synapse: types exc_primary, inh_primary (source
ids offset by N_E), n2a (targets prefixed "A") and a2n (sources prefixed "A").
For each type this prints the edge count and, over the sources and the
targets that occur in it, the out- and in-degree statistics (count, mean,
population standard deviation, minimum, upper median, maximum), in the line
format of the reference's scripts/degree_stats.py so that the two outputs
can be compared directly.

:func:`csv_degrees` reads a dump; :func:`csr_degrees` computes the same
degrees from a network's in-CSRs without a file (a second route to the same
numbers), and :func:`dump_rows` writes the rows of a dump from in-CSRs.
"""
from __future__ import annotations

import sys
from collections import Counter

import numpy as np

TYPES = ("exc_primary", "inh_primary", "n2a", "a2n")


def csv_degrees(path) -> dict:
    """{type: (Counter of sources, Counter of targets)} of a dump."""
    out = {}
    with open(path, newline="") as f:
        header = f.readline().strip()
        if header != "type,source,target":
            raise ValueError(f"unexpected header {header!r}")
        for line in f:
            t, s, d = line.rstrip("\n").split(",")
            src, dst = out.setdefault(t, (Counter(), Counter()))
            src[s] += 1
            dst[d] += 1
    return out


def _csr_of(net, which):
    rs, src = net.csr(which)
    return np.asarray(rs, dtype=np.int64), np.asarray(src, dtype=np.int64)


def csr_degrees(net) -> dict:
    """The same degrees as csv_degrees, from a Network's in-CSRs (row =
    target): {type: (source id -> count, target id -> count)} with the
    dump's id spelling (offsets and "A" prefixes)."""
    spec = (("exc_primary", "exc", 0, "", ""), ("inh_primary", "inh", net.n_exc, "", ""),
            ("n2a", "e2a", 0, "", "A"), ("a2n", "sic", 0, "A", ""))
    out = {}
    for t, which, off, sp, tp in spec:
        rs, src = _csr_of(net, which)
        if len(src) == 0:
            continue
        s_ids, s_cnt = np.unique(src + off, return_counts=True)
        indeg = np.diff(rs)
        tgt = np.nonzero(indeg)[0]
        out[t] = (Counter({f"{sp}{i}": int(c) for i, c in zip(s_ids, s_cnt)}),
                  Counter({f"{tp}{i}": int(indeg[i]) for i in tgt}))
    return out


def dump_rows(net):
    """The rows of the network's dump, in the writer's order (types in order,
    targets ascending, each target's in-row in CSR order)."""
    spec = (("exc_primary", "exc", 0, "", ""), ("inh_primary", "inh", net.n_exc, "", ""),
            ("n2a", "e2a", 0, "", "A"), ("a2n", "sic", 0, "A", ""))
    yield "type,source,target"
    for t, which, off, sp, tp in spec:
        rs, src = _csr_of(net, which)
        for row in range(len(rs) - 1):
            for k in range(rs[row], rs[row + 1]):
                yield f"{t},{sp}{src[k] + off},{tp}{row}"


def summary_line(counts: Counter) -> str:
    vals = sorted(counts.values())
    n = len(vals)
    if n == 0:
        return "n=0"
    mean = sum(vals) / n
    std = (sum((v - mean) ** 2 for v in vals) / n) ** 0.5
    return (f"n={n:>7d}  mean={mean:8.3f}  std={std:7.3f}  "
            f"min={vals[0]:>6d}  median={vals[n // 2]:>6d}  max={vals[-1]:>6d}")


def report(degrees: dict) -> str:
    lines = []
    for t in sorted(degrees):
        src, dst = degrees[t]
        lines.append(f"== {t}: {sum(src.values())} edges ==")
        lines.append(f"  out-degree (per source): {summary_line(src)}")
        lines.append(f"  in-degree  (per target): {summary_line(dst)}")
    return "\n".join(lines)


def main(argv=None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    if len(argv) != 1:
        print("usage: python -m astrosimgpu_cuda_amd.degree_stats edges.csv", file=sys.stderr)
        return 1
    print(report(csv_degrees(argv[0])))
    return 0


if __name__ == "__main__":
    sys.exit(main())
