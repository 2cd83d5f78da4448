"""Reader of AMReX single-level plotfiles (test helper).

Written from AMReX's documented plotfile layout (Header, Level_0/Cell_H,
Level_0/Cell_D_*), independently of the library's writer, so the tests can
read what fx_write_plotfile wrote.  Returns the variable names, the header
values and each variable as an (nz, ny, nx) array."""
from __future__ import annotations

import os
import re

import numpy as np


def _box(s):
    m = re.match(r"\(\((-?\d+),(-?\d+),(-?\d+)\) \((-?\d+),(-?\d+),(-?\d+)\) \((\d+),(\d+),(\d+)\)\)", s.strip())
    assert m, s
    v = [int(g) for g in m.groups()]
    return tuple(v[:3]), tuple(v[3:6])


def read(path):
    with open(os.path.join(path, "Header")) as f:
        lines = f.read().split("\n")
    it = iter(lines)
    hdr = {"version": next(it)}
    nvar = int(next(it))
    names = [next(it) for _ in range(nvar)]
    hdr["dim"] = int(next(it))
    hdr["time"] = float(next(it))
    hdr["finest_level"] = int(next(it))
    hdr["prob_lo"] = [float(v) for v in next(it).split()]
    hdr["prob_hi"] = [float(v) for v in next(it).split()]
    hdr["ref_ratio"] = next(it).split()
    hdr["domain"] = _box(next(it))
    hdr["level_steps"] = [int(v) for v in next(it).split()]
    hdr["dx"] = [float(v) for v in next(it).split()]
    hdr["coord"] = int(next(it))
    hdr["bwidth"] = int(next(it))
    lev, nbox, t = next(it).split()
    hdr["nboxes"] = int(nbox)
    hdr["level_step"] = int(next(it))
    hdr["box_bounds"] = [[float(v) for v in next(it).split()] for _ in range(hdr["dim"] * hdr["nboxes"])]
    hdr["cell_path"] = next(it)
    lo, hi = hdr["domain"]
    nx, ny, nz = (hi[d] - lo[d] + 1 for d in range(3))
    with open(os.path.join(path, hdr["cell_path"] + "_H")) as f:
        ch = f.read().split("\n")
    hdr["cell_h"] = ch
    fod = [l for l in ch if l.startswith("FabOnDisk:")]
    assert len(fod) == 1
    _, fname, off = fod[0].split()
    with open(os.path.join(path, os.path.dirname(hdr["cell_path"]), fname), "rb") as f:
        f.seek(int(off))
        head = f.readline().decode()
        m = re.match(r"FAB \(\((\d+), \(([\d ]+)\)\),\((\d+), \(([\d ]+)\)\)\)(\(\(.*\)\)) (\d+)", head.strip())
        assert m, head
        nbytes = int(m.group(1))
        order = [int(v) for v in m.group(4).split()]
        blo, bhi = _box(m.group(5))
        ncomp = int(m.group(6))
        assert nbytes == 8 and ncomp == nvar and (blo, bhi) == (lo, hi)
        dtype = "<f8" if order == [8, 7, 6, 5, 4, 3, 2, 1] else ">f8"
        raw = np.frombuffer(f.read(8 * nx * ny * nz * ncomp), dtype=dtype)
    data = {n: raw[q * nx * ny * nz:(q + 1) * nx * ny * nz].reshape(nz, ny, nx) for q, n in enumerate(names)}
    hdr["fab_header"] = head
    return names, hdr, data
