"""Boundary-condition series of a CATHY project: the non-atmospheric,
non-seepage-face Dirichlet and Neumann files (input/nansfdirbc,
input/nansfneubc) and the seepage-face file (input/sfbc).

The files are Fortran list-directed text read record by record by the
reference (every READ starts a new line, its items may run over several
lines, whatever follows its last item on that line is skipped):

* nansfdirbc / nansfneubc (bcone.f:53-57, bcnxt.f:37-40, rdndbc.f,
  readbc.f), per time record::

      TIME
      NODIN2 NODINF
      [NODIN2 surface node ids]            if NODIN2 > 0
      [NODINF 3-d node ids]                if NODINF > 0
      [NODIN2 values]                      if NODIN2 > 0
      [NODINF values]                      if NODINF > 0

  NODIN2 > 0 surface nodes are replicated down every layer (node + l*NNOD,
  l = 1..NSTR, the same value); NODIN2 < 0 (Neumann only) adds the NNOD
  base nodes NSTR*NNOD + i, whose fluxes NEUMANN computes as free drainage
  (neumann.f:19-24), ahead of the NODINF listed nodes.
* sfbc (sfvone.f, sfvnxt.f), per time record: TIME, NSF, then per face its
  node count and its node ids (top to bottom).

``read_bc`` / ``read_sfbc`` give the expanded records the reference's
BCNOD/BCINP and SFVNOD arrays hold; ``write_bc`` / ``write_sfbc`` write the
same layout (node ids 1-based, values with 17 significant digits).  The
series go to the device with ``gpu.FlowGPU.set_bc`` / ``set_seepage``.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np


class _ListDirected:
    """Fortran list-directed READ over the lines of a text file."""

    def __init__(self, text: str, name: str):
        self.lines = text.splitlines()
        self.pos = 0
        self.name = name

    def eof(self) -> bool:
        return self.pos >= len(self.lines)

    @staticmethod
    def _tokens(line: str):
        out = []
        for tok in line.replace(",", " ").replace("\t", " ").split():
            if tok.startswith("/"):
                break
            if "*" in tok:   # r*c repeat count
                r, c = tok.split("*", 1)
                if r.isdigit():
                    out.extend([c] * int(r))
                    continue
            out.append(tok)
        return out

    def read(self, k: int) -> list[str]:
        """k items from the next record(s); the rest of the last line read
        is skipped, as a list-directed READ does."""
        got: list[str] = []
        while len(got) < k:
            if self.eof():
                raise EOFError(f"{self.name}: file ends where {k} items were expected (line {self.pos + 1})")
            got.extend(self._tokens(self.lines[self.pos]))
            self.pos += 1
        return got[:k]

    def read_float(self, k: int) -> np.ndarray:
        try:
            return np.array([float(t.replace("d", "e").replace("D", "e")) for t in self.read(k)], dtype=np.float64)
        except ValueError as e:
            raise ValueError(f"{self.name}: bad number near line {self.pos}: {e}") from None

    def read_int(self, k: int) -> np.ndarray:
        toks = self.read(k)
        try:
            return np.array([int(t) for t in toks], dtype=np.int64)
        except ValueError as e:
            raise ValueError(f"{self.name}: bad integer near line {self.pos}: {e}") from None


@dataclass
class BCRecord:
    """One time record, expanded as rdndbc.f/readbc.f expand it: ``nodes``
    0-based node ids, ``values`` the head (m) or flux (m^3/s) per node."""
    time: float
    nodin2: int
    nodinf: int
    nodes: np.ndarray
    values: np.ndarray


@dataclass
class BCSeries:
    """A nansfdirbc / nansfneubc file: records in file order."""
    kind: str                      # "dirichlet" or "neumann"
    records: list[BCRecord] = field(default_factory=list)

    @property
    def free_drainage(self) -> bool:
        """Every record asks for free drainage at the base (NODIN2 < 0)."""
        return bool(self.records) and all(r.nodin2 < 0 for r in self.records)

    def file_records(self, nnod: int, nstr: int):
        """The records as ``write_bc`` takes them (the inverse of the
        expansion: NODIN2 > 0 records give back their surface part)."""
        out = []
        for r in self.records:
            if r.nodin2 > 0:
                k, rep = r.nodin2, r.nodin2 * (nstr + 1)
                out.append((r.time, k, r.nodes[:k], r.values[:k], r.nodes[rep:], r.values[rep:]))
            elif r.nodin2 < 0:
                out.append((r.time, r.nodin2, [], [], r.nodes[nnod:], r.values[nnod:]))
            else:
                out.append((r.time, 0, [], [], r.nodes, r.values))
        return out

    def explicit(self, nnod: int):
        """(times, counts, nodes, values) of the listed entries, the ones
        after the NNOD free-drainage base nodes of NODIN2 < 0 records: what
        ``cathy_gpu_flow3d_set_bc`` takes (node ids 1-based)."""
        times, counts, nodes, vals = [], [], [], []
        for r in self.records:
            skip = nnod if r.nodin2 < 0 else 0
            times.append(r.time)
            counts.append(len(r.nodes) - skip)
            nodes.append(r.nodes[skip:] + 1)
            vals.append(r.values[skip:])
        cat = lambda xs, dt: np.concatenate(xs).astype(dt) if xs else np.zeros(0, dtype=dt)
        return (np.array(times, dtype=np.float64), np.array(counts, dtype=np.int32), cat(nodes, np.int32),
                cat(vals, np.float64))


def _expand(kind: str, t: float, nodin2: int, surf, extra, sv, xv, nnod: int, nstr: int, where: str) -> BCRecord:
    """One record as rdndbc.f / readbc.f expand it (0-based node ids)."""
    n = nnod * (nstr + 1)
    surf = np.asarray(surf, dtype=np.int64)
    extra = np.asarray(extra, dtype=np.int64)
    sv = np.asarray(sv, dtype=np.float64)
    xv = np.asarray(xv, dtype=np.float64)
    if nodin2 < 0 and kind == "dirichlet":
        raise ValueError(f"{where}: NODIN2 < 0 (free drainage) is a Neumann-only record")
    if nodin2 > 0 and (len(surf) != nodin2 or len(sv) != nodin2):
        raise ValueError(f"{where}: NODIN2 = {nodin2} surface nodes and values expected")
    if len(extra) != len(xv):
        raise ValueError(f"{where}: {len(extra)} nodes but {len(xv)} values")
    if nodin2 < 0:
        nodes = np.concatenate([nstr * nnod + np.arange(nnod), extra])
        values = np.concatenate([np.zeros(nnod), xv])
    elif nodin2 > 0:
        nodes = np.concatenate([surf + l * nnod for l in range(nstr + 1)] + [extra])
        values = np.concatenate([sv] * (nstr + 1) + [xv])
    else:
        nodes, values = extra, xv
    if len(nodes) and (nodes.min() < 0 or nodes.max() >= n):
        raise ValueError(f"{where}: node id out of 1..{n} in the record at t={t}")
    return BCRecord(float(t), int(nodin2), len(extra), nodes.astype(np.int64), values)


def make_series(kind: str, records, nnod: int, nstr: int) -> BCSeries:
    """A series from records ``(time, nodin2, surface_nodes, surface_values,
    nodes3d, values3d)`` (0-based ids), expanded as the reference expands a
    file holding them (``write_bc`` writes that file)."""
    if kind not in ("dirichlet", "neumann"):
        raise ValueError("kind is 'dirichlet' or 'neumann'")
    out = BCSeries(kind, [_expand(kind, t, nodin2, surf, extra, sv, xv, nnod, nstr, f"record {i}")
                          for i, (t, nodin2, surf, sv, extra, xv) in enumerate(records)])
    if not out.records:
        raise ValueError("no time record")
    return out


def read_bc(path: str, kind: str, nnod: int, nstr: int) -> BCSeries:
    """Read a nansfdirbc (kind "dirichlet") or nansfneubc ("neumann") file
    for a mesh of ``nnod`` surface nodes and ``nstr`` layers."""
    if kind not in ("dirichlet", "neumann"):
        raise ValueError("kind is 'dirichlet' or 'neumann'")
    with open(path) as f:
        rd = _ListDirected(f.read(), os.path.basename(path))
    out = BCSeries(kind)
    while True:
        # a READ of TIME that meets the end of the file ends the series
        # (bcnxt.f:37, END=700)
        while not rd.eof() and not rd._tokens(rd.lines[rd.pos]):
            rd.pos += 1
        if rd.eof():
            break
        t = float(rd.read_float(1)[0])
        nodin2, nodinf = (int(v) for v in rd.read_int(2))
        if nodinf < 0:
            raise ValueError(f"{path}: negative NODINF {nodinf}")
        # rdndbc.f: surface nodes (NODIN2 > 0), then the NODINF 3-d nodes;
        # readbc.f: their values in the same order
        surf = rd.read_int(nodin2) - 1 if nodin2 > 0 else np.zeros(0, dtype=np.int64)
        extra = rd.read_int(nodinf) - 1 if nodinf else np.zeros(0, dtype=np.int64)
        sv = rd.read_float(nodin2) if nodin2 > 0 else np.zeros(0)
        xv = rd.read_float(nodinf) if nodinf else np.zeros(0)
        out.records.append(_expand(kind, t, nodin2, surf, extra, sv, xv, nnod, nstr, path))
    if not out.records:
        raise ValueError(f"{path}: no time record")
    return out


def write_bc(path: str, records, nnod: int | None = None) -> None:
    """Write records ``(time, nodin2, surface_nodes, surface_values,
    nodes3d, values3d)`` (0-based node ids) in the nansfdirbc/nansfneubc
    layout.  For NODIN2 < 0 pass ``surface_nodes``/``surface_values`` empty
    (the base nodes are implied)."""
    with open(path, "w") as f:
        for t, nodin2, sn, sv, xn, xv in records:
            sn, sv, xn, xv = (np.asarray(a) for a in (sn, sv, xn, xv))
            if nodin2 > 0 and len(sn) != nodin2:
                raise ValueError("NODIN2 must equal the number of surface nodes")
            f.write(f"{t:.17g}\ttime\n{nodin2}\t{len(xn)}\tNODIN2\tNODINF\n")
            for v in (sn if nodin2 > 0 else []):
                f.write(f"{int(v) + 1}\n")
            for v in xn:
                f.write(f"{int(v) + 1}\n")
            for v in (sv if nodin2 > 0 else []):
                f.write(f"{float(v):.17g}\n")
            for v in xv:
                f.write(f"{float(v):.17g}\n")


@dataclass
class SFRecord:
    """One seepage-face record: the faces' node lists (0-based, top to
    bottom) from ``time`` on."""
    time: float
    faces: list[np.ndarray]


def read_sfbc(path: str, n: int, z: np.ndarray | None = None) -> list[SFRecord]:
    """Read an sfbc file (sfvone.f / sfvnxt.f).  With ``z`` the faces are
    checked to run downwards, as sfvnxt.f:52-62 does."""
    with open(path) as f:
        rd = _ListDirected(f.read(), os.path.basename(path))
    out: list[SFRecord] = []
    while True:
        while not rd.eof() and not rd._tokens(rd.lines[rd.pos]):
            rd.pos += 1
        if rd.eof():
            break
        t = float(rd.read_float(1)[0])
        nsf = int(rd.read_int(1)[0])
        faces = []
        for _ in range(nsf):
            k = int(rd.read_int(1)[0])
            nodes = rd.read_int(k) - 1
            if len(nodes) and (nodes.min() < 0 or nodes.max() >= n):
                raise ValueError(f"{path}: seepage node out of 1..{n} at t={t}")
            if z is not None and np.any(z[nodes[:-1]] < z[nodes[1:]]):
                raise ValueError(f"{path}: seepage face nodes not in descending elevation at t={t}")
            faces.append(nodes.astype(np.int64))
        out.append(SFRecord(t, faces))
    if not out:
        raise ValueError(f"{path}: no time record")
    return out


def write_sfbc(path: str, records: list[SFRecord]) -> None:
    with open(path, "w") as f:
        for r in records:
            f.write(f"{r.time:.17g}\n{len(r.faces)}\n")
            for nodes in r.faces:
                f.write(f"{len(nodes)}\n" + " ".join(str(int(v) + 1) for v in nodes) + "\n")


def seepage_faces(records: list[SFRecord]) -> list[np.ndarray]:
    """The faces of an sfbc series for the device, which takes one fixed set
    (cathy_gpu_flow3d_set_seepage): every record must hold the same faces
    (a last record far beyond the run, which the reference needs because
    SFVONE reads the next record's header at once, repeats them)."""
    first = records[0].faces
    for r in records[1:]:
        if len(r.faces) != len(first) or any(not np.array_equal(a, b) for a, b in zip(r.faces, first)):
            raise ValueError(f"seepage faces change at t={r.time}: time-varying face sets are not supported")
    return [np.asarray(f, dtype=np.int64) for f in first]
