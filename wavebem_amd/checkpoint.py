"""Solution dumps of the time loop in the reference's on-disk formats.

FreeSurface::dump_solution (source/free_surface.cc:10577-10590) writes the DAE state and
its time derivative with deal.II's ``Vector<double>::block_write`` to ``<fname>_y.dat`` and
``<fname>_yp.dat``, and the refinement history of the PersistentTriangulation with
``write_flags`` to ``<fname>_tria.dat``; FreeSurface::restore_solution (:10593-10610) reads
them back.  Dumps are taken every ``dumping_period`` of simulated time under the name
``<output_file_name>_dump_<k>`` (source/free_surface.cc:1003-1010).

deal.II (8.4/8.5, the reference's pinned version) is not in the reference tree, so its two
formats are restated here from its published documentation and source:

* ``Vector::block_write``: the size in decimal, ``'\\n'``, ``'['``, the entries as raw
  native (little-endian) IEEE doubles, ``']'``; ``block_read`` reads the size up to the
  newline and checks both brackets.
* ``PersistentTriangulation::write_flags``: the magic number 0xa200 and the number of
  refinement cycles on one line; for every cycle the refine flags and then the coarsen
  flags of that cycle, each written as ``<begin magic> <N>``, a line of the N flags packed
  eight to a byte (bit i % 8 of byte i / 8, N / 8 + 1 bytes, each as a decimal followed by
  a space) and a line with the end magic (0xa000 / 0xa001 refine, 0xa010 / 0xa011 coarsen);
  then 0xa201 on its own line.

The meshes of this library carry no refinement history, so their triangulation dump holds
zero cycles; the readers accept any number of cycles.  The state vectors are the caller's
(for the time-loop stand-in, its (eta, phi) layout, not the reference's DAE layout).
"""
from __future__ import annotations

import numpy as np

MN_TRIA_REFINE_FLAGS_BEGIN = 0xA000
MN_TRIA_REFINE_FLAGS_END = 0xA001
MN_TRIA_COARSEN_FLAGS_BEGIN = 0xA010
MN_TRIA_COARSEN_FLAGS_END = 0xA011
MN_PERSISTENT_TRIA_FLAGS_BEGIN = 0xA200
MN_PERSISTENT_TRIA_FLAGS_END = 0xA201


def block_write(f, vec):
    """Writes a float64 vector to the binary stream f in the block_write layout."""
    v = np.ascontiguousarray(vec, dtype=np.float64)
    if v.ndim != 1:
        raise ValueError("block_write takes a vector")
    f.write(f"{v.size}\n[".encode("ascii"))
    f.write(v.tobytes())  # native byte order, as the reference's raw write
    f.write(b"]")


def block_read(f):
    """Reads a vector written by block_write from the binary stream f."""
    head = bytearray()
    while True:
        ch = f.read(1)
        if not ch:
            raise ValueError("block_read: end of stream before the size line")
        if ch == b"\n":
            break
        head += ch
        if len(head) > 20:
            raise ValueError("block_read: size line too long")
    n = int(head.decode("ascii"))
    if f.read(1) != b"[":
        raise ValueError("block_read: missing '['")
    data = f.read(8 * n)
    if len(data) != 8 * n:
        raise ValueError("block_read: truncated data")
    if f.read(1) != b"]":
        raise ValueError("block_read: missing ']'")
    return np.frombuffer(data, dtype=np.float64).copy()


def _write_bool_vector(out, begin, flags, end):
    flags = [bool(b) for b in flags]
    n = len(flags)
    packed = [0] * (n // 8 + 1)
    for i, b in enumerate(flags):
        if b:
            packed[i // 8] |= 1 << (i % 8)
    out.append(f"{begin} {n}\n")
    out.append("".join(f"{b} " for b in packed))
    out.append(f"\n{end}\n")


def write_flags(f, refine_cycles=(), coarsen_cycles=()):
    """PersistentTriangulation::write_flags to the text stream f: one refine and one
    coarsen flag vector per refinement cycle."""
    if len(refine_cycles) != len(coarsen_cycles):
        raise ValueError("one refine and one coarsen vector per cycle")
    out = [f"{MN_PERSISTENT_TRIA_FLAGS_BEGIN} {len(refine_cycles)}\n"]
    for r, c in zip(refine_cycles, coarsen_cycles):
        _write_bool_vector(out, MN_TRIA_REFINE_FLAGS_BEGIN, r, MN_TRIA_REFINE_FLAGS_END)
        _write_bool_vector(out, MN_TRIA_COARSEN_FLAGS_BEGIN, c, MN_TRIA_COARSEN_FLAGS_END)
    out.append(f"{MN_PERSISTENT_TRIA_FLAGS_END}\n")
    f.write("".join(out))


def read_flags(f):
    """Reads write_flags output; returns (refine_cycles, coarsen_cycles) as bool lists."""
    tok = f.read().split()
    pos = 0

    def take():
        nonlocal pos
        if pos >= len(tok):
            raise ValueError("read_flags: unexpected end of stream")
        pos += 1
        return int(tok[pos - 1])

    def bools(begin, end):
        if take() != begin:
            raise ValueError("read_flags: bad begin magic number")
        n = take()
        packed = [take() for _ in range(n // 8 + 1)]
        if take() != end:
            raise ValueError("read_flags: bad end magic number")
        return [bool((packed[i // 8] >> (i % 8)) & 1) for i in range(n)]

    if take() != MN_PERSISTENT_TRIA_FLAGS_BEGIN:
        raise ValueError("read_flags: not a persistent triangulation flag stream")
    cycles = take()
    refine, coarsen = [], []
    for _ in range(cycles):
        refine.append(bools(MN_TRIA_REFINE_FLAGS_BEGIN, MN_TRIA_REFINE_FLAGS_END))
        coarsen.append(bools(MN_TRIA_COARSEN_FLAGS_BEGIN, MN_TRIA_COARSEN_FLAGS_END))
    if take() != MN_PERSISTENT_TRIA_FLAGS_END:
        raise ValueError("read_flags: bad end of stream")
    return refine, coarsen


def dump_name(output_file_name, counter):
    """``<output_file_name>_dump_<counter>`` (source/free_surface.cc:1006-1007)."""
    return f"{output_file_name}_dump_{int(counter)}"


def dump_solution(fname, y, yp, refine_cycles=(), coarsen_cycles=()):
    """FreeSurface::dump_solution: <fname>_y.dat, <fname>_yp.dat, <fname>_tria.dat."""
    with open(fname + "_y.dat", "wb") as f:
        block_write(f, y)
    with open(fname + "_yp.dat", "wb") as f:
        block_write(f, yp)
    with open(fname + "_tria.dat", "w") as f:
        write_flags(f, refine_cycles, coarsen_cycles)


def restore_solution(fname):
    """FreeSurface::restore_solution: (y, yp, (refine_cycles, coarsen_cycles))."""
    with open(fname + "_y.dat", "rb") as f:
        y = block_read(f)
    with open(fname + "_yp.dat", "rb") as f:
        yp = block_read(f)
    with open(fname + "_tria.dat") as f:
        flags = read_flags(f)
    return y, yp, flags


class Dumper:
    """The dump schedule of source/free_surface.cc:1003-1010: the k-th dump (k = 1, 2, ...)
    is written at the first step with t >= k * dumping_period."""

    def __init__(self, output_file_name, dumping_period):
        self.output_file_name = output_file_name
        self.period = float(dumping_period)
        self.counter = 1
        self.written = []

    def step(self, t, y, yp):
        if self.period > 0 and t >= self.period * self.counter:
            name = dump_name(self.output_file_name, self.counter)
            self.counter += 1
            dump_solution(name, y, yp)
            self.written.append(name)
            return name
        return None
