"""The reference's binary files restated from etc.cu, for the tests only.

Test infrastructure: only tests/ may import this module.  It writes the
byte streams of ``output_data`` (etc.cu:449-961) from an ``OracleSim`` and
parses them back, independently of the library's writers, so that the tests
compare the library's files with these byte for byte in layout and within
tolerance in value.  Native-endian int32 / FP64, no headers, arrays in the
order of the fwrite calls.

Fields the oracle does not model are written as the reference initialises
them: the VOF indicator vf, vf2 is 0 (init1.cu:396-411).  cel->rx, ry, rz
are never assigned by the reference (allocated at init1.cu:547, 937 and only
copied); they are written as each cell's vertex mean, the convention of the
library.
"""
from __future__ import annotations

import numpy as np

Q = 19


def expand2pow2(n: int) -> int:
    """etc.cu:425-446: halve n (rounding odd values up) until it is 1; 2^count."""
    cnt = 0
    while n > 1:
        if n % 2 == 1:
            n += 1
        n //= 2
        cnt += 1
    return 2 ** cnt


def _cell_dims(o):
    voff, eoff = o.cells["voff"], o.cells["eoff"]
    nc = len(voff) - 1
    V, E = int(voff[1] - voff[0]), int(eoff[1] - eoff[0])
    assert np.all(np.diff(voff) == V) and np.all(np.diff(eoff) == E), "the formats need one mesh"
    return nc, V, E


def _means(a, nc, V):
    return a.reshape(nc, V).mean(axis=1)


def restart_bytes(o, iter_, fnum, fn=None) -> bytes:
    """restart.bin (etc.cu:738-789) of the oracle state; fn overrides the
    populations (AoS [n*19])."""
    n = o.n
    parts = [np.array([iter_, fnum], np.int32).tobytes(),
             np.ascontiguousarray(o.fn if fn is None else fn, np.float64).tobytes()]
    for a in (o.dm, o.um, o.vm, o.wm, np.zeros(n), np.zeros(n)):
        parts.append(np.ascontiguousarray(a, np.float64).tobytes())
    if o.cells is not None:
        c = o.cells
        nc, V, E = _cell_dims(o)
        mn = expand2pow2(3 * V)
        for k in ("x", "y", "z"):
            parts.append(c[k].tobytes())
        for k in ("x", "y", "z"):
            parts.append(_means(c[k], nc, V).tobytes())
        for k in ("u", "v", "w"):
            parts.append(c[k].tobytes())
        qi = np.zeros(nc * mn)
        qi[:3 * nc * V] = c["qi"]  # substitute_load: cell c at c*3V (stress.cu:292-300)
        parts.append(qi.tobytes())
        parts.append(c["tp"].tobytes())
    return b"".join(parts)


def prof_bytes(o, iter_) -> bytes:
    """prof%04d.bin (etc.cu:547-560)."""
    n = o.n
    parts = [np.array([iter_], np.int32).tobytes()]
    for a in (o.dm, o.um, o.vm, o.wm, np.zeros(n), np.zeros(n)):
        parts.append(np.ascontiguousarray(a, np.float64).tobytes())
    parts.append(np.ascontiguousarray(o.bc, np.int32).tobytes())
    return b"".join(parts)


def cell_bytes(o) -> bytes:
    """cell%04d.bin (etc.cu:565-611)."""
    c = o.cells
    nc, V, E = _cell_dims(o)
    parts = [np.array([nc, V, E], np.int32).tobytes()]
    for k in ("x", "y", "z", "u", "v", "w"):
        parts.append(c[k].tobytes())
    parts.append(c["qi"][:nc * V * 3].tobytes())
    parts.append(c["tp"].tobytes())
    parts.append(np.ascontiguousarray(c["ele"], np.int32).tobytes())
    return b"".join(parts)


class _Reader:
    def __init__(self, buf: bytes):
        self.buf, self.off = buf, 0

    def take(self, dtype, count):
        a = np.frombuffer(self.buf, dtype=dtype, count=count, offset=self.off)
        self.off += a.nbytes
        return a

    def done(self):
        return self.off == len(self.buf)


def restart_size(n, nc=0, V=0, E=0) -> int:
    """Bytes of restart.bin for n padded nodes and nc cells of V vertices and
    E elements."""
    size = 2 * 4 + 8 * (n * Q + 6 * n)
    if nc:
        size += 8 * (6 * nc * V + 3 * nc + nc * expand2pow2(3 * V) + 2 * nc * E)
    return size


def parse_restart(buf: bytes, n, nc=0, V=0, E=0) -> dict:
    r = _Reader(buf)
    it, fnum = (int(v) for v in r.take(np.int32, 2))
    out = dict(iter=it, fnum=fnum, fn=r.take(np.float64, n * Q))
    for k in ("dm", "um", "vm", "wm", "vf", "vf2"):
        out[k] = r.take(np.float64, n)
    if nc:
        for k in ("x", "y", "z"):
            out[k] = r.take(np.float64, nc * V)
        for k in ("rx", "ry", "rz"):
            out[k] = r.take(np.float64, nc)
        for k in ("u", "v", "w"):
            out[k] = r.take(np.float64, nc * V)
        out["qi"] = r.take(np.float64, nc * expand2pow2(3 * V))
        out["tp"] = r.take(np.float64, 2 * nc * E)
    assert r.done(), "trailing bytes"
    return out


def parse_prof(buf: bytes, n) -> dict:
    r = _Reader(buf)
    out = dict(iter=int(r.take(np.int32, 1)[0]))
    for k in ("dm", "um", "vm", "wm", "vf", "vf2"):
        out[k] = r.take(np.float64, n)
    out["bc"] = r.take(np.int32, n)
    assert r.done(), "trailing bytes"
    return out


def parse_cell(buf: bytes) -> dict:
    r = _Reader(buf)
    nc, V, E = (int(v) for v in r.take(np.int32, 3))
    out = dict(n=nc, vertex=V, element=E)
    for k in ("x", "y", "z", "u", "v", "w"):
        out[k] = r.take(np.float64, nc * V)
    out["qi"] = r.take(np.float64, 3 * nc * V)
    out["tp"] = r.take(np.float64, 2 * nc * E)
    out["ele"] = r.take(np.int32, 3 * nc * E)
    assert r.done(), "trailing bytes"
    return out


def adhesion_bytes(fnum_ad, iter_, f_adh, attach, is_knob, is_attached, r_ij) -> bytes:
    """adhe%05d.bin of Adhesion() (adhesion.cu:785-810) from the adhesion
    arrays: fnum_ad, iter (int32), f_adh [3 Nv] (AoS, FP64), attachment_point
    [Nv NS] (int32), is_knob [Nv] (bool, one byte), the receptors'
    is_attached (int32) and r_ij [Nv NS] (FP64)."""
    return b"".join([np.array([fnum_ad, iter_], np.int32).tobytes(),
                     np.ascontiguousarray(f_adh, np.float64).tobytes(),
                     np.ascontiguousarray(attach, np.int32).tobytes(),
                     np.ascontiguousarray(is_knob, np.uint8).tobytes(),
                     np.ascontiguousarray(is_attached, np.int32).tobytes(),
                     np.ascontiguousarray(r_ij, np.float64).tobytes()])


def parse_adhesion(buf: bytes, nv, ns, nrec) -> dict:
    """Inverse of adhesion_bytes for Nv vertices, NS springs per knob and
    nrec receptors."""
    assert len(buf) == 8 + 24 * nv + 4 * nv * ns + nv + 4 * nrec + 8 * nv * ns, len(buf)
    o = 0

    def take(dtype, count):
        nonlocal o
        a = np.frombuffer(buf, dtype, count, o)
        o += a.nbytes
        return a

    fnum_ad, iter_ = (int(v) for v in take(np.int32, 2))
    return dict(fnum_ad=fnum_ad, iter=iter_, f_adh=take(np.float64, 3 * nv), attach=take(np.int32, nv * ns),
                is_knob=take(np.uint8, nv), is_attached=take(np.int32, nrec), r_ij=take(np.float64, nv * ns))
