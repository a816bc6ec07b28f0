"""Edit journal files (VXJ1, src/journal.rs): read a journal into edit records for
World.apply_edits / the server's edit batches, and write one as EditJournal::save does.

The byte work is in libvxhost (vxw_journal_parse / vxw_journal_write); this module is the
Python face of it.
"""
from __future__ import annotations

import ctypes as C
import dataclasses

import numpy as np

from . import _abi as A

E_JOURNAL = -9


class JournalError(ValueError):
    """The file is not a journal this world can load (the reason EditJournal::load gives)."""


@dataclasses.dataclass
class Journal:
    version: int
    flags: int
    seed: int
    sea_level: int
    edits: np.ndarray                      # EDIT_DTYPE records, file order
    player: dict | None = None             # pos, yaw, pitch, hotbar, fly
    bar: list[int] | None = None           # 10 block ids
    recovered: int = 0                     # appended edits past the header's count
    torn_bytes: int = 0                    # trailing bytes of a half-written edit, dropped


def parse(data: bytes, seed: int, sea_level: int) -> Journal:
    """EditJournal::load (src/journal.rs:391-475) of a journal's bytes."""
    L = A.host()
    buf = C.create_string_buffer(bytes(data), len(data))
    info = A.JournalInfo()
    err = C.create_string_buffer(512)
    rc = L.vxw_journal_parse(buf, len(data), seed, sea_level, C.byref(info), None, 0, err, len(err))
    if rc == E_JOURNAL:
        raise JournalError(err.value.decode())
    A.check(rc, "vxw_journal_parse")
    edits = np.zeros(info.count, dtype=A.EDIT_DTYPE)
    if info.count:
        A.check(L.vxw_journal_parse(buf, len(data), seed, sea_level, C.byref(info), A.ptr(edits), edits.size,
                                    err, len(err)), "vxw_journal_parse")
    player = None
    if info.has_player:
        p = info.player
        player = {"pos": (p.pos[0], p.pos[1], p.pos[2]), "yaw": p.yaw, "pitch": p.pitch, "hotbar": p.hotbar,
                  "fly": bool(p.fly)}
    bar = list(info.bar) if info.has_bar else None
    return Journal(info.version, info.flags, info.seed, info.sea_level, edits, player, bar, info.recovered,
                   info.torn_bytes)


def read(path: str, seed: int, sea_level: int) -> Journal:
    with open(path, "rb") as f:
        return parse(f.read(), seed, sea_level)


def dumps(edits: np.ndarray, seed: int, sea_level: int, player: dict | None = None,
          bar: list[int] | None = None) -> bytes:
    """EditJournal::save (src/journal.rs:477-508): the last edit of each voxel, sorted by
    (x, z, y), after the header and the optional player and bar blocks."""
    L = A.host()
    e = np.ascontiguousarray(edits, dtype=A.EDIT_DTYPE)
    pl = None
    if player is not None:
        x, y, z = player["pos"]
        pl = A.Player((C.c_float * 3)(x, y, z), player["yaw"], player["pitch"], player["hotbar"],
                      int(player["fly"]), 0)
    b = None
    if bar is not None:
        if len(bar) != 10:
            raise ValueError("a bar holds 10 block ids")
        b = np.ascontiguousarray(bar, dtype=np.uint16)
    n = C.c_size_t()
    args = (A.ptr(e) if e.size else None, e.size, seed, sea_level, C.byref(pl) if pl is not None else None,
            A.ptr(b) if b is not None else None)
    A.check(L.vxw_journal_write(*args, None, 0, C.byref(n)), "vxw_journal_write")
    out = C.create_string_buffer(n.value)
    A.check(L.vxw_journal_write(*args, out, n.value, C.byref(n)), "vxw_journal_write")
    return out.raw[:n.value]


def write(path: str, edits: np.ndarray, seed: int, sea_level: int, player: dict | None = None,
          bar: list[int] | None = None) -> None:
    with open(path, "wb") as f:
        f.write(dumps(edits, seed, sea_level, player, bar))


def headless_spawn_y(terrain_height: int, sea_level: int) -> float:
    """The player y of a headless run's journal: spawn_position (src/scene.rs:17-21) puts
    the player 2 blocks above max(height(8, 8), sea level), and the headless camera sits
    40 blocks above that (src/headless.rs:196-198)."""
    return float(max(terrain_height, sea_level) + 2 + 40)
