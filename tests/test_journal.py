"""Edit journal files (VXJ1, src/journal.rs:13-139, :391-508).

tests/golden/journal_spawn.vxj is the journal the reference's headless run wrote
(uploadme_runs/journal.journal, 80 bytes, kept unchanged): seed 1337, sea level 128, a
player block and a bar block, no edits. The parser must read it as EditJournal::load
does, and the writer must give back the same bytes from what was read.
"""
import math
import os
import struct

import numpy as np
import pytest

from voxelcraft_menu_amd import _abi as A
from voxelcraft_menu_amd import journal as J
from voxelcraft_menu_amd.world import World

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "journal_spawn.vxj")


def _golden() -> bytes:
    with open(GOLDEN, "rb") as f:
        return f.read()


def test_reference_journal_reads_as_load_does():
    j = J.parse(_golden(), 1337, 128)
    assert (j.version, j.flags, j.seed, j.sea_level) == (3, 5, 1337, 128)  # PLAYER | BAR
    assert j.edits.size == 0 and j.recovered == 0 and j.torn_bytes == 0
    p = j.player
    assert p["pos"] == (8.5, 170.0, 8.5)
    # the headless camera: spawn column (8.5, 8.5), 2 + 40 blocks above max(height, sea);
    # height(8, 8) is under the sea here, so y = 128 + 42
    assert p["pos"][1] == J.headless_spawn_y(0, 128)
    # yaw 40 and pitch -14 degrees (src/config.rs:385-387), stored as f32 radians
    assert p["yaw"] == np.float32(math.radians(40.0)) and p["pitch"] == np.float32(math.radians(-14.0))
    assert p["hotbar"] == 4 and p["fly"] is True
    assert j.bar == [1, 2, 3, 4, 8, 6, 7, 9, 10, 11]


def test_writer_gives_back_the_reference_bytes():
    j = J.parse(_golden(), 1337, 128)
    assert J.dumps(j.edits, 1337, 128, j.player, j.bar) == _golden()


def _edits(rows):
    e = np.zeros(len(rows), dtype=A.EDIT_DTYPE)
    for i, (x, y, z, b) in enumerate(rows):
        e[i] = (x, y, z, b, 0)
    return e


def test_save_keeps_last_edit_per_voxel_sorted_x_z_y():
    rows = [(5, 1, 2, 1), (-3, 70, 9, 2), (5, 1, 2, 4), (5, 0, 2, 3), (5, 1, -1, 8), (-3, 70, 9, 0)]
    data = J.dumps(_edits(rows), 7, 96)
    head = struct.unpack("<4sIiiII", data[:24])
    assert head == (b"VXJ1", 3, 7, 96, 4, 0)
    recs = np.frombuffer(data[24:], dtype=A.EDIT_DTYPE)
    got = [(int(r["x"]), int(r["y"]), int(r["z"]), int(r["id"])) for r in recs]
    assert got == [(-3, 70, 9, 0), (5, 1, -1, 8), (5, 0, 2, 3), (5, 1, 2, 4)]
    assert all(int(r["pad"]) == 0 for r in recs)
    back = J.parse(data, 7, 96)
    assert back.edits.size == 4 and back.player is None and back.bar is None


def test_append_log_recovery_and_torn_tail():
    base = J.dumps(_edits([(1, 2, 3, 1)]), 7, 96)
    extra = _edits([(4, 5, 6, 2), (7, 8, 9, 3)]).tobytes()
    appended = bytearray(base + extra + b"\x01\x02\x03")
    struct.pack_into("<I", appended, 20, 2)  # FLAG_APPEND; the header still counts 1
    j = J.parse(bytes(appended), 7, 96)
    assert j.edits.size == 3 and j.recovered == 2 and j.torn_bytes == 3
    assert [int(e["id"]) for e in j.edits] == [1, 2, 3]
    with pytest.raises(J.JournalError, match="header says 1 edits"):
        J.parse(base + extra, 7, 96)  # the same bytes without FLAG_APPEND


@pytest.mark.parametrize("mutate,reason", [
    (lambda b: b[:10], "too short"),
    (lambda b: b"VXJ2" + b[4:], "not an edit journal"),
    (lambda b: b[:4] + struct.pack("<I", 0) + b[8:], "version 0"),
    (lambda b: b[:4] + struct.pack("<I", 4) + b[8:], "version 4"),
    (lambda b: b[:24] + b[24:40], "too short to hold"),
    (lambda b: b + b"\x00" * 16, "header says 0 edits"),
    (lambda b: b[:24] + struct.pack("<f", float("nan")) + b[28:], "not a finite number"),
    (lambda b: b[:44] + struct.pack("<I", 10) + b[48:], "hotbar slot 10"),
    (lambda b: b[:56] + struct.pack("<H", 13) + b[58:], "block 13 in bar slot 0"),
    (lambda b: b[:58] + struct.pack("<H", 25) + b[60:], "tuft"),
    (lambda b: b[:60] + struct.pack("<H", 26) + b[62:], "no block has that id"),
    (lambda b: b[:62] + struct.pack("<H", 0) + b[64:], "air"),
])
def test_load_refuses_what_the_reference_refuses(mutate, reason):
    with pytest.raises(J.JournalError, match=reason):
        J.parse(mutate(_golden()), 1337, 128)


def test_wrong_world_is_refused():
    with pytest.raises(J.JournalError, match="seed 1337 sea-level 128"):
        J.parse(_golden(), 1337, 96)


def test_journal_edits_apply_to_a_world(tmp_path):
    w = World()
    w.insert_dense((0, 0, 0), np.ones(64 ** 3, np.uint16))
    rows = [(3, 4, 5, 0), (10, 20, 30, 8), (3, 4, 5, 10), (63, 63, 63, 0)]
    path = str(tmp_path / "edits.vxj")
    J.write(path, _edits(rows), 7, 96)
    j = J.read(path, 7, 96)
    assert w.apply_edits(j.edits) >= 0
    assert w.get_block(3, 4, 5) == 10 and w.get_block(10, 20, 30) == 8 and w.get_block(63, 63, 63) == 0
    assert w.get_block(0, 0, 0) == 1
