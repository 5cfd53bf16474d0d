"""The science-file readers (athenak_amd.science_io) on files assembled here byte by byte from the
format the reference's writer documents (src/outputs/pbin.cpp:15-60, phst.cpp:80-163)."""
import struct
import zlib

import numpy as np
import pytest

from athenak_amd import science_io as sio


def fnv_by_hand(rows):
    h = 1469598103934665603
    for row in rows:
        for v in row:
            for b in struct.pack("<q", v):
                h ^= b
                h = (h * 1099511628211) % (1 << 64)
    return h


def test_topology_hash():
    rng = np.random.default_rng(1)
    rows = rng.integers(-3, 40, size=(7, 4))
    assert sio.topology_hash(*rows.T) == fnv_by_hand(rows.tolist())
    assert sio.topology_hash([], [], [], []) == 1469598103934665603
    # order matters (gid order is part of the hash)
    assert sio.topology_hash(*rows[::-1].T) != sio.topology_hash(*rows.T)


def assemble(cols, consts, text_extra="", version=2, chunk_rows=None, corrupt=None):
    """A pbin file: cols = [(table, name, array)], consts = [(name, dtype, value)]."""
    data, lines, chunk_lines, at = b"", [], [], 0
    for table, name, arr in cols:
        arr = np.ascontiguousarray(arr)
        enc = "raw"
        if version == 3 and table == "particles":
            enc = "shuffle_deflate"
            stored = b""
            rows = chunk_rows or max(1, arr.size)
            for r0 in range(0, arr.size, rows):
                part = arr[r0:r0 + rows]
                sh = part.view(np.uint8).reshape(part.size, arr.itemsize).T.tobytes()
                z = zlib.compress(sh, 1)
                chunk_lines.append("chunk = %s 0 %d %d %d %d" % (name, r0, part.size,
                                                                  at + len(stored), len(z)))
                stored += z
            raw = stored
        else:
            raw = arr.tobytes()
        line = "column = %s %s %s %d %d %d" % (table, name, arr.dtype.str, arr.size, at,
                                                len(raw))
        lines.append(line + (" " + enc if version == 3 else ""))
        data += raw + b"\0" * ((64 - len(raw) % 64) % 64)
        at = len(data)
    text = ("AthenaK particle science file (pbin) format_version = %d\n[file]\nformat = AKPBIN\n"
            "[snapshot]\ntime = 0.25  # a comment\ncycle = 7\n%s[columns]\n%s\n" %
            (version, text_extra, "\n".join(lines)))
    if chunk_lines:
        text += "[chunks]\n" + "\n".join(chunk_lines) + "\n"
    text += "[constants]\n" + "".join("constant = particles %s %s %r\n" % c for c in consts)
    text += "[input]\n<job>\nbasename = x\n"
    tb = text.encode()
    data_start = 64 + len(tb) + (64 - (64 + len(tb)) % 64) % 64
    n_part = next(a.size for t, n, a in cols if t == "particles")
    n_blk = next(a.size for t, n, a in cols if t == "blocks")
    size = data_start + len(data) + 16
    pre = b"AKPBIN01" + struct.pack("<II", 0x01020304, version) + struct.pack(
        "<6Q", len(tb), data_start, len(data), n_part, n_blk, 0)
    blob = pre + tb + b"\0" * (data_start - 64 - len(tb)) + data + b"AKPBEND1" + \
        struct.pack("<Q", size if corrupt != "size" else size + 1)
    return blob


@pytest.mark.parametrize("version", [2, 3])
def test_pbin_columns_constants_and_blocks(tmp_path, version):
    rng = np.random.default_rng(5)
    n = 1000
    ids = rng.integers(0, 2 ** 40, n).astype("<u8")
    x = rng.random(n)
    v = rng.normal(size=n).astype("<f4")
    gid = np.arange(6, dtype="<i8")
    blob = assemble([("particles", "id", ids), ("particles", "x1", x), ("particles", "v1", v),
                     ("blocks", "gid", gid)],
                    [("mass", "<f8", 0.1 / 3), ("species", "<u4", 0)], version=version,
                    chunk_rows=300)
    p = tmp_path / "a.pbin"
    p.write_bytes(blob)
    d = sio.read_pbin(str(p))
    assert d["preamble"]["version"] == version and d["preamble"]["n_particles"] == n
    np.testing.assert_array_equal(d["particles"]["id"], ids)
    np.testing.assert_array_equal(d["particles"]["x1"], x)
    np.testing.assert_array_equal(d["particles"]["v1"], v)
    assert d["particles"]["v1"].dtype == np.float32
    np.testing.assert_array_equal(d["blocks"]["gid"], gid)
    assert d["particles"]["mass"].shape == (n,) and d["particles"]["mass"][17] == 0.1 / 3
    assert int(d["particles"]["species"][0]) == 0
    assert d["header"]["snapshot"]["time"] == "0.25" and d["header"]["snapshot"]["cycle"] == "7"
    assert d["input"].startswith("<job>")


def test_pbin_incomplete_files_are_refused(tmp_path):
    cols = [("particles", "x1", np.arange(4.0)), ("blocks", "gid", np.arange(1, dtype="<i8"))]
    p = tmp_path / "b.pbin"
    p.write_bytes(assemble(cols, [], corrupt="size"))
    with pytest.raises(sio.ScienceFileError, match="footer"):
        sio.read_pbin(str(p))
    p.write_bytes(assemble(cols, [])[:-20])
    with pytest.raises(sio.ScienceFileError):
        sio.read_pbin(str(p))
    p.write_bytes(b"NOTPBIN!" + assemble(cols, [])[8:])
    with pytest.raises(sio.ScienceFileError, match="not a pbin"):
        sio.read_pbin(str(p))


def test_phst_restart_blocks(tmp_path):
    head = "# AthenaK dust particle history (phst) version 1\n# velocity_frame = lab\n" \
           "# Athena++ history data\n#" + "".join("  [%d]=%s" % (k + 1, c) for k, c in
                                                  enumerate(sio.PHST_COLUMNS)) + "\n"
    rows1 = [[t, 0.1] + [float(k) for k in range(14)] for t in (0.0, 0.1, 0.2, 0.3)]
    rows2 = [[t, 0.1] + [float(-k) for k in range(14)] for t in (0.2, 0.3, 0.4)]
    body = head + "".join(" ".join("%.16e" % v for v in r) + "\n" for r in rows1)
    body += head + "".join(" ".join("%.16e" % v for v in r) + "\n" for r in rows2)
    p = tmp_path / "h.phst"
    p.write_text(body)
    h = sio.read_phst(str(p))
    np.testing.assert_array_equal(h["time"], [0.0, 0.1, 0.2, 0.3, 0.4])
    np.testing.assert_array_equal(h["mom1"], [2.0, 2.0, -2.0, -2.0, -2.0])
    assert "# velocity_frame = lab" in h["_comments"]
