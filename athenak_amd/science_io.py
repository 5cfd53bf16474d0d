"""Readers of the dust science outputs: the particle file (pbin, format versions 1-3) and the
particle history (phst).  Files written by this library (dk_write_pbin, dk_write_phst) and by
AthenaK's writers (src/outputs/pbin.cpp, src/outputs/phst.cpp) read alike; only numpy and the
standard library are needed.

pbin layout (src/outputs/pbin.cpp:15-60): a 64-byte preamble ("AKPBIN01", the byte-order mark
0x01020304, the format version, then text_bytes, data_start, data_bytes, n_particles, n_blocks
as uint64), the text header ("[section]" lines of "key = value"; [columns] lists every column as
"column = table name dtype count offset bytes [encoding]"; [constants] the particle quantities
stored once; [chunks] the compressed chunks of version 3; [input] last, the run's parameters),
the columns from data_start on, and a 16-byte footer ("AKPBEND1", the file size).
"""

import os
import re
import zlib

import numpy as np

PBIN_MAGIC, PBIN_END = b"AKPBIN01", b"AKPBEND1"
PHST_COLUMNS = ("time", "dt", "n_particles", "mass", "mom1", "mom2", "mom3", "ekin", "z_rms",
                "rho_max", "rho_max_x1", "rho_max_x2", "rho_max_x3", "rho_max_level",
                "chi_max", "dt_stiff_min")


class ScienceFileError(ValueError):
    pass


def topology_hash(level, lx1, lx2, lx3):
    """The pbin header's block_topology_hash: FNV-1a 64 (with the writer's offset basis) over
    (level, lx1, lx2, lx3) of every block in gid order as little-endian int64."""
    h, mask = 1469598103934665603, (1 << 64) - 1
    arr = np.stack([np.asarray(a, np.int64) for a in (level, lx1, lx2, lx3)], axis=1)
    for byte in arr.astype("<i8").tobytes():
        h = ((h ^ byte) * 1099511628211) & mask
    return h


def _sections(text):
    """{section: {key: value}}, the column list, the constants, the chunks, the input dump."""
    head, sep, dump = text.partition("\n[input]\n")
    if not sep:
        raise ScienceFileError("pbin text header without an [input] section")
    sec, cols, consts, chunks = {}, [], {}, []
    cur = None
    for line in head.split("\n")[1:]:
        s = line.strip()
        if not s or s.startswith("#"):
            continue
        if s.startswith("[") and s.endswith("]"):
            cur = s[1:-1]
            sec.setdefault(cur, {})
            continue
        if cur is None or "=" not in line:
            continue
        key, value = (x.strip() for x in line.split("=", 1))
        f = value.split()
        if cur == "columns" and key == "column":
            cols.append(dict(table=f[0], name=f[1], dtype=f[2], count=int(f[3]),
                             offset=int(f[4]), nbytes=int(f[5]),
                             encoding=f[6] if len(f) > 6 else "raw"))
        elif cur == "constants" and key == "constant":
            consts[(f[0], f[1])] = (f[2], f[3])
        elif cur == "chunks" and key == "chunk":
            chunks.append(dict(name=f[0], first_row=int(f[2]), rows=int(f[3]),
                               offset=int(f[4]), stored=int(f[5])))
        else:
            sec[cur][key] = value.split("  #")[0].strip()
    return sec, cols, consts, chunks, dump


def read_pbin(path):
    """{"header": sections, "preamble": dict, "particles": {name: array}, "blocks": {name:
    array}, "input": text}; header constants become read-only broadcast columns."""
    size = os.path.getsize(path)
    with open(path, "rb") as f:
        blob = f.read()
    if len(blob) < 80 or blob[:8] != PBIN_MAGIC:
        raise ScienceFileError("%s: not a pbin file" % path)
    mark = int.from_bytes(blob[8:12], "little")
    bo = "<" if mark == 0x01020304 else ">" if mark == 0x04030201 else None
    if bo is None:
        raise ScienceFileError("%s: bad byte-order mark" % path)
    version = int(np.frombuffer(blob[12:16], bo + "u4")[0])
    if version not in (1, 2, 3):
        raise ScienceFileError("%s: format version %d" % (path, version))
    text_bytes, data_start, data_bytes, n_part, n_blocks = (
        int(v) for v in np.frombuffer(blob[16:56], bo + "u8"))
    foot = blob[data_start + data_bytes:data_start + data_bytes + 16]
    if foot[:8] != PBIN_END or int(np.frombuffer(foot[8:16], bo + "u8")[0]) != size:
        raise ScienceFileError("%s: incomplete file (footer)" % path)
    sec, cols, consts, chunks, dump = _sections(blob[64:64 + text_bytes].decode("utf-8"))
    out = {"particles": {}, "blocks": {}}
    for c in cols:
        dt = np.dtype(c["dtype"])
        at = data_start + c["offset"]
        if c["offset"] + c["nbytes"] > data_bytes:
            raise ScienceFileError("%s: column %s past the data section" % (path, c["name"]))
        if c["encoding"] == "raw":
            if c["nbytes"] != c["count"] * dt.itemsize:
                raise ScienceFileError("%s: column %s has an inconsistent size"
                                       % (path, c["name"]))
            arr = np.frombuffer(blob, dt, c["count"], at)
        elif c["encoding"] == "shuffle_deflate":
            parts = []
            for ch in (x for x in chunks if x["name"] == c["name"]):
                if ch["rows"] == 0:
                    continue
                raw = zlib.decompress(blob[data_start + ch["offset"]:
                                           data_start + ch["offset"] + ch["stored"]])
                sh = np.frombuffer(raw, np.uint8).reshape(dt.itemsize, ch["rows"])
                parts.append(np.ascontiguousarray(sh.T).view(dt).reshape(-1))
            arr = np.concatenate(parts) if parts else np.zeros(0, dt)
        else:
            raise ScienceFileError("%s: encoding %s" % (path, c["encoding"]))
        out[c["table"]][c["name"]] = arr
    for (table, name), (dtype, value) in consts.items():
        if table != "particles":
            continue
        if name in out["particles"]:
            raise ScienceFileError("%s: %s is a column and a constant" % (path, name))
        dt = np.dtype(dtype)
        v = float(value) if dt.kind == "f" else int(value)
        out["particles"][name] = np.broadcast_to(np.array(v, dt), (n_part,))
    out["header"] = sec
    out["preamble"] = dict(version=version, byte_order=bo, data_start=data_start,
                           data_bytes=data_bytes, n_particles=n_part, n_blocks=n_blocks)
    out["input"] = dump
    return out


def read_phst(path):
    """{column: array} of a phst file (the last header block's names; rows that a restart
    repeated, time >= the first time of a later block, dropped)."""
    blocks, names, rows, comments = [], None, [], []
    with open(path, "r", encoding="utf-8") as f:
        for line in f:
            if line.startswith("#"):
                if names is not None and rows:
                    blocks.append((names, rows))
                    names, rows = None, []
                found = re.findall(r"\[(\d+)\]=(\S+)", line)
                if found:
                    names = [n for _, n in found]
                elif not line.startswith("# Athena++ history data"):
                    comments.append(line.rstrip("\n"))
                continue
            if names is not None and line.strip():
                rows.append([float(v) for v in line.split()])
    if names is not None:
        blocks.append((names, rows))
    if not blocks:
        raise ScienceFileError("%s: no phst header" % path)
    names = blocks[-1][0]
    data = []
    for n, r in blocks:
        if n != names:
            raise ScienceFileError("%s: header blocks with different columns" % path)
        a = np.array(r, float).reshape(-1, len(names))
        if data and len(a):
            data = [d[d[:, 0] < a[0, 0]] for d in data]
        data.append(a)
    table = np.concatenate(data)
    out = {n: table[:, k] for k, n in enumerate(names)}
    out["_comments"] = comments
    return out
