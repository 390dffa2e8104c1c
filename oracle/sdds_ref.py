"""Pure-Python reader and writer for the SDDS1 subset spiffe's files use.

TEST INFRASTRUCTURE ONLY: the checker of the library's SDDS reader and of
the files its deck runner writes; only tests/ load it.  Written from the
SDDS1 protocol: a text header of &description/&parameter/&column/&data
commands ('!' lines are comments), then pages.  A binary page is an int32
row count, the values of the parameters without fixed_value, then the rows
in order, each row holding every column (strings as int32 length + bytes);
an ASCII page has one line per such parameter, a row-count line (unless
no_row_counts=1) and one line per row.
"""
import re
import struct

import numpy as np

_FMT = {"double": "d", "float": "f", "long": "i", "ulong": "I", "short": "h", "ushort": "H", "character": "c"}


def _fields(cmd):
    out = {}
    for k, v in re.findall(r'(\w+)\s*=\s*("(?:[^"\\]|\\.)*"|[^,\s&]+)', cmd):
        if v.startswith('"'):
            v = v[1:-1].replace('\\"', '"')
        out[k.lower()] = v
    return out


def read(path):
    """{'params': [dict], 'columns': [dict], 'pages': [{'params': {...}, 'columns': {name: array}}]}"""
    data = open(path, "rb").read()
    assert data.startswith(b"SDDS"), path
    pos = data.index(b"\n") + 1
    header = []
    while True:
        end = data.index(b"\n", pos)
        line = data[pos:end].decode()
        pos = end + 1
        if line.lstrip().startswith("!"):
            continue
        header.append(line)
        if "&data" in line and "&end" in line:
            break
    cmds = re.findall(r"&(\w+)(.*?)&end", "\n".join(header), flags=re.S)
    params, cols, mode = [], [], {}
    for name, body in cmds:
        f = _fields(body)
        if name == "parameter":
            params.append(f)
        elif name == "column":
            cols.append(f)
        elif name == "data":
            mode = f
    pages = []
    if mode.get("mode", "binary") == "binary":
        while pos + 4 <= len(data):
            (rows,) = struct.unpack_from("<i", data, pos)
            pos += 4
            pv = {}
            for p in params:
                if "fixed_value" in p:
                    pv[p["name"]] = p["fixed_value"]
                    continue
                t = p.get("type", "double")
                if t == "string":
                    (n,) = struct.unpack_from("<i", data, pos)
                    pv[p["name"]] = data[pos + 4:pos + 4 + n].decode()
                    pos += 4 + n
                else:
                    (v,) = struct.unpack_from("<" + _FMT[t], data, pos)
                    pos += struct.calcsize(_FMT[t])
                    pv[p["name"]] = v
            types = [c.get("type", "double") for c in cols]
            if "string" in types:
                vals = {c["name"]: [] for c in cols}
                for _ in range(rows):
                    for c, t in zip(cols, types):
                        if t == "string":
                            (n,) = struct.unpack_from("<i", data, pos)
                            vals[c["name"]].append(data[pos + 4:pos + 4 + n].decode())
                            pos += 4 + n
                        else:
                            (v,) = struct.unpack_from("<" + _FMT[t], data, pos)
                            pos += struct.calcsize(_FMT[t])
                            vals[c["name"]].append(v)
                colv = {k: np.array(v) for k, v in vals.items()}
            else:
                dt = np.dtype([(c["name"], "<" + _FMT[t]) for c, t in zip(cols, types)])
                arr = np.frombuffer(data, dtype=dt, count=rows, offset=pos)
                pos += rows * dt.itemsize
                colv = {c["name"]: arr[c["name"]].copy() for c in cols}
            pages.append({"params": pv, "columns": colv})
    else:
        lines = [l for l in data[pos:].decode().split("\n") if not l.lstrip().startswith("!")]
        i = 0
        while i < len(lines):
            while i < len(lines) and not lines[i].strip():
                i += 1
            if i >= len(lines):
                break
            pv = {}
            for p in params:
                if "fixed_value" in p:
                    pv[p["name"]] = p["fixed_value"]
                    continue
                pv[p["name"]] = lines[i].strip()
                i += 1
            if mode.get("no_row_counts", "0") == "1":
                rows = []
                while i < len(lines) and lines[i].strip():
                    rows.append(lines[i].split())
                    i += 1
            else:
                n = int(lines[i].split()[0])
                i += 1
                rows = [lines[i + k].split() for k in range(n)]
                i += n
            colv = {c["name"]: np.array([float(r[j]) for r in rows]) for j, c in enumerate(cols)}
            pages.append({"params": pv, "columns": colv})
    return {"params": params, "columns": cols, "pages": pages}


def write_binary(path, columns, pages, params=()):
    """columns: [(name, type)], params: [(name, type)] (no fixed values);
    pages: [(param_values, {name: array})]."""
    with open(path, "wb") as fp:
        fp.write(b"SDDS1\n!# little-endian\n")
        for n, t in params:
            fp.write(("&parameter name=%s, type=%s, &end\n" % (n, t)).encode())
        for n, t in columns:
            fp.write(("&column name=%s, type=%s, &end\n" % (n, t)).encode())
        fp.write(b"&data mode=binary, &end\n")
        for pv, cv in pages:
            rows = len(next(iter(cv.values()))) if cv else 0
            fp.write(struct.pack("<i", rows))
            for (n, t), v in zip(params, pv):
                fp.write(struct.pack("<" + _FMT[t], v))
            dt = np.dtype([(n, "<" + _FMT[t]) for n, t in columns])
            rec = np.zeros(rows, dtype=dt)
            for n, _ in columns:
                rec[n] = cv[n]
            fp.write(rec.tobytes())
