"""Oracle (test infrastructure only): the cell-parameter file of Setka::Save_cell_parameters and
Setka::Download_cell_parameters (TITAN/source/Setka_physics.cpp:4107-4285), restated with
`struct` record by record.

write() follows Save_cell_parameters: per cell (All_Cell order) the number of keys as size_t and
each (key length as size_t, key bytes, double value) in the order the cell's dict gives (the
reference iterates an unordered_map, whose order is unspecified); then bool true, the node count
as size_t and every node's coord[0] (x, y, z); then 1000 bool false.  read() follows
Download_cell_parameters: keys listed in `mk_keys` come back as 0.0 (the reference resets the
Monte-Carlo moments, Setka_physics.cpp:4233-4237); the node block is read when its bool is true.
Little-endian, size_t of 8 bytes and bool of 1 byte: the reference's x86-64 build.

write_mk() / read_mk() restate the Monte-Carlo parameter file of Setka::Save_cell_MK_parameters and
Setka::Download_cell_MK_parameters (Setka_physics.cpp:4165-4209, 4287-4336): the key count as
size_t and each key once (length as size_t, bytes), then per cell one double per key in key order,
then 1000 bool false.  The reader assigns every key of the file to every cell whose MK zone is not
`zone_except`.
"""
import struct


def write(path, cells, xyz=None, trailer=1000):
    with open(path, "wb") as f:
        for kv in cells:
            f.write(struct.pack("<Q", len(kv)))
            for k, v in kv.items():
                kb = k.encode()
                f.write(struct.pack("<Q", len(kb)))
                f.write(kb)
                f.write(struct.pack("<d", float(v)))
        if xyz is not None:
            f.write(struct.pack("<?", True))
            f.write(struct.pack("<Q", len(xyz)))
            for x, y, z in xyz:
                f.write(struct.pack("<ddd", x, y, z))
        else:
            f.write(struct.pack("<?", False))
        f.write(b"\x00" * trailer)


def read(path, n_cells, mk_keys=()):
    """Returns (cells: list of dicts, xyz: list of (x, y, z) or None)."""
    with open(path, "rb") as f:
        data = f.read()
    pos = 0

    def take(fmt):
        nonlocal pos
        size = struct.calcsize(fmt)
        if pos + size > len(data):
            raise EOFError("truncated at byte %d" % pos)
        out = struct.unpack_from(fmt, data, pos)
        pos += size
        return out

    cells = []
    for _ in range(n_cells):
        (cnt,) = take("<Q")
        kv = {}
        for _ in range(cnt):
            (ks,) = take("<Q")
            if pos + ks > len(data):
                raise EOFError("truncated key at byte %d" % pos)
            key = data[pos:pos + ks].decode()
            pos += ks
            (v,) = take("<d")
            kv[key] = 0.0 if key in mk_keys else v
        cells.append(kv)
    (follow,) = take("<?")
    xyz = None
    if follow:
        (nn,) = take("<Q")
        xyz = [take("<ddd") for _ in range(nn)]
    return cells, xyz


def write_mk(path, keys, values, trailer=1000):
    """values: per cell a sequence of one float per key (Setka_physics.cpp:4173-4208)."""
    with open(path, "wb") as f:
        f.write(struct.pack("<Q", len(keys)))
        for k in keys:
            kb = k.encode()
            f.write(struct.pack("<Q", len(kb)))
            f.write(kb)
        for row in values:
            for v in row:
                f.write(struct.pack("<d", float(v)))
        f.write(b"\x00" * trailer)


def read_mk(path, cells, cell_zone=None, zone_except=-1):
    """Updates `cells` (list of dicts) in place, Setka_physics.cpp:4296-4331: every key of the file
    for each cell whose zone (cell_zone[c], or none) is not zone_except.  Returns the file's keys."""
    with open(path, "rb") as f:
        data = f.read()
    (cnt,) = struct.unpack_from("<Q", data, 0)
    pos = 8
    keys = []
    for _ in range(cnt):
        (ks,) = struct.unpack_from("<Q", data, pos)
        pos += 8
        if pos + ks > len(data):
            raise EOFError("truncated key at byte %d" % pos)
        keys.append(data[pos:pos + ks].decode())
        pos += ks
    for c, kv in enumerate(cells):
        for k in keys:
            (v,) = struct.unpack_from("<d", data, pos)
            pos += 8
            if cell_zone is None or cell_zone[c] != zone_except:
                kv[k] = v
    return keys
