"""Reader for the binary record dumps written by the oracle programs
(oracle/dump_format.hpp): u32 name length, name, u8 kind (0 f64, 1 i32),
u64 count, payload."""
import struct
import numpy as np


def read_dump(path):
    out = {}
    with open(path, "rb") as f:
        data = f.read()
    pos = 0
    while pos < len(data):
        (l,) = struct.unpack_from("<I", data, pos); pos += 4
        name = data[pos:pos + l].decode(); pos += l
        kind = data[pos]; pos += 1
        (n,) = struct.unpack_from("<Q", data, pos); pos += 8
        if kind == 0:
            out[name] = np.frombuffer(data, dtype="<f8", count=n, offset=pos).copy(); pos += 8 * n
        else:
            out[name] = np.frombuffer(data, dtype="<i4", count=n, offset=pos).copy(); pos += 4 * n
    return out
