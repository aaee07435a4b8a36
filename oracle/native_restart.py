"""CPU restatement of NATIVE_RESTART.patch
(src/model/climate/gcm/checkpoint/native/index.ts:3-46): TEST
INFRASTRUCTURE ONLY (the tests compare the engine's exo_native_patch with
it; nothing shipped imports it).

A native PlaSim restart file is a run of Fortran unformatted sequential
records: uint32 byte count (little-endian), the bytes, the count again.
Each field is a 16-byte blank-padded name record and a values record (FP32
or FP64). patch() overwrites the values of every named field and reports
the reference's errors.
"""
import struct


class NativeRestartError(ValueError):
    pass


def _record(buf, offset):
    # native/index.ts:9-20
    if offset + 8 > len(buf):
        raise NativeRestartError("Truncated native restart record")
    (nbytes,) = struct.unpack_from("<I", buf, offset)
    start = offset + 4
    if start + nbytes + 4 > len(buf) or struct.unpack_from("<I", buf, start + nbytes)[0] != nbytes:
        raise NativeRestartError("Invalid native restart record")
    return start + nbytes + 4, start, nbytes


def patch(buffer, fields):
    """buffer: bytes; fields: {name: sequence of floats} (insertion order
    is the order missing names are listed in)."""
    result = bytearray(buffer)
    pending = list(fields)
    offset = 0
    while offset < len(result):  # native/index.ts:21-41
        offset, start, nbytes = _record(result, offset)
        if nbytes != 16:
            raise NativeRestartError("Invalid native restart field name")
        name = result[start:start + 16].decode("utf-8").strip()
        offset, dstart, dbytes = _record(result, offset)
        values = fields.get(name)
        if values is None:
            continue
        precision = dbytes / len(values) if len(values) else float("inf")
        if precision not in (4, 8):
            raise NativeRestartError(f"Native restart shape mismatch for {name}")
        fmt = "<d" if precision == 8 else "<f"
        for i, v in enumerate(values):
            struct.pack_into(fmt, result, dstart + i * int(precision), float(v))
        if name in pending:
            pending.remove(name)
    if pending:
        raise NativeRestartError(f"Native restart fields missing: {', '.join(pending)}")
    return bytes(result)
