"""RIFF reader for the AVI tests: the file as nested (tag, body, offset)."""
import struct


def parse(data: bytes):
    """Top-level chunks of a RIFF file.  RIFF and LIST chunks become
    (b"RIFF:<form>" / b"LIST:<type>", [children], offset); others become
    (fourcc, payload bytes, offset).  Every size must fit its parent and
    chunks are padded to even length."""
    def walk(off, end):
        out = []
        while off + 8 <= end:
            tag, size = data[off:off + 4], struct.unpack("<I", data[off + 4:off + 8])[0]
            assert off + 8 + size <= end, (tag, size)
            if tag in (b"RIFF", b"LIST"):
                out.append((tag + b":" + data[off + 8:off + 12], walk(off + 12, off + 8 + size),
                            off))
            else:
                out.append((tag, data[off + 8:off + 8 + size], off))
            off += 8 + size + (size & 1)
        assert off == end
        return out
    return walk(0, len(data))


def avi_frames(data: bytes):
    """(frame count in avih, [JPEG payloads of the '00dc' chunks]) of an AVI."""
    (top,) = parse(data)
    assert top[0] == b"RIFF:AVI "
    hdrl = next(c for c in top[1] if c[0] == b"LIST:hdrl")
    movi = next(c for c in top[1] if c[0] == b"LIST:movi")
    total = struct.unpack("<I", hdrl[1][0][1][16:20])[0]
    return total, [c[1] for c in movi[1] if c[0] == b"00dc"]
