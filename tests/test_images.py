"""Image files: the library's PNG and JPEG decoders (bh_png_decode,
bh_jpeg_decode) with cv::imread IMREAD_COLOR semantics (config.h:35-37), the
PPM / PNG / JPEG frame writer and the Motion-JPEG AVI writer that stands in
for cv::VideoWriter (blackhole_solution_test.cc:69-72,330).

The JPEG decoder is pinned byte for byte against Pillow, whose JPEG decoder is
libjpeg(-turbo), the library imread uses: on the reference's own
resource/*.jpeg textures when the reference tree is present, and on files
Pillow writes over 4:4:4 / 4:2:2 / 4:2:0, greyscale, two qualities, odd sizes,
restart intervals and optimized Huffman tables; progressive files are refused.

The decoder is pinned by
  - Pillow (an independent PNG decoder, where installed) on 8-bit files it
    writes itself and on the reference's own resource/acc_disc.png when the
    reference tree is present;
  - files from the encoder below, written from the PNG specification, over
    every colour type, bit depth, scanline filter and Adam7 interlacing,
    whose expected output follows from the samples it was given (an encoder
    and a decoder share no code, so a wrong filter or pass layout in either
    fails the round trip);
  - corrupt inputs (bad signature, CRC, truncated stream), which must fail.
Runs on the CPU: the codec is host code in the library.
"""
import io
import struct
import zlib
from pathlib import Path

import numpy as np
import pytest

from blackhole_8_amd import abi, images
from tests import riff

ACC_DISC = Path("/root/reference/resource/acc_disc.png")
ADAM7 = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2),
         (0, 1, 1, 2)]
CHANNELS = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}


def _chunk(t: bytes, data: bytes) -> bytes:
    return struct.pack(">I", len(data)) + t + data + struct.pack(">I", zlib.crc32(t + data))


def _pack(samples: np.ndarray, depth: int) -> bytes:
    """One scanline of samples (flat, in file order) packed MSB first."""
    s = samples.astype(np.uint32).ravel()
    if depth == 16:
        return s.astype(">u2").tobytes()
    if depth == 8:
        return s.astype(np.uint8).tobytes()
    per = 8 // depth
    out = bytearray((len(s) + per - 1) // per)
    for k, v in enumerate(s):
        out[k // per] |= int(v) << (8 - depth * (k % per + 1))
    return bytes(out)


def _paeth(a, b, c):
    # PNG specification 9.4: the neighbour nearest to a + b - c, ties a, b, c
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    return a if pa <= pb and pa <= pc else (b if pb <= pc else c)


def _filter(raw: bytes, prev: bytes, ftype: int, bpp: int) -> bytes:
    out = bytearray(len(raw))
    for i, x in enumerate(raw):
        a = raw[i - bpp] if i >= bpp else 0
        b = prev[i] if prev else 0
        c = prev[i - bpp] if (prev and i >= bpp) else 0
        pred = [0, a, b, (a + b) // 2, _paeth(a, b, c)][ftype]
        out[i] = (x - pred) & 0xFF
    return bytes(out)


def encode(samples: np.ndarray, ctype: int, depth: int, interlace=False, filters=(0,),
           palette=None, extra_chunks=()):
    """samples: rows x cols x channels integers < 2**depth (indices for ctype 3)."""
    h, w, ch = samples.shape
    assert ch == CHANNELS[ctype]
    bpp = max(1, ch * depth // 8)
    passes = ADAM7 if interlace else [(0, 0, 1, 1)]
    raw = bytearray()
    k = 0
    for x0, y0, dx, dy in passes:
        sub = samples[y0::dy, x0::dx]
        if sub.size == 0:
            continue
        prev = None
        for row in sub:
            line = _pack(row, depth)
            ft = filters[k % len(filters)]
            k += 1
            raw += bytes([ft]) + _filter(line, prev, ft, bpp)
            prev = line
    out = b"\x89PNG\r\n\x1a\n"
    out += _chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, int(interlace)))
    for t, d in extra_chunks:
        out += _chunk(t, d)
    if palette is not None:
        out += _chunk(b"PLTE", bytes(np.asarray(palette, np.uint8).ravel()))
    z = zlib.compress(bytes(raw), 6)
    out += _chunk(b"IDAT", z[: len(z) // 2]) + _chunk(b"IDAT", z[len(z) // 2:])  # split stream
    out += _chunk(b"IEND", b"")
    return out


def expected_bgr(samples, ctype, depth, palette=None):
    """IMREAD_COLOR: 3 x 8-bit B, G, R; alpha dropped; grey replicated and
    scaled to 0..255 below 8 bits; 16-bit keeps the high byte."""
    s = samples.astype(np.int64)

    def eight(v):
        if depth == 16:
            return v >> 8
        return v * 255 // ((1 << depth) - 1)

    if ctype in (0, 4):
        g = eight(s[..., 0])
        rgb = np.stack([g, g, g], -1)
    elif ctype == 3:
        pal = np.asarray(palette, np.int64)
        idx = s[..., 0]
        rgb = np.where((idx < len(pal))[..., None], pal[np.minimum(idx, len(pal) - 1)], 0)
    else:
        rgb = eight(s[..., :3])
    return rgb[..., ::-1].astype(np.uint8)


CASES = [(0, 1), (0, 2), (0, 4), (0, 8), (0, 16), (2, 8), (2, 16), (3, 1), (3, 2), (3, 4),
         (3, 8), (4, 8), (4, 16), (6, 8), (6, 16)]


@pytest.mark.parametrize("ctype,depth", CASES)
@pytest.mark.parametrize("interlace", [False, True])
def test_decode_all_formats_filters_interlace(ctype, depth, interlace):
    rng = np.random.default_rng(ctype * 100 + depth + 7 * interlace)
    h, w = 13, 19  # odd sizes: partial bytes, short Adam7 passes
    hi = (1 << depth) if ctype != 3 else min(1 << depth, 200)
    samples = rng.integers(0, hi, (h, w, CHANNELS[ctype]))
    pal = rng.integers(0, 256, (min(1 << depth, 256), 3)) if ctype == 3 else None
    if ctype == 3 and depth >= 2:
        pal = pal[: max(1, len(pal) - 1)]  # the top index falls outside: reads black
        samples[0, 0, 0] = (1 << depth) - 1
    data = encode(samples, ctype, depth, interlace, filters=(0, 1, 2, 3, 4), palette=pal)
    got = images.decode_png(data)
    assert got.shape == (h, w, 3)
    np.testing.assert_array_equal(got, expected_bgr(samples, ctype, depth, pal))


@pytest.mark.parametrize("ftype", [0, 1, 2, 3, 4])
def test_each_filter_alone_on_smooth_and_edge_data(ftype):
    # gradients make Sub/Up/Average/Paeth predictions matter; saturated edges
    # exercise the modulo-256 wrap
    h, w = 24, 40
    y, x = np.mgrid[0:h, 0:w]
    r = (x * 7 + y * 3) % 256
    g = np.where((x // 5 + y // 4) % 2 == 0, 255, 0)
    b = (255 - x * 6) % 256
    a = (y * 11) % 256
    samples = np.stack([r, g, b, a], -1)
    got = images.decode_png(encode(samples, 6, 8, filters=(ftype,)))
    np.testing.assert_array_equal(got, expected_bgr(samples, 6, 8))


def test_paeth_predictor_known_answers():
    # the encoder's predictor against hand-worked values (a, b, c) -> pick
    assert _paeth(10, 20, 15) == 15  # p = 15: c exact
    assert _paeth(10, 20, 10) == 20  # p = 20: b
    assert _paeth(20, 10, 10) == 20  # p = 20: a
    assert _paeth(5, 5, 5) == 5
    assert _paeth(0, 255, 255) == 0  # p = 0: a
    assert _paeth(100, 0, 50) == 50  # p = 50: c; pa = pb = 50 > pc = 0


def test_ancillary_chunks_are_skipped_and_gamma_not_applied():
    rng = np.random.default_rng(3)
    samples = rng.integers(0, 256, (8, 9, 3))
    extra = [(b"gAMA", struct.pack(">I", 45455)), (b"tEXt", b"Comment\x00hello"),
             (b"tRNS", b"\x00\x01\x00\x02\x00\x03")]
    got = images.decode_png(encode(samples, 2, 8, extra_chunks=extra))
    np.testing.assert_array_equal(got, expected_bgr(samples, 2, 8))


def test_corrupt_files_fail():
    rng = np.random.default_rng(4)
    good = encode(rng.integers(0, 256, (6, 7, 3)), 2, 8)
    bad_sig = b"\x88" + good[1:]
    bad_crc = bytearray(good)
    bad_crc[33 + 8] ^= 0xFF  # inside the first IDAT's data (IHDR is 8 + 25 bytes)
    truncated = good[: len(good) // 2]
    for data in (bad_sig, bytes(bad_crc), truncated, b"", b"\x89PNG\r\n\x1a\n"):
        with pytest.raises(images.ImageError) as e:
            images.decode_png(data)
        assert e.value.code == abi.BH_ERR_INVALID_ARG
    # an unknown critical chunk is refused, an unknown ancillary one skipped
    samples = rng.integers(0, 256, (4, 4, 3))
    with pytest.raises(images.ImageError) as e:
        images.decode_png(encode(samples, 2, 8, extra_chunks=[(b"ABCD", b"x")]))
    assert e.value.code == abi.BH_ERR_UNSUPPORTED
    images.decode_png(encode(samples, 2, 8, extra_chunks=[(b"abCD", b"x")]))


def test_decode_matches_pillow_on_its_own_files():
    Image = pytest.importorskip("PIL.Image")
    rng = np.random.default_rng(5)
    for mode, shape in (("RGB", (31, 45, 3)), ("RGBA", (17, 23, 4)), ("L", (20, 33)),
                        ("P", (15, 16))):
        arr = rng.integers(0, 256, shape).astype(np.uint8)
        im = Image.fromarray(arr, mode="L" if mode == "P" else mode)
        if mode == "P":
            im = im.convert("P", palette=Image.Palette.ADAPTIVE, colors=37)
        buf = io.BytesIO()
        im.save(buf, format="PNG", optimize=True)
        got = images.decode_png(buf.getvalue())
        want = np.asarray(Image.open(io.BytesIO(buf.getvalue())).convert("RGB"))[..., ::-1]
        np.testing.assert_array_equal(got, want, err_msg=mode)


@pytest.mark.skipif(not ACC_DISC.exists(), reason="reference resources not present")
def test_reference_disc_texture_matches_pillow():
    Image = pytest.importorskip("PIL.Image")
    got = images.load(ACC_DISC)
    im = Image.open(ACC_DISC)
    assert got.shape == (598, 578, 3)  # rows x cols (the texture is 578 wide)
    want = np.asarray(im.convert("RGBA"))[..., :3][..., ::-1]  # alpha dropped, not composited
    np.testing.assert_array_equal(got, want)


def test_write_ppm_and_png_round_trip(tmp_path):
    rng = np.random.default_rng(6)
    bgr = rng.integers(0, 256, (37, 53, 3)).astype(np.uint8)
    for ext in ("ppm", "png"):
        p = tmp_path / f"frame.{ext}"
        images.write(p, bgr)
        np.testing.assert_array_equal(images.load(p), bgr)
    # the files store R, G, B
    raw = (tmp_path / "frame.ppm").read_bytes()
    assert raw.startswith(b"P6\n53 37\n255\n")
    body = np.frombuffer(raw[len(b"P6\n53 37\n255\n"):], np.uint8).reshape(37, 53, 3)
    np.testing.assert_array_equal(body, bgr[..., ::-1])
    Image = pytest.importorskip("PIL.Image")
    np.testing.assert_array_equal(np.asarray(Image.open(tmp_path / "frame.png")), bgr[..., ::-1])


def test_write_rejects_unknown_extension(tmp_path):
    with pytest.raises(images.ImageError) as e:
        images.write(tmp_path / "x.bmp", np.zeros((2, 2, 3), np.uint8))
    assert e.value.code == abi.BH_ERR_UNSUPPORTED


# ---------------------------------------------------------------------------
# JPEG (baseline): against Pillow, whose decoder is libjpeg(-turbo), the
# library imread decodes JPEG with; the decoder is meant to be byte-exact
# ---------------------------------------------------------------------------
REF_JPEGS = [Path("/root/reference/resource") / f for f in ("mooni.jpeg", "karina.jpeg",
                                                              "winter.jpg")]


def _pillow_bgr(data: bytes):
    Image = pytest.importorskip("PIL.Image")
    return np.asarray(Image.open(io.BytesIO(data)).convert("RGB"))[..., ::-1]


def _smooth_image(h, w, seed, channels=3):
    # gradients plus noise: smooth areas exercise the DC path, noise the AC
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w]
    base = np.stack([(x * 255 // max(1, w - 1)), (y * 255 // max(1, h - 1)),
                     ((x + y) * 127 // max(1, w + h - 2))], -1)[..., :channels]
    return np.clip(base + rng.integers(-40, 41, base.shape), 0, 255).astype(np.uint8)


@pytest.mark.skipif(not all(p.exists() for p in REF_JPEGS), reason="reference resources absent")
@pytest.mark.parametrize("path", REF_JPEGS, ids=lambda p: p.name)
def test_reference_jpegs_match_libjpeg(path):
    data = path.read_bytes()
    got = images.decode_jpeg(data)
    np.testing.assert_array_equal(got, _pillow_bgr(data))
    np.testing.assert_array_equal(images.load(path), got)


def test_jpeg_narrow_images_match_libjpeg():
    # libjpeg upsamples 2:1 chroma with the triangle filter only when the
    # chroma plane is more than 2 samples wide, else by replication: every
    # width 1..9 and height 1..9 over 4:4:4 / 4:2:2 / 4:2:0
    Image = pytest.importorskip("PIL.Image")
    for h in range(1, 10):
        for w in range(1, 10):
            arr = _smooth_image(h, w, 31 * h + w)
            for sub in (0, 1, 2):
                buf = io.BytesIO()
                Image.fromarray(arr, "RGB").save(buf, "JPEG", quality=90, subsampling=sub)
                data = buf.getvalue()
                np.testing.assert_array_equal(images.decode_jpeg(data), _pillow_bgr(data),
                                              err_msg=f"{h}x{w} subsampling {sub}")


@pytest.mark.parametrize("subsampling", [0, 1, 2])  # 4:4:4, 4:2:2, 4:2:0
@pytest.mark.parametrize("size", [(1, 1), (7, 13), (16, 16), (37, 53), (120, 97)])
@pytest.mark.parametrize("quality", [35, 95])
def test_jpeg_matches_libjpeg(subsampling, size, quality):
    Image = pytest.importorskip("PIL.Image")
    h, w = size
    arr = _smooth_image(h, w, h * 1000 + w + quality)
    buf = io.BytesIO()
    Image.fromarray(arr, "RGB").save(buf, "JPEG", quality=quality, subsampling=subsampling)
    data = buf.getvalue()
    np.testing.assert_array_equal(images.decode_jpeg(data), _pillow_bgr(data))


def test_jpeg_grey_restart_and_optimized_tables():
    Image = pytest.importorskip("PIL.Image")
    arr = _smooth_image(45, 61, 9)
    cases = [(Image.fromarray(arr[..., 0], "L"), {}),
             (Image.fromarray(arr, "RGB"), {"optimize": True}),
             (Image.fromarray(arr, "RGB"), {"restart_marker_blocks": 3, "subsampling": 2}),
             (Image.fromarray(arr, "RGB"), {"restart_marker_blocks": 1, "subsampling": 0})]
    for im, kw in cases:
        buf = io.BytesIO()
        im.save(buf, "JPEG", quality=80, **kw)
        data = buf.getvalue()
        if kw.get("restart_marker_blocks"):
            assert b"\xff\xd0" in data
        got = images.decode_jpeg(data)
        np.testing.assert_array_equal(got, _pillow_bgr(data), err_msg=str(kw))
        if im.mode == "L":
            assert (got[..., 0] == got[..., 1]).all() and (got[..., 1] == got[..., 2]).all()


def test_jpeg_refuses_progressive_and_corrupt(tmp_path):
    Image = pytest.importorskip("PIL.Image")
    arr = _smooth_image(32, 32, 3)
    buf = io.BytesIO()
    Image.fromarray(arr, "RGB").save(buf, "JPEG", progressive=True)
    with pytest.raises(images.ImageError) as e:
        images.decode_jpeg(buf.getvalue())
    assert e.value.code == abi.BH_ERR_UNSUPPORTED
    buf = io.BytesIO()
    Image.fromarray(arr, "RGB").save(buf, "JPEG")
    good = buf.getvalue()
    for bad in (b"", good[:2], b"\xff\xd8\x00" + good[3:], good[:200]):
        with pytest.raises(images.ImageError):
            images.decode_jpeg(bad)
    # files are told apart by signature
    p = tmp_path / "t.bin"
    p.write_bytes(good)
    np.testing.assert_array_equal(images.load(p), images.decode_jpeg(good))


def test_jpeg_solid_colour_known_answer():
    # a flat 8x8-aligned image has only DC terms: every pixel decodes to the
    # colour the encoder quantised it to, the same for all pixels
    Image = pytest.importorskip("PIL.Image")
    arr = np.zeros((32, 48, 3), np.uint8)
    arr[...] = (200, 40, 90)
    buf = io.BytesIO()
    Image.fromarray(arr, "RGB").save(buf, "JPEG", quality=100, subsampling=0)
    got = images.decode_jpeg(buf.getvalue())
    assert (got == got[0, 0]).all()
    np.testing.assert_allclose(got[0, 0].astype(int), [90, 40, 200], atol=3)


# ---------------------------------------------------------------------------
# JPEG encoder (frames of the MJPG video) and the AVI container
# ---------------------------------------------------------------------------
def _psnr(a, b):
    return 10 * np.log10(255.0 ** 2 / np.mean((a.astype(np.float64) - b) ** 2))


@pytest.mark.parametrize("size", [(1, 1), (2, 3), (9, 17), (16, 16), (33, 2), (45, 61),
                                  (270, 480)])
@pytest.mark.parametrize("quality", [10, 75, 95, 100])
def test_jpeg_encoder_decodes_under_libjpeg(size, quality):
    # libjpeg reads every file the encoder writes to the same pixels as the
    # library's own decoder, and the error is what libjpeg's own encoder
    # gets at the same quality and 4:2:0 (within 0.5 dB)
    Image = pytest.importorskip("PIL.Image")
    h, w = size
    arr = _smooth_image(h, w, h * 7 + w + quality)
    bgr = arr[..., ::-1].copy()
    data = images.encode_jpeg(bgr, quality)
    assert data[:2] == b"\xff\xd8" and data[-2:] == b"\xff\xd9"
    got = images.decode_jpeg(data)
    np.testing.assert_array_equal(got, _pillow_bgr(data))
    buf = io.BytesIO()
    Image.fromarray(arr, "RGB").save(buf, "JPEG", quality=quality, subsampling=2)
    theirs = _pillow_bgr(buf.getvalue())
    if h * w >= 64:
        assert _psnr(got, bgr) >= _psnr(theirs, bgr) - 0.5


def test_jpeg_encoder_structure_and_known_answers():
    # a flat image holds only DC terms: every pixel decodes to one colour,
    # the input colour within the DC quantiser step
    flat = np.zeros((48, 80, 3), np.uint8)
    flat[...] = (90, 40, 200)
    got = images.decode_jpeg(images.encode_jpeg(flat, 95))
    assert (got == got[0, 0]).all()
    np.testing.assert_allclose(got[0, 0].astype(int), [90, 40, 200], atol=3)
    # one restart interval per MCU row: DRI = MCUs per row, RST0..7 cycling
    arr = _smooth_image(100, 70, 5)[..., ::-1].copy()
    data = images.encode_jpeg(arr, 80)
    i = data.index(b"\xff\xdd")
    assert struct.unpack(">HH", data[i + 2:i + 6]) == (4, (70 + 15) // 16)
    sos = data.index(b"\xff\xda")
    body = data[sos:]
    rst = [body[k + 1] for k in range(len(body) - 1)
           if body[k] == 0xFF and 0xD0 <= body[k + 1] <= 0xD7]
    assert rst == [0xD0 + (k & 7) for k in range((100 + 15) // 16 - 1)]
    # quality is clamped to 1..100; higher quality never loses to lower
    p = [_psnr(images.decode_jpeg(images.encode_jpeg(arr, q)), arr) for q in (1, 30, 70, 100)]
    assert p == sorted(p)
    with pytest.raises(ValueError):
        images.encode_jpeg(np.zeros((4, 4), np.uint8))


def test_write_jpeg(tmp_path):
    bgr = _smooth_image(37, 53, 4)
    for ext in ("jpg", "jpeg"):
        p = tmp_path / f"frame.{ext}"
        images.write(p, bgr)
        assert p.read_bytes() == images.encode_jpeg(bgr, 95)  # imwrite's default quality
        np.testing.assert_array_equal(images.load(p), images.decode_jpeg(p.read_bytes()))


@pytest.mark.parametrize("fps,scale,rate", [(29, 1, 29), (29.97, 1000, 29970)])
def test_mjpg_avi_layout(tmp_path, fps, scale, rate):
    h, w, n = 36, 50, 5
    frames = [_smooth_image(h, w, 100 + k)[..., ::-1].copy() for k in range(n)]
    path = tmp_path / "video.avi"
    with images.Video(path, h, w, fps=fps) as v:
        for f in frames:
            v.write(f)
    data = path.read_bytes()
    (top,) = riff.parse(data)
    assert top[0] == b"RIFF:AVI "
    assert struct.unpack("<I", data[4:8])[0] == len(data) - 8
    hdrl, movi, idx1 = top[1]
    assert hdrl[0] == b"LIST:hdrl" and movi[0] == b"LIST:movi" and idx1[0] == b"idx1"
    avih, strl = hdrl[1]
    us, _, _, flags, total, _, streams, _, aw, ah = struct.unpack("<10I", avih[1][:40])
    assert (total, streams, aw, ah) == (n, 1, w, h) and flags & 0x10
    assert us == round(1e6 / fps)
    strh, strf = strl[1]
    assert strh[1][:8] == b"vidsMJPG"
    assert struct.unpack("<II", strh[1][20:28]) == (scale, rate)
    assert struct.unpack("<I", strh[1][32:36])[0] == n
    assert struct.unpack("<4h", strh[1][48:56]) == (0, 0, w, h)
    bi = struct.unpack("<IiiHH4s", strf[1][:20])
    assert bi == (40, w, h, 1, 24, b"MJPG")
    chunks = movi[1]
    assert [c[0] for c in chunks] == [b"00dc"] * n
    movi_tag = movi[2] + 8  # offset of the 'movi' fourcc
    entries = [struct.unpack("<4sIII", idx1[1][16 * k:16 * k + 16]) for k in range(n)]
    for (tag, payload, off), (etag, eflags, eoff, esize), f in zip(chunks, entries, frames):
        assert (etag, eflags, eoff, esize) == (b"00dc", 0x10, off - movi_tag, len(payload))
        assert payload == images.encode_jpeg(f, 95)
        np.testing.assert_array_equal(_pillow_bgr(payload), images.decode_jpeg(payload))


def test_video_errors(tmp_path):
    with pytest.raises(images.ImageError):
        images.Video(tmp_path / "a.avi", 10, 10, fps=0)
    with pytest.raises(images.ImageError):
        images.Video(tmp_path / "a.avi", 10, 10, quality=0)
    with pytest.raises(images.ImageError):
        images.Video(tmp_path / "no_such_dir" / "a.avi", 10, 10)
    v = images.Video(tmp_path / "b.avi", 10, 12)
    with pytest.raises(ValueError):
        v.write(np.zeros((12, 10, 3), np.uint8))
    v.close()
    v.close()  # closing twice is harmless
    with pytest.raises(ValueError):
        v.write(np.zeros((10, 12, 3), np.uint8))
    # an empty video is still a well-formed file
    (top,) = riff.parse((tmp_path / "b.avi").read_bytes())
    assert [c[0] for c in top[1]] == [b"LIST:hdrl", b"LIST:movi", b"idx1"]
