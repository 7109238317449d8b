"""PNG files: dgpu_read_png and dgpu_write_png (host code of libdortgpu.so, no GPU
needed) against dort's read_image_8 / write_image_png (src/dort/image.cpp, stb_image),
compiled from the reference checkout into oracle/_ref, and against what the format
fixes independently of both:

* the reader, on PNGs written here with Python's zlib in every colour type, bit depth,
  filter type, zlib level (stored, fixed and dynamic Huffman blocks), with Adam7
  interlacing, several IDAT chunks and ancillary chunks; the expected 8-bit RGB follows
  stb's conversion (16-bit: high byte; 1/2/4-bit grey times 255/85/17; palette lookup;
  grey replicated; alpha and tRNS dropped), and dort's reader must agree with it where
  its stb (v2.10) reads the file correctly: it refuses 16-bit samples, and it misreads
  1/2/4-bit rows filtered with up, average or Paeth;
* the writer, by parsing its files here: chunk CRCs (zlib.crc32), the zlib stream
  (zlib.decompress checks the deflate data and the Adler-32), and unfiltering the rows
  in numpy must give back the image; dort's reader must read the same pixels."""
import struct
import zlib

import numpy as np
import pytest

from dort_amd import gpu
from oracle import ref

SIG = b"\x89PNG\r\n\x1a\n"
CHANNELS = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}
ADAM7 = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)]


def _chunk(kind, data):
    return struct.pack(">I", len(data)) + kind + data + struct.pack(">I", zlib.crc32(kind + data) & 0xFFFFFFFF)


def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    return a if pa <= pb and pa <= pc else (b if pb <= pc else c)


def _filter_row(t, row, prior, bpp):
    out = bytearray(len(row))
    for i in range(len(row)):
        a = row[i - bpp] if i >= bpp else 0
        b = prior[i]
        c = prior[i - bpp] if i >= bpp else 0
        pred = [0, a, b, (a + b) >> 1, _paeth(a, b, c)][t]
        out[i] = (row[i] - pred) & 0xFF
    return bytes(out)


def _pack_row(samples, depth):
    """One scanline of unsigned samples (already in channel order) at `depth` bits."""
    if depth == 8:
        return bytes(int(v) for v in samples)
    if depth == 16:
        return b"".join(struct.pack(">H", int(v)) for v in samples)
    bits = "".join(format(int(v), f"0{depth}b") for v in samples)
    bits += "0" * (-len(bits) % 8)
    return bytes(int(bits[k:k + 8], 2) for k in range(0, len(bits), 8))


def encode_png(samples, ctype, depth, interlace=False, level=6, filters=(0, 1, 2, 3, 4), palette=None,
               idat_parts=1, extra_chunks=()):
    """A PNG of `samples` (H, W, channels) written here, independently of the library."""
    h, w, ch = samples.shape
    bpp = max(1, ch * depth // 8)
    raw = bytearray()
    k = 0
    passes = ADAM7 if interlace else [(0, 0, 1, 1)]
    for x0, y0, dx, dy in passes:
        sub = samples[y0::dy, x0::dx]
        if sub.shape[0] == 0 or sub.shape[1] == 0:
            continue
        prior = bytes(len(_pack_row(sub[0].reshape(-1), depth)))
        for r in range(sub.shape[0]):
            row = _pack_row(sub[r].reshape(-1), depth)
            t = filters[k % len(filters)]
            k += 1
            raw += bytes([t]) + _filter_row(t, row, prior, bpp)
            prior = row
    z = zlib.compress(bytes(raw), level)
    out = SIG + _chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, 1 if interlace else 0))
    for kind, data in extra_chunks:
        out += _chunk(kind, data)
    if palette is not None:
        out += _chunk(b"PLTE", bytes(np.asarray(palette, np.uint8).reshape(-1)))
    step = -(-len(z) // idat_parts)
    for s in range(0, len(z), step):
        out += _chunk(b"IDAT", z[s:s + step])
    return out + _chunk(b"IEND", b"")


def expected_rgb(samples, ctype, depth, palette=None):
    """stb_image's 8-bit, 3-channel result for these samples."""
    s = samples.astype(np.int64)
    if depth == 16:
        s = s >> 8
    elif depth < 8 and ctype == 0:
        s = s * {1: 255, 2: 85, 4: 17}[depth]
    if ctype == 3:
        return np.asarray(palette, np.uint8)[s[..., 0]]
    if ctype in (0, 4):
        return np.repeat(s[..., :1], 3, axis=2).astype(np.uint8)
    return s[..., :3].astype(np.uint8)


CASES = [
    # (colour type, bit depth)
    (0, 1), (0, 2), (0, 4), (0, 8), (0, 16), (2, 8), (2, 16), (3, 1), (3, 2), (3, 4), (3, 8),
    (4, 8), (4, 16), (6, 8), (6, 16),
]


def _samples(rng, h, w, ctype, depth, npal=None):
    top = npal if ctype == 3 else (1 << depth)
    s = rng.integers(0, top, size=(h, w, CHANNELS[ctype]))
    s[h // 3] = s[h // 3, :1]  # a flat row: long matches
    return s


@pytest.mark.parametrize("ctype,depth", CASES)
@pytest.mark.parametrize("interlace", [False, True], ids=["flat", "adam7"])
def test_read_png_formats(tmp_path, ctype, depth, interlace):
    rng = np.random.default_rng(ctype * 100 + depth + (7 if interlace else 0))
    h, w = 13, 29  # odd sizes: partial bytes, short and empty Adam7 passes
    npal = min(1 << depth, 200) if ctype == 3 else None
    palette = rng.integers(0, 256, size=(npal, 3)) if ctype == 3 else None
    s = _samples(rng, h, w, ctype, depth, npal)
    path = tmp_path / "f.png"
    path.write_bytes(encode_png(s, ctype, depth, interlace=interlace, palette=palette))
    want = expected_rgb(s, ctype, depth, palette)
    got = gpu.read_png(str(path))
    assert got.shape == (h, w, 3) and got.dtype == np.uint8
    assert np.array_equal(got, want)
    if ref.available():
        if depth < 8:
            # stb v2.10 (dort's reader) takes the previous row of a 1/2/4-bit image from
            # the wrong offset of its buffer (stb_image.h:4000-4011), so the up, average
            # and Paeth filters give undefined values there: compare with dort on rows
            # filtered with none and sub only
            path.write_bytes(encode_png(s, ctype, depth, interlace=interlace, palette=palette, filters=(0, 1)))
            assert np.array_equal(gpu.read_png(str(path)), want)
        if depth == 16:
            # stb v2.10 reads 1/2/4/8-bit PNGs only (stb_image.h:4349); the library also
            # reads 16-bit ones, keeping each sample's high byte as later stb versions do
            with pytest.raises(ValueError):
                ref.read_png(str(path))
        else:
            assert np.array_equal(ref.read_png(str(path)), want)


@pytest.mark.parametrize("level", [0, 1, 6, 9])
@pytest.mark.parametrize("filters", [(0,), (1,), (2,), (3,), (4,), (4, 3, 2, 1, 0)], ids=lambda f: "f" + "".join(map(str, f)))
def test_read_png_filters_and_blocks(tmp_path, level, filters):
    """Every filter type alone and mixed, with stored (level 0), fixed and dynamic Huffman
    blocks, the stream split over 3 IDAT chunks, ancillary chunks before the data."""
    rng = np.random.default_rng(level * 10 + len(filters))
    h, w = 40, 67
    y, x = np.mgrid[0:h, 0:w]
    s = np.stack([(x * 3 + y) % 256, (y * 5) % 256, rng.integers(0, 256, size=(h, w))], axis=2)
    extra = [(b"gAMA", struct.pack(">I", 45455)), (b"tEXt", b"Comment\x00dort_amd test")]
    path = tmp_path / "f.png"
    path.write_bytes(encode_png(s, 2, 8, level=level, filters=filters, idat_parts=3, extra_chunks=extra))
    assert np.array_equal(gpu.read_png(str(path)), s.astype(np.uint8))
    if ref.available():
        assert np.array_equal(ref.read_png(str(path)), s.astype(np.uint8))


def test_read_png_transparency_dropped(tmp_path):
    """tRNS and alpha do not change the RGB that stb returns for 3 channels."""
    rng = np.random.default_rng(5)
    pal = rng.integers(0, 256, size=(16, 3))
    s = rng.integers(0, 16, size=(9, 11, 1))
    path = tmp_path / "p.png"
    data = bytearray(encode_png(s, 3, 4, palette=pal, filters=(0, 1)))
    # insert a tRNS chunk after PLTE
    at = data.index(b"IDAT") - 4
    data[at:at] = _chunk(b"tRNS", bytes(range(16)))
    path.write_bytes(bytes(data))
    assert np.array_equal(gpu.read_png(str(path)), expected_rgb(s, 3, 4, pal))
    if ref.available():
        assert np.array_equal(ref.read_png(str(path)), expected_rgb(s, 3, 4, pal))


def _parse_written(data):
    """Chunks of a PNG written by the library: CRCs checked, IDAT zlib-decompressed
    (Adler-32 checked), rows unfiltered; returns the (H, W, 3) image."""
    assert data[:8] == SIG
    p, idat, hdr = 8, b"", None
    kinds = []
    while p < len(data):
        (n,) = struct.unpack(">I", data[p:p + 4])
        kind, body = data[p + 4:p + 8], data[p + 8:p + 8 + n]
        (crc,) = struct.unpack(">I", data[p + 8 + n:p + 12 + n])
        assert crc == zlib.crc32(kind + body) & 0xFFFFFFFF, kind
        kinds.append(kind)
        if kind == b"IHDR":
            hdr = struct.unpack(">IIBBBBB", body)
        elif kind == b"IDAT":
            idat += body
        p += 12 + n
    assert kinds[0] == b"IHDR" and kinds[-1] == b"IEND"
    w, h, depth, ctype, comp, filt, inter = hdr
    assert (depth, ctype, comp, filt, inter) == (8, 2, 0, 0, 0)
    raw = zlib.decompress(idat)
    stride = 3 * w
    assert len(raw) == h * (stride + 1)
    img = np.zeros((h, stride), np.int64)
    prior = np.zeros(stride, np.int64)
    for y in range(h):
        t = raw[y * (stride + 1)]
        row = np.frombuffer(raw, np.uint8, stride, y * (stride + 1) + 1).astype(np.int64)
        out = np.zeros(stride, np.int64)
        for i in range(stride):
            a = out[i - 3] if i >= 3 else 0
            b = prior[i]
            c = prior[i - 3] if i >= 3 else 0
            pred = [0, a, b, (a + b) >> 1, _paeth(a, b, c)][t]
            out[i] = (row[i] + pred) & 0xFF
        img[y] = out
        prior = out
    return img.reshape(h, w, 3).astype(np.uint8)


def _images():
    rng = np.random.default_rng(11)
    y, x = np.mgrid[0:64, 0:96]
    yield "noise", rng.integers(0, 256, size=(31, 45, 3)).astype(np.uint8)
    yield "gradient", np.stack([x * 2 % 256, y * 4 % 256, (x + y) % 256], axis=2).astype(np.uint8)
    yield "flat", np.full((20, 300, 3), 77, np.uint8)
    yield "1x1", np.array([[[1, 2, 3]]], np.uint8)
    yield "row", rng.integers(0, 4, size=(1, 5000, 3)).astype(np.uint8)
    yield "column", rng.integers(0, 256, size=(700, 1, 3)).astype(np.uint8)
    img = np.tile(rng.integers(0, 256, size=(1, 40, 3)), (50, 30, 1)).astype(np.uint8)  # matches of length 258
    yield "repeats", img


@pytest.mark.parametrize("name,img", list(_images()), ids=[n for n, _ in _images()])
def test_write_png(tmp_path, name, img):
    path = str(tmp_path / "w.png")
    gpu.write_png(path, img)
    data = open(path, "rb").read()
    assert np.array_equal(_parse_written(data), img)
    assert np.array_equal(gpu.read_png(path), img)
    if ref.available():
        assert np.array_equal(ref.read_png(path), img)
        theirs = str(tmp_path / "stb.png")
        ref.write_png(theirs, img)
        assert np.array_equal(gpu.read_png(theirs), img)


def test_png_errors(tmp_path):
    p = tmp_path / "bad.png"
    with pytest.raises(gpu.DgpuError):
        gpu.read_png(str(tmp_path / "missing.png"))
    p.write_bytes(b"GIF89a")
    with pytest.raises(gpu.DgpuError):
        gpu.read_png(str(p))
    s = np.random.default_rng(3).integers(0, 256, size=(8, 8, 3))
    good = encode_png(s, 2, 8)
    p.write_bytes(good[:len(good) // 2])  # truncated
    with pytest.raises(gpu.DgpuError):
        gpu.read_png(str(p))
    raw = bytearray()
    for r in range(8):
        raw += bytes([7]) + bytes(int(v) for v in s[r].reshape(-1))  # filter type 7
    p.write_bytes(SIG + _chunk(b"IHDR", struct.pack(">IIBBBBB", 8, 8, 8, 2, 0, 0, 0)) +
                  _chunk(b"IDAT", zlib.compress(bytes(raw))) + _chunk(b"IEND", b""))
    with pytest.raises(gpu.DgpuError):
        gpu.read_png(str(p))
    # corrupt deflate data (an invalid block type 3 right after the zlib header)
    p.write_bytes(SIG + _chunk(b"IHDR", struct.pack(">IIBBBBB", 8, 8, 8, 2, 0, 0, 0)) +
                  _chunk(b"IDAT", b"\x78\x01\x07\x00\x00") + _chunk(b"IEND", b""))
    with pytest.raises(gpu.DgpuError):
        gpu.read_png(str(p))
    # bad zlib header, unknown critical chunk, palette image without PLTE, bad depth
    p.write_bytes(good.replace(b"IDAT\x78", b"IDAT\x77", 1))
    with pytest.raises(gpu.DgpuError):
        gpu.read_png(str(p))
    at = good.index(b"IDAT") - 4
    p.write_bytes(good[:at] + _chunk(b"ZZZZ", b"x") + good[at:])
    with pytest.raises(gpu.DgpuError):
        gpu.read_png(str(p))
    p.write_bytes(encode_png(s[..., :1] % 4, 3, 8, palette=None))
    with pytest.raises(gpu.DgpuError):
        gpu.read_png(str(p))
    p.write_bytes(SIG + _chunk(b"IHDR", struct.pack(">IIBBBBB", 8, 8, 3, 2, 0, 0, 0)) + _chunk(b"IEND", b""))
    with pytest.raises(gpu.DgpuError):
        gpu.read_png(str(p))
    with pytest.raises(ValueError):
        gpu.write_png(str(p), np.zeros((4, 4), np.uint8))
