"""Scene and image files: dgpu_read_ply, dgpu_read_rgbe, dgpu_write_rgbe (host code of
libdortgpu.so, no GPU needed) against dort's own readers and writers (ply_mesh.cpp over
rply, image.cpp over stb), compiled from the reference checkout into oracle/_ref, and
against what the formats fix: the fan triangulation of each face, the RGBE rule
m * 2^(e - 136) and Ward's encoding of the largest channel's mantissa."""
import glob
import os
import struct

import numpy as np
import pytest

from dort_amd import abi, gpu
from oracle import ref, rgbe

HERE = os.path.dirname(os.path.abspath(__file__))

_FMT = {"char": "b", "uchar": "B", "short": "h", "ushort": "H", "int": "i", "uint": "I",
        "float": "f", "double": "d", "int8": "b", "uint8": "B", "int32": "i", "float32": "f"}


def write_ply(path, fmt, verts, faces, vprops=("x", "y", "z"), vtypes=None, count_type="uchar",
              index_type="int", extra_face_prop=False, extra_element=False):
    """A PLY writer for the tests: verts (V, len(vprops)) values, faces: lists of indices."""
    vtypes = vtypes or ["float"] * len(vprops)
    lines = ["ply", f"format {fmt} 1.0", "comment written by the dort_amd tests",
             f"element vertex {len(verts)}"]
    lines += [f"property {t} {p}" for p, t in zip(vprops, vtypes)]
    lines += [f"element face {len(faces)}", f"property list {count_type} {index_type} vertex_indices"]
    if extra_face_prop:
        lines += ["property uchar flags"]
    if extra_element:
        lines += ["element edge 2", "property int vertex1", "property int vertex2"]
    lines += ["end_header"]
    head = ("\n".join(lines) + "\n").encode()
    with open(path, "wb") as f:
        f.write(head)
        if fmt == "ascii":
            body = []
            for v in verts:
                body.append(" ".join(repr(float(x)) if t in ("float", "double", "float32") else str(int(x))
                                     for x, t in zip(v, vtypes)))
            for fc in faces:
                body.append(" ".join(str(x) for x in [len(fc)] + list(fc)) + (" 7" if extra_face_prop else ""))
            if extra_element:
                body += ["0 1", "1 2"]
            f.write(("\n".join(body) + "\n").encode())
        else:
            e = "<" if fmt == "binary_little_endian" else ">"
            for v in verts:
                for x, t in zip(v, vtypes):
                    f.write(struct.pack(e + _FMT[t], x if _FMT[t] in "fd" else int(x)))
            for fc in faces:
                f.write(struct.pack(e + _FMT[count_type], len(fc)))
                for x in fc:
                    f.write(struct.pack(e + _FMT[index_type], x))
                if extra_face_prop:
                    f.write(struct.pack(e + "B", 7))
            if extra_element:
                f.write(struct.pack(e + "iiii", 0, 1, 1, 2))


def fan(faces):
    out = []
    for fc in faces:
        for k in range(2, len(fc)):
            out.append((fc[0], fc[k - 1], fc[k]))
    return np.array(out, np.uint32).reshape(-1, 3)


def _mesh(seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(40, 3)).astype(np.float32)
    faces = []
    for n in (3, 4, 5, 3, 6, 4, 3):
        faces.append(list(rng.choice(40, size=n, replace=False)))
    faces.append([5, 6])  # a degenerate face: no triangle
    return v, faces


@pytest.mark.parametrize("fmt", ["ascii", "binary_little_endian", "binary_big_endian"])
@pytest.mark.parametrize("variant", ["plain", "extra"])
def test_read_ply(tmp_path, fmt, variant):
    v, faces = _mesh(3)
    path = str(tmp_path / "m.ply")
    if variant == "plain":
        write_ply(path, fmt, v, faces)
    else:
        # extra vertex properties between and after the coordinates, other index types,
        # a face property after the list and an element the readers skip
        rng = np.random.default_rng(4)
        cols = np.concatenate([rng.normal(size=(40, 1)), v[:, :1], rng.integers(0, 255, (40, 1)), v[:, 1:],
                               rng.normal(size=(40, 1))], axis=1)
        write_ply(path, fmt, cols, faces, vprops=("nx", "x", "red", "y", "z", "u"),
                  vtypes=["float", "float", "uchar", "float", "float", "double"], count_type="uchar",
                  index_type="uint", extra_face_prop=True, extra_element=True)
    p, t = gpu.read_ply(path)
    assert np.array_equal(p, v)
    assert np.array_equal(t, fan(faces))
    if ref.available():
        rp, rt = ref.read_ply(path)
        assert np.array_equal(p, rp)
        assert np.array_equal(t, rt)


def test_read_ply_errors(tmp_path):
    with pytest.raises(gpu.DgpuError):
        gpu.read_ply(str(tmp_path / "missing.ply"))
    bad = tmp_path / "bad.ply"
    bad.write_bytes(b"ply\nformat ascii 1.0\nelement vertex 1\nproperty quad x\nend_header\n0\n")
    with pytest.raises(gpu.DgpuError):
        gpu.read_ply(str(bad))
    v, faces = _mesh(5)
    faces[0] = [0, 1, 99]  # index beyond the vertices: an error here (undefined in dort)
    path = str(tmp_path / "range.ply")
    write_ply(path, "ascii", v, faces)
    with pytest.raises(gpu.DgpuError):
        gpu.read_ply(path)
    trunc = tmp_path / "trunc.ply"
    write_ply(str(trunc), "binary_little_endian", *_mesh(6))
    trunc.write_bytes(trunc.read_bytes()[:-5])
    with pytest.raises(gpu.DgpuError):
        gpu.read_ply(str(trunc))


GOLDEN = sorted(glob.glob(os.path.join(HERE, "golden", "*.hdr")))


@pytest.mark.parametrize("path", GOLDEN, ids=os.path.basename)
def test_read_rgbe_golden(path):
    """dort's reference images (written by stb_image_write, run-length encoded)."""
    img = gpu.read_rgbe(path)
    assert np.array_equal(img, rgbe.read_hdr(path))
    if ref.available():
        assert np.array_equal(img, ref.read_hdr(path))


def _hdr_image(h, w, seed):
    rng = np.random.default_rng(seed)
    img = np.exp(rng.normal(0, 3, size=(h, w, 3))).astype(np.float32)
    img[0, : w // 2] = 0.0                       # black
    img[1 % h, :] = img[1 % h, :1]               # one long run per channel
    img[:, -1] = 1e-35                           # below the encodable range
    img[-1, 0] = (3.0e38, 1.0, 0.5)              # near the float maximum
    return img


def _rgbe_decode(img):
    """Ward's encoding then stb's decoding, in numpy: the value a round trip must give."""
    m = img.max(axis=2)
    out = np.zeros_like(img)
    ok = m >= 1e-32
    mant, ex = np.frexp(m[ok])
    s = (mant * np.float32(256.0) / m[ok]).astype(np.float32)
    q = np.floor(img[ok] * s[:, None]).astype(np.float32)
    e = (ex + 128) & 0xFF  # the exponent is a byte: 2^128 and above wrap to 0 (black)
    out[ok] = np.where((e > 0)[:, None], q * np.ldexp(np.float32(1.0), e.astype(np.int32) - 136)[:, None], 0)
    return out


@pytest.mark.parametrize("shape", [(5, 7), (16, 300), (3, 8)])
def test_write_rgbe_round_trip(tmp_path, shape):
    """dgpu_write_rgbe encodes as stb's writer does: both files decode to the same values,
    in dgpu_read_rgbe and in dort's reader; flat scanlines below width 8, RLE above."""
    img = _hdr_image(*shape, seed=shape[1])
    mine = str(tmp_path / "mine.hdr")
    gpu.write_rgbe(mine, img)
    back = gpu.read_rgbe(mine)
    assert np.array_equal(back, _rgbe_decode(img))
    if ref.available():
        theirs = str(tmp_path / "theirs.hdr")
        ref.write_hdr(theirs, img)
        assert np.array_equal(ref.read_hdr(mine), ref.read_hdr(theirs))
        assert np.array_equal(back, ref.read_hdr(theirs))
        assert np.array_equal(gpu.read_rgbe(theirs), back)


def test_read_rgbe_errors(tmp_path):
    with pytest.raises(gpu.DgpuError):
        gpu.read_rgbe(str(tmp_path / "missing.hdr"))
    p = tmp_path / "bad.hdr"
    p.write_bytes(b"#?RADIANCE\nFORMAT=32-bit_rle_xyze\n\n-Y 2 +X 2\n" + bytes(16))
    with pytest.raises(gpu.DgpuError) as e:
        gpu.read_rgbe(str(p))
    assert e.value.status == abi.DGPU_EUNSUPPORTED
    p.write_bytes(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 2 +X 2\n" + bytes(9))
    with pytest.raises(gpu.DgpuError):
        gpu.read_rgbe(str(p))
