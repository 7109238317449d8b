"""Radiance RGBE (.hdr) reader (TEST INFRASTRUCTURE ONLY).

Decodes the fixtures in tests/golden/ (dort's test/ref images written by
stbi_write_hdr, src/dort/image.cpp:58-67): header lines, a "-Y h +X w"
resolution line, then scanlines either flat or new-style RLE
(2, 2, hi(w), lo(w), then four run-length coded channel planes).
Pixel value = mantissa byte * 2^(e-136), exactly as stb_image's HDR loader
(extern/stb/stb_image.h, stbi__hdr_convert) decodes it for dort's own comparisons;
the writer truncates mantissas, so decoded values sit up to one mantissa unit low.
"""
import numpy as np


def read_hdr(path):
    with open(path, "rb") as f:
        data = f.read()
    pos = 0
    lines = []
    while True:
        end = data.index(b"\n", pos)
        line = data[pos:end]
        pos = end + 1
        if line.startswith(b"-Y") or line.startswith(b"+Y"):
            res = line.split()
            break
        lines.append(line)
    if not lines or not lines[0].startswith(b"#?"):
        raise ValueError("not a Radiance file")
    h, w = int(res[1]), int(res[3])
    out = np.zeros((h, w, 4), np.uint8)
    for y in range(h):
        if w >= 8 and w < 32768 and data[pos] == 2 and data[pos + 1] == 2 and data[pos + 2] < 128:
            if (data[pos + 2] << 8 | data[pos + 3]) != w:
                raise ValueError("bad scanline width")
            pos += 4
            for c in range(4):
                x = 0
                while x < w:
                    cnt = data[pos]
                    pos += 1
                    if cnt > 128:
                        cnt -= 128
                        out[y, x:x + cnt, c] = data[pos]
                        pos += 1
                    else:
                        out[y, x:x + cnt, c] = np.frombuffer(data, np.uint8, cnt, pos)
                        pos += cnt
                    x += cnt
        else:
            out[y] = np.frombuffer(data, np.uint8, w * 4, pos).reshape(w, 4)
            pos += w * 4
    e = out[..., 3].astype(np.int32)
    scale = np.where(e > 0, np.ldexp(1.0, e - 136), 0.0)
    rgb = out[..., :3].astype(np.float64) * scale[..., None]
    rgb[e == 0] = 0.0
    return rgb.astype(np.float32)
