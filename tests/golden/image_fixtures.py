"""Writers of small OpenEXR, Radiance RGBE and PFM files, and the values
TextureUtility.LoadImageFromPath (Radiantium.Core/TextureUtility.cs:118-203)
reads from them through Magick.NET-Q16-HDRI.

The reference's own .exr / .hdr assets are not in its repository, and no
OpenEXR library is installed here, so these encoders are written from the
published file specifications, independently of the C++ decoders in
radiantium_amd/csrc/rd_image.cpp: OpenEXR scanline files with NONE, RLE,
ZIPS, ZIP and PIZ compression (PIZ = value bitmap + LUT, the 2D wavelet with
its 14-bit and 16-bit lifting steps, canonical Huffman with zero runs in the
code table and a run-length symbol), Radiance RGBE with run-length and flat
scanlines, PFM in both byte orders.

The expected values follow Magick's pipeline: an OpenEXR sample goes through
the half RGBA interface (FLOAT and UINT rounded to half), a Radiance pixel is
QuantumRange * 2^(e - 136) * mantissa, and every float sample is stored as a
float quantum and divided back: (float)(65535 * v) / 65535.

Usage as a script: python tests/golden/image_fixtures.py <dir> writes one
file of each kind there.
"""
import heapq
import os
import struct
import sys
import zlib

import numpy as np

HALF, FLOAT, UINT = 1, 2, 0
NONE, RLE, ZIPS, ZIP, PIZ = 0, 1, 2, 3, 4
LINES = {NONE: 1, RLE: 1, ZIPS: 1, ZIP: 16, PIZ: 32}


def via_quantum(v):
    """A float sample after Magick's float quantum and / Quantum.Max."""
    return (np.float32(65535.0 * np.asarray(v, np.float64)) / np.float32(65535.0)).astype(np.float32)


def through_half(v, ptype):
    """The sample as Magick's half RGBA interface delivers it, as float32."""
    if ptype == HALF:
        return np.asarray(v, np.float16).astype(np.float32)
    return np.asarray(np.asarray(v, np.float32), np.float16).astype(np.float32)


# ---------------------------------------------------------------- OpenEXR
def _interleave_predict(raw):
    b = np.frombuffer(raw, np.uint8)
    t = np.concatenate([b[0::2], b[1::2]]).astype(np.int32)
    d = t.copy()
    d[1:] = (t[1:] - t[:-1] + 128) & 255
    return d.astype(np.uint8).tobytes()


def _rle(data):
    out = bytearray()
    i, n = 0, len(data)
    while i < n:
        j = i + 1
        while j < n and data[j] == data[i] and j - i < 128:
            j += 1
        if j - i >= 3:
            out += struct.pack("b", j - i - 1) + bytes([data[i]])
            i = j
            continue
        j = i
        while j < n and j - i < 127 and not (j + 2 < n and data[j] == data[j + 1] == data[j + 2]):
            j += 1
        if j == i:
            j = i + 1
        out += struct.pack("b", -(j - i)) + data[i:j]
        i = j
    return bytes(out)


def _wenc14(a, b):
    a_s = a - 65536 if a >= 32768 else a   # as 16-bit signed values
    b_s = b - 65536 if b >= 32768 else b
    return ((a_s + b_s) >> 1) & 0xffff, (a_s - b_s) & 0xffff


def _wenc16(a, b):
    ao = (a + 0x8000) & 0xffff
    m = (ao + b) >> 1
    d = ao - b
    if d < 0:
        m = (m + 0x8000) & 0xffff
    return m & 0xffff, d & 0xffff


def _wav2_encode(buf, base, nx, ox, ny, oy, mx):
    """In-place forward wavelet of the plane at buf[base], element stride ox,
    row stride oy (mirrors the decoder's loops in the opposite direction)."""
    enc = _wenc14 if mx < (1 << 14) else _wenc16
    n = min(nx, ny)
    p, p2 = 1, 2
    while p2 <= n:
        oy1, oy2, ox1, ox2 = oy * p, oy * p2, ox * p, ox * p2
        py = base
        ey = base + oy * (ny - p2)
        while py <= ey:
            px = py
            ex = py + ox * (nx - p2)
            while px <= ex:
                p01, p10 = px + ox1, px + oy1
                p11 = p10 + ox1
                i00, i01 = enc(buf[px], buf[p01])
                i10, i11 = enc(buf[p10], buf[p11])
                buf[px], buf[p10] = enc(i00, i10)
                buf[p01], buf[p11] = enc(i01, i11)
                px += ox2
            if nx & p:
                p10 = px + oy1
                i00, buf[p10] = enc(buf[px], buf[p10])
                buf[px] = i00
            py += oy2
        if ny & p:
            px = py
            ex = py + ox * (nx - p2)
            while px <= ex:
                p01 = px + ox1
                i00, buf[p01] = enc(buf[px], buf[p01])
                buf[px] = i00
                px += ox2
        p = p2
        p2 <<= 1


def _huffman_lengths(freq):
    """Code length per symbol (dict symbol -> frequency), plain Huffman."""
    heap = [(f, i, (s,)) for i, (s, f) in enumerate(sorted(freq.items()))]
    heapq.heapify(heap)
    length = {s: 0 for s in freq}
    if len(heap) == 1:
        length[heap[0][2][0]] = 1
        return length
    k = len(heap)
    while len(heap) > 1:
        f1, _, s1 = heapq.heappop(heap)
        f2, _, s2 = heapq.heappop(heap)
        for s in s1 + s2:
            length[s] += 1
        heapq.heappush(heap, (f1 + f2, k, s1 + s2))
        k += 1
    assert max(length.values()) <= 58
    return length


class _Bits:
    def __init__(self):
        self.out = bytearray()
        self.c = 0
        self.lc = 0

    def put(self, n, v):
        self.c = (self.c << n) | (v & ((1 << n) - 1))
        self.lc += n
        while self.lc >= 8:
            self.lc -= 8
            self.out.append((self.c >> self.lc) & 0xff)
        self.c &= (1 << self.lc) - 1 if self.lc else 0

    def flush(self):
        if self.lc:
            self.out.append((self.c << (8 - self.lc)) & 0xff)
        return bytes(self.out)


def _huf_compress(data):
    data = [int(v) for v in data]
    if not data:
        return b""
    freq = {}
    for v in data:
        freq[v] = freq.get(v, 0) + 1
    im = min(freq)
    rlc = max(freq) + 1   # the run-length symbol follows the largest value
    freq[rlc] = 1
    iM = rlc
    length = _huffman_lengths(freq)
    # canonical codes, longest length first, one length's codes in symbol order
    count = [0] * 59
    for s, l in length.items():
        count[l] += 1
    start = [0] * 59
    c = 0
    for l in range(58, 0, -1):
        start[l] = c
        c = (c + count[l]) >> 1
    code = {}
    for s in sorted(length):
        code[s] = start[length[s]]
        start[length[s]] += 1
    # code table: 6-bit lengths from im to iM, zero runs 59-62 (2-5) and 63 + 8 bits (6-261)
    t = _Bits()
    s = im
    while s <= iM:
        l = length.get(s, 0)
        if l == 0:
            z = 1
            while s + z <= iM and length.get(s + z, 0) == 0 and z < 261:
                z += 1
            if z >= 6:
                t.put(6, 63)
                t.put(8, z - 6)
                s += z
                continue
            if z >= 2:
                t.put(6, 59 + z - 2)
                s += z
                continue
        t.put(6, l)
        s += 1
    table = t.flush()
    b = _Bits()

    def send(sym, run):
        if length[sym] + length[rlc] + 8 < length[sym] * run:
            b.put(length[sym], code[sym])
            b.put(length[rlc], code[rlc])
            b.put(8, run)
        else:
            for _ in range(run + 1):
                b.put(length[sym], code[sym])

    cs = 0
    sym = data[0]
    for v in data[1:]:
        if v == sym and cs < 255:
            cs += 1
        else:
            send(sym, cs)
            cs = 0
        sym = v
    send(sym, cs)
    nbits = len(b.out) * 8 + b.lc
    bits = b.flush()
    return struct.pack("<IIIII", im, iM, len(table), nbits, 0) + table + bits


def _piz_compress(raw, width, lines, sizes):
    words = np.frombuffer(raw, "<u2")
    planes = [[] for _ in sizes]
    pos = 0
    for _ in range(lines):
        for ci, sz in enumerate(sizes):
            nw = width * (sz // 2)
            planes[ci].append(words[pos:pos + nw])
            pos += nw
    tmp = np.concatenate([np.concatenate(p) for p in planes]).astype(np.int64)
    bitmap = np.zeros(8192, np.uint8)
    for v in np.unique(tmp):
        bitmap[v >> 3] |= 1 << (v & 7)
    bitmap[0] &= 0xfe
    nz = np.nonzero(bitmap)[0]
    lo, hi = (int(nz[0]), int(nz[-1])) if len(nz) else (8191, 0)
    lut = np.zeros(65536, np.int64)
    k = 0
    for i in range(65536):
        if i == 0 or bitmap[i >> 3] & (1 << (i & 7)):
            lut[i] = k
            k += 1
    maxv = k - 1
    buf = [int(v) for v in lut[tmp]]
    base = 0
    for sz in sizes:
        nw = sz // 2
        for j in range(nw):
            _wav2_encode(buf, base + j, width, nw, lines, width * nw, maxv)
        base += width * lines * nw
    huf = _huf_compress(buf)
    out = struct.pack("<HH", lo, hi)
    if lo <= hi:
        out += bitmap[lo:hi + 1].tobytes()
    return out + struct.pack("<i", len(huf)) + huf


def write_exr(path, channels, compression):
    """channels: {name: (ptype, array[h, w])} (same shape).  Returns nothing;
    the file stores the channels sorted by name."""
    names = sorted(channels)
    h, w = next(iter(channels.values()))[1].shape
    hdr = bytearray(struct.pack("<II", 20000630, 2))

    def attr(name, typ, val):
        hdr.extend(name.encode() + b"\0" + typ.encode() + b"\0" + struct.pack("<i", len(val)) + val)

    cl = b"".join(n.encode() + b"\0" + struct.pack("<i", channels[n][0]) + b"\0\0\0\0" + struct.pack("<ii", 1, 1)
                  for n in names) + b"\0"
    attr("channels", "chlist", cl)
    attr("compression", "compression", bytes([compression]))
    box = struct.pack("<iiii", 0, 0, w - 1, h - 1)
    attr("dataWindow", "box2i", box)
    attr("displayWindow", "box2i", box)
    attr("lineOrder", "lineOrder", b"\0")
    attr("pixelAspectRatio", "float", struct.pack("<f", 1.0))
    attr("screenWindowCenter", "v2f", struct.pack("<ff", 0.0, 0.0))
    attr("screenWindowWidth", "float", struct.pack("<f", 1.0))
    hdr.append(0)
    lpb = LINES[compression]
    sizes = [2 if channels[n][0] == HALF else 4 for n in names]
    chunks = []
    for y0 in range(0, h, lpb):
        lines = min(lpb, h - y0)
        raw = bytearray()
        for y in range(y0, y0 + lines):
            for n in names:
                pt, a = channels[n]
                row = a[y]
                raw += (np.asarray(row, "<f2") if pt == HALF else np.asarray(row, "<f4") if pt == FLOAT
                        else np.asarray(row, "<u4")).tobytes()
        raw = bytes(raw)
        if compression == NONE:
            data = raw
        elif compression == RLE:
            data = _rle(_interleave_predict(raw))
        elif compression in (ZIPS, ZIP):
            data = zlib.compress(_interleave_predict(raw))
        else:
            data = _piz_compress(raw, w, lines, sizes)
        if len(data) >= len(raw):
            data = raw
        chunks.append(struct.pack("<ii", y0, len(data)) + data)
    off = len(hdr) + 8 * len(chunks)
    table = bytearray()
    for c in chunks:
        table += struct.pack("<Q", off)
        off += len(c)
    with open(path, "wb") as f:
        f.write(bytes(hdr) + bytes(table) + b"".join(chunks))


def exr_expected(channels):
    """Rows top first, 4 channels (R, G, B, A) as LoadImageFromPath's pixel source."""
    h, w = next(iter(channels.values()))[1].shape
    out = np.zeros((h, w, 4), np.float32)
    have = set(channels)
    if not have & {"R", "G", "B"} and "Y" in have:
        pt, a = channels["Y"]
        v = via_quantum(through_half(a, pt))
        out[..., 0] = out[..., 1] = out[..., 2] = v
    for k, n in enumerate("RGB"):
        if n in channels:
            pt, a = channels[n]
            out[..., k] = via_quantum(through_half(a, pt))
    if "A" in channels:
        pt, a = channels["A"]
        out[..., 3] = via_quantum(through_half(a, pt))
    else:
        out[..., 3] = via_quantum(1.0)
    return out


# ---------------------------------------------------------------- Radiance RGBE
def _rgbe_runs(comp):
    out = bytearray()
    i, n = 0, len(comp)
    while i < n:
        j = i + 1
        while j < n and comp[j] == comp[i] and j - i < 127:
            j += 1
        if j - i >= 4:
            out += bytes([128 + j - i, comp[i]])
            i = j
            continue
        j = i
        while j < n and j - i < 128 and not (j + 3 < n and comp[j] == comp[j + 1] == comp[j + 2] == comp[j + 3]):
            j += 1
        if j == i:
            j = i + 1
        out += bytes([j - i]) + bytes(comp[i:j])
        i = j
    return bytes(out)


def write_hdr(path, rgbe, top_first=True):
    """rgbe: uint8 [h, w, 4] in image rows top first."""
    h, w, _ = rgbe.shape
    out = bytearray(b"#?RADIANCE\n# test image\nFORMAT=32-bit_rle_rgbe\n\n")
    out += (f"-Y {h} +X {w}\n" if top_first else f"+Y {h} +X {w}\n").encode()
    rows = range(h) if top_first else range(h - 1, -1, -1)
    for y in rows:
        px = rgbe[y]
        if 8 <= w < 0x8000:
            out += bytes([2, 2, w >> 8, w & 255])
            for c in range(4):
                out += _rgbe_runs(bytes(px[:, c]))
        else:
            out += px.tobytes()
    with open(path, "wb") as f:
        f.write(bytes(out))


def hdr_expected(rgbe):
    m = rgbe[..., :3].astype(np.float64)
    e = rgbe[..., 3:4].astype(np.float64)
    g = np.where(e != 0, np.power(2.0, e - 136.0), 0.0)
    return np.where(e != 0, (np.float32(65535.0 * g * m) / np.float32(65535.0)), 0.0).astype(np.float32)


# ---------------------------------------------------------------- PFM
def write_pfm(path, img, little=True):
    """img: float32 [h, w, 3] or [h, w] rows top first; PFM stores bottom row first."""
    gray = img.ndim == 2
    h, w = img.shape[:2]
    head = f"{'Pf' if gray else 'PF'}\n{w} {h}\n{-1.0 if little else 1.0}\n".encode()
    data = np.ascontiguousarray(img[::-1]).astype("<f4" if little else ">f4")
    with open(path, "wb") as f:
        f.write(head + data.tobytes())


def pfm_expected(img):
    return via_quantum(np.asarray(img, np.float32)).reshape(img.shape[0], img.shape[1], -1)


def sample_channels(h, w, seed, ptype, names="RGB", spread=1.0):
    rng = np.random.default_rng(seed)
    out = {}
    for n in names:
        if ptype == UINT:
            a = rng.integers(0, 70000, (h, w)).astype(np.uint32)
        else:
            a = (rng.lognormal(0.0, 1.5, (h, w)) * spread).astype(np.float32)
            a[rng.random((h, w)) < 0.2] = 0.0          # runs of equal values
            a[:, : w // 3] = a[:, :1]                    # flat blocks
        out[n] = (ptype, a)
    return out


if __name__ == "__main__":
    d = sys.argv[1] if len(sys.argv) > 1 else "."
    os.makedirs(d, exist_ok=True)
    write_exr(os.path.join(d, "piz_half.exr"), sample_channels(45, 13, 0, HALF), PIZ)
    write_exr(os.path.join(d, "zip_float.exr"), sample_channels(37, 21, 1, FLOAT), ZIP)
    write_hdr(os.path.join(d, "sky.hdr"), np.random.default_rng(2).integers(0, 256, (9, 17, 4)).astype(np.uint8))
    write_pfm(os.path.join(d, "img.pfm"), np.random.default_rng(3).random((7, 5, 3)).astype(np.float32))
    print("written to", d)
