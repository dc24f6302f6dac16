"""Reader for USD binary "crate" files (``.usd`` / ``.usdc``, PXR-USDC
format 0.4-0.9) without the USD library.

The reference loads its scenes with ``pxr.Usd`` (``modelDO.py:545-640``,
``laparoscope.py:302-363``); ``pxr`` is not available here, so this module
reads the file directly.  It reads the file's own tables only and executes
nothing from it.  Supported: the token, string, field, field-set, path and
spec tables, and the value types that scene geometry uses (scalars, vectors,
matrices, tokens, strings, arrays of numbers and vectors, token and path
vectors, specifiers).  Other values (dictionaries, list ops, time samples)
read as ``None``.

Format notes:
* The tables are compressed with LZ4 in chunks: a chunk-count byte, then
  for each chunk an int32 size and an LZ4 block.
* Integer tables use the USD integer coding (a common delta, 2-bit codes per
  value, then 8/16/32-bit or 16/32/64-bit deltas).
* A value is a 64-bit reference: array / inlined / compressed bits, an 8-bit
  type, and a 48-bit payload (an inline value or a file offset).
"""
from __future__ import annotations

import struct
from typing import Dict, List, Optional

import numpy as np


# ---------------------------------------------------------------- compression
def _lz4_block(src: bytes) -> bytes:
    out = bytearray()
    i, n = 0, len(src)
    while i < n:
        tok = src[i]
        i += 1
        lit = tok >> 4
        if lit == 15:
            while True:
                c = src[i]
                i += 1
                lit += c
                if c != 255:
                    break
        out += src[i:i + lit]
        i += lit
        if i >= n:
            break
        off = src[i] | (src[i + 1] << 8)
        i += 2
        ml = tok & 15
        if ml == 15:
            while True:
                c = src[i]
                i += 1
                ml += c
                if c != 255:
                    break
        ml += 4
        if off == 0 or off > len(out):
            raise ValueError("corrupt LZ4 block")
        start = len(out) - off
        if off >= ml:
            out += out[start:start + ml]
        else:  # overlapping copy repeats the last `off` bytes
            for k in range(ml):
                out.append(out[start + k])
    return bytes(out)


def decompress(src: bytes) -> bytes:
    """Chunked LZ4 as written by USD's fast compression."""
    nch = src[0]
    if nch == 0:
        return _lz4_block(src[1:])
    out = bytearray()
    o = 1
    for _ in range(nch):
        size = struct.unpack_from("<i", src, o)[0]
        o += 4
        out += _lz4_block(src[o:o + size])
        o += size
    return bytes(out)


def decode_ints(data: bytes, count: int, width: int = 4) -> np.ndarray:
    """USD integer coding: a common delta, 2-bit codes (0 common, 1/2/3 a
    delta of 1/2/4 bytes for 32-bit ints, 2/4/8 for 64-bit), then the deltas;
    values are the running sum of the deltas."""
    fmt = {4: ("<i", "<b", "<h", "<i"), 8: ("<q", "<h", "<i", "<q")}[width]
    size = {4: (4, 1, 2, 4), 8: (8, 2, 4, 8)}[width]
    common = struct.unpack_from(fmt[0], data, 0)[0]
    ncode = (count * 2 + 7) // 8
    codes = np.frombuffer(data, np.uint8, ncode, size[0])
    v = size[0] + ncode
    out = np.empty(count, np.int64)
    prev = 0
    unpack = struct.unpack_from
    for i in range(count):
        c = (int(codes[i >> 2]) >> (2 * (i & 3))) & 3
        if c == 0:
            prev += common
        else:
            prev += unpack(fmt[c], data, v)[0]
            v += size[c]
        out[i] = prev
    return out


# --------------------------------------------------------------- value types
_SCALAR = {  # type id: (struct format, numpy dtype, components)
    1: ("<?", np.bool_, 1), 2: ("<B", np.uint8, 1), 3: ("<i", np.int32, 1), 4: ("<I", np.uint32, 1),
    5: ("<q", np.int64, 1), 6: ("<Q", np.uint64, 1), 7: ("<e", np.float16, 1), 8: ("<f", np.float32, 1),
    9: ("<d", np.float64, 1),
    13: ("<4d", np.float64, 4), 14: ("<9d", np.float64, 9), 15: ("<16d", np.float64, 16),
    16: ("<4d", np.float64, 4), 17: ("<4f", np.float32, 4), 18: ("<4e", np.float16, 4),
    19: ("<2d", np.float64, 2), 20: ("<2f", np.float32, 2), 21: ("<2e", np.float16, 2), 22: ("<2i", np.int32, 2),
    23: ("<3d", np.float64, 3), 24: ("<3f", np.float32, 3), 25: ("<3e", np.float16, 3), 26: ("<3i", np.int32, 3),
    27: ("<4d", np.float64, 4), 28: ("<4f", np.float32, 4), 29: ("<4e", np.float16, 4), 30: ("<4i", np.int32, 4),
}
_MATRIX = {13: 2, 14: 3, 15: 4}
T_STRING, T_TOKEN, T_ASSET, T_TOKENVEC, T_SPECIFIER, T_VARIABILITY, T_PATHVEC = 10, 11, 12, 41, 42, 44, 40
T_DOUBLEVEC, T_STRINGVEC, T_TIMECODE = 48, 50, 56
SPECIFIERS = ("def", "over", "class")


class CrateFile:
    """The specs of one crate file: ``prims`` maps a prim path to its
    fields, ``attributes`` maps "prim.attr" paths to theirs."""

    def __init__(self, path: str):
        with open(path, "rb") as f:
            self.data = f.read()
        b = self.data
        if b[:8] != b"PXR-USDC":
            raise ValueError(f"{path}: not a USD crate file")
        self.version = tuple(b[8:11])
        if self.version < (0, 4, 0):
            raise ValueError(f"{path}: crate version {self.version} < 0.4.0 is not supported")
        toc = struct.unpack_from("<q", b, 16)[0]
        nsec = struct.unpack_from("<Q", b, toc)[0]
        self.sections = {}
        for i in range(nsec):
            o = toc + 8 + 32 * i
            name = b[o:o + 16].split(b"\0")[0].decode()
            self.sections[name] = struct.unpack_from("<qq", b, o + 16)
        self._read_tokens()
        self._read_strings()
        self._read_fields()
        self._read_fieldsets()
        self._read_paths()
        self._read_specs()

    # -------------------------------------------------------------- tables
    def _u64(self, o):
        return struct.unpack_from("<Q", self.data, o)[0]

    def _ints(self, o, n, width=4):
        """Compressed integer table at o: (values, next offset)."""
        size = self._u64(o)
        return decode_ints(decompress(self.data[o + 8:o + 8 + size]), n, width), o + 8 + size

    def _read_tokens(self):
        o = self.sections["TOKENS"][0]
        n, _, csize = struct.unpack_from("<QQQ", self.data, o)
        raw = decompress(self.data[o + 24:o + 24 + csize])
        self.tokens = [t.decode("utf-8") for t in raw.split(b"\0")[:n]]

    def _read_strings(self):
        o = self.sections["STRINGS"][0]
        n = self._u64(o)
        self.strings = list(struct.unpack_from(f"<{n}I", self.data, o + 8))

    def _read_fields(self):
        o = self.sections["FIELDS"][0]
        n = self._u64(o)
        tok, o = self._ints(o + 8, n)
        size = self._u64(o)
        reps = decompress(self.data[o + 8:o + 8 + size])
        self.fields = list(zip((self.tokens[t] for t in tok), struct.unpack_from(f"<{n}Q", reps)))

    def _read_fieldsets(self):
        o = self.sections["FIELDSETS"][0]
        n = self._u64(o)
        self.fieldsets, _ = self._ints(o + 8, n)

    def _read_paths(self):
        o = self.sections["PATHS"][0]
        n = self._u64(o)
        nenc = self._u64(o + 8)
        pidx, o2 = self._ints(o + 16, nenc)
        etok, o2 = self._ints(o2, nenc)
        jumps, _ = self._ints(o2, nenc)
        self.paths: List[Optional[str]] = [None] * n
        # iterative form of the depth-first path tree: a child follows its
        # parent; jumps[i] > 0 is the offset of the sibling, -1 a child only,
        # 0 a sibling only (next), -2 neither
        stack = [(0, None)]
        while stack:
            i, parent = stack.pop()
            while True:
                if parent is None:
                    path = "/"
                else:
                    t = int(etok[i])
                    name = self.tokens[abs(t)]
                    if t < 0:
                        path = parent + "." + name
                    else:
                        path = (parent if parent != "/" else "") + "/" + name
                self.paths[int(pidx[i])] = path
                j = int(jumps[i])
                has_child = j > 0 or j == -1
                has_sibling = j >= 0
                if has_child and has_sibling:
                    stack.append((i + j, parent))
                if has_child:
                    parent = path
                if not (has_child or has_sibling):
                    break
                i += 1

    def _read_specs(self):
        o = self.sections["SPECS"][0]
        n = self._u64(o)
        paths, o = self._ints(o + 8, n)
        fsets, o = self._ints(o, n)
        types, _ = self._ints(o, n)
        self.specs: Dict[str, dict] = {}
        self.spec_types: Dict[str, int] = {}
        fs = self.fieldsets
        for p, f, t in zip(paths, fsets, types):
            fields = {}
            k = int(f)
            while k < len(fs) and fs[k] >= 0:
                name, rep = self.fields[int(fs[k])]
                fields[name] = rep
                k += 1
            path = self.paths[int(p)]
            self.specs[path] = fields
            self.spec_types[path] = int(t)

    # -------------------------------------------------------------- values
    def value(self, rep: int):
        """Decode one value reference."""
        is_array = bool(rep >> 63 & 1)
        inlined = bool(rep >> 62 & 1)
        compressed = bool(rep >> 61 & 1)
        t = (rep >> 48) & 0xFF
        payload = rep & ((1 << 48) - 1)
        b = self.data
        if is_array:
            return self._array(t, payload, compressed)
        if inlined:
            return self._inline(t, payload)
        if t in _SCALAR:
            fmt, _, k = _SCALAR[t]
            v = struct.unpack_from(fmt, b, payload)
            if t in _MATRIX:
                m = _MATRIX[t]
                return np.array(v, np.float64).reshape(m, m)
            return v[0] if k == 1 else np.array(v)
        if t == T_TOKEN:
            return self.tokens[struct.unpack_from("<I", b, payload)[0]]
        if t in (T_STRING, T_ASSET):
            return self.tokens[self.strings[struct.unpack_from("<I", b, payload)[0]]]
        if t in (T_TOKENVEC, T_PATHVEC):
            n = self._u64(payload)
            idx = struct.unpack_from(f"<{n}I", b, payload + 8)
            return [self.tokens[i] for i in idx] if t == T_TOKENVEC else [self.paths[i] for i in idx]
        if t == T_DOUBLEVEC:
            n = self._u64(payload)
            return list(struct.unpack_from(f"<{n}d", b, payload + 8))
        if t == T_TIMECODE:
            return struct.unpack_from("<d", b, payload)[0]
        return None

    def _inline(self, t, payload):
        if t == T_TOKEN:
            return self.tokens[payload & 0xFFFFFFFF]
        if t in (T_STRING, T_ASSET):
            return self.tokens[self.strings[payload & 0xFFFFFFFF]]
        if t == T_SPECIFIER:
            return SPECIFIERS[payload & 0xFF] if (payload & 0xFF) < 3 else int(payload & 0xFF)
        if t in (T_VARIABILITY,):
            return int(payload & 0xFF)
        raw = struct.pack("<Q", payload)
        if t in (1, 2, 3, 4, 8):
            return struct.unpack_from(_SCALAR[t][0], raw)[0]
        if t in (5, 6):  # 64-bit ints inlined from their low 32 bits
            return struct.unpack_from("<i" if t == 5 else "<I", raw)[0]
        if t == 9:  # doubles inline as floats
            return float(struct.unpack_from("<f", raw)[0])
        if t == 7:
            return float(struct.unpack_from("<e", raw)[0])
        if t == T_TIMECODE:
            return float(struct.unpack_from("<f", raw)[0])
        if t in _MATRIX:  # diagonal as int8s
            m = _MATRIX[t]
            d = struct.unpack_from(f"<{m}b", raw)
            return np.diag(np.array(d, np.float64))
        if t in _SCALAR:  # vectors and quaternions as int8 components
            k = _SCALAR[t][2]
            return np.array(struct.unpack_from(f"<{k}b", raw), _SCALAR[t][1])
        return None

    def _array(self, t, o, compressed):
        b = self.data
        if o == 0:
            return np.zeros(0)
        if self.version >= (0, 7, 0):
            n = self._u64(o)
            o += 8
        else:
            n = struct.unpack_from("<I", b, o)[0]
            o += 4
        if t == T_TOKEN:
            return [self.tokens[i] for i in struct.unpack_from(f"<{n}I", b, o)]
        if t in (T_STRING, T_ASSET):
            return [self.tokens[self.strings[i]] for i in struct.unpack_from(f"<{n}I", b, o)]
        if t not in _SCALAR:
            return None
        fmt, dt, k = _SCALAR[t]
        if compressed and t in (3, 4, 5, 6):
            width = 8 if t in (5, 6) else 4
            vals, _ = self._ints(o, n, width)
            return vals.astype(dt)
        if compressed and t in (7, 8, 9):
            code = chr(b[o])
            o += 1
            if code == "i":
                vals, _ = self._ints(o, n)
                return vals.astype(dt)
            if code == "t":
                lut_n = struct.unpack_from("<I", b, o)[0]
                o += 4
                lut = np.frombuffer(b, dt, lut_n, o)
                o += lut_n * np.dtype(dt).itemsize
                idx, _ = self._ints(o, n)
                return lut[idx]
            raise ValueError(f"unknown float array coding {code!r}")
        arr = np.frombuffer(b, dt, n * k, o).copy()
        if t in _MATRIX:
            m = _MATRIX[t]
            return arr.reshape(n, m, m)
        return arr.reshape(n, k) if k > 1 else arr

    # -------------------------------------------------------------- access
    def prim_paths(self) -> List[str]:
        return [p for p in self.specs if p and "." not in p.rsplit("/", 1)[-1]]

    def fields_of(self, path: str) -> Dict[str, object]:
        """Decoded fields of a spec (prim or attribute)."""
        return {k: self.value(v) for k, v in self.specs.get(path, {}).items()}

    def get(self, prim: str, attr: str, field: str = "default"):
        """Value of a prim's attribute (its default value) or None."""
        rep = self.specs.get(f"{prim}.{attr}", {}).get(field)
        return None if rep is None else self.value(rep)

    def attributes(self, prim: str) -> List[str]:
        pre = prim + "."
        return [p[len(pre):] for p in self.specs if p.startswith(pre)]
