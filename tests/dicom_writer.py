"""Minimal DICOM writer for synthetic test files: explicit or implicit VR
little endian, elements given as (group, element, VR, value bytes), plus
pre-encoded chunks for sequences of undefined length.  Independent of the
library's reader (plain struct packing)."""
import struct

import numpy as np

EXPLICIT = "1.2.840.10008.1.2.1"
IMPLICIT = "1.2.840.10008.1.2"
BIG_ENDIAN = "1.2.840.10008.1.2.2"
JPEG_LOSSLESS = "1.2.840.10008.1.2.4.70"
LONG_VR = {"OB", "OD", "OF", "OL", "OV", "OW", "SQ", "SV", "UC", "UN", "UR", "UT", "UV"}
UNDEFINED = 0xFFFFFFFF


def element(g, e, vr, value: bytes, explicit=True, length=None) -> bytes:
    if length is None and len(value) % 2:
        value += b"\0" if vr in ("UI", "OB", "UN") else b" "
    L = len(value) if length is None else length
    head = struct.pack("<HH", g, e)
    if explicit and g != 0xFFFE:
        if vr in LONG_VR:
            head += vr.encode() + b"\0\0" + struct.pack("<I", L)
        else:
            head += vr.encode() + struct.pack("<H", L)
    else:
        head += struct.pack("<I", L)
    return head + value


def sequence_undefined(g, e, items, explicit=True, vr="SQ") -> bytes:
    """A sequence of undefined length whose items (lists of elements) have
    undefined length too, with item and sequence delimiters."""
    inner_explicit = explicit and vr != "UN"
    body = b""
    for it in items:
        body += element(0xFFFE, 0xE000, "", b"", explicit, UNDEFINED)
        body += b"".join(element(*x, explicit=inner_explicit) for x in it)
        body += element(0xFFFE, 0xE00D, "", b"", explicit, 0)
    body += element(0xFFFE, 0xE0DD, "", b"", explicit, 0)
    return element(g, e, vr, b"", explicit, UNDEFINED) + body


def write(path, elements, ts=EXPLICIT, raw=()):
    """elements: (g, e, vr, bytes); raw: (g, e, bytes) chunks inserted in tag
    order (pre-encoded elements such as undefined-length sequences)."""
    explicit = ts != IMPLICIT
    tsb = ts.encode()
    meta_body = element(0x0002, 0x0001, "OB", b"\0\1") + element(0x0002, 0x0010, "UI", tsb)
    meta = element(0x0002, 0x0000, "UL", struct.pack("<I", len(meta_body))) + meta_body
    chunks = [((g, e), element(g, e, vr, v, explicit)) for g, e, vr, v in elements]
    chunks += [((g, e), b) for g, e, b in raw]
    body = b"".join(b for _, b in sorted(chunks, key=lambda c: c[0]))
    with open(path, "wb") as f:
        f.write(b"\0" * 128 + b"DICM" + meta + body)


def f32(*x):
    return struct.pack(f"<{len(x)}f", *x)


def u16(x):
    return struct.pack("<H", x)


def ctpd_elements(g: dict, frame: np.ndarray, photons=None):
    """A DICOM-CT-PD projection: the private geometry elements CollectReconCfg
    reads (values as float32 / float64 / uint16 bytes) and one frame of
    NumberofDetectorColumns rows x NumberofDetectorRows columns."""
    W, H = frame.shape
    els = [
        (0x0008, 0x0060, "CS", b"CT"),
        (0x0018, 0x0061, "DS", g.get("hu", b"0.0192")),
        (0x0018, 0x0090, "DS", b"500"),
        (0x0018, 0x9311, "FD", struct.pack("<d", g["pitch"])),
        (0x0028, 0x0002, "US", u16(1)),
        (0x0028, 0x0010, "US", u16(W)),
        (0x0028, 0x0011, "US", u16(H)),
        (0x0028, 0x0100, "US", u16(16)),
        (0x0028, 0x0101, "US", u16(16)),
        (0x7029, 0x0010, "LO", b"DLX_DETECTOR_GEOMETRY"),
        (0x7029, 0x1002, "FL", f32(g["PerDetW"])),
        (0x7029, 0x1006, "FL", f32(g["PerDetH"])),
        (0x7029, 0x100B, "CS", b"CYLINDRICAL"),
        (0x7029, 0x1010, "US", u16(H)),
        (0x7029, 0x1011, "US", u16(W)),
        (0x7031, 0x0010, "LO", b"DLX_FOCAL_SPOT"),
        (0x7031, 0x1001, "FL", f32(g["angle"])),
        (0x7031, 0x1002, "FL", f32(g.get("z0", -12.5))),
        (0x7031, 0x1003, "FL", f32(g["SO"])),
        (0x7031, 0x1031, "FL", f32(g["SD"])),
        (0x7031, 0x1033, "FL", f32(g["cx"], g["cy"])),
        (0x7033, 0x0010, "LO", b"DLX_SOURCE"),
        (0x7033, 0x100B, "FL", f32(0.25)),
        (0x7033, 0x100C, "FL", f32(3.3613e-4)),
        (0x7033, 0x100D, "FL", f32(-0.5)),
        (0x7033, 0x1013, "US", u16(g["DefTimes"])),
        (0x7FE0, 0x0010, "OW", np.ascontiguousarray(frame, "<u2").tobytes()),
    ]
    if photons is not None:
        els.append((0x7033, 0x1065, "FL", f32(*photons)))
    return els
