"""Writes ct_image_header.dcm: the header of the reference's sample CT image
(AAPM_script/Info.IMA, which reconscript.m:7 reads through
CollectImageCfg.m), cut just before its pixel data element, so that no image
content is kept, and with every person-name (PN) value overwritten by 'X's of
the same length (lengths and offsets unchanged).  Run in a checkout that has
the reference next to it:

    python tests/golden/make_ct_image_header.py /root/reference/AAPM_script/Info.IMA
"""
import os
import struct
import sys


def header_end(d: bytearray) -> int:
    """Offset of the top-level (7FE0,0010) element (explicit VR LE data set);
    top-level PN values are blanked on the way."""
    o = 132
    while o + 8 <= len(d):
        g, e = struct.unpack_from("<HH", d, o)
        if (g, e) == (0x7FE0, 0x0010):
            return o
        vr = bytes(d[o + 4:o + 6])
        if vr in (b"OB", b"OD", b"OF", b"OL", b"OV", b"OW", b"SQ", b"SV", b"UC", b"UN", b"UR",
                  b"UT", b"UV"):
            n = struct.unpack_from("<I", d, o + 8)[0]
            o += 12 + n
        else:
            n = struct.unpack_from("<H", d, o + 6)[0]
            if vr == b"PN":
                d[o + 8:o + 8 + n] = b"X" * n
            o += 8 + n
    raise SystemExit("no pixel data element")


if __name__ == "__main__":
    src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/AAPM_script/Info.IMA"
    d = bytearray(open(src, "rb").read())
    assert d[128:132] == b"DICM"
    end = header_end(d)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ct_image_header.dcm")
    with open(out, "wb") as f:
        f.write(d[:end])
    print(f"{out}: {end} bytes")
