"""Writes the file-format detection fixtures from the reference's own test data.

The reference test unit_tests/spk_detection/problem.c:14-46 embeds byte samples
of a real DE440 .bsp header (64 bytes) and of a linux_p1550p2650.440 file
(301 bytes, including the constant-name signature at 0xFC), and an ASCII source
line that must be rejected.  This script copies those bytes (data, not code)
into the fixture files below; it is run once in a tree where the reference is
available:

    python tests/golden/make_detection_fixtures.py /root/reference
"""
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def extract(src, name):
    m = re.search(r"static const unsigned char %s\[\] = \{(.*?)\};" % name, src, re.S)
    return bytes(int(x, 16) for x in re.findall(r"0x([0-9a-fA-F]{2})", m.group(1)))


def main(ref):
    src = open(os.path.join(ref, "unit_tests", "spk_detection", "problem.c")).read()
    bsp = extract(src, "valid_bsp_sample")
    asc = extract(src, "ascii_bin_sample")
    line = re.search(r'static const char ascii_source_sample\[\] = "(.*?)\\n";', src).group(1).encode() + b"\n"
    for fname, data in (("detect_valid_bsp_sample.bin", bsp), ("detect_ascii_bin_sample.bin", asc),
                        ("detect_ascii_source_sample.txt", line)):
        with open(os.path.join(HERE, fname), "wb") as f:
            f.write(data)
        print(fname, len(data))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
