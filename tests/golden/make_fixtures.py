"""Writes fixtures from the reference test suite's data to tests/golden/.

* fonts/font_[a-g].ttf -- the reference's tests embed seven fonts as base64
  string literals (test/test.cpp:114-304: font_a..font_c valid, font_d..font_g
  malformed in the ways the comments there describe) and decode them at
  start-up with their own decoder (test.cpp:2285-2310).  They are decoded here
  the same way: the symbols A-Z a-z 0-9 + / carry six bits each, any other
  symbol counts as zero, and decoding stops at the first '='.
* reference_hashes.json -- the expected image hash, width and height of each
  test in the reference's table (test.cpp:2186-2261); a test passes when the
  hash of its sRGB8 image (test.cpp:2365-2406) is within Hamming distance 5
  (test.cpp:2614-2618).

Only data is taken: the decoded bytes and the table's numbers.

Usage: python tests/golden/make_fixtures.py [path/to/test.cpp]
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT = "/root/reference/test/test.cpp"
NAMES = "abcdefg"


def decode(text):
    out, data, held = bytearray(), 0, 0
    for ch in text:
        if ch == "=":
            break
        if "A" <= ch <= "Z":
            value = ord(ch) - ord("A")
        elif "a" <= ch <= "z":
            value = ord(ch) - ord("a") + 26
        elif "0" <= ch <= "9":
            value = ord(ch) - ord("0") + 52
        elif ch == "+":
            value = 62
        elif ch == "/":
            value = 63
        else:
            value = 0
        data = data << 6 | value
        held += 6
        if held >= 8:
            held -= 8
            out.append((data >> held) & 0xFF)
            data &= (1 << held) - 1
    return bytes(out)


def literals(source):
    fonts = {}
    for name in NAMES:
        m = re.search(r"char const font_%s_base64\[\] =((?:\s*\"[^\"]*\")+)\s*;" % name, source)
        if not m:
            raise SystemExit("font_%s_base64 not found" % name)
        fonts[name] = "".join(re.findall(r"\"([^\"]*)\"", m.group(1)))
    return fonts


def hash_table(source):
    rows = re.findall(r"\{\s*0x([0-9a-fA-F]{8}),\s*(\d+),\s*(\d+),\s*(\w+),\s*\"(\w+)\"\s*\}",
                      source)
    if not rows:
        raise SystemExit("test table not found")
    return {name: {"hash": int(h, 16), "width": int(w), "height": int(hh)}
            for h, w, hh, _, name in rows}


def main():
    path = sys.argv[1] if len(sys.argv) > 1 else DEFAULT
    source = open(path, encoding="utf-8", errors="replace").read()
    out_dir = os.path.join(HERE, "fonts")
    os.makedirs(out_dir, exist_ok=True)
    for name, text in literals(source).items():
        data = decode(text)
        with open(os.path.join(out_dir, "font_%s.ttf" % name), "wb") as f:
            f.write(data)
        print("font_%s.ttf %d bytes" % (name, len(data)))
    table = hash_table(source)
    with open(os.path.join(HERE, "reference_hashes.json"), "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
        f.write("\n")
    print("reference_hashes.json %d tests" % len(table))


if __name__ == "__main__":
    main()
