"""Extracts the known-answer QR test data of the reference's test/qr.cc
(a_data: a 16x10 matrix with quadratic SPR monomial columns, x_data: the
solution; test/qr.cc:9-39) into qr_kat.json.  Reads the reference source as
text; run in a container where /root/reference exists."""
import json
import os
import re
import sys

SRC = "/root/reference/test/qr.cc"
HERE = os.path.dirname(os.path.abspath(__file__))


def block(text, name):
    m = re.search(r"%s\s*\[[^\]]*\](?:\[[^\]]*\])?\s*=\s*\{(.*?)\};" % name, text, re.S)
    if not m:
        sys.exit("cannot find %s in %s" % (name, SRC))
    return [float(v) for v in re.findall(r"[-+]?\d+\.\d+e[-+]\d+", m.group(1))]


def main():
    text = open(SRC).read()
    a = block(text, "a_data")
    x = block(text, "x_data")
    assert len(a) == 160 and len(x) == 10
    out = {"source": "test/qr.cc:9-56 (solveQR(A, A*x) must return x to 1e-15)",
           "a": [a[10 * i:10 * i + 10] for i in range(16)], "x": x, "tol": 1e-15}
    with open(os.path.join(HERE, "qr_kat.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
