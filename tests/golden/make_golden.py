"""Writes tests/golden/products.json: SHA-256 of product limbs (little-endian
bytes) for the seeded workloads, computed with libgmp's mpn_mul / mpn_sqr,
plus a few full small vectors.  The integer product is unique, so these pin
the reference's mul!/sqr! output (test/test_fpntt.jl compares with BigInt).

    python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "nativebigint.jl_amd"))

from nbhip import workloads as wl  # noqa: E402
from tests import gmp_ref  # noqa: E402


def sha(r: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(r, dtype="<u8").tobytes()).hexdigest()


def main():
    assert gmp_ref.available(), "libgmp is required to write the fixtures"
    out = {"generator": "libgmp mpn_mul/mpn_sqr via ctypes", "products": {}, "small": {}}
    shapes = [("single", 0, 4096, 4096, False), ("single", 1, 4096, 4096, True),
              ("large", 0, 1 << 20, 1 << 20, False), ("large", 1, 1 << 20, 1 << 18, False),
              ("large", 2, 1 << 22, 1 << 22, True),
              ("test", 10, 152, 152, False), ("test", 11, 160, 160, True), ("test", 12, 65536, 16384, False),
              ("test", 13, 144000, 144000, False)]
    for wname, job, m, n, sq in shapes:
        a = wl.operand(wname, job, 0, m)
        if sq:
            r = gmp_ref.sqr(a)
        else:
            b = wl.operand(wname, job, 1, n)
            r = gmp_ref.mul(a, b)
        key = f"{wname}/{job}/{m}x{n}{'/sqr' if sq else ''}"
        out["products"][key] = sha(r)
        print(key, out["products"][key][:16])
    for m, n in [(160, 152), (300, 64)]:
        a = wl.operand("test", 20 + m, 0, m)
        b = wl.operand("test", 20 + m, 1, n)
        out["small"][f"test/{20 + m}/{m}x{n}"] = [format(int(v), "016x") for v in gmp_ref.mul(a, b)]
    with open(os.path.join(HERE, "products.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
