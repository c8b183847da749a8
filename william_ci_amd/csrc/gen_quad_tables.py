#!/usr/bin/env python3
"""Writes tcg_quad_tables.inc: the reference's literal 64- and 128-point
t-quadrature tables (nodes t_k and weights w_k of gauss_legendre_points,
src/module/precompute_t.f90:495-627 and :631-891), the rules used when
fast_quad is off.  The 64-point table is not a textbook rule (it deviates
from Gauss-Legendre by ~4e-5 relative), so its digits are carried as data.

Usage: python gen_quad_tables.py /root/reference > tcg_quad_tables.inc
"""
import re
import sys


def literal_lists(src: str):
    """(name, [decimal strings]) of every `name = [ ... ]` array literal in
    the subroutine gauss_legendre_points (values written as 0.123d0)."""
    body = src[src.index("subroutine gauss_legendre_points"):src.index("end subroutine gauss_legendre_points")]
    return [(m.group(1), re.findall(r"[0-9]+\.[0-9]+", m.group(2)))
            for m in re.finditer(r"(t_val|w_val)\s*=\s*\[(.*?)\]", body, re.S)]


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    src = open(f"{ref}/src/module/precompute_t.f90").read()
    lists = literal_lists(src)
    tables = {(name, len(digits)): digits for name, digits in lists}
    print("// Literal t-quadrature tables of the reference (gauss_legendre_points,")
    print("// src/module/precompute_t.f90:495-891), written by gen_quad_tables.py.")
    print("#pragma once")
    for n in (64, 128):
        for name, cname in (("t_val", "t"), ("w_val", "w")):
            d = tables[(name, n)]
            assert len(d) == n, (name, n, len(d))
            print(f"static const double kQuad{n}_{cname}[{n}] = {{")
            for i in range(0, n, 4):
                print("    " + ", ".join(d[i:i + 4]) + ",")
            print("};")


if __name__ == "__main__":
    main()
