"""Text files of the reference's integral keywords Details and Read.

* ``write_details`` follows details_integrals / details_integrals_SAO
  (src/Tools/print/Details.f90:1-160), called by the driver when the keyword
  Details is set (main.f90:455-461), into a directory (the reference's
  tmp_file_name):
    OV.dat, KI.dat, NA.dat    one line per element (i, j), j >= i, Fortran format
                              '(i3,i3,6x,200es24.16e3)'; the full layout leaves out
                              |x| <= 1e-15, the SAO layout writes every element
    OV_matrix.dat, KI_matrix.dat, NA_matrix.dat
                              header '(15x,1000(i3,15x))', then row i as
                              '(i3,6x,1000(f16.12,2x))'
    ERI.dat                   'i j k l ERI(i,j,k,l)' for |ERI| > 1e-15, i outermost,
                              i over the first index's range (nBas, or nfuc for SAO),
                              list-directed (``write(1,*)``)
  List-directed output is laid out by the compiler; ERI.dat follows gfortran's
  layout (the compiler of the reference's Makefile): each integer in 12
  columns, each real as G25.17E3 after a separator blank.
* ``read_integrals`` follows read_integrals_from_file
  (src/Energies/Energies_tools.f90:9-78), keyword Read: S, T, V symmetric from
  the (i, j) lines of OV.dat, KI.dat, NA.dat, Hc = T + V, and
  G(mu, nu, la, si) from ERI.dat (list-directed input: any spacing).

Indices in the files are 1-based (Fortran); arrays here are numpy arrays
indexed the same way as the reference's, 0-based: S[i, j] = S(i+1, j+1),
and the ERI as eri[i, j, k, l] = ERI(i+1, j+1, k+1, l+1).
"""
from __future__ import annotations

import math
import os

import numpy as np

__all__ = ["write_details", "read_integrals", "fortran_es", "fortran_f", "list_directed_real"]


def fortran_es(x: float, w: int = 24, d: int = 16, e: int = 3) -> str:
    """Fortran ESw.dEe: one digit before the point, d after, exponent of e digits."""
    x = float(x)
    if x == 0.0:
        m, ex = "0." + "0" * d, 0
        sign = "-" if math.copysign(1.0, x) < 0 else ""
    else:
        txt = f"{abs(x):.{d}e}"
        mant, exs = txt.split("e")
        m, ex = mant, int(exs)
        sign = "-" if x < 0 else ""
    s = f"{sign}{m}E{'-' if ex < 0 else '+'}{abs(ex):0{e}d}"
    return s.rjust(w) if len(s) <= w else "*" * w


def fortran_f(x: float, w: int = 16, d: int = 12) -> str:
    """Fortran Fw.d."""
    s = f"{float(x):.{d}f}"
    if s.startswith("-") and float(s) == 0.0 and math.copysign(1.0, float(x)) > 0:
        s = s[1:]
    return s.rjust(w) if len(s) <= w else "*" * w


def list_directed_real(x: float) -> str:
    """gfortran's list-directed REAL(8) item: a separator blank, then G25.17E3
    (F editing for 0.1 <= |x| < 1e17 with the exponent field left blank, E
    editing otherwise)."""
    x = float(x)
    if x == 0.0:
        return " " + "0.0000000000000000".rjust(20) + " " * 5
    mant, exs = f"{x:.16e}".split("e")
    ex = int(exs)                          # |x| rounded to 17 digits = m 10^ex, 1 <= |m| < 10
    if -1 <= ex < 17:
        body = f"{x:.{16 - ex}f}"
        return " " + body.rjust(20) + " " * 5
    body = f"{mant}E{'-' if ex < 0 else '+'}{abs(ex):03d}"
    return " " + body.rjust(25)


def _pairs_file(path: str, M: np.ndarray, skip_small: bool) -> None:
    n = M.shape[0]
    with open(path, "w") as f:
        for i in range(n):
            for j in range(i, n):
                v = M[i, j]
                if skip_small and not abs(v) > 1e-15:
                    continue
                f.write(f"{i + 1:3d}{j + 1:3d}      {fortran_es(v)}\n")


def _matrix_file(path: str, M: np.ndarray) -> None:
    # a position edit (15x, 2x) after the last item writes nothing: the record
    # ends with the last value
    n = M.shape[0]
    with open(path, "w") as f:
        f.write(" " * 15 + (" " * 15).join(f"{i + 1:3d}" for i in range(n)) + "\n")
        for i in range(n):
            f.write(f"{i + 1:3d}      " + "  ".join(fortran_f(M[i, j]) for j in range(n)) + "\n")


def write_details(dirpath: str, S: np.ndarray, T: np.ndarray, V: np.ndarray, eri: np.ndarray,
                  sao: bool = False) -> None:
    """The files of the keyword Details.  eri[i, j, k, l] = ERI(i+1, j+1, k+1, l+1)
    with i over nBas (full layout) or nfuc (SAO layout, sao=True)."""
    os.makedirs(dirpath, exist_ok=True)
    skip = not sao
    for name, M in (("OV", S), ("KI", T), ("NA", V)):
        M = np.asarray(M, dtype=np.float64)
        _pairs_file(os.path.join(dirpath, f"{name}.dat"), M, skip)
        _matrix_file(os.path.join(dirpath, f"{name}_matrix.dat"), M)
    E = np.asarray(eri, dtype=np.float64)
    n1, nb = E.shape[0], E.shape[1]
    ints = lambda *v: "".join(f"{x:12d}" for x in v)
    with open(os.path.join(dirpath, "ERI.dat"), "w") as f:
        for i in range(n1):
            Ei = E[i]
            idx = np.argwhere(np.abs(Ei) > 1e-15)            # (j, k, l) in C order: l fastest, as the loops
            for j, k, l in idx:
                f.write(ints(i + 1, j + 1, k + 1, l + 1) + list_directed_real(Ei[j, k, l]) + "\n")


def read_integrals(dirpath: str, nbas: int):
    """read_integrals_from_file: S, T, V (symmetric), Hc = T + V and the ERI
    G[mu, nu, la, si] (nBas^4, zero where ERI.dat has no line)."""
    def pairs(name):
        M = np.zeros((nbas, nbas))
        with open(os.path.join(dirpath, f"{name}.dat")) as f:
            for line in f:
                t = line.split()
                if len(t) < 3:
                    continue
                mu, nu, v = int(t[0]) - 1, int(t[1]) - 1, float(t[2].replace("D", "E").replace("d", "e"))
                M[mu, nu] = v
                M[nu, mu] = v
        return M
    S, T, V = pairs("OV"), pairs("KI"), pairs("NA")
    G = np.zeros((nbas, nbas, nbas, nbas))
    with open(os.path.join(dirpath, "ERI.dat")) as f:
        for line in f:
            t = line.split()
            if len(t) < 5:
                continue
            G[int(t[0]) - 1, int(t[1]) - 1, int(t[2]) - 1, int(t[3]) - 1] = float(t[4].replace("D", "E"))
    return S, T, V, T + V, G
