"""Torus systems: supercell geometry and the Cartesian AO list.

Host-side mirror of the reference's input layer, which stays as it is in
the reference; the engine only needs its result, the AO list.

* ``read_unitcell``   follows ``build_super_molecule`` / ``read_2nd_line``
  (src/geometry/super_molecule.f90:1,189): line 1 the type, line 2
  ``nx [ny nz] d Lx [Ly Lz]``, keywords up to the first ``$$``, then the
  unit cell atoms up to the next ``$$``.  Cells are generated x fastest,
  then y, then z (super_molecule.f90:121-135).
* ``read_basis``      follows ``read_atom_shells_gen`` (src/basis/basis.f90:21):
  Dalton format, ``a Z`` header, ``$ S-TYPE FUNCTIONS`` / ``ne nc`` blocks.
* ``classification``  follows src/basis/class.f90:1-70: one AO per Cartesian
  component, order s; px,py,pz; dxx,dxy,dxz,dyy,dyz,dzz; ... per atom, and
  zero-coefficient primitives dropped (src/module/filter.f90).

The AO list is kept as a structure of arrays (centre, powers, primitive
offsets, exponents, coefficients), the layout the C ABI takes.
"""
from __future__ import annotations

import dataclasses
import math
import re
from typing import Dict, List, Sequence, Tuple

import numpy as np

# element label -> nuclear charge (the labels the reference's geometry reader knows
# for the light elements; only the charge is used on this path)
_Z = {"H": 1, "HE": 2, "LI": 3, "BE": 4, "B": 5, "C": 6, "N": 7, "O": 8, "F": 9, "NE": 10}

# Cartesian component order of src/basis/class.f90:15-21
_CART = {
    0: [(0, 0, 0)],
    1: [(1, 0, 0), (0, 1, 0), (0, 0, 1)],
    2: [(2, 0, 0), (1, 1, 0), (1, 0, 1), (0, 2, 0), (0, 1, 1), (0, 0, 2)],
    3: [(3, 0, 0), (2, 1, 0), (2, 0, 1), (1, 2, 0), (1, 1, 1), (1, 0, 2), (0, 3, 0),
        (0, 2, 1), (0, 1, 2), (0, 0, 3)],
}
_SHELL_L = {"S": 0, "P": 1, "D": 2, "F": 3}
_NDP = {"Tori1D": 1, "Tori2D": 2, "Tori3D": 3}


@dataclasses.dataclass
class AtomBasis:
    """Shells of one element: l -> (exponents[ne], coefficients[ne, nc])."""
    shells: Dict[int, Tuple[np.ndarray, np.ndarray]]


@dataclasses.dataclass
class TorusSystem:
    kind: str                     # Tori1D / Tori2D / Tori3D
    ndp: int                      # periodic directions
    L: Tuple[float, float, float]
    ncell: Tuple[int, int, int]   # nx, ny, nz
    spacing: float
    keywords: List[str]
    geom: np.ndarray              # (natom, 3)
    Z: np.ndarray                 # (natom,) int
    nauc: int                     # atoms per unit cell
    # AO list (structure of arrays)
    center: np.ndarray            # (nbas, 3) f64
    lxyz: np.ndarray              # (nbas, 3) int32
    prim_off: np.ndarray          # (nbas+1,) int32
    expo: np.ndarray              # (nprim_total,) f64
    coef: np.ndarray              # (nprim_total,) f64, raw library coefficients

    @property
    def nbas(self) -> int:
        return int(self.center.shape[0])

    @property
    def n_cell(self) -> int:
        return int(np.prod(self.ncell))

    @property
    def nfuc(self) -> int:
        return self.nbas // self.n_cell

    @property
    def natom(self) -> int:
        return int(self.geom.shape[0])

    @property
    def n_electrons(self) -> int:
        return int(self.Z.sum())

    @property
    def n_occ(self) -> int:
        return self.n_electrons // 2

    @property
    def a(self) -> Tuple[float, float, float]:
        return tuple(2.0 * math.pi / Lv if Lv > 0 else 0.0 for Lv in self.L)

    def nprim(self) -> np.ndarray:
        return np.diff(self.prim_off).astype(np.int32)

    def with_coef(self, coef: np.ndarray) -> "TorusSystem":
        return dataclasses.replace(self, coef=np.ascontiguousarray(coef, dtype=np.float64))


def read_basis(text: str) -> Dict[int, AtomBasis]:
    """Parse a Dalton-format ``Basis`` file (src/basis/basis.f90:21-90)."""
    out: Dict[int, AtomBasis] = {}
    lines = text.splitlines()
    i = 0
    cur = None
    while i < len(lines):
        s = lines[i].strip()
        i += 1
        if not s:
            continue
        if s[0] in "Aa" and re.match(r"^[Aa]\s+\d+", s):
            cur = int(s.split()[1])
            out[cur] = AtomBasis(shells={})
            continue
        if cur is None or s[0] != "$" or "-TYPE FUNCTIONS" not in s.upper():
            continue
        l = _SHELL_L[s[2].upper()]
        ne, nc = (int(v) for v in lines[i].split()[:2])
        i += 1
        ex = np.zeros(ne)
        co = np.zeros((ne, nc))
        for r in range(ne):
            vals = [float(v.replace("D", "E").replace("d", "e")) for v in lines[i].split()]
            i += 1
            ex[r] = vals[0]
            co[r, :] = vals[1:1 + nc]
        out[cur].shells[l] = (ex, co)
    return out


def read_unitcell(text: str):
    """Parse ``unitcell.mol`` (src/geometry/super_molecule.f90:189-259,281-402)."""
    lines = [ln for ln in text.splitlines()]
    kind = lines[0].split()[0]
    if kind not in _NDP:
        raise ValueError(f"unsupported calculation type {kind!r} (Tori1D/2D/3D only)")
    toks = lines[1].split()
    if kind == "Tori1D":
        ncell = (int(toks[0]), 1, 1)
        d = float(toks[1])
        L = (float(toks[2]), 0.0, 0.0)
    elif kind == "Tori2D":
        ncell = (int(toks[0]), int(toks[1]), 1)
        d = float(toks[2])
        L = (float(toks[3]), float(toks[4]), 0.0)
    else:
        ncell = (int(toks[0]), int(toks[1]), int(toks[2]))
        d = float(toks[3])
        L = (float(toks[4]), float(toks[5]), float(toks[6]))
    keywords = []
    j = 2
    while j < len(lines) and lines[j].strip() != "$$":
        if lines[j].strip():
            keywords.append(lines[j].strip())
        j += 1
    j += 1
    atoms = []
    while j < len(lines) and lines[j].strip() != "$$":
        t = lines[j].split()
        if t:
            atoms.append((t[0], float(t[1]), float(t[2]), float(t[3])))
        j += 1
    return kind, ncell, d, L, keywords, atoms


def classification(geom: np.ndarray, Z: Sequence[int], basis: Dict[int, AtomBasis]):
    """AO list of the supercell (src/basis/class.f90:1-70)."""
    center, lxyz, off, expo, coef = [], [], [0], [], []
    for k in range(geom.shape[0]):
        ab = basis[int(Z[k])]
        for l in sorted(ab.shells):
            ex, co = ab.shells[l]
            for fn in range(co.shape[1]):
                keep = co[:, fn] != 0.0                # filter: drop zero coefficients
                for comp in _CART[l]:
                    center.append(geom[k])
                    lxyz.append(comp)
                    expo.extend(ex[keep].tolist())
                    coef.extend(co[keep, fn].tolist())
                    off.append(len(expo))
    return (np.ascontiguousarray(center, dtype=np.float64).reshape(-1, 3),
            np.ascontiguousarray(lxyz, dtype=np.int32).reshape(-1, 3),
            np.ascontiguousarray(off, dtype=np.int32),
            np.ascontiguousarray(expo, dtype=np.float64),
            np.ascontiguousarray(coef, dtype=np.float64))


def build_system(unitcell_text: str, basis_text: str) -> TorusSystem:
    kind, ncell, d, L, keywords, atoms = read_unitcell(unitcell_text)
    basis = read_basis(basis_text)
    geom, Z = [], []
    nx, ny, nz = ncell
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                for (lab, x, y, z) in atoms:
                    geom.append((x + i * d, y + j * d, z + k * d))
                    Z.append(_Z[lab.upper()])
    geom = np.asarray(geom, dtype=np.float64)
    Z = np.asarray(Z, dtype=np.int32)
    center, lxyz, off, expo, coef = classification(geom, Z, basis)
    return TorusSystem(kind=kind, ndp=_NDP[kind], L=L, ncell=ncell, spacing=d,
                       keywords=keywords, geom=geom, Z=Z, nauc=len(atoms),
                       center=center, lxyz=lxyz, prim_off=off, expo=expo, coef=coef)


# ---------------------------------------------------------------------------
# The benchmark inputs (deterministic).  H basis: contracted s (13.01, 1.962,
# 0.4446), uncontracted s 0.122, uncontracted p 0.727; +d 0.5 for the 3D case.
H_BASIS = """a 1
$ S-TYPE FUNCTIONS
4 2
13.01 0.0197 0.0
1.962 0.138  0.0
0.4446 0.478 0.0
0.122  0.0   1.0
$ P-TYPE FUNCTIONS
1 1
0.727 1.0
"""

H_BASIS_D = H_BASIS + """$ D-TYPE FUNCTIONS
1 1
0.5 1.0
"""


def chain_1d(n: int, d: float, L: float, keywords=("SAO", "DIIS"), basis: str = H_BASIS) -> TorusSystem:
    """Tori1D ring of n H atoms, one per cell."""
    kw = "\n".join(keywords)
    text = f"Tori1D\n{n} {d} {L}\n{kw}\n$$\nH 0 0 0\n$$\n"
    return build_system(text, basis)


def lattice_2d_h2(n: int, d: float, L: float, bond: float = 1.4,
                  keywords=("SAO", "direct", "DIIS"), basis: str = H_BASIS) -> TorusSystem:
    kw = "\n".join(keywords)
    text = f"Tori2D\n{n} {n} {d} {L} {L}\n{kw}\n$$\nH 0 0 0\nH {bond} 0 0\n$$\n"
    return build_system(text, basis)


def lattice_3d_h2(n: int, d: float, L: float, bond: float = 1.4,
                  keywords=("SAO", "direct", "DIIS"), basis: str = H_BASIS_D) -> TorusSystem:
    kw = "\n".join(keywords)
    text = f"Tori3D\n{n} {n} {n} {d} {L} {L} {L}\n{kw}\n$$\nH 0 0 0\nH {bond} 0 0\n$$\n"
    return build_system(text, basis)


def workload(name: str) -> TorusSystem:
    """The named benchmark inputs."""
    if name == "h10_ring":          # Tori1D H10, d=1.8, L=18.0, 50 functions
        return chain_1d(10, 1.8, 18.0)
    if name == "h64_chain":         # Tori1D H64, d=1.8, L=115.2, 320 functions
        return chain_1d(64, 1.8, 115.2)
    if name == "h2_8x8":            # Tori2D 8x8 H2, d=3.6, L=28.8, 640 functions
        return lattice_2d_h2(8, 3.6, 28.8)
    if name == "h2_4x4x4_d":        # Tori3D 4^3 H2 + d, d=4.0, L=16.0, 1408 functions
        return lattice_3d_h2(4, 4.0, 16.0)
    raise KeyError(name)
