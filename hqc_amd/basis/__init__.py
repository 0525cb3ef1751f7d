"""CP2K-format Gaussian basis sets for hydrogen (gth-szv, gth-dzv, gth-dzvp).

Mirrors hqc/basis/parse.py (load_as_str :32, parse_quant_num :116, parse_gto
:143, normalize_gto_coeff :192): same file format, same normalisation, same
exceptions (FileNotFoundError for an unknown basis name, ValueError for a
malformed block).
"""
import math
import os

import numpy as np

DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


def load_as_str(element, basis):
    """Text of one element's block, starting with '#BASIS SET' (parse.py:32-61)."""
    path = basis if os.path.isfile(basis) else os.path.join(DATA_DIR, basis + ".dat")
    if not os.path.isfile(path):
        raise FileNotFoundError(f"No '{basis}.dat' in '{DATA_DIR}'.")
    with open(path) as f:
        chunks = f.read().split("#BASIS SET\n")
    for c in chunks:
        if c.startswith(element):
            return "#BASIS SET\n" + c
    raise ValueError(f"No '{element}' in '{path}'.")


def _lines(text):
    return [ln.strip() for ln in text.split("\n") if ln.strip()]


def parse_sets(element, basis):
    """[(quant_num, table)] with table rows 'alpha c_1 .. c_k' (parse.py:116-175)."""
    lines = _lines(load_as_str(element, basis))
    out, pos = [], 3
    for _ in range(int(lines[2])):
        q = np.array(lines[pos].split(), dtype=int)
        if len(q) != 5 + q[2] - q[1]:
            raise ValueError(f"Line {lines[pos]} has incorrect number of l orbitals.")
        tab = np.array([ln.split() for ln in lines[pos + 1:pos + 1 + q[3]]], dtype=float)
        if tab.shape != (q[3], q[4:].sum() + 1):
            raise ValueError(f"Line {lines[pos + 1]} has incorrect number of coefficents.")
        out.append((q, tab))
        pos += q[3] + 1
    return out


def _double_factorial(k):
    r = 1
    while k > 1:
        r *= k
        k -= 2
    return r


def normalized_functions(element, basis):
    """List of (l, alphas, normalised radial coefficients) per contracted function,
    in file order (parse.py:192-239)."""
    funcs = []
    for q, tab in parse_sets(element, basis):
        a = tab[:, 0]
        lo = int(q[1])
        col = 1
        for l in range(lo, int(q[2]) + 1):
            for _ in range(int(q[4 + l - lo])):
                c = tab[:, col]
                col += 1
                overlap = 2.0 ** (l + 1.5) * np.sum(
                    np.outer(c, c) * (a[:, None] + a[None, :]) ** (-l - 1.5)
                    * np.outer(a, a) ** (l / 2 + 0.75))
                prim = 2.0 ** (l + 1.75) / np.pi ** 0.25 / math.sqrt(_double_factorial(2 * l + 1))
                funcs.append((l, a, c * a ** (l / 2 + 0.75) * prim / math.sqrt(overlap)))
    return funcs


class HBasis:
    """Per-atom AO layout for the GPU kernels.

    alpha (na,) all primitive exponents; coef (na, nloc) coefficient of each
    primitive in each local AO, with the real solid-harmonic constant folded in
    (1/sqrt(4 pi) for s, sqrt(3/(4 pi)) for p; hqc/pbc/solver.py:128-135);
    comp (nloc,) -1 for s, 0/1/2 for p_x/p_y/p_z.  AO order within an atom:
    contracted functions in file order, p functions as x, y, z
    (hqc/pbc/gto.py:268-300, hqc/pbc/solver.py:316-321).
    """

    def __init__(self, basis, element="H"):
        funcs = normalized_functions(element, basis)
        ls = [f[0] for f in funcs]
        if min(ls) != 0:
            raise ValueError("l_min != 0")
        if max(ls) > 1:
            raise ValueError("l_max > 1 is not supported yet.")
        if any(b < a for a, b in zip(ls, ls[1:])):
            raise ValueError("basis with p before s functions is not supported")
        # primitives: concatenate each set's exponents once (shared by its functions)
        alphas, owner = [], []
        for si, (q, tab) in enumerate(parse_sets(element, basis)):
            alphas.extend(tab[:, 0].tolist())
            owner.extend([si] * tab.shape[0])
        self.alpha = np.array(alphas)
        cols, comp = [], []
        fi = 0
        for si, (q, tab) in enumerate(parse_sets(element, basis)):
            nf = int(q[4:].sum())
            for _ in range(nf):
                l, _, c = funcs[fi]
                fi += 1
                full = np.zeros(len(self.alpha))
                full[np.array(owner) == si] = c
                if l == 0:
                    cols.append(full / math.sqrt(4 * math.pi))
                    comp.append(-1)
                else:
                    for d in range(3):
                        cols.append(full * math.sqrt(3 / (4 * math.pi)))
                        comp.append(d)
        self.coef = np.stack(cols, axis=1)
        self.comp = np.array(comp, dtype=np.int32)
        self.l_max = max(ls)
        self.nloc = self.coef.shape[1]
        self.na = len(self.alpha)
