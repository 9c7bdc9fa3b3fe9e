"""Writes the Matrix Market fixtures under tests/golden/mm/ and, where the
reference's own reader is built (oracle/_ref/libmmref.so, compiled from the
reference's CG/src/mmio_wrapper.cpp and CG/src/mmio.c by `make -C oracle ref`),
records what it returns for each file in tests/golden/mm_expected.json.

The files are small hand-made cases of what the reference's
loadMMSparseMatrix (CG/src/mmio_wrapper.cpp:170-369) accepts and refuses:
symmetric / general / integer / skew-symmetric storage, 0- and 1-based
indices, comments and a blank size line, upper-case banner words; pattern,
complex, array, repeated entries and both-based indices are refused.

Run: python tests/golden/make_mm_fixtures.py
"""
import ctypes
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "mm")
REF_LIB = os.path.join(REPO, "oracle", "_ref", "libmmref.so")


def spd(n, seed):
    rng = np.random.default_rng(seed)
    a = rng.uniform(-1, 1, (n, n)).round(3)
    a = (a + a.T) / 2
    a[np.abs(a) < 0.4] = 0.0
    a += np.diag(np.abs(a).sum(axis=1) + 1.0)
    return a


def entries(a, lower_only):
    out = []
    for i in range(a.shape[0]):
        for j in range(a.shape[1]):
            if a[i, j] != 0 and (not lower_only or j <= i):
                out.append((i, j, a[i, j]))
    return out


def fmt_value(v, exp):
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    return format(float(v), ".15e") if exp else repr(float(v))


def write(name, banner, n, rows, comments=("% written by make_mm_fixtures.py",), size_line=None, base=1,
          exp=False):
    path = os.path.join(OUT, name)
    with open(path, "w") as f:
        f.write(banner + "\n")
        for c in comments:
            f.write(c + "\n")
        f.write(size_line if size_line is not None else f"{n} {n} {len(rows)}\n")
        for r in rows:
            i, j = r[0] + base, r[1] + base
            f.write(f"{i} {j}" + "".join(" " + fmt_value(v, exp) for v in r[2:]) + "\n")
    return path


def make():
    os.makedirs(OUT, exist_ok=True)
    rng = np.random.default_rng(7)
    a5 = spd(5, 1)
    write("spd_sym_real.mtx", "%%MatrixMarket matrix coordinate real symmetric", 5, entries(a5, True),
          comments=("% 5x5 SPD, lower triangle", "%", "% values with exponents below"), exp=True)
    a6 = spd(6, 2)
    e6 = entries(a6, False)
    rng.shuffle(e6)
    write("spd_general_shuffled.mtx", "%%MatrixMarket matrix coordinate real general", 6, [tuple(e) for e in e6])
    ai = np.round(spd(7, 3) * 4)
    write("spd_sym_integer.mtx", "%%MatrixMarket matrix coordinate integer symmetric", 7,
          [(i, j, int(v)) for i, j, v in entries(ai, True)])
    s4 = np.array([[0, 1.5, 0, -2], [-1.5, 0, 0.25, 0], [0, -0.25, 0, 3], [2, 0, -3, 0]])
    write("skew.mtx", "%%MatrixMarket matrix coordinate real skew-symmetric", 4,
          [(i, j, v) for i, j, v in entries(s4, True) if i != j])
    a4 = spd(4, 4)
    write("base0_general.mtx", "%%MatrixMarket matrix coordinate real general", 4, entries(a4, False), base=0)
    write("upper_banner_blank_size.mtx", "%%MatrixMarket MATRIX Coordinate REAL General", 5,
          entries(a5, False), comments=("% the size line is preceded by an empty line",),
          size_line=f"\n5 5 {len(entries(a5, False))}\n")
    # the dimension index never occurs (last row and column empty): the reference's base test says 0
    z = np.zeros((5, 5))
    z[:4, :4] = spd(4, 5)
    write("no_last_index.mtx", "%%MatrixMarket matrix coordinate real general", 5, entries(z, False))
    # refused ones
    write("refuse_pattern.mtx", "%%MatrixMarket matrix coordinate pattern symmetric", 3,
          [(0, 0), (1, 1), (2, 1), (2, 2)])
    write("refuse_complex.mtx", "%%MatrixMarket matrix coordinate complex general", 2,
          [(0, 0, 1.0, 0.5), (1, 1, 2.0, -0.5)])
    with open(os.path.join(OUT, "refuse_array.mtx"), "w") as f:
        f.write("%%MatrixMarket matrix array real general\n2 2\n1.0\n0.0\n0.0\n1.0\n")
    write("refuse_duplicate.mtx", "%%MatrixMarket matrix coordinate real general", 3,
          [(0, 0, 2.0), (1, 1, 2.0), (2, 2, 2.0), (1, 1, 0.5)])
    write("refuse_both_bases.mtx", "%%MatrixMarket matrix coordinate real general", 3,
          [(0, 0, 2.0), (1, 1, 2.0), (2, 2, 2.0)], base=0, size_line="3 3 4\n")
    with open(os.path.join(OUT, "refuse_both_bases.mtx"), "a") as f:
        f.write("3 3 1.0\n")


def reference_outputs():
    if not os.path.exists(REF_LIB):
        raise SystemExit(f"{REF_LIB} is not built (make -C oracle ref)")
    L = ctypes.CDLL(REF_LIB)
    ip, dp = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
    L.ref_load_mm.argtypes = [ctypes.c_char_p, ip, ip, ip, ctypes.POINTER(dp), ctypes.POINTER(ip), ctypes.POINTER(ip)]
    res = {}
    for name in sorted(os.listdir(OUT)):
        m, n, nnz = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        v, rp, cp = dp(), ip(), ip()
        rc = L.ref_load_mm(os.path.join(OUT, name).encode(), ctypes.byref(m), ctypes.byref(n), ctypes.byref(nnz),
                           ctypes.byref(v), ctypes.byref(rp), ctypes.byref(cp))
        if rc != 0:
            res[name] = {"accepted": False}
            continue
        k = nnz.value
        res[name] = {"accepted": True, "m": m.value, "n": n.value, "nnz": k,
                     "rowptr": [rp[i] for i in range(m.value + 1)], "col": [cp[i] for i in range(k)],
                     "val": [v[i] for i in range(k)]}
    return res


if __name__ == "__main__":
    make()
    exp = reference_outputs()
    with open(os.path.join(HERE, "mm_expected.json"), "w") as f:
        json.dump(exp, f, indent=1)
        f.write("\n")
    print({k: v["accepted"] for k, v in exp.items()})
