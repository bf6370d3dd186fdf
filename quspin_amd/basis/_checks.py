"""Hermiticity, symmetry and particle-conservation checks of operator lists
(the ``check_hermitian`` / ``check_symm`` / ``check_pcon`` calls made by the
operator constructors, e.g. quantum_LinearOperator_core.py:129-136).

Terms are expanded into the +/-/z/n alphabet, sorted by site (with the
fermionic exchange sign for fermion bases) and summed per
(operator string, sites, drive); the checks compare these sums.
"""

import numpy as np


def _expand_chars(basis, opstr, indx, J):
    """x/y -> +/- combinations (spin.py:359-428 for spins; fermion x/y are
    Majorana operators: x = c + c^+, y = -i(c - c^+))."""
    terms = [("", (), complex(J))]
    fermion = getattr(basis, "_fermion", False)
    for ch, i in zip(opstr, indx):
        new = []
        for s, ix, c in terms:
            if ch == "I":
                new.append((s, ix, c))
            elif ch == "x":
                f = 1.0 if fermion else 0.5
                new.append((s + "+", ix + (i,), c * f))
                new.append((s + "-", ix + (i,), c * f))
            elif ch == "y":
                if fermion:
                    new.append((s + "+", ix + (i,), c * 1j))
                    new.append((s + "-", ix + (i,), c * -1j))
                else:
                    new.append((s + "+", ix + (i,), c * -0.5j))
                    new.append((s + "-", ix + (i,), c * 0.5j))
            else:
                new.append((s + ch, ix + (i,), c))
        terms = new
    return terms


def _canonical(basis, static, dynamic):
    """dict (opstr, sites, drive key) -> summed coupling."""
    out = {}

    def add(opstr, indx, J, key):
        opstr, indx, J = basis._opstr_prepare(opstr, indx, J)
        for s, ix, c in _expand_chars(basis, opstr, indx, J):
            s2, ix2, c2 = basis._sort_opstr((s, ix, c))
            k = (s2, tuple(int(v) for v in ix2), key)
            out[k] = out.get(k, 0) + c2

    for opstr, bonds in static:
        for b in bonds:
            add(opstr, b[1:], b[0], None)
    for item in dynamic:
        opstr, bonds, f, f_args = item[0], item[1], item[2], tuple(item[3])
        for b in bonds:
            add(opstr, b[1:], b[0], (f, f_args))
    return {k: v for k, v in out.items() if abs(v) > 1e-13}


def _equal(d1, d2, tol=1e-11):
    keys = set(d1) | set(d2)
    return all(abs(d1.get(k, 0) - d2.get(k, 0)) <= tol for k in keys)


def check_hermitian(basis, static, dynamic):
    d = _canonical(basis, static, dynamic)
    h = {}
    for (s, ix, key), c in d.items():
        s2, ix2, c2 = basis._hc_opstr((s, ix, c))
        k = (s2, tuple(ix2), key)
        h[k] = h.get(k, 0) + c2
    if not _equal(d, h):
        raise TypeError("Hermiticity check failed! To turn off check, use check_herm=False in hamiltonian.")
    print("Hermiticity check passed!")


def check_pcon(basis, static, dynamic):
    if not basis._check_pcon:
        return
    d = _canonical(basis, static, dynamic)
    spinful = getattr(basis, "_spinful", False)
    half = basis.N // 2
    for (s, ix, key), c in d.items():
        if spinful:
            du = sum((1 if ch == "+" else -1) for ch, i in zip(s, ix) if ch in "+-" and i < half)
            dd = sum((1 if ch == "+" else -1) for ch, i in zip(s, ix) if ch in "+-" and i >= half)
            bad = du != 0 or dd != 0
        else:
            bad = s.count("+") != s.count("-")
        if bad:
            raise TypeError("Particle conservation check failed! To turn off check, use check_pcon=False in "
                            "hamiltonian.")
    print("Particle conservation check passed!")


def check_symm(basis, static, dynamic):
    if not len(basis._pers):
        return
    d = _canonical(basis, static, dynamic)
    fermion = getattr(basis, "_fermion", False)
    for m in basis._maps:
        m = np.asarray(m)
        inv = m < 0
        perm = np.where(inv, -(m + 1), m)
        pinv = np.empty_like(perm)
        pinv[perm] = np.arange(len(perm))
        t = {}
        for (s, ix, key), c in d.items():
            s2, ix2 = [], []
            cc = c
            for ch, j in zip(s, ix):
                i = int(pinv[j])
                if inv[i]:
                    if ch == "z":
                        cc = -cc
                    elif ch == "+":
                        ch = "-"
                    elif ch == "-":
                        ch = "+"
                    elif ch == "n":
                        return  # population inversion of n is not a single term; skip the check
                s2.append(ch)
                ix2.append(i)
            s3, ix3, c3 = basis._sort_opstr(("".join(s2), tuple(ix2), cc))
            k = (s3, tuple(ix3), key)
            t[k] = t.get(k, 0) + c3
        if not _equal(d, t):
            raise TypeError("Symmetry check failed! To turn off check, use check_symm=False in hamiltonian.")
    print("Symmetry checks passed!")


def is_symmetric(basis, op_list):
    """True when the (opstr, indx, J) list commutes with every symmetry map
    of ``basis`` (then the gather form of H.v is exact); False when it does
    not or when the check cannot decide."""
    if not len(basis._pers):
        return True
    static = [[opstr, [[J] + list(indx)]] for opstr, indx, J in op_list]
    try:
        d = _canonical(basis, static, [])
    except Exception:
        return False
    for m in basis._maps:
        m = np.asarray(m)
        inv = m < 0
        perm = np.where(inv, -(m + 1), m)
        pinv = np.empty_like(perm)
        pinv[perm] = np.arange(len(perm))
        t = {}
        for (s, ix, key), c in d.items():
            s2, ix2 = [], []
            cc = c
            for ch, j in zip(s, ix):
                i = int(pinv[j])
                if inv[i]:
                    if ch == "z":
                        cc = -cc
                    elif ch == "+":
                        ch = "-"
                    elif ch == "-":
                        ch = "+"
                    elif ch == "n":
                        return False
                s2.append(ch)
                ix2.append(i)
            s3, ix3, c3 = basis._sort_opstr(("".join(s2), tuple(ix2), cc))
            k = (s3, tuple(ix3), key)
            t[k] = t.get(k, 0) + c3
        if not _equal(d, t):
            return False
    return True
