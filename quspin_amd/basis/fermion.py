"""Fermion bases (src/quspin/basis/basis_general/fermion.py:20-1174) on the
device cores.

Jordan-Wigner convention: modes are ordered by site index (spinful: up sites
0..L-1, then down sites L..2L-1), and c_j / c_j^+ carry (-1)^(number of
occupied modes of lower index).  Spectra and other gauge-invariant quantities
match the reference; the per-element sign convention of the reference's
native cores is not observable from its sources.
"""

import numpy as np
from scipy.special import comb

from .. import _lib as L
from .base_general import basis_general, get_basis_type


def _anticommute_sort(op):
    """Sort an operator string by site with the fermionic exchange sign of
    each swapped pair of +/- characters (fermion.py:206-241)."""
    opstr, indx, J = op[0], list(op[1]), op[2]
    z = list(zip(opstr, indx))
    n = len(z)
    swapped = True
    sign = 1
    while swapped:
        swapped = False
        for i in range(n - 1):
            if z[i][1] > z[i + 1][1]:
                z[i], z[i + 1] = z[i + 1], z[i]
                swapped = True
                if z[i][0] in "+-" and z[i + 1][0] in "+-":
                    sign *= -1
    if z:
        a, b = zip(*z)
        return ("".join(a), tuple(b), J * sign)
    return (opstr, tuple(indx), J)


class spinless_fermion_basis_general(basis_general):
    """Spinless fermions with user-defined site-permutation symmetries.
    Operators ``I + - n z x y`` (z = n - 1/2, x = c + c^+, y = -i(c - c^+))."""

    _fermion = True

    def __init__(self, N, Nf=None, nf=None, Ns_block_est=None, make_basis=True, block_order=None, **blocks):
        blocks.pop("_Np", None)
        basis_general.__init__(self, N, block_order=block_order, **blocks)
        self._check_pcon = False
        self._count_particles = False
        if Nf is not None and nf is not None:
            raise ValueError("cannot use 'nf' and 'Nf' simultaineously.")
        if Nf is None and nf is not None:
            Nf = int(nf * N)
        if Nf is None:
            self._Ns = 1 << N
        elif type(Nf) is int:
            self._check_pcon = True
            self._get_proj_pcon = True
            self._Ns = comb(N, Nf, exact=True)
        else:
            Nf = list(Nf)
            self._Ns = 0
            for _nf in Nf:
                if _nf > N or _nf < 0:
                    raise ValueError("particle number Nf must satisfy: 0 <= Nf <= N")
                self._Ns += comb(N, _nf, exact=True)
        self._pcon_args = dict(N=N, Nf=Nf)
        if len(self._pers) > 0:
            if Ns_block_est is None:
                self._Ns = int(float(self._Ns) / np.multiply.reduce(self._pers)) * 4
            else:
                if type(Ns_block_est) is not int:
                    raise TypeError("Ns_block_est must be integer value.")
                if Ns_block_est <= 0:
                    raise ValueError("Ns_block_est must be an integer > 0")
                self._Ns = Ns_block_est
        self._basis_dtype = get_basis_type(N, None, 2)
        self._sps = 2
        self._Ns_block_est = self._Ns
        self._N = N
        self._Np = Nf
        self._allowed_ops = set(["I", "n", "+", "-", "z", "x", "y"])
        self._core = self._make_core(L.KIND_SPINLESS, N)
        if make_basis:
            self.make()
        else:
            self._Ns = 1

    def _hc_opstr(self, op):
        op = list(op)
        op[0] = list(op[0].replace("+", "%").replace("-", "+").replace("%", "-"))
        op[0].reverse()
        op[0] = "".join(op[0])
        op[1] = tuple(reversed(list(op[1])))
        op[2] = np.conjugate(op[2])
        return _anticommute_sort(op)

    def _sort_opstr(self, op):
        return _anticommute_sort(op)


def process_spinful_map(N, map, q):
    """Simple (length-N) map -> advanced (length-2N) map; an inverted site
    swaps the species (fermion.py:275-284)."""
    map = np.asarray(map, dtype=np.int32)
    if len(map) == N:
        i_map = map.copy()
        i_map[map < 0] = -(i_map[map < 0] + 1) + N
        adv_map = np.hstack((i_map, (i_map + N) % (2 * N)))
        return adv_map, q
    return map, q


class spinful_fermion_basis_general(spinless_fermion_basis_general):
    """Spinful fermions: state = down + (up << L); up sites 0..L-1, down sites
    L..2L-1 in the advanced notation.  Operator strings in the simple notation
    carry one '|' separating the up and down parts (fermion.py:287-1174)."""

    _spinful = True

    def __init__(self, N, Nf=None, nf=None, Ns_block_est=None, simple_symm=True, make_basis=True, block_order=None,
                 double_occupancy=True, **blocks):
        blocks.pop("_Np", None)
        if Nf is not None and nf is not None:
            raise ValueError("cannot use 'nf' and 'Nf' simultaineously.")
        if Nf is None and nf is not None:
            Nf = (int(nf[0] * N), int(nf[1] * N))
        if any((type(m) is not tuple) and (len(m) != 2) for m in blocks.values() if m is not None):
            raise ValueError("blocks must contain tuple: (map,q).")
        self._simple_symm = simple_symm
        if simple_symm:
            if any(len(item[0]) != N for item in blocks.values() if item is not None):
                raise ValueError("Simple symmetries must have mapping of length N.")
        else:
            if any(len(item[0]) != 2 * N for item in blocks.values() if item is not None):
                raise ValueError("Advanced symmetries must have mapping of length 2*N.")
        blocks = {k: process_spinful_map(N, *v) for k, v in blocks.items() if v is not None}
        basis_general.__init__(self, 2 * N, block_order=block_order, **blocks)
        self._count_particles = False
        if Nf is None:
            self._Ns = (1 << N) ** 2
        else:
            if type(Nf) is tuple and len(Nf) == 2 and all(type(x) is int for x in Nf):
                self._get_proj_pcon = True
                self._Ns = comb(N, Nf[0], exact=True) * comb(N, Nf[1], exact=True)
                Nf = [Nf]
            else:
                Nf = list(Nf)
                if any((type(tup) is not tuple) and len(tup) != 2 for tup in Nf):
                    raise ValueError("Nf must be tuple of integers or iterable object of tuples.")
                self._Ns = 0
                for Nup, Ndown in Nf:
                    if Nup > N or Nup < 0 or Ndown > N or Ndown < 0:
                        raise ValueError("particle numbers in Nf must satisfy: 0 <= n <= N")
                    self._Ns += comb(N, Nup, exact=True) * comb(N, Ndown, exact=True)
        self._pcon_args = dict(N=N, Nf=Nf)
        if len(self._pers) > 0 or not double_occupancy:
            if Ns_block_est is None:
                self._Ns = int(float(self._Ns) / np.multiply.reduce(self._pers)) * 4
            else:
                if type(Ns_block_est) is not int:
                    raise TypeError("Ns_block_est must be integer value.")
                if Ns_block_est <= 0:
                    raise ValueError("Ns_block_est must be an integer > 0")
                self._Ns = Ns_block_est
        self._basis_dtype = get_basis_type(2 * N, None, 2)
        self._sps = 2
        self._N = 2 * N
        self._Ns_block_est = self._Ns
        self._Np = Nf
        self._double_occupancy = double_occupancy
        self._check_pcon = Nf is not None
        self._allowed_ops = set(["I", "n", "+", "-", "z", "x", "y"])
        self._core = self._make_core(L.KIND_SPINFUL, 2 * N, double_occupancy)
        if make_basis:
            self.make()
        else:
            self._Ns = 1

    def _simple_to_adv(self, op):
        opstr, indx = op[0], op[1]
        if opstr.count("|") == 0:
            raise ValueError("missing '|' charactor in: {0}, {1}".format(opstr, indx))
        i = opstr.index("|")
        indx = list(indx)
        left = tuple(indx[:i])
        right = tuple(j + self._N // 2 for j in indx[i:])
        a, b = opstr.split("|", 1)
        return (a + b, left + right) + tuple(op[2:])

    def _opstr_prepare(self, opstr, indx, J):
        if self._simple_symm:
            opstr, indx = self._simple_to_adv((opstr, indx))[:2]
        return opstr, indx, J

    def int_to_state(self, state, bracket_notation=True):
        state = int(state)
        Lh = self.N // 2
        up = [(state >> (self.N - 1 - i)) & 1 for i in range(Lh)]
        dn = [(state >> (Lh - 1 - i)) & 1 for i in range(Lh)]
        s = "|" + " ".join(map(str, up)) + ">|" + " ".join(map(str, dn)) + ">"
        return s if bracket_notation else s.replace(" ", "").replace("|", "").replace(">", "")

    def index(self, up_state, down_state=None):
        if down_state is None:
            return basis_general.index(self, up_state)
        if isinstance(up_state, str):
            up_state = int(up_state.replace(" ", ""), 2)
        if isinstance(down_state, str):
            down_state = int(down_state.replace(" ", ""), 2)
        return basis_general.index(self, (int(up_state) << (self.N // 2)) | int(down_state))
