"""``spin_basis_general`` for spin-1/2 (src/quspin/basis/basis_general/
spin.py:17-447) on the device hcb core."""

import numpy as np

from .base_general import hcb_basis_general

S_dict = {(str(i) + "/2" if i % 2 == 1 else str(i // 2)): (i + 1, i / 2.0) for i in range(1, 10001)}


class spin_basis_general(hcb_basis_general):
    """Spin-1/2 basis with user-defined symmetries (maps and quantum numbers).

    Operator strings ``I + - x y z`` with the ``pauli`` convention of the
    reference: ``pauli=0`` spin operators, ``pauli=1`` Pauli matrices (every
    non-identity character scales by 2), ``pauli=-1`` Pauli matrices with
    sigma^+- un-scaled (spin.py:216-276).  Higher spins are out of scope.
    """

    def __init__(self, N, Nup=None, m=None, S="1/2", pauli=1, Ns_block_est=None, make_basis=True,
                 block_order=None, **blocks):
        self._S = S
        self._pauli = int(pauli)
        sps, S_val = S_dict[S]
        if self._pauli not in [-1, 0, 1]:
            raise ValueError("Invalid value for optional argument pauli. Allowed values are the integers [-1,0,1].")
        if sps != 2:
            raise NotImplementedError("only S='1/2' is supported by the device spin core")
        blocks.pop("_Np", None)
        if Nup is not None and m is not None:
            raise ValueError("Cannot use Nup and m at the same time")
        if m is not None and Nup is None:
            if m < -S_val or m > S_val:
                raise ValueError("m must be between -S and S")
            Nup = int((m + S_val) * N)
        try:
            Nup_iter = iter(Nup)
            M = int(2 * S_val * N)
            Nup = [(M + (n + 1) if n < 0 else n) for n in Nup_iter]
        except TypeError:
            if Nup is not None and Nup < 0:
                Nup = int(2 * S_val * N) + Nup + 1
        self._pcon_args = dict(N=N, Nup=Nup, S=self._S)
        hcb_basis_general.__init__(self, N, Nb=Nup, Ns_block_est=Ns_block_est, _make_basis=make_basis,
                                   block_order=block_order, **blocks)
        self._allowed_ops = set(["I", "+", "-", "x", "y", "z"])
        self._operators = ("availible operators for spin_basis_1d:\n\tI: identity \n\t+: raising operator"
                           "\n\t-: lowering operator\n\tx: x pauli/spin operator\n\ty: y pauli/spin operator"
                           "\n\tz: z pauli/spin operator")

    def _scale(self, opstr):
        if self._pauli == 1:
            return 1 << len(opstr.replace("I", ""))
        if self._pauli == -1:
            return 1 << len(opstr.replace("I", "").replace("+", "").replace("-", ""))
        return 1

    def _opstr_prepare(self, opstr, indx, J):
        return opstr, indx, J * self._scale(opstr)

    def _hc_opstr(self, op):
        op = list(op)
        op[0] = list(op[0].replace("+", "%").replace("-", "+").replace("%", "-"))
        op[0].reverse()
        op[0] = "".join(op[0])
        op[1] = list(op[1])
        op[1].reverse()
        op[1] = tuple(op[1])
        op[2] = np.conjugate(op[2])
        return self._sort_opstr(op)

    def _sort_opstr(self, op):
        op = list(op)
        z = list(zip(op[0], op[1]))
        if z:
            z.sort(key=lambda x: x[1])
            a, b = zip(*z)
            op[0] = "".join(a)
            op[1] = tuple(b)
        return tuple(op)

    def __repr__(self):
        return "< instance of 'qspin.basis.general_hcb' with {0} states >".format(self._Ns)
