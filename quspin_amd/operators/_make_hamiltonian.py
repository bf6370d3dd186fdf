"""Term-list consolidation and drive functions
(src/quspin/operators/_make_hamiltonian.py:19-80, _functions.py)."""

import numpy as np


def _consolidate_static(static_list):
    """Sum couplings of repeated (opstr, indx); drop |J| <= 10 eps
    (_make_hamiltonian.py:19-41).  Returns [(opstr, indx, J), ...]."""
    eps = 10 * np.finfo(np.float64).eps
    static_dict = {}
    for opstr, bonds in static_list:
        d = static_dict.setdefault(opstr, {})
        for bond in bonds:
            J = bond[0]
            indx = tuple(int(i) for i in bond[1:])
            d[indx] = d.get(indx, 0) + J
    out = []
    for opstr, d in static_dict.items():
        for indx, J in d.items():
            if np.abs(J) > eps:
                out.append((opstr, indx, J))
    return out


def _consolidate_dynamic(dynamic_list):
    """Group dynamic terms per (f, f_args) (_make_hamiltonian.py:44-73)."""
    eps = 10 * np.finfo(np.float64).eps
    dd = {}
    for opstr, bonds, f, f_args in dynamic_list:
        f_args = tuple(f_args)
        d = dd.setdefault((opstr, f, f_args), {})
        for bond in bonds:
            J = bond[0]
            indx = tuple(int(i) for i in bond[1:])
            d[indx] = d.get(indx, 0) + J
    out = {}
    for (opstr, f, f_args), d in dd.items():
        out.setdefault((f, f_args), []).extend((opstr, indx, J) for indx, J in d.items() if np.abs(J) > eps)
    return out


class function:
    """Callable drive f(t, *args), hashable by (f, args) (operators/_functions.py)."""

    def __init__(self, func, args):
        self._func = func
        self._args = tuple(args)

    def __call__(self, t):
        return self._func(t, *self._args)

    def __eq__(self, other):
        return isinstance(other, function) and self._func == other._func and self._args == other._args

    def __hash__(self):
        return hash((self._func, self._args))

    def __repr__(self):
        return "function({}, {})".format(getattr(self._func, "__name__", self._func), self._args)

    def conjugate(self):
        """t -> conj(f(t)) (the drive of the conjugated operator, _functions.py conjugate_function)."""
        return _conj_function(self)

    conj = conjugate


class _conj_function(function):
    """Complex conjugate of a drive; its own conjugate is the drive again."""

    def __init__(self, base):
        function.__init__(self, base._func, base._args)
        self._base = base

    def __call__(self, t):
        return np.conj(self._base(t))

    def conjugate(self):
        return self._base

    conj = conjugate

    def __eq__(self, other):
        return isinstance(other, _conj_function) and self._base == other._base

    def __hash__(self):
        return hash(("conj", hash(self._base)))

    def __repr__(self):
        return "conj({!r})".format(self._base)
