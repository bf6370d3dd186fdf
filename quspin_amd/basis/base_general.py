"""General-symmetry basis on the device.

Host mirror of ``basis_general`` (src/quspin/basis/basis_general/
base_general.py:101-1181) and ``hcb_basis_general`` (base_hcb.py:15-95).  The
native ``_core`` object is replaced by ``_Core``, a handle on the C-ABI core
of ``libquspin_hip.so``; the basis (``uintX[Ns]``, descending), the
normalisations ``n`` and the prefix blocks live in HBM as torch buffers.
"""

import warnings

import numpy as np

from .. import _lib as L
from .._lib import ptr, stream_handle


class GeneralBasisWarning(Warning):
    pass


MAXPRINT = 50


def _get_index_type(Ns):
    """int32 unless Ns reaches 2^31-1 (src/quspin/basis/base.py:7-11)."""
    return np.int32 if Ns < np.iinfo(np.int32).max else np.int64


def get_basis_type(N, Np, sps):
    """Smallest unsigned word holding N sites of ``sps`` states (uint32 up to
    32 bits, uint64 up to 64; wider words are not supported here)."""
    bits = int(np.ceil(N * np.log2(sps))) if sps > 1 else N
    if bits <= 32:
        return np.uint32
    if bits <= 64:
        return np.uint64
    raise NotImplementedError("bases wider than 64 bits are not supported")


def _act(labels, m):
    """One application of the signed site map ``m`` to signed site labels
    (label ``-(j+1)`` = site ``j`` inverted): labels sitting at the inverted
    entries of ``m`` are negated, then pulled through the site permutation."""
    inv = m < 0
    src = np.where(inv, -1 - m, m)
    return np.where(inv, -1 - labels, labels)[src]


def process_map(map, q):
    """Validate a site map; return (map, period, q, set of its powers as label
    tuples).  Same period and errors as base_general.py:22-51."""
    m = np.asarray(map, dtype=np.int32)
    src = np.where(m < 0, -1 - m, m)
    ident = np.arange(m.size, dtype=np.int32)
    if not np.array_equal(np.sort(src), ident):
        raise ValueError("map must be a one-to-one site mapping.")
    powers = {tuple(ident)}
    cur, per = _act(ident, m), 1
    while not np.array_equal(cur, ident):
        powers.add(tuple(cur))
        cur, per = _act(cur, m), per + 1
    if per == 1:
        warnings.warn("identity mapping found in set of transformations.", GeneralBasisWarning, stacklevel=5)
    return m, per, q, powers


def _check_symmetry_maps(item1, item2):
    """Warnings of base_general.py:54-98: two blocks generating the same
    cyclic group, or maps that do not commute."""
    (name1, (m1, _, _, grp1)), (name2, (m2, _, _, grp2)) = item1, item2
    if grp1 == grp2:
        warnings.warn("mappings for block {} and block {} produce the same symmetry.".format(name1, name2),
                      GeneralBasisWarning, stacklevel=5)
    ident = np.arange(m1.size)
    if not np.array_equal(_act(_act(ident, m1), m2), _act(_act(ident, m2), m1)):
        warnings.warn("using non-commuting symmetries can lead to unwanted behaviour of general basis, make sure "
                      "that quantum numbers are invariant under non-commuting symmetries!",
                      GeneralBasisWarning, stacklevel=5)


class _Core:
    """Device basis core (replaces ``*_basis_core_wrap``; C-ABI
    ``qsx_core_create``)."""

    def __init__(self, kind, N, maps, pers, qs, double_occupancy=True):
        L.require_device()
        lib = L.load()
        self.kind, self.N = kind, N
        maps = np.ascontiguousarray(maps, dtype=np.int32).reshape(-1, N) if len(pers) else np.zeros((0, N), np.int32)
        self.maps = maps
        self.pers = np.ascontiguousarray(pers, dtype=np.int32)
        self.qs = np.ascontiguousarray(qs, dtype=np.int32)
        h = L._vp()
        rc = lib.qsx_core_create(kind, N, maps.ctypes.data_as(L._vp), self.pers.ctypes.data_as(L._vp),
                                 self.qs.ctypes.data_as(L._vp), len(self.pers), 1 if double_occupancy else 0,
                                 L.C.byref(h))
        L.check(rc, "qsx_core_create")
        self.h = h

    def __del__(self):
        try:
            if getattr(self, "h", None) is not None and L._lib is not None:
                L._lib.qsx_core_destroy(self.h)
        except Exception:
            pass

    @property
    def G(self):
        return L.load().qsx_core_group_order(self.h)


class basis_general:
    """Shared machinery of the general bases (see module docstring)."""

    _fermion = False
    _spinful = False

    def __init__(self, N, block_order=None, **kwargs):
        self._unique_me = True
        self._check_herm = True
        self._check_pcon = None
        self._basis_pcon = None
        self._get_proj_pcon = False
        self._made_basis = False
        self._Ns_block_est = 0
        self._dev = None
        self._host_basis = None
        self._host_n = None
        if self.__class__ is basis_general:
            raise TypeError("general_basis class is not to be instantiated.")
        kwargs = {key: value for key, value in kwargs.items() if value is not None}
        n_maps = len(kwargs)
        if n_maps > 32:
            raise ValueError("general basis can only support up to 32 symmetries.")
        self._conserved = "custom symmetries" if n_maps > 0 else ""
        if any((type(m) is not tuple) and (len(m) != 2) for m in kwargs.values()):
            raise ValueError("blocks must contain tuple: (map,q).")
        kwargs = {block: process_map(*item) for block, item in kwargs.items()}
        if block_order is None:
            sorted_items = sorted(kwargs.items(), key=lambda x: x[1][1])
        else:
            block_order = list(block_order)
            missing = set(kwargs.keys()) - set(block_order)
            if missing:
                raise ValueError("{} names found in block names but missing from block_order.".format(missing))
            missing = set(block_order) - set(kwargs.keys())
            if missing:
                raise ValueError("{} names found in block_order but missing from block names.".format(missing))
            block_order.reverse()
            sorted_items = [(key, kwargs[key]) for key in block_order]
        self._blocks = {block: ((-1) ** q if per == 2 else q) for block, (_, per, q, _) in sorted_items}
        self._maps_dict = {block: m for block, (m, _, _, _) in sorted_items}
        remove_index = []
        for i, item1 in enumerate(sorted_items[:-1]):
            if item1[1][1] == 1:
                remove_index.append(i)
            for item2 in sorted_items[i + 1:]:
                _check_symmetry_maps(item1, item2)
        remove_index.sort()
        if sorted_items:
            _, items = zip(*sorted_items)
            items = list(items)
            for i in remove_index:  # ascending pops, as the reference does
                items.pop(i)
            maps, pers, qs, _ = zip(*items)
            self._maps = np.vstack(maps)
            self._qs = np.asarray(qs, dtype=np.int32)
            self._pers = np.asarray(pers, dtype=np.int32)
            if any(m.shape[0] != N for m in self._maps[:]):
                raise ValueError("size of map is not equal to N.")
        else:
            self._maps = np.array([[]], dtype=np.int32)
            self._qs = np.array([], dtype=np.int32)
            self._pers = np.array([], dtype=np.int32)
        nmax = self._pers.prod()
        self._n_dtype = np.min_scalar_type(nmax)

    # ------------------------------------------------------------------ core
    def _make_core(self, kind, N, double_occupancy=True):
        maps = self._maps if len(self._pers) else np.zeros((0, N), np.int32)
        return _Core(kind, N, maps, self._pers, self._qs, double_occupancy)

    # ------------------------------------------------------------ properties
    @property
    def Ns(self):
        """int: number of states in the (symmetry-reduced) basis."""
        return self._Ns

    @property
    def N(self):
        """int: number of sites (bits) of the lattice."""
        return self._N

    @property
    def sps(self):
        return self._sps

    @property
    def blocks(self):
        return dict(self._blocks)

    @property
    def dtype(self):
        return self._basis_dtype

    @property
    def states(self):
        """numpy array of the representative states, descending."""
        if self._host_basis is None and self._made_basis:
            if self._Ns > 0:
                self._host_basis = L.to_host(self._dev["basis"]).view(self._basis_dtype)
            else:
                self._host_basis = np.array([], dtype=self._basis_dtype)
        return self._host_basis

    @property
    def _basis(self):
        return self.states

    @property
    def _n(self):
        if self._host_n is None and self._made_basis:
            if self._Ns > 0:
                a = L.to_host(self._dev["n"])
                self._host_n = a.view(np.dtype(self._n_dtype)) if a.dtype.itemsize == np.dtype(self._n_dtype).itemsize else a.astype(self._n_dtype)
            else:
                self._host_n = np.array([], dtype=self._n_dtype)
        return self._host_n

    @property
    def description(self):
        blocks = ", ".join("{} = {}".format(k, v) for k, v in self._blocks.items())
        return "general basis for lattice of N = {0} sites containing {1} states \n\tsymmetries: {2}\n".format(
            self._N, self._Ns, blocks or "none")

    def __str__(self):
        s = self.description
        if self._made_basis:
            st = self.states
            n = len(st)
            idx = range(n) if n <= MAXPRINT else list(range(MAXPRINT // 2)) + list(range(n - MAXPRINT // 2, n))
            s += "reference states: \n"
            for i in idx:
                s += "\t{:>8d}.  {}\n".format(i, self.int_to_state(int(st[i])))
        return s

    # ---------------------------------------------------------- make basis
    def _Np_spec(self):
        """Particle sectors as the int32 array the C-ABI expects."""
        Np = self._Np
        if Np is None:
            return np.zeros(0, dtype=np.int32), 0
        if self._spinful:
            arr = np.asarray(Np, dtype=np.int32).reshape(-1, 2)
            return np.ascontiguousarray(arr.ravel()), arr.shape[0]
        arr = np.atleast_1d(np.asarray(Np, dtype=np.int32))
        return np.ascontiguousarray(arr), arr.size

    def make(self, Ns_block_est=None, N_p=None):
        """Build the representative basis on the device (``make``
        base_general.py:817-899): enumerate the particle sector(s), keep orbit
        maxima with positive normalisation, store them descending."""
        L.require_device()
        lib = L.load()
        if Ns_block_est is not None:
            Ns_cap = Ns_block_est if Ns_block_est > self._Ns_block_est else self._Ns_block_est
        else:
            Ns_cap = max([self._Ns, 1000, self._Ns_block_est])
        Np, nNp = self._Np_spec()
        st = stream_handle()
        Ns_c = L.C.c_int64(0)
        L.check(lib.qsx_make_basis(self._core.h, Np.ctypes.data_as(L._vp), nNp, None, None, 1, 0, L.C.byref(Ns_c),
                                   st), "qsx_make_basis")
        Ns = int(Ns_c.value)
        if Ns > Ns_cap:
            raise ValueError(
                "estimate for size of reduced Hilbert-space is too low, please double check that transformation "
                "mappings are correct or use 'Ns_block_est' argument to give an upper bound of the block size.")
        nb = np.dtype(self._n_dtype).itemsize
        nb = 1 if nb <= 1 else (2 if nb <= 2 else 4)
        word = np.dtype(self._basis_dtype)
        basis_d = L.empty(max(Ns, 1), np.int32 if word.itemsize == 4 else np.int64)
        n_d = L.empty(max(Ns, 1), {1: np.uint8, 2: np.int16, 4: np.int32}[nb])
        if Ns > 0:
            rc = lib.qsx_make_basis(self._core.h, Np.ctypes.data_as(L._vp), nNp, ptr(basis_d), ptr(n_d), nb, Ns,
                                    L.C.byref(Ns_c), st)
            L.check(rc, "qsx_make_basis")
        self._dev = {"basis": basis_d[:Ns], "n": n_d[:Ns], "n_bytes": nb}
        self._host_basis = None
        self._host_n = None
        self._Ns = int(Ns)
        self._Ns_block_est = int(Ns)
        self._index_type = _get_index_type(self._Ns)
        self._made_basis = True
        # n dtype shrinks to what the stored values need (base_general.py:287-290)
        if Ns > 0 and len(self._pers):
            nmax = int(self._dev["n"].to(L.torch().int32).max().item())
            self._n_dtype = np.min_scalar_type(nmax)
        self.make_basis_blocks(N_p=N_p)

    def make_basis_blocks(self, N_p=None):
        """Prefix blocks bounding the search for a representative
        (``make_basis_blocks`` base_general.py:901-950)."""
        if not self._made_basis:
            raise ValueError("reference states are not constructed yet. basis must be constructed before "
                             "calculating blocks")
        sps = self.sps if self.sps is not None else 2
        if N_p is None:
            N_p = int(np.floor(np.log(self._Ns // 2 + 1) / np.log(sps)))
        else:
            N_p = int(N_p)
        if len(self._pers) == 0 and self._Np is None:
            N_p = 0
        self._N_p = min(max(N_p, 0), self.N)
        if self._N_p > 0:
            P = 1 << self._N_p
            begin = L.empty(P, np.int64)
            end = L.empty(P, np.int64)
            rc = L.load().qsx_make_basis_blocks(self._core.h, self._N_p, ptr(self._dev["basis"]), self._Ns,
                                                ptr(begin), ptr(end), stream_handle())
            L.check(rc, "qsx_make_basis_blocks")
            self._dev["begin"], self._dev["end"] = begin, end
        else:
            self._dev["begin"], self._dev["end"] = None, None

    @property
    def _basis_begin(self):
        b = self._dev.get("begin")
        return np.array([], np.intp) if b is None else L.to_host(b)

    @property
    def _basis_end(self):
        b = self._dev.get("end")
        return np.array([], np.intp) if b is None else L.to_host(b)

    def lookup_table(self, max_bytes=None, entry_bytes=None):
        """Device table mapping every state of the particle sector (by
        combinadic rank) to (basis index, group element) of its
        representative, built on first use; None when the basis has no
        symmetry, is not a single sector, or the table would exceed
        ``max_bytes`` (default: 40 GiB).  ``entry_bytes=8`` builds 8-byte
        entries where 4 would do (the layout large bases need); the first
        build is kept for the basis."""
        if not self._made_basis or not len(self._pers):
            return None
        if entry_bytes not in (None, 4, 8):
            raise ValueError("entry_bytes must be 4 or 8")
        if "table" in self._dev:
            return self._dev["table"]
        lib = L.load()
        ent, eb, ib, nl = L._i64(), L._i32(), L._i32(), L._i64()
        d = self._dev
        rc = lib.qsx_table_info(self._core.h, self._Ns, ptr(d["n"]), d["n_bytes"], L.C.byref(ent), L.C.byref(eb),
                                L.C.byref(ib), L.C.byref(nl))
        if rc != 0:
            self._dev["table"] = None
            return None
        if entry_bytes == 8:
            eb.value = 8
        elif entry_bytes == 4 and eb.value != 4:
            raise ValueError("this basis needs 8-byte table entries")
        nbytes = ent.value * eb.value
        if nbytes > (max_bytes if max_bytes is not None else 40 * 2 ** 30):
            return None
        tab = L.empty(max(ent.value, 1), np.int32 if eb.value == 4 else np.int64)
        # coefficient table chi(g) * sqrt(n) (complex128 pairs), when the entries carry its ids
        lut = L.empty(nl.value, np.complex128) if nl.value > 0 else None
        rc = lib.qsx_table_build(self._core.h, ptr(d["basis"]), self._Ns, ptr(d.get("begin")), ptr(d.get("end")),
                                 self._N_p, ptr(d["n"]), d["n_bytes"], ptr(tab), eb.value, ib.value,
                                 ptr(lut) if lut is not None else None, nl.value, stream_handle())
        L.check(rc, "qsx_table_build")
        self._dev["table"] = (tab, eb.value, ib.value, lut)
        return self._dev["table"]

    def _dev_args(self):
        """Positional basis arguments of the C-ABI calls."""
        d = self._dev
        return (ptr(d["basis"]), ptr(d["n"]), d["n_bytes"], self._Ns, ptr(d.get("begin")), ptr(d.get("end")),
                self._N_p)

    # ------------------------------------------------------------- queries
    def _word_dev(self, states):
        t = L.require_device()
        a = np.ascontiguousarray(np.atleast_1d(np.asarray(states)).astype(self._basis_dtype))
        return L.to_device(a), a.shape

    def representative(self, states, out=None, return_g=False, return_sign=False):
        """Orbit maxima of ``states`` (base_general.py:1035-1117)."""
        states = np.asarray(states, order="C", dtype=self._basis_dtype)
        states = np.atleast_1d(states)
        if states.ndim != 1:
            raise TypeError("dimension of array_like states must not exceed 1.")
        s_d, _ = self._word_dev(states)
        o_d = L.empty(states.shape[0], np.int32 if np.dtype(self._basis_dtype).itemsize == 4 else np.int64)
        nm = self._qs.shape[0]
        g_d = L.zeros((states.shape[0], max(nm, 1)), np.int32) if return_g else None
        sg_d = L.zeros(states.shape[0], np.int8) if return_sign else None
        rc = L.load().qsx_representative(self._core.h, ptr(s_d), ptr(o_d), ptr(g_d), ptr(sg_d), states.shape[0],
                                         stream_handle())
        L.check(rc, "qsx_representative")
        reps = L.to_host(o_d).view(self._basis_dtype)
        g_out = L.to_host(g_d)[:, :nm] if return_g else None
        sign_out = L.to_host(sg_d) if return_sign else None
        if out is not None:
            if not isinstance(out, np.ndarray):
                raise TypeError("out must be a numpy.ndarray")
            if states.shape != out.shape:
                raise TypeError("states and out must have same shape.")
            if out.dtype != self._basis_dtype:
                raise TypeError("out must have same type as basis")
            out[...] = reps
            res = tuple(x for x in (g_out, sign_out) if x is not None)
            return res if len(res) > 1 else (res[0] if res else None)
        res = [reps] + [x for x in (g_out, sign_out) if x is not None]
        return tuple(res) if len(res) > 1 else reps

    def normalization(self, states, out=None):
        """Character sum over the stabiliser of each state, 0 outside the
        sector (base_general.py:1119-1181)."""
        states = np.atleast_1d(np.asarray(states, order="C", dtype=self._basis_dtype))
        s_d, _ = self._word_dev(states)
        o_d = L.zeros(states.shape[0], np.int64)
        rc = L.load().qsx_normalization(self._core.h, ptr(s_d), ptr(o_d), 8, states.shape[0], stream_handle())
        L.check(rc, "qsx_normalization")
        vals = L.to_host(o_d).astype(np.uint64)
        if out is None:
            return vals.astype(np.min_scalar_type(vals.max(initial=0))).squeeze()
        if states.shape != out.shape:
            raise TypeError("states and out must have same shape.")
        if np.issubdtype(out.dtype, np.signedinteger):
            raise TypeError("out must have datatype numpy.uint8, numpy.uint16, numpy.uint32, or numpy.uint64.")
        out[...] = vals.astype(out.dtype)

    def index(self, *args):
        """Position of a state (int or string) in the basis."""
        if len(args) == 1:
            s = args[0]
        else:
            s = "".join(str(a) for a in args)
        if isinstance(s, str):
            s = int(s.replace("|", "").replace(">", "").replace("<", "").replace(" ", ""), self.sps)
        st = self.states[::-1]
        pos = np.searchsorted(st, np.asarray(s, dtype=self._basis_dtype))
        if pos >= st.size or st[pos] != s:
            raise ValueError("state {} not found in basis".format(s))
        return self._Ns - 1 - int(pos)

    def int_to_state(self, state, bracket_notation=True):
        state = int(state)
        bits = [(state >> (self.N - 1 - i)) & 1 for i in range(self.N)]
        s_str = " ".join(str(b) for b in bits)
        return "|" + s_str + ">" if bracket_notation else s_str.replace(" ", "")

    def state_to_int(self, state):
        state = state.replace("|", "").replace(">", "").replace("<", "").replace(" ", "")
        return int(state, 2)

    # ---------------------------------------------------------- operators
    def _opstr_prepare(self, opstr, indx, J):
        """Subclasses map (opstr, indx, J) to the core's bit alphabet
        (Pauli scaling, spinful '|' notation)."""
        return opstr, indx, J

    def _validate_op(self, opstr, indx):
        indx = np.asarray(indx, dtype=np.int32)
        if len(opstr) != len(indx):
            raise ValueError("length of opstr does not match length of indx")
        if np.any(indx >= self._N) or np.any(indx < 0):
            raise ValueError("values in indx falls outside of system")
        extra_ops = set(opstr) - self._allowed_ops
        if extra_ops:
            raise ValueError("unrecognized characters {} in operator string.".format(extra_ops))

    def Op(self, opstr, indx, J, dtype):
        """(ME, row, col) of one operator string in the reduced basis
        (``_Op`` base_general.py:292-352): one element per basis state,
        vanishing elements removed."""
        if not self._made_basis:
            raise AttributeError("this function requires the basis to be constructed first; use basis.make().")
        opstr, indx, J = self._opstr_prepare(opstr, indx, J)
        self._validate_op(opstr, indx)
        dtype = np.dtype(dtype)
        if self._Ns <= 0:
            return (np.array([], dtype=dtype), np.array([], dtype=self._index_type),
                    np.array([], dtype=self._index_type))
        idx_bytes = np.dtype(self._index_type).itemsize
        row = L.empty(self._Ns, self._index_type)
        col = L.empty(self._Ns, self._index_type)
        me = L.empty(self._Ns, np.complex128)
        indx = np.ascontiguousarray(indx, dtype=np.int32)
        Jc = complex(J)
        rc = L.load().qsx_op(self._core.h, ptr(row), ptr(col), idx_bytes, ptr(me), opstr.encode(),
                             indx.ctypes.data_as(L._vp), len(opstr), Jc.real, Jc.imag, *self._dev_args(),
                             stream_handle())
        L.check(rc, "qsx_op")
        ME = L.to_host(me)
        if dtype.kind != "c":
            if np.any(ME.imag != 0):
                raise TypeError("attempting to use real dtype for complex matrix elements")
            ME = ME.real
        ME = ME.astype(dtype)
        mask = ME != 0
        return ME[mask], L.to_host(row)[mask], L.to_host(col)[mask]

    def _compile_oplist(self, op_list):
        """Device operator list (``qsx_oplist_create``) of prepared terms."""
        opstrs, lens, idx, Js = [], [], [], []
        for opstr, indx, J in op_list:
            opstr, indx, J = self._opstr_prepare(opstr, indx, J)
            self._validate_op(opstr, indx)
            opstrs.append(opstr)
            lens.append(len(opstr))
            idx.extend(int(i) for i in indx)
            Jc = complex(J)
            Js.extend([Jc.real, Jc.imag])
        from ._checks import is_symmetric
        sym = is_symmetric(self, op_list)
        return _OpList(self._core, "".join(opstrs), np.asarray(lens, np.int32), np.asarray(idx, np.int32),
                       np.asarray(Js, np.float64), sym)

    def inplace_Op(self, v_in, op_list, dtype, transposed=False, conjugated=False, a=1.0, v_out=None):
        """v_out += a * Op v_in for the operator list (``inplace_Op``
        src/quspin/basis/base.py:164-223 -> base_general.py:354-410), all
        terms fused in one device launch."""
        if not self._made_basis:
            raise AttributeError("this function requires the basis to be constructed first; use basis.make().")
        on_dev = L.is_device_tensor(v_in)
        if on_dev:
            vin_np_dtype = L.np_dtype(v_in.dtype)
        else:
            v_in = np.asanyarray(v_in)
            vin_np_dtype = v_in.dtype
        result_dtype = np.result_type(vin_np_dtype, dtype)
        shape = tuple(v_in.shape)
        if shape[0] != self.Ns:
            raise ValueError("dimension mismatch")
        if v_out is not None:
            out_dt = L.np_dtype(v_out.dtype) if L.is_device_tensor(v_out) else v_out.dtype
            if out_dt != result_dtype:
                raise TypeError("v_out does not have the correct data type.")
            if tuple(v_out.shape) != shape:
                raise ValueError("invalid shape for v_out and v_in: v_in.shape != v_out.shape")
            if not L.is_device_tensor(v_out) and not v_out.flags["CARRAY"]:
                raise ValueError("v_out is not a writable C-contiguous array")
        vin_d = L.to_device(v_in, result_dtype)
        if v_out is None:
            vout_d = L.zeros(shape, result_dtype)
        elif L.is_device_tensor(v_out):
            vout_d = v_out
        else:
            vout_d = L.to_device(v_out, result_dtype)
        nvec = int(np.prod(shape[1:])) if len(shape) > 1 else 1
        ol = self._compile_oplist(op_list)
        ac = complex(a)
        rc = L.load().qsx_hv(self._core.h, ol.h, *self._dev_args(), ptr(vin_d), ptr(vout_d),
                             L.dtype_code(result_dtype), nvec, int(bool(transposed)), int(bool(conjugated)),
                             int(bool(conjugated)), ac.real, ac.imag, ac.real, ac.imag, 1, None, 0,
                             stream_handle())
        L.check(rc, "qsx_hv")
        if on_dev or L.is_device_tensor(v_out):
            return vout_d.squeeze() if v_out is None else v_out
        res = L.to_host(vout_d)
        if v_out is not None:
            v_out[...] = res
            return v_out.squeeze()
        return res.squeeze()

    # ------------------------------------------ sectors and projections
    def _require_made(self):
        if not self._made_basis:
            raise AttributeError("this function requires the basis to be constructed first; use basis.make().")

    def Op_shift_sector(self, other_basis, op_list, v_in, v_out=None, dtype=None):
        """Apply an operator that moves states of ``other_basis``' sector
        into this basis' sector: v_out = P_self^H Op P_other v_in
        (base_general.py:412-519), one device pass per term as the reference
        calls its core per term.  Both bases must carry the same symmetry
        maps.  ``v_in``: (other_basis.Ns,) or (other_basis.Ns, nvec), host or
        device; the result lives where ``v_in`` lives."""
        if not isinstance(other_basis, self.__class__):
            raise ValueError("other_basis must be the same type as the given basis.")
        self._require_made()
        other_basis._require_made()
        _, _, J_list = zip(*op_list)
        J_list = np.asarray(J_list)
        if dtype is not None:
            J_list = J_list.astype(dtype)
        on_dev = L.is_device_tensor(v_in)
        vdt = L.np_dtype(v_in.dtype) if on_dev else np.asanyarray(v_in).dtype
        result_dtype = np.result_type(np.float32, J_list.dtype, vdt)
        shape = tuple(v_in.shape)
        nvecs = int(np.prod(shape[1:])) if len(shape) > 1 else 1
        if shape[0] != other_basis.Ns:
            raise ValueError("invalid shape for v_in")
        if v_out is not None:
            odt = L.np_dtype(v_out.dtype) if L.is_device_tensor(v_out) else v_out.dtype
            if odt != result_dtype:
                raise TypeError("v_out does not have the correct data type.")
            if not L.is_device_tensor(v_out) and not v_out.flags["CARRAY"]:
                raise ValueError("v_out is not a writable C-contiguous array")
            if tuple(v_out.shape) not in ((self.Ns, nvecs), (self.Ns,) if nvecs == 1 else (self.Ns, nvecs)):
                raise ValueError("invalid shape for v_out")
        # the kernels accumulate in float64 / complex128
        kdt = np.dtype(np.complex128) if result_dtype.kind == "c" else np.dtype(np.float64)
        vin_d = L.to_device(v_in, kdt).reshape(other_basis.Ns, nvecs).contiguous()
        if v_out is None:
            vout_d = L.zeros((self.Ns, nvecs), kdt)
        else:
            vout_d = L.to_device(v_out, kdt).reshape(self.Ns, nvecs).contiguous()
        lib = L.load()
        st = stream_handle()
        for (opstr, indx, _), J in zip(op_list, J_list):
            opstr, indx, J = self._opstr_prepare(opstr, indx, J)
            self._validate_op(opstr, indx)
            idx = np.ascontiguousarray(indx, dtype=np.int32)
            Jc = complex(J)
            od = other_basis._dev
            rc = lib.qsx_op_shift_sector(self._core.h, other_basis._core.h, ptr(vin_d), ptr(vout_d), L.dtype_code(kdt),
                                         nvecs, opstr.encode(), idx.ctypes.data_as(L._vp), len(opstr), Jc.real, Jc.imag,
                                         *self._dev_args(), ptr(od["basis"]), ptr(od["n"]), od["n_bytes"],
                                         other_basis.Ns, st)
            L.check(rc, "qsx_op_shift_sector")
        res = vout_d.to(L.torch_dtype(result_dtype))
        if v_out is not None:
            if L.is_device_tensor(v_out):
                v_out.copy_(res.reshape(v_out.shape))
                return v_out
            v_out[...] = L.to_host(res).reshape(v_out.shape)
            return v_out.squeeze() if nvecs == 1 else v_out
        if on_dev:
            return res.squeeze(1) if nvecs == 1 else res
        out = L.to_host(res)
        return out.squeeze() if nvecs == 1 else out

    def _full_space(self, pcon):
        """(pcon flag, size of the space the projector maps into): the 2^N
        states, or the basis' particle sector (get_proj, base_general.py:553-566)."""
        if pcon:
            if not self._get_proj_pcon:
                raise TypeError("pcon=True only works for basis of a single particle number sector.")
            info = np.zeros(8, np.int64)
            L.check(L.load().qsx_sector_rank(self._core.h, None, 0, None, info.ctypes.data_as(L._vp), None),
                    "qsx_sector_rank")
            if not info[5] or info[0] > 1:
                raise TypeError("pcon=True only works for basis of a single particle number sector.")
            return 1, int(info[4])
        return 0, int(self._sps) ** int(self._N)

    def get_proj(self, dtype, pcon=False):
        """Projector P (full x reduced, scipy CSC) from the reduced basis to
        the full space or its particle sector (get_proj,
        base_general.py:521-582): column r holds c_r chi(g) sign(g, r) on the
        rows of the orbit images g.r, c_r = 1/sqrt(n_r |G|), built on the
        device."""
        import scipy.sparse as sp
        self._require_made()
        pc, Ns_full = self._full_space(pcon)
        dtype = np.dtype(dtype)
        index_type = np.result_type(np.min_scalar_type(-Ns_full), np.int32)
        idx_bytes = np.dtype(index_type).itemsize
        if self._Ns == 0:
            return sp.csc_matrix((Ns_full, 0), dtype=dtype)
        lib = L.load()
        st = stream_handle()
        nnz = L.C.c_int64(0)
        d = self._dev
        args = (self._core.h, ptr(d["basis"]), ptr(d["n"]), d["n_bytes"], self._Ns, pc, Ns_full,
                L.dtype_code(dtype), idx_bytes)
        L.check(lib.qsx_get_proj(*args, None, None, None, L.C.byref(nnz), st), "qsx_get_proj")
        indptr = L.empty(self._Ns + 1, index_type)
        indices = L.empty(max(nnz.value, 1), index_type)
        data = L.empty(max(nnz.value, 1), dtype)
        L.check(lib.qsx_get_proj(*args, ptr(indptr), ptr(indices), ptr(data), L.C.byref(nnz), st), "qsx_get_proj")
        n = nnz.value
        return sp.csc_matrix((L.to_host(data)[:n], L.to_host(indices)[:n], L.to_host(indptr)),
                             shape=(Ns_full, self._Ns))

    def _project(self, v0, sparse, pcon, to_reduced):
        import scipy.sparse as sp
        self._require_made()
        pc, Ns_full = self._full_space(pcon)
        on_dev = L.is_device_tensor(v0)
        if not on_dev and not hasattr(v0, "shape"):
            v0 = np.asanyarray(v0)
        n_in, n_out = (Ns_full, self._Ns) if to_reduced else (self._Ns, Ns_full)
        squeeze = v0.ndim == 1
        if v0.ndim not in (1, 2):
            raise ValueError("excpecting v0 to have ndim > 0 and at most 2")
        nv = 1 if squeeze else int(v0.shape[1])
        vdt = L.np_dtype(v0.dtype) if on_dev else v0.dtype
        if self._Ns <= 0:
            if sparse:
                return sp.csr_matrix(([], ([], [])), shape=(n_out, 0), dtype=vdt)
            return np.zeros((n_out, 0), dtype=vdt)
        if v0.shape[0] != n_in:
            raise ValueError("v0 shape {0} not compatible with Ns_full={1}".format(tuple(v0.shape), Ns_full)
                             if to_reduced else
                             "v0 shape {0} not compatible with Ns={1}".format(tuple(v0.shape), self._Ns))
        if sp.issparse(v0):
            raise ValueError
        if sparse and not on_dev:
            P = self.get_proj(vdt, pcon=pcon)
            # the reference's sparse forms: P^T (no conjugate) and P
            # (base_general.py:671, 781)
            return P.T.dot(sp.csr_matrix(v0.reshape(n_in, nv))) if to_reduced else P.dot(sp.csc_matrix(v0.reshape(n_in, nv)))
        kdt = np.dtype(vdt)
        if kdt not in (np.float32, np.float64, np.complex64, np.complex128):
            kdt = np.result_type(kdt, np.float64)
        x = L.to_device(v0, kdt).reshape(n_in, nv).contiguous()
        y = L.zeros((n_out, nv), kdt)
        fn = L.load().qsx_project_to_dense if to_reduced else L.load().qsx_project_from_dense
        d = self._dev
        L.check(fn(self._core.h, ptr(d["basis"]), ptr(d["n"]), d["n_bytes"], self._Ns, ptr(x), ptr(y),
                   L.dtype_code(kdt), nv, pc, Ns_full, stream_handle()),
                "qsx_project_to_dense" if to_reduced else "qsx_project_from_dense")
        if on_dev:
            return y.squeeze(1) if squeeze else y
        out = L.to_host(y)
        return out[:, 0] if squeeze else out

    def project_to(self, v0, sparse=True, pcon=False):
        """Full-space (or particle-sector) states to the reduced basis:
        P^H v0 for the dense form (project_to_dense, base_general.py:584-694;
        test/test_project_to.py:24-29), the reference's P^T v0 for the
        sparse form."""
        return self._project(v0, sparse, pcon, True)

    def project_from(self, v0, sparse=True, pcon=False):
        """Reduced-basis states to the full space (or particle sector):
        P v0 (project_from_dense, base_general.py:696-794)."""
        return self._project(v0, sparse, pcon, False)

    def get_vec(self, v0, sparse=True, pcon=False):
        """Deprecated name of ``project_from`` (base_general.py:684-694)."""
        return self.project_from(v0, sparse=sparse, pcon=pcon)

    def get_amp(self, states, out=None, amps=None, mode="representative"):
        """Rescale factor C = psi_s / psi_r^sym between full-basis and
        reduced amplitudes (get_amp, base_general.py:1183-1269), computed on
        the device: C = sign conj(chi(g)) sqrt(n_r / |G|) for g.s = r.
        ``mode="full_basis"`` replaces ``states`` in place by their
        representatives."""
        states_a = np.atleast_1d(np.asarray(states, order="C", dtype=self._basis_dtype))
        states_shape = states_a.shape
        if out is not None:
            if states_shape != out.shape:
                raise TypeError("states and out must have same shape.")
            if out.dtype not in [np.float32, np.float64, np.complex64, np.complex128]:
                raise TypeError("out must have datatype numpy.float32, numpy.float64, numpy.complex64, or "
                                "numpy.complex128.")
            if not out.flags["CARRAY"]:
                raise ValueError("out must be C-contiguous array.")
        elif amps is not None:
            out = np.zeros(states_shape, dtype=amps.dtype)
        else:
            out = np.zeros(states_shape, dtype=np.complex128)
        if mode not in ("representative", "full_basis"):
            raise ValueError("mode accepts only the values 'representative' and 'full_basis'.")
        s_d, _ = self._word_dev(states_a.ravel())
        o_d = L.empty(states_a.size, out.dtype)
        rc = L.load().qsx_get_amp(self._core.h, ptr(s_d), ptr(o_d), L.dtype_code(out.dtype), states_a.size,
                                  1 if mode == "full_basis" else 0, stream_handle())
        L.check(rc, "qsx_get_amp")
        out[...] = L.to_host(o_d).reshape(states_shape)
        if mode == "full_basis" and isinstance(states, np.ndarray) and states.dtype == self._basis_dtype:
            states[...] = L.to_host(s_d).view(self._basis_dtype).reshape(states.shape)
        if amps is not None:
            if states_shape != amps.shape:
                raise TypeError("states and amps must have same shape.")
            if mode == "representative":
                amps *= out
            else:
                amps /= out
        return out.squeeze()

    # --------------------------------------------- hermiticity/symmetry checks
    def _expand_term(self, opstr, indx, J):
        return [(opstr, tuple(indx), J)]

    def check_hermitian(self, static, dynamic):
        from ._checks import check_hermitian
        check_hermitian(self, static, dynamic)

    def check_symm(self, static, dynamic):
        from ._checks import check_symm
        check_symm(self, static, dynamic)

    def check_pcon(self, static, dynamic):
        from ._checks import check_pcon
        check_pcon(self, static, dynamic)


class _OpList:
    """Device operator list handle."""

    def __init__(self, core, opstrs, lens, idx, Js, symmetric=True):
        self.core = core
        self.symmetric = bool(symmetric)
        h = L._vp()
        rc = L.load().qsx_oplist_create(core.h, len(lens), opstrs.encode(), lens.ctypes.data_as(L._vp),
                                        idx.ctypes.data_as(L._vp), Js.ctypes.data_as(L._vp), int(symmetric),
                                        L.C.byref(h))
        L.check(rc, "qsx_oplist_create")
        self.h = h
        no, nd, cx = L._i32(), L._i32(), L._i32()
        L.load().qsx_oplist_info(h, L.C.byref(no), L.C.byref(nd), L.C.byref(cx))
        self.n_off, self.n_diag, self.complex_me = no.value, nd.value, cx.value

    def __del__(self):
        try:
            if getattr(self, "h", None) is not None and L._lib is not None:
                L._lib.qsx_oplist_destroy(self.h)
        except Exception:
            pass


class hcb_basis_general(basis_general):
    """Hard-core bosons / spin-1/2 (base_hcb.py:15-95)."""

    def __init__(self, N, Nb=None, Ns_block_est=None, _Np=None, _make_basis=True, block_order=None, **kwargs):
        from scipy.special import comb
        basis_general.__init__(self, N, block_order=block_order, **kwargs)
        self._check_pcon = False
        self._count_particles = False
        if Nb is None:
            self._Ns = 1 << N
        elif type(Nb) is int:
            self._check_pcon = True
            self._get_proj_pcon = True
            self._Ns = comb(N, Nb, exact=True)
        else:
            try:
                iter(Nb)
            except TypeError:
                raise TypeError("Nb must be integer or iteratable object.")
            Nb = list(Nb)
            self._Ns = 0
            for b in Nb:
                if b > N or b < 0:
                    raise ValueError("particle number Nb must satisfy: 0 <= Nb <= N")
                self._Ns += comb(N, b, exact=True)
        if len(self._pers) > 0:
            if Ns_block_est is None:
                self._Ns = int(float(self._Ns) / np.multiply.reduce(self._pers)) * 2
            else:
                if type(Ns_block_est) is not int:
                    raise TypeError("Ns_block_est must be integer value.")
                self._Ns = Ns_block_est
        self._basis_dtype = get_basis_type(N, None, 2)
        self._N = N
        self._Ns_block_est = self._Ns
        self._Np = Nb
        self._sps = 2
        self._allowed_ops = set(["I", "x", "y", "z", "+", "-", "n"])
        self._core = self._make_core(L.KIND_HCB, N)
        if _make_basis:
            self.make()
        else:
            self._Ns = 1
