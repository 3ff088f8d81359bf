"""ctypes binding of oracle/_ref/libsadmrg_ref.so -- TEST INFRASTRUCTURE ONLY.

The shared object is the reference library itself, compiled from its own
sources by oracle/Makefile together with oracle/ref_harness.cpp.  Tests use
it to generate and cross-check golden vectors; bench.py uses it as the timed
CPU baseline.  The product never loads it.

Tensor exchange format (identical to the product's C ABI):
  header  = [nbond, locspin, {dir, nmom, q..., dim...} per bond]
  present = uint8 per block (column-major block index over bonds)
  data    = present blocks concatenated in block order, column-major;
            complex data interleaved (re, im).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_ref", "libsadmrg_ref.so")

_lib = None


def available() -> bool:
    return os.path.exists(LIB_PATH)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not available():
        raise FileNotFoundError(LIB_PATH + " (build it with `make -C oracle`)")
    L = C.CDLL(LIB_PATH)
    vp, ip, dp, bp = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_ubyte)
    sig = {
        "ref_last_error": (C.c_char_p, []),
        "ref_quiet": (None, [C.c_int]),
        "ref_init": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int]),
        "ref_set_max_dcut": (None, [C.c_int]),
        "ref_set_psize": (None, [C.c_int]),
        "ref_set_blas_threading": (None, [C.c_int, C.c_int]),
        "ref_cg": (C.c_double, [C.c_int] * 6),
        "ref_fac_hamilt_vec": (C.c_int, [dp]),
        "ref_fac_table5": (C.c_int, [C.c_int, dp]),
        "ref_fac_transformation": (C.c_int, [C.c_int, dp]),
        "ref_tensor_new": (vp, [C.c_int, C.c_int, ip, bp, dp, C.c_int]),
        "ref_tensor_free": (None, [vp]),
        "ref_tensor_header": (C.c_int, [vp, ip, C.c_int]),
        "ref_tensor_nten": (C.c_int, [vp]),
        "ref_tensor_get": (C.c_longlong, [vp, bp, dp, ip]),
        "ref_tensor_copy": (vp, [vp]),
        "ref_hvm": (vp, [vp, vp, vp]),
        "ref_contract": (vp, [vp, C.c_int, vp, C.c_int]),
        "ref_env": (vp, [C.c_int, vp, vp, vp, vp, C.c_int]),
        "ref_svd": (C.c_int, [vp, vp, C.c_double, vp, C.c_double, dp]),
        "ref_fuse": (vp, [C.c_int, C.c_int, ip, ip]),
        "ref_tensor_op": (C.c_int, [vp, C.c_int, C.c_double, C.c_double]),
        "ref_tensor_binop": (C.c_int, [vp, C.c_int, vp]),
        "ref_inner_prod_c": (C.c_int, [vp, vp, dp]),
        "ref_dmrg_new": (vp, [C.c_int] * 7),
        "ref_dmrg_new_exci": (vp, [C.c_int] * 8),
        "ref_dmrg_state_energy": (C.c_double, [vp, C.c_int]),
        "ref_dmrg_free": (None, [vp]),
        "ref_dmrg_ns": (C.c_int, [vp]),
        "ref_dmrg_set_couplings": (C.c_int, [vp, ip, dp, C.c_int]),
        "ref_dmrg_get_couplings": (C.c_int, [vp, ip, dp]),
        "ref_dmrg_idmrg": (C.c_int, [vp]),
        "ref_dmrg_sweep": (C.c_int, [vp]),
        "ref_dmrg_energy": (C.c_double, [vp]),
        "ref_dmrg_tensor": (vp, [vp, C.c_int, C.c_int, C.c_int]),
        "ref_dmrg_set_tensor": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, vp]),
        "ref_dmrg_call": (C.c_int, [vp, C.c_int, C.c_int, C.c_int]),
        "ref_dmrg_input_vector": (vp, [vp, C.c_int, C.c_int]),
        "ref_dmrg_heff": (vp, [vp, C.c_int, C.c_int, vp]),
        "ref_dmrg_lanczos": (C.c_double, [vp, C.c_int, C.c_int, vp]),
        "ref_dmrg_svd": (C.c_int, [vp, vp, C.c_int, C.c_int, C.c_double, C.c_double, dp, C.c_int]),
        "ref_dmrg_time_heff": (C.c_double, [vp, C.c_int, C.c_int, vp, C.c_int]),
        "ref_dmrg_save": (C.c_int, [vp, C.c_char_p, C.c_int]),
        "ref_dmrg_read": (C.c_int, [vp, C.c_char_p, C.c_int, C.c_int]),
        "ref_dmrg_ww": (C.c_int, [vp, dp, C.c_int]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _iptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _bptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_ubyte))


@dataclass
class Bond:
    """One tensor leg: direction, sorted sector labels (2J or charge) and dims."""
    dir: int
    q: list
    dim: list

    def header(self):
        return [self.dir, len(self.q)] + list(self.q) + list(self.dim)


def parse_header(h):
    h = list(h)
    nb, ls = h[0], h[1]
    bonds, p = [], 2
    for _ in range(nb):
        d, nm = h[p], h[p + 1]
        bonds.append(Bond(d, h[p + 2:p + 2 + nm], h[p + 2 + nm:p + 2 + 2 * nm]))
        p += 2 + 2 * nm
    return nb, ls, bonds


def make_header(locspin, bonds):
    h = [len(bonds), locspin]
    for b in bonds:
        h += b.header()
    return np.asarray(h, dtype=np.int32)


@dataclass
class Packed:
    """A block tensor in exchange format (host numpy arrays)."""
    sym: int
    cplx: int
    header: np.ndarray
    present: np.ndarray
    data: np.ndarray            # float64 (interleaved if complex)
    cgc_kind: np.ndarray = field(default=None)

    @property
    def bonds(self):
        return parse_header(self.header)[2]

    @property
    def locspin(self):
        return int(self.header[1])


class RefTensor:
    """Owning handle to a reference BlockTensor<Scalar,Sym>."""

    def __init__(self, ptr, sym, cplx):
        self.ptr = C.c_void_p(ptr)
        self.sym = sym
        self.cplx = cplx

    def __del__(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.ref_tensor_free(self.ptr)
            self.ptr = None

    @classmethod
    def from_packed(cls, p: Packed, cgc2_kind=0):
        L = lib()
        hdr = np.ascontiguousarray(p.header, dtype=np.int32)
        pres = np.ascontiguousarray(p.present, dtype=np.uint8)
        data = np.ascontiguousarray(p.data, dtype=np.float64)
        if data.size == 0:
            data = np.zeros(1)
        ptr = L.ref_tensor_new(p.sym, p.cplx, _iptr(hdr), _bptr(pres), _dptr(data), cgc2_kind)
        return cls(ptr, p.sym, p.cplx)

    def packed(self) -> Packed:
        L = lib()
        n = L.ref_tensor_header(self.ptr, None, 0)
        hdr = np.zeros(max(n, 1), dtype=np.int32)
        L.ref_tensor_header(self.ptr, _iptr(hdr), n)
        hdr = hdr[:n]
        nten = L.ref_tensor_nten(self.ptr)
        pres = np.zeros(max(nten, 1), dtype=np.uint8)
        kind = np.zeros(max(nten, 1), dtype=np.int32)
        nd = L.ref_tensor_get(self.ptr, _bptr(pres), None, _iptr(kind))
        if nd < 0:
            raise RuntimeError("reference tensor block dims inconsistent")
        w = 2 if self.cplx else 1
        data = np.zeros(max(nd * w, 1))
        L.ref_tensor_get(self.ptr, _bptr(pres), _dptr(data), _iptr(kind))
        return Packed(self.sym, self.cplx, hdr, pres[:nten].copy(), data[:nd * w].copy(), kind[:nten].copy())

    def copy(self):
        return RefTensor(lib().ref_tensor_copy(self.ptr), self.sym, self.cplx)


def init(physpn=1, max_angm=20, max_dcut=64, psize=1, quiet=True):
    L = lib()
    if os.environ.get("SADMRG_REF_VERBOSE"):
        quiet = False
    L.ref_quiet(1 if quiet else 0)
    L.ref_init(physpn, max_angm, max_dcut, psize)


def hvm(vec, op1, op2):
    return RefTensor(lib().ref_hvm(vec.ptr, op1.ptr, op2.ptr), vec.sym, vec.cplx)


def contract(t1, i1, t2, i2):
    return RefTensor(lib().ref_contract(t1.ptr, i1, t2.ptr, i2), t1.sym, t1.cplx)


ENV_KIND = {"overlap_initial": 0, "overlap_transformation": 1, "operator_initial": 2,
            "operator_transformation": 3, "operator_pairup": 4}


def env(kind, uu, vv, op1=None, op2=None, flag=0):
    k = ENV_KIND[kind]
    p1 = op1.ptr if op1 is not None else None
    p2 = op2.ptr if op2 is not None else None
    return RefTensor(lib().ref_env(k, uu.ptr, vv.ptr, p1, p2, flag), uu.sym, uu.cplx)


class RefDMRG:
    """Handle to the reference DMRG<Scalar,Sym> (dmrg.hpp:29-133)."""

    def __init__(self, sym, cplx, xx, yy, bondd, physpn=1, totspin=0, exci=0):
        self.sym, self.cplx = sym, cplx
        self.ptr = C.c_void_p(lib().ref_dmrg_new_exci(sym, cplx, xx, yy, bondd, physpn, totspin, exci))
        self.ns = lib().ref_dmrg_ns(self.ptr)

    def __del__(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.ref_dmrg_free(self.ptr)
            self.ptr = None

    def set_couplings(self, hmap, coup):
        h = np.ascontiguousarray(hmap, dtype=np.int32).reshape(-1)
        c = np.ascontiguousarray(coup, dtype=np.float64)
        rc = lib().ref_dmrg_set_couplings(self.ptr, _iptr(h), _dptr(c), len(c))
        if rc != 0:
            raise ValueError("too many couplings")

    def couplings(self):
        h = np.zeros(self.ns * self.ns, dtype=np.int32)
        c = np.zeros(self.ns + 1)
        lib().ref_dmrg_get_couplings(self.ptr, _iptr(h), _dptr(c))
        return h.reshape(self.ns, self.ns), c

    def idmrg(self):
        lib().ref_dmrg_idmrg(self.ptr)

    def sweep(self):
        lib().ref_dmrg_sweep(self.ptr)

    def energy(self):
        return lib().ref_dmrg_energy(self.ptr)

    SLOTS = {"uu": 0, "hh": 1, "opr": 2, "orth": 3, "ovlp": 4, "overlapvec": 5, "sigma": 6}

    def tensor(self, which, i, j=0):
        """uu[i], hh[i], opr[i][j]; orth[j][i], ovlp[j][i], overlapvec[j]; sigma[i]."""
        code = self.SLOTS[which]
        return RefTensor(lib().ref_dmrg_tensor(self.ptr, code, i, j), self.sym, self.cplx)

    def state_energy(self, j):
        return lib().ref_dmrg_state_energy(self.ptr, j)

    def set_tensor(self, which, i, j, t):
        code = self.SLOTS[which]
        lib().ref_dmrg_set_tensor(self.ptr, code, i, j, t.ptr)

    def call(self, name, a=0, b=0):
        code = {"wavefunc_transformation": 0, "left2": 1, "right2": 2, "left1": 3,
                "right1": 4, "prepare_sweep": 5}[name]
        lib().ref_dmrg_call(self.ptr, code, a, b)

    def input_vector(self, il, ir):
        return RefTensor(lib().ref_dmrg_input_vector(self.ptr, il, ir), self.sym, self.cplx)

    def heff(self, il, ir, vin):
        return RefTensor(lib().ref_dmrg_heff(self.ptr, il, ir, vin.ptr), self.sym, self.cplx)

    def lanczos(self, il, ir, vec):
        return lib().ref_dmrg_lanczos(self.ptr, il, ir, vec.ptr)

    def svd(self, vec, a, b, p1, p2, cap=4096):
        w = np.zeros(cap)
        lib().ref_dmrg_svd(self.ptr, vec.ptr, a, b, p1, p2, _dptr(w), cap)
        return w

    def time_heff(self, il, ir, vin, reps):
        return lib().ref_dmrg_time_heff(self.ptr, il, ir, vin.ptr, reps)

    def save(self, directory, exci=0):
        """save_mps2(exci); save_ww(); save_enr() into `directory`."""
        if lib().ref_dmrg_save(self.ptr, str(directory).encode(), exci) != 0:
            raise OSError(f"cannot enter {directory}")

    def read(self, directory, read1, read2=0):
        """read_mps(read1, read2) from `directory`; returns the pass flag."""
        rc = lib().ref_dmrg_read(self.ptr, str(directory).encode(), read1, read2)
        if rc < 0:
            raise OSError(f"cannot enter {directory}")
        return rc == 1

    def ww(self, cap):
        w = np.zeros(cap)
        lib().ref_dmrg_ww(self.ptr, _dptr(w), cap)
        return w
