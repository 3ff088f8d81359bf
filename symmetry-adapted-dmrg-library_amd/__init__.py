"""ctypes binding of libsadmrg.so (include/sadmrg.h).

Host-side mirror of the reference's BlockTensor / Lanczos / DMRG interface
for tests and benchmarks.  All arithmetic runs in the HIP library; this module
only moves tensors across the C ABI.  If the shared library is missing, the
import of any entry point raises: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsadmrg.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include", "sadmrg.h")

U1, SU2 = 0, 1
FP64, FP32, BF16 = 0, 1, 2

_lib = None


class LanczosOptions(C.Structure):
    """sadmrg_lanczos_options (include/sadmrg.h)."""
    _fields_ = [("mlanc", C.c_int), ("precision", C.c_int), ("refine", C.c_int), ("refine_tol", C.c_double),
                ("refine_krylov", C.c_int), ("max_restarts", C.c_int), ("check_every", C.c_int),
                ("device_steps", C.c_int), ("graph", C.c_int)]


class SadmrgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"sadmrg error {code}: {msg}")
        self.code = code


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(f"{LIB_PATH} is not built (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    vp, ip, dp, bp = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_ubyte)
    pp = C.POINTER(C.c_void_p)
    llp = C.POINTER(C.c_longlong)
    i, d = C.c_int, C.c_double
    sig = {
        "sadmrg_last_error": (C.c_char_p, []),
        "sadmrg_version": (i, [ip, ip]),
        "sadmrg_ctx_create": (i, [i, pp]),
        "sadmrg_ctx_destroy": (i, [vp]),
        "sadmrg_ctx_sync": (i, [vp]),
        "sadmrg_ctx_memory": (i, [vp, llp, llp]),
        "sadmrg_ctx_stats": (i, [vp, C.POINTER(C.c_longlong), dp, C.POINTER(C.c_longlong), dp]),
        "sadmrg_ctx_stream": (i, [vp, pp]),
        "sadmrg_clebsch_gordan": (i, [i, i, i, i, i, i, dp]),
        "sadmrg_cgc_create": (i, [i, pp]),
        "sadmrg_cgc_destroy": (i, [vp]),
        "sadmrg_cgc_hamilt_vec": (i, [vp, i, i, i, dp]),
        "sadmrg_cgc_onsite": (i, [vp, i, i, i, i, i, i, dp]),
        "sadmrg_cgc_transformation": (i, [vp, i, i, i, i, i, i, i, dp]),
        "sadmrg_cgc_pairup": (i, [vp, i, i, i, i, i, i, dp]),
        "sadmrg_tensor_create": (i, [vp, i, i, ip, i, bp, dp, i, pp]),
        "sadmrg_tensor_null": (i, [vp, i, i, pp]),
        "sadmrg_tensor_destroy": (i, [vp]),
        "sadmrg_tensor_header": (i, [vp, ip, i, ip]),
        "sadmrg_tensor_nten": (i, [vp, ip]),
        "sadmrg_tensor_nelem": (i, [vp, llp]),
        "sadmrg_tensor_download": (i, [vp, bp, dp]),
        "sadmrg_tensor_copy": (i, [vp, vp]),
        "sadmrg_tensor_device_ptr": (i, [vp, pp]),
        "sadmrg_hamiltonian_vector_multiplication": (i, [vp, vp, vp, vp, vp]),
        "sadmrg_contract": (i, [vp, vp, i, vp, i]),
        "sadmrg_env_update": (i, [vp, i, vp, vp, vp, vp, i, vp]),
        "sadmrg_svd": (i, [vp, vp, d, vp, d, i, dp, i, ip]),
        "sadmrg_fuse": (i, [vp, ip, ip]),
        "sadmrg_fuse_to_singlet": (i, [vp, ip, ip]),
        "sadmrg_make_spinor_start": (i, [vp, i]),
        "sadmrg_tensor_transform": (i, [vp, i, d, d]),
        "sadmrg_tensor_axpy": (i, [vp, d, d, vp]),
        "sadmrg_tensor_dot": (i, [vp, vp, dp]),
        "sadmrg_rng_seed": (i, [C.c_ulonglong]),
        "sadmrg_heff_create": (i, [vp, vp, vp, vp, i, pp, pp, vp, pp]),
        "sadmrg_heff_destroy": (i, [vp]),
        "sadmrg_heff_apply": (i, [vp, vp, vp, i]),
        "sadmrg_heff_apply_device": (i, [vp, vp, vp, i]),
        "sadmrg_heff_apply_stage": (i, [vp, vp, vp, i, i]),
        "sadmrg_heff_apply_unbatched": (i, [vp, vp, vp, i]),
        "sadmrg_heff_stage_flops": (i, [vp, i, dp]),
        "sadmrg_heff_issued_flops": (i, [vp, i, i, dp]),
        "sadmrg_heff_stats": (i, [vp, dp, dp, llp, ip]),
        "sadmrg_heff_merge": (i, [C.POINTER(vp), i, pp]),
        "sadmrg_lanczos": (i, [vp, vp, i, i, i, d, dp, ip, dp]),
        "sadmrg_lanczos_default_options": (i, [C.POINTER(LanczosOptions)]),
        "sadmrg_eigh_batched": (i, [vp, i, ip, pp, ip, pp, ip, pp, ip]),
        "sadmrg_eigvalsh_batched_noalloc": (i, [vp, i, ip, pp, ip, pp]),
        "sadmrg_svd_eigensolver": (i, [i, ip]),
        "sadmrg_lanczos_ex": (i, [vp, vp, C.POINTER(LanczosOptions), dp, ip, dp]),
        "sadmrg_lanczos_batched": (i, [vp, vp, i, C.POINTER(LanczosOptions), dp, ip, ip, dp, ip]),
        "sadmrg_autotune": (i, [i, i, ip]),
        "sadmrg_svd_gram_min": (i, [i, ip]),
        "sadmrg_svd_jacobi": (i, [i, ip, llp]),
        "sadmrg_eigh_clusters": (i, [i, ip, llp]),
        "sadmrg_svd_stats": (i, [C.POINTER(C.c_longlong), dp, i]),
        "sadmrg_lanczos_mode": (i, [i, i, ip, C.POINTER(C.c_longlong), i]),
        "sadmrg_gemm_vec_staging": (i, [i, ip]),
        "sadmrg_gemm_persistent": (i, [i, ip]),
        "sadmrg_heff_per_block": (i, [i, ip]),
        "sadmrg_heff_complex_planar": (i, [i, ip]),
        "sadmrg_heff_launch_plan": (i, [i, ip]),
        "sadmrg_gemm_balanced_split": (i, [i, ip]),
        "sadmrg_dmrg_create": (i, [vp, i, i, i, i, i, i, i, pp]),
        "sadmrg_dmrg_create_exci": (i, [vp, i, i, i, i, i, i, i, i, pp]),
        "sadmrg_dmrg_exci": (i, [vp, ip]),
        "sadmrg_dmrg_set_channel_weights": (i, [vp, dp, i]),
        "sadmrg_dmrg_set_lower_state": (i, [vp, i, vp]),
        "sadmrg_dmrg_lower_energy": (i, [vp, i, dp]),
        "sadmrg_dmrg_destroy": (i, [vp]),
        "sadmrg_dmrg_ns": (i, [vp, ip]),
        "sadmrg_dmrg_copy_state": (i, [vp, vp]),
        "sadmrg_dmrg_set_couplings": (i, [vp, ip, dp, i]),
        "sadmrg_dmrg_get_couplings": (i, [vp, ip, dp, i]),
        "sadmrg_dmrg_set_max_dcut": (i, [vp, i]),
        "sadmrg_dmrg_set_precision": (i, [vp, i, i, d]),
        "sadmrg_dmrg_set_refinement": (i, [vp, i, i]),
        "sadmrg_dmrg_set_check_interval": (i, [vp, i]),
        "sadmrg_dmrg_set_fp64_energy_every": (i, [vp, i]),
        "sadmrg_dmrg_set_verbose": (i, [vp, i]),
        "sadmrg_dmrg_save_mps": (i, [vp, C.c_char_p, i]),
        "sadmrg_dmrg_save_ww": (i, [vp, C.c_char_p]),
        "sadmrg_dmrg_save_enr": (i, [vp, C.c_char_p]),
        "sadmrg_dmrg_read_mps": (i, [vp, C.c_char_p, i, i, ip]),
        "sadmrg_dmrg_read_ww": (i, [vp, C.c_char_p, i]),
        "sadmrg_dmrg_read_enr": (i, [vp, C.c_char_p, i]),
        "sadmrg_dmrg_set_checkpoint_dir": (i, [vp, C.c_char_p]),
        "sadmrg_dmrg_ww": (i, [vp, dp, i, ip]),
        "sadmrg_dmrg_idmrg": (i, [vp]),
        "sadmrg_dmrg_prepare_sweep": (i, [vp]),
        "sadmrg_dmrg_sweep": (i, [vp]),
        "sadmrg_dmrg_energy": (i, [vp, dp]),
        "sadmrg_dmrg_records": (i, [vp, dp, i, ip]),
        "sadmrg_dmrg_get_tensor": (i, [vp, i, i, i, vp]),
        "sadmrg_dmrg_set_tensor": (i, [vp, i, i, i, vp]),
        "sadmrg_dmrg_call": (i, [vp, i, i, i]),
        "sadmrg_dmrg_input_vector": (i, [vp, i, i, vp]),
        "sadmrg_dmrg_coupling_gradient": (i, [vp, dp, i, dp]),
        "sadmrg_dmrg_heff": (i, [vp, i, i, vp, vp, i]),
        "sadmrg_dmrg_heff_create": (i, [vp, i, i, vp, C.POINTER(vp)]),
        "sadmrg_dmrg_lanczos": (i, [vp, i, i, vp, dp]),
        "sadmrg_dmrg_svd": (i, [vp, vp, i, i, d, d, dp, i, ip]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def exported_symbols():
    """Names declared in include/sadmrg.h (parsed from the header)."""
    import re
    text = open(INCLUDE).read()
    return sorted(set(re.findall(r"\b(sadmrg_[a-z0-9_]+)\s*\(", text)))


def check(rc):
    if rc != 0:
        msg = lib().sadmrg_last_error().decode()
        raise SadmrgError(rc, msg)


def _iptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _bptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_ubyte))


class Context:
    """One device, one stream (sadmrg_ctx)."""

    def __init__(self, device=0):
        p = C.c_void_p()
        check(lib().sadmrg_ctx_create(device, C.byref(p)))
        self.ptr = p

    def close(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.sadmrg_ctx_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        self.close()

    def sync(self):
        check(lib().sadmrg_ctx_sync(self.ptr))

    def stream(self):
        p = C.c_void_p()
        check(lib().sadmrg_ctx_stream(self.ptr, C.byref(p)))
        return p.value

    def memory(self):
        a, b = C.c_longlong(), C.c_longlong()
        check(lib().sadmrg_ctx_memory(self.ptr, C.byref(a), C.byref(b)))
        return a.value, b.value

    def host_stats(self):
        """(hipMalloc calls, their seconds, stream synchronisations, their seconds)."""
        a, b, c, d = C.c_longlong(), C.c_double(), C.c_longlong(), C.c_double()
        check(lib().sadmrg_ctx_stats(self.ptr, C.byref(a), C.byref(b), C.byref(c), C.byref(d)))
        return {"mallocs": a.value, "malloc_s": b.value, "syncs": c.value, "sync_s": d.value}


class Cgc:
    """SU(2) recoupling tables."""

    def __init__(self, physpn=1):
        p = C.c_void_p()
        check(lib().sadmrg_cgc_create(physpn, C.byref(p)))
        self.ptr = p

    def __del__(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.sadmrg_cgc_destroy(self.ptr)

    def hamilt_vec(self, i, j, k):
        o = C.c_double()
        check(lib().sadmrg_cgc_hamilt_vec(self.ptr, i, j, k, C.byref(o)))
        return o.value

    def onsite(self, side, j, m, l, i, p):
        o = C.c_double()
        check(lib().sadmrg_cgc_onsite(self.ptr, side, j, m, l, i, p, C.byref(o)))
        return o.value

    def transformation(self, side, j, m, l, i, k, p):
        o = C.c_double()
        check(lib().sadmrg_cgc_transformation(self.ptr, side, j, m, l, i, k, p, C.byref(o)))
        return o.value

    def pairup(self, side, j, i, k, m, p):
        o = C.c_double()
        check(lib().sadmrg_cgc_pairup(self.ptr, side, j, i, k, m, p, C.byref(o)))
        return o.value


def clebsch_gordan(j1, m1, j2, m2, j3, m3):
    o = C.c_double()
    check(lib().sadmrg_clebsch_gordan(j1, m1, j2, m2, j3, m3, C.byref(o)))
    return o.value


@dataclass
class Packed:
    """Block tensor in the C ABI exchange format (host arrays)."""
    sym: int
    cplx: int
    header: np.ndarray
    present: np.ndarray
    data: np.ndarray


class Tensor:
    """Device block tensor (sadmrg_tensor)."""

    def __init__(self, ctx, ptr, sym, cplx):
        self.ctx, self.ptr, self.sym, self.cplx = ctx, ptr, sym, cplx

    def __del__(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.sadmrg_tensor_destroy(self.ptr)
            self.ptr = None

    @classmethod
    def null(cls, ctx, sym, cplx=0):
        p = C.c_void_p()
        check(lib().sadmrg_tensor_null(ctx.ptr, sym, cplx, C.byref(p)))
        return cls(ctx, p, sym, cplx)

    @classmethod
    def from_packed(cls, ctx, pk, cgc2_kind=0):
        hdr = np.ascontiguousarray(pk.header, dtype=np.int32)
        pres = np.ascontiguousarray(pk.present, dtype=np.uint8)
        data = np.ascontiguousarray(pk.data, dtype=np.float64)
        if data.size == 0:
            data = np.zeros(1)
        if pres.size == 0:
            pres = np.zeros(1, dtype=np.uint8)
        p = C.c_void_p()
        check(lib().sadmrg_tensor_create(ctx.ptr, pk.sym, pk.cplx, _iptr(hdr), len(hdr), _bptr(pres), _dptr(data),
                                         cgc2_kind, C.byref(p)))
        return cls(ctx, p, pk.sym, pk.cplx)

    def header(self):
        n = C.c_int()
        check(lib().sadmrg_tensor_header(self.ptr, None, 0, C.byref(n)))
        h = np.zeros(max(n.value, 1), dtype=np.int32)
        check(lib().sadmrg_tensor_header(self.ptr, _iptr(h), len(h), C.byref(n)))
        return h[:n.value]

    def nten(self):
        n = C.c_int()
        check(lib().sadmrg_tensor_nten(self.ptr, C.byref(n)))
        return n.value

    def nelem(self):
        n = C.c_longlong()
        check(lib().sadmrg_tensor_nelem(self.ptr, C.byref(n)))
        return n.value

    def packed(self):
        nt = self.nten()
        ne = self.nelem()
        pres = np.zeros(max(nt, 1), dtype=np.uint8)
        w = 2 if self.cplx else 1
        data = np.zeros(max(ne * w, 1))
        check(lib().sadmrg_tensor_download(self.ptr, _bptr(pres), _dptr(data)))
        return Packed(self.sym, self.cplx, self.header(), pres[:nt].copy(), data[:ne * w].copy())

    def copy(self):
        t = Tensor.null(self.ctx, self.sym, self.cplx)
        check(lib().sadmrg_tensor_copy(t.ptr, self.ptr))
        return t

    def device_ptr(self):
        p = C.c_void_p()
        check(lib().sadmrg_tensor_device_ptr(self.ptr, C.byref(p)))
        return p.value

    def transform(self, op, a=0.0, b=0.0):
        check(lib().sadmrg_tensor_transform(self.ptr, op, a, b))
        return self

    def axpy(self, re, x, im=0.0):
        check(lib().sadmrg_tensor_axpy(self.ptr, re, im, x.ptr))
        return self

    def dot(self, y):
        o = np.zeros(2)
        check(lib().sadmrg_tensor_dot(self.ptr, y.ptr, _dptr(o)))
        return complex(o[0], o[1])


def hvm(vec, op1, op2, cgc=None):
    out = Tensor.null(vec.ctx, vec.sym, vec.cplx)
    check(lib().sadmrg_hamiltonian_vector_multiplication(out.ptr, vec.ptr, op1.ptr, op2.ptr,
                                                         cgc.ptr if cgc else None))
    return out


def contract(t1, i1, t2, i2):
    out = Tensor.null(t1.ctx, t1.sym, t1.cplx)
    check(lib().sadmrg_contract(out.ptr, t1.ptr, i1, t2.ptr, i2))
    return out


ENV_KIND = {"overlap_initial": 0, "overlap_transformation": 1, "operator_initial": 2, "overlap_adjoint": 5,
            "operator_transformation": 3, "operator_pairup": 4}


def env(kind, uu, vv, op1=None, op2=None, flag=0, cgc=None):
    out = Tensor.null(uu.ctx, uu.sym, uu.cplx)
    check(lib().sadmrg_env_update(out.ptr, ENV_KIND[kind], uu.ptr, vv.ptr, op1.ptr if op1 else None,
                                  op2.ptr if op2 else None, flag, cgc.ptr if cgc else None))
    return out


def svd(vec, tu, p1, tv, p2, max_dcut, cap=1 << 16):
    w = np.zeros(cap)
    n = C.c_int()
    check(lib().sadmrg_svd(vec.ptr, tu.ptr, p1, tv.ptr, p2, max_dcut, _dptr(w), cap, C.byref(n)))
    return w[:n.value].copy()


def _bond_arr(b):
    return np.asarray([b[0], len(b[1])] + list(b[1]) + list(b[2]), dtype=np.int32)


def fuse(ctx, sym, cplx, b1, b2):
    """b = (dir, labels, dims)"""
    out = Tensor.null(ctx, sym, cplx)
    a1, a2 = _bond_arr(b1), _bond_arr(b2)
    check(lib().sadmrg_fuse(out.ptr, _iptr(a1), _iptr(a2)))
    return out


class Heff:
    """Fused H_eff matvec plan (sadmrg_heff)."""

    def __init__(self, ctx, cgc, hl, hr, o1, o2, psi_layout):
        n = len(o1)
        A = (C.c_void_p * max(n, 1))(*[t.ptr.value if isinstance(t.ptr, C.c_void_p) else t.ptr for t in o1])
        B = (C.c_void_p * max(n, 1))(*[t.ptr.value if isinstance(t.ptr, C.c_void_p) else t.ptr for t in o2])
        p = C.c_void_p()
        check(lib().sadmrg_heff_create(ctx.ptr, cgc.ptr if cgc else None, hl.ptr if hl else None,
                                       hr.ptr if hr else None, n, A, B, psi_layout.ptr, C.byref(p)))
        self.ptr, self.ctx = p, ctx
        self._keep = (hl, hr, list(o1), list(o2), psi_layout)

    def __del__(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.sadmrg_heff_destroy(self.ptr)

    @classmethod
    def merge(cls, parts):
        """One plan over independent plans: every stage of every part in one
        launch; apply_device on the parts' vectors back to back."""
        arr = (C.c_void_p * len(parts))(*[p.ptr.value for p in parts])
        q = C.c_void_p()
        check(lib().sadmrg_heff_merge(arr, len(parts), C.byref(q)))
        h = cls.__new__(cls)
        h.ptr, h.ctx = q, parts[0].ctx
        h._keep = ()
        return h

    def apply(self, psi, precision=FP64):
        out = Tensor.null(psi.ctx, psi.sym, psi.cplx)
        check(lib().sadmrg_heff_apply(self.ptr, psi.ptr, out.ptr, precision))
        return out

    def apply_device(self, d_in, d_out, precision=FP64, unbatched=False):
        """H_eff psi on raw device arrays; unbatched: one launch per output
        block (the per-block baseline) instead of the grouped launch."""
        f = lib().sadmrg_heff_apply_unbatched if unbatched else lib().sadmrg_heff_apply_device
        check(f(self.ptr, C.c_void_p(d_in), C.c_void_p(d_out), precision))

    def apply_stage(self, d_in, d_out, precision=FP64, stage=0):
        check(lib().sadmrg_heff_apply_stage(self.ptr, C.c_void_p(d_in), C.c_void_p(d_out), precision, stage))

    def stage_flops(self, stage):
        f = C.c_double()
        check(lib().sadmrg_heff_stage_flops(self.ptr, stage, C.byref(f)))
        return f.value

    def issued_flops(self, precision=FP64, stage=0):
        """MFMA flops the last application's launch plans issue (edge tiles
        whole, K rounded up to the K step)."""
        f = C.c_double()
        check(lib().sadmrg_heff_issued_flops(self.ptr, stage, precision, C.byref(f)))
        return f.value

    def stats(self):
        f, b = C.c_double(), C.c_double()
        n = C.c_longlong()
        info = np.zeros(8, dtype=np.int32)
        check(lib().sadmrg_heff_stats(self.ptr, C.byref(f), C.byref(b), C.byref(n), _iptr(info)))
        return {"flops": f.value, "bytes": b.value, "n": n.value, "blocks": int(info[0]),
                "stage1_problems": int(info[1]), "stage2_problems": int(info[2]), "tiles1": int(info[3]),
                "tiles2": int(info[4]), "tile_cfg1": int(info[5]), "tile_cfg2": int(info[6]),
                "split_tiles": int(info[7])}

    def lanczos_ex(self, psi, **options):
        """sadmrg_lanczos_ex: every LanczosOptions field may be given by name
        (mlanc, precision, refine, refine_tol, refine_krylov, max_restarts,
        check_every, device_steps, graph); the rest keep their defaults."""
        o = LanczosOptions()
        check(lib().sadmrg_lanczos_default_options(C.byref(o)))
        for k, v in options.items():
            if k not in dict(LanczosOptions._fields_):
                raise TypeError(f"unknown Lanczos option {k}")
            setattr(o, k, type(getattr(o, k))(v))
        e = C.c_double()
        cnt = np.zeros(4, dtype=np.int32)
        res = C.c_double()
        check(lib().sadmrg_lanczos_ex(self.ptr, psi.ptr, C.byref(o), C.byref(e), _iptr(cnt), C.byref(res)))
        return e.value, {"matvec64": int(cnt[0]), "matvec32": int(cnt[1]), "steps": int(cnt[2]),
                         "restarts": int(cnt[3]), "residual64": res.value}

    def lanczos_batched(self, d_psi, n_inst, **options):
        """sadmrg_lanczos_batched on a merged plan: d_psi (device pointer) holds
        the instances' start vectors back to back and receives their Ritz
        vectors; options as lanczos_ex.  Returns (energies, info)."""
        o = LanczosOptions()
        check(lib().sadmrg_lanczos_default_options(C.byref(o)))
        for k, v in options.items():
            if k not in dict(LanczosOptions._fields_):
                raise TypeError(f"unknown Lanczos option {k}")
            setattr(o, k, type(getattr(o, k))(v))
        e = np.zeros(n_inst)
        steps = np.zeros(n_inst, dtype=np.int32)
        rst = np.zeros(n_inst, dtype=np.int32)
        res = np.zeros(n_inst)
        cnt = np.zeros(2, dtype=np.int32)
        check(lib().sadmrg_lanczos_batched(self.ptr, C.c_void_p(d_psi), n_inst, C.byref(o), _dptr(e), _iptr(steps),
                                           _iptr(rst), _dptr(res), _iptr(cnt)))
        return e, {"steps": steps, "restarts": rst, "residual64": res, "matvec64": int(cnt[0]),
                   "matvec32": int(cnt[1])}

    def lanczos(self, psi, mlanc=40, precision=FP64, refine=False, refine_tol=0.0):
        e = C.c_double()
        cnt = np.zeros(4, dtype=np.int32)
        res = C.c_double()
        check(lib().sadmrg_lanczos(self.ptr, psi.ptr, mlanc, precision, 1 if refine else 0, refine_tol,
                                   C.byref(e), _iptr(cnt), C.byref(res)))
        return e.value, {"matvec64": int(cnt[0]), "matvec32": int(cnt[1]), "steps": int(cnt[2]),
                         "restarts": int(cnt[3]), "residual64": res.value}


def autotune(enable=None, clear=False):
    """Tile-configuration autotuner (process-wide); returns the table size."""
    n = C.c_int()
    check(lib().sadmrg_autotune(-1 if enable is None else int(bool(enable)), int(bool(clear)), C.byref(n)))
    return n.value


def vec_staging(enable=None):
    """16-byte operand staging of the H_eff GEMMs (process-wide); returns the
    previous setting."""
    prev = C.c_int()
    check(lib().sadmrg_gemm_vec_staging(-1 if enable is None else int(bool(enable)), C.byref(prev)))
    return bool(prev.value)


def heff_launch_plan(plan=None):
    """Force every H_eff stage onto one launch plan (-1: tuned/default);
    returns the previous setting."""
    prev = C.c_int()
    check(lib().sadmrg_heff_launch_plan(-2 if plan is None else int(plan), C.byref(prev)))
    return prev.value


def heff_complex_planar(enable=None):
    """Complex H_eff plans as real tables on planar operands (process-wide,
    default off: the complex kernel); returns the previous setting."""
    prev = C.c_int()
    check(lib().sadmrg_heff_complex_planar(-1 if enable is None else int(bool(enable)), C.byref(prev)))
    return bool(prev.value)


def heff_per_block(enable=None):
    """Every H_eff application as one launch per output block (process-wide);
    returns the previous setting."""
    prev = C.c_int()
    check(lib().sadmrg_heff_per_block(-1 if enable is None else int(bool(enable)), C.byref(prev)))
    return bool(prev.value)


def persistent(mode=None):
    """Persistent tile-queue launches of the real grouped GEMMs
    (process-wide): 0 never, 1 always, 2 autotuned per H_eff stage (the
    default); returns the previous mode."""
    prev = C.c_int()
    check(lib().sadmrg_gemm_persistent(-1 if mode is None else int(mode), C.byref(prev)))
    return prev.value


def balanced_split(enable=None):
    """Split-reduction policy of the grouped GEMMs for plans built afterwards
    (process-wide): balanced (parts of equal depth, the default) or uniform
    (the same number of parts per problem).  Returns the previous setting."""
    prev = C.c_int()
    check(lib().sadmrg_gemm_balanced_split(-1 if enable is None else int(bool(enable)), C.byref(prev)))
    return bool(prev.value)


def eigh_batched(ctx, n, a, lda, w, k=None, z=None, ldz=None):
    """sadmrg_eigh_batched on device buffers: n, lda, k, ldz are lists of ints,
    a, w, z lists of device addresses (ints).  Eigenvalues ascending into w;
    eigenvectors of the k largest into z."""
    cnt = len(n)
    arr = lambda xs: (C.c_int * max(cnt, 1))(*xs)  # noqa: E731
    ptrs = lambda xs: (C.c_void_p * max(cnt, 1))(*[C.c_void_p(x) for x in xs])  # noqa: E731
    k = k or [0] * cnt
    z = z or [0] * cnt
    ldz = ldz or [1] * cnt
    check(lib().sadmrg_eigh_batched(ctx.ptr, cnt, arr(n), ptrs(a), arr(lda), ptrs(w), arr(k), ptrs(z), arr(ldz)))


def eigvalsh_batched_noalloc(ctx, n, a, lda, w):
    """sadmrg_eigvalsh_batched_noalloc: eigenvalues only, ascending into w,
    with the library's static per-device scratch (no allocation)."""
    cnt = len(n)
    arr = lambda xs: (C.c_int * max(cnt, 1))(*xs)  # noqa: E731
    ptrs = lambda xs: (C.c_void_p * max(cnt, 1))(*[C.c_void_p(x) for x in xs])  # noqa: E731
    check(lib().sadmrg_eigvalsh_batched_noalloc(ctx.ptr, cnt, arr(n), ptrs(a), arr(lda), ptrs(w)))


def svd_eigensolver(value=None):
    """Eigensolver of the truncation's Gram route (process-wide): 1 the
    hand-written batched solver, 0 rocSOLVER syevd on padded batches.
    Returns the previous setting."""
    prev = C.c_int()
    check(lib().sadmrg_svd_eigensolver(-1 if value is None else int(value), C.byref(prev)))
    return prev.value


def lanczos_mode(device_steps=None, graph=None, reset=False):
    """Process-wide Lanczos step mode.  Sets the given switches and returns
    (device_steps, graph, graph replays, graphs captured) as they were before
    the call."""
    prev = (C.c_int * 2)()
    counts = (C.c_longlong * 2)()
    check(lib().sadmrg_lanczos_mode(-1 if device_steps is None else int(bool(device_steps)),
                                    -1 if graph is None else int(bool(graph)),
                                    prev, counts, 1 if reset else 0))
    return bool(prev[0]), bool(prev[1]), int(counts[0]), int(counts[1])


def svd_stats(reset=False):
    """Truncation counters since the last reset (process-wide)."""
    c = (C.c_longlong * 4)()
    sec = C.c_double()
    check(lib().sadmrg_svd_stats(c, C.byref(sec), int(bool(reset))))
    return {"calls": c[0], "gram_sectors": c[1], "gesdd_sectors": c[2], "gram_redone": c[3],
            "seconds": sec.value}


def _path(directory):
    return b"" if directory is None else os.fsencode(str(directory))


def eigh_clusters(mode=-1):
    """Inverse iteration over clusters of close eigenvalues: 0 members in
    parallel lanes (default), 1 every cluster serially, -1 only reads.
    Process-wide.  Returns (previous mode, clusters run again serially)."""
    prev = C.c_int()
    cnt = C.c_longlong()
    check(lib().sadmrg_eigh_clusters(int(mode), C.byref(prev), C.byref(cnt)))
    return prev.value, int(cnt.value)


def svd_jacobi(max_order=-1):
    """Largest min(m, n) of the real sectors the truncation decomposes by the
    one-sided Jacobi kernel (0 off, -1 only reads; process-wide).  Returns
    (previous value, sectors decomposed by it, sectors handed to gesdd)."""
    prev = C.c_int()
    cnt = (C.c_longlong * 2)()
    check(lib().sadmrg_svd_jacobi(int(max_order), C.byref(prev), cnt))
    return prev.value, int(cnt[0]), int(cnt[1])


def svd_gram_min(value=-1):
    """Smallest sector (rows of its larger side) that sweep-step truncations
    decompose through the Gram matrix; 0 turns the route off, -1 only reads.
    Process-wide.  Returns the previous value."""
    prev = C.c_int()
    check(lib().sadmrg_svd_gram_min(int(value), C.byref(prev)))
    return prev.value


REC_FIELDS = ("il", "ir", "energy", "trunc_err", "kept", "matvec64", "matvec32", "t_env", "t_lanczos", "t_svd",
              "heff_flops", "ndata", "ritz32", "energy64")


class DMRG:
    """DMRG<Scalar,Sym> on the device engine (sadmrg_dmrg)."""

    def __init__(self, ctx, sym, cplx, xx, yy, bondd, physpn=1, totspin=0, exci=0):
        p = C.c_void_p()
        check(lib().sadmrg_dmrg_create_exci(ctx.ptr, sym, cplx, xx, yy, bondd, physpn, totspin, exci, C.byref(p)))
        self.ptr, self.ctx, self.sym, self.cplx = p, ctx, sym, cplx
        self._xx, self._yy, self._bondd, self._physpn, self._totspin = xx, yy, bondd, physpn, totspin
        self.exci = exci
        n = C.c_int()
        check(lib().sadmrg_dmrg_ns(p, C.byref(n)))
        self.ns = n.value

    def __del__(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.sadmrg_dmrg_destroy(self.ptr)

    def clone(self):
        """A new driver holding a device-side deep copy of this one's state."""
        p = C.c_void_p()
        check(lib().sadmrg_dmrg_create_exci(self.ctx.ptr, self.sym, self.cplx, self._xx, self._yy, self._bondd,
                                            self._physpn, self._totspin, self.exci, C.byref(p)))
        other = DMRG.__new__(DMRG)
        other.ptr, other.ctx, other.sym, other.cplx, other.ns = p, self.ctx, self.sym, self.cplx, self.ns
        other.exci = self.exci
        other._xx, other._yy, other._bondd, other._physpn, other._totspin = (
            self._xx, self._yy, self._bondd, self._physpn, self._totspin)
        check(lib().sadmrg_dmrg_copy_state(p, self.ptr))
        return other

    def set_lower_state(self, j, other):
        """Stored state j (< exci) := other's MPS and energy."""
        check(lib().sadmrg_dmrg_set_lower_state(self.ptr, j, other.ptr))

    def lower_energy(self, j):
        e = C.c_double()
        check(lib().sadmrg_dmrg_lower_energy(self.ptr, j, C.byref(e)))
        return e.value

    def set_couplings(self, hmap, coup):
        h = np.ascontiguousarray(hmap, dtype=np.int32).reshape(-1)
        c = np.ascontiguousarray(coup, dtype=np.float64)
        check(lib().sadmrg_dmrg_set_couplings(self.ptr, _iptr(h), _dptr(c), len(c)))

    def set_channel_weights(self, w):
        """w: (ncoup, 3) complex weights for the channel labels q = -2, 0, +2
        (None restores the isotropic exchange).  See xxz_dm_weights."""
        if w is None:
            check(lib().sadmrg_dmrg_set_channel_weights(self.ptr, None, 0))
            return
        w = np.asarray(w, dtype=np.complex128).reshape(-1, 3)
        flat = np.ascontiguousarray(np.stack([w.real, w.imag], axis=-1).reshape(-1))
        check(lib().sadmrg_dmrg_set_channel_weights(self.ptr, _dptr(flat), len(flat)))

    def couplings(self):
        h = np.zeros(self.ns * self.ns, dtype=np.int32)
        c = np.zeros(self.ns + 1)
        check(lib().sadmrg_dmrg_get_couplings(self.ptr, _iptr(h), _dptr(c), len(c)))
        return h.reshape(self.ns, self.ns), c

    def set_max_dcut(self, d):
        check(lib().sadmrg_dmrg_set_max_dcut(self.ptr, d))

    def set_precision(self, precision=FP64, refine=False, refine_tol=0.0):
        check(lib().sadmrg_dmrg_set_precision(self.ptr, precision, 1 if refine else 0, refine_tol))

    def set_refinement(self, krylov_steps=6, max_restarts=6):
        check(lib().sadmrg_dmrg_set_refinement(self.ptr, krylov_steps, max_restarts))

    def set_verbose(self, level=1):
        """One line per sweep step on stderr (sadmrg_dmrg_set_verbose)."""
        check(lib().sadmrg_dmrg_set_verbose(self.ptr, level))

    def set_check_interval(self, every=4):
        """Lanczos convergence check after every `every`-th step (reference: 4)."""
        check(lib().sadmrg_dmrg_set_check_interval(self.ptr, every))

    def set_fp64_energy_every(self, every=1):
        """Mixed precision: FP64 energies at every step (1), every k-th step
        and the last (k > 1) or the last step of a sweep only (0)."""
        check(lib().sadmrg_dmrg_set_fp64_energy_every(self.ptr, every))

    # checkpoint files in the reference's format (dmrg_inst.cpp:947-1177)
    def save_mps(self, directory=None, exci=0):
        """save_mps2(exci): uu-{ns}-{bondd}-{totspin}-{exci}-{i}.dat/.bin per site."""
        check(lib().sadmrg_dmrg_save_mps(self.ptr, _path(directory), exci))

    def save_ww(self, directory=None):
        check(lib().sadmrg_dmrg_save_ww(self.ptr, _path(directory)))

    def save_enr(self, directory=None):
        check(lib().sadmrg_dmrg_save_enr(self.ptr, _path(directory)))

    def read_mps(self, directory, read1, read2=None):
        """read_mps(read1) or, with read2, read_mps(read1, read2) (also ww and
        the energy); returns the reference's pass flag."""
        ok = C.c_int()
        check(lib().sadmrg_dmrg_read_mps(self.ptr, _path(directory), read1, -1 if read2 is None else read2,
                                         C.byref(ok)))
        return bool(ok.value)

    def read_ww(self, directory, read1):
        check(lib().sadmrg_dmrg_read_ww(self.ptr, _path(directory), read1))

    def read_enr(self, directory, read1):
        check(lib().sadmrg_dmrg_read_enr(self.ptr, _path(directory), read1))

    def set_checkpoint_dir(self, directory):
        """sweep() writes the reference's checkpoint files into directory
        (None: no files, the default)."""
        check(lib().sadmrg_dmrg_set_checkpoint_dir(self.ptr, None if directory is None else _path(directory)))

    def ww(self):
        """Singular values of the last truncation (or what read_ww loaded)."""
        n = C.c_int()
        check(lib().sadmrg_dmrg_ww(self.ptr, None, 0, C.byref(n)))
        w = np.zeros(n.value)
        if n.value:
            check(lib().sadmrg_dmrg_ww(self.ptr, _dptr(w), n.value, C.byref(n)))
        return w

    def idmrg(self):
        check(lib().sadmrg_dmrg_idmrg(self.ptr))

    def prepare_sweep(self):
        check(lib().sadmrg_dmrg_prepare_sweep(self.ptr))

    def sweep(self):
        check(lib().sadmrg_dmrg_sweep(self.ptr))

    def energy(self):
        e = C.c_double()
        check(lib().sadmrg_dmrg_energy(self.ptr, C.byref(e)))
        return e.value

    def records(self):
        out = []
        while True:
            nf = len(REC_FIELDS)
            buf = np.zeros(nf * 256)
            n = C.c_int()
            check(lib().sadmrg_dmrg_records(self.ptr, _dptr(buf), 256, C.byref(n)))
            for k in range(n.value):
                out.append(dict(zip(REC_FIELDS, buf[nf * k:nf * k + nf].tolist())))
            if n.value < 256:
                return out

    _SLOTS = {"uu": 0, "hh": 1, "opr": 2, "orth": 3, "ovlp": 4, "overlapvec": 5}

    def tensor(self, which, i, j=0):
        """uu[i], hh[i], opr[i][j]; orth[j][i], ovlp[j][i], overlapvec[j] of
        stored state j."""
        code = self._SLOTS[which]
        t = Tensor.null(self.ctx, self.sym, self.cplx)
        check(lib().sadmrg_dmrg_get_tensor(self.ptr, code, i, j, t.ptr))
        return t

    def set_tensor(self, which, i, j, t):
        code = self._SLOTS[which]
        check(lib().sadmrg_dmrg_set_tensor(self.ptr, code, i, j, t.ptr))

    def call(self, name, a=0, b=0):
        code = {"wavefunc_transformation": 0, "left2": 1, "right2": 2, "left1": 3, "right1": 4,
                "prepare_sweep": 5}[name]
        check(lib().sadmrg_dmrg_call(self.ptr, code, a, b))

    def coupling_gradient(self):
        """(dE/dJ_c for c = 0..ns, <psi|H|psi>) at the last Lanczos solve."""
        g = np.zeros(self.ns + 1)
        e = C.c_double()
        check(lib().sadmrg_dmrg_coupling_gradient(self.ptr, _dptr(g), len(g), C.byref(e)))
        return g, e.value

    def input_vector(self, il, ir):
        t = Tensor.null(self.ctx, self.sym, self.cplx)
        check(lib().sadmrg_dmrg_input_vector(self.ptr, il, ir, t.ptr))
        return t

    def heff(self, il, ir, vin, precision=FP64):
        t = Tensor.null(self.ctx, self.sym, self.cplx)
        check(lib().sadmrg_dmrg_heff(self.ptr, il, ir, vin.ptr, t.ptr, precision))
        return t

    def heff_plan(self, il, ir, layout):
        """The H_eff plan of step (il, ir) as the sweep builds it (a Heff)."""
        q = C.c_void_p()
        check(lib().sadmrg_dmrg_heff_create(self.ptr, il, ir, layout.ptr, C.byref(q)))
        h = Heff.__new__(Heff)
        h.ptr, h.ctx = q, self.ctx
        h._keep = (self, layout)
        return h

    def lanczos(self, il, ir, vec):
        e = C.c_double()
        check(lib().sadmrg_dmrg_lanczos(self.ptr, il, ir, vec.ptr, C.byref(e)))
        return e.value

    def svd(self, vec, a, b, p1, p2, cap=1 << 16):
        w = np.zeros(cap)
        n = C.c_int()
        check(lib().sadmrg_dmrg_svd(self.ptr, vec.ptr, a, b, p1, p2, _dptr(w), cap, C.byref(n)))
        return w[:n.value].copy()


# ---- lattice maps for DMRG.set_couplings -------------------------------------
def xxz_dm_weights(ncoup, params):
    """Channel weights of XXZ exchange with a z-axis Dzyaloshinskii-Moriya
    term, coup[c] * [Jxy (S^x S^x + S^y S^y) + Jz S^z S^z + D (S^x S^y - S^y S^x)]
    for the coupling indices in params = {c: (Jxy, Jz, D)}; other indices
    keep the isotropic exchange.  Returns an (ncoup, 3) complex array for
    DMRG.set_channel_weights."""
    w = np.ones((ncoup, 3), dtype=np.complex128)
    for c, (jxy, jz, dm) in params.items():
        w[c] = (jxy - 1j * dm, jz, jxy + 1j * dm)
    return w


def snake_site(x, y, width):
    """Snake ordering of the reference constructor (dmrg_inst.cpp:121-140):
    even columns run up, odd columns run down."""
    return x * width + (y if x % 2 == 0 else width - 1 - y)


def reference_lattice(xx, yy):
    """hmap built by DMRG(xx, yy, ...) (dmrg_inst.cpp:117-144): nearest
    neighbours, periodic in both directions, coupling index 1."""
    ns = xx * yy
    h = np.zeros((ns, ns), dtype=np.int32)
    for i in range(ns):
        x1 = i // yy
        y1 = i % yy if x1 % 2 == 0 else yy - 1 - (i % yy)
        for x2, y2 in (((x1 + 1) % xx, y1), (x1, (y1 + 1) % yy)):
            j = snake_site(x2, y2, yy)
            if i != j:
                h[i, j] = h[j, i] = 1
    return h


def open_chain(length):
    """Open chain, coupling index 1 on every bond."""
    h = np.zeros((length, length), dtype=np.int32)
    for i in range(length - 1):
        h[i, i + 1] = h[i + 1, i] = 1
    return h


def j1j2_cylinder(width, length):
    """Snake-ordered width x length cylinder, periodic around the width and
    open along the length: index 1 on nearest-neighbour bonds (J1), index 2 on
    diagonal bonds between neighbouring columns (J2)."""
    ns = width * length
    h = np.zeros((ns, ns), dtype=np.int32)
    for x in range(length):
        for y in range(width):
            i = snake_site(x, y, width)
            nb = [(x, (y + 1) % width)] + ([(x + 1, y)] if x + 1 < length else [])
            for a, b in nb:
                j = snake_site(a, b, width)
                if i != j:
                    h[i, j] = h[j, i] = 1
            if x + 1 < length:
                for dy in (1, -1):
                    j = snake_site(x + 1, (y + dy) % width, width)
                    if i != j and h[i, j] == 0:
                        h[i, j] = h[j, i] = 2
    return h
