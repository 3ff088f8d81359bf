"""CPU restatement of the reference's two-site DMRG inner loop -- TEST
INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg as the checker; never by the product.

numpy, FP64.  Each function follows the reference function cited in its
docstring (file:line in lingwangqs/symmetry-adapted-DMRG-library).  Block
tensors are kept as {block index: ndarray} with numpy axes in bond order and
Fortran storage, so .ravel(order="F") is the reference's flat layout.

Pinning (what fixes each function independently of this file):
  * cg, fac tables        -> tests/golden/cgc_su2.npz (reference tables, all
                             2J < 20) and closed forms (orthonormality, known
                             values) in tests/test_oracle.py
  * hvm, contract         -> tests/golden/hvm_*.npz (reference outputs)
  * env updates, H_eff,
    Lanczos, svd          -> tests/golden/dmrg_*.npz (reference mid-sweep
                             states, H_eff*psi, Lanczos energy, singular values)
                             and exact diagonalisation of small lattices
"""
from __future__ import annotations

import math
from functools import lru_cache

import numpy as np


# ---- bookkeeping (src/symmetry/su2.cpp, src/symmetry/u1.cpp) -------------------
class BT:
    """Block tensor: sym 1 = SU(2), 0 = U(1); bonds = [(dir, labels, dims)]."""

    def __init__(self, sym, cplx, locspin, bonds, blocks=None):
        self.sym, self.cplx, self.locspin = sym, cplx, locspin
        self.bonds = [(int(d), [int(x) for x in q], [int(x) for x in n]) for d, q, n in bonds]
        self.blocks = dict(blocks or {})

    @property
    def nbond(self):
        return len(self.bonds)

    def nten(self):
        n = 1
        for b in self.bonds:
            n *= len(b[1])
        return n

    def idx(self, blk):
        out = []
        for b in self.bonds:
            out.append(blk % len(b[1]))
            blk //= len(b[1])
        return out

    def labels(self, blk):
        return [self.bonds[k][1][i] for k, i in enumerate(self.idx(blk))]

    def dims(self, blk):
        return [self.bonds[k][2][i] for k, i in enumerate(self.idx(blk))]

    def index(self, labels):
        """column-major block index (su2.cpp:432-450); None if a label is missing"""
        b = 0
        for k in range(self.nbond - 1, -1, -1):
            q = self.bonds[k][1]
            if labels[k] not in q:
                return None
            b = b * len(q) + q.index(labels[k])
        return b

    def allowed_labels(self, labels):
        if self.sym == 0:   # u1.cpp:367-378
            if any(labels[k] not in self.bonds[k][1] for k in range(self.nbond)):
                return False
            return sum(self.bonds[k][0] * labels[k] for k in range(self.nbond)) == self.locspin
        if self.nbond == 1:  # su2.cpp:483-507
            return labels[0] == self.locspin
        reach = {labels[0]}
        for j in labels[1:]:
            nxt = set()
            for J in reach:
                nxt.update(range(abs(J - j), J + j + 1, 2))
            reach = nxt
        return self.locspin in reach

    def allowed(self, blk):
        return self.allowed_labels(self.labels(blk))

    def copy(self):
        return BT(self.sym, self.cplx, self.locspin, self.bonds, {k: v.copy() for k, v in self.blocks.items()})


def inv(b):
    return (-b[0], list(b[1]), list(b[2]))


def fuse_bonds(sym, a, b):
    """SU2Bond::fuse (su2.cpp:182-217) / U1Bond::fuse (u1.cpp:111-141)."""
    acc = {}
    order = []
    if sym == 1:
        assert a[0] == b[0]
        for j1, d1 in zip(a[1], a[2]):
            for j2, d2 in zip(b[1], b[2]):
                for jk in range(abs(j1 - j2), j1 + j2 + 1, 2):
                    if jk not in acc:
                        order.append(jk)
                        acc[jk] = 0
                    acc[jk] += d1 * d2
        ks = sorted(acc)
        return (-a[0], ks, [acc[k] for k in ks])
    for q2, d2 in zip(b[1], b[2]):
        for q1, d1 in zip(a[1], a[2]):
            c = q1 + q2 if a[0] * b[0] == 1 else -q1 + q2
            acc[c] = acc.get(c, 0) + d1 * d2
    ks = sorted(acc)
    return (-b[0], ks, [acc[k] for k in ks])


# ---- exchange format ------------------------------------------------------------------
def from_packed(d):
    h = [int(x) for x in d["header"]]
    nb, ls = h[0], h[1]
    bonds, p = [], 2
    for _ in range(nb):
        dr, nm = h[p], h[p + 1]
        bonds.append((dr, h[p + 2:p + 2 + nm], h[p + 2 + nm:p + 2 + 2 * nm]))
        p += 2 + 2 * nm
    t = BT(int(d["sym"]), int(d["cplx"]), ls, bonds)
    off = 0
    w = 2 if t.cplx else 1
    for b in range(t.nten() if nb else 0):
        if b < len(d["present"]) and d["present"][b]:
            dims = t.dims(b)
            n = int(np.prod(dims))
            x = np.asarray(d["data"][off * w:(off + n) * w], dtype=np.float64)
            x = x[0::2] + 1j * x[1::2] if t.cplx else x
            t.blocks[b] = x.reshape(dims, order="F")
            off += n
    return t


def to_packed(t):
    h = [t.nbond, t.locspin]
    for dr, q, n in t.bonds:
        h += [dr, len(q)] + list(q) + list(n)
    nt = t.nten() if t.nbond else 0
    pres = np.zeros(nt, dtype=np.uint8)
    chunks = []
    for b in range(nt):
        if b in t.blocks:
            pres[b] = 1
            x = t.blocks[b].ravel(order="F")
            if t.cplx:
                y = np.empty(2 * x.size)
                y[0::2], y[1::2] = x.real, x.imag
                x = y
            chunks.append(np.asarray(x, dtype=np.float64))
    data = np.concatenate(chunks) if chunks else np.zeros(0)
    return dict(sym=t.sym, cplx=t.cplx, header=np.asarray(h, np.int32), present=pres, data=data)


# ---- dense helpers (dense_inst.cpp) ------------------------------------------------------
def dcontract(A, ia, B, ib):
    """Dense::contract (dense_inst.cpp:369-408): remaining axes of A cyclically
    after ia, then those of B after ib."""
    ra = [(ia + 1 + i) % A.ndim for i in range(A.ndim - 1)]
    rb = [(ib + 1 + i) % B.ndim for i in range(B.ndim - 1)]
    return np.tensordot(A.transpose(ra + [ia]), B.transpose([ib] + rb), axes=1)


# ---- Clebsch-Gordan and recoupling factors (su2.cpp:41-93, cgc_tables.cpp) ---------
def cg(j1, m1, j2, m2, j3, m3):
    """<j1 m1; j2 m2 | j3 m3>, doubled arguments (Racah's formula)."""
    if m1 + m2 != m3 or j3 < abs(j1 - j2) or j3 > j1 + j2 or (j1 + j2 + j3) % 2:
        return 0.0
    if abs(m1) > j1 or abs(m2) > j2 or abs(m3) > j3:
        return 0.0
    f = math.factorial
    pre = math.sqrt((j3 + 1) * f((j3 + j1 - j2) // 2) * f((j3 - j1 + j2) // 2) * f((j1 + j2 - j3) // 2)
                    / f((j1 + j2 + j3) // 2 + 1))
    mf = math.sqrt(f((j1 + m1) // 2) * f((j1 - m1) // 2) * f((j2 + m2) // 2) * f((j2 - m2) // 2)
                   * f((j3 + m3) // 2) * f((j3 - m3) // 2))
    a1, a2, a3 = (j1 + j2 - j3) // 2, (j1 - m1) // 2, (j2 + m2) // 2
    b1, b2 = (j3 - j2 + m1) // 2, (j3 - j1 - m2) // 2
    s = 0.0
    for k in range((j1 + j2 + j3) // 2 + 1):
        if min(a1 - k, a2 - k, a3 - k, b1 + k, b2 + k) < 0:
            continue
        s += (-1) ** k / (f(k) * f(a1 - k) * f(a2 - k) * f(a3 - k) * f(b1 + k) * f(b2 + k))
    return pre * mf * s


@lru_cache(maxsize=None)
def cgc_tensor(j1, j2, j3):
    """CGCTable::fill_block (su2.cpp:79-93): T[i, j, k], m = 2i - j1 etc."""
    T = np.zeros((j1 + 1, j2 + 1, j3 + 1))
    for i in range(j1 + 1):
        for j in range(j2 + 1):
            for k in range(j3 + 1):
                T[i, j, k] = cg(j1, 2 * i - j1, j2, 2 * j - j2, j3, 2 * k - j3)
    return T


@lru_cache(maxsize=None)
def singlet(j):
    """Dense::make_singlet (dense_inst.cpp:685-701)."""
    S = np.zeros((j + 1, j + 1))
    for p in range(j + 1):
        S[p, j - p] = (1.0 if (j - p) % 2 == 0 else -1.0) / math.sqrt(j + 1)
    return S


def _ratio(r, o):
    """Dense::is_proportional_to (dense_inst.cpp:800-844) as used for the tables."""
    if r.shape != o.shape:
        return 0.0
    if np.all(np.abs(r) <= 1e-12):
        return 0.0
    flat_o = o.ravel(order="F")
    pos = int(np.argmax(np.abs(flat_o)))
    return float(r.ravel(order="F")[pos] / flat_o[pos])


@lru_cache(maxsize=None)
def fac_hamilt_vec(i, j, k):
    """cgc_tables.cpp:156-164."""
    if not (abs(i - j) <= k <= i + j and i % 2 == k % 2):
        return 0.0
    C = cgc_tensor(i, j, k)
    t = np.einsum("abc,pa,qb,rc->pqr", C, singlet(i), singlet(j), singlet(k))
    return float(np.sum(C * t) * math.sqrt(j + 1))


def _tri(t1, t2, t3, pattern_b):
    # A: sum t1(a,b,c) t3(a,y,z) t2(z,b,c');  B: sum t1(a,b,c) t3(b,y,z) t2(a,z,c')
    if not pattern_b:
        return np.einsum("abc,ayz,zbd->cyd", t1, t3, t2)
    return np.einsum("abc,byz,azd->cyd", t1, t3, t2)


@lru_cache(maxsize=None)
def fac_onsite(side, j, m, l, i, pidx, physpn=1):
    """fac_operator_onsite_{left,rght}[j][m][l][i][p0+p1*P2] (cgc_tables.cpp:190-227)."""
    P2 = 2 * physpn + 1
    p0, p1, k = pidx % P2, pidx // P2, i
    if m % 2 or not 0 <= m <= 6:
        return 0.0
    if not (abs(i - p0) <= j <= i + p0 and abs(k - p1) <= l <= k + p1 and abs(j - m) <= l <= j + m
            and j % 2 == l % 2 and abs(p0 - m) <= p1 <= p0 + m):
        return 0.0
    t3 = cgc_tensor(p0, m, p1)
    if side == 0:
        r = _tri(cgc_tensor(i, p0, j), cgc_tensor(k, p1, l), t3, True)
    else:
        r = _tri(cgc_tensor(p0, i, j), cgc_tensor(p1, k, l), t3, False)
    return _ratio(r, cgc_tensor(j, m, l))


@lru_cache(maxsize=None)
def fac_transformation(side, j, m, l, i, k, p0):
    """fac_operator_transformation_{left,rght}[j][m][l][i][k][p0] (cgc_tables.cpp:257-297)."""
    p1 = p0
    if m % 2 or not 0 <= m <= 6:
        return 0.0
    if not (abs(i - p0) <= j <= i + p0 and abs(k - p1) <= l <= k + p1 and abs(i - m) <= k <= i + m
            and k % 2 == i % 2 and abs(j - m) <= l <= j + m and j % 2 == l % 2):
        return 0.0
    t3 = cgc_tensor(i, m, k)
    if side == 0:
        r = _tri(cgc_tensor(i, p0, j), cgc_tensor(k, p1, l), t3, False)
    else:
        r = _tri(cgc_tensor(p0, i, j), cgc_tensor(p1, k, l), t3, True)
    return _ratio(r, cgc_tensor(j, m, l))


@lru_cache(maxsize=None)
def fac_pairup(side, j, i, k, m, pidx, physpn=1):
    """fac_operator_pairup_{left,rght}[j][i][k][m][p0+p1*P2] (cgc_tables.cpp:323-357)."""
    P2 = 2 * physpn + 1
    p0, p1, l = pidx % P2, pidx // P2, j
    if m % 2 or not 0 <= m <= 6:
        return 0.0
    if not (abs(i - p0) <= j <= i + p0 and abs(k - p1) <= l <= k + p1 and abs(i - m) <= k <= i + m
            and abs(p0 - m) <= p1 <= p0 + m):
        return 0.0
    op2 = np.einsum("abc,pb->apc", cgc_tensor(p0, m, p1), singlet(m)) * math.sqrt(m + 1)
    op1 = cgc_tensor(i, m, k)
    if side == 0:
        uu, vv = cgc_tensor(i, p0, j), cgc_tensor(k, p1, l)
        r = np.einsum("abc,ayz,byr,zrd->cd", uu, op1, op2, vv)
    else:
        uu, vv = cgc_tensor(p0, i, j), cgc_tensor(p1, k, l)
        r = np.einsum("abc,byz,ayr,rzd->cd", uu, op1, op2, vv)
    return _ratio(r, np.eye(j + 1))


# ---- block operations --------------------------------------------------------------
def hvm(vec, op1, op2):
    """BlockTensor::hamiltonian_vector_multiplication (block_tensor_inst.cpp:2170-2318)."""
    if op1.nbond == 0 or op2.nbond == 0:
        return BT(vec.sym, vec.cplx, 0, [])
    out = BT(vec.sym, vec.cplx, 0, [op1.bonds[2], op2.bonds[2]])
    nt1, nt2, nt3 = op1.nten(), vec.nten(), op2.nten()
    for i in range(nt1 * nt2):
        j, k = i % nt1, (i // nt1) % nt2
        if not op1.allowed(j) or not vec.allowed(k):
            continue
        a1, a2 = op1.labels(j), vec.labels(k)
        if a1[0] != a2[0] or j not in op1.blocks or k not in vec.blocks:
            continue
        step1 = dcontract(vec.blocks[k], 0, op1.blocks[j], 0)
        for l in range(nt3):
            if not op2.allowed(l):
                continue
            a3 = op2.labels(l)
            if a3[0] != a2[1] or a1[1] != a3[1]:
                continue
            lab = [a1[2], a3[2]]
            if not out.allowed_labels(lab):
                continue
            m = out.index(lab)
            if l not in op2.blocks:
                continue
            tmp = np.einsum("xyc,xyd->cd", step1, op2.blocks[l])
            if vec.sym == 1:
                fac = fac_hamilt_vec(a1[0], a1[1], a1[2])
                if abs(fac) < 1e-12:
                    continue
                tmp = tmp * fac
            out.blocks[m] = out.blocks[m] + tmp if m in out.blocks else tmp
    return out


def contract(t1, i1, t2, i2):
    """BlockTensor::contract (block_tensor_inst.cpp:1046-1126); U(1) keeps the
    first shared-sector contribution only, as the reference does."""
    assert t1.bonds[i1] == inv(t2.bonds[i2]) or (t1.bonds[i1][1] == t2.bonds[i2][1] and
                                                t1.bonds[i1][2] == t2.bonds[i2][2] and
                                                t1.bonds[i1][0] == -t2.bonds[i2][0])
    n1, n2 = t1.nbond, t2.nbond
    bonds = [t1.bonds[(i1 + 1 + i) % n1] for i in range(n1 - 1)] + \
            [t2.bonds[(i2 + 1 + i) % n2] for i in range(n2 - 1)]
    out = BT(t1.sym, t1.cplx, t1.locspin + t2.locspin, bonds)
    sh = t1.bonds[i1]
    for i in range(out.nten()):
        if not out.allowed(i):
            continue
        lab = out.labels(i)
        l1, l2 = [0] * n1, [0] * n2
        for k in range(n1 - 1):
            l1[(i1 + 1 + k) % n1] = lab[k]
        for k in range(n2 - 1):
            l2[(i2 + 1 + k) % n2] = lab[k + n1 - 1]
        for q in sh[1]:
            l1[i1] = l2[i2] = q
            if not t1.allowed_labels(l1) or not t2.allowed_labels(l2):
                continue
            kb, lb = t1.index(l1), t2.index(l2)
            if kb not in t1.blocks or lb not in t2.blocks:
                continue
            r = dcontract(t1.blocks[kb], i1, t2.blocks[lb], i2)
            if i not in out.blocks:
                out.blocks[i] = r
            elif t1.sym == 1:
                out.blocks[i] = out.blocks[i] + r
    return out


def axpy(y, a, x):
    """y += a x with null + present = scaled copy (block_tensor_inst.cpp:225-278)."""
    if x.nbond == 0:
        return y
    if y is None or y.nbond == 0:
        z = x.copy()
        z.blocks = {k: a * v for k, v in z.blocks.items()}
        return z
    for k, v in x.blocks.items():
        y.blocks[k] = y.blocks[k] + a * v if k in y.blocks else a * v
    return y


def env(kind, uu, vv, op=None, sg=None, flag=0):
    """Environment updates (block_tensor_inst.cpp:1282-2167).
    kind: overlap_initial, overlap_transformation, operator_initial,
    operator_transformation, operator_pairup."""
    su2 = uu.sym == 1
    cj = np.conj if uu.cplx else (lambda x: x)
    if kind == "overlap_initial":          # :1661-1788
        out = BT(uu.sym, uu.cplx, 0, [uu.bonds[2], inv(vv.bonds[2])])
        for j in sorted(uu.blocks):
            if not uu.allowed(j):
                continue
            a1 = uu.labels(j)
            for k in sorted(vv.blocks):
                if not vv.allowed(k):
                    continue
                a2 = vv.labels(k)
                if a1[0] != a2[0] or a1[1] != a2[1]:
                    continue
                m = out.index([a1[2], a2[2]])
                if m is None or not out.allowed(m):
                    continue
                r = np.einsum("abc,abd->cd", uu.blocks[j], cj(vv.blocks[k]))
                out.blocks[m] = out.blocks[m] + r if m in out.blocks else r
        return out
    nt1, nt2 = op.nten(), uu.nten()
    if kind == "overlap_transformation":   # :1790-1951
        out = BT(uu.sym, uu.cplx, 0, [uu.bonds[2], inv(vv.bonds[2])])
        for i in range(nt1 * nt2):
            j, k = i % nt1, (i // nt1) % nt2
            if not op.allowed(j) or not uu.allowed(k) or j not in op.blocks or k not in uu.blocks:
                continue
            a1, a2 = op.labels(j), uu.labels(k)
            if a1[0] != a2[flag]:
                continue
            for l in range(vv.nten()):
                if not vv.allowed(l) or l not in vv.blocks:
                    continue
                a3 = vv.labels(l)
                if a1[1] != a3[flag] or a2[1 - flag] != a3[1 - flag]:
                    continue
                lab = [a2[2], a3[2]]
                if not out.allowed_labels(lab):
                    continue
                m = out.index(lab)
                if flag == 0:
                    r = np.einsum("ay,abc,ybd->cd", op.blocks[j], uu.blocks[k], cj(vv.blocks[l]))
                else:
                    r = np.einsum("by,abc,ayd->cd", op.blocks[j], uu.blocks[k], cj(vv.blocks[l]))
                out.blocks[m] = out.blocks[m] + r if m in out.blocks else r
        return out
    if kind in ("operator_initial", "operator_transformation"):
        init = kind == "operator_initial"
        ib = 1 - flag if init else flag
        out = BT(uu.sym, uu.cplx, 0, [uu.bonds[2], op.bonds[1], inv(vv.bonds[2])])
        for i in range(nt1 * nt2):
            j, k = i % nt1, (i // nt1) % nt2
            if not op.allowed(j) or not uu.allowed(k) or j not in op.blocks or k not in uu.blocks:
                continue
            a1, a2 = op.labels(j), uu.labels(k)
            if a1[0] != a2[ib]:
                continue
            for l in range(vv.nten()):
                if not vv.allowed(l) or l not in vv.blocks:
                    continue
                a3 = vv.labels(l)
                if init:   # :1392-1424
                    if a1[2] != a3[1 - flag] or a2[flag] != a3[flag]:
                        continue
                else:      # :1572-1606
                    if a1[2] != a3[flag] or a2[1 - flag] != a3[1 - flag]:
                        continue
                lab = [a2[2], a1[1], a3[2]]
                if not out.allowed_labels(lab):
                    continue
                m = out.index(lab)
                U, O, V = uu.blocks[k], op.blocks[j], cj(vv.blocks[l])
                if init:
                    assert O.size == 1
                    r = O.ravel()[0] * np.einsum("abc,abd->cd", U, V)[:, None, :]
                    r = np.broadcast_to(r, (U.shape[2], O.shape[1], V.shape[2])).copy()
                elif flag == 0:
                    r = np.einsum("abc,ayz,zbd->cyd", U, O, V)
                else:
                    r = np.einsum("abc,byz,azd->cyd", U, O, V)
                if su2:
                    if init:
                        fac = fac_onsite(flag, a2[2], a1[1], a3[2], a2[flag], a1[0] + a1[2] * 3)
                    else:
                        fac = fac_transformation(flag, a2[2], a1[1], a3[2], a2[flag], a3[flag], a2[1 - flag])
                    if abs(fac) < 1e-8:
                        continue
                    r = r * fac
                out.blocks[m] = out.blocks[m] + r if m in out.blocks else r
        return out
    if kind == "operator_pairup":          # :1953-2167
        out = BT(uu.sym, uu.cplx, 0, [uu.bonds[2], inv(vv.bonds[2])])
        for i in range(nt1 * nt2):
            j, k = i % nt1, (i // nt1) % nt2
            if not op.allowed(j) or not uu.allowed(k) or j not in op.blocks or k not in uu.blocks:
                continue
            a1, a2 = op.labels(j), uu.labels(k)
            if a1[0] != a2[flag]:
                continue
            for l in range(vv.nten()):
                if not vv.allowed(l) or l not in vv.blocks:
                    continue
                a3 = vv.labels(l)
                if a1[2] != a3[flag]:
                    continue
                for n in range(sg.nten()):
                    if not sg.allowed(n):
                        continue
                    a4 = sg.labels(n)
                    if a4[0] != a2[1 - flag] or a4[2] != a3[1 - flag] or a4[1] != a1[1] or n not in sg.blocks:
                        continue
                    lab = [a2[2], a3[2]]
                    if not out.allowed_labels(lab):
                        continue
                    m = out.index(lab)
                    U, O, V, S = uu.blocks[k], op.blocks[j], cj(vv.blocks[l]), sg.blocks[n]
                    if flag == 0:
                        r = np.einsum("abc,ayz,zed,bye->cd", U, O, V, S)
                    else:
                        r = np.einsum("abc,byz,ezd,aye->cd", U, O, V, S)
                    if su2:
                        fac = fac_pairup(flag, a2[2], a2[flag], a3[flag], a1[1], a4[0] + a4[2] * 3)
                        if abs(fac) < 1e-8:
                            continue
                        r = r * fac
                    out.blocks[m] = out.blocks[m] + r if m in out.blocks else r
        return out
    raise ValueError(kind)


def heff(hl, hr, terms, vec):
    """DMRG::hamiltonian_vector_multiplication_idmrg (dmrg_inst.cpp:283-371)
    given the block Hamiltonians and the (op1, op2) bond terms."""
    out = None
    if hl is not None and hl.nbond:
        out = contract(hl, 0, vec, 0)
    if hr is not None and hr.nbond:
        r = contract(vec, 1, hr, 0)
        out = r if out is None else axpy(out, 1.0, r)
    if out is None:
        out = BT(vec.sym, vec.cplx, 0, vec.bonds, {k: np.zeros_like(v) for k, v in vec.blocks.items()})
    for o1, o2 in terms:
        r = hvm(vec, o1, o2)
        if r.nbond:
            out = axpy(out, 1.0, r)
    return out


def heff_terms(state, hmap, coup, ns, il, ir):
    """Bond terms of dmrg_inst.cpp:310-351 from a state dict {'opr_i_j': BT, ...}."""
    def opr(i, j):
        return state.get(f"opr_{i}_{j}")
    terms = []
    left = il < ns // 2
    m1 = il + 1 if left else ns - ir
    m4 = ns - ir if left else il + 1
    for m0 in range(m1):
        op = None
        if left:
            j = m0
            for m in range(m4):
                k = ir + m
                if hmap[j][k] > 0 and opr(ir, k) is not None:
                    op = axpy(op, coup[hmap[j][k]], opr(ir, k))
            if op is None:
                continue
            terms.append((opr(il, j), op))
        else:
            k = ir + m0
            for m in range(m4):
                j = m
                if hmap[j][k] > 0 and opr(il, j) is not None:
                    op = axpy(op, coup[hmap[j][k]], opr(il, j))
            if op is None:
                continue
            terms.append((op, opr(ir, k)))
    return [(a, b) for a, b in terms if a is not None and b is not None]


def heff_from_state(state_packed, hmap, coup, ns, il, ir, vec_packed):
    """H_eff*psi from packed tensors; returns a packed tensor."""
    state = {k: from_packed(v) for k, v in state_packed.items()}
    vec = from_packed(vec_packed)
    hmap = np.asarray(hmap)
    out = heff(state.get(f"hh_{il}"), state.get(f"hh_{ir}"), heff_terms(state, hmap, coup, ns, il, ir), vec)
    return to_packed(out)


def lanczos1(apply, v0, mlanc=40):
    """Lanczos::lanczos1 (dmrg_inst.cpp:1286-1348) on flat vectors; returns
    (energy, Ritz vector, matvec count)."""
    nmv = 0

    def H(x):
        nonlocal nmv
        nmv += 1
        return apply(x)

    f = [v0 / np.linalg.norm(v0)]
    aal, nnl = np.zeros(mlanc), np.zeros(mlanc + 1)
    tol = 1e-6

    def tri(m):
        T = np.diag(aal[:m]) + np.diag(nnl[1:m], 1) + np.diag(nnl[1:m], -1)
        w, z = np.linalg.eigh(T)
        return w, z

    def ritz(m, z):
        return sum(z[j, 0] * f[j] for j in range(m))

    def check(v):
        v = v / np.linalg.norm(v)
        hv = H(v)
        return v, abs(np.vdot(v, hv)) / np.linalg.norm(hv)

    w1 = H(f[0])
    aal[0] = np.vdot(w1, f[0]).real
    w1 = w1 - aal[0] * f[0]
    nnl[1] = np.linalg.norm(w1)
    f.append(w1 / nnl[1])
    if mlanc == 2:
        if nnl[1] < tol:
            w, z = tri(1)
            return w[0], ritz(1, z), nmv
        w2 = H(f[1])
        aal[1] = np.vdot(w2, f[1]).real
        w, z = tri(2)
        return w[0], ritz(2, z), nmv
    for m in range(2, mlanc):
        wm = H(f[m - 1])
        aal[m - 1] = np.vdot(wm, f[m - 1]).real
        wm = wm - aal[m - 1] * f[m - 1] - nnl[m - 1] * f[m - 2]
        for j in range(m):
            wm = wm - np.vdot(f[j], wm) * f[j]
        nnl[m] = np.linalg.norm(wm)
        f.append(wm / nnl[m])
        w, z = tri(m)
        if nnl[m] < tol:
            v, _ = check(ritz(m, z))
            return w[0], v, nmv
        if m % 4 == 0:
            v, ov = check(ritz(m, z))
            if abs(ov - 1.0) < 1e-8:
                return w[0], v, nmv
    wl = H(f[mlanc - 1])
    aal[mlanc - 1] = np.vdot(wl, f[mlanc - 1]).real
    w, z = tri(mlanc)
    v, _ = check(ritz(mlanc, z))
    return w[0], v, nmv


def svd_values(vec, max_dcut):
    """Singular values of BlockTensor::svd (block_tensor_inst.cpp:501-602,
    dense_inst.cpp:447-541): kept (sector order) then discarded."""
    n0 = len(vec.bonds[0][1])
    per = {}
    for i, q0 in enumerate(vec.bonds[0][1]):
        for j, q1 in enumerate(vec.bonds[1][1]):
            b = i + j * n0
            if q0 == q1 and b in vec.blocks:
                s = np.linalg.svd(vec.blocks[b], compute_uv=False)
                smax = s[0] if s[0] > 1e-16 else 1.0
                kept = max(1, int(np.sum(s / smax > 1e-10)))
                per[i] = s[:min(kept, max_dcut)]
    allv = sorted([(v, i, k) for i, s in per.items() for k, v in enumerate(s)], key=lambda t: -t[0])
    dcut = min(len(allv), max_dcut)
    dim = {i: 0 for i in per}
    for v, i, k in allv[:dcut]:
        dim[i] += 1
    kept, disc = [], []
    for i in sorted(per):
        kept += list(per[i][:dim[i]])
        disc += list(per[i][dim[i]:])
    return np.asarray(kept + disc), dim


# ---- checkpoint files (dmrg_inst.cpp:947-1177) ---------------------------------------
def _nelement(t):
    """SU2Struct::get_nelement (su2.cpp:535-548) / U1Struct (u1.cpp:406-417):
    reduced and CGC sizes summed over every allowed block."""
    allowed = [b for b in range(t.nten()) if t.allowed(b)]
    ndata = sum(math.prod(t.dims(b)) for b in allowed)
    ncgc = sum(math.prod(j + 1 for j in t.labels(b)) for b in allowed) if t.sym == 1 else 0
    return allowed, ndata, ncgc


def mps_site_files(t):
    """save_mps2's two files for one site tensor (dmrg_inst.cpp:961-988):
    the .dat text and the .bin bytes.  The .bin holds ndata Scalars then ncgc
    doubles (SU(2)); present blocks come first in block order
    (BlockTensor::get_elems, block_tensor_inst.cpp:643-660), the rest of each
    section is zero (the buffers are value-initialised, dmrg_inst.cpp:982-983).
    CGC blocks of 3-leg tensors are the canonical fill_block tensors.  For
    U(1) the reference leaves ncgc uninitialised (U1Struct::get_nelement
    never sets it), so its CGC section has no defined length; here it is
    empty."""
    lines = [f"{t.nbond}\t{t.locspin}\n"]
    for d, q, n in t.bonds:
        lines.append(f"{d}\t{len(q)}\n")
        lines.append("".join(f"{x}\t" for x in q) + "\n")
        lines.append("".join(f"{x}\t" for x in n) + "\n")
    allowed, ndata, ncgc = _nelement(t)
    dt = np.complex128 if t.cplx else np.float64
    buf = np.zeros(ndata, dt)
    off = 0
    for b in allowed:
        if b in t.blocks:
            x = np.asarray(t.blocks[b], dt).ravel(order="F")
            buf[off:off + x.size] = x
            off += x.size
    out = buf.tobytes()
    if t.sym == 1:
        assert t.nbond == 3, "MPS site tensors have three bonds"
        cg = np.zeros(ncgc)
        off = 0
        for b in allowed:
            if b in t.blocks:
                x = cgc_tensor(*t.labels(b)).ravel(order="F")
                cg[off:off + x.size] = x
                off += x.size
        out += cg.tobytes()
    return "".join(lines), out


def read_mps_site(dat, binary, sym, cplx):
    """read_mps for one site (dmrg_inst.cpp:1046-1072) with
    BlockTensor::set(struct, data, cgc) (block_tensor_inst.cpp:120-159):
    every allowed block of the structure is filled, in block order, from
    consecutive elements of the file.  Returns the tensor and the CGC section
    (one flat array, SU(2); empty for U(1))."""
    tok = [int(x) for x in dat.split()]
    nb, ls = tok[0], tok[1]
    p, bonds = 2, []
    for _ in range(nb):
        dr, nm = tok[p], tok[p + 1]
        q, n = tok[p + 2:p + 2 + nm], tok[p + 2 + nm:p + 2 + 2 * nm]
        order = sorted(range(nm), key=lambda k: q[k])  # SU2Bond::set sorts by label (su2.cpp:125)
        bonds.append((dr, [q[k] for k in order], [n[k] for k in order]))
        p += 2 + 2 * nm
    t = BT(sym, cplx, ls, bonds)
    allowed, ndata, ncgc = _nelement(t)
    dt = np.complex128 if cplx else np.float64
    data = np.frombuffer(binary, dt, count=ndata)
    off = 0
    for b in allowed:
        dims = t.dims(b)
        n = math.prod(dims)
        t.blocks[b] = data[off:off + n].reshape(dims, order="F").copy()
        off += n
    cgc = np.frombuffer(binary, np.float64, count=ncgc, offset=ndata * data.itemsize) if sym == 1 else np.zeros(0)
    return t, cgc
