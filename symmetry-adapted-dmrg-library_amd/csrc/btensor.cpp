// btensor.cpp -- see btensor.hpp.
#include "btensor.hpp"
#include "eigen.hpp"
#include <atomic>
#include <chrono>
#include <functional>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <rocsolver/rocsolver.h>
#include <array>
#include <map>
#include <unordered_map>
#include "kernels.hpp"

namespace sadmrg {

namespace {


GemmPrec prec_of(bool cplx) { return cplx ? GemmPrec::Z64 : GemmPrec::F64; }

void labels_of(const BTensor& t, int b, int* lab) { t.st.labels(b, lab); }

// Column-major matrix view of one block as an operand.
Operand view(int arena, const BTensor& t, int b, int64_t ld, bool trans, bool conj = false) {
  Operand o;
  o.arena = arena;
  o.off = t.off[b];
  o.ld = ld;
  o.trans = trans;
  o.conj = conj;
  return o;
}

struct Arenas {
  ArenaTable at{};
  int64_t n[kMaxArenas];
  Arenas() {
    std::memset(&at, 0, sizeof(at));
    for (auto& x : n) x = -1;
  }
  // base pointer and element count (the count bounds every table access)
  void set(int i, const void* p, int64_t elems) {
    at.base[i] = p;
    n[i] = p ? std::max<int64_t>(elems, 0) : 0;
  }
};

void run_batch(GemmBatch& b, Context* ctx, bool cplx, const Arenas& a) {
  if (b.empty()) return;
  for (int i = 0; i < kMaxArenas; ++i)
    if (a.n[i] >= 0) b.limit(i, a.n[i]);
  b.finalize(ctx, cplx);
  b.run(ctx, prec_of(cplx), a.at);
}

}  // namespace

// ---------------------------------------------------------------------------
void BTensor::layout(Context* c, bool complex_, const Struct& s, const std::vector<uint8_t>& mask, uint8_t tag0) {
  ctx = c;
  cplx = complex_;
  st = s;
  const int n = s.nten();
  off.assign(n, -1);
  tag.assign(n, tag0);
  nelem = 0;
  fuse_id = false;
  host.clear();
  for (int b = 0; b < n; ++b)
    if (mask[b]) {
      off[b] = nelem;
      nelem += s.block_size(b);
    }
  buf = DevBuf(c, std::max<int64_t>(nelem, 1) * sizeof(double) * w());
  SADMRG_HIP(launch_fill(std::max<int64_t>(nelem, 1) * w(), 0.0, dptr(), c->stream()));
}

void BTensor::clean() {
  st = Struct{};
  off.clear();
  tag.clear();
  buf.release();
  nelem = 0;
  host.clear();
}

void BTensor::keep_host() {
  host.assign(static_cast<size_t>(nelem) * w(), 0.0);
  if (nelem > 0) {
    SADMRG_HIP(hipMemcpyAsync(host.data(), dptr(), host.size() * sizeof(double), hipMemcpyDeviceToHost,
                              ctx->stream()));
    ctx->sync();
  }
}

const double* BTensor::host_values(std::vector<double>& tmp) const {
  if (!host.empty() && host.size() == static_cast<size_t>(nelem) * w()) return host.data();
  tmp.assign(static_cast<size_t>(nelem) * w(), 0.0);
  if (nelem > 0) {
    SADMRG_HIP(hipMemcpyAsync(tmp.data(), dptr(), tmp.size() * sizeof(double), hipMemcpyDeviceToHost,
                              ctx->stream()));
    ctx->sync();
  }
  return tmp.data();
}

Struct parse_header(SymKind sym, const int* h, int len) {
  require(len >= 2, "tensor header too short");
  Struct s;
  s.sym = sym;
  const int nb = h[0];
  s.locspin = h[1];
  require(nb >= 0 && nb <= kMaxLeg, "tensor header: bad nbond");
  int p = 2;
  for (int i = 0; i < nb; ++i) {
    require(p + 2 <= len, "tensor header truncated");
    const int dir = h[p], nm = h[p + 1];
    require(nm >= 0 && p + 2 + 2 * nm <= len, "tensor header truncated");
    std::vector<int> q(h + p + 2, h + p + 2 + nm), d(h + p + 2 + nm, h + p + 2 + 2 * nm);
    for (int x : d) require(x >= 0, "tensor header: negative dimension");
    s.bonds.push_back(Bond::make(dir, q, d));
    p += 2 + 2 * nm;
  }
  return s;
}

std::vector<int> make_header(const Struct& s) {
  std::vector<int> h{s.nbond(), s.locspin};
  for (const auto& b : s.bonds) {
    h.push_back(b.dir);
    h.push_back(b.nmom());
    h.insert(h.end(), b.q.begin(), b.q.end());
    h.insert(h.end(), b.dim.begin(), b.dim.end());
  }
  return h;
}

void tensor_from_host(BTensor& t, Context* ctx, SymKind sym, bool cplx, const int* header, int len,
                      const unsigned char* present, const double* data, int cgc2_kind) {
  Struct s = parse_header(sym, header, len);
  std::vector<uint8_t> mask(s.nten(), 0);
  for (int b = 0; b < s.nten(); ++b) mask[b] = (present[b] && s.allowed(b)) ? 1 : 0;
  const uint8_t tg = s.nbond() == 3 ? kCgcCanon : (cgc2_kind == 1 ? kCgcSinglet : kCgcIdent);
  t.clean();
  if (s.nbond() == 0) return;
  t.layout(ctx, cplx, s, mask, tg);
  if (t.nelem > 0) {
    // input packs present blocks in block order: identical to the layout
    SADMRG_HIP(hipMemcpyAsync(t.dptr(), data, t.nelem * sizeof(double) * t.w(), hipMemcpyHostToDevice,
                              ctx->stream()));
    ctx->sync();
  }
}

void tensor_to_host(const BTensor& t, unsigned char* present, double* data) {
  for (int b = 0; b < t.nten(); ++b) present[b] = t.present(b) ? 1 : 0;
  if (t.nelem > 0 && data) {
    SADMRG_HIP(hipMemcpyAsync(data, t.dptr(), t.nelem * sizeof(double) * t.w(), hipMemcpyDeviceToHost,
                              t.ctx->stream()));
    t.ctx->sync();
  }
}

void copy_tensor(BTensor& dst, const BTensor& src) {
  if (&dst == &src) return;
  dst.clean();
  if (src.is_null()) return;
  std::vector<uint8_t> mask(src.nten());
  for (int b = 0; b < src.nten(); ++b) mask[b] = src.present(b);
  dst.layout(src.ctx, src.cplx, src.st, mask, kCgcIdent);
  dst.tag = src.tag;
  dst.fuse_id = src.fuse_id;
  dst.host = src.host;
  if (src.nelem > 0)
    SADMRG_HIP(hipMemcpyAsync(dst.dptr(), src.dptr(), src.nelem * sizeof(double) * src.w(),
                              hipMemcpyDeviceToDevice, src.ctx->stream()));
}

// ---------------------------------------------------------------------------
// hamiltonian_vector_multiplication: out[a', b'] += fac * O1^T psi O2 per
// (op1 block j, psi block k, op2 block l) triple; stage 1 = psi^T O1
// (block_tensor_inst.cpp:2258), stage 2 = step1^T O2 (:2267), fac from
// fac_hamilt_vec[J_in][chan][J_out] with |fac| < 1e-12 skipped (:2268-2272).
void op_hvm(BTensor& out, const BTensor& vec, const BTensor& op1, const BTensor& op2, CgcTables* cgc) {
  out.clean();
  if (op1.is_null() || op2.is_null()) return;
  const SymKind sym = vec.sym();
  Context* ctx = vec.ctx;
  require(vec.nbond() == 2 && op1.nbond() == 3 && op2.nbond() == 3, "hamiltonian_vector_multiplication: ranks");
  const Bond ch2 = sym == SymKind::SU2 ? op2.st.bonds[1] : op2.st.bonds[1].inverted();
  require(op1.st.bonds[0] == vec.st.bonds[0].inverted() && op2.st.bonds[0] == vec.st.bonds[1].inverted() &&
              op1.st.bonds[1] == ch2,
          "hamiltonian_vector_multiplication bb wrong");
  Struct os;
  os.sym = sym;
  os.locspin = 0;
  os.bonds = {op1.st.bonds[2], op2.st.bonds[2]};
  const int nt1 = op1.nten(), nt2 = vec.nten(), nt3 = op2.nten(), nto = os.nten();
  struct S1 { int j, k; int64_t off; int M, N, K; };
  struct Con { int s1, l; double fac; };
  std::vector<S1> s1;
  std::vector<std::vector<Con>> per(nto);
  int64_t scratch = 0;
  int a1[3], a2[2], a3[3], d1[3], d2[2], d3[3];
  for (int i = 0; i < nt1 * nt2; ++i) {
    const int j = i % nt1, k = (i / nt1) % nt2;
    if (!op1.st.allowed(j) || !vec.st.allowed(k)) continue;
    labels_of(op1, j, a1);
    labels_of(vec, k, a2);
    if (a1[0] != a2[0]) continue;
    if (!op1.present(j) || !vec.present(k)) continue;
    op1.st.block_dims(j, d1);
    vec.st.block_dims(k, d2);
    int si = -1;
    for (int l = 0; l < nt3; ++l) {
      if (!op2.st.allowed(l)) continue;
      labels_of(op2, l, a3);
      if (a3[0] != a2[1] || a1[1] != a3[1]) continue;
      const int lab[2] = {a1[2], a3[2]};
      if (!os.allowed_labels(lab)) continue;
      const int m = os.index_of_labels(lab);
      if (m < 0) continue;
      if (!op2.present(l)) continue;
      double fac = 1.0;
      if (sym == SymKind::SU2) {
        fac = cgc->hamilt_vec(a1[0], a1[1], a1[2]);
        if (std::fabs(fac) < 1e-12) continue;
      }
      if (si < 0) {
        si = static_cast<int>(s1.size());
        s1.push_back(S1{j, k, scratch, d2[1], d1[1] * d1[2], d2[0]});
        scratch += int64_t(d2[1]) * d1[1] * d1[2];
      }
      per[m].push_back(Con{si, l, fac});
    }
  }
  std::vector<uint8_t> mask(nto, 0);
  for (int m = 0; m < nto; ++m) mask[m] = per[m].empty() ? 0 : 1;
  out.layout(ctx, vec.cplx, os, mask, kCgcSinglet);
  if (s1.empty()) return;
  DevBuf sbuf(ctx, std::max<int64_t>(scratch, 1) * sizeof(double) * vec.w());
  Arenas ar;
  ar.set(0, vec.dptr(), vec.nelem);
  ar.set(1, op1.dptr(), op1.nelem);
  ar.set(2, op2.dptr(), op2.nelem);
  ar.set(3, sbuf.get(), scratch);
  ar.set(4, out.dptr(), out.nelem);
  GemmBatch g1, g2;
  for (const auto& s : s1) {
    const int p = g1.add_problem(3, s.off, s.M, s.M, s.N, false);
    g1.add_seg(p, view(0, vec, s.k, s.K, true), view(1, op1, s.j, s.K, false), s.K, 1.0);
  }
  for (int m = 0; m < nto; ++m) {
    if (per[m].empty()) continue;
    int dm[2];
    os.block_dims(m, dm);
    const int p = g2.add_problem(4, out.off[m], dm[0], dm[0], dm[1], false);
    for (const auto& c : per[m]) {
      const S1& s = s1[c.s1];
      op2.st.block_dims(c.l, d3);
      const int kk = d3[0] * d3[1];
      require(kk == s.M * (s.N / dm[0]), "hvm: inner dimension mismatch");
      Operand A;
      A.arena = 3;
      A.off = s.off;
      A.ld = kk;
      A.trans = true;
      g2.add_seg(p, A, view(2, op2, c.l, kk, false), kk, c.fac);
    }
  }
  run_batch(g1, ctx, vec.cplx, ar);
  run_batch(g2, ctx, vec.cplx, ar);
}

// ---------------------------------------------------------------------------
// Generic contraction of bond i1 of t1 with bond i2 of t2.  Output bonds are
// t1's remaining bonds cyclically after i1, then t2's after i2; each block is
// one GEMM per shared sector (Dense::contract, dense_inst.cpp:369-408).  For
// U(1) only the first matching shared sector contributes, as in the reference
// (block_tensor_inst.cpp:1099-1117 has no U(1) accumulation branch).
void op_contract(BTensor& out, const BTensor& t1, int i1, const BTensor& t2, int i2) {
  Context* ctx = t1.ctx;
  const SymKind sym = t1.sym();
  const int n1 = t1.nbond(), n2 = t2.nbond();
  require(i1 >= 0 && i1 < n1 && i2 >= 0 && i2 < n2, "contract: bond index out of range");
  require(t1.st.bonds[i1] == t2.st.bonds[i2].inverted(),
          "BlockTensor::contract, two bonds can not be contracted, check bond parameters");
  Struct os;
  os.sym = sym;
  os.locspin = t1.st.locspin + t2.st.locspin;
  for (int i = 0; i < n1 - 1; ++i) os.bonds.push_back(t1.st.bonds[(i1 + 1 + i) % n1]);
  for (int i = 0; i < n2 - 1; ++i) os.bonds.push_back(t2.st.bonds[(i2 + 1 + i) % n2]);
  const int nto = os.nten();
  const Bond& sh = t1.st.bonds[i1];
  std::vector<std::vector<std::pair<int, int>>> per(nto);
  int lab[2 * kMaxLeg], l1[kMaxLeg], l2[kMaxLeg];
  for (int i = 0; i < nto; ++i) {
    if (!os.allowed(i)) continue;
    os.labels(i, lab);
    for (int k = 0; k < n1 - 1; ++k) l1[(i1 + 1 + k) % n1] = lab[k];
    for (int k = 0; k < n2 - 1; ++k) l2[(i2 + 1 + k) % n2] = lab[k + n1 - 1];
    for (int j = 0; j < sh.nmom(); ++j) {
      l1[i1] = sh.q[j];
      l2[i2] = sh.q[j];
      if (!t1.st.allowed_labels(l1) || !t2.st.allowed_labels(l2)) continue;
      const int kb = t1.st.index_of_labels(l1), lb = t2.st.index_of_labels(l2);
      if (kb < 0 || lb < 0) continue;
      if (!t1.present(kb) || !t2.present(lb)) continue;
      if (sym == SymKind::U1 && !per[i].empty()) continue;
      per[i].push_back({kb, lb});
    }
  }
  std::vector<uint8_t> mask(nto, 0);
  for (int i = 0; i < nto; ++i) mask[i] = per[i].empty() ? 0 : 1;
  uint8_t tg = kCgcOther;
  if (os.nbond() == 3) tg = kCgcCanon;
  else if (n1 == 2 && n2 == 2) {
    const uint8_t ta = t1.tag.empty() ? kCgcIdent : t1.tag[0], tb = t2.tag.empty() ? kCgcIdent : t2.tag[0];
    tg = ta == kCgcIdent ? tb : (tb == kCgcIdent ? ta : kCgcOther);
  }
  BTensor res;
  res.layout(ctx, t1.cplx, os, mask, tg);

  // Interior contracted legs need a transposed copy (Dense::shift).
  auto prod = [](const int* d, int a, int b) {
    int64_t p = 1;
    for (int i = a; i < b; ++i) p *= d[i];
    return p;
  };
  int64_t scratch = 0;
  struct Item { int out, kb, lb; int64_t aoff, boff; };
  std::vector<Item> items;
  for (int i = 0; i < nto; ++i)
    for (auto& pr : per[i]) {
      Item it{i, pr.first, pr.second, -1, -1};
      int d[kMaxLeg];
      if (i1 != 0 && i1 != n1 - 1) { it.aoff = scratch; scratch += t1.block_size(pr.first); }
      if (i2 != 0 && i2 != n2 - 1) { it.boff = scratch; scratch += t2.block_size(pr.second); }
      (void)d;
      items.push_back(it);
    }
  DevBuf sbuf(ctx, std::max<int64_t>(scratch, 1) * sizeof(double) * t1.w());
  for (const auto& it : items) {
    int d[kMaxLeg];
    if (it.aoff >= 0) {
      t1.st.block_dims(it.kb, d);
      const int64_t B = prod(d, 0, i1 + 1), At = prod(d, i1 + 1, n1);
      SADMRG_HIP(launch_copy2d(static_cast<int>(At), static_cast<int>(B), t1.block_ptr(it.kb), B, 1,
                               sbuf.as<double>() + it.aoff * t1.w(), At, t1.w(), 1.0, nullptr, 0, ctx->stream()));
    }
    if (it.boff >= 0) {
      t2.st.block_dims(it.lb, d);
      const int64_t B = prod(d, 0, i2), At = prod(d, i2, n2);
      SADMRG_HIP(launch_copy2d(static_cast<int>(At), static_cast<int>(B), t2.block_ptr(it.lb), B, 1,
                               sbuf.as<double>() + it.boff * t1.w(), At, t1.w(), 1.0, nullptr, 0, ctx->stream()));
    }
  }
  Arenas ar;
  ar.set(0, t1.dptr(), t1.nelem);
  ar.set(1, t2.dptr(), t2.nelem);
  ar.set(2, sbuf.get(), scratch);
  ar.set(3, res.dptr(), res.nelem);
  GemmBatch g;
  for (int i = 0; i < nto; ++i) {
    if (per[i].empty()) continue;
    int dout[2 * kMaxLeg];
    os.block_dims(i, dout);
    int64_t M = 1, N = 1;
    for (int k = 0; k < n1 - 1; ++k) M *= dout[k];
    for (int k = 0; k < n2 - 1; ++k) N *= dout[n1 - 1 + k];
    const int p = g.add_problem(3, res.off[i], M, static_cast<int>(M), static_cast<int>(N), false);
    for (const auto& it : items) {
      if (it.out != i) continue;
      int da[kMaxLeg], db[kMaxLeg];
      t1.st.block_dims(it.kb, da);
      t2.st.block_dims(it.lb, db);
      const int K = da[i1];
      Operand A, B;
      if (i1 == 0) { A = view(0, t1, it.kb, K, true); }
      else if (i1 == n1 - 1) { A = view(0, t1, it.kb, M, false); }
      else { A.arena = 2; A.off = it.aoff; A.ld = M; A.trans = false; }
      if (i2 == 0) { B = view(1, t2, it.lb, K, false); }
      else if (i2 == n2 - 1) { B = view(1, t2, it.lb, N, true); }
      else { B.arena = 2; B.off = it.boff; B.ld = K; B.trans = false; }
      g.add_seg(p, A, B, K, 1.0);
    }
  }
  run_batch(g, ctx, t1.cplx, ar);
  out = std::move(res);
}

// ---------------------------------------------------------------------------
// Environment updates.  The reduced physical leg (spin-1/2: dimension 1 per
// sector) lets every block chain reduce to plain GEMMs; a larger physical or
// channel dimension is rejected as unsupported.
namespace {

int dims3(const BTensor& t, int b, int k) {
  int d[kMaxLeg];
  t.st.block_dims(b, d);
  return d[k];
}

// Row offset of each block of an identity fusion inside its fused sector
// (the pos[] bookkeeping of BlockTensor::fuse, block_tensor_inst.cpp:722-738).
std::vector<int64_t> fuse_positions(const BTensor& f) {
  std::vector<int64_t> pos(f.nten(), -1), acc(f.st.bonds[2].nmom(), 0);
  int idx[3], d[3];
  for (int b = 0; b < f.nten(); ++b) {
    if (!f.st.allowed(b)) continue;
    f.st.decode(b, idx);
    f.st.block_dims(b, d);
    pos[b] = acc[idx[2]];
    acc[idx[2]] += int64_t(d[0]) * d[1];
  }
  return pos;
}

// Scatter table builder with a host-side bounds check of every task.
struct ScatterBuilder {
  std::vector<ScatterTask> tasks;
  void copy(int64_t src_off, int64_t lds, int64_t dst_off, int64_t ldd, int rows, int cols, double scale) {
    // about 16 K elements per task (one workgroup each), at most 64 columns:
    // enough workgroups to spread a few large blocks over every CU
    const int cw = std::max(1, std::min(64, 16384 / std::max(rows, 1)));
    for (int c0 = 0; c0 < cols; c0 += cw) {
      ScatterTask t{};
      t.src_off = src_off + int64_t(c0) * lds;
      t.dst_off = dst_off + int64_t(c0) * ldd;
      t.lds = lds;
      t.ldd = ldd;
      t.rows = rows;
      t.cols = std::min(cw, cols - c0);
      t.kind = 0;
      t.scale = scale;
      tasks.push_back(t);
    }
  }
  void stripe(int64_t dst_off, int64_t ldd, int n, double scale) {
    ScatterTask t{};
    t.dst_off = dst_off;
    t.ldd = ldd;
    t.rows = n;
    t.cols = 1;
    t.kind = 1;
    t.scale = scale;
    tasks.push_back(t);
  }
  void run(Context* ctx, const double* src, int64_t src_n, double* dst, int64_t dst_n, int w) {
    run_many(ctx, {{src, dst}}, src_n, dst_n, w);
  }
  // The same table on several (source, destination) pairs of equal sizes.
  void run_many(Context* ctx, const std::vector<std::pair<const double*, double*>>& pairs, int64_t src_n,
                int64_t dst_n, int w) {
    if (tasks.empty() || pairs.empty()) return;
    for (const auto& t : tasks) {
      if (t.rows <= 0 || t.cols <= 0) continue;
      if (t.kind == 0) {
        require(t.src_off >= 0 && t.src_off + (t.rows - 1) + int64_t(t.cols - 1) * t.lds < src_n,
                "scatter table out of bounds (source)");
        require(t.dst_off >= 0 && t.dst_off + (t.rows - 1) + int64_t(t.cols - 1) * t.ldd < dst_n,
                "scatter table out of bounds (destination)");
      } else {
        require(t.dst_off >= 0 && t.dst_off + int64_t(t.rows - 1) * (t.ldd + 1) < dst_n,
                "scatter table out of bounds (stripe)");
      }
    }
    DevBuf d(ctx, tasks.size() * sizeof(ScatterTask));
    ctx->upload(d.get(), tasks.data(), tasks.size() * sizeof(ScatterTask));
    for (const auto& pr : pairs)
      SADMRG_HIP(launch_block_scatter(d.as<ScatterTask>(), static_cast<int>(tasks.size()), pr.first, pr.second, w,
                                      ctx->stream()));
  }
};

// Further (output, operator) pairs executed with the table of the first.
typedef std::vector<std::pair<BTensor*, const BTensor*>> EnvExtra;

void op_env_impl(BTensor& out, int kind, const BTensor& uu, const BTensor& vv, const BTensor* op1,
                 const BTensor* op2, int flag, CgcTables* cgc, const EnvExtra* extra) {
  Context* ctx = uu.ctx;
  const SymKind sym = uu.sym();
  const bool su2 = sym == SymKind::SU2;
  const bool cx = uu.cplx;
  require(flag == 0 || flag == 1, "environment update: flag must be 0 or 1");
  require(uu.nbond() == 3 && vv.nbond() == 3, "environment update: uu/vv must be 3-leg");
  Struct os;
  os.sym = sym;
  os.locspin = 0;
  int a1[3], a2[3], a3[3], a4[3];
  Arenas ar;
  ar.set(0, uu.dptr(), uu.nelem);
  ar.set(1, vv.dptr(), vv.nelem);
  ar.set(2, op1 ? op1->dptr() : nullptr, op1 ? op1->nelem : 0);
  ar.set(3, op2 ? op2->dptr() : nullptr, op2 ? op2->nelem : 0);
  GemmBatch g1, g2;
  int64_t scratch = 0;
  std::vector<std::vector<std::pair<int, int>>> dummy;

  if (kind == kOverlapInitial) {
    require(uu.st.bonds[flag] == vv.st.bonds[flag] && uu.st.bonds[1 - flag] == vv.st.bonds[1 - flag],
            "overlap initial wrong bb");
    os.bonds = {uu.st.bonds[2], vv.st.bonds[2].inverted()};
    const int nto = os.nten();
    struct C { int j, k; };
    std::vector<std::vector<C>> per(nto);
    const BlockList ul(uu), vl(vv);
    const LabelIndex vix(vl, 0, 1);
    for (int pj = 0; pj < ul.size(); ++pj) {
      const int j = ul.b[pj];
      for (int q = 0; q < 3; ++q) a1[q] = ul.lab[pj][q];
      for (const int pk : vix.get(a1[0], a1[1])) {
        const int k = vl.b[pk];
        for (int q = 0; q < 3; ++q) a2[q] = vl.lab[pk][q];
        const int lab[2] = {a1[2], a2[2]};
        const int m = os.index_of_labels(lab);
        // blocks outside the symmetry-allowed set are never read back
        if (m < 0 || !os.allowed(m)) continue;
        per[m].push_back(C{j, k});
      }
    }
    std::vector<uint8_t> mask(nto);
    for (int m = 0; m < nto; ++m) mask[m] = !per[m].empty();
    out.clean();
    out.layout(ctx, cx, os, mask, kCgcIdent);
    ar.set(5, out.dptr(), out.nelem);
    for (int m = 0; m < nto; ++m) {
      if (per[m].empty()) continue;
      int dm[2];
      os.block_dims(m, dm);
      const int p = g2.add_problem(5, out.off[m], dm[0], dm[0], dm[1], false);
      for (auto& c : per[m]) {
        const int K = dims3(uu, c.j, 0) * dims3(uu, c.j, 1);
        g2.add_seg(p, view(0, uu, c.j, K, true), view(1, vv, c.k, K, false, true), K, 1.0);
      }
    }
    run_batch(g2, ctx, cx, ar);
    return;
  }

  require(op1 != nullptr, "environment update: missing operator");
  const BTensor& op = *op1;
  // Stage-1 product of one (op block j, uu block k) pair, reused by every vv
  // block that matches it.
  struct S1 { int j, k; int64_t off; int M, N, K; };
  std::vector<S1> s1;
  struct Con { int s1, l; double scale; int K; };
  std::vector<std::vector<Con>> per;

  // Enlarged basis (uu = vv = an identity fusion, dmrg_inst.cpp:595,651):
  // every block chain is a scaled copy of one operator block, or a scaled
  // identity stripe, placed at the embedded position.  The GEMM lowering
  // gives the same numbers (products with exact 0/1 entries); the scatter
  // moves only the operator bytes.
  bool scatter = uu.fuse_id && vv.fuse_id && uu.st == vv.st;
  if (scatter && op.nbond() == 3)
    for (int j = 0; j < op.nten(); ++j)
      if (op.present(j) && dims3(op, j, 1) != 1) scatter = false;
  std::vector<int64_t> fpos;
  if (scatter) fpos = fuse_positions(uu);
  auto scatter_run = [&](bool stripe) {
    ScatterBuilder sb;
    for (int m = 0; m < os.nten(); ++m) {
      if (per[m].empty()) continue;
      int dm[3];
      os.block_dims(m, dm);
      const int64_t ldd = os.nbond() == 3 ? int64_t(dm[0]) * dm[1] : dm[0];
      for (const auto& c : per[m]) {
        const int k = stripe ? c.s1 : s1[c.s1].k;
        const int64_t dst = out.off[m] + fpos[k] + ldd * fpos[c.l];
        if (stripe) {
          sb.stripe(dst, ldd, c.K, c.scale);
          continue;
        }
        const int j = s1[c.s1].j;
        int d[3];
        op.st.block_dims(j, d);
        const int rows = d[0], cols = op.nbond() == 3 ? d[2] : d[1];
        sb.copy(op.off[j], rows, dst, ldd, rows, cols, c.scale);
      }
    }
    std::vector<std::pair<const double*, double*>> pairs{{stripe ? nullptr : op.dptr(), out.dptr()}};
    if (extra && !stripe)
      for (const auto& e : *extra) pairs.emplace_back(e.second->dptr(), e.first->dptr());
    sb.run_many(ctx, pairs, stripe ? 0 : op.nelem, out.nelem, uu.w());
  };

  auto stage1_uu_op = [&](int j, int k) -> S1 {
    // step1(c, y, z) = sum_a uu(a@flag, c) op(a, y, z): uu is (l x r) once
    // the unit physical leg is dropped.
    const int l = dims3(uu, k, flag), r = dims3(uu, k, 2);
    require(dims3(uu, k, 1 - flag) == 1, "environment update: physical leg dimension > 1 is unsupported");
    int d[3];
    op.st.block_dims(j, d);
    require(d[0] == l, "environment update: operator / site tensor dimension mismatch");
    S1 s{j, k, scratch, r, d[1] * d[2], l};
    scratch += int64_t(r) * d[1] * d[2];
    return s;
  };

  if (kind == kOverlapTransformation) {
    require(uu.st.bonds[flag] == op.st.bonds[0].inverted() && vv.st.bonds[flag] == op.st.bonds[1],
            "overlap transformation wrong bb");
    os.bonds = {uu.st.bonds[2], vv.st.bonds[2].inverted()};
    per.resize(os.nten());
    const BlockList ol(op), ul(uu), vl(vv);
    const LabelIndex oix(ol, 0, -1), vix(vl, flag, 1 - flag);
    for (int pk = 0; pk < ul.size(); ++pk) {
      const int k = ul.b[pk];
      for (int q = 0; q < 3; ++q) a2[q] = ul.lab[pk][q];
      for (const int pj : oix.get(a2[flag])) {
      const int j = ol.b[pj];
      for (int q = 0; q < 3; ++q) a1[q] = ol.lab[pj][q];
      int si = -1;
      for (const int pl : vix.get(a1[1], a2[1 - flag])) {
        const int l = vl.b[pl];
        for (int q = 0; q < 3; ++q) a3[q] = vl.lab[pl][q];
        const int lab[2] = {a2[2], a3[2]};
        if (!os.allowed_labels(lab)) continue;
        const int m = os.index_of_labels(lab);
        if (si < 0) {
          // step1(y, c) = op^T (y x l) uu (l x r)
          const int l0 = dims3(uu, k, flag), r = dims3(uu, k, 2);
          require(dims3(uu, k, 1 - flag) == 1, "overlap transformation: physical leg dimension > 1 is unsupported");
          const int y = dims3(op, j, 1);
          si = static_cast<int>(s1.size());
          s1.push_back(S1{j, k, scratch, y, r, l0});
          scratch += int64_t(y) * r;
        }
        per[m].push_back(Con{si, l, 1.0, s1[si].M});
      }
      }
    }
    std::vector<uint8_t> mask(os.nten());
    for (int m = 0; m < os.nten(); ++m) mask[m] = !per[m].empty();
    out.clean();
    out.layout(ctx, cx, os, mask, kCgcIdent);
    if (scatter) {
      scatter_run(false);
      return;
    }
    DevBuf sbuf(ctx, std::max<int64_t>(scratch, 1) * sizeof(double) * uu.w());
    ar.set(4, sbuf.get(), scratch);
    ar.set(5, out.dptr(), out.nelem);
    for (const auto& s : s1) {
      const int p = g1.add_problem(4, s.off, s.M, s.M, s.N, false);
      g1.add_seg(p, view(2, op, s.j, s.K, true), view(0, uu, s.k, s.K, false), s.K, 1.0);
    }
    for (int m = 0; m < os.nten(); ++m) {
      if (per[m].empty()) continue;
      int dm[2];
      os.block_dims(m, dm);
      const int p = g2.add_problem(5, out.off[m], dm[0], dm[0], dm[1], false);
      for (const auto& c : per[m]) {
        const S1& s = s1[c.s1];
        require(dims3(vv, c.l, 1 - flag) == 1, "overlap transformation: physical leg dimension > 1");
        Operand A;
        A.arena = 4;
        A.off = s.off;
        A.ld = s.M;
        A.trans = true;
        g2.add_seg(p, A, view(1, vv, c.l, s.M, false, true), s.M, 1.0);
      }
    }
    run_batch(g1, ctx, cx, ar);
    run_batch(g2, ctx, cx, ar);
    return;
  }

  if (kind == kOperatorInitial) {
    require(uu.st.bonds[1 - flag] == op.st.bonds[0].inverted() && vv.st.bonds[1 - flag] == op.st.bonds[2],
            "operator initial wrong bb");
    os.bonds = {uu.st.bonds[2], op.st.bonds[1], vv.st.bonds[2].inverted()};
    per.resize(os.nten());
    // operator values are needed on the host as scalar weights
    std::vector<double> hop_tmp;
    const double* hop = op.host_values(hop_tmp);
    const int P2 = cgc ? cgc->p2() : 3;
    const BlockList ol(op), ul(uu), vl(vv);
    const LabelIndex oix(ol, 0, -1), vix(vl, 1 - flag, flag);
    for (int pk = 0; pk < ul.size(); ++pk) {
      const int k = ul.b[pk];
      for (int q = 0; q < 3; ++q) a2[q] = ul.lab[pk][q];
      for (const int pj : oix.get(a2[1 - flag])) {
      const int j = ol.b[pj];
      for (int q = 0; q < 3; ++q) a1[q] = ol.lab[pj][q];
      require(op.block_size(j) == 1, "operator initial: multi-element site operators are unsupported");
      for (const int pl : vix.get(a1[2], a2[flag])) {
        const int l = vl.b[pl];
        for (int q = 0; q < 3; ++q) a3[q] = vl.lab[pl][q];
        const int lab[3] = {a2[2], a1[1], a3[2]};
        if (!os.allowed_labels(lab)) continue;
        const int m = os.index_of_labels(lab);
        double fac = 1.0;
        if (su2) {
          fac = cgc->onsite(flag, a2[2], a1[1], a3[2], a2[flag], a1[0] + a1[2] * P2);
          if (std::fabs(fac) < 1e-8) continue;
        }
        require(!cx || hop[op.off[j] * 2 + 1] == 0.0, "operator initial: complex site operator unsupported");
        const double sv = hop[op.off[j] * op.w()];
        const int K = dims3(uu, k, 0) * dims3(uu, k, 1);
        per[m].push_back(Con{k, l, fac * sv, K});
      }
      }
    }
    std::vector<uint8_t> mask(os.nten());
    for (int m = 0; m < os.nten(); ++m) mask[m] = !per[m].empty();
    out.clean();
    out.layout(ctx, cx, os, mask, kCgcCanon);
    if (scatter) {
      scatter_run(true);
      return;
    }
    ar.set(5, out.dptr(), out.nelem);
    for (int m = 0; m < os.nten(); ++m) {
      if (per[m].empty()) continue;
      int dm[3];
      os.block_dims(m, dm);
      const int p = g2.add_problem(5, out.off[m], int64_t(dm[0]) * dm[1], dm[0] * dm[1], dm[2], false);
      for (const auto& c : per[m])
        g2.add_seg(p, view(0, uu, c.s1, c.K, true), view(1, vv, c.l, c.K, false, true), c.K, c.scale);
    }
    run_batch(g2, ctx, cx, ar);
    return;
  }

  if (kind == kOperatorTransformation) {
    require(uu.st.bonds[flag] == op.st.bonds[0].inverted() && vv.st.bonds[flag] == op.st.bonds[2],
            "operator transformation wrong bb");
    os.bonds = {uu.st.bonds[2], op.st.bonds[1], vv.st.bonds[2].inverted()};
    per.resize(os.nten());
    const BlockList ol(op), ul(uu), vl(vv);
    const LabelIndex oix(ol, 0, -1), vix(vl, flag, 1 - flag);
    for (int pk = 0; pk < ul.size(); ++pk) {
      const int k = ul.b[pk];
      for (int q = 0; q < 3; ++q) a2[q] = ul.lab[pk][q];
      for (const int pj : oix.get(a2[flag])) {
      const int j = ol.b[pj];
      for (int q = 0; q < 3; ++q) a1[q] = ol.lab[pj][q];
      int si = -1;
      for (const int pl : vix.get(a1[2], a2[1 - flag])) {
        const int l = vl.b[pl];
        for (int q = 0; q < 3; ++q) a3[q] = vl.lab[pl][q];
        const int lab[3] = {a2[2], a1[1], a3[2]};
        if (!os.allowed_labels(lab)) continue;
        const int m = os.index_of_labels(lab);
        require(dims3(vv, l, 1 - flag) == 1, "operator transformation: physical leg dimension > 1 is unsupported");
        double fac = 1.0;
        if (su2) {
          fac = cgc->transformation(flag, a2[2], a1[1], a3[2], a2[flag], a3[flag], a2[1 - flag]);
          if (std::fabs(fac) < 1e-8) continue;
        }
        if (si < 0) {
          si = static_cast<int>(s1.size());
          s1.push_back(stage1_uu_op(j, k));
        }
        per[m].push_back(Con{si, l, fac, dims3(op, j, 2)});
      }
      }
    }
    std::vector<uint8_t> mask(os.nten());
    for (int m = 0; m < os.nten(); ++m) mask[m] = !per[m].empty();
    out.clean();
    out.layout(ctx, cx, os, mask, kCgcCanon);
    if (extra)
      for (const auto& e : *extra) {
        e.first->clean();
        e.first->layout(ctx, cx, os, mask, kCgcCanon);
      }
    if (scatter) {
      scatter_run(false);
      return;
    }
    DevBuf sbuf(ctx, std::max<int64_t>(scratch, 1) * sizeof(double) * uu.w());
    ar.set(4, sbuf.get(), scratch);
    ar.set(5, out.dptr(), out.nelem);
    for (const auto& s : s1) {
      const int p = g1.add_problem(4, s.off, s.M, s.M, s.N, false);
      g1.add_seg(p, view(0, uu, s.k, s.K, true), view(2, op, s.j, s.K, false), s.K, 1.0);
    }
    for (int m = 0; m < os.nten(); ++m) {
      if (per[m].empty()) continue;
      int dm[3];
      os.block_dims(m, dm);
      const int M = dm[0] * dm[1];
      const int p = g2.add_problem(5, out.off[m], M, M, dm[2], false);
      for (const auto& c : per[m]) {
        const S1& s = s1[c.s1];
        // step1 as (r*chan) x z, times conj(vv) (z x r')
        Operand A;
        A.arena = 4;
        A.off = s.off;
        A.ld = M;
        A.trans = false;
        g2.add_seg(p, A, view(1, vv, c.l, c.K, false, true), c.K, c.scale);
      }
    }
    run_batch(g1, ctx, cx, ar);
    run_batch(g2, ctx, cx, ar);
    if (extra)
      for (const auto& e : *extra) {  // same tables, other operator and output
        Arenas ae = ar;
        ae.set(2, e.second->dptr(), e.second->nelem);
        ae.set(5, e.first->dptr(), e.first->nelem);
        if (!g1.empty()) g1.run(ctx, prec_of(cx), ae.at);
        if (!g2.empty()) g2.run(ctx, prec_of(cx), ae.at);
      }
    return;
  }

  if (kind == kOperatorPairup) {
    require(op2 != nullptr, "operator pairup: missing site operator");
    const BTensor& sg = *op2;
    const Bond chk = su2 ? sg.st.bonds[1] : sg.st.bonds[1].inverted();
    require(uu.st.bonds[flag] == op.st.bonds[0].inverted() && vv.st.bonds[flag] == op.st.bonds[2] &&
                op.st.bonds[1] == chk,
            "operator pairup wrong bb");
    os.bonds = {uu.st.bonds[2], vv.st.bonds[2].inverted()};
    per.resize(os.nten());
    std::vector<double> hsg_tmp;
    const double* hsg = sg.host_values(hsg_tmp);
    const int P2 = cgc ? cgc->p2() : 3;
    const BlockList ol(op), ul(uu), vl(vv), gl(sg);
    const LabelIndex oix(ol, 0, -1), vix(vl, flag, -1);
    for (int pk = 0; pk < ul.size(); ++pk) {
      const int k = ul.b[pk];
      for (int q = 0; q < 3; ++q) a2[q] = ul.lab[pk][q];
      for (const int pj : oix.get(a2[flag])) {
      const int j = ol.b[pj];
      for (int q = 0; q < 3; ++q) a1[q] = ol.lab[pj][q];
      int si = -1;
      for (const int pl : vix.get(a1[2])) {
        const int l = vl.b[pl];
        for (int q = 0; q < 3; ++q) a3[q] = vl.lab[pl][q];
        for (int pn = 0; pn < gl.size(); ++pn) {
          const int n = gl.b[pn];
          for (int q = 0; q < 3; ++q) a4[q] = gl.lab[pn][q];
          if (a4[0] != a2[1 - flag] || a4[2] != a3[1 - flag] || a4[1] != a1[1]) continue;
          const int lab[2] = {a2[2], a3[2]};
          if (!os.allowed_labels(lab)) continue;
          const int m = os.index_of_labels(lab);
          require(sg.block_size(n) == 1 && dims3(op, j, 1) == 1 && dims3(vv, l, 1 - flag) == 1,
                  "operator pairup: physical/channel leg dimension > 1 is unsupported");
          double fac = 1.0;
          if (su2) {
            fac = cgc->pairup(flag, a2[2], a2[1 - flag == 1 ? 0 : 1], a3[1 - flag == 1 ? 0 : 1], a1[1],
                              a4[0] + a4[2] * P2);
            if (std::fabs(fac) < 1e-8) continue;
          }
          require(!cx || hsg[sg.off[n] * 2 + 1] == 0.0, "operator pairup: complex site operator unsupported");
          const double sv = hsg[sg.off[n] * sg.w()];
          if (si < 0) {
            si = static_cast<int>(s1.size());
            s1.push_back(stage1_uu_op(j, k));
          }
          per[m].push_back(Con{si, l, fac * sv, dims3(op, j, 2)});
        }
      }
      }
    }
    std::vector<uint8_t> mask(os.nten());
    for (int m = 0; m < os.nten(); ++m) mask[m] = !per[m].empty();
    out.clean();
    out.layout(ctx, cx, os, mask, kCgcIdent);
    if (scatter) {
      scatter_run(false);
      return;
    }
    DevBuf sbuf(ctx, std::max<int64_t>(scratch, 1) * sizeof(double) * uu.w());
    ar.set(4, sbuf.get(), scratch);
    ar.set(5, out.dptr(), out.nelem);
    for (const auto& s : s1) {
      const int p = g1.add_problem(4, s.off, s.M, s.M, s.N, false);
      g1.add_seg(p, view(0, uu, s.k, s.K, true), view(2, op, s.j, s.K, false), s.K, 1.0);
    }
    for (int m = 0; m < os.nten(); ++m) {
      if (per[m].empty()) continue;
      int dm[2];
      os.block_dims(m, dm);
      const int p = g2.add_problem(5, out.off[m], dm[0], dm[0], dm[1], false);
      for (const auto& c : per[m]) {
        const S1& s = s1[c.s1];
        Operand A;
        A.arena = 4;
        A.off = s.off;
        A.ld = dm[0];
        A.trans = false;
        g2.add_seg(p, A, view(1, vv, c.l, c.K, false, true), c.K, c.scale);
      }
    }
    run_batch(g1, ctx, cx, ar);
    run_batch(g2, ctx, cx, ar);
    return;
  }
  fail(kErrInvalid, "environment update: unknown kind");
}

}  // namespace

void op_env(BTensor& out, int kind, const BTensor& uu, const BTensor& vv, const BTensor* op1, const BTensor* op2,
            int flag, CgcTables* cgc) {
  op_env_impl(out, kind, uu, vv, op1, op2, flag, cgc, nullptr);
}

void op_env_transform_many(const std::vector<BTensor*>& outs, const BTensor& uu, const BTensor& vv,
                           const std::vector<const BTensor*>& ops, int flag, CgcTables* cgc) {
  require(outs.size() == ops.size(), "operator transformations: one output per operator");
  auto same_layout = [](const BTensor& a, const BTensor& b) {
    return a.cplx == b.cplx && a.nelem == b.nelem && a.st == b.st && a.off == b.off;
  };
  std::vector<uint8_t> done(ops.size(), 0);
  for (size_t q = 0; q < ops.size(); ++q) {
    if (done[q]) continue;
    require(ops[q] != nullptr && outs[q] != nullptr && outs[q] != ops[q], "operator transformations: bad operand");
    EnvExtra extra;
    for (size_t r = q + 1; r < ops.size(); ++r)
      if (!done[r] && ops[r] && outs[r] && outs[r] != ops[r] && same_layout(*ops[r], *ops[q])) {
        extra.emplace_back(outs[r], ops[r]);
        done[r] = 1;
      }
    op_env_impl(*outs[q], kOperatorTransformation, uu, vv, ops[q], nullptr, flag, cgc, &extra);
    done[q] = 1;
  }
}

// ---------------------------------------------------------------------------
void op_fuse(BTensor& out, Context* ctx, bool cplx, SymKind sym, const Bond& b1, const Bond& b2) {
  Struct s;
  s.sym = sym;
  s.locspin = 0;
  s.bonds = {b1, b2, fuse_bonds(sym, b1, b2)};
  std::vector<uint8_t> mask(s.nten());
  for (int b = 0; b < s.nten(); ++b) mask[b] = s.allowed(b);
  out.clean();
  out.layout(ctx, cplx, s, mask, kCgcCanon);
  std::vector<int64_t> pos(s.bonds[2].nmom(), 0);
  int idx[3], d[3];
  for (int b = 0; b < s.nten(); ++b) {
    if (!mask[b]) continue;
    s.decode(b, idx);
    s.block_dims(b, d);
    const int64_t n = int64_t(d[0]) * d[1];
    require(pos[idx[2]] + n <= d[2], "BlockTensor::fuse(bondtype, bondtype) wrong");
    const int w = out.w();
    SADMRG_HIP(launch_strided_fill(n, 1.0, out.block_ptr(b) + pos[idx[2]] * n * w, (1 + n) * w, ctx->stream()));
    pos[idx[2]] += n;
  }
  out.fuse_id = true;
}

void op_fuse_to_singlet(BTensor& out, Context* ctx, bool cplx, SymKind sym, const Bond& b1, const Bond& b2) {
  Struct s;
  s.sym = sym;
  s.locspin = 0;
  s.bonds = {b1, b2};
  std::vector<uint8_t> mask(s.nten(), 0);
  int lab[2];
  for (int b = 0; b < s.nten(); ++b) {
    s.labels(b, lab);
    mask[b] = sym == SymKind::SU2 ? (lab[0] == lab[1]) : (lab[0] * b1.dir + lab[1] * b2.dir == 0);
  }
  out.clean();
  out.layout(ctx, cplx, s, mask, kCgcSinglet);
  for (int b = 0; b < s.nten(); ++b)
    if (mask[b]) SADMRG_HIP(launch_strided_fill(s.block_size(b), 1.0, out.block_ptr(b), out.w(), ctx->stream()));
}

// Cyclic bond shift (Dense::shift is a 2-D transpose of a B x A split).
static void shift_impl(BTensor& t, int i0, int i1, bool vectran_dirs) {
  if (t.is_null()) return;
  const int n = t.nbond();
  const int ishift = (i1 > i0) ? (i1 - i0) : (n - (i0 - i1));
  Struct ns = t.st;
  for (int i = 0; i < n; ++i) ns.bonds[(i + ishift) % n] = t.st.bonds[i];
  if (vectran_dirs && t.sym() == SymKind::SU2) {
    ns.bonds[0].dir = 1;
    ns.bonds[1].dir = 1;
    ns.bonds[2].dir = -1;
  }
  std::vector<uint8_t> mask(ns.nten(), 0);
  std::vector<int> map(t.nten(), -1);
  int lab[kMaxLeg], lab1[kMaxLeg];
  for (int b = 0; b < t.nten(); ++b) {
    if (!t.st.allowed(b) || !t.present(b)) continue;
    t.st.labels(b, lab);
    for (int k = 0; k < n; ++k) lab1[(k + ishift) % n] = lab[k];
    const int j = ns.index_of_labels(lab1);
    if (j < 0) continue;
    map[b] = j;
    mask[j] = 1;
  }
  BTensor r;
  r.layout(t.ctx, t.cplx, ns, mask, n == 3 ? kCgcCanon : (t.tag.empty() ? kCgcIdent : t.tag[0]));
  const int p = n - ishift;
  for (int b = 0; b < t.nten(); ++b) {
    if (map[b] < 0) continue;
    int d[kMaxLeg];
    t.st.block_dims(b, d);
    int64_t B = 1, A = 1;
    for (int i = 0; i < p; ++i) B *= d[i];
    for (int i = p; i < n; ++i) A *= d[i];
    // new[a + A*b] = old[b + B*a]
    SADMRG_HIP(launch_copy2d(static_cast<int>(A), static_cast<int>(B), t.block_ptr(b), B, 1, r.block_ptr(map[b]), A,
                             t.w(), 1.0, nullptr, 0, t.ctx->stream()));
  }
  t = std::move(r);
}

void op_vectran(BTensor& t, bool left_to_right) {
  require(t.nbond() == 3, "vectran: 3-leg tensor expected");
  if (left_to_right) shift_impl(t, 1, 0, true);
  else shift_impl(t, 2, 0, true);
}

void op_shift(BTensor& t, int i0, int i1) {
  if (i0 == i1) return;
  shift_impl(t, i0, i1, false);
}

void op_take_conjugate(BTensor& t, int bond) {
  require(bond >= 0 && bond < t.nbond(), "take_conjugate: bad bond");
  t.st.bonds[bond].dir = -t.st.bonds[bond].dir;
  if (t.sym() == SymKind::SU2 && t.nbond() == 2)
    for (auto& g : t.tag) g = (g == kCgcIdent) ? kCgcSinglet : kCgcOther;
  if (t.sym() != SymKind::U1 || t.is_null()) return;
  // U(1): the charge rule sum dir*q = locspin changes with the direction, so
  // blocks it no longer allows drop out.  The reference keeps their data in
  // memory but every block loop skips them (U1Struct::take_conjugate only
  // flips the direction, u1.cpp:295-298); the visible tensor is the same.
  std::vector<uint8_t> mask(t.nten(), 0);
  bool drop = false;
  for (int b = 0; b < t.nten(); ++b) {
    mask[b] = t.present(b) && t.st.allowed(b);
    if (t.present(b) && !mask[b]) drop = true;
  }
  if (!drop) return;
  BTensor r;
  r.layout(t.ctx, t.cplx, t.st, mask, t.tag.empty() ? kCgcIdent : t.tag[0]);
  for (int b = 0; b < t.nten(); ++b)
    if (mask[b] && t.block_size(b) > 0)
      SADMRG_HIP(hipMemcpyAsync(r.block_ptr(b), t.block_ptr(b), t.block_size(b) * sizeof(double) * t.w(),
                                hipMemcpyDeviceToDevice, t.ctx->stream()));
  t = std::move(r);
}

void op_makeup_input_vector(BTensor& t) {
  require(t.nbond() == 2, "BlockTensor::makeup_input_vector called by wrong tensor");
  const int n0 = t.st.bonds[0].nmom(), n1 = t.st.bonds[1].nmom();
  std::vector<uint8_t> mask(t.nten(), 0);
  bool change = false;
  for (int b = 0; b < t.nten(); ++b) mask[b] = t.present(b);
  for (int i = 0; i < n0; ++i)
    for (int j = 0; j < n1; ++j)
      if (t.st.bonds[0].q[i] == t.st.bonds[1].q[j] && !mask[i + j * n0]) {
        mask[i + j * n0] = 1;
        change = true;
      }
  if (!change) return;
  BTensor r;
  r.layout(t.ctx, t.cplx, t.st, mask, kCgcSinglet);
  for (int b = 0; b < t.nten(); ++b)
    if (t.present(b) && t.block_size(b) > 0)
      SADMRG_HIP(hipMemcpyAsync(r.block_ptr(b), t.block_ptr(b), t.block_size(b) * sizeof(double) * t.w(),
                                hipMemcpyDeviceToDevice, t.ctx->stream()));
  t = std::move(r);
}

void op_make_spinor_start(BTensor& t, Context* ctx, bool cplx, SymKind sym, int physpn) {
  Struct s;
  s.sym = sym;
  s.locspin = 0;
  std::vector<uint8_t> mask;
  std::vector<double> vals;
  if (sym == SymKind::SU2) {
    s.bonds = {Bond::make(-1, {physpn}, {1}), Bond::make(-1, {2}, {1}), Bond::make(1, {physpn}, {1})};
    mask.assign(s.nten(), 0);
    mask[0] = 1;
    vals = {1.0};
  } else {
    std::vector<int> q, one;
    for (int i = 0; i <= physpn; ++i) { q.push_back(-physpn + 2 * i); one.push_back(1); }
    s.bonds = {Bond::make(-1, q, one), Bond::make(-1, {-2, 0, 2}, {1, 1, 1}), Bond::make(1, q, one)};
    mask.assign(s.nten(), 0);
    int lab[3];
    for (int b = 0; b < s.nten(); ++b) {
      if (!s.allowed(b)) continue;
      s.labels(b, lab);
      if (lab[0] == lab[2] && lab[1] == 0) { mask[b] = 1; vals.push_back(lab[0] / 2.0); }
      else if (lab[0] != lab[2]) { mask[b] = 1; vals.push_back(1.0 / std::sqrt(2.0)); }
    }
  }
  t.clean();
  t.layout(ctx, cplx, s, mask, kCgcCanon);
  std::vector<double> h(t.nelem * t.w(), 0.0);
  for (size_t i = 0; i < vals.size(); ++i) h[i * t.w()] = vals[i];
  if (!h.empty()) {
    SADMRG_HIP(hipMemcpyAsync(t.dptr(), h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream()));
    ctx->sync();
  }
}

// ---------------------------------------------------------------------------
void op_scale(BTensor& t, double re, double im) {
  if (t.is_null() || t.nelem == 0) return;
  if (!t.cplx) {
    SADMRG_HIP(launch_axpby(t.nelem, re, 0.0, nullptr, t.dptr(), t.ctx->stream()));
  } else {
    SADMRG_HIP(launch_zscal(t.nelem, re, im, t.dptr(), t.dptr(), t.ctx->stream()));
  }
}

void op_axpy_channels(BTensor& y, double s, const dcmplex (&w)[3], const BTensor& x, int bond) {
  if (x.is_null()) return;
  require(x.sym() == SymKind::U1, "channel weights: U(1) tensors only");
  require(bond >= 0 && bond < x.nbond(), "channel weights: bond out of range");
  for (const dcmplex& c : w)
    require(x.cplx || c.imag() == 0.0, "channel weights: complex weights need complex tensors");
  BTensor t;
  copy_tensor(t, x);
  std::vector<int> lab(t.nbond());
  for (int b = 0; b < t.nten(); ++b) {
    if (!t.present(b) || t.block_size(b) == 0) continue;
    t.st.labels(b, lab.data());
    const int q = lab[bond];
    require(q == -2 || q == 0 || q == 2, "channel weights: channel label outside {-2, 0, 2}");
    const dcmplex c = s * w[q / 2 + 1];
    if (!t.cplx)
      SADMRG_HIP(launch_axpby(t.block_size(b), c.real(), 0.0, nullptr, t.block_ptr(b), t.ctx->stream()));
    else
      SADMRG_HIP(launch_zscal(t.block_size(b), c.real(), c.imag(), t.block_ptr(b), t.block_ptr(b), t.ctx->stream()));
  }
  op_axpy(y, 1.0, 0.0, t);  // t's memory returns to the stream-ordered cache
}

void op_fill(BTensor& t, double v) {
  if (t.is_null() || t.nelem == 0) return;
  if (!t.cplx) SADMRG_HIP(launch_fill(t.nelem, v, t.dptr(), t.ctx->stream()));
  else {
    SADMRG_HIP(launch_fill(t.nelem * 2, 0.0, t.dptr(), t.ctx->stream()));
    SADMRG_HIP(launch_strided_fill(t.nelem, v, t.dptr(), 2, t.ctx->stream()));
  }
}

// y += (re, im) * x over the blocks of y's structure; a null block of y that
// is present in x becomes a scaled copy (block_tensor_inst.cpp:225-278).
void op_axpy(BTensor& y, double re, double im, const BTensor& x) {
  if (x.is_null()) return;
  if (y.is_null()) {
    copy_tensor(y, x);
    op_scale(y, re, im);
    return;
  }
  require(y.st == x.st, "BlockTensor::operator+=: structures differ");
  std::vector<uint8_t> mask(y.nten());
  bool grow = false;
  for (int b = 0; b < y.nten(); ++b) {
    mask[b] = y.present(b) || (x.present(b) && y.st.allowed(b));
    if (mask[b] && !y.present(b)) grow = true;
  }
  if (grow) {
    BTensor r;
    r.layout(y.ctx, y.cplx, y.st, mask, y.tag.empty() ? kCgcIdent : y.tag[0]);
    for (int b = 0; b < y.nten(); ++b)
      if (y.present(b) && y.block_size(b) > 0)
        SADMRG_HIP(hipMemcpyAsync(r.block_ptr(b), y.block_ptr(b), y.block_size(b) * sizeof(double) * y.w(),
                                  hipMemcpyDeviceToDevice, y.ctx->stream()));
    y = std::move(r);
  }
  Context* ctx = y.ctx;
  // identical layouts: one fused update
  if (y.off == x.off) {
    if (!y.cplx) SADMRG_HIP(launch_axpby(y.nelem, 1.0, re, x.dptr(), y.dptr(), ctx->stream()));
    else {
      double* c = ctx->dscal();
      const double h[2] = {re, im};
      ctx->upload(c, h, sizeof(h));
      SADMRG_HIP(launch_gemv_n_update(x.dptr(), 0, 1, c, y.dptr(), y.nelem, 1, ctx->stream()));
    }
    return;
  }
  for (int b = 0; b < y.nten(); ++b) {
    if (!y.st.allowed(b) || !x.present(b) || !y.present(b)) continue;
    const int64_t n = y.block_size(b);
    if (!y.cplx) {
      SADMRG_HIP(launch_axpby(n, 1.0, re, x.block_ptr(b), y.block_ptr(b), ctx->stream()));
    } else {
      double* c = ctx->dscal();
      const double h[2] = {re, im};
      ctx->upload(c, h, sizeof(h));
      SADMRG_HIP(launch_gemv_n_update(x.block_ptr(b), 0, 1, c, y.block_ptr(b), n, 1, ctx->stream()));
    }
  }
}

dcmplex op_dot(const BTensor& x, const BTensor& y) {
  if (x.is_null() || y.is_null()) return dcmplex(0, 0);
  Context* ctx = x.ctx;
  require(x.cplx == y.cplx && x.ctx == y.ctx, "inner_prod_c: scalar types or contexts differ");
  // The reference walks x's blocks and indexes y's by the same block number,
  // exiting when two present blocks differ in size (dense_inst.cpp:637); a
  // present block of x past y's block count would be read out of range there.
  std::vector<int> blocks;
  for (int b = 0; b < x.nten(); ++b) {
    if (!x.st.allowed(b) || !x.present(b)) continue;
    require(b < y.nten(), "inner_prod_c: structures differ (block past the other tensor's block count)");
    if (!y.present(b)) continue;
    require(x.block_size(b) == y.block_size(b), "inner_prod_c: size mismatch of block " + std::to_string(b));
    blocks.push_back(b);
  }
  const int nb = static_cast<int>(blocks.size());
  require(nb <= Context::kMaxDscal / 2, "inner_prod_c: too many blocks");
  for (int i = 0; i < nb; ++i) {
    const int b = blocks[i];
    SADMRG_HIP(launch_gemv_t(x.block_ptr(b), 0, 1, y.block_ptr(b), x.block_size(b), x.cplx ? 1 : 0, ctx->partial(),
                             1.0, ctx->dscal() + 2 * i, ctx->stream()));
  }
  std::vector<double> h(2 * std::max(nb, 1), 0.0);
  if (nb) {
    SADMRG_HIP(hipMemcpyAsync(h.data(), ctx->dscal(), 2 * nb * sizeof(double), hipMemcpyDeviceToHost, ctx->stream()));
  }
  ctx->sync();
  dcmplex s(0, 0);
  for (int i = 0; i < nb; ++i) s += dcmplex(h[2 * i], h[2 * i + 1]);
  return s;
}

double op_normalize(BTensor& t) {
  double nor = std::sqrt(std::fabs(op_dot(t, t).real()));
  if (nor > 1e-12) op_scale(t, 1.0 / nor);
  return nor;
}

std::mt19937_64& global_rng() {
  static std::mt19937_64 g(42);
  return g;
}

void reset_global_rng(uint64_t seed) { global_rng().seed(seed); }

void op_random_init(BTensor& t) {
  if (t.is_null() || t.nelem == 0) return;
  std::uniform_real_distribution<double> dist(-1.0, 1.0);
  auto& g = global_rng();
  std::vector<double> h(t.nelem * t.w());
  for (int b = 0; b < t.nten(); ++b) {
    if (!t.st.allowed(b) || !t.present(b)) continue;
    const int64_t n = t.block_size(b);
    double* p = h.data() + t.off[b] * t.w();
    for (int64_t e = 0; e < n; ++e) {
      if (!t.cplx) p[e] = dist(g);
      else {
        // dcmplex(dist(g), dist(g)) as compiled by GCC evaluates the
        // imaginary argument first
        const double im = dist(g);
        const double re = dist(g);
        p[2 * e] = re;
        p[2 * e + 1] = im;
      }
    }
  }
  SADMRG_HIP(hipMemcpyAsync(t.dptr(), h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, t.ctx->stream()));
  t.ctx->sync();
}

// ---------------------------------------------------------------------------
// Adjoint of overlap_transformation.  The forward update is, per output
// block m and contributing (X block j, uu block k, vv block l),
//   T(X)_m += uu_k^T X_j vv_l            (uu_k: l x r, X_j: l x y, vv_l: y x r')
// so <lam, T(X)> = sum_j <X_j, sum_{k,l,m} uu_k lam_m vv_l^T>, with the same
// block matching as the forward loop nest.  Lowered as two grouped launches:
// t_{k,m} = uu_k lam_m, then out_j = sum t_{k,m} vv_l^T.
void op_env_adjoint(BTensor& out, const BTensor& lam, const BTensor& uu, const BTensor& vv, const BTensor& like,
                    int flag) {
  Context* ctx = uu.ctx;
  require(!uu.cplx && !vv.cplx && !lam.cplx, "adjoint environment update: real tensors only");
  require(flag == 0 || flag == 1, "adjoint environment update: flag must be 0 or 1");
  require(uu.nbond() == 3 && vv.nbond() == 3 && lam.nbond() == 2 && like.nbond() == 2,
          "adjoint environment update: ranks");
  require(lam.st.bonds[0] == uu.st.bonds[2] && lam.st.bonds[1] == vv.st.bonds[2].inverted(),
          "adjoint environment update: lam does not match the output of the forward update");
  require(uu.st.bonds[flag] == like.st.bonds[0].inverted() && vv.st.bonds[flag] == like.st.bonds[1],
          "adjoint environment update: `like` does not match the input of the forward update");
  int a1[3], a2[3], a3[3];
  struct T1 { int k, m; int64_t off; int M, N, K; };
  std::vector<T1> t1;
  std::map<std::pair<int, int>, int> t1_of;
  struct Con { int t, l; };
  std::vector<std::vector<Con>> per(like.nten());
  int64_t scratch = 0;
  for (int j = 0; j < like.nten(); ++j) {
    if (!like.st.allowed(j)) continue;
    labels_of(like, j, a1);
    for (int k = 0; k < uu.nten(); ++k) {
      if (!uu.st.allowed(k) || !uu.present(k)) continue;
      labels_of(uu, k, a2);
      if (a1[0] != a2[flag]) continue;
      for (int l = 0; l < vv.nten(); ++l) {
        if (!vv.st.allowed(l) || !vv.present(l)) continue;
        labels_of(vv, l, a3);
        if (a1[1] != a3[flag] || a2[1 - flag] != a3[1 - flag]) continue;
        const int lab[2] = {a2[2], a3[2]};
        if (!lam.st.allowed_labels(lab)) continue;
        const int m = lam.st.index_of_labels(lab);
        if (m < 0 || !lam.present(m)) continue;
        require(dims3(uu, k, 1 - flag) == 1 && dims3(vv, l, 1 - flag) == 1,
                "adjoint environment update: physical leg dimension > 1 is unsupported");
        auto it = t1_of.find({k, m});
        int ti;
        if (it == t1_of.end()) {
          const int l0 = dims3(uu, k, flag), r = dims3(uu, k, 2), r2 = dims3(vv, l, 2);
          ti = static_cast<int>(t1.size());
          t1.push_back(T1{k, m, scratch, l0, r2, r});
          t1_of.emplace(std::make_pair(k, m), ti);
          scratch += int64_t(l0) * r2;
        } else {
          ti = it->second;
        }
        per[j].push_back(Con{ti, l});
      }
    }
  }
  std::vector<uint8_t> mask(like.nten());
  for (int j = 0; j < like.nten(); ++j) mask[j] = !per[j].empty();
  out.clean();
  out.layout(ctx, false, like.st, mask, kCgcIdent);
  DevBuf sbuf(ctx, std::max<int64_t>(scratch, 1) * sizeof(double));
  Arenas ar;
  ar.set(0, uu.dptr(), uu.nelem);
  ar.set(1, vv.dptr(), vv.nelem);
  ar.set(2, lam.dptr(), lam.nelem);
  ar.set(4, sbuf.get(), scratch);
  ar.set(5, out.dptr(), out.nelem);
  GemmBatch g1, g2;
  for (const auto& t : t1) {
    const int p = g1.add_problem(4, t.off, t.M, t.M, t.N, false);
    g1.add_seg(p, view(0, uu, t.k, t.M, false), view(2, lam, t.m, t.K, false), t.K, 1.0);
  }
  for (int j = 0; j < like.nten(); ++j) {
    if (per[j].empty()) continue;
    int d[2];
    like.st.block_dims(j, d);
    const int p = g2.add_problem(5, out.off[j], d[0], d[0], d[1], false);
    for (const auto& c : per[j]) {
      const T1& t = t1[c.t];
      require(t.M == d[0] && dims3(vv, c.l, flag) == d[1], "adjoint environment update: dimension mismatch");
      Operand A;
      A.arena = 4;
      A.off = t.off;
      A.ld = t.M;
      g2.add_seg(p, A, view(1, vv, c.l, d[1], true), t.N, 1.0);
    }
  }
  run_batch(g1, ctx, false, ar);
  run_batch(g2, ctx, false, ar);
  ctx->sync();
}

namespace {
std::atomic<int> g_gram_min{256};
std::atomic<int> g_jacobi_max{64};
// sweeps before a Jacobi sector is handed to gesdd (quadratic convergence:
// about 6 to 12 are needed)
constexpr int kJacobiSweeps = 60;
std::atomic<int> g_svd_eig{1};
SvdStats g_svd_stats;
}
void set_gram_min(int n) { g_gram_min.store(n); }
void set_svd_jacobi_max(int q) { g_jacobi_max.store(q); }
int svd_jacobi_max() { return g_jacobi_max.load(); }
int gram_min() { return g_gram_min.load(); }
void set_svd_eigensolver(int v) { g_svd_eig.store(v); }
int svd_eigensolver() { return g_svd_eig.load(); }
SvdStats& svd_stats() { return g_svd_stats; }

// ---------------------------------------------------------------------------
// Block SVD truncation (block_tensor_inst.cpp:501-637, dense_inst.cpp:447-541).
void op_svd(const BTensor& vec, BTensor& tu, double p1, BTensor& tv, double p2, int max_dcut,
            std::vector<double>& wout, std::vector<int>* kept_out) {
  const auto t_start = std::chrono::steady_clock::now();
  g_svd_stats.calls++;
  Context* ctx = vec.ctx;
  const SymKind sym = vec.sym();
  const bool cx = vec.cplx;
  require(vec.nbond() == 2 && tu.nbond() == 3 && tv.nbond() == 3, "svd: ranks");
  const Bond bb2_0 = tu.st.bonds[0], bb2_1 = tu.st.bonds[1], bb2_2 = tv.st.bonds[0], bb2_3 = tv.st.bonds[1];
  const Bond bb0 = fuse_bonds(sym, bb2_0, bb2_1);
  const Bond bb1 = fuse_bonds(sym, bb2_2, bb2_3);
  const int nm1 = bb0.nmom(), nm2 = bb1.nmom();
  require(vec.st.bonds[0].q == bb0.q && vec.st.bonds[1].q == bb1.q, "svd: vector structure does not match tu/tv");
  struct Sec { int i, j; int m, n, mn; int64_t u_off, vt_off, s_off; };
  std::vector<Sec> secs;
  std::vector<int> sec_of(nm1, -1);
  int64_t u_tot = 0, vt_tot = 0, s_tot = 0, a_max = 0;
  for (int i = 0; i < nm1; ++i)
    for (int j = 0; j < nm2; ++j) {
      if (bb0.q[i] != bb1.q[j]) continue;
      const int b = i + j * nm1;
      if (!vec.present(b)) continue;
      Sec s{i, j, bb0.dim[i], bb1.dim[j], std::min(bb0.dim[i], bb1.dim[j]), u_tot, vt_tot, s_tot};
      sec_of[i] = static_cast<int>(secs.size());
      secs.push_back(s);
      u_tot += int64_t(s.m) * s.mn;
      vt_tot += int64_t(s.mn) * s.n;
      s_tot += s.mn;
      a_max = std::max<int64_t>(a_max, int64_t(s.m) * s.n);
    }
  const int w = vec.w();
  DevBuf du(ctx, std::max<int64_t>(u_tot, 1) * 8 * w), dvt(ctx, std::max<int64_t>(vt_tot, 1) * 8 * w);
  DevBuf ds(ctx, std::max<int64_t>(s_tot, 1) * 8), dwork(ctx, std::max<int64_t>(a_max, 1) * 8 * w);
  DevBuf dinfo(ctx, sizeof(int) * std::max<size_t>(secs.size(), 1));
  SADMRG_HIP(hipMemsetAsync(dinfo.get(), 0, sizeof(int) * std::max<size_t>(secs.size(), 1), ctx->stream()));
  rocblas_handle h = static_cast<rocblas_handle>(ctx->solver());
  // small real sectors of sweep steps (one side absorbs S): one-sided
  // Jacobi in LDS, every such sector in one launch (jacobi_svd_kernel); the
  // rest by gesdd or the Gram route below.  iDMRG steps, (p1, p2) = (0, 0),
  // stay on gesdd: their blocks are rank deficient, the kept rule's 1e-10
  // threshold then decides on round-off values (gesdd's sqrt(eps) S_max
  // against the kernel's exact zeros), and the growth trajectory (the start
  // vectors drawn per kept block) follows that decision.
  std::vector<size_t> jac, rest;
  const int jmax = svd_jacobi_max();
  const bool sweep_step = p1 + p2 > 0.0;
  for (size_t k = 0; k < secs.size(); ++k)
    (!cx && sweep_step && jmax > 0 && jacobi_svd_fits(secs[k].m, secs[k].n, jmax) ? jac : rest).push_back(k);
  DevBuf jjobs, jflags;
  std::vector<int> jf(jac.size(), 0);
  bool jac_pending = false;
  // One sector, unpadded, straight into its U / V' / S slots.
  auto single = [&](size_t si) {
    const Sec& s = secs[si];
    const int b = s.i + s.j * nm1;
    SADMRG_HIP(hipMemcpyAsync(dwork.get(), vec.block_ptr(b), int64_t(s.m) * s.n * 8 * w, hipMemcpyDeviceToDevice,
                              ctx->stream()));
    rocblas_status st;
    if (!cx)
      st = rocsolver_dgesdd(h, rocblas_svect_singular, rocblas_svect_singular, s.m, s.n, dwork.as<double>(), s.m,
                            ds.as<double>() + s.s_off, du.as<double>() + s.u_off, s.m, dvt.as<double>() + s.vt_off,
                            s.mn, dinfo.as<int>() + si);
    else
      st = rocsolver_zgesdd(h, rocblas_svect_singular, rocblas_svect_singular, s.m, s.n,
                            dwork.as<rocblas_double_complex>(), s.m, ds.as<double>() + s.s_off,
                            du.as<rocblas_double_complex>() + s.u_off, s.m,
                            dvt.as<rocblas_double_complex>() + s.vt_off, s.mn, dinfo.as<int>() + si);
    if (st != rocblas_status_success) fail(kErrNumerics, "rocsolver gesdd failed");
  };
  // Sectors of similar size share one strided-batched gesdd, each padded with
  // zeros to the batch's largest (m, n).  Zero padding appends zero singular
  // values and leaves the leading singular triplets unchanged, their vectors
  // vanishing on the padding, so a sector's U (m x mn) and V' (mn x n) are the
  // leading blocks of its padded factors.  A sector whose largest singular
  // value is round-off (kept as one state by the reference's rule) is redone
  // unpadded, because its singular vectors are then not determined by the
  // data and a padded solve may return one that lives on the padding.
  std::vector<uint8_t> padded(secs.size(), 0);
  constexpr int kMaxBatch = 32;
  auto svd_batch = [&](const std::vector<size_t>& list, int mp, int np) {
    const int cnt = static_cast<int>(list.size());
    const int kp = std::min(mp, np);
    const int64_t sa = int64_t(mp) * np, su = int64_t(mp) * kp, sv = int64_t(kp) * np;
    DevBuf ba(ctx, sa * cnt * 8 * w), bs(ctx, int64_t(kp) * cnt * 8), bu(ctx, su * cnt * 8 * w),
        bv(ctx, sv * cnt * 8 * w), binfo(ctx, sizeof(int) * cnt);
    SADMRG_HIP(launch_fill(sa * cnt * w, 0.0, ba.as<double>(), ctx->stream()));
    for (int q = 0; q < cnt; ++q) {
      const Sec& s = secs[list[q]];
      SADMRG_HIP(launch_copy2d(s.m, s.n, vec.block_ptr(s.i + s.j * nm1), 1, s.m, ba.as<double>() + sa * q * w, mp, w,
                               1.0, nullptr, 0, ctx->stream()));
    }
    rocblas_status st;
    if (!cx)
      st = rocsolver_dgesdd_strided_batched(h, rocblas_svect_singular, rocblas_svect_singular, mp, np, ba.as<double>(),
                                            mp, sa, bs.as<double>(), kp, bu.as<double>(), mp, su, bv.as<double>(), kp,
                                            sv, binfo.as<int>(), cnt);
    else
      st = rocsolver_zgesdd_strided_batched(
          h, rocblas_svect_singular, rocblas_svect_singular, mp, np, ba.as<rocblas_double_complex>(), mp, sa,
          bs.as<double>(), kp, bu.as<rocblas_double_complex>(), mp, su, bv.as<rocblas_double_complex>(), kp, sv,
          binfo.as<int>(), cnt);
    if (st != rocblas_status_success) fail(kErrNumerics, "rocsolver gesdd (batched) failed");
    for (int q = 0; q < cnt; ++q) {
      const size_t si = list[q];
      const Sec& s = secs[si];
      padded[si] = 1;
      SADMRG_HIP(hipMemcpyAsync(ds.as<double>() + s.s_off, bs.as<double>() + int64_t(kp) * q, int64_t(s.mn) * 8,
                                hipMemcpyDeviceToDevice, ctx->stream()));
      SADMRG_HIP(hipMemcpyAsync(dinfo.as<int>() + si, binfo.as<int>() + q, sizeof(int), hipMemcpyDeviceToDevice,
                                ctx->stream()));
      SADMRG_HIP(launch_copy2d(s.m, s.mn, bu.as<double>() + su * q * w, 1, mp, du.as<double>() + s.u_off * w, s.m, w,
                               1.0, nullptr, 0, ctx->stream()));
      SADMRG_HIP(launch_copy2d(s.mn, s.n, bv.as<double>() + sv * q * w, 1, kp, dvt.as<double>() + s.vt_off * w, s.mn,
                               w, 1.0, nullptr, 0, ctx->stream()));
    }
    ctx->sync();  // batch buffers are released at the end of this scope
  };
  // Density-matrix route for large sectors of sweep steps, where one side
  // absorbs S, (p1, p2) = (1, 0) or (0, 1): the isometric side is the
  // eigenbasis of the Gram matrix of that side (psi^T psi for V, psi psi^T for
  // U; strided-batched syevd), the other side is psi times it divided by S --
  // the projector builder multiplies S back in, so the division is undone
  // exactly and small S cost no accuracy there.  A Gram eigenvalue lambda_k
  // determines S_k = sqrt(lambda_k) only while lambda_k > kGramRes lambda_max;
  // below that ("unresolved") S_k is reported as the norm of psi times the
  // eigenvector, which is exact to round-off of S_max but, inside a cluster of
  // unresolved eigenvalues, belongs to an arbitrary mixture of those singular
  // vectors.  The route is therefore only accepted when no unresolved value
  // can take part in the kept set (see the check after the decompositions);
  // otherwise the sector is decomposed again by gesdd.
  const int gmin = gram_min();
  constexpr double kGramRes = 1e-13;
  const bool right_gram = (p1 == 1.0 && p2 == 0.0), left_gram = (p1 == 0.0 && p2 == 1.0);
  // the route can only be accepted when the cap can bind: more singular
  // values in total than max_dcut
  const bool gram = gmin > 0 && !cx && (right_gram || left_gram) && max_dcut > 0 && s_tot > max_dcut;
  std::vector<uint8_t> gram_sec(secs.size(), 0);
  std::vector<double> gram_s(s_tot, 0.0);
  std::vector<double> gram_res(secs.size(), 0.0);  // > 0: the sector has unresolved values below this
  std::vector<int> gram_nres(secs.size(), 0);        // resolved values of a Gram sector
  auto gram_batch = [&](const std::vector<size_t>& list, int mp, int np) {
    const int cnt = static_cast<int>(list.size());
    const int g = right_gram ? np : mp;  // Gram order
    const int64_t sa = int64_t(mp) * np, sg = int64_t(g) * g;
    DevBuf ba(ctx, sa * cnt * 8), bg(ctx, sg * cnt * 8), by(ctx, sa * cnt * 8), bd(ctx, int64_t(g) * cnt * 8),
        be(ctx, int64_t(g) * cnt * 8), bn(ctx, int64_t(g) * cnt * 8), binfo(ctx, sizeof(int) * cnt);
    SADMRG_HIP(launch_fill(sa * cnt, 0.0, ba.as<double>(), ctx->stream()));
    for (int q = 0; q < cnt; ++q) {
      const Sec& sc = secs[list[q]];
      SADMRG_HIP(launch_copy2d(sc.m, sc.n, vec.block_ptr(sc.i + sc.j * nm1), 1, sc.m, ba.as<double>() + sa * q, mp,
                               1, 1.0, nullptr, 0, ctx->stream()));
    }
    rocblas_handle rb = h;
    const double one = 1.0, zero = 0.0;
    rocblas_status st;
    if (right_gram)
      st = rocblas_dgemm_strided_batched(rb, rocblas_operation_transpose, rocblas_operation_none, np, np, mp, &one,
                                         ba.as<double>(), mp, sa, ba.as<double>(), mp, sa, &zero, bg.as<double>(),
                                         np, sg, cnt);
    else
      st = rocblas_dgemm_strided_batched(rb, rocblas_operation_none, rocblas_operation_transpose, mp, mp, np, &one,
                                         ba.as<double>(), mp, sa, ba.as<double>(), mp, sa, &zero, bg.as<double>(),
                                         mp, sg, cnt);
    if (st != rocblas_status_success) fail(kErrNumerics, "rocblas gemm (Gram) failed");
    if (rocsolver_dsyevd_strided_batched(h, rocblas_evect_original, rocblas_fill_upper, g, bg.as<double>(), g, sg,
                                         bd.as<double>(), g, be.as<double>(), g, binfo.as<int>(), cnt) !=
        rocblas_status_success)
      fail(kErrNumerics, "rocsolver syevd (Gram) failed");
    // Y = psi times the eigenvectors: right, Y = A V (mp x np, column c of Y
    // belongs to eigenvector c); left, Y = G^T A (row c of Y to eigenvector c).
    // bn receives the norms of those columns / rows.
    if (right_gram) {
      if (rocblas_dgemm_strided_batched(rb, rocblas_operation_none, rocblas_operation_none, mp, np, np, &one,
                                        ba.as<double>(), mp, sa, bg.as<double>(), np, sg, &zero, by.as<double>(), mp,
                                        sa, cnt) != rocblas_status_success)
        fail(kErrNumerics, "rocblas gemm (Gram) failed");
      SADMRG_HIP(launch_strided_norms(np * cnt, mp, by.as<double>(), 1, mp, bn.as<double>(), ctx->stream()));
    } else {
      if (rocblas_dgemm_strided_batched(rb, rocblas_operation_transpose, rocblas_operation_none, mp, np, mp, &one,
                                        bg.as<double>(), mp, sg, ba.as<double>(), mp, sa, &zero, by.as<double>(), mp,
                                        sa, cnt) != rocblas_status_success)
        fail(kErrNumerics, "rocblas gemm (Gram) failed");
      for (int q = 0; q < cnt; ++q)
        SADMRG_HIP(launch_strided_norms(mp, np, by.as<double>() + sa * q, mp, 1, bn.as<double>() + int64_t(g) * q,
                                        ctx->stream()));
    }
    std::vector<double> hd(int64_t(g) * cnt), hn(int64_t(g) * cnt);
    std::vector<int> hi(cnt);
    SADMRG_HIP(hipMemcpyAsync(hd.data(), bd.get(), hd.size() * 8, hipMemcpyDeviceToHost, ctx->stream()));
    SADMRG_HIP(hipMemcpyAsync(hn.data(), bn.get(), hn.size() * 8, hipMemcpyDeviceToHost, ctx->stream()));
    SADMRG_HIP(hipMemcpyAsync(hi.data(), binfo.get(), hi.size() * sizeof(int), hipMemcpyDeviceToHost, ctx->stream()));
    ctx->sync();
    // S (descending rank k <-> eigenvector column g-1-k) and the scale 1/S
    // of the non-isometric factor, by rank (kinv) and by eigenvector column (cinv)
    std::vector<double> kinv(int64_t(g) * cnt, 0.0), cinv(int64_t(g) * cnt, 0.0);
    for (int q = 0; q < cnt; ++q) {
      const size_t si = list[q];
      const Sec& sc = secs[si];
      if (hi[q] != 0) fail(kErrNumerics, "Dense::svd: syevd info=" + std::to_string(hi[q]));
      const double* lam = hd.data() + int64_t(g) * q;
      const double* nrm = hn.data() + int64_t(g) * q;
      const double lmax = std::max(lam[g - 1], 0.0);
      double* S = gram_s.data() + sc.s_off;
      int nres = 0;
      for (int k = 0; k < sc.mn; ++k) {
        const int c = g - 1 - k;
        const double l = lam[c];
        const bool resolved = l > kGramRes * lmax && l > 0.0;
        S[k] = resolved ? std::sqrt(l) : nrm[c];
        if (resolved) nres = k + 1;
        cinv[int64_t(g) * q + c] = S[k] > 0.0 ? 1.0 / S[k] : 0.0;
        kinv[int64_t(g) * q + k] = cinv[int64_t(g) * q + c];
      }
      gram_nres[si] = nres;
      if (nres < sc.mn) {
        // unresolved tail: values only, in descending order
        std::sort(S + nres, S + sc.mn, std::greater<double>());
        gram_res[si] = std::sqrt(kGramRes * lmax) + 1e-15 * std::sqrt(lmax);
      }
      gram_sec[si] = 1;
    }
    DevBuf dcinv(ctx, cinv.size() * 8), dkinv(ctx, kinv.size() * 8);
    SADMRG_HIP(hipMemcpyAsync(dcinv.get(), cinv.data(), cinv.size() * 8, hipMemcpyHostToDevice, ctx->stream()));
    SADMRG_HIP(hipMemcpyAsync(dkinv.get(), kinv.data(), kinv.size() * 8, hipMemcpyHostToDevice, ctx->stream()));
    for (int q = 0; q < cnt; ++q) {
      const Sec& sc = secs[list[q]];
      const double* ci = dcinv.as<double>() + int64_t(g) * q;
      const double* ki = dkinv.as<double>() + int64_t(g) * q;
      if (right_gram) {
        // U = Y[:, desc] / S, V' = V[:, desc]^T
        SADMRG_HIP(launch_copy2d(sc.m, sc.mn, by.as<double>() + sa * q + int64_t(np - 1) * mp, 1, -int64_t(mp),
                                 du.as<double>() + sc.u_off, sc.m, 1, 1.0, ki, 0, ctx->stream()));
        SADMRG_HIP(launch_copy2d(sc.mn, sc.n, bg.as<double>() + sg * q + int64_t(np - 1) * np, -int64_t(np), 1,
                                 dvt.as<double>() + sc.vt_off, sc.mn, 1, 1.0, nullptr, 0, ctx->stream()));
      } else {
        // U = G[:, desc]; V' row k = (U_k^T A) / S_k
        SADMRG_HIP(launch_copy2d(sc.m, sc.mn, bg.as<double>() + sg * q + int64_t(mp - 1) * mp, 1, -int64_t(mp),
                                 du.as<double>() + sc.u_off, sc.m, 1, 1.0, nullptr, 0, ctx->stream()));
        SADMRG_HIP(launch_strided_scale(mp, np, by.as<double>() + sa * q, mp, 1, ci, ctx->stream()));
        SADMRG_HIP(launch_copy2d(sc.mn, sc.n, by.as<double>() + sa * q + (mp - 1), -1, mp,
                                 dvt.as<double>() + sc.vt_off, sc.mn, 1, 1.0, nullptr, 0, ctx->stream()));
      }
    }
    ctx->sync();  // batch buffers are released at the end of this scope
  };
  // Unpadded Gram route on the hand-written batched eigensolver: every
  // sector's Gram matrix (lower triangle by syrk, straight from the block)
  // is tridiagonalised and all its eigenvalues found here; the eigenvectors
  // of the kept values only are computed once the global selection is final
  // (eig_vectors below).  S_k = sqrt(lambda_k) for resolved eigenvalues; an
  // unresolved one is reported as sqrt(max(lambda, 0)), within
  // sqrt(eps lambda_max) of the singular value.
  std::vector<uint8_t> eig_sec(secs.size(), 0);
  std::vector<DevBuf> eig_g, eig_w;
  std::vector<EigJob> eig_jobs;
  std::vector<size_t> eig_of_job;
  std::unique_ptr<SymEigBatch> eig;
  auto eig_batch = [&](const std::vector<size_t>& list) {
    eig.reset(new SymEigBatch(ctx));
    eig_g.clear();
    eig_w.clear();
    eig_jobs.clear();
    eig_of_job.clear();
    const double one = 1.0, zero = 0.0;
    for (size_t si : list) {
      const Sec& sc = secs[si];
      const int g = right_gram ? sc.n : sc.m;
      eig_g.emplace_back(ctx, int64_t(g) * g * 8);
      eig_w.emplace_back(ctx, int64_t(g) * 8);
      const double* A = vec.block_ptr(sc.i + sc.j * nm1);
      rocblas_status st = right_gram
          ? rocblas_dsyrk(h, rocblas_fill_lower, rocblas_operation_transpose, g, sc.m, &one, A, sc.m, &zero,
                          eig_g.back().as<double>(), g)
          : rocblas_dsyrk(h, rocblas_fill_lower, rocblas_operation_none, g, sc.n, &one, A, sc.m, &zero,
                          eig_g.back().as<double>(), g);
      if (st != rocblas_status_success) fail(kErrNumerics, "rocblas syrk (Gram) failed");
      EigJob J;
      J.a = eig_g.back().as<double>();
      J.n = g;
      J.lda = g;
      J.w = eig_w.back().as<double>();
      eig_jobs.push_back(J);
      eig_of_job.push_back(si);
    }
    eig->eigenvalues(eig_jobs);
    std::vector<std::vector<double>> lamv(list.size());
    for (size_t q = 0; q < list.size(); ++q) {
      lamv[q].resize(eig_jobs[q].n);
      SADMRG_HIP(hipMemcpyAsync(lamv[q].data(), eig_jobs[q].w, eig_jobs[q].n * 8, hipMemcpyDeviceToHost,
                                ctx->stream()));
    }
    ctx->sync();
    for (size_t q = 0; q < list.size(); ++q) {
      const size_t si = list[q];
      const Sec& sc = secs[si];
      std::vector<double>& lam = lamv[q];
      const int g = eig_jobs[q].n;
      std::sort(lam.begin(), lam.end());  // bisection values are monotone up to its tolerance
      const double lmax = std::max(lam[g - 1], 0.0);
      double* S = gram_s.data() + sc.s_off;
      int nres = 0;
      for (int k = 0; k < sc.mn; ++k) {
        const double l = lam[g - 1 - k];
        const bool resolved = l > kGramRes * lmax && l > 0.0;
        S[k] = std::sqrt(std::max(l, 0.0));
        if (resolved) nres = k + 1;
      }
      gram_nres[si] = nres;
      gram_res[si] = nres < sc.mn ? std::sqrt(kGramRes * lmax) + 1e-15 * std::sqrt(lmax) : 0.0;
      gram_sec[si] = 1;
      eig_sec[si] = 1;
      SADMRG_HIP(hipMemcpyAsync(eig_jobs[q].w, lam.data(), g * 8, hipMemcpyHostToDevice, ctx->stream()));
    }
    ctx->sync();
  };
  // Decomposes the listed sectors: groups of similar size, largest first.
  // The Gram route's eigensolver is launch-latency bound (one chain of small
  // kernels per matrix column, batched over the matrices), so a step whose
  // largest sector qualifies sends every sector through one padded batch.
  constexpr int kMaxGramBatch = 64;
  auto decompose = [&](std::vector<size_t> list, bool allow_gram) {
    std::stable_sort(list.begin(), list.end(), [&](size_t x, size_t y) {
      return std::max(secs[x].m, secs[x].n) > std::max(secs[y].m, secs[y].n);
    });
    const auto td = std::chrono::steady_clock::now();
    if (!list.empty())
      g_svd_stats.max_sector = std::max<int64_t>(g_svd_stats.max_sector, std::max(secs[list[0]].m, secs[list[0]].n));
    if (allow_gram && gram && !list.empty() && std::max(secs[list[0]].m, secs[list[0]].n) >= gmin &&
        svd_eigensolver() == 1) {
      bool fits = true;
      for (size_t si : list) fits = fits && (right_gram ? secs[si].n : secs[si].m) <= SymEigBatch::max_order();
      if (fits) {
        eig_batch(list);
        g_svd_stats.gram_sectors += static_cast<int64_t>(list.size());
        g_svd_stats.t_gram += std::chrono::duration<double>(std::chrono::steady_clock::now() - td).count();
        return;
      }
    }
    if (allow_gram && gram && !list.empty() && std::max(secs[list[0]].m, secs[list[0]].n) >= gmin) {
      struct Tm {
        std::chrono::steady_clock::time_point t0;
        ~Tm() { g_svd_stats.t_gram += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
      } tm{td};
      for (size_t k0 = 0; k0 < list.size(); k0 += kMaxGramBatch) {
        std::vector<size_t> part(list.begin() + k0, list.begin() + std::min(list.size(), k0 + kMaxGramBatch));
        int mp = 0, np = 0;
        for (size_t si : part) {
          mp = std::max(mp, secs[si].m);
          np = std::max(np, secs[si].n);
        }
        gram_batch(part, mp, np);
        g_svd_stats.gram_sectors += static_cast<int64_t>(part.size());
      }
      return;
    }
    for (size_t k0 = 0; k0 < list.size();) {
      const int lead = std::max(secs[list[k0]].m, secs[list[k0]].n);
      size_t k1 = k0 + 1;
      while (k1 < list.size() && static_cast<int>(k1 - k0) < kMaxBatch &&
             10 * std::max(secs[list[k1]].m, secs[list[k1]].n) >= 6 * lead)
        ++k1;
      std::vector<size_t> part(list.begin() + k0, list.begin() + k1);
      int mp = 0, np = 0;
      for (size_t si : part) {
        mp = std::max(mp, secs[si].m);
        np = std::max(np, secs[si].n);
      }
      if (part.size() == 1) single(part[0]);
      else svd_batch(part, mp, np);
      g_svd_stats.svd_sectors += static_cast<int64_t>(part.size());
      k0 = k1;
    }
    ctx->sync();
    g_svd_stats.t_gesdd += std::chrono::duration<double>(std::chrono::steady_clock::now() - td).count();
  };
  // S of every sector on the host; padded sectors whose largest value is
  // round-off and Gram sectors without a resolved value are redone unpadded.
  std::vector<double> hs(std::max<int64_t>(s_tot, 1));
  auto finish = [&]() {
    std::vector<double> lead(secs.size(), 0.0);
    for (size_t si = 0; si < secs.size(); ++si)
      if (padded[si])
        SADMRG_HIP(hipMemcpyAsync(&lead[si], ds.as<double>() + secs[si].s_off, 8, hipMemcpyDeviceToHost,
                                  ctx->stream()));
    if (jac_pending)
      SADMRG_HIP(hipMemcpyAsync(jf.data(), jflags.get(), jac.size() * sizeof(int), hipMemcpyDeviceToHost,
                                ctx->stream()));
    ctx->sync();
    if (jac_pending) {
      // a Jacobi sector that did not converge is decomposed by gesdd
      jac_pending = false;
      for (size_t t = 0; t < jac.size(); ++t)
        if (jf[t]) {
          single(jac[t]);
          g_svd_stats.jacobi_redo++;
        }
    }
    for (size_t si = 0; si < secs.size(); ++si)
      if (padded[si] && lead[si] <= 1e-16) {
        padded[si] = 0;
        single(si);
      }
    for (size_t si = 0; si < secs.size(); ++si)
      if (gram_sec[si] && gram_nres[si] == 0) {
        gram_sec[si] = 0;
        gram_res[si] = 0.0;
        eig_sec[si] = 0;
        single(si);
      }
    for (size_t si = 0; si < secs.size(); ++si)
      if (gram_sec[si]) {
        SADMRG_HIP(hipMemcpyAsync(ds.as<double>() + secs[si].s_off, gram_s.data() + secs[si].s_off,
                                  int64_t(secs[si].mn) * 8, hipMemcpyHostToDevice, ctx->stream()));
        SADMRG_HIP(hipMemsetAsync(dinfo.as<int>() + si, 0, sizeof(int), ctx->stream()));
      }
    std::vector<int> hinfo(std::max<size_t>(secs.size(), 1));
    SADMRG_HIP(hipMemcpyAsync(hs.data(), ds.get(), s_tot * 8, hipMemcpyDeviceToHost, ctx->stream()));
    SADMRG_HIP(hipMemcpyAsync(hinfo.data(), dinfo.get(), secs.size() * sizeof(int), hipMemcpyDeviceToHost,
                              ctx->stream()));
    ctx->sync();
    for (size_t si = 0; si < secs.size(); ++si)
      if (hinfo[si] != 0) fail(kErrNumerics, "Dense::svd: gesdd info=" + std::to_string(hinfo[si]));
  };
  // kept per sector: S/S_max > 1e-10, at least one, at most max_dcut
  std::vector<int> dc(nm1, 0);
  struct SV { double v; int sec; int idx; };
  std::vector<SV> all;
  auto select = [&]() {
    all.clear();
    for (size_t si = 0; si < secs.size(); ++si) {
      const Sec& s = secs[si];
      const double* S = hs.data() + s.s_off;
      const double smax = S[0] > 1e-16 ? S[0] : 1.0;
      int kept = 0;
      for (int k = 0; k < s.mn; ++k)
        if (S[k] / smax > 1e-10) ++kept;
      if (kept == 0) kept = 1;
      if (max_dcut > 0 && kept > max_dcut) kept = max_dcut;
      dc[s.i] = kept;
      for (int k = 0; k < kept; ++k) all.push_back(SV{S[k], s.i, k});
    }
    // global descending order; ties broken by sector then index (deterministic)
    std::stable_sort(all.begin(), all.end(), [](const SV& a, const SV& b) { return a.v > b.v; });
    // Values equal to within the decompositions' rounding (1e-12 relative)
    // are ordered as exact ties, by sector then index, so that the cut
    // through a degenerate multiplet (symmetry-related sectors) does not
    // depend on which SVD routine produced the last bits.
    for (size_t a = 0; a < all.size();) {
      size_t b = a + 1;
      while (b < all.size() && all[a].v - all[b].v <= 1e-12 * all[a].v) ++b;
      if (b - a > 1)
        std::stable_sort(all.begin() + a, all.begin() + b, [&](const SV& x, const SV& y) {
          return sec_of[x.sec] != sec_of[y.sec] ? sec_of[x.sec] < sec_of[y.sec] : x.idx < y.idx;
        });
      a = b;
    }
  };
  if (!jac.empty()) {
    const auto tj = std::chrono::steady_clock::now();
    std::vector<JacobiSvdJob> jobs(jac.size());
    int words = 0;
    for (size_t t = 0; t < jac.size(); ++t) {
      const Sec& sc = secs[jac[t]];
      JacobiSvdJob& J = jobs[t];
      J.a = vec.block_ptr(sc.i + sc.j * nm1);
      J.m = sc.m;
      J.n = sc.n;
      J.lda = sc.m;
      J.s = ds.as<double>() + sc.s_off;
      J.u = du.as<double>() + sc.u_off;
      J.ldu = sc.m;
      J.vt = dvt.as<double>() + sc.vt_off;
      J.ldvt = sc.mn;
      const int qe = sc.mn + (sc.mn & 1);
      words = std::max(words, std::max(sc.m, sc.n) * qe + qe * qe);
    }
    jjobs = DevBuf(ctx, jobs.size() * sizeof(JacobiSvdJob));
    ctx->upload(jjobs.get(), jobs.data(), jobs.size() * sizeof(JacobiSvdJob));
    jflags = DevBuf(ctx, jac.size() * sizeof(int));
    SADMRG_HIP(launch_jacobi_svd(jjobs.as<JacobiSvdJob>(), static_cast<int>(jac.size()), words, kJacobiSweeps,
                                 jflags.as<int>(), ctx->stream()));
    jac_pending = true;
    g_svd_stats.jacobi_sectors += static_cast<int64_t>(jac.size());
    g_svd_stats.t_jacobi += std::chrono::duration<double>(std::chrono::steady_clock::now() - tj).count();
  }
  decompose(rest, true);
  finish();
  select();
  {
    // Gram sectors are accepted when the kept set cannot depend on their
    // unresolved values: the cap max_dcut is binding and the cut (the
    // smallest kept value) lies above twice the sector's resolution limit,
    // which bounds every unresolved value.
    const int n0 = static_cast<int>(all.size());
    const bool binding = max_dcut > 0 && n0 > max_dcut;
    const double vcut = binding ? all[max_dcut - 1].v : 0.0;
    std::vector<size_t> redo;
    for (size_t si = 0; si < secs.size(); ++si)
      if (gram_sec[si] && gram_res[si] > 0.0 && !(binding && vcut > 2.0 * gram_res[si])) redo.push_back(si);
    if (!redo.empty()) {
      g_svd_stats.gram_redo += static_cast<int64_t>(redo.size());
      for (size_t si : redo) {
        gram_sec[si] = 0;
        gram_res[si] = 0.0;
        eig_sec[si] = 0;
      }
      decompose(redo, false);
      finish();
      select();
    }
  }
  const int n = static_cast<int>(all.size());
  const int dcut = n > max_dcut ? max_dcut : n;
  std::vector<int> dim(nm1, 0);
  for (int k = 0; k < dcut; ++k) dim[all[k].sec]++;
  wout.assign(n, 0.0);
  int mk = 0, md = dcut;
  std::vector<int> kept_q, kept_d;
  for (int i = 0; i < nm1; ++i) {
    if (sec_of[i] < 0) continue;
    const double* S = hs.data() + secs[sec_of[i]].s_off;
    for (int k = 0; k < dim[i]; ++k) wout[mk + k] = S[k];
    for (int k = dim[i]; k < dc[i]; ++k) wout[md + k - dim[i]] = S[k];
    mk += dim[i];
    md += dc[i] - dim[i];
    if (dim[i] > 0) { kept_q.push_back(bb0.q[i]); kept_d.push_back(dim[i]); }
  }
  if (kept_out) *kept_out = dim;
  // Eigen route: eigenvectors of the kept values of each accepted sector, then
  // the other factor.  Right Gram: V = eigenvectors of psi^T psi (n x k),
  // U = psi V / S.  Left Gram: U = eigenvectors of psi psi^T (m x k), written
  // straight into U, V' rows = U^T psi / S.  Only the first dim columns of U /
  // rows of V' are filled: the projectors take no more.
  if (eig) {
    const auto tv = std::chrono::steady_clock::now();
    struct Tm {
      std::chrono::steady_clock::time_point t0;
      ~Tm() { g_svd_stats.t_vectors += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
    } tm{tv};
    std::vector<EigJob> jobs = eig_jobs;
    std::vector<DevBuf> zb(jobs.size());
    bool any = false;
    for (size_t q = 0; q < jobs.size(); ++q) {
      const size_t si = eig_of_job[q];
      const Sec& sc = secs[si];
      const int k = (eig_sec[si] && gram_sec[si]) ? dim[sc.i] : 0;
      jobs[q].k = k;
      if (k == 0) continue;
      any = true;
      if (right_gram) {
        zb[q] = DevBuf(ctx, int64_t(sc.n) * k * 8);
        jobs[q].z = zb[q].as<double>();
        jobs[q].ldz = sc.n;
      } else {
        jobs[q].z = du.as<double>() + sc.u_off;
        jobs[q].ldz = sc.m;
      }
    }
    if (any) {
      eig->top_vectors(jobs);
      std::vector<double> inv(s_tot, 0.0);
      for (size_t q = 0; q < jobs.size(); ++q) {
        const Sec& sc = secs[eig_of_job[q]];
        for (int c = 0; c < jobs[q].k; ++c) {
          const double sv = hs[sc.s_off + c];
          inv[sc.s_off + c] = sv > 0.0 ? 1.0 / sv : 0.0;
        }
      }
      DevBuf dinv(ctx, std::max<int64_t>(s_tot, 1) * 8);
      SADMRG_HIP(hipMemcpyAsync(dinv.get(), inv.data(), s_tot * 8, hipMemcpyHostToDevice, ctx->stream()));
      const double one = 1.0, zero = 0.0;
      for (size_t q = 0; q < jobs.size(); ++q) {
        const int k = jobs[q].k;
        if (k == 0) continue;
        const Sec& sc = secs[eig_of_job[q]];
        const double* A = vec.block_ptr(sc.i + sc.j * nm1);
        const double* iv = dinv.as<double>() + sc.s_off;
        if (right_gram) {
          // V' rows c = V column c; U = A V / S
          SADMRG_HIP(launch_copy2d(k, sc.n, jobs[q].z, sc.n, 1, dvt.as<double>() + sc.vt_off, sc.mn, 1, 1.0, nullptr,
                                   0, ctx->stream()));
          if (rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, sc.m, k, sc.n, &one, A, sc.m, jobs[q].z,
                            sc.n, &zero, du.as<double>() + sc.u_off, sc.m) != rocblas_status_success)
            fail(kErrNumerics, "rocblas gemm (Gram) failed");
          SADMRG_HIP(launch_strided_scale(k, sc.m, du.as<double>() + sc.u_off, 1, sc.m, iv, ctx->stream()));
        } else {
          if (rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, k, sc.n, sc.m, &one,
                            du.as<double>() + sc.u_off, sc.m, A, sc.m, &zero, dvt.as<double>() + sc.vt_off,
                            sc.mn) != rocblas_status_success)
            fail(kErrNumerics, "rocblas gemm (Gram) failed");
          SADMRG_HIP(launch_strided_scale(k, sc.n, dvt.as<double>() + sc.vt_off, sc.mn, 1, iv, ctx->stream()));
        }
      }
      ctx->sync();  // z buffers and the scale table leave scope
    }
  }
  // projectors: proju = (bb0^-1, new), projv = (bb1^-1, new')
  const Bond nb = Bond::make(bb0.dir, kept_q, kept_d);
  Bond nbv = nb;
  if (sym == SymKind::U1) nbv.dir = -nbv.dir;
  Struct su, sv;
  su.sym = sv.sym = sym;
  su.locspin = sv.locspin = 0;
  su.bonds = {bb0.inverted(), nb};
  sv.bonds = {bb1.inverted(), nbv};
  std::vector<uint8_t> mu(su.nten(), 0), mv(sv.nten(), 0);
  // the k-th allowed block of each projector takes the k-th kept sector
  std::vector<int> ublk, vblk;
  for (int b = 0; b < su.nten(); ++b)
    if (su.allowed(b)) ublk.push_back(b);
  for (int b = 0; b < sv.nten(); ++b)
    if (sv.allowed(b)) vblk.push_back(b);
  std::vector<int> kept_secs;
  for (int i = 0; i < nm1; ++i)
    if (dim[i] > 0) kept_secs.push_back(i);
  require(ublk.size() >= kept_secs.size() && vblk.size() >= kept_secs.size(), "svd: projector blocks");
  for (size_t k = 0; k < kept_secs.size(); ++k) { mu[ublk[k]] = 1; mv[vblk[k]] = 1; }
  BTensor pu, pv;
  pu.layout(ctx, cx, su, mu, kCgcIdent);
  pv.layout(ctx, cx, sv, mv, kCgcIdent);
  // column scales S^p
  std::vector<double> cs;
  std::vector<int64_t> cs_off_u(nm1, 0), cs_off_v(nm1, 0);
  for (int i : kept_secs) {
    const double* S = hs.data() + secs[sec_of[i]].s_off;
    cs_off_u[i] = static_cast<int64_t>(cs.size());
    for (int k = 0; k < dim[i]; ++k) cs.push_back(p1 != 0.0 ? std::pow(S[k], p1) : 1.0);
    cs_off_v[i] = static_cast<int64_t>(cs.size());
    for (int k = 0; k < dim[i]; ++k) cs.push_back(p2 != 0.0 ? std::pow(S[k], p2) : 1.0);
  }
  DevBuf dcs(ctx, std::max<size_t>(cs.size(), 1) * 8);
  if (!cs.empty())
    SADMRG_HIP(hipMemcpyAsync(dcs.get(), cs.data(), cs.size() * 8, hipMemcpyHostToDevice, ctx->stream()));
  for (size_t k = 0; k < kept_secs.size(); ++k) {
    const int i = kept_secs[k];
    const Sec& s = secs[sec_of[i]];
    // U block (m x dim): first dim columns of U scaled by S^p1
    SADMRG_HIP(launch_copy2d(s.m, dim[i], du.as<double>() + s.u_off * w, 1, s.m, pu.block_ptr(ublk[k]), s.m, w, 1.0,
                             dcs.as<double>() + cs_off_u[i], 0, ctx->stream()));
    // V block (n x dim): V(j, c) = VT(c, j) * S_c^p2 (no conjugation, dense_inst.cpp:531-535)
    SADMRG_HIP(launch_copy2d(s.n, dim[i], dvt.as<double>() + s.vt_off * w, s.mn, 1, pv.block_ptr(vblk[k]), s.n, w,
                             1.0, dcs.as<double>() + cs_off_v[i], 0, ctx->stream()));
  }
  BTensor left, rght;
  op_fuse(left, ctx, cx, sym, bb2_0, bb2_1);
  op_fuse(rght, ctx, cx, sym, bb2_2, bb2_3);
  op_contract(tu, left, 2, pu, 0);
  op_contract(tv, rght, 2, pv, 0);
  ctx->sync();
  g_svd_stats.seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
}

}  // namespace sadmrg
