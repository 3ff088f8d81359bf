// capi.cpp -- C ABI (include/sadmrg.h) over the device engine.
#include "../../include/sadmrg.h"
#include <cstring>
#include <exception>
#include <memory>
#include "btensor.hpp"
#include "dmrg.hpp"
#include "eigen.hpp"
#include "heff.hpp"
#include "lanczos.hpp"

using namespace sadmrg;

struct sadmrg_ctx {
  std::unique_ptr<Context> c;
};
struct sadmrg_cgc {
  std::unique_ptr<CgcTables> t;
};
struct sadmrg_tensor {
  Context* ctx;
  SymKind sym;
  bool cplx;
  BTensor t;
};
struct sadmrg_heff {
  std::unique_ptr<HeffPlan> p;
};
struct sadmrg_dmrg {
  Context* ctx;
  std::unique_ptr<Dmrg> d;
};

namespace {

template <class F>
int guard(F&& f) {
  try {
    f();
    return SADMRG_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    set_last_error("host allocation failed");
    return SADMRG_ERR_INVALID;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return SADMRG_ERR_INVALID;
  }
}

void need(const void* p, const char* what) {
  if (!p) fail(kErrInvalid, std::string("null argument: ") + what);
}

SymKind sym_of(int s) {
  if (s != SADMRG_U1 && s != SADMRG_SU2) fail(kErrInvalid, "symmetry must be SADMRG_U1 or SADMRG_SU2");
  return s == SADMRG_SU2 ? SymKind::SU2 : SymKind::U1;
}

Prec prec_of(int p) {
  if (p != SADMRG_FP64 && p != SADMRG_FP32 && p != SADMRG_BF16)
    fail(kErrInvalid, "precision must be SADMRG_FP64, SADMRG_FP32 or SADMRG_BF16");
  return p == SADMRG_FP32 ? Prec::FP32 : (p == SADMRG_BF16 ? Prec::BF16 : Prec::FP64);
}

Bond bond_of(const int* b) {
  need(b, "bond");
  const int nm = b[1];
  if (nm < 0) fail(kErrInvalid, "bond: negative sector count");
  return Bond::make(b[0], std::vector<int>(b + 2, b + 2 + nm), std::vector<int>(b + 2 + nm, b + 2 + 2 * nm));
}

void same_kind(const sadmrg_tensor* a, const sadmrg_tensor* b) {
  if (a->sym != b->sym || a->cplx != b->cplx || a->ctx != b->ctx)
    fail(kErrInvalid, "tensors differ in symmetry, scalar type or context");
}

CgcTables* cg_of(sadmrg_cgc* c, SymKind s) {
  if (s == SymKind::SU2) {
    need(c, "cgc (required for SU(2))");
    return c->t.get();
  }
  return c ? c->t.get() : nullptr;
}

void fix_kind(sadmrg_tensor* t) {
  // ops rebuild the BTensor; keep the handle's symmetry consistent
  if (!t->t.is_null()) t->t.st.sym = t->sym;
  t->t.ctx = t->ctx;
  t->t.cplx = t->cplx;
}

}  // namespace

extern "C" {

const char* sadmrg_last_error(void) { return last_error(); }

int sadmrg_version(int* major, int* minor) {
  return guard([&] {
    need(major, "major");
    need(minor, "minor");
    *major = 0;
    *minor = 1;
  });
}

int sadmrg_ctx_create(int device, sadmrg_ctx** out) {
  return guard([&] {
    need(out, "out");
    auto c = std::make_unique<sadmrg_ctx>();
    c->c = std::make_unique<Context>(device);
    *out = c.release();
  });
}

int sadmrg_ctx_destroy(sadmrg_ctx* ctx) {
  return guard([&] { delete ctx; });
}

int sadmrg_ctx_sync(sadmrg_ctx* ctx) {
  return guard([&] {
    need(ctx, "ctx");
    ctx->c->sync();
  });
}

int sadmrg_ctx_stats(sadmrg_ctx* ctx, long long* mallocs, double* malloc_seconds, long long* syncs,
                     double* sync_seconds) {
  return guard([&] {
    need(ctx, "ctx");
    if (mallocs) *mallocs = ctx->c->mallocs();
    if (malloc_seconds) *malloc_seconds = ctx->c->malloc_seconds();
    if (syncs) *syncs = ctx->c->syncs();
    if (sync_seconds) *sync_seconds = ctx->c->sync_seconds();
  });
}

int sadmrg_ctx_memory(sadmrg_ctx* ctx, long long* in_use, long long* cached) {
  return guard([&] {
    need(ctx, "ctx");
    if (in_use) *in_use = static_cast<long long>(ctx->c->bytes_in_use());
    if (cached) *cached = static_cast<long long>(ctx->c->bytes_cached());
  });
}

int sadmrg_ctx_stream(sadmrg_ctx* ctx, void** stream) {
  return guard([&] {
    need(ctx, "ctx");
    need(stream, "stream");
    *stream = static_cast<void*>(ctx->c->stream());
  });
}

int sadmrg_clebsch_gordan(int j1, int m1, int j2, int m2, int j3, int m3, double* out) {
  return guard([&] {
    need(out, "out");
    *out = clebsch_gordan(j1, m1, j2, m2, j3, m3);
  });
}

int sadmrg_cgc_create(int physpn, sadmrg_cgc** out) {
  return guard([&] {
    need(out, "out");
    auto c = std::make_unique<sadmrg_cgc>();
    c->t = std::make_unique<CgcTables>(physpn);
    *out = c.release();
  });
}

int sadmrg_cgc_destroy(sadmrg_cgc* cgc) {
  return guard([&] { delete cgc; });
}

int sadmrg_cgc_hamilt_vec(sadmrg_cgc* c, int i, int j, int k, double* out) {
  return guard([&] {
    need(c, "cgc");
    need(out, "out");
    *out = c->t->hamilt_vec(i, j, k);
  });
}

int sadmrg_cgc_onsite(sadmrg_cgc* c, int side, int j, int m, int l, int i, int p, double* out) {
  return guard([&] {
    need(c, "cgc");
    need(out, "out");
    *out = c->t->onsite(side, j, m, l, i, p);
  });
}

int sadmrg_cgc_transformation(sadmrg_cgc* c, int side, int j, int m, int l, int i, int k, int p, double* out) {
  return guard([&] {
    need(c, "cgc");
    need(out, "out");
    *out = c->t->transformation(side, j, m, l, i, k, p);
  });
}

int sadmrg_cgc_pairup(sadmrg_cgc* c, int side, int j, int i, int k, int m, int p, double* out) {
  return guard([&] {
    need(c, "cgc");
    need(out, "out");
    *out = c->t->pairup(side, j, i, k, m, p);
  });
}

int sadmrg_tensor_create(sadmrg_ctx* ctx, int sym, int cplx, const int* header, int header_len,
                         const unsigned char* present, const double* data, int cgc2_kind, sadmrg_tensor** out) {
  return guard([&] {
    need(ctx, "ctx");
    need(header, "header");
    need(out, "out");
    auto t = std::make_unique<sadmrg_tensor>();
    t->ctx = ctx->c.get();
    t->sym = sym_of(sym);
    t->cplx = cplx != 0;
    Struct s = parse_header(t->sym, header, header_len);
    if (s.nbond() > 0) {
      need(present, "present");
      int64_t n = 0;
      for (int b = 0; b < s.nten(); ++b)
        if (present[b] && s.allowed(b)) n += s.block_size(b);
      if (n > 0) need(data, "data");
    }
    tensor_from_host(t->t, t->ctx, t->sym, t->cplx, header, header_len, present, data, cgc2_kind);
    fix_kind(t.get());
    *out = t.release();
  });
}

int sadmrg_tensor_null(sadmrg_ctx* ctx, int sym, int cplx, sadmrg_tensor** out) {
  return guard([&] {
    need(ctx, "ctx");
    need(out, "out");
    auto t = std::make_unique<sadmrg_tensor>();
    t->ctx = ctx->c.get();
    t->sym = sym_of(sym);
    t->cplx = cplx != 0;
    fix_kind(t.get());
    *out = t.release();
  });
}

int sadmrg_tensor_destroy(sadmrg_tensor* t) {
  return guard([&] { delete t; });
}

int sadmrg_tensor_header(const sadmrg_tensor* t, int* header, int cap, int* len) {
  return guard([&] {
    need(t, "tensor");
    const std::vector<int> h = make_header(t->t.st);
    if (len) *len = static_cast<int>(h.size());
    if (header) {
      if (cap < static_cast<int>(h.size())) fail(kErrInvalid, "header buffer too small");
      std::memcpy(header, h.data(), h.size() * sizeof(int));
    }
  });
}

int sadmrg_tensor_nten(const sadmrg_tensor* t, int* nten) {
  return guard([&] {
    need(t, "tensor");
    need(nten, "nten");
    *nten = t->t.nten();
  });
}

int sadmrg_tensor_nelem(const sadmrg_tensor* t, long long* n) {
  return guard([&] {
    need(t, "tensor");
    need(n, "n");
    *n = t->t.nelem;
  });
}

int sadmrg_tensor_download(const sadmrg_tensor* t, unsigned char* present, double* data) {
  return guard([&] {
    need(t, "tensor");
    if (t->t.is_null()) return;
    need(present, "present");
    tensor_to_host(t->t, present, data);
  });
}

int sadmrg_tensor_copy(sadmrg_tensor* dst, const sadmrg_tensor* src) {
  return guard([&] {
    need(dst, "dst");
    need(src, "src");
    same_kind(dst, src);
    copy_tensor(dst->t, src->t);
    fix_kind(dst);
  });
}

int sadmrg_tensor_device_ptr(const sadmrg_tensor* t, void** ptr) {
  return guard([&] {
    need(t, "tensor");
    need(ptr, "ptr");
    *ptr = t->t.dptr();
  });
}

int sadmrg_hamiltonian_vector_multiplication(sadmrg_tensor* out, const sadmrg_tensor* vec, const sadmrg_tensor* op1,
                                             const sadmrg_tensor* op2, sadmrg_cgc* cgc) {
  return guard([&] {
    need(out, "out");
    need(vec, "vec");
    need(op1, "op1");
    need(op2, "op2");
    same_kind(out, vec);
    same_kind(vec, op1);
    same_kind(vec, op2);
    BTensor r;
    op_hvm(r, vec->t, op1->t, op2->t, cg_of(cgc, vec->sym));
    out->t = std::move(r);
    fix_kind(out);
  });
}

int sadmrg_contract(sadmrg_tensor* out, const sadmrg_tensor* t1, int i1, const sadmrg_tensor* t2, int i2) {
  return guard([&] {
    need(out, "out");
    need(t1, "t1");
    need(t2, "t2");
    same_kind(t1, t2);
    same_kind(out, t1);
    if (t1->t.is_null() || t2->t.is_null()) fail(kErrInvalid, "contract: null tensor");
    BTensor r;
    op_contract(r, t1->t, i1, t2->t, i2);
    out->t = std::move(r);
    fix_kind(out);
  });
}

int sadmrg_env_update(sadmrg_tensor* out, int kind, const sadmrg_tensor* uu, const sadmrg_tensor* vv,
                      const sadmrg_tensor* op1, const sadmrg_tensor* op2, int flag, sadmrg_cgc* cgc) {
  return guard([&] {
    need(out, "out");
    need(uu, "uu");
    need(vv, "vv");
    same_kind(out, uu);
    same_kind(uu, vv);
    if (kind != 0) {
      need(op1, "op1");
      same_kind(uu, op1);
    }
    if (kind == 4 || kind == 5) {
      need(op2, "op2");
      same_kind(uu, op2);
    }
    if (kind < 0 || kind > 5) fail(kErrInvalid, "env_update: kind must be 0..5");
    BTensor r;
    if (kind == 5) op_env_adjoint(r, op1->t, uu->t, vv->t, op2->t, flag);
    else op_env(r, kind, uu->t, vv->t, op1 ? &op1->t : nullptr, op2 ? &op2->t : nullptr, flag, cg_of(cgc, uu->sym));
    out->t = std::move(r);
    fix_kind(out);
  });
}

int sadmrg_svd(const sadmrg_tensor* vec, sadmrg_tensor* tu, double p1, sadmrg_tensor* tv, double p2, int max_dcut,
               double* wout, int wout_cap, int* nw) {
  return guard([&] {
    need(vec, "vec");
    need(tu, "tu");
    need(tv, "tv");
    same_kind(vec, tu);
    same_kind(vec, tv);
    if (max_dcut < 1) fail(kErrInvalid, "svd: max_dcut must be >= 1");
    std::vector<double> w;
    op_svd(vec->t, tu->t, p1, tv->t, p2, max_dcut, w);
    fix_kind(tu);
    fix_kind(tv);
    if (nw) *nw = static_cast<int>(w.size());
    if (wout) {
      for (int k = 0; k < wout_cap && k < static_cast<int>(w.size()); ++k) wout[k] = w[k];
    }
  });
}

int sadmrg_fuse(sadmrg_tensor* out, const int* b1, const int* b2) {
  return guard([&] {
    need(out, "out");
    op_fuse(out->t, out->ctx, out->cplx, out->sym, bond_of(b1), bond_of(b2));
    fix_kind(out);
  });
}

int sadmrg_fuse_to_singlet(sadmrg_tensor* out, const int* b1, const int* b2) {
  return guard([&] {
    need(out, "out");
    op_fuse_to_singlet(out->t, out->ctx, out->cplx, out->sym, bond_of(b1), bond_of(b2));
    fix_kind(out);
  });
}

int sadmrg_make_spinor_start(sadmrg_tensor* out, int physpn) {
  return guard([&] {
    need(out, "out");
    op_make_spinor_start(out->t, out->ctx, out->cplx, out->sym, physpn);
    fix_kind(out);
  });
}

int sadmrg_tensor_transform(sadmrg_tensor* t, int op, double a, double b) {
  return guard([&] {
    need(t, "tensor");
    BTensor& x = t->t;
    if (x.is_null() && op != 6 && op != 4 && op != 8) fail(kErrInvalid, "transform: null tensor");
    switch (op) {
      case 0: op_vectran(x, true); break;
      case 1: op_vectran(x, false); break;
      case 2: op_take_conjugate(x, static_cast<int>(a)); break;
      case 3: op_makeup_input_vector(x); break;
      case 4: op_scale(x, a, b); break;
      case 5: op_normalize(x); break;
      case 6: op_random_init(x); break;
      case 7: op_shift(x, static_cast<int>(a), static_cast<int>(b)); break;
      case 8: op_fill(x, a); break;
      default: fail(kErrInvalid, "transform: unknown op");
    }
    t->ctx->sync();
    fix_kind(t);
  });
}

int sadmrg_tensor_axpy(sadmrg_tensor* y, double re, double im, const sadmrg_tensor* x) {
  return guard([&] {
    need(y, "y");
    need(x, "x");
    same_kind(y, x);
    op_axpy(y->t, re, im, x->t);
    y->ctx->sync();
    fix_kind(y);
  });
}

int sadmrg_tensor_dot(const sadmrg_tensor* x, const sadmrg_tensor* y, double* out2) {
  return guard([&] {
    need(x, "x");
    need(y, "y");
    need(out2, "out");
    same_kind(x, y);
    const dcmplex v = op_dot(x->t, y->t);
    out2[0] = v.real();
    out2[1] = v.imag();
  });
}

int sadmrg_rng_seed(unsigned long long seed) {
  return guard([&] { reset_global_rng(seed); });
}

int sadmrg_heff_create(sadmrg_ctx* ctx, sadmrg_cgc* cgc, const sadmrg_tensor* hl, const sadmrg_tensor* hr, int nterm,
                       const sadmrg_tensor* const* o1, const sadmrg_tensor* const* o2,
                       const sadmrg_tensor* psi_layout, sadmrg_heff** out) {
  return guard([&] {
    need(ctx, "ctx");
    need(psi_layout, "psi_layout");
    need(out, "out");
    if (nterm < 0) fail(kErrInvalid, "heff: negative term count");
    if (nterm > 0) {
      need(o1, "o1");
      need(o2, "o2");
    }
    static const BTensor kNull;
    std::vector<const BTensor*> a, b;
    for (int t = 0; t < nterm; ++t) {
      need(o1[t], "o1[t]");
      need(o2[t], "o2[t]");
      same_kind(o1[t], psi_layout);
      same_kind(o2[t], psi_layout);
      a.push_back(&o1[t]->t);
      b.push_back(&o2[t]->t);
    }
    if (hl) same_kind(hl, psi_layout);
    if (hr) same_kind(hr, psi_layout);
    auto h = std::make_unique<sadmrg_heff>();
    h->p = std::make_unique<HeffPlan>(ctx->c.get(), cg_of(cgc, psi_layout->sym), hl ? hl->t : kNull,
                                      hr ? hr->t : kNull, a, b, psi_layout->t);
    *out = h.release();
  });
}

int sadmrg_heff_destroy(sadmrg_heff* h) {
  return guard([&] { delete h; });
}

int sadmrg_heff_merge(const sadmrg_heff* const* parts, int n, sadmrg_heff** out) {
  return guard([&] {
    need(parts, "parts");
    need(out, "out");
    if (n < 1) fail(kErrInvalid, "heff_merge: at least one plan");
    std::vector<const HeffPlan*> ps;
    for (int i = 0; i < n; ++i) {
      need(parts[i], "parts[i]");
      ps.push_back(parts[i]->p.get());
    }
    auto h = std::make_unique<sadmrg_heff>();
    h->p = HeffPlan::merge(ps);
    *out = h.release();
  });
}

int sadmrg_heff_apply(sadmrg_heff* h, const sadmrg_tensor* in, sadmrg_tensor* out, int precision) {
  return guard([&] {
    need(h, "heff");
    if (h->p->merged()) fail(kErrInvalid, "heff_apply: a merged plan takes device vectors (heff_apply_device)");
    need(in, "psi_in");
    need(out, "psi_out");
    const BTensor& L = h->p->layout();
    if (in->t.st != L.st || in->t.off != L.off) fail(kErrInvalid, "heff_apply: psi_in layout differs from the plan");
    if (out->t.st != L.st || out->t.off != L.off) {
      std::vector<uint8_t> mask(L.nten());
      for (int b = 0; b < L.nten(); ++b) mask[b] = L.present(b);
      out->t.clean();
      out->t.layout(L.ctx, L.cplx, L.st, mask, kCgcSinglet);
      fix_kind(out);
    }
    h->p->apply(in->t.dptr(), out->t.dptr(), prec_of(precision));
    h->p->ctx()->sync();
  });
}

int sadmrg_heff_apply_device(sadmrg_heff* h, const void* d_in, void* d_out, int precision) {
  return guard([&] {
    need(h, "heff");
    need(d_in, "d_in");
    need(d_out, "d_out");
    h->p->apply(static_cast<const double*>(d_in), static_cast<double*>(d_out), prec_of(precision));
  });
}

int sadmrg_heff_apply_unbatched(sadmrg_heff* h, const void* d_in, void* d_out, int precision) {
  return guard([&] {
    need(h, "heff");
    need(d_in, "d_in");
    need(d_out, "d_out");
    h->p->apply(static_cast<const double*>(d_in), static_cast<double*>(d_out), prec_of(precision), 0, true);
  });
}

int sadmrg_heff_apply_stage(sadmrg_heff* h, const void* d_in, void* d_out, int precision, int stage) {
  return guard([&] {
    need(h, "heff");
    need(d_in, "d_in");
    need(d_out, "d_out");
    if (stage < 0 || stage > 2) fail(kErrInvalid, "heff_apply_stage: stage must be 0, 1 or 2");
    h->p->apply(static_cast<const double*>(d_in), static_cast<double*>(d_out), prec_of(precision), stage);
  });
}

int sadmrg_heff_stage_flops(sadmrg_heff* h, int stage, double* flops) {
  return guard([&] {
    need(h, "heff");
    need(flops, "flops");
    *flops = h->p->stage_flops(stage);
  });
}

int sadmrg_heff_issued_flops(sadmrg_heff* h, int stage, int precision, double* issued) {
  return guard([&] {
    need(h, "heff");
    need(issued, "issued");
    require(stage >= 0 && stage <= 2, "heff_issued_flops: stage must be 0, 1 or 2");
    *issued = h->p->issued_flops(stage, prec_of(precision));
  });
}

int sadmrg_heff_stats(sadmrg_heff* h, double* flops, double* bytes, long long* n, int* info) {
  return guard([&] {
    need(h, "heff");
    const HeffStats& s = h->p->stats();
    if (flops) *flops = s.flops;
    if (bytes) *bytes = s.bytes64;
    if (n) *n = s.n;
    if (info) {
      info[0] = s.nsec_in;
      info[1] = s.stage1_probs;
      info[2] = s.stage2_probs;
      info[3] = s.tiles1;
      info[4] = s.tiles2;
      info[5] = s.tile_cfg1;
      info[6] = s.tile_cfg2;
      info[7] = s.split_tiles;
    }
  });
}

int sadmrg_heff_launch_plan(int plan, int* previous) {
  return guard([&] {
    if (previous) *previous = heff_launch_plan();
    if (plan >= -1) set_heff_launch_plan(plan);
  });
}

int sadmrg_heff_complex_planar(int enable, int* previous) {
  return guard([&] {
    if (previous) *previous = heff_complex_planar() ? 1 : 0;
    if (enable >= 0) set_heff_complex_planar(enable != 0);
  });
}

int sadmrg_heff_per_block(int enable, int* previous) {
  return guard([&] {
    if (previous) *previous = heff_per_block() ? 1 : 0;
    if (enable >= 0) set_heff_per_block(enable != 0);
  });
}

int sadmrg_autotune(int enable, int clear, int* entries) {
  return guard([&] {
    if (enable >= 0) set_autotune(enable != 0);
    if (clear) autotune_clear();
    if (entries) *entries = autotune_entries();
  });
}

int sadmrg_lanczos_mode(int device_steps, int graph, int* prev, long long* counts, int reset) {
  return guard([&] {
    bool ds, gr;
    lanczos_mode(&ds, &gr);
    if (prev) {
      prev[0] = ds ? 1 : 0;
      prev[1] = gr ? 1 : 0;
    }
    set_lanczos_mode(device_steps < 0 ? ds : device_steps != 0, graph < 0 ? gr : graph != 0);
    long long r = 0, c = 0;
    lanczos_graph_counts(&r, &c, reset != 0);
    if (counts) {
      counts[0] = r;
      counts[1] = c;
    }
  });
}

int sadmrg_gemm_vec_staging(int enable, int* previous) {
  return guard([&] {
    if (previous) *previous = GemmBatch::vec_staging() ? 1 : 0;
    if (enable >= 0) GemmBatch::set_vec_staging(enable != 0);
  });
}

int sadmrg_gemm_persistent(int mode, int* previous) {
  return guard([&] {
    if (previous) *previous = GemmBatch::persistent();
    if (mode > 2) fail(kErrInvalid, "gemm_persistent: mode must be -1, 0, 1 or 2");
    if (mode >= 0) GemmBatch::set_persistent(mode);
  });
}

int sadmrg_gemm_balanced_split(int enable, int* previous) {
  return guard([&] {
    if (previous) *previous = GemmBatch::balanced_split() ? 1 : 0;
    if (enable >= 0) GemmBatch::set_balanced_split(enable != 0);
  });
}

int sadmrg_svd_stats(long long* counts, double* seconds, int reset) {
  return guard([&] {
    SvdStats& st = svd_stats();
    if (counts) {
      counts[0] = st.calls;
      counts[1] = st.gram_sectors;
      counts[2] = st.svd_sectors;
      counts[3] = st.gram_redo;
    }
    if (seconds) *seconds = st.seconds;
    if (reset) st = SvdStats{};
  });
}

int sadmrg_eigh_batched(sadmrg_ctx* ctx, int count, const int* n, double* const* a, const int* lda,
                        double* const* w, const int* k, double* const* z, const int* ldz) {
  return guard([&] {
    need(ctx, "ctx");
    if (count < 0) fail(kErrInvalid, "eigh: count must be >= 0");
    if (count == 0) return;
    need(n, "n");
    need(a, "a");
    need(lda, "lda");
    need(w, "w");
    std::vector<EigJob> jobs(count);
    for (int q = 0; q < count; ++q) {
      EigJob& J = jobs[q];
      if (n[q] < 0 || n[q] > SymEigBatch::max_order()) fail(kErrInvalid, "eigh: matrix order out of range");
      if (n[q] > 0 && (a[q] == nullptr || w[q] == nullptr || lda[q] < n[q])) fail(kErrInvalid, "eigh: bad matrix");
      J.a = a[q];
      J.n = n[q];
      J.lda = lda[q];
      J.w = w[q];
      J.k = k ? k[q] : 0;
      if (J.k < 0 || J.k > J.n) fail(kErrInvalid, "eigh: k out of range");
      if (J.k > 0) {
        need(z, "z");
        need(ldz, "ldz");
        if (z[q] == nullptr || ldz[q] < J.n) fail(kErrInvalid, "eigh: bad eigenvector buffer");
        J.z = z[q];
        J.ldz = ldz[q];
      }
    }
    SymEigBatch b(ctx->c.get());
    b.eigenvalues(jobs);
    b.top_vectors(jobs);
  });
}

int sadmrg_eigvalsh_batched_noalloc(sadmrg_ctx* ctx, int count, const int* n, double* const* a, const int* lda,
                                    double* const* w) {
  return guard([&] {
    need(ctx, "ctx");
    if (count < 0) fail(kErrInvalid, "eigvalsh: count must be >= 0");
    if (count == 0) return;
    need(n, "n");
    need(a, "a");
    need(lda, "lda");
    need(w, "w");
    std::vector<EigJob> jobs(count);
    for (int q = 0; q < count; ++q) {
      EigJob& J = jobs[q];
      if (n[q] < 0 || n[q] > SymEigBatch::max_order()) fail(kErrInvalid, "eigvalsh: matrix order out of range");
      if (n[q] > 0 && (a[q] == nullptr || w[q] == nullptr || lda[q] < n[q])) fail(kErrInvalid, "eigvalsh: bad matrix");
      J.a = a[q];
      J.n = n[q];
      J.lda = lda[q];
      J.w = w[q];
    }
    SymEigBatch::eigenvalues_noalloc(ctx->c.get(), jobs);
  });
}

int sadmrg_eigh_clusters(int mode, int* previous, long long* serial_redo) {
  return guard([&] {
    if (previous) *previous = SymEigBatch::cluster_serial() ? 1 : 0;
    if (serial_redo) *serial_redo = static_cast<long long>(SymEigBatch::profile().serial_clusters);
    if (mode == 0 || mode == 1) SymEigBatch::set_cluster_serial(mode == 1);
    else if (mode != -1) fail(kErrInvalid, "eigh clusters: 0, 1 or -1");
  });
}

int sadmrg_svd_eigensolver(int value, int* previous) {
  return guard([&] {
    if (previous) *previous = svd_eigensolver();
    if (value == 0 || value == 1) set_svd_eigensolver(value);
    else if (value != -1) fail(kErrInvalid, "svd eigensolver: 0, 1 or -1");
  });
}

int sadmrg_svd_jacobi(int max_order, int* previous, long long* counts) {
  return guard([&] {
    if (max_order < -1 || max_order > kJacobiMaxOrder)
      fail(kErrInvalid, "svd_jacobi: max_order must be -1 .. " + std::to_string(kJacobiMaxOrder));
    if (previous) *previous = svd_jacobi_max();
    if (counts) {
      counts[0] = svd_stats().jacobi_sectors;
      counts[1] = svd_stats().jacobi_redo;
    }
    if (max_order >= 0) set_svd_jacobi_max(max_order);
  });
}

int sadmrg_svd_gram_min(int value, int* previous) {
  return guard([&] {
    if (value < -1) fail(kErrInvalid, "svd_gram_min: value must be >= -1");
    if (previous) *previous = gram_min();
    if (value >= 0) set_gram_min(value);
  });
}

int sadmrg_lanczos(sadmrg_heff* h, sadmrg_tensor* psi, int mlanc, int precision, int refine, double refine_tol,
                   double* energy, int* counters, double* residual) {
  return guard([&] {
    need(h, "heff");
    need(psi, "psi");
    if (h->p->merged()) fail(kErrInvalid, "lanczos: a merged plan has no single wave function");
    const BTensor& L = h->p->layout();
    if (psi->t.st != L.st || psi->t.off != L.off) fail(kErrInvalid, "lanczos: psi layout differs from the plan");
    if (mlanc < 1 || mlanc > 2047) fail(kErrInvalid, "lanczos: mlanc must be in 1..2047");
    LanczosOpts o;
    o.mlanc = mlanc;
    o.prec = prec_of(precision);
    o.refine = refine != 0;
    if (refine_tol > 0) o.refine_tol = refine_tol;
    const LanczosOut r = lanczos_solve(*h->p, psi->t.dptr(), o);
    h->p->ctx()->sync();
    if (energy) *energy = r.energy;
    if (counters) {
      counters[0] = r.matvec64;
      counters[1] = r.matvec32;
      counters[2] = r.steps;
      counters[3] = r.restarts;
    }
    if (residual) *residual = r.residual64;
  });
}

int sadmrg_lanczos_default_options(sadmrg_lanczos_options* opt) {
  return guard([&] {
    need(opt, "opt");
    const LanczosOpts d;
    opt->mlanc = d.mlanc;
    opt->precision = SADMRG_FP64;
    opt->refine = d.refine ? 1 : 0;
    opt->refine_tol = d.refine_tol;
    opt->refine_krylov = d.refine_krylov;
    opt->max_restarts = d.max_restarts;
    opt->check_every = d.check_every;
    opt->device_steps = d.device_steps ? 1 : 0;
    opt->graph = d.graph ? 1 : 0;
  });
}

int sadmrg_lanczos_ex(sadmrg_heff* h, sadmrg_tensor* psi, const sadmrg_lanczos_options* opt, double* energy,
                      int* counters, double* residual) {
  return guard([&] {
    need(h, "heff");
    need(psi, "psi");
    need(opt, "opt");
    if (h->p->merged()) fail(kErrInvalid, "lanczos: a merged plan has no single wave function");
    const BTensor& L = h->p->layout();
    if (psi->t.st != L.st || psi->t.off != L.off) fail(kErrInvalid, "lanczos: psi layout differs from the plan");
    if (opt->mlanc < 1 || opt->mlanc > 2047) fail(kErrInvalid, "lanczos: mlanc must be in 1..2047");
    if (opt->check_every < 1) fail(kErrInvalid, "lanczos: check_every must be >= 1");
    if (opt->refine_krylov < 1 || opt->max_restarts < 0) fail(kErrInvalid, "lanczos: bad refinement options");
    if (!(opt->refine_tol > 0)) fail(kErrInvalid, "lanczos: refine_tol must be > 0");
    LanczosOpts o;
    o.mlanc = opt->mlanc;
    o.prec = prec_of(opt->precision);
    o.refine = opt->refine != 0;
    o.refine_tol = opt->refine_tol;
    o.refine_krylov = opt->refine_krylov;
    o.max_restarts = opt->max_restarts;
    o.check_every = opt->check_every;
    o.device_steps = opt->device_steps != 0;
    o.graph = opt->graph != 0;
    const LanczosOut r = lanczos_solve(*h->p, psi->t.dptr(), o);
    h->p->ctx()->sync();
    if (energy) *energy = r.energy;
    if (counters) {
      counters[0] = r.matvec64;
      counters[1] = r.matvec32;
      counters[2] = r.steps;
      counters[3] = r.restarts;
    }
    if (residual) *residual = r.residual64;
  });
}

int sadmrg_lanczos_batched(sadmrg_heff* h, void* d_psi, int n_inst, const sadmrg_lanczos_options* opt,
                           double* energies, int* steps, int* restarts, double* residuals, int* counters) {
  return guard([&] {
    need(h, "heff");
    need(d_psi, "psi");
    need(opt, "opt");
    need(energies, "energies");
    const std::vector<int64_t> seg = h->p->part_offsets();
    if (n_inst != static_cast<int>(seg.size()) - 1)
      fail(kErrInvalid, "lanczos_batched: n_inst differs from the plan's instance count");
    if (opt->mlanc < 2 || opt->mlanc > 62) fail(kErrInvalid, "lanczos_batched: mlanc must be in 2..62");
    if (opt->check_every < 1) fail(kErrInvalid, "lanczos_batched: check_every must be >= 1");
    if (opt->refine_krylov < 1 || opt->max_restarts < 0) fail(kErrInvalid, "lanczos_batched: bad refinement options");
    if (!(opt->refine_tol > 0)) fail(kErrInvalid, "lanczos_batched: refine_tol must be > 0");
    if (h->p->projectors() > 0) fail(kErrInvalid, "lanczos_batched: plans with projectors are not batched");
    LanczosOpts o;
    o.mlanc = opt->mlanc;
    o.prec = prec_of(opt->precision);
    o.refine = opt->refine != 0;
    o.refine_tol = opt->refine_tol;
    o.refine_krylov = opt->refine_krylov;
    o.max_restarts = opt->max_restarts;
    o.check_every = opt->check_every;
    const BatchedLanczosOut r = lanczos_batched(*h->p, seg, static_cast<double*>(d_psi), o);
    for (int i = 0; i < n_inst; ++i) {
      energies[i] = r.energy[i];
      if (steps) steps[i] = r.steps[i];
      if (restarts) restarts[i] = r.restarts[i];
      if (residuals) residuals[i] = r.residual64[i];
    }
    if (counters) {
      counters[0] = r.matvec64;
      counters[1] = r.matvec32;
    }
  });
}

int sadmrg_dmrg_create(sadmrg_ctx* ctx, int sym, int cplx, int xx, int yy, int bondd, int physpn, int totspin,
                       sadmrg_dmrg** out) {
  return guard([&] {
    need(ctx, "ctx");
    need(out, "out");
    if (xx < 1 || yy < 1 || bondd < 1) fail(kErrInvalid, "dmrg: lattice sizes and bondd must be positive");
    auto d = std::make_unique<sadmrg_dmrg>();
    d->ctx = ctx->c.get();
    d->d = std::make_unique<Dmrg>(ctx->c.get(), sym_of(sym), cplx != 0, xx, yy, bondd, physpn, totspin);
    *out = d.release();
  });
}

int sadmrg_dmrg_create_exci(sadmrg_ctx* ctx, int sym, int cplx, int xx, int yy, int bondd, int physpn, int totspin,
                            int max_exci, sadmrg_dmrg** out) {
  return guard([&] {
    need(ctx, "ctx");
    need(out, "out");
    if (xx < 1 || yy < 1 || bondd < 1) fail(kErrInvalid, "dmrg: lattice sizes and bondd must be positive");
    auto d = std::make_unique<sadmrg_dmrg>();
    d->ctx = ctx->c.get();
    d->d = std::make_unique<Dmrg>(ctx->c.get(), sym_of(sym), cplx != 0, xx, yy, bondd, physpn, totspin, max_exci);
    *out = d.release();
  });
}

int sadmrg_dmrg_exci(sadmrg_dmrg* d, int* exci) {
  return guard([&] {
    need(d, "dmrg");
    need(exci, "exci");
    *exci = d->d->exci();
  });
}

int sadmrg_dmrg_set_lower_state(sadmrg_dmrg* d, int j, const sadmrg_dmrg* src) {
  return guard([&] {
    need(d, "dmrg");
    need(src, "src");
    d->d->set_lower_state(j, *src->d);
  });
}

int sadmrg_dmrg_lower_energy(sadmrg_dmrg* d, int j, double* e) {
  return guard([&] {
    need(d, "dmrg");
    need(e, "e");
    *e = d->d->lower_energy(j);
  });
}

int sadmrg_dmrg_set_channel_weights(sadmrg_dmrg* d, const double* w, int n) {
  return guard([&] {
    need(d, "dmrg");
    if (n < 0) fail(kErrInvalid, "channel weights: n >= 0");
    if (n > 0) need(w, "w");
    d->d->set_channel_weights(std::vector<double>(w, w + n));
  });
}

int sadmrg_dmrg_destroy(sadmrg_dmrg* d) {
  return guard([&] { delete d; });
}

int sadmrg_dmrg_ns(sadmrg_dmrg* d, int* ns) {
  return guard([&] {
    need(d, "dmrg");
    need(ns, "ns");
    *ns = d->d->ns();
  });
}

int sadmrg_dmrg_copy_state(sadmrg_dmrg* dst, const sadmrg_dmrg* src) {
  return guard([&] {
    need(dst, "dst");
    need(src, "src");
    if (dst->ctx != src->ctx) fail(kErrInvalid, "copy_state: drivers belong to different contexts");
    dst->d->copy_state_from(*src->d);
  });
}

int sadmrg_dmrg_set_couplings(sadmrg_dmrg* d, const int* hmap, const double* coup, int ncoup) {
  return guard([&] {
    need(d, "dmrg");
    need(hmap, "hmap");
    need(coup, "coup");
    const int ns = d->d->ns();
    if (ncoup < 1) fail(kErrInvalid, "set_couplings: ncoup must be >= 1");
    d->d->set_couplings(std::vector<int>(hmap, hmap + ns * ns), std::vector<double>(coup, coup + ncoup));
  });
}

int sadmrg_dmrg_get_couplings(sadmrg_dmrg* d, int* hmap, double* coup, int cap) {
  return guard([&] {
    need(d, "dmrg");
    const auto& h = d->d->hmap();
    const auto& c = d->d->coup();
    if (hmap) std::memcpy(hmap, h.data(), h.size() * sizeof(int));
    if (coup)
      for (int i = 0; i < cap && i < static_cast<int>(c.size()); ++i) coup[i] = c[i];
  });
}

int sadmrg_dmrg_set_max_dcut(sadmrg_dmrg* d, int max_dcut) {
  return guard([&] {
    need(d, "dmrg");
    if (max_dcut < 1) fail(kErrInvalid, "max_dcut must be >= 1");
    d->d->set_max_dcut(max_dcut);
  });
}

int sadmrg_dmrg_set_precision(sadmrg_dmrg* d, int precision, int refine, double refine_tol) {
  return guard([&] {
    need(d, "dmrg");
    LanczosOpts o = d->d->lanczos_opts();
    o.prec = prec_of(precision);
    o.refine = refine != 0;
    if (refine_tol > 0) o.refine_tol = refine_tol;
    d->d->set_lanczos(o);
  });
}

int sadmrg_dmrg_set_refinement(sadmrg_dmrg* d, int krylov_steps, int max_restarts) {
  return guard([&] {
    need(d, "dmrg");
    if (krylov_steps < 1 || max_restarts < 0) fail(kErrInvalid, "set_refinement: krylov_steps >= 1, max_restarts >= 0");
    LanczosOpts o = d->d->lanczos_opts();
    o.refine_krylov = krylov_steps;
    o.max_restarts = max_restarts;
    d->d->set_lanczos(o);
  });
}

int sadmrg_dmrg_set_verbose(sadmrg_dmrg* d, int level) {
  return guard([&] {
    need(d, "dmrg");
    if (level < 0) fail(kErrInvalid, "verbose level must be >= 0");
    d->d->set_verbose(level);
  });
}

int sadmrg_dmrg_set_check_interval(sadmrg_dmrg* d, int every) {
  return guard([&] {
    need(d, "dmrg");
    if (every < 1) fail(kErrInvalid, "set_check_interval: every >= 1");
    LanczosOpts o = d->d->lanczos_opts();
    o.check_every = every;
    d->d->set_lanczos(o);
  });
}

int sadmrg_dmrg_set_fp64_energy_every(sadmrg_dmrg* d, int every) {
  return guard([&] {
    need(d, "dmrg");
    require(every >= 0, "set_fp64_energy_every: every must be >= 0");
    d->d->set_fp64_energy_every(every);
  });
}

extern "C++" {
static std::string dir_arg(const char* dir) { return dir ? std::string(dir) : std::string(); }
}

int sadmrg_dmrg_save_mps(sadmrg_dmrg* d, const char* dir, int exci) {
  return guard([&] {
    need(d, "dmrg");
    if (exci < 0) fail(kErrInvalid, "save_mps: exci >= 0");
    d->d->save_mps(dir_arg(dir), exci);
  });
}

int sadmrg_dmrg_save_ww(sadmrg_dmrg* d, const char* dir) {
  return guard([&] {
    need(d, "dmrg");
    d->d->save_ww(dir_arg(dir));
  });
}

int sadmrg_dmrg_save_enr(sadmrg_dmrg* d, const char* dir) {
  return guard([&] {
    need(d, "dmrg");
    d->d->save_enr(dir_arg(dir));
  });
}

int sadmrg_dmrg_read_mps(sadmrg_dmrg* d, const char* dir, int read1, int read2, int* pass) {
  return guard([&] {
    need(d, "dmrg");
    if (read1 < 1) fail(kErrInvalid, "read_mps: read1 >= 1");
    const bool ok = read2 >= 0 ? d->d->read_mps(dir_arg(dir), read1, read2) : d->d->read_mps(dir_arg(dir), read1);
    if (pass) *pass = ok ? 1 : 0;
  });
}

int sadmrg_dmrg_read_ww(sadmrg_dmrg* d, const char* dir, int read1) {
  return guard([&] {
    need(d, "dmrg");
    d->d->read_ww(dir_arg(dir), read1);
  });
}

int sadmrg_dmrg_read_enr(sadmrg_dmrg* d, const char* dir, int read1) {
  return guard([&] {
    need(d, "dmrg");
    d->d->read_enr(dir_arg(dir), read1);
  });
}

int sadmrg_dmrg_set_checkpoint_dir(sadmrg_dmrg* d, const char* dir) {
  return guard([&] {
    need(d, "dmrg");
    d->d->set_checkpoint(dir_arg(dir), dir != nullptr);
  });
}

int sadmrg_dmrg_ww(sadmrg_dmrg* d, double* w, int cap, int* n) {
  return guard([&] {
    need(d, "dmrg");
    const auto& v = d->d->ww();
    if (n) *n = static_cast<int>(v.size());
    if (w)
      for (int k = 0; k < cap && k < static_cast<int>(v.size()); ++k) w[k] = v[k];
  });
}

int sadmrg_dmrg_idmrg(sadmrg_dmrg* d) {
  return guard([&] {
    need(d, "dmrg");
    d->d->do_idmrg();
  });
}

int sadmrg_dmrg_prepare_sweep(sadmrg_dmrg* d) {
  return guard([&] {
    need(d, "dmrg");
    d->d->prepare_sweep();
  });
}

int sadmrg_dmrg_sweep(sadmrg_dmrg* d) {
  return guard([&] {
    need(d, "dmrg");
    d->d->sweep();
  });
}

int sadmrg_dmrg_energy(sadmrg_dmrg* d, double* e) {
  return guard([&] {
    need(d, "dmrg");
    need(e, "e");
    *e = d->d->energy();
  });
}

int sadmrg_dmrg_records(sadmrg_dmrg* d, double* rec, int cap, int* n) {
  return guard([&] {
    need(d, "dmrg");
    auto& r = d->d->records();
    const int m = std::min<int>(cap, static_cast<int>(r.size()));
    for (int i = 0; i < m; ++i) {
      const StepRecord& s = r[i];
      double* p = rec + 14 * i;
      p[0] = s.il;
      p[1] = s.ir;
      p[2] = s.energy;
      p[3] = s.trunc_err;
      p[4] = s.kept;
      p[5] = s.matvec64;
      p[6] = s.matvec32;
      p[7] = s.t_env;
      p[8] = s.t_lanczos;
      p[9] = s.t_svd;
      p[10] = s.heff_flops;
      p[11] = static_cast<double>(s.ndata);
      p[12] = s.ritz32;
      p[13] = s.energy64;
    }
    r.erase(r.begin(), r.begin() + m);
    if (n) *n = m;
  });
}

static BTensor& dmrg_slot(sadmrg_dmrg* d, int which, int i, int j) {
  const int ns = d->d->ns();
  if (which == 5) i = 0;
  if (i < 0 || i >= ns || j < 0 || j >= ns) fail(kErrInvalid, "dmrg tensor index out of range");
  if (which >= 3 && which <= 5 && j >= d->d->exci()) fail(kErrInvalid, "dmrg tensor: stored state index out of range");
  if (which == 0) return d->d->uu(i);
  if (which == 1) return d->d->hh(i);
  if (which == 2) return d->d->opr(i, j);
  if (which == 3) return d->d->orth(j, i);
  if (which == 4) return d->d->ovlp(j, i);
  if (which == 5) return d->d->overlapvec(j);
  fail(kErrInvalid, "dmrg tensor: which must be 0 (uu), 1 (hh), 2 (opr), 3 (orth), 4 (ovlp) or 5 (overlapvec)");
}

int sadmrg_dmrg_get_tensor(sadmrg_dmrg* d, int which, int i, int j, sadmrg_tensor* out) {
  return guard([&] {
    need(d, "dmrg");
    need(out, "out");
    copy_tensor(out->t, dmrg_slot(d, which, i, j));
    d->ctx->sync();
    fix_kind(out);
  });
}

int sadmrg_dmrg_set_tensor(sadmrg_dmrg* d, int which, int i, int j, const sadmrg_tensor* t) {
  return guard([&] {
    need(d, "dmrg");
    need(t, "tensor");
    copy_tensor(dmrg_slot(d, which, i, j), t->t);
    d->ctx->sync();
  });
}

int sadmrg_dmrg_call(sadmrg_dmrg* d, int code, int a, int b) {
  return guard([&] {
    need(d, "dmrg");
    switch (code) {
      case 0: d->d->wavefunc_transformation(a, b); break;
      case 1: d->d->prepare_site_operator_from_left(a, b); break;
      case 2: d->d->prepare_site_operator_from_right(a, b); break;
      case 3: d->d->prepare_site_operator_from_left(a); break;
      case 4: d->d->prepare_site_operator_from_right(a); break;
      case 5: d->d->prepare_sweep(); break;
      default: fail(kErrInvalid, "dmrg_call: unknown code");
    }
    d->ctx->sync();
  });
}

int sadmrg_dmrg_coupling_gradient(sadmrg_dmrg* d, double* grad, int cap, double* energy) {
  return guard([&] {
    need(d, "dmrg");
    need(grad, "grad");
    const std::vector<double> g = d->d->coupling_gradient(energy);
    for (int c = 0; c < cap; ++c) grad[c] = c < static_cast<int>(g.size()) ? g[c] : 0.0;
  });
}

int sadmrg_dmrg_input_vector(sadmrg_dmrg* d, int il, int ir, sadmrg_tensor* out) {
  return guard([&] {
    need(d, "dmrg");
    need(out, "out");
    BTensor v;
    d->d->prepare_input_vector(il, ir, v);
    out->t = std::move(v);
    fix_kind(out);
  });
}

int sadmrg_dmrg_heff(sadmrg_dmrg* d, int il, int ir, const sadmrg_tensor* vin, sadmrg_tensor* vout, int precision) {
  return guard([&] {
    need(d, "dmrg");
    need(vin, "vin");
    need(vout, "vout");
    auto plan = d->d->make_heff(il, ir, vin->t, nullptr);
    std::vector<uint8_t> mask(vin->t.nten());
    for (int b = 0; b < vin->t.nten(); ++b) mask[b] = vin->t.present(b);
    BTensor r;
    r.layout(d->ctx, vin->t.cplx, vin->t.st, mask, kCgcSinglet);
    plan->apply(vin->t.dptr(), r.dptr(), prec_of(precision));
    d->ctx->sync();
    vout->t = std::move(r);
    fix_kind(vout);
  });
}

int sadmrg_dmrg_heff_create(sadmrg_dmrg* d, int il, int ir, const sadmrg_tensor* layout, sadmrg_heff** out) {
  return guard([&] {
    need(d, "dmrg");
    need(layout, "layout");
    need(out, "out");
    auto h = std::make_unique<sadmrg_heff>();
    h->p = d->d->make_heff(il, ir, layout->t, nullptr);
    *out = h.release();
  });
}

int sadmrg_dmrg_lanczos(sadmrg_dmrg* d, int il, int ir, sadmrg_tensor* vec, double* energy) {
  return guard([&] {
    need(d, "dmrg");
    need(vec, "vec");
    const double e = d->d->lanczos_solve(il, ir, vec->t);
    if (energy) *energy = e;
  });
}

int sadmrg_dmrg_svd(sadmrg_dmrg* d, const sadmrg_tensor* vec, int a, int b, double p1, double p2, double* wout,
                    int cap, int* nw) {
  return guard([&] {
    need(d, "dmrg");
    need(vec, "vec");
    d->d->svd(vec->t, a, b, p1, p2);
    const auto& w = d->d->last_wout();
    if (nw) *nw = static_cast<int>(w.size());
    if (wout)
      for (int k = 0; k < cap && k < static_cast<int>(w.size()); ++k) wout[k] = w[k];
  });
}

}  // extern "C"
