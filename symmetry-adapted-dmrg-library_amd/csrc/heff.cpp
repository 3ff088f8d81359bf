// heff.cpp -- see heff.hpp.
#include "heff.hpp"
#include <atomic>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include "kernels.hpp"

namespace sadmrg {

namespace {
enum { kArIn = 0, kArOut = 1, kArScratch = 2, kArPack = 3 };
inline int64_t pad4(int64_t x) { return (x + 3) / 4 * 4; }
// zero a device buffer of `bytes` (allocations are rounded to 256 bytes)
void zero_bytes(Context* ctx, void* p, size_t bytes) {
  if (bytes) SADMRG_HIP(launch_fill(static_cast<int64_t>((bytes + 7) / 8), 0.0, static_cast<double*>(p), ctx->stream()));
}
// column chunks of one block copy (one workgroup each)
void add_copy(std::vector<ScatterTask>& v, int64_t src_off, int64_t lds, int64_t dst_off, int64_t ldd, int rows,
              int cols, int chunk = 64) {
  for (int c0 = 0; c0 < cols; c0 += chunk) {
    ScatterTask t{};
    t.src_off = src_off + int64_t(c0) * lds;
    t.dst_off = dst_off + int64_t(c0) * ldd;
    t.lds = lds;
    t.ldd = ldd;
    t.rows = rows;
    t.cols = std::min(chunk, cols - c0);
    t.kind = 0;
    t.scale = 1.0;
    v.push_back(t);
  }
}
}  // namespace

HeffPlan::HeffPlan(Context* ctx, CgcTables* cgc, const BTensor& hl, const BTensor& hr,
                   const std::vector<const BTensor*>& o1, const std::vector<const BTensor*>& o2,
                   const BTensor& psi)
    : ctx_(ctx), cplx_(psi.cplx) {
  require(o1.size() == o2.size(), "H_eff: term operator lists differ in length");
  require(!psi.is_null() && psi.nbond() == 2, "H_eff: psi must be a 2-leg tensor");
  const SymKind sym = psi.sym();
  const bool su2 = sym == SymKind::SU2;
  const int w = psi.w();
  layout_.ctx = ctx;
  layout_.cplx = psi.cplx;
  layout_.st = psi.st;
  layout_.off = psi.off;
  layout_.tag = psi.tag;
  layout_.nelem = psi.nelem;
  n_ = psi.nelem;
  const Struct& ps = psi.st;
  const int ntp = ps.nten();

  // ---- pack operator blocks -------------------------------------------------
  std::vector<const BTensor*> tens;
  auto tid_of = [&](const BTensor* t) {
    for (size_t i = 0; i < tens.size(); ++i)
      if (tens[i] == t) return static_cast<int>(i);
    tens.push_back(t);
    return static_cast<int>(tens.size()) - 1;
  };
  if (!hl.is_null()) tid_of(&hl);
  if (!hr.is_null()) tid_of(&hr);
  for (size_t t = 0; t < o1.size(); ++t) {
    if (o1[t] && !o1[t]->is_null()) tid_of(o1[t]);
    if (o2[t] && !o2[t]->is_null()) tid_of(o2[t]);
  }
  // padded pack: block b of tensor i at poff[i][b], leading dimension pld[i][b]
  // (a block's rows are its elements over its last dimension)
  std::vector<std::vector<int64_t>> poff(tens.size()), pld(tens.size());
  for (size_t i = 0; i < tens.size(); ++i) {
    require(tens[i]->cplx == cplx_ && tens[i]->sym() == sym, "H_eff: operator scalar/symmetry mismatch");
    const BTensor& t = *tens[i];
    poff[i].assign(t.nten(), -1);
    pld[i].assign(t.nten(), 0);
    int dd[4];
    for (int b = 0; b < t.nten(); ++b) {
      if (!t.present(b) || t.block_size(b) == 0) continue;
      t.st.block_dims(b, dd);
      const int64_t cols = dd[t.nbond() - 1];
      const int64_t rows = t.block_size(b) / cols;
      poff[i][b] = pack_elems_;
      pld[i][b] = pad4(rows);
      pack_elems_ += pld[i][b] * cols;
    }
  }
  pack64_ = DevBuf(ctx, std::max<int64_t>(pack_elems_, 1) * 8 * w);
  {
    // one launch for every block of every tensor (task.pad indexes the
    // source bases); the copies write the zero padding
    std::vector<ScatterTask> ti;
    std::vector<const double*> bases(tens.size());
    for (size_t i = 0; i < tens.size(); ++i) {
      bases[i] = tens[i]->dptr();
      for (int b = 0; b < tens[i]->nten(); ++b) {
        if (poff[i][b] < 0) continue;
        int dd[4];
        tens[i]->st.block_dims(b, dd);
        const int64_t cols = dd[tens[i]->nbond() - 1];
        const size_t first = ti.size();
        add_copy(ti, tens[i]->off[b], tens[i]->block_size(b) / cols, poff[i][b], pld[i][b],
                 static_cast<int>(tens[i]->block_size(b) / cols), static_cast<int>(cols));
        for (size_t k = first; k < ti.size(); ++k) ti[k].pad = static_cast<int32_t>(i);
      }
    }
    if (!ti.empty()) {
      const size_t tb_bytes = ti.size() * sizeof(ScatterTask);
      DevBuf tb(ctx, tb_bytes + bases.size() * sizeof(double*));
      ctx->upload(tb.get(), ti.data(), tb_bytes);
      const double* const* dbases = reinterpret_cast<const double* const*>(static_cast<char*>(tb.get()) + tb_bytes);
      ctx->upload(const_cast<const double**>(dbases), bases.data(), bases.size() * sizeof(double*));
      SADMRG_HIP(launch_repack(tb.as<ScatterTask>(), static_cast<int>(ti.size()), nullptr, dbases, pack64_.get(), w, 0,
                               ctx->stream()));
      // tb returns to the stream-ordered cache; the staged sources are copies
    }
  }
  auto pk = [&](const BTensor* t, int b) -> Operand {
    const int i = tid_of(t);
    require(poff[i][b] >= 0, "H_eff: operator block missing from the pack");
    Operand o;
    o.arena = kArPack;
    o.off = poff[i][b];
    o.ld = pld[i][b];
    return o;
  };
  // padded input copy: psi block b at in_off[b], leading dimension pad4(rows)
  std::vector<int64_t> in_off(ntp, -1);
  for (int b = 0; b < ntp; ++b) {
    if (!psi.present(b)) continue;
    int dd[2];
    ps.block_dims(b, dd);
    in_off[b] = in_elems_;
    add_copy(in_tasks_, psi.off[b], dd[0], in_elems_, pad4(dd[0]), dd[0], dd[1], 8);  // many workgroups
    in_elems_ += pad4(dd[0]) * dd[1];
  }
  auto pin = [&](int b) -> Operand {
    int dd[2];
    ps.block_dims(b, dd);
    Operand o;
    o.arena = kArIn;
    o.off = in_off[b];
    o.ld = pad4(dd[0]);
    return o;
  };

  // ---- stage-2 problems: one per psi block ----------------------------------
  std::vector<int> prob_of(ntp, -1);
  int lab[4], d[4];
  for (int b = 0; b < ntp; ++b) {
    if (!psi.present(b)) continue;
    ps.block_dims(b, d);
    prob_of[b] = g2_.add_problem(kArOut, psi.off[b], d[0], d[0], d[1], false);
    stats_.nsec_in++;
  }
  std::vector<int> contributions(ntp, 0);

  // diagonal: contract(hh_L, 0, psi, 0) and contract(psi, 1, hh_R, 0)
  if (!hl.is_null()) {
    require(hl.nbond() == 2 && hl.st.bonds[0] == ps.bonds[0].inverted() && hl.st.bonds[1] == ps.bonds[0],
            "H_eff: left block Hamiltonian bonds do not match psi");
    for (int b = 0; b < ntp; ++b) {
      if (!psi.present(b)) continue;
      ps.labels(b, lab);
      const int l1[2] = {lab[0], lab[0]};
      if (!hl.st.allowed_labels(l1)) continue;
      const int hb = hl.st.index_of_labels(l1);
      if (hb < 0 || !hl.present(hb)) continue;
      ps.block_dims(b, d);
      Operand A = pk(&hl, hb), B = pin(b);
      A.trans = true;
      g2_.add_seg(prob_of[b], A, B, d[0], 1.0);
      contributions[b]++;
    }
  }
  if (!hr.is_null()) {
    require(hr.nbond() == 2 && hr.st.bonds[0] == ps.bonds[1].inverted() && hr.st.bonds[1] == ps.bonds[1],
            "H_eff: right block Hamiltonian bonds do not match psi");
    for (int b = 0; b < ntp; ++b) {
      if (!psi.present(b)) continue;
      ps.labels(b, lab);
      const int l2[2] = {lab[1], lab[1]};
      if (!hr.st.allowed_labels(l2)) continue;
      const int hb = hr.st.index_of_labels(l2);
      if (hb < 0 || !hr.present(hb)) continue;
      ps.block_dims(b, d);
      Operand A = pin(b), B = pk(&hr, hb);
      g2_.add_seg(prob_of[b], A, B, d[1], 1.0);
      contributions[b]++;
    }
  }

  // bond terms
  const BlockList lp(psi);
  for (size_t t = 0; t < o1.size(); ++t) {
    const BTensor* op1 = o1[t];
    const BTensor* op2 = o2[t];
    if (!op1 || !op2 || op1->is_null() || op2->is_null()) continue;
    const Bond ch2 = su2 ? op2->st.bonds[1] : op2->st.bonds[1].inverted();
    require(op1->st.bonds[0] == ps.bonds[0].inverted() && op2->st.bonds[0] == ps.bonds[1].inverted() &&
                op1->st.bonds[1] == ch2 && op1->st.bonds[2] == ps.bonds[0] && op2->st.bonds[2] == ps.bonds[1],
            "H_eff: bond-term operator bonds do not match psi");
    int a1[3], a3[3], d1[3], d3[3];
    // the reference's loop nest (k over psi, j over op1, l over op2) over
    // matching blocks only
    const BlockList l1(*op1), l3(*op2);
    const LabelIndex ix1(l1, 0, -1), ix3(l3, 0, 1);
    for (int ip = 0; ip < lp.size(); ++ip) {
      const int k = lp.b[ip];
      lab[0] = lp.lab[ip][0];
      lab[1] = lp.lab[ip][1];
      for (const int pj : ix1.get(lab[0])) {
      const int j = l1.b[pj];
      for (int q = 0; q < 3; ++q) a1[q] = l1.lab[pj][q];
      op1->st.block_dims(j, d1);
      ps.block_dims(k, d);
      int64_t xoff = -1;
      int xm = 0;
      for (const int pl : ix3.get(lab[1], a1[1])) {
        const int l = l3.b[pl];
        for (int q = 0; q < 3; ++q) a3[q] = l3.lab[pl][q];
        const int ol[2] = {a1[2], a3[2]};
        if (!ps.allowed_labels(ol)) continue;
        const int m = ps.index_of_labels(ol);
        if (m < 0) continue;
        double fac = 1.0;
        if (su2) {
          fac = cgc->hamilt_vec(a1[0], a1[1], a1[2]);
          if (std::fabs(fac) < 1e-12) continue;
        }
        require(psi.present(m), "H_eff: output block missing from the psi layout");
        if (xoff < 0) {
          // X = psi_k^T O1_j : (b) x (c * a'), leading dimension padded
          require(d1[1] == 1, "H_eff: operator channel dimension must be 1");
          xoff = scratch_elems_;
          xm = d[1];
          scratch_elems_ += pad4(d[1]) * d1[1] * d1[2];
          const int p = g1_.add_problem(kArScratch, xoff, pad4(d[1]), d[1], d1[1] * d1[2], false);
          Operand A = pin(k), B = pk(op1, j);
          A.trans = true;
          g1_.add_seg(p, A, B, d[0], 1.0);
        }
        op2->st.block_dims(l, d3);
        const int kk = d3[0] * d3[1];
        require(kk == xm * d1[1], "H_eff: inner dimension mismatch");
        Operand A, B = pk(op2, l);
        A.arena = kArScratch;
        A.off = xoff;
        A.ld = pad4(kk);
        A.trans = true;
        g2_.add_seg(prob_of[m], A, B, kk, fac);
        contributions[m]++;
      }
      }
    }
  }
  out_null_.assign(ntp, 0);
  for (int b = 0; b < ntp; ++b) out_null_[b] = psi.present(b) && contributions[b] == 0;

  for (GemmBatch* g : {&g1_, &g2_}) {
    g->limit(kArIn, in_elems_);
    g->limit(kArOut, n_);
    g->limit(kArScratch, scratch_elems_);
    g->limit(kArPack, pack_elems_);
    g->request_vec();  // the padded layout above
  }
  g1_.finalize(ctx, cplx_);
  g2_.finalize(ctx, cplx_);
  scratch64_ = DevBuf(ctx, std::max<int64_t>(scratch_elems_, 1) * 8 * w);
  zero_bytes(ctx, scratch64_.get(), scratch64_.bytes());  // padding rows stay zero
  in64_ = DevBuf(ctx, std::max<int64_t>(in_elems_, 1) * 8 * w);
  if (!in_tasks_.empty()) {
    in_tasks_dev_ = DevBuf(ctx, in_tasks_.size() * sizeof(ScatterTask));
    ctx->upload(in_tasks_dev_.get(), in_tasks_.data(), in_tasks_.size() * sizeof(ScatterTask));
  }
  stats_.n = n_;
  stats_.flops = g1_.flops(cplx_) + g2_.flops(cplx_);
  stats_.bytes64 = 8.0 * w * (2.0 * n_ + pack_elems_);
  stats_.stage1_probs = g1_.nprob();
  stats_.stage2_probs = g2_.nprob();
  stats_.stage1_segs = g1_.nseg();
  stats_.stage2_segs = g2_.nseg();
  const GemmPrec p64 = cplx_ ? GemmPrec::Z64 : GemmPrec::F64;
  stats_.tile_cfg1 = g1_.default_cfg(p64);
  stats_.tile_cfg2 = g2_.default_cfg(p64);
  stats_.tiles1 = g1_.ntiles(ctx, stats_.tile_cfg1);
  stats_.tiles2 = g2_.ntiles(ctx, stats_.tile_cfg2);
  stats_.split_tiles = g1_.nsplit_tiles(ctx, stats_.tile_cfg1) + g2_.nsplit_tiles(ctx, stats_.tile_cfg2);
}

std::unique_ptr<HeffPlan> HeffPlan::merge(const std::vector<const HeffPlan*>& parts) {
  require(!parts.empty(), "H_eff merge: no plans");
  Context* ctx = parts[0]->ctx_;
  const bool cx = parts[0]->cplx_;
  const int w = cx ? 2 : 1;
  std::unique_ptr<HeffPlan> m(new HeffPlan(ctx, cx));
  m->merged_ = true;
  int64_t pack = 0, scratch = 0, n = 0;
  m->part_off_.assign(1, 0);
  for (const HeffPlan* p : parts) m->part_off_.push_back(m->part_off_.back() + p->n_);
  for (const HeffPlan* p : parts) {
    require(p && p->ctx_ == ctx && p->cplx_ == cx, "H_eff merge: plans differ in context or scalar type");
    require(p->projectors() == 0, "H_eff merge: plans with projectors are not merged");
    int64_t shift[kMaxArenas] = {};
    shift[kArIn] = m->in_elems_;
    shift[kArOut] = n;
    shift[kArScratch] = scratch;
    shift[kArPack] = pack;
    m->g1_.absorb(p->g1_, shift);
    m->g2_.absorb(p->g2_, shift);
    for (ScatterTask t : p->in_tasks_) {
      t.src_off += n;
      t.dst_off += m->in_elems_;
      m->in_tasks_.push_back(t);
    }
    n += p->n_;
    m->in_elems_ += p->in_elems_;
    pack += p->pack_elems_;
    scratch += p->scratch_elems_;
  }
  m->n_ = n;
  m->pack_elems_ = pack;
  m->scratch_elems_ = scratch;
  m->pack64_ = DevBuf(ctx, std::max<int64_t>(pack, 1) * 8 * w);
  int64_t at = 0;
  for (const HeffPlan* p : parts) {
    if (p->pack_elems_)
      SADMRG_HIP(hipMemcpyAsync(m->pack64_.as<double>() + at * w, p->pack64_.get(), p->pack_elems_ * 8 * w,
                                hipMemcpyDeviceToDevice, ctx->stream()));
    at += p->pack_elems_;
  }
  m->scratch64_ = DevBuf(ctx, std::max<int64_t>(scratch, 1) * 8 * w);
  zero_bytes(ctx, m->scratch64_.get(), m->scratch64_.bytes());
  m->in64_ = DevBuf(ctx, std::max<int64_t>(m->in_elems_, 1) * 8 * w);
  if (!m->in_tasks_.empty()) {
    m->in_tasks_dev_ = DevBuf(ctx, m->in_tasks_.size() * sizeof(ScatterTask));
    ctx->upload(m->in_tasks_dev_.get(), m->in_tasks_.data(), m->in_tasks_.size() * sizeof(ScatterTask));
  }
  for (GemmBatch* g : {&m->g1_, &m->g2_}) {
    g->request_vec();
    g->limit(kArIn, m->in_elems_);
    g->limit(kArOut, n);
    g->limit(kArScratch, scratch);
    g->limit(kArPack, pack);
  }
  m->g1_.finalize(ctx, cx);
  m->g2_.finalize(ctx, cx);
  HeffStats& st = m->stats_;
  st.n = n;
  st.flops = m->g1_.flops(cx) + m->g2_.flops(cx);
  st.bytes64 = 8.0 * w * (2.0 * n + pack);
  st.stage1_probs = m->g1_.nprob();
  st.stage2_probs = m->g2_.nprob();
  st.stage1_segs = m->g1_.nseg();
  st.stage2_segs = m->g2_.nseg();
  const GemmPrec p64 = cx ? GemmPrec::Z64 : GemmPrec::F64;
  st.tile_cfg1 = m->g1_.default_cfg(p64);
  st.tile_cfg2 = m->g2_.default_cfg(p64);
  st.tiles1 = m->g1_.ntiles(ctx, st.tile_cfg1);
  st.tiles2 = m->g2_.ntiles(ctx, st.tile_cfg2);
  st.split_tiles = m->g1_.nsplit_tiles(ctx, st.tile_cfg1) + m->g2_.nsplit_tiles(ctx, st.tile_cfg2);
  for (const HeffPlan* p : parts) st.nsec_in += p->stats_.nsec_in;
  return m;
}

namespace {
struct TuneKey {
  int prec, stage;
  uint64_t shape;
  bool operator<(const TuneKey& o) const {
    return prec != o.prec ? prec < o.prec : stage != o.stage ? stage < o.stage : shape < o.shape;
  }
};
std::mutex g_tune_mu;
std::map<TuneKey, int> g_tune;
bool g_autotune = true;
}  // namespace

void set_autotune(bool on) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_autotune = on;
}
bool autotune() {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  return g_autotune;
}
int autotune_entries() {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  return static_cast<int>(g_tune.size());
}
void autotune_clear() {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune.clear();
}

// Tile configuration for one stage: tuned (cached per key) or the default.
namespace {
std::atomic<int> g_forced_plan{-1};
}  // namespace

void set_heff_launch_plan(int plan) {
  require(plan >= -1 && plan < 2 * GemmBatch::kPlans, "H_eff launch plan out of range");
  g_forced_plan = plan;
}
int heff_launch_plan() { return g_forced_plan; }

int HeffPlan::cfg_for(GemmBatch& g, GemmPrec prec, int stage, const ArenaTable& at) {
  const int forced = g_forced_plan;
  if (forced >= 0) {
    require(!g.complex() || complex_tile_cfg(forced % GemmBatch::kPlans % kTileConfigs),
            "H_eff launch plan: complex tables have tile configurations 0 and 3 only");
    return forced;
  }
  if (!autotune()) return g.default_cfg(prec);
  const TuneKey key{static_cast<int>(prec), stage, g.shape_key()};
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find(key);
    if (it != g_tune.end()) return it->second;
  }
  hipEvent_t e0, e1;
  SADMRG_HIP(hipEventCreate(&e0));
  SADMRG_HIP(hipEventCreate(&e1));
  int best = g.default_cfg(prec);
  float best_ms = 0;
  // plans >= kPlans: the persistent launch of the same plan (mode 2 only)
  const int nplans = GemmBatch::persistent() == 2 ? 2 * GemmBatch::kPlans : GemmBatch::kPlans;
  for (int cfg = 0; cfg < nplans; ++cfg) {
    // split variants only differ for tables that split their reductions
    const int plan = cfg % GemmBatch::kPlans;
    // complex tables: no split reductions or persistent launches, two tile
    // configurations
    if (g.complex() && (cfg != plan || plan >= kTileConfigs || !complex_tile_cfg(plan))) continue;
    if (plan >= kTileConfigs && g.nsplit_tiles(ctx_, plan % kTileConfigs) == 0) continue;
    g.run(ctx_, prec, at, cfg);  // builds the tile set and warms caches
    SADMRG_HIP(hipEventRecord(e0, ctx_->stream()));
    g.run(ctx_, prec, at, cfg);
    SADMRG_HIP(hipEventRecord(e1, ctx_->stream()));
    SADMRG_HIP(hipEventSynchronize(e1));
    float ms = 0;
    SADMRG_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (cfg == 0 || ms < best_ms) {
      best_ms = ms;
      best = cfg;
    }
  }
  SADMRG_HIP(hipEventDestroy(e0));
  SADMRG_HIP(hipEventDestroy(e1));
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune[key] = best;
  return best;
}

double HeffPlan::stage_flops(int stage) const {
  if (stage == 1) return g1_.flops(cplx_);
  if (stage == 2) return g2_.flops(cplx_);
  return g1_.flops(cplx_) + g2_.flops(cplx_);
}

double HeffPlan::issued_flops(int stage, Prec prec) const {
  const bool planar = cplx_ && planar_built_ && heff_complex_planar();
  const GemmBatch& a = planar ? g1p_ : g1_;
  const GemmBatch& b = planar ? g2p_ : g2_;
  GemmPrec p1, p2;
  if (prec == Prec::FP64) {
    p1 = p2 = (cplx_ && !planar) ? GemmPrec::Z64 : GemmPrec::F64;
  } else if (prec == Prec::FP32) {
    p1 = GemmPrec::F32;
    p2 = GemmPrec::F32_OUT64;
  } else {
    p1 = GemmPrec::BF16;
    p2 = GemmPrec::BF16_OUT64;
  }
  double f = 0;
  if (stage != 2) f += a.issued_flops(stats_.tile_cfg1, p1);
  if (stage != 1) f += b.issued_flops(stats_.tile_cfg2, p2);
  return f;
}

namespace {
std::atomic<bool> g_per_block{false};
std::atomic<bool> g_complex_planar{false};
int64_t round4(int64_t n) { return (n + 3) / 4 * 4; }
}  // namespace

void set_heff_per_block(bool on) { g_per_block = on; }
bool heff_per_block() { return g_per_block; }
void set_heff_complex_planar(bool on) { g_complex_planar = on; }
bool heff_complex_planar() { return g_complex_planar; }

void HeffPlan::build_planar() {
  planar_built_ = true;
  for (auto& x : n4_) x = 0;
  n4_[kArIn] = round4(in_elems_);
  n4_[kArOut] = round4(n_);
  n4_[kArScratch] = round4(scratch_elems_);
  n4_[kArPack] = round4(pack_elems_);
  for (GemmBatch* g : {&g1p_, &g2p_}) {
    g->planar_from(g == &g1p_ ? g1_ : g2_, n4_);
    g->request_vec();
    for (int a : {kArIn, kArOut, kArScratch, kArPack}) g->limit(a, 2 * n4_[a]);
    g->finalize(ctx_, false);
  }
  op64_ = DevBuf(ctx_, std::max<int64_t>(2 * n4_[kArOut], 1) * 8);
  zero_bytes(ctx_, op64_.get(), op64_.bytes());
}

// Complex H_eff on planar operands: the input is repacked (interleaved, the
// padded layout of the complex tables) and split into planes, both stages
// run as real tables, and the planar output is interleaved into d_out.
void HeffPlan::apply_planar(const double* d_in, double* d_out, Prec prec, int stage, bool unbatched) {
  require(prec == Prec::FP64 || prec == Prec::FP32, "H_eff: the BF16 path is real-only");
  if (!planar_built_) build_planar();
  const bool f32 = prec == Prec::FP32;
  const int64_t np = 2 * n4_[kArPack], ns = 2 * n4_[kArScratch], ni = 2 * n4_[kArIn];
  DevBuf& pp = f32 ? pp32_ : pp64_;
  DevBuf& sp = f32 ? sp32_ : sp64_;
  DevBuf& ip = f32 ? ip32_ : ip64_;
  const size_t es = f32 ? 4 : 8;
  if (!pp.get()) {
    pp = DevBuf(ctx_, std::max<int64_t>(np, 1) * es);
    sp = DevBuf(ctx_, std::max<int64_t>(ns, 1) * es);
    ip = DevBuf(ctx_, std::max<int64_t>(ni, 1) * es);
    zero_bytes(ctx_, sp.get(), sp.bytes());  // padding rows stay zero
    SADMRG_HIP(launch_deinterleave(pack_elems_, n4_[kArPack], pack64_.as<double>(), pp.get(), f32 ? 1 : 0,
                                   ctx_->stream()));
  }
  if (stage != 2) {
    repack_input(d_in, in64_.get(), 0);
    SADMRG_HIP(launch_deinterleave(in_elems_, n4_[kArIn], in64_.as<double>(), ip.get(), f32 ? 1 : 0,
                                   ctx_->stream()));
  }
  ArenaTable at;
  std::memset(&at, 0, sizeof(at));
  at.base[kArIn] = ip.get();
  at.base[kArOut] = op64_.get();
  at.base[kArScratch] = sp.get();
  at.base[kArPack] = pp.get();
  auto exec = [&](GemmBatch& g, GemmPrec p, int cfg) {
    if (unbatched) g.run_per_problem(ctx_, p, at, cfg);
    else g.run(ctx_, p, at, cfg);
  };
  if (stage != 2) {
    const GemmPrec p = f32 ? GemmPrec::F32 : GemmPrec::F64;
    stats_.tile_cfg1 = cfg_for(g1p_, p, 1, at);
    exec(g1p_, p, stats_.tile_cfg1);
  }
  if (stage != 1) {
    const GemmPrec p = f32 ? GemmPrec::F32_OUT64 : GemmPrec::F64;
    stats_.tile_cfg2 = cfg_for(g2p_, p, 2, at);
    exec(g2p_, p, stats_.tile_cfg2);
    SADMRG_HIP(launch_interleave(n_, n4_[kArOut], op64_.as<double>(), d_out, ctx_->stream()));
  }
  if (stage == 0) apply_projectors(d_in, d_out);
}

void HeffPlan::apply(const double* d_in, double* d_out, Prec prec, int stage, bool unbatched) {
  unbatched = unbatched || g_per_block;
  if (cplx_ && g_complex_planar) {
    apply_planar(d_in, d_out, prec, stage, unbatched);
    return;
  }
  auto exec = [&](GemmBatch& g, GemmPrec p, const ArenaTable& at, int cfg) {
    if (unbatched) g.run_per_problem(ctx_, p, at, cfg);
    else g.run(ctx_, p, at, cfg);
  };
  ArenaTable at;
  std::memset(&at, 0, sizeof(at));
  if (cplx_ && prec == Prec::FP32) {
    // complex<float> operands, FP32 MFMA, complex<double> output
    if (!pack32_.get()) {
      pack32_ = DevBuf(ctx_, std::max<int64_t>(pack_elems_, 1) * 8);
      scratch32_ = DevBuf(ctx_, std::max<int64_t>(scratch_elems_, 1) * 8);
      in32_ = DevBuf(ctx_, std::max<int64_t>(in_elems_, 1) * 8);
      zero_bytes(ctx_, scratch32_.get(), scratch32_.bytes());
      SADMRG_HIP(launch_f64_to_f32(2 * pack_elems_, pack64_.as<double>(), pack32_.as<float>(), ctx_->stream()));
    }
    if (stage != 2) repack_input(d_in, in32_.get(), 1);
    at.base[kArIn] = in32_.get();
    at.base[kArOut] = d_out;
    at.base[kArScratch] = scratch32_.get();
    at.base[kArPack] = pack32_.get();
    if (stage != 2) {
      stats_.tile_cfg1 = cfg_for(g1_, GemmPrec::Z32, 1, at);
      exec(g1_, GemmPrec::Z32, at, stats_.tile_cfg1);
    }
    if (stage != 1) {
      stats_.tile_cfg2 = cfg_for(g2_, GemmPrec::Z32_OUT64, 2, at);
      exec(g2_, GemmPrec::Z32_OUT64, at, stats_.tile_cfg2);
    }
    if (stage == 0) apply_projectors(d_in, d_out);
    return;
  }
  if (prec == Prec::FP64 || cplx_) {
    require(prec == Prec::FP64, "H_eff: the BF16 path is real-only");
    if (stage != 2) repack_input(d_in, in64_.get(), 0);
    at.base[kArIn] = in64_.get();
    at.base[kArOut] = d_out;
    at.base[kArScratch] = scratch64_.get();
    at.base[kArPack] = pack64_.get();
    const GemmPrec p = cplx_ ? GemmPrec::Z64 : GemmPrec::F64;
    if (stage != 2) {
      stats_.tile_cfg1 = cfg_for(g1_, p, 1, at);
      exec(g1_, p, at, stats_.tile_cfg1);
    }
    if (stage != 1) {
      stats_.tile_cfg2 = cfg_for(g2_, p, 2, at);
      exec(g2_, p, at, stats_.tile_cfg2);
    }
    if (stage == 0) apply_projectors(d_in, d_out);
    return;
  }
  if (prec == Prec::BF16) {
    if (!pack16_.get()) {
      pack16_ = DevBuf(ctx_, std::max<int64_t>(pack_elems_, 1) * 2);
      scratch16_ = DevBuf(ctx_, std::max<int64_t>(scratch_elems_, 1) * 2);
      in16_ = DevBuf(ctx_, std::max<int64_t>(in_elems_, 1) * 2);
      zero_bytes(ctx_, scratch16_.get(), scratch16_.bytes());
      SADMRG_HIP(launch_f64_to_bf16(pack_elems_, pack64_.as<double>(), pack16_.as<uint16_t>(), ctx_->stream()));
    }
    if (stage != 2) repack_input(d_in, in16_.get(), 2);
    at.base[kArIn] = in16_.get();
    at.base[kArOut] = d_out;
    at.base[kArScratch] = scratch16_.get();
    at.base[kArPack] = pack16_.get();
    if (stage != 2) {
      stats_.tile_cfg1 = cfg_for(g1_, GemmPrec::BF16, 1, at);
      exec(g1_, GemmPrec::BF16, at, stats_.tile_cfg1);
    }
    if (stage != 1) {
      stats_.tile_cfg2 = cfg_for(g2_, GemmPrec::BF16_OUT64, 2, at);
      exec(g2_, GemmPrec::BF16_OUT64, at, stats_.tile_cfg2);
    }
    if (stage == 0) apply_projectors(d_in, d_out);
    return;
  }
  if (!pack32_.get()) {
    pack32_ = DevBuf(ctx_, std::max<int64_t>(pack_elems_, 1) * 4);
    scratch32_ = DevBuf(ctx_, std::max<int64_t>(scratch_elems_, 1) * 4);
    in32_ = DevBuf(ctx_, std::max<int64_t>(in_elems_, 1) * 4);
    zero_bytes(ctx_, scratch32_.get(), scratch32_.bytes());
    SADMRG_HIP(launch_f64_to_f32(pack_elems_, pack64_.as<double>(), pack32_.as<float>(), ctx_->stream()));
  }
  if (stage != 2) repack_input(d_in, in32_.get(), 1);
  at.base[kArIn] = in32_.get();
  at.base[kArOut] = d_out;
  at.base[kArScratch] = scratch32_.get();
  at.base[kArPack] = pack32_.get();
  if (stage != 2) {
    stats_.tile_cfg1 = cfg_for(g1_, GemmPrec::F32, 1, at);
    exec(g1_, GemmPrec::F32, at, stats_.tile_cfg1);
  }
  if (stage != 1) {
    stats_.tile_cfg2 = cfg_for(g2_, GemmPrec::F32_OUT64, 2, at);
    exec(g2_, GemmPrec::F32_OUT64, at, stats_.tile_cfg2);
  }
  if (stage == 0) apply_projectors(d_in, d_out);
}

void HeffPlan::repack_input(const double* d_in, void* dst, int out_type) {
  if (in_tasks_.empty()) return;
  SADMRG_HIP(launch_repack(in_tasks_dev_.as<ScatterTask>(), static_cast<int>(in_tasks_.size()), d_in, nullptr, dst,
                           cplx_ ? 2 : 1, out_type, ctx_->stream()));
}

void HeffPlan::add_projector(const BTensor& phi, double weight) {
  require(!merged_, "H_eff projector: merged plans have no single wave function");
  require(phi.st == layout_.st && phi.off == layout_.off && phi.nelem == n_ && phi.cplx == cplx_,
          "H_eff projector: phi must have the layout of the wave function");
  const int P = static_cast<int>(proj_w_.size()) + 1;
  require(P <= 64, "H_eff projector: at most 64 projectors");
  const int w = cplx_ ? 2 : 1;
  const size_t col = static_cast<size_t>(std::max<int64_t>(n_, 1)) * w * 8;
  DevBuf nb(ctx_, 2 * P * col);
  char* base = nb.as<char>();
  const char* old = proj_.as<char>();
  for (int j = 0; j + 1 < P; ++j) {  // keep the earlier columns
    SADMRG_HIP(hipMemcpyAsync(base + j * col, old + j * col, col, hipMemcpyDeviceToDevice, ctx_->stream()));
    SADMRG_HIP(hipMemcpyAsync(base + (P + j) * col, old + (P - 1 + j) * col, col, hipMemcpyDeviceToDevice,
                              ctx_->stream()));
  }
  if (n_ > 0) {
    double* wcol = reinterpret_cast<double*>(base + (2 * P - 1) * col);
    SADMRG_HIP(hipMemcpyAsync(base + (P - 1) * col, phi.dptr(), n_ * w * 8, hipMemcpyDeviceToDevice, ctx_->stream()));
    SADMRG_HIP(hipMemcpyAsync(wcol, phi.dptr(), n_ * w * 8, hipMemcpyDeviceToDevice, ctx_->stream()));
    SADMRG_HIP(launch_axpby(n_ * w, weight, 0.0, nullptr, wcol, ctx_->stream()));
  }
  ctx_->sync();
  proj_ = std::move(nb);
  proj_dot_ = DevBuf(ctx_, 2 * P * 8);
  proj_w_.push_back(weight);
}

// y += sum_j (weight_j phi_j) <phi_j|x>: one reduction over all phi_j, one update.
void HeffPlan::apply_projectors(const double* d_in, double* d_out) {
  const int P = projectors();
  if (P == 0 || n_ == 0) return;
  const int64_t ld = n_;
  const double* phi = proj_.as<double>();
  const double* wphi = phi + static_cast<int64_t>(P) * ld * (cplx_ ? 2 : 1);
  SADMRG_HIP(launch_gemv_t(phi, ld, P, d_in, n_, cplx_ ? 1 : 0, ctx_->partial(), 1.0, proj_dot_.as<double>(),
                           ctx_->stream()));
  SADMRG_HIP(launch_gemv_n_update(wphi, ld, P, proj_dot_.as<double>(), d_out, n_, cplx_ ? 1 : 0, ctx_->stream()));
}

}  // namespace sadmrg
