// gemm_plan.cpp -- see gemm_plan.hpp.
#include "gemm_plan.hpp"
#include <algorithm>
#include <atomic>
#include <map>
#include <utility>

namespace sadmrg {

namespace {
std::atomic<bool> g_vec_staging{true};
std::atomic<int> g_persistent{2};
std::atomic<bool> g_balanced_split{true};
}

void GemmBatch::set_vec_staging(bool on) { g_vec_staging = on; }
bool GemmBatch::vec_staging() { return g_vec_staging; }
void GemmBatch::set_persistent(int mode) { g_persistent = mode; }
int GemmBatch::persistent() { return g_persistent; }
void GemmBatch::set_balanced_split(bool on) { g_balanced_split = on; }
bool GemmBatch::balanced_split() { return g_balanced_split; }

int GemmBatch::add_problem(int c_arena, int64_t c_off, int64_t ldc, int M, int N, bool accumulate) {
  require(!finalized_, "GemmBatch: add after finalize");
  P p{};
  p.p.c_off = arena_off(c_arena, c_off);
  p.p.ldc = ldc;
  p.p.M = M;
  p.p.N = N;
  p.p.accumulate = accumulate ? 1 : 0;
  probs_.push_back(std::move(p));
  return static_cast<int>(probs_.size()) - 1;
}

void GemmBatch::add_seg(int prob, const Operand& A, const Operand& B, int K, double scale, double scale_im) {
  require(prob >= 0 && prob < nprob(), "GemmBatch: bad problem index");
  if (K <= 0) return;
  GemmSeg s{};
  s.a_off = arena_off(A.arena, A.off);
  s.b_off = arena_off(B.arena, B.off);
  s.lda = A.ld;
  s.ldb = B.ld;
  s.K = K;
  s.flags = (A.trans ? kTransA : 0) | (B.trans ? kTransB : 0) | (A.conj ? kConjA : 0) | (B.conj ? kConjB : 0);
  s.scale = scale;
  s.scale_im = scale_im;
  probs_[prob].segs.push_back(s);
}

void GemmBatch::absorb(const GemmBatch& o, const int64_t (&shift)[kMaxArenas]) {
  require(!finalized_, "GemmBatch: absorb after finalize");
  require(o.finalized_, "GemmBatch: absorb needs a finalized source");
  constexpr int64_t mask = (int64_t(1) << kArenaShift) - 1;
  auto move = [&](int64_t off) {
    const int a = static_cast<int>(static_cast<uint64_t>(off) >> kArenaShift);
    return arena_off(a, (off & mask) + shift[a]);
  };
  for (const GemmProb& q : o.hp_) {
    P p{};
    p.p = q;
    p.p.c_off = move(q.c_off);
    for (int s = q.seg_begin; s < q.seg_end; ++s) {
      GemmSeg g = o.hs_[s];
      g.a_off = move(g.a_off);
      g.b_off = move(g.b_off);
      p.segs.push_back(g);
    }
    probs_.push_back(std::move(p));
  }
}

int64_t GemmBatch::nseg() const {
  if (finalized_) return static_cast<int64_t>(hs_.size());
  int64_t n = 0;
  for (const auto& p : probs_) n += static_cast<int64_t>(p.segs.size());
  return n;
}

double GemmBatch::flops(bool cplx) const {
  double f = 0;
  if (finalized_) {
    for (const auto& p : hp_)
      for (int s = p.seg_begin; s < p.seg_end; ++s) f += 2.0 * p.M * double(p.N) * hs_[s].K;
  } else {
    for (const auto& p : probs_)
      for (const auto& s : p.segs) f += 2.0 * p.p.M * double(p.p.N) * s.K;
  }
  return cplx ? 4.0 * f : f;
}

double GemmBatch::issued_flops(int cfg, GemmPrec prec) const {
  require(finalized_, "GemmBatch::issued_flops: table not finalized");
  const int tc = cfg < 0 ? default_cfg(prec) % kTileConfigs : (cfg % kPlans) % kTileConfigs;
  const int tm = tile_m(tc), tn = tile_n(tc);
  int tk = 16;
  double per = 2.0;
  if (complex_)
    per = 8.0;  // four real products per complex step, K step 16
  else
    tk = (prec == GemmPrec::F64) ? 16 : (prec == GemmPrec::BF16 || prec == GemmPrec::BF16_OUT64) ? 64 : 32;
  double f = 0;
  for (const auto& p : hp_) {
    double k = 0;
    for (int s = p.seg_begin; s < p.seg_end; ++s) k += double((hs_[s].K + tk - 1) / tk) * tk;
    const double tiles = double((p.M + tm - 1) / tm) * double((p.N + tn - 1) / tn);
    f += per * tiles * tm * tn * k;
  }
  return f;
}

// Launch order of the tiles.  Workgroups are dealt to the XCDs round-robin
// (workgroup w runs on XCD w % kXcds), and each XCD has its own L2.  Tiles
// that share operand panels -- the tiles of one problem and one reduction
// split -- form a group; groups are balanced over kXcds queues by
// longest-processing-time-first on their reduction work, and the queues are
// interleaved so that queue x occupies workgroup slots x, x + kXcds, ...
// Within a queue the longest reductions come first, and a group's tiles are
// consecutive, so the workgroups resident on one XCD read the same panels.
std::vector<GemmTile> GemmBatch::xcd_order(const std::vector<GemmTile>& tiles, const std::vector<GemmSeg>& segs) {
  constexpr int kXcds = 8;
  struct Group {
    int64_t k = 0;      // reduction length of each tile
    int64_t work = 0;   // k * tiles
    std::vector<int> idx;
  };
  std::vector<Group> groups;
  std::map<std::pair<int, int>, int> gid;
  for (size_t i = 0; i < tiles.size(); ++i) {
    const auto key = std::make_pair(tiles[i].prob, tiles[i].seg_lo);
    auto it = gid.find(key);
    int g;
    if (it == gid.end()) {
      g = static_cast<int>(groups.size());
      gid.emplace(key, g);
      groups.emplace_back();
      for (int s = tiles[i].seg_lo; s < tiles[i].seg_hi; ++s) groups[g].k += segs[s].K;
    } else {
      g = it->second;
    }
    groups[g].idx.push_back(static_cast<int>(i));
  }
  std::vector<int> order(groups.size());
  for (size_t g = 0; g < groups.size(); ++g) {
    order[g] = static_cast<int>(g);
    groups[g].work = groups[g].k * static_cast<int64_t>(groups[g].idx.size());
  }
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return groups[a].k > groups[b].k; });
  std::vector<std::vector<int>> queue(kXcds);
  std::vector<int64_t> load(kXcds, 0);
  for (int g : order) {
    const int x = static_cast<int>(std::min_element(load.begin(), load.end()) - load.begin());
    load[x] += groups[g].work + 1;
    for (int i : groups[g].idx) queue[x].push_back(i);
  }
  std::vector<GemmTile> out;
  out.reserve(tiles.size());
  std::vector<size_t> pos(kXcds, 0);
  for (bool more = true; more;) {
    more = false;
    for (int x = 0; x < kXcds; ++x)
      if (pos[x] < queue[x].size()) {
        out.push_back(tiles[queue[x][pos[x]++]]);
        more = true;
      }
  }
  return out;
}

void GemmBatch::finalize(Context* ctx, bool complex) {
  require(!finalized_, "GemmBatch: finalize twice");
  finalized_ = true;
  complex_ = complex;
  // large tiles when the table is dominated by big problems
  double big = 0, all = 0;
  for (const auto& p : probs_) {
    int64_t k = 0;
    for (const auto& s : p.segs) k += s.K;
    const double w = double(p.p.M) * p.p.N * double(k);
    all += w;
    if (p.p.M >= 256 && p.p.N >= 256) big += w;
  }
  big_ = all > 0 && big / all > 0.6;
  {
    // work fractions (eighths) in min(M, N) classes <64, <128, <256, <512, >=512
    double cls[5] = {0, 0, 0, 0, 0};
    for (const auto& p : probs_) {
      int64_t k = 0;
      for (const auto& s : p.segs) k += s.K;
      const int mn = std::min(p.p.M, p.p.N);
      const int c = mn < 64 ? 0 : mn < 128 ? 1 : mn < 256 ? 2 : mn < 512 ? 3 : 4;
      cls[c] += double(p.p.M) * p.p.N * double(k);
    }
    small_ = all > 0 && cls[0] / all > 0.5;
    uint64_t key = 0;
    for (int c = 0; c < 5; ++c) key = key * 9 + (all > 0 ? static_cast<uint64_t>(8.0 * cls[c] / all + 0.5) : 0);
    int lg = 0;
    for (double f = all; f >= 2.0 && lg < 63; f /= 2.0) ++lg;
    key_ = key * 64 + static_cast<uint64_t>(lg / 2);
  }
  // bounds check of every access against the declared arena sizes
  auto check = [&](int64_t aoff, int64_t rows, int64_t cols, int64_t ld, bool trans, const char* what) {
    const int a = static_cast<int>(static_cast<uint64_t>(aoff) >> kArenaShift);
    const int64_t off = aoff & ((int64_t(1) << kArenaShift) - 1);
    if (a < 0 || a >= kMaxArenas || lim_[a] < 0 || rows <= 0 || cols <= 0) return;
    const int64_t last = trans ? off + (rows - 1) * ld + (cols - 1) : off + (rows - 1) + (cols - 1) * ld;
    if (off < 0 || last >= lim_[a] || (trans ? ld < cols : ld < rows))
      fail(kErrInvalid, std::string("block table out of bounds (") + what + ", arena " + std::to_string(a) + ")");
    // the kernels address an operand with 32-bit byte offsets from its first
    // element (16-byte complex elements at most)
    if ((last - off + 1) > (int64_t(1) << 27))
      fail(kErrUnsupported, std::string("operand block above 2^27 elements (") + what + ")");
  };
  for (const auto& p : probs_) {
    if (p.p.M <= 0 || p.p.N <= 0) continue;
    check(p.p.c_off, p.p.M, p.p.N, p.p.ldc, false, "C");
    for (const auto& s : p.segs) {
      check(s.a_off, p.p.M, s.K, s.lda, (s.flags & kTransA) != 0, "A");
      check(s.b_off, s.K, p.p.N, s.ldb, (s.flags & kTransB) != 0, "B");
    }
  }
  if (vec_req_) {
    constexpr int64_t mask = (int64_t(1) << kArenaShift) - 1;
    for (const auto& p : probs_)
      for (const auto& g : p.segs)
        require((g.a_off & mask) % 4 == 0 && (g.b_off & mask) % 4 == 0 && g.lda % 4 == 0 && g.ldb % 4 == 0,
                "GemmBatch: 16-byte staging requested for an operand that is not 4-element aligned");
    vec_ = true;
  }
  hp_.reserve(probs_.size());
  for (size_t i = 0; i < probs_.size(); ++i) {
    GemmProb p = probs_[i].p;
    if (p.M <= 0 || p.N <= 0) continue;
    p.seg_begin = static_cast<int32_t>(hs_.size());
    int64_t k = 0;
    for (const auto& s : probs_[i].segs) {
      hs_.push_back(s);
      k += s.K;
    }
    p.seg_end = static_cast<int32_t>(hs_.size());
    hp_.push_back(p);
    klen_.push_back(k);
  }
  probs_.clear();
  probs_.shrink_to_fit();
  if (hp_.empty()) return;
  d_probs_ = DevBuf(ctx, hp_.size() * sizeof(GemmProb));
  d_segs_ = DevBuf(ctx, std::max<size_t>(hs_.size(), 1) * sizeof(GemmSeg));
  ctx->upload(d_probs_.get(), hp_.data(), hp_.size() * sizeof(GemmProb));
  if (!hs_.empty()) ctx->upload(d_segs_.get(), hs_.data(), hs_.size() * sizeof(GemmSeg));
}

void GemmBatch::planar_from(const GemmBatch& cx, const int64_t (&n4)[kMaxArenas]) {
  require(cx.finalized_ && cx.complex_, "GemmBatch::planar_from: needs a finalized complex table");
  require(!finalized_ && probs_.empty(), "GemmBatch::planar_from: target must be empty");
  const int64_t mask = (int64_t(1) << kArenaShift) - 1;
  auto opnd = [&](int64_t aoff, int64_t ld, bool trans, bool im) {
    Operand o;
    o.arena = static_cast<int>(static_cast<uint64_t>(aoff) >> kArenaShift);
    o.off = (aoff & mask) + (im ? n4[o.arena] : 0);
    o.ld = ld;
    o.trans = trans;
    o.conj = false;
    return o;
  };
  for (const GemmProb& p : cx.hp_) {
    const int ca = static_cast<int>(static_cast<uint64_t>(p.c_off) >> kArenaShift);
    const int64_t co = p.c_off & mask;
    const int pr = add_problem(ca, co, p.ldc, p.M, p.N, p.accumulate != 0);
    const int pi = add_problem(ca, n4[ca] + co, p.ldc, p.M, p.N, p.accumulate != 0);
    for (int si = p.seg_begin; si < p.seg_end; ++si) {
      const GemmSeg& g = cx.hs_[si];
      const bool ta = (g.flags & kTransA) != 0, tb = (g.flags & kTransB) != 0;
      const double sa = (g.flags & kConjA) ? -1.0 : 1.0, sb = (g.flags & kConjB) ? -1.0 : 1.0;
      const Operand Ar = opnd(g.a_off, g.lda, ta, false), Ai = opnd(g.a_off, g.lda, ta, true);
      const Operand Br = opnd(g.b_off, g.ldb, tb, false), Bi = opnd(g.b_off, g.ldb, tb, true);
      // s A B with A = Ar + i sa Ai, B = Br + i sb Bi, s = sr + i sm:
      //   re = sr (Ar Br - sa sb Ai Bi) - sm (sa Ai Br + sb Ar Bi)
      //   im = sr (sa Ai Br + sb Ar Bi) + sm (Ar Br - sa sb Ai Bi)
      const double sr = g.scale, sm = g.scale_im;
      if (sr != 0.0) {
        add_seg(pr, Ar, Br, g.K, sr);
        add_seg(pr, Ai, Bi, g.K, -sr * sa * sb);
        add_seg(pi, Ai, Br, g.K, sr * sa);
        add_seg(pi, Ar, Bi, g.K, sr * sb);
      }
      if (sm != 0.0) {
        add_seg(pr, Ai, Br, g.K, -sm * sa);
        add_seg(pr, Ar, Bi, g.K, -sm * sb);
        add_seg(pi, Ar, Br, g.K, sm);
        add_seg(pi, Ai, Bi, g.K, -sm * sa * sb);
      }
    }
  }
}

int GemmBatch::default_cfg(GemmPrec prec) const {
  if (complex_) return small_ ? 3 : 0;
  // FP64 parts take about twice as long per K step as FP32 ones, so FP64
  // tables balance best with the finer split variant
  const int variant = prec == GemmPrec::F64 ? kTileConfigs : 0;
  if (!big_) return variant;
  return (prec == GemmPrec::F64 ? 2 : 1) + variant;
}

GemmBatch::TileSet& GemmBatch::tileset(Context* ctx, int cfg) {
  require(finalized_, "GemmBatch: run before finalize");
  require(cfg >= 0 && cfg < kPlans, "GemmBatch: bad launch plan");
  require(!complex_ || complex_tile_cfg(cfg % kTileConfigs), "GemmBatch: complex tables use tile configurations 0 and 3");
  TileSet& ts = sets_[cfg];
  if (ts.built) return ts;
  ts.built = true;
  const int tm = tile_m(cfg % kTileConfigs), tn = tile_n(cfg % kTileConfigs);
  int64_t base_tiles = 0;
  for (const auto& p : hp_) base_tiles += int64_t((p.M + tm - 1) / tm) * ((p.N + tn - 1) / tn);
  // Split reductions, for tables with fewer output tiles than
  // kTargetWorkgroups.  Each problem's reduction is cut at segment boundaries,
  // by cumulative K, into parts of about equal length L (at least kMinSplitK),
  // L chosen so that the table yields about `rounds` parts per workgroup slot
  // of the device: every slot then runs about the same reduction work and the
  // launch has no long tail of the deepest problems.  Uniform policy
  // (rounds = 0): every problem is cut into the same number of parts (enough
  // workgroups to fill the GPU), whatever its depth.
  constexpr int64_t kMinSplitK = 512;
  const bool splitting = !complex_ && base_tiles > 0 && base_tiles < kTargetWorkgroups;
  const int rounds = balanced_split() ? kSplitRounds[cfg / kTileConfigs] : 0;
  int want = 1;
  int64_t part_len = 0;
  if (splitting && rounds == 0) {
    want = static_cast<int>((kTargetWorkgroups + base_tiles - 1) / base_tiles);
  } else if (splitting) {
    int64_t total = 0;
    for (size_t pi = 0; pi < hp_.size(); ++pi)
      total += int64_t((hp_[pi].M + tm - 1) / tm) * ((hp_[pi].N + tn - 1) / tn) * klen_[pi];
    part_len = std::max<int64_t>(kMinSplitK, total / (int64_t(kWorkgroupSlots) * rounds));
  }
  std::vector<GemmTile> ht;
  std::vector<GemmReduce> hr;
  int nparts = 0;
  for (size_t pi = 0; pi < hp_.size(); ++pi) {
    const GemmProb& p = hp_[pi];
    const int nseg = p.seg_end - p.seg_begin;
    const int64_t parts = part_len > 0 ? (klen_[pi] + part_len / 2) / part_len : int64_t(want);
    int split = static_cast<int>(std::min<int64_t>({parts, int64_t(nseg), klen_[pi] / kMinSplitK}));
    // segment groups [cut[g], cut[g+1]) with near-equal K
    std::vector<int> cut{p.seg_begin};
    if (split > 1) {
      int64_t acc = 0;
      int g = 1;
      for (int s = p.seg_begin; s < p.seg_end && g < split; ++s) {
        acc += hs_[s].K;
        if (acc * split >= klen_[pi] * g && s + 1 < p.seg_end) {
          cut.push_back(s + 1);
          ++g;
        }
      }
    }
    cut.push_back(p.seg_end);
    split = static_cast<int>(cut.size()) - 1;
    for (int n0 = 0; n0 < p.N; n0 += tn)
      for (int m0 = 0; m0 < p.M; m0 += tm) {
        if (split <= 1) {
          ht.push_back(GemmTile{static_cast<int32_t>(pi), m0, n0, p.seg_begin, p.seg_end, -1, -1, 0});
          continue;
        }
        const int red = static_cast<int>(hr.size());
        hr.push_back(GemmReduce{static_cast<int32_t>(pi), m0, n0, nparts, split, 0});
        for (int g = 0; g < split; ++g)
          ht.push_back(GemmTile{static_cast<int32_t>(pi), m0, n0, cut[g], cut[g + 1], nparts + g, red, 0});
        nparts += split;
      }
  }
  ht = xcd_order(ht, hs_);
  ts.ntiles = static_cast<int>(ht.size());
  ts.nreduce = static_cast<int>(hr.size());
  ts.nparts = nparts;
  if (ts.ntiles == 0) return ts;
  ts.tiles = DevBuf(ctx, ht.size() * sizeof(GemmTile));
  ctx->upload(ts.tiles.get(), ht.data(), ht.size() * sizeof(GemmTile));
  if (ts.nreduce > 0) {
    ts.reduce = DevBuf(ctx, hr.size() * sizeof(GemmReduce));
    ctx->upload(ts.reduce.get(), hr.data(), hr.size() * sizeof(GemmReduce));
    // room for double partials (the widest output type)
    ts.partial = DevBuf(ctx, int64_t(nparts) * tm * tn * sizeof(double));
    ts.counters = DevBuf(ctx, hr.size() * sizeof(int));
    SADMRG_HIP(hipMemsetAsync(ts.counters.get(), 0, hr.size() * sizeof(int), ctx->stream()));
  }
  // ticket and leaver counters of persistent launches (left zero by every launch)
  ts.queue = DevBuf(ctx, 2 * sizeof(int));
  SADMRG_HIP(hipMemsetAsync(ts.queue.get(), 0, 2 * sizeof(int), ctx->stream()));
  return ts;
}

int GemmBatch::ntiles(Context* ctx, int cfg) { return tileset(ctx, cfg).ntiles; }
int GemmBatch::nsplit_tiles(Context* ctx, int cfg) { return tileset(ctx, cfg).nreduce; }

void GemmBatch::run(Context* ctx, GemmPrec prec, const ArenaTable& arenas, int cfg) {
  require(finalized_, "GemmBatch: run before finalize");
  if (hp_.empty()) return;
  if (cfg < 0) cfg = default_cfg(prec);
  require(cfg < 2 * kPlans, "GemmBatch: bad launch plan");
  const int mode = g_persistent;
  const bool pers = !complex_ && (mode == 1 || (mode == 2 && cfg >= kPlans));
  cfg %= kPlans;
  const bool zprec = prec == GemmPrec::Z64 || prec == GemmPrec::Z32 || prec == GemmPrec::Z32_OUT64;
  require(zprec == complex_, "GemmBatch: precision and table disagree on complex data");
  const TileSet& ts = tileset(ctx, cfg);
  if (ts.ntiles == 0) return;
  GemmLaunch L;
  L.tiles = ts.tiles.as<GemmTile>();
  L.probs = d_probs_.as<GemmProb>();
  L.segs = d_segs_.as<GemmSeg>();
  L.reduce = ts.nreduce > 0 ? ts.reduce.as<GemmReduce>() : nullptr;
  L.partial = ts.nreduce > 0 ? ts.partial.get() : nullptr;
  L.counters = ts.nreduce > 0 ? ts.counters.as<int>() : nullptr;
  L.arenas = arenas;
  L.ntiles = ts.ntiles;
  L.nreduce = ts.nreduce;
  L.vec = vec_ && g_vec_staging && !complex_;  // every real operand type
  L.queue = pers ? ts.queue.as<int>() : nullptr;
  SADMRG_HIP(launch_grouped_gemm(prec, cfg % kTileConfigs, L, ctx->stream()));
}

void GemmBatch::run_per_problem(Context* ctx, GemmPrec prec, const ArenaTable& arenas, int cfg) {
  require(finalized_, "GemmBatch: run before finalize");
  if (hp_.empty()) return;
  if (cfg < 0) cfg = default_cfg(prec);
  require(cfg >= 0 && cfg < 2 * kPlans, "GemmBatch: bad launch plan");
  cfg %= kTileConfigs;  // no split reductions or persistence: the tile configuration only
  require(!complex_ || complex_tile_cfg(cfg), "GemmBatch: complex tables use tile configurations 0 and 3");
  const bool zprec = prec == GemmPrec::Z64 || prec == GemmPrec::Z32 || prec == GemmPrec::Z32_OUT64;
  require(zprec == complex_, "GemmBatch: precision and table disagree on complex data");
  ProbTiles& pt = pp_[cfg];
  if (!pt.built) {
    pt.built = true;
    const int tm = tile_m(cfg), tn = tile_n(cfg);
    std::vector<GemmTile> ht;
    pt.begin.assign(hp_.size() + 1, 0);
    for (size_t pi = 0; pi < hp_.size(); ++pi) {
      pt.begin[pi] = static_cast<int>(ht.size());
      const GemmProb& p = hp_[pi];
      for (int n0 = 0; n0 < p.N; n0 += tn)
        for (int m0 = 0; m0 < p.M; m0 += tm)
          ht.push_back(GemmTile{static_cast<int32_t>(pi), m0, n0, p.seg_begin, p.seg_end, -1, -1, 0});
    }
    pt.begin[hp_.size()] = static_cast<int>(ht.size());
    if (!ht.empty()) {
      pt.tiles = DevBuf(ctx, ht.size() * sizeof(GemmTile));
      ctx->upload(pt.tiles.get(), ht.data(), ht.size() * sizeof(GemmTile));
    }
  }
  for (size_t pi = 0; pi < hp_.size(); ++pi) {
    const int n = pt.begin[pi + 1] - pt.begin[pi];
    if (n <= 0) continue;
    GemmLaunch L;
    L.tiles = pt.tiles.as<GemmTile>() + pt.begin[pi];
    L.probs = d_probs_.as<GemmProb>();
    L.segs = d_segs_.as<GemmSeg>();
    L.reduce = nullptr;
    L.partial = nullptr;
    L.counters = nullptr;
    L.arenas = arenas;
    L.ntiles = n;
    L.nreduce = 0;
    L.vec = vec_ && g_vec_staging && !complex_;  // every real operand type
    SADMRG_HIP(launch_grouped_gemm(prec, cfg, L, ctx->stream()));
  }
}

}  // namespace sadmrg
