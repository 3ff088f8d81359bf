// gemm_plan.cpp -- see gemm_plan.hpp.
#include "gemm_plan.hpp"
#include <algorithm>

namespace sadmrg {

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

int64_t GemmBatch::nseg() const {
  int64_t n = 0;
  for (const auto& p : probs_) n += static_cast<int64_t>(p.segs.size());
  return n;
}

double GemmBatch::flops(bool cplx) const {
  double f = 0;
  for (const auto& p : probs_)
    for (const auto& s : p.segs) f += 2.0 * p.p.M * double(p.p.N) * s.K;
  return cplx ? 4.0 * f : f;
}

void GemmBatch::finalize(Context* ctx, int tile_cfg) {
  require(!finalized_, "GemmBatch: finalize twice");
  finalized_ = true;
  if (tile_cfg < 0) {
    // large tiles when the table is dominated by big problems
    double big = 0, all = 0;
    for (const auto& p : probs_) {
      int64_t k = 0;
      for (const auto& s : p.segs) k += s.K;
      const double w = double(p.p.M) * p.p.N * double(k);
      all += w;
      if (p.p.M >= 256 && p.p.N >= 256) big += w;
    }
    tile_cfg = (all > 0 && big / all > 0.6) ? 1 : 0;
  }
  tile_cfg_ = tile_cfg;
  // bounds check of every access against the declared arena sizes
  auto check = [&](int64_t aoff, int64_t rows, int64_t cols, int64_t ld, bool trans, const char* what) {
    const int a = static_cast<int>(static_cast<uint64_t>(aoff) >> kArenaShift);
    const int64_t off = aoff & ((int64_t(1) << kArenaShift) - 1);
    if (a < 0 || a >= kMaxArenas || lim_[a] < 0 || rows <= 0 || cols <= 0) return;
    const int64_t last = trans ? off + (rows - 1) * ld + (cols - 1) : off + (rows - 1) + (cols - 1) * ld;
    if (off < 0 || last >= lim_[a] || (trans ? ld < cols : ld < rows))
      fail(kErrInvalid, std::string("block table out of bounds (") + what + ", arena " + std::to_string(a) + ")");
  };
  for (const auto& p : probs_) {
    if (p.p.M <= 0 || p.p.N <= 0) continue;
    check(p.p.c_off, p.p.M, p.p.N, p.p.ldc, false, "C");
    for (const auto& s : p.segs) {
      check(s.a_off, p.p.M, s.K, s.lda, (s.flags & kTransA) != 0, "A");
      check(s.b_off, s.K, p.p.N, s.ldb, (s.flags & kTransB) != 0, "B");
    }
  }
  const int tm = tile_m(tile_cfg), tn = tile_n(tile_cfg);
  std::vector<GemmProb> hp;
  std::vector<GemmSeg> hs;
  std::vector<GemmTile> ht;
  hp.reserve(probs_.size());
  for (size_t i = 0; i < probs_.size(); ++i) {
    GemmProb p = probs_[i].p;
    if (p.M <= 0 || p.N <= 0) continue;
    p.seg_begin = static_cast<int32_t>(hs.size());
    for (const auto& s : probs_[i].segs) hs.push_back(s);
    p.seg_end = static_cast<int32_t>(hs.size());
    const int pi = static_cast<int>(hp.size());
    hp.push_back(p);
    for (int n0 = 0; n0 < p.N; n0 += tn)
      for (int m0 = 0; m0 < p.M; m0 += tm) ht.push_back(GemmTile{pi, m0, n0, 0});
  }
  // Longest reductions first so the tail of the launch is short tiles.
  std::vector<int64_t> klen(hp.size(), 0);
  for (size_t i = 0; i < hp.size(); ++i)
    for (int s = hp[i].seg_begin; s < hp[i].seg_end; ++s) klen[i] += hs[s].K;
  std::stable_sort(ht.begin(), ht.end(),
                   [&](const GemmTile& a, const GemmTile& b) { return klen[a.prob] > klen[b.prob]; });
  ntiles_ = static_cast<int>(ht.size());
  if (ntiles_ == 0) return;
  d_probs_ = DevBuf(ctx, hp.size() * sizeof(GemmProb));
  d_segs_ = DevBuf(ctx, std::max<size_t>(hs.size(), 1) * sizeof(GemmSeg));
  d_tiles_ = DevBuf(ctx, ht.size() * sizeof(GemmTile));
  SADMRG_HIP(hipMemcpyAsync(d_probs_.get(), hp.data(), hp.size() * sizeof(GemmProb), hipMemcpyHostToDevice,
                            ctx->stream()));
  if (!hs.empty())
    SADMRG_HIP(hipMemcpyAsync(d_segs_.get(), hs.data(), hs.size() * sizeof(GemmSeg), hipMemcpyHostToDevice,
                              ctx->stream()));
  SADMRG_HIP(hipMemcpyAsync(d_tiles_.get(), ht.data(), ht.size() * sizeof(GemmTile), hipMemcpyHostToDevice,
                            ctx->stream()));
  // host vectors go out of scope: make the uploads complete first
  ctx->sync();
}

void GemmBatch::run(Context* ctx, GemmPrec prec, const ArenaTable& arenas) const {
  require(finalized_, "GemmBatch: run before finalize");
  if (ntiles_ == 0) return;
  SADMRG_HIP(launch_grouped_gemm(prec, tile_cfg_, d_tiles_.as<GemmTile>(), d_probs_.as<GemmProb>(),
                                 d_segs_.as<GemmSeg>(), arenas, ntiles_, ctx->stream()));
}

}  // namespace sadmrg
