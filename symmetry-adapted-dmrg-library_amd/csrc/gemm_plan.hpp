// gemm_plan.hpp -- host builder for grouped GEMM block tables.
//
// A GemmBatch collects output problems (one per output block) and, per
// problem, an ordered list of K segments.  finalize() flattens them into the
// device tables of grouped_gemm.hpp, picks the tile shape and uploads; run()
// launches one grouped kernel.  Tables persist, so a batch built once (e.g.
// the H_eff plan of one Lanczos solve) is re-run with new arena bases.
#pragma once
#include <cstdint>
#include <vector>
#include "core.hpp"
#include "grouped_gemm.hpp"
#include "kernels.hpp"

namespace sadmrg {

struct Operand {
  int arena = 0;
  int64_t off = 0;  // element offset
  int64_t ld = 0;
  bool trans = false;
  bool conj = false;
};

class GemmBatch {
 public:
  int add_problem(int c_arena, int64_t c_off, int64_t ldc, int M, int N, bool accumulate);
  void add_seg(int prob, const Operand& A, const Operand& B, int K, double scale, double scale_im = 0.0);
  // Declares the element count of an arena; finalize() rejects any problem or
  // segment whose extent leaves a declared arena (host-side bounds check).
  void limit(int arena, int64_t elems) { lim_[arena] = elems; }
  int nprob() const { return static_cast<int>(probs_.size()); }
  int64_t nseg() const;
  bool empty() const { return probs_.empty(); }
  // Algorithmic flops (real multiply-adds x2; complex x4) of the table.
  double flops(bool cplx) const;

  // tile_cfg < 0 chooses from the problem shapes.  allow_split lets the
  // planner split long reductions over several workgroups when the table
  // has too few output tiles to fill the GPU (real precisions only).
  void finalize(Context* ctx, int tile_cfg = -1, bool allow_split = true);
  void run(Context* ctx, GemmPrec prec, const ArenaTable& arenas) const;
  int tile_cfg() const { return tile_cfg_; }
  int ntiles() const { return ntiles_; }
  int nsplit_tiles() const { return nreduce_; }
  int nparts() const { return nparts_; }

  // Workgroups the planner aims for before splitting reductions.
  static constexpr int kTargetWorkgroups = 2048;

 private:
  struct P {
    GemmProb p;
    std::vector<GemmSeg> segs;
  };
  static std::vector<GemmTile> xcd_order(const std::vector<GemmTile>& tiles, const std::vector<GemmSeg>& segs);
  std::vector<P> probs_;
  int64_t lim_[kMaxArenas] = {-1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1};
  DevBuf d_tiles_, d_probs_, d_segs_, d_reduce_, d_partial_, d_counters_;
  int ntiles_ = 0;
  int nreduce_ = 0;
  int nparts_ = 0;
  int tile_cfg_ = 0;
  bool finalized_ = false;
};

}  // namespace sadmrg
