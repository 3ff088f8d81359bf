// gemm_plan.hpp -- host builder for grouped GEMM block tables.
//
// A GemmBatch collects output problems (one per output block) and, per
// problem, an ordered list of K segments.  finalize() checks every access
// against the declared arena sizes and uploads the problem and segment
// tables; the tile list of a tile configuration (grouped_gemm.hpp) is built
// and uploaded the first time that configuration runs.  Tables persist, so a
// batch built once (e.g. the H_eff plan of one Lanczos solve) is re-run with
// new arena bases, in either precision and any tile configuration.
#pragma once
#include <cstdint>
#include <memory>
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
  int nprob() const { return finalized_ ? static_cast<int>(hp_.size()) : static_cast<int>(probs_.size()); }
  int64_t nseg() const;
  // no problem to run (before finalize: none added; after: none with work)
  bool empty() const { return finalized_ ? hp_.empty() : probs_.empty(); }
  // Algorithmic flops (real multiply-adds x2; complex x4) of the table.
  double flops(bool cplx) const;

  // The real table of a finalized complex table on planar operands: an
  // arena of n complex elements becomes two real planes, re at the complex
  // offsets and im at n4[arena] plus them (n4 >= n, a multiple of 4).  Each
  // complex problem becomes two real ones (re and im planes of C), each
  // complex segment s * op(A) op(B) (conjugations and a complex scale
  // included) four real segments, eight when the scale is complex.  This
  // batch must be empty; it is left unfinalized.
  void planar_from(const GemmBatch& cx, const int64_t (&n4)[kMaxArenas]);
  // Appends every problem and segment of a finalized batch, each element
  // offset moved by shift[arena] (independent tables merged into one launch).
  void absorb(const GemmBatch& o, const int64_t (&shift)[kMaxArenas]);
  // The caller guarantees the layout of 16-byte staging: every operand
  // block starts on a multiple of 4 elements, leading dimensions are
  // multiples of 4, and the elements between an operand's contiguous extent
  // and that rounded up to 4 are zero.  finalize() checks the alignment and
  // fails if it does not hold.  Real FP64 / FP32 runs then stage 16-byte
  // vectors.
  void request_vec() { vec_req_ = true; }
  bool vec() const { return vec_; }
  // complex: complex tables (tile configurations 0 and 3, no split
  // reductions).
  void finalize(Context* ctx, bool complex);
  // cfg: a launch plan id (tile configuration + split variant, see kPlans);
  // cfg < 0: the default configuration for prec (default_cfg).
  void run(Context* ctx, GemmPrec prec, const ArenaTable& arenas, int cfg = -1);
  // One launch per problem, over that problem's tiles only and without split
  // reductions: the per-block baseline the grouped launch is measured against.
  void run_per_problem(Context* ctx, GemmPrec prec, const ArenaTable& arenas, int cfg = -1);
  // Default launch plan for a precision.  Tile configuration from the
  // problem shapes: large tiles when most of the work is in problems
  // >= 256 x 256 (128 x 128 for FP32 operands, 128 x 64 for FP64, whose wider
  // accumulators would otherwise cost occupancy).  Split variant: FP64 the
  // finer one (about 8 parts per workgroup slot), otherwise about 3.
  // Complex tables: 32 x 32 tiles when most of the work is in problems with
  // min(M, N) < 64, else 64 x 64.
  int default_cfg(GemmPrec prec) const;
  int big_tiles() const { return big_; }
  // Coarse signature of the problem shapes (work fractions per size class of
  // min(M, N), plus the table's size class): equal keys mark tables for which
  // one tuned tile configuration serves.
  uint64_t shape_key() const { return key_; }
  bool complex() const { return complex_; }
  // MFMA flops a launch plan issues (finalized tables): each output tile
  // computes its whole TM x TN block over every segment of its problem, K
  // rounded up to the precision's K step.  flops() / issued_flops() is the
  // share of issued MFMA work that is not edge-tile padding.
  double issued_flops(int cfg, GemmPrec prec) const;
  // Tiles / split output tiles of a configuration (builds it if needed).
  int ntiles(Context* ctx, int cfg);
  int nsplit_tiles(Context* ctx, int cfg);

  // Workgroups the planner aims for before splitting reductions.
  static constexpr int kTargetWorkgroups = 2048;
  // Workgroups resident at once on the device: 256 CUs, two 256-thread
  // workgroups each (the large-tile configurations' register and LDS budget).
  static constexpr int kWorkgroupSlots = 512;
  // Launch plans: tile configuration (kernels.hpp) x split variant.  Split
  // variant v cuts reductions into parts of equal depth, about
  // kSplitRounds[v] parts per workgroup slot; plan id = cfg + kTileConfigs * v.
  static constexpr int kSplitVariants = 2;
  static constexpr int kSplitRounds[kSplitVariants] = {3, 8};
  static constexpr int kPlans = kTileConfigs * kSplitVariants;
  // Process-wide split policy for tile sets built afterwards: balanced (the
  // default, parts of equal depth) or uniform (every problem cut into the
  // same number of parts whatever its depth; the comparison baseline).
  static void set_balanced_split(bool on);
  static bool balanced_split();
  // Process-wide switch for 16-byte staging of tables that allow it
  // (default on; off: element staging, for comparison).
  static void set_vec_staging(bool on);
  static bool vec_staging();
  // Persistent launches of real tables: a grid of about the resident
  // workgroup count taking tiles from a device ticket counter, instead of
  // one workgroup per tile.  A launch plan id cfg >= kPlans is the
  // persistent launch of plan cfg - kPlans.  Process-wide mode: 0 never
  // (plans >= kPlans run as cfg - kPlans), 1 every launch, 2 (default) the
  // autotuner times both launch kinds per stage (persistent(2) tells it to).
  static void set_persistent(int mode);
  static int persistent();

 private:
  struct P {
    GemmProb p;
    std::vector<GemmSeg> segs;
  };
  struct TileSet {
    bool built = false;
    DevBuf tiles, reduce, partial, counters, queue;
    int ntiles = 0, nreduce = 0, nparts = 0;
  };
  TileSet& tileset(Context* ctx, int cfg);
  struct ProbTiles {
    bool built = false;
    DevBuf tiles;
    std::vector<int> begin;  // tiles of problem p: [begin[p], begin[p + 1])
  };
  ProbTiles pp_[kTileConfigs];
  static std::vector<GemmTile> xcd_order(const std::vector<GemmTile>& tiles, const std::vector<GemmSeg>& segs);

  std::vector<P> probs_;
  int64_t lim_[kMaxArenas] = {-1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1};
  // flattened tables (host copies kept for building tile sets)
  std::vector<GemmProb> hp_;
  std::vector<GemmSeg> hs_;
  std::vector<int64_t> klen_;
  DevBuf d_probs_, d_segs_;
  TileSet sets_[kPlans];
  bool complex_ = false;
  bool big_ = false;
  bool small_ = false;  // most work in problems with min(M, N) < 64
  bool vec_req_ = false, vec_ = false;
  uint64_t key_ = 0;
  bool finalized_ = false;
};

}  // namespace sadmrg
