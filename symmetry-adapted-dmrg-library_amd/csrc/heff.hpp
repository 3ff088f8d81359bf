// heff.hpp -- fused effective-Hamiltonian matvec plan.
//
// H_eff psi = H_L^T psi + psi H_R + sum_t fac * O1_t^T psi O2_t
// (DMRG::hamiltonian_vector_multiplication_idmrg, dmrg_inst.cpp:283-371, with
// every bond term t evaluated by the block loops of
// BlockTensor::hamiltonian_vector_multiplication, block_tensor_inst.cpp:2170-2318).
//
// All sector GEMMs of one matvec run as two grouped launches:
//   stage 1: one problem per (term, op1 block, psi block):  X = psi_s^T O1_t
//   stage 2: one problem per output block s', whose K segments are
//            H_L^T psi, psi H_R and every (term, s -> s') product X^T O2_t,
//            so the whole right-hand side of an output block accumulates in
//            MFMA registers of the workgroup that owns the tile.
// Operator blocks are packed once per plan into one device arena (and one
// FP32 copy for the mixed-precision path); the block tables are built once
// and reused by every Lanczos iteration.
#pragma once
#include <memory>
#include <vector>
#include "btensor.hpp"

namespace sadmrg {

// Operand precision of the bulk products: FP64; FP32 operands with FP32
// accumulation; BF16 operands with FP32 accumulation (an ablation: 8
// significant bits).  Outputs are FP64 in every case.
enum class Prec : int { FP64 = 0, FP32 = 1, BF16 = 2 };

struct HeffStats {
  double flops = 0;        // algorithmic flops per application
  double bytes64 = 0;      // algorithmic HBM bytes per application (FP64 operands)
  int64_t n = 0;           // psi elements
  int nsec_in = 0;         // psi blocks
  int stage1_probs = 0, stage2_probs = 0;
  int64_t stage1_segs = 0, stage2_segs = 0;
  int tiles1 = 0, tiles2 = 0;
  int tile_cfg1 = 0, tile_cfg2 = 0;  // launch plans of the last application (GemmBatch::kPlans)
  int split_tiles = 0;  // output tiles whose reduction is split over workgroups (both stages)
};

// Tile-configuration autotuner.  On by default: the first application of a
// plan whose (precision, stage, shape signature) is not in the process-wide
// table times every legal configuration once (HIP events on the engine's
// stream) and records the fastest; later plans with the same key reuse it.
// Off: each table uses its shape-based default.
void set_autotune(bool on);
bool autotune();
// Number of entries in the tuning table / forget all entries.
int autotune_entries();
void autotune_clear();

// Process-wide: every H_eff stage runs launch plan `plan` (a GemmBatch plan
// id, persistent ids included) instead of the tuned or default one; -1
// (default) restores tuning.  For tests and measurements of single plans.
void set_heff_launch_plan(int plan);
int heff_launch_plan();

// Process-wide: complex plans run as real tables on planar operands
// (GemmBatch::planar_from; the real kernel's 16-byte staging, split
// reductions, persistent launches and autotuner) instead of the complex
// kernel on interleaved data.  Default off: each complex product becomes
// four real ones that load every operand panel twice, and on small blocks
// the complex kernel, which loads each once for its four MFMA products, is
// faster.
void set_heff_complex_planar(bool on);
bool heff_complex_planar();

// Process-wide: every H_eff application (Lanczos included) runs one launch
// per output block of each stage instead of the grouped launch (the per-block
// baseline, as the reference issues one GEMM per block).  Default off.
void set_heff_per_block(bool on);
bool heff_per_block();

class HeffPlan {
 public:
  // hl / hr may be null tensors.  o1[t] / o2[t] are the bond-term operators
  // (one of each pair is usually a coupling-weighted sum, built by the caller).
  HeffPlan(Context* ctx, CgcTables* cgc, const BTensor& hl, const BTensor& hr, const std::vector<const BTensor*>& o1,
           const std::vector<const BTensor*>& o2, const BTensor& psi_layout);

  // One plan over independent plans (instances of different sizes): their
  // block tables merged into one stage-1 and one stage-2 table, so each
  // stage of every instance runs in a single launch.  Vectors are the
  // instances' vectors back to back, in the order of parts.  The parts must
  // share the context and scalar type and carry no projectors; they may be
  // destroyed afterwards (operators are copied).  No reference counterpart.
  static std::unique_ptr<HeffPlan> merge(const std::vector<const HeffPlan*>& parts);

  // d_in / d_out: device arrays in the layout of psi_layout (n elements).
  // unbatched: one launch per output block of each stage instead of one
  // grouped launch (the per-block baseline).
  // The full application (stage 0) ends with the rank-one terms added by
  // add_projector.
  void apply(const double* d_in, double* d_out, Prec prec, int stage = 0, bool unbatched = false);
  // Adds weight * phi <phi|x> to every full application: the orthogonality
  // penalty of the excited-state solve (dmrg_inst.cpp:352-363, weight =
  // -E_j, phi = the stored state j projected on the current basis).  phi is
  // copied; it must have the structure and block offsets of psi_layout.
  // Evaluated in FP64 at both precisions.
  void add_projector(const BTensor& phi, double weight);
  int projectors() const { return static_cast<int>(proj_w_.size()); }
  bool merged() const { return merged_; }
  // merged plans: element offsets of the instances' vectors (parts + 1
  // entries); a plain plan: {0, n}
  std::vector<int64_t> part_offsets() const { return merged_ ? part_off_ : std::vector<int64_t>{0, n_}; }
  double stage_flops(int stage) const;
  // MFMA flops issued by the launch plans of the last application (stage 1
  // or 2, 0: both) for operands of precision prec: edge tiles computed whole,
  // K rounded up to the K step (GemmBatch::issued_flops).
  double issued_flops(int stage, Prec prec) const;
  const HeffStats& stats() const { return stats_; }
  Context* ctx() const { return ctx_; }
  const BTensor& layout() const { return layout_; }
  bool complex() const { return cplx_; }
  int64_t n() const { return n_; }
  // output blocks that receive no contribution (null in the reference)
  const std::vector<uint8_t>& out_null() const { return out_null_; }

 private:
  explicit HeffPlan(Context* ctx, bool cplx) : ctx_(ctx), cplx_(cplx) {}
  Context* ctx_;
  bool cplx_;
  int64_t n_ = 0;
  BTensor layout_;  // structure / offsets only (no data)
  // Operand layout: every block of the operator pack, of the stage-1
  // intermediate and of the input copy is column-major with its leading
  // dimension padded to a multiple of 4 elements and starts on a multiple of
  // 4 elements (16 bytes in FP32), so the GEMMs stage 16-byte vectors; the
  // padding is zero (written by the pack and input copies, filled once for
  // the intermediate, whose padding rows the GEMMs never write).  The input
  // vector is repacked into that layout (and converted) once per
  // application; the output is written in the caller's layout.
  DevBuf pack64_, pack32_, scratch64_, scratch32_, in32_, in64_;
  DevBuf pack16_, scratch16_, in16_;  // BF16 operand copies (built on first BF16 use)
  int64_t pack_elems_ = 0, scratch_elems_ = 0, in_elems_ = 0;
  std::vector<ScatterTask> in_tasks_;  // input repack table (host copy)
  DevBuf in_tasks_dev_;
  void repack_input(const double* d_in, void* dst, int out_type);
  GemmBatch g1_, g2_;
  int cfg_for(GemmBatch& g, GemmPrec prec, int stage, const ArenaTable& at);
  // complex plans on planar operands: the real tables, the planes' offsets
  // (n4 per arena) and planar copies of the pack, input, scratch and output
  GemmBatch g1p_, g2p_;
  bool planar_built_ = false;
  int64_t n4_[kMaxArenas] = {};
  DevBuf pp64_, pp32_, ip64_, ip32_, sp64_, sp32_, op64_;
  void build_planar();
  void apply_planar(const double* d_in, double* d_out, Prec prec, int stage, bool unbatched);
  HeffStats stats_;
  std::vector<uint8_t> out_null_;
  // projector columns: phi_0..phi_{P-1}, then weight_j * phi_j (n elements each)
  DevBuf proj_, proj_dot_;
  std::vector<double> proj_w_;
  bool merged_ = false;
  std::vector<int64_t> part_off_;
  void apply_projectors(const double* d_in, double* d_out);
};

}  // namespace sadmrg
