// grouped_gemm.hpp -- device-side block table for the grouped GEMM launch.
//
// Every block contraction on the DMRG inner loop (H_eff*psi stage 1 and 2,
// the environment updates, the truncation projectors) is lowered by the host
// planner into this table and executed by one grouped launch:
//
//   C_p  (=|+=)  sum_{s in segs(p)}  scale_s * opA_s(A_s) * opB_s(B_s)
//
// opA(A)[m,k] = transA ? A[k + m*lda] : A[m + k*lda]   (column-major storage)
// opB(B)[k,n] = transB ? B[n + k*ldb] : B[k + n*ldb]
// conjA / conjB conjugate the complex operand before use.
//
// A "tile" is one TM x TN output tile of one problem; the kernel grid walks
// the tile list, so problems of very different shapes share one launch.
#pragma once
#include <cstdint>

namespace sadmrg {

enum GemmFlags : int {
  kTransA = 1,
  kTransB = 2,
  kConjA = 4,
  kConjB = 8,
};

struct GemmSeg {
  int64_t a_off;   // element offset of A inside the A arena
  int64_t b_off;   // element offset of B inside the B arena
  int64_t lda;
  int64_t ldb;
  int32_t K;
  int32_t flags;
  double scale;    // real scale folded into the A operand
  double scale_im; // imaginary part of the scale (complex only)
};

struct GemmProb {
  int64_t c_off;   // element offset of C inside the C arena
  int64_t ldc;
  int32_t M, N;
  int32_t seg_begin, seg_end;
  int32_t accumulate;  // 1: C += result, 0: C = result
  int32_t pad;
};

// One workgroup's work: output tile (m0, n0) of problem `prob`, reduced over
// the problem's segments [seg_lo, seg_hi).  part < 0: the tile owns the whole
// reduction and writes C directly.  part >= 0 (split reduction): the tile
// writes its partial sum, dense TM x TN column-major, into partial slot
// `part` and signs in on the counter of GemmReduce task `red`; the workgroup
// that signs in last folds all slots of the task into C in slot order, so the
// result does not depend on which workgroup finishes last.
struct GemmTile {
  int32_t prob;
  int32_t m0;
  int32_t n0;
  int32_t seg_lo;
  int32_t seg_hi;
  int32_t part;
  int32_t red;
  int32_t pad;
};

// C tile (m0, n0) of problem `prob` (=|+=) sum of partial slots
// [part_begin, part_begin + nparts) in increasing slot order.
struct GemmReduce {
  int32_t prob;
  int32_t m0;
  int32_t n0;
  int32_t part_begin;
  int32_t nparts;
  int32_t pad;
};

// Arenas: each segment addresses A and B relative to an arena base; the
// arena is selected per segment by the top bits of the offset so that one
// launch can read operands from several tensors.  Up to kMaxArenas bases.
constexpr int kMaxArenas = 16;
constexpr int kArenaShift = 56;

inline int64_t arena_off(int arena, int64_t off) {
  return (static_cast<int64_t>(arena) << kArenaShift) | off;
}

struct ArenaTable {
  const void* base[kMaxArenas];
};

}  // namespace sadmrg
