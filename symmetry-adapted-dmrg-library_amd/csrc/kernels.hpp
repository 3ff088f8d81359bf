// kernels.hpp -- host-callable launchers for the kernels in kernels.hip.
#pragma once
#include <hip/hip_runtime_api.h>
#include <cstdint>
#include "grouped_gemm.hpp"

namespace sadmrg {

enum class GemmPrec : int {
  F64 = 0,        // double operands, FP64 MFMA, double output
  F32 = 1,        // float operands, FP32 MFMA, float output
  F32_OUT64 = 2,  // float operands, FP32 MFMA, double output
  Z64 = 3,        // complex<double> operands (interleaved), 4 FP64 MFMA products
  BF16 = 4,       // bfloat16 operands, BF16 MFMA with FP32 accumulation, bfloat16 output
  BF16_OUT64 = 5, // bfloat16 operands, BF16 MFMA with FP32 accumulation, double output
  Z32 = 6,        // complex<float> operands, 4 FP32 MFMA products, complex<float> output
  Z32_OUT64 = 7,  // complex<float> operands, 4 FP32 MFMA products, complex<double> output
};

// Device tables of one grouped launch (see grouped_gemm.hpp).  With split
// reductions (nreduce > 0): partial holds nparts dense TM x TN tiles of the
// output type, counters holds nreduce zero-initialised ints that the kernel
// returns to zero, so one table must not run on two streams at once.
struct GemmLaunch {
  const GemmTile* tiles = nullptr;
  const GemmProb* probs = nullptr;
  const GemmSeg* segs = nullptr;
  const GemmReduce* reduce = nullptr;
  void* partial = nullptr;
  int* counters = nullptr;
  ArenaTable arenas{};
  int ntiles = 0;
  int nreduce = 0;
  // Every operand of the table starts on a 16-byte boundary, has a leading
  // dimension that is a multiple of 16 bytes, and is zero between its
  // contiguous extent and that rounded up to 16 bytes: operands are then
  // staged in 16-byte vectors (FP64 and FP32 tables).
  bool vec = false;
  // Persistent launch (real tables): two zeroed ints, the tile ticket and
  // leaver counters; the kernel leaves them zero.  nullptr: one workgroup
  // per tile.
  int* queue = nullptr;
};
// CUs the persistent grid is sized for (MI355X)
constexpr int kPersistCUs = 256;

// Output tile per configuration: 0: 64x64, 1: 128x128, 2: 128x64, 3: 32x32
// (K steps of 16 for FP64, 32 for FP32, 64 for BF16 operands).  Complex
// tables use configurations 0 and 3 (K step 16).
// One launch, split reductions included.
constexpr int kTileConfigs = 4;
hipError_t launch_grouped_gemm(GemmPrec prec, int tile_cfg, const GemmLaunch& launch, hipStream_t st);

inline int tile_m(int tile_cfg) { return tile_cfg == 0 ? 64 : tile_cfg == 3 ? 32 : 128; }
inline int tile_n(int tile_cfg) { return tile_cfg == 1 ? 128 : tile_cfg == 3 ? 32 : 64; }
// tile configurations the complex kernel has
inline bool complex_tile_cfg(int tile_cfg) { return tile_cfg == 0 || tile_cfg == 3; }

// out[c] = sign * sum_i conj(X[i,c]) * y[i] for c < ncol (complex pairs, 2*ncol doubles).
// `partial` must hold 2 * 1024 * ncol doubles.
hipError_t launch_gemv_t(const double* X, int64_t ldx, int ncol, const double* y, int64_t n, int cplx,
                         double* partial, double sign, double* out, hipStream_t st);
// y += sum_c coef[c] * X[:,c]
hipError_t launch_gemv_n_update(const double* X, int64_t ldx, int ncol, const double* coef, double* y, int64_t n,
                                int cplx, hipStream_t st);
// The same with the column count min(*dncol, maxcol) read on the device
// (maxcol <= 64), so a captured launch serves every step of a solve.
hipError_t launch_gemv_t_dev(const double* X, int64_t ldx, int maxcol, const int* dncol, const double* y, int64_t n,
                             int cplx, double* partial, double sign, double* out, hipStream_t st);
hipError_t launch_gemv_n_update_dev(const double* X, int64_t ldx, int maxcol, const int* dncol, const double* coef,
                                    double* y, int64_t n, int cplx, hipStream_t st);

// Device state of a Lanczos solve whose Krylov steps run without host
// synchronisation (and can be captured once in a graph and replayed).
constexpr int kLanczosDevMax = 64;
struct LanczosDev {
  int m;     // column the next step produces
  int stop;  // 1: the step at m had beta below the breakdown threshold
  int pad[2];
  double dot[2];  // <x, w> of the current step (complex pair)
  double nrm[2];  // <w, w>
  double alpha[kLanczosDevMax];
  double beta[kLanczosDevMax + 1];
};
// nd = doubles per vector, ldv = doubles per column of V
hipError_t launch_lanczos_update(LanczosDev* s, double* w, const double* x, const double* V, int64_t ldv, int64_t nd,
                                 hipStream_t st);
hipError_t launch_lanczos_finish(LanczosDev* s, const double* w, double* x, double* V, int64_t ldv, int64_t nd,
                                 double tol, hipStream_t st);
hipError_t launch_lanczos_advance(LanczosDev* s, hipStream_t st);
hipError_t launch_axpby(int64_t n, double a, double b, const double* x, double* y, hipStream_t st);
hipError_t launch_zscal(int64_t n, double sr, double si, const double* x, double* y, hipStream_t st);
hipError_t launch_f64_to_f32(int64_t n, const double* x, float* y, hipStream_t st);
hipError_t launch_f32_to_f64(int64_t n, const float* x, double* y, hipStream_t st);
// round to nearest even; y holds bfloat16 bit patterns
hipError_t launch_f64_to_bf16(int64_t n, const double* x, uint16_t* y, hipStream_t st);
hipError_t launch_fill(int64_t n, double v, double* y, hipStream_t st);
// dst[i + j*ldd] (+)= scale * colscale[j] * src[i*s_row + j*s_col]; elements of w doubles.
hipError_t launch_copy2d(int rows, int cols, const double* src, int64_t s_row, int64_t s_col, double* dst,
                         int64_t ldd, int w, double scale, const double* colscale, int accumulate, hipStream_t st);
// p[t*stride] = v for t < n
hipError_t launch_strided_fill(int64_t n, double v, double* p, int64_t stride, hipStream_t st);
// out[v] = 2-norm of (A[v*s_vec + e*s_elem], e < len) for v < count
hipError_t launch_strided_norms(int count, int len, const double* a, int64_t s_elem, int64_t s_vec, double* out,
                                hipStream_t st);
// A[v*s_vec + e*s_elem] *= f[v] for v < count, e < len
hipError_t launch_strided_scale(int count, int len, double* a, int64_t s_elem, int64_t s_vec, const double* f,
                                hipStream_t st);

constexpr int kPartialGrid = 1024;

// One rectangular piece of a block scatter: dst[r + c*ldd] (+)= scale * src[r + c*lds]
// for r < rows, c < cols (kind 0), or dst[t*(ldd+1)] = scale for t < rows
// (kind 1: a scaled identity stripe).  Offsets are in elements.
// Segmented vector operations of batched solves: a vector holds the
// instances' vectors back to back; SegBlock b covers elements [begin, end)
// of instance `inst` (blocks of one instance consecutive, in element order).
struct SegBlock {
  int64_t begin, end;
  int32_t inst, pad;
};
// partial[(b * ncol + c) * 2 + {0, 1}] = conj(X[:, c]) . y over block b
// (ldx in elements; w = 2 for complex data, conjugate-linear in X).
hipError_t launch_seg_dot(const double* X, int64_t ldx, int ncol, const double* y, const SegBlock* blocks, int nblocks,
                          int cplx, double* partial, hipStream_t st);
// out[(i * ncol + c) * 2 + {0, 1}] = sum of the partials of instance i's
// blocks [first[i], first[i + 1]) in block order (deterministic).
hipError_t launch_seg_reduce(const double* partial, const int* first, int ninst, int ncol, double* out,
                             hipStream_t st);
// y = beta[i] y + sum_c coef[(i * ncol + c) * 2 + {re, im}] X[:, c] on the
// elements of instance i (complex products when cplx); X may be null when
// ncol == 0.
hipError_t launch_seg_combine(const double* X, int64_t ldx, int ncol, const double* coef, const double* beta,
                              double* y, const SegBlock* blocks, int nblocks, int cplx, hipStream_t st);

// Planar complex layout (complex tables lowered onto the real grouped GEMM):
// n complex elements as two real planes, re at [0, n) and im at [n4, n4 + n),
// n4 = n rounded up to a multiple of 4, the tails zero.
// deinterleave: interleaved complex (double) -> planar (out_type 0: double,
// 1: float); interleave: planar double -> interleaved complex.
hipError_t launch_deinterleave(int64_t n, int64_t n4, const double* src, void* dst, int out_type, hipStream_t st);
hipError_t launch_interleave(int64_t n, int64_t n4, const double* src, double* dst, hipStream_t st);

struct ScatterTask {
  int64_t src_off, dst_off, lds, ldd;
  int32_t rows, cols, kind, pad;
  double scale;
};
// Executes a scatter table: one workgroup per task (tasks are pre-split into
// column chunks); w = doubles per element (2 for complex).
hipError_t launch_block_scatter(const ScatterTask* tasks, int ntasks, const double* src, double* dst, int w,
                                hipStream_t st);
// Block copies from FP64 into a padded layout of out_type 0: double, 1: float,
// 2: bfloat16 (round to nearest even), w values per element: rows [0, rows)
// copied and [rows, ldd) zeroed in each column.  srcs (optional): per-task
// source bases indexed by the task's pad field.
hipError_t launch_repack(const ScatterTask* tasks, int ntasks, const double* src, const double* const* srcs,
                         void* dst, int w, int out_type, hipStream_t st);

}  // namespace sadmrg
