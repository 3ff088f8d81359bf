// kernels.hip -- hand-written gfx950 kernels for the two-site DMRG inner loop.
//
//  * grouped_gemm_kernel : one launch executes a whole block table (see
//    grouped_gemm.hpp).  Each workgroup owns one TM x TN output tile of one
//    problem and walks that problem's K segments (all sector pairs / bond
//    terms that feed the tile), accumulating in MFMA registers.  No atomics:
//    every output element is produced by exactly one workgroup in a fixed
//    order, so results are bitwise reproducible run to run.
//      FP64: v_mfma_f64_16x16x4f64, FP32: v_mfma_f32_16x16x4f32,
//      BF16 (FP32 accumulation): v_mfma_f32_16x16x16_bf16.
//  * vector kernels used by Lanczos: deterministic two-stage reductions for
//    dot products / norms, axpy with a device-resident coefficient, block
//    Gram-Schmidt (V^T w and w -= V c), precision conversion.
#include <hip/hip_runtime.h>
#include <cstdint>
#include <type_traits>
#include "grouped_gemm.hpp"
#include "kernels.hpp"

namespace sadmrg {

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf16_t;
typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));

__device__ inline d4_t mfma_16x16x4(double a, double b, d4_t c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ inline f4_t mfma_16x16x4(float a, float b, f4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

template <class T> struct Acc4;
template <> struct Acc4<double> { typedef d4_t type; };
template <> struct Acc4<float> { typedef f4_t type; };
template <> struct Acc4<bf16_t> { typedef f4_t type; };

// Segment scale factors are applied in this type when a K step is staged
// (BF16 operands are scaled in FP32 and rounded once); split-reduction
// partial tiles are held in PartT.
template <class T> struct ScaleT { typedef T type; };
template <> struct ScaleT<bf16_t> { typedef float type; };
template <class T> struct PartT { typedef T type; };
template <> struct PartT<bf16_t> { typedef float type; };
template <class T> __device__ inline T scaled(T v, typename ScaleT<T>::type s) { return v * s; }
template <> __device__ inline bf16_t scaled<bf16_t>(bf16_t v, float s) {
  return static_cast<bf16_t>(static_cast<float>(v) * s);
}
// Values of one k-slot group a lane reads per fragment: four FP32 or BF16,
// two FP64 (16 bytes, or 8 for BF16).
template <class T> struct FragW { static constexpr int NF = sizeof(T) == 8 ? 2 : 4; };

// Output row of accumulator element r held by lane group g = lane / 16 (the
// column is lane % 16).  The FP64 16x16x4 MFMA interleaves rows across lane
// groups (row = g + 4r); the FP32 16x16x4 MFMA gives each lane group four
// consecutive rows (row = 4g + r).
template <class T> __device__ inline int acc_row(int g, int r);
template <> __device__ inline int acc_row<double>(int g, int r) { return g + 4 * r; }
template <> __device__ inline int acc_row<float>(int g, int r) { return 4 * g + r; }
template <> __device__ inline int acc_row<bf16_t>(int g, int r) { return 4 * g + r; }

// The arena bases are device allocations: the pointer is formed in the
// global address space, so accesses through it compile to global (not FLAT)
// memory instructions.
__device__ inline const void* arena_ptr(const ArenaTable& at, int64_t off, int elem_size) {
  typedef __attribute__((address_space(1))) const char gchar;
  const int a = static_cast<int>(static_cast<uint64_t>(off) >> kArenaShift);
  const int64_t o = off & ((int64_t(1) << kArenaShift) - 1);
  gchar* g = (gchar*)(at.base[a]);
  return (const void*)(g + o * elem_size);
}

// ---------------------------------------------------------------------------
// Real grouped GEMM.
//
// Workgroup = 256 threads = 2 x 2 waves, output tile TM x TN, K steps of TK.
// Operand tiles are staged through LDS in the orientation they have in HBM,
// so every global load is coalesced along the contiguous index:
//   RK: element (r, k) at r + k*ld  (A plain, B transposed) -> lds[k*LR + r]
//   KR: element (r, k) at k + r*ld  (A transposed, B plain) -> lds[kr_off(r, k)]
// ("r" is m for A and n for B).  MFMA k-slot order: in each 16-deep chunk,
// lane group g = lane/16 feeds k = 4g + j to the j-th of four MFMAs, for A
// and B alike, so a KR tile is read with one 16-byte LDS load per fragment.
// RK rows are padded, LR = R + 4, which puts the rows 4g (g = 0..3) read by
// one instruction on distinct bank groups.  KR rows are TK elements with the
// 16-byte groups of row r XOR-swizzled by (r/2) mod 8: the two rows a wave
// stores per instruction land in opposite bank halves, and the 16 rows of a
// fragment read land in 16 distinct 16-byte bank slots.  LDS is double
// buffered with one barrier per K step; global loads run two steps ahead of
// the MFMAs.
template <class T, int R, int TK>
struct OperandTile {
  static constexpr int E = R * TK / 256;  // elements staged per thread and step
  static constexpr int NQ = 16 / sizeof(T);  // elements per 16-byte group
  static constexpr int LR = R + 4;
  static constexpr int LK = TK;
  static constexpr int ELEMS = (TK * LR > R * LK) ? TK * LR : R * LK;
  static_assert(256 % R == 0 && 256 % TK == 0 && TK % 16 == 0, "tile shape");
  static_assert(TK / NQ == 8, "KR swizzle assumes 8 groups of 16 bytes per row");
  // element offset of (r, k) in a KR tile
  __device__ static inline int kr_off(int r, int k) { return r * LK + (((k / NQ) ^ ((r >> 1) & 7)) * NQ) + k % NQ; }
};

// Per-thread view of one operand of the current segment.
template <class T>
struct OperandCursor {
  const T* safe;    // an element that always exists (the operand's first)
  const T* p;       // this thread's element 0 of the current step
  int64_t estride;  // element e is p[e * estride]
  int64_t kstep;    // p advance per K step
  int rr;           // rows left for this thread's first element (<= 0: none)
  int kk;           // k left for this thread's first element at the current step
  int epos;         // per-e increment of the varying index (rows for KR, k for RK)
  bool kr;
};

template <class T, int R, int TK>
__device__ inline void cursor_init(OperandCursor<T>& c, const T* base, int64_t ld, bool kr, int r0, int rows, int K,
                                   int tid) {
  c.kr = kr;
  c.safe = base;
  if (!kr) {  // RK
    const int rt = tid % R, kt = tid / R;
    c.p = base + r0 + rt + int64_t(kt) * ld;
    c.epos = 256 / R;
    c.estride = int64_t(c.epos) * ld;
    c.kstep = int64_t(TK) * ld;
    c.rr = rows - r0 - rt;
    c.kk = K - kt;
  } else {    // KR
    const int kt = tid % TK, rt = tid / TK;
    c.p = base + kt + int64_t(r0 + rt) * ld;
    c.epos = 256 / TK;
    c.estride = int64_t(c.epos) * ld;
    c.kstep = TK;
    c.rr = rows - r0 - rt;
    c.kk = K - kt;
  }
}

template <class T, int E>
__device__ inline void cursor_load(const OperandCursor<T>& c, T (&v)[E]) {
#pragma unroll
  for (int e = 0; e < E; ++e) {
    // branch-free: out-of-range lanes read a valid element and drop it
    const bool ok = c.kr ? (c.kk > 0 && e * c.epos < c.rr) : (c.rr > 0 && e * c.epos < c.kk);
    const T x = *(ok ? c.p + e * c.estride : c.safe);
    v[e] = ok ? x : T(0);
  }
}

template <class T>
__device__ inline void cursor_advance(OperandCursor<T>& c, int TK) {
  c.p += c.kstep;
  c.kk -= TK;
}

// ---- vector staging (tables whose operands are aligned to 4 elements and
// zero-padded to a leading dimension of 4) -----------------------------------
// A thread stages VW consecutive elements along the operand's contiguous
// index per vector, NV vectors per K step: 16 bytes (4 FP32 or 2 FP64
// values), or 8 bytes (4 BF16 values, the 4-element alignment of the
// layout).  Vec16 is the 16-byte view of an LDS row used by fragment reads.
template <class T> struct Vec16;
template <> struct Vec16<float> { typedef float4 type; };
template <> struct Vec16<double> { typedef double2 type; };
template <> struct Vec16<bf16_t> { typedef uint4 type; };
template <class T> struct VecS { typedef typename Vec16<T>::type type; };
template <> struct VecS<bf16_t> { typedef uint2 type; };
// the segment scale applied to a staged vector (BF16: in FP32, rounded once
// per element, as scaled() does for single elements)
template <class V, class S> __device__ inline V vscaled(const V& v, S s) { return v * s; }
template <> __device__ inline uint2 vscaled<uint2, float>(const uint2& v, float s) {
  auto one = [&](unsigned short h) {
    const bf16_t x = static_cast<bf16_t>(static_cast<float>(__builtin_bit_cast(bf16_t, h)) * s);
    return static_cast<unsigned>(__builtin_bit_cast(unsigned short, x));
  };
  uint2 r;
  r.x = one(static_cast<unsigned short>(v.x & 0xffffu)) | (one(static_cast<unsigned short>(v.x >> 16)) << 16);
  r.y = one(static_cast<unsigned short>(v.y & 0xffffu)) | (one(static_cast<unsigned short>(v.y >> 16)) << 16);
  return r;
}

template <class T, int R, int TK>
struct VecTile {
  static constexpr int VW = sizeof(typename VecS<T>::type) / sizeof(T);
  static constexpr int NV = R * TK / (256 * VW);  // vectors per thread and step
  static_assert(NV >= 1 && (R / VW) <= 256 && 256 % (R / VW) == 0 && 256 % (TK / VW) == 0, "vector tile shape");
};

template <class T, int R, int TK>
__device__ inline void vcursor_init(OperandCursor<T>& c, const T* base, int64_t ld, bool kr, int r0, int rows, int K,
                                    int tid) {
  constexpr int VW = VecTile<T, R, TK>::VW;
  c.kr = kr;
  c.safe = base;
  if (!kr) {  // RK: VW consecutive rows at one k
    constexpr int NR = R / VW;
    const int rv = (tid % NR) * VW, kt = tid / NR;
    c.p = base + r0 + rv + int64_t(kt) * ld;
    c.epos = 256 / NR;
    c.estride = int64_t(c.epos) * ld;
    c.kstep = int64_t(TK) * ld;
    c.rr = rows - r0 - rv;  // the vector is used if it starts inside the rows
    c.kk = K - kt;
  } else {    // KR: VW consecutive k of one row
    constexpr int NK = TK / VW;
    const int kv = (tid % NK) * VW, rt = tid / NK;
    c.p = base + kv + int64_t(r0 + rt) * ld;
    c.epos = 256 / NK;
    c.estride = int64_t(c.epos) * ld;
    c.kstep = TK;
    c.rr = rows - r0 - rt;
    c.kk = K - kv;          // elements past K in the vector are the zero padding
  }
}

template <class T, int NV>
__device__ inline void vcursor_load(const OperandCursor<T>& c, typename VecS<T>::type (&v)[NV]) {
  typedef typename VecS<T>::type V;
#pragma unroll
  for (int e = 0; e < NV; ++e) {
    const bool ok = c.kr ? (c.kk > 0 && e * c.epos < c.rr) : (c.rr > 0 && e * c.epos < c.kk);
    const V x = *reinterpret_cast<const V*>(ok ? c.p + e * c.estride : c.safe);
    v[e] = ok ? x : V{};
  }
}

template <class T, int R, int TK>
__device__ inline void vtile_store(T* lds, bool kr, const typename VecS<T>::type (&v)[VecTile<T, R, TK>::NV],
                                   typename ScaleT<T>::type scale, int tid) {
  using OT = OperandTile<T, R, TK>;
  typedef typename VecS<T>::type V;
  constexpr int VW = VecTile<T, R, TK>::VW;
  constexpr int NV = VecTile<T, R, TK>::NV;
  if (!kr) {
    constexpr int NR = R / VW;
    const int rv = (tid % NR) * VW, kt = tid / NR;
#pragma unroll
    for (int e = 0; e < NV; ++e)
      *reinterpret_cast<V*>(&lds[(kt + e * (256 / NR)) * OT::LR + rv]) = vscaled(v[e], scale);
  } else {
    // a vector of VW <= NQ elements at a multiple of VW stays inside one
    // swizzled 16-byte group
    constexpr int NK = TK / VW;
    const int kv = (tid % NK) * VW, rt = tid / NK;
#pragma unroll
    for (int e = 0; e < NV; ++e)
      *reinterpret_cast<V*>(&lds[OT::kr_off(rt + e * (256 / NK), kv)]) = vscaled(v[e], scale);
  }
}

template <class T, int R, int TK>
__device__ inline void tile_store(T* lds, bool kr, const T (&v)[OperandTile<T, R, TK>::E],
                                  typename ScaleT<T>::type scale, int tid) {
  using OT = OperandTile<T, R, TK>;
  if (!kr) {
    const int rt = tid % R, kt = tid / R;
#pragma unroll
    for (int e = 0; e < OT::E; ++e) lds[(kt + e * (256 / R)) * OT::LR + rt] = scaled(v[e], scale);
  } else {
    const int kt = tid % TK, rt = tid / TK;
#pragma unroll
    for (int e = 0; e < OT::E; ++e) lds[OT::kr_off(rt + e * (256 / TK), kt)] = scaled(v[e], scale);
  }
}

// KR fragment reads use the 16-byte view Vec16 of LDS (rows of LK elements,
// LK a multiple of 16 bytes, so a 4-slot group never straddles a 16-byte word).

// NF = FragW<T>::NF consecutive k-slot values (j = j0 .. j0+NF-1) of
// fragment row `row` in chunk kc.
template <class T, int R, int TK, bool KR>
__device__ inline void tile_frag(const T* lds, int row, int kc, int g, int j0, T (&f)[FragW<T>::NF]) {
  using OT = OperandTile<T, R, TK>;
  constexpr int NQ = 16 / sizeof(T);
  constexpr int NF = FragW<T>::NF;
  if constexpr (KR) {
    static_assert((OT::LK * sizeof(T)) % 16 == 0, "KR rows must be 16-byte multiples");
    if constexpr (NF == NQ) {
      const auto* v = reinterpret_cast<const typename Vec16<T>::type*>(lds);
      const auto x = v[OT::kr_off(row, kc + 4 * g + j0) / NQ];
      if constexpr (NQ == 4) { f[0] = x.x; f[1] = x.y; f[2] = x.z; f[3] = x.w; }
      else { f[0] = x.x; f[1] = x.y; }
    } else {
      // BF16: four values, 8 bytes, inside one 16-byte group (k = kc + 4g is a
      // multiple of 4, so kr_off is too)
      static_assert(NF * sizeof(T) == 8, "8-byte fragment");
      const uint2 x = reinterpret_cast<const uint2*>(lds)[OT::kr_off(row, kc + 4 * g + j0) / 4];
      f[0] = __builtin_bit_cast(T, static_cast<unsigned short>(x.x & 0xffffu));
      f[1] = __builtin_bit_cast(T, static_cast<unsigned short>(x.x >> 16));
      f[2] = __builtin_bit_cast(T, static_cast<unsigned short>(x.y & 0xffffu));
      f[3] = __builtin_bit_cast(T, static_cast<unsigned short>(x.y >> 16));
    }
  } else {
#pragma unroll
    for (int j = 0; j < NF; ++j) f[j] = lds[(kc + 4 * g + j0 + j) * OT::LR + row];
  }
}

// All MFMAs of one K step on an LDS buffer, operand orientations fixed.
template <class T, int TM, int TN, int TK, bool KRA, bool KRB, class Acc, int FM, int FN>
__device__ inline void mfma_step(const T* sa, const T* sb, Acc (&acc)[FM][FN], int wm, int wn, int lane) {
  constexpr int NF = FragW<T>::NF;
  constexpr int NGRP = (TK / 16) * (4 / NF);  // fragment groups per K step
  const int g = lane >> 4;
  if constexpr (sizeof(T) == 2) {
    // BF16: gfx950's 32-deep MFMA takes eight k-slot values per lane, the
    // four of lane group g in the 16-deep chunks kc and kc + 16 (A and B
    // alike, so the products sum over all 32 k of the pair of chunks)
    static_assert(TK % 32 == 0, "BF16 K steps are multiples of 32");
    // B fragments of the chunk pair held packed, A fragments one at a time
    auto frag8 = [&](const T* lds, int row, int kc, bool is_a) {
      T lo[NF], hi[NF];
      if (is_a) {
        tile_frag<T, TM, TK, KRA>(lds, row, kc, g, 0, lo);
        tile_frag<T, TM, TK, KRA>(lds, row, kc + 16, g, 0, hi);
      } else {
        tile_frag<T, TN, TK, KRB>(lds, row, kc, g, 0, lo);
        tile_frag<T, TN, TK, KRB>(lds, row, kc + 16, g, 0, hi);
      }
      bf8_t v;
#pragma unroll
      for (int q = 0; q < NF; ++q) {
        v[q] = lo[q];
        v[NF + q] = hi[q];
      }
      return v;
    };
#pragma unroll 1
    for (int kc = 0; kc < TK; kc += 32) {
      bf8_t pb[FN];
#pragma unroll
      for (int j = 0; j < FN; ++j) pb[j] = frag8(sb, wn * (TN / 2) + j * 16 + (lane & 15), kc, false);
#pragma unroll
      for (int i = 0; i < FM; ++i) {
        const bf8_t pa = frag8(sa, wm * (TM / 2) + i * 16 + (lane & 15), kc, true);
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(pa, pb[j], acc[i][j], 0, 0, 0);
      }
    }
    return;
  }
  // one fragment group live at a time keeps the accumulators, the staged
  // next step and the fragments inside the 256-register budget
#pragma unroll 1
  for (int gi = 0; gi < NGRP; ++gi) {
    const int kc = (gi / (4 / NF)) * 16, j0 = (gi % (4 / NF)) * NF;
    T af[FM][NF], bf[FN][NF];
#pragma unroll
    for (int i = 0; i < FM; ++i) tile_frag<T, TM, TK, KRA>(sa, wm * (TM / 2) + i * 16 + (lane & 15), kc, g, j0, af[i]);
#pragma unroll
    for (int j = 0; j < FN; ++j) tile_frag<T, TN, TK, KRB>(sb, wn * (TN / 2) + j * 16 + (lane & 15), kc, g, j0, bf[j]);
#pragma unroll
    for (int q = 0; q < NF; ++q)
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = mfma_16x16x4(af[i][q], bf[j][q], acc[i][j]);
  }
}

// WPE: waves per SIMD the register allocation must allow (2: two workgroups
// per CU, 256 registers per lane; 1: one workgroup, 512 registers).
// PF: K steps that global loads run ahead of the LDS stage (1 or 2; 2 holds
// two register sets of staged operands).
// One output tile (or one split part of it) of the grouped GEMM, executed by
// the whole workgroup.
template <class T, class TOut, int TM, int TN, int TK, int PF, bool VEC>
__device__ __forceinline__ void gemm_tile(const GemmTile tile, const GemmProb* __restrict__ probs,
                                          const GemmSeg* __restrict__ segs, const ArenaTable& arenas,
                                          typename PartT<TOut>::type* __restrict__ partial,
                                          const GemmReduce* __restrict__ reduce, int* __restrict__ counters,
                                          const int tid) {
  using TA = OperandTile<T, TM, TK>;
  using TB = OperandTile<T, TN, TK>;
  constexpr int FM = TM / 32;  // 16x16 fragments per wave (2 x 2 waves)
  constexpr int FN = TN / 32;
  typedef typename Acc4<T>::type acc_t;
  typedef typename ScaleT<T>::type scale_t;
  typedef typename PartT<TOut>::type part_t;

  __shared__ __align__(16) T lds[2][TA::ELEMS + TB::ELEMS];

  const GemmProb prob = probs[tile.prob];
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave & 1;
  const int wn = wave >> 1;
  const int g = lane >> 4;

  acc_t acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

  // loader: segment cursor two steps ahead of the MFMAs
  int lseg = tile.seg_lo;
  while (lseg < tile.seg_hi && segs[lseg].K <= 0) ++lseg;
  int lk0 = 0, lK = 0;
  OperandCursor<T> ca, cb;
  scale_t lscale = 1;
  bool lkra = false, lkrb = false;
  auto enter = [&](int si) {
    const GemmSeg sg = segs[si];
    const T* A = static_cast<const T*>(arena_ptr(arenas, sg.a_off, sizeof(T)));
    const T* B = static_cast<const T*>(arena_ptr(arenas, sg.b_off, sizeof(T)));
    lkra = (sg.flags & kTransA) != 0;
    lkrb = (sg.flags & kTransB) == 0;
    if constexpr (VEC) {
      vcursor_init<T, TM, TK>(ca, A, sg.lda, lkra, tile.m0, prob.M, sg.K, tid);
      vcursor_init<T, TN, TK>(cb, B, sg.ldb, lkrb, tile.n0, prob.N, sg.K, tid);
    } else {
      cursor_init<T, TM, TK>(ca, A, sg.lda, lkra, tile.m0, prob.M, sg.K, tid);
      cursor_init<T, TN, TK>(cb, B, sg.ldb, lkrb, tile.n0, prob.N, sg.K, tid);
    }
    lscale = static_cast<scale_t>(sg.scale);
    lK = sg.K;
  };
  // advance the loader past the step it just loaded; false when exhausted
  auto next = [&]() {
    lk0 += TK;
    if (lk0 < lK) {
      cursor_advance(ca, TK);
      cursor_advance(cb, TK);
      return true;
    }
    lk0 = 0;
    ++lseg;
    while (lseg < tile.seg_hi && segs[lseg].K <= 0) ++lseg;
    if (lseg >= tile.seg_hi) return false;
    enter(lseg);
    return true;
  };

  unsigned skra = 0, skrb = 0;  // bit b: orientation of LDS buffer b
  auto compute = [&](int buf) {
    const T* sa = &lds[buf][0];
    const T* sb = &lds[buf][TA::ELEMS];
    switch (((skra >> buf) & 1u) | (((skrb >> buf) & 1u) << 1)) {
      case 0: mfma_step<T, TM, TN, TK, false, false>(sa, sb, acc, wm, wn, lane); break;
      case 1: mfma_step<T, TM, TN, TK, true, false>(sa, sb, acc, wm, wn, lane); break;
      case 2: mfma_step<T, TM, TN, TK, false, true>(sa, sb, acc, wm, wn, lane); break;
      default: mfma_step<T, TM, TN, TK, true, true>(sa, sb, acc, wm, wn, lane); break;
    }
  };
  // one staged K step in registers, with the orientation and scale it was
  // loaded with (elements, or 16-byte vectors when VEC)
  typedef typename VecS<T>::type vec_t;
  struct StagedS {
    T a[TA::E], b[TB::E];
    scale_t scale;
    bool kra, krb, valid;
  };
  struct StagedV {
    vec_t a[VecTile<T, TM, TK>::NV], b[VecTile<T, TN, TK>::NV];
    scale_t scale;
    bool kra, krb, valid;
  };
  typedef typename std::conditional<VEC, StagedV, StagedS>::type Staged;
  bool more = lseg < tile.seg_hi;
  if (more) enter(lseg);
  auto fetch = [&](Staged& r) {
    r.valid = more;
    if (!more) return;
    if constexpr (VEC) {
      vcursor_load<T, VecTile<T, TM, TK>::NV>(ca, r.a);
      vcursor_load<T, VecTile<T, TN, TK>::NV>(cb, r.b);
    } else {
      cursor_load<T, TA::E>(ca, r.a);
      cursor_load<T, TB::E>(cb, r.b);
    }
    r.scale = lscale;
    r.kra = lkra;
    r.krb = lkrb;
    more = next();
  };
  auto put = [&](const Staged& r, int buf) {
    if constexpr (VEC) {
      vtile_store<T, TM, TK>(&lds[buf][0], r.kra, r.a, r.scale, tid);
      vtile_store<T, TN, TK>(&lds[buf][TA::ELEMS], r.krb, r.b, scale_t(1), tid);
    } else {
      tile_store<T, TM, TK>(&lds[buf][0], r.kra, r.a, r.scale, tid);
      tile_store<T, TN, TK>(&lds[buf][TA::ELEMS], r.krb, r.b, scale_t(1), tid);
    }
    skra = (skra & ~(1u << buf)) | (r.kra ? 1u << buf : 0u);
    skrb = (skrb & ~(1u << buf)) | (r.krb ? 1u << buf : 0u);
  };
  if constexpr (PF == 1) {
    // global loads one K step ahead of the LDS stage
    Staged r;
    fetch(r);
    if (r.valid) {
      put(r, 0);
      fetch(r);
      __syncthreads();
      for (int s = 0;; ++s) {
        const int cur = s & 1, nb = cur ^ 1;
        const bool has_next = r.valid;
        if (has_next) {
          put(r, nb);
          fetch(r);
        }
        compute(cur);
        if (!has_next) break;
        __syncthreads();
      }
    }
  } else {
    // global loads two K steps ahead of the LDS stage, alternating between
    // two register sets (the loop is unrolled by two so that each set is a
    // fixed set of registers)
    Staged r0, r1;
    fetch(r0);
    if (r0.valid) {
      put(r0, 0);
      fetch(r0);
      fetch(r1);
      __syncthreads();
      for (;;) {
        bool nxt = r0.valid;
        if (nxt) {
          put(r0, 1);
          fetch(r0);
        }
        compute(0);
        if (!nxt) break;
        __syncthreads();
        nxt = r1.valid;
        if (nxt) {
          put(r1, 0);
          fetch(r1);
        }
        compute(1);
        if (!nxt) break;
        __syncthreads();
      }
    }
  }

  if (tile.part >= 0) {
    // split reduction: dense partial tile; the last workgroup of the task to
    // sign in folds every slot into C in slot order
    part_t* P = partial + int64_t(tile.part) * (TM * TN);
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        const int n = wn * (TN / 2) + j * 16 + (lane & 15);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int m = wm * (TM / 2) + i * 16 + acc_row<T>(g, r);
          P[m + n * TM] = static_cast<part_t>(acc[i][j][r]);
        }
      }
    __shared__ int s_last;
    __threadfence();  // partial visible device-wide before signing in
    __syncthreads();
    const GemmReduce rt = reduce[tile.red];
    if (tid == 0) {
      const int prev = __hip_atomic_fetch_add(&counters[tile.red], 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
      s_last = (prev == rt.nparts - 1) ? 1 : 0;
    }
    __syncthreads();
    if (!s_last) return;
    __threadfence();  // every other slot's partial is visible from here on
    TOut* C = static_cast<TOut*>(const_cast<void*>(arena_ptr(arenas, prob.c_off, sizeof(TOut))));
    const int rows = min(TM, prob.M - rt.m0);
    const int cols = min(TN, prob.N - rt.n0);
    for (int idx = tid; idx < TM * TN; idx += 256) {
      const int m = idx % TM, n = idx / TM;
      if (m >= rows || n >= cols) continue;
      const part_t* Q = partial + int64_t(rt.part_begin) * (TM * TN) + idx;
      part_t v = Q[0];
      for (int q = 1; q < rt.nparts; ++q) v += Q[int64_t(q) * (TM * TN)];
      TOut* p = C + (rt.m0 + m) + int64_t(rt.n0 + n) * prob.ldc;
      if (prob.accumulate) *p = static_cast<TOut>(static_cast<part_t>(*p) + v);
      else *p = static_cast<TOut>(v);
    }
    if (tid == 0) __hip_atomic_store(&counters[tile.red], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return;
  }
  TOut* C = static_cast<TOut*>(const_cast<void*>(arena_ptr(arenas, prob.c_off, sizeof(TOut))));
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int gn = tile.n0 + wn * (TN / 2) + j * 16 + (lane & 15);
      if (gn >= prob.N) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gm = tile.m0 + wm * (TM / 2) + i * 16 + acc_row<T>(g, r);
        if (gm >= prob.M) continue;
        TOut* p = C + gm + int64_t(gn) * prob.ldc;
        if constexpr (sizeof(TOut) == 2) {
          // BF16 output: one rounding of the FP32 sum
          const float v = acc[i][j][r];
          *p = static_cast<TOut>(prob.accumulate ? static_cast<float>(*p) + v : v);
        } else {
          const TOut v = static_cast<TOut>(acc[i][j][r]);
          if (prob.accumulate) *p += v;
          else *p = v;
        }
      }
    }
}

// The grouped GEMM launch.  queue == nullptr: one workgroup per tile
// (tiles[tile_base + blockIdx.x]), placed by the hardware dispatcher.
// Persistent (PERSIST): a grid of about the resident workgroup count; each
// workgroup takes tile tickets from queue[0] until they exceed ntiles, and
// the last workgroup to leave (queue[1] counts the leavers) resets both
// counters, so the next launch (or graph replay) starts from zero without a
// host-side fill.  Every workgroup takes exactly one failing ticket before it
// signs out, so no ticket is taken after the reset.
template <class T, class TOut, int TM, int TN, int TK, int WPE, int PF, bool VEC = false, bool PERSIST = false>
__global__ __launch_bounds__(256, WPE) void grouped_gemm_kernel(const GemmTile* __restrict__ tiles,
                                                              const GemmProb* __restrict__ probs,
                                                              const GemmSeg* __restrict__ segs,
                                                              ArenaTable arenas,
                                                              typename PartT<TOut>::type* __restrict__ partial,
                                                              const GemmReduce* __restrict__ reduce,
                                                              int* __restrict__ counters, int tile_base,
                                                              int* __restrict__ queue, int ntiles) {
  if constexpr (!PERSIST) {
    gemm_tile<T, TOut, TM, TN, TK, PF, VEC>(tiles[tile_base + blockIdx.x], probs, segs, arenas, partial, reduce,
                                            counters, static_cast<int>(threadIdx.x));
  } else {
    __shared__ int s_ticket;
    for (;;) {
      __syncthreads();  // the previous tile is done with LDS and s_ticket
      if (threadIdx.x == 0)
        s_ticket = __hip_atomic_fetch_add(&queue[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __syncthreads();
      // uniform across the workgroup: the tile and everything derived from it
      // stay in scalar registers, as with blockIdx.x
      const int t = __builtin_amdgcn_readfirstlane(s_ticket);
      if (t >= ntiles) break;
      // an opaque copy of the thread index per tile: nothing derived from it
      // is hoisted out of the loop and kept live across tiles (the body needs
      // the registers of the one-tile kernel)
      int tid = static_cast<int>(threadIdx.x);
      asm volatile("" : "+v"(tid));
      gemm_tile<T, TOut, TM, TN, TK, PF, VEC>(tiles[tile_base + t], probs, segs, arenas, partial, reduce, counters,
                                              tid);
    }
    if (threadIdx.x == 0) {
      const int left = __hip_atomic_fetch_add(&queue[1], 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
      if (left == static_cast<int>(gridDim.x) - 1) {
        __hip_atomic_store(&queue[0], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&queue[1], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
  }
}

// Complex grouped GEMM on interleaved (re, im) data, FP64 or FP32 operands
// (FP32 accumulation, FP32 or FP64 output).  Four real MFMA products per
// step; conjugation flags and the segment scale are applied when a K step is
// written to LDS.  One K step is held in registers ahead of the MFMAs: the
// loads of step s+1 (the next K block, or the first of the next segment) are
// issued before the products of step s, so their latency overlaps the
// MFMAs.  Each complex element is one 8- or 16-byte load.
template <class T> struct Cplx;
template <> struct Cplx<float> { typedef float2 type; };
template <> struct Cplx<double> { typedef double2 type; };

template <class T, class TOut, int TM, int TN, int TK, int WM, int WN>
__global__ __launch_bounds__(64 * WM * WN) void grouped_zgemm_kernel(const GemmTile* __restrict__ tiles,
                                                                     const GemmProb* __restrict__ probs,
                                                                     const GemmSeg* __restrict__ segs,
                                                                     ArenaTable arenas, int tile_base) {
  constexpr int NT = 64 * WM * WN;  // threads: WM x WN waves
  constexpr int FM = TM / WM / 16;
  constexpr int FN = TN / WN / 16;
  constexpr int PAD = 2;
  constexpr int A_ELEMS = TM * TK / NT;
  constexpr int B_ELEMS = TN * TK / NT;
  static_assert(A_ELEMS >= 1 && B_ELEMS >= 1 && NT % TM == 0 && NT % TN == 0 && NT % TK == 0, "tile shape");
  typedef typename Acc4<T>::type acc_t;
  typedef typename Cplx<T>::type V;
  __shared__ T sAr[TK][TM + PAD], sAi[TK][TM + PAD];
  __shared__ T sBr[TK][TN + PAD], sBi[TK][TN + PAD];

  const GemmTile tile = tiles[tile_base + blockIdx.x];
  const GemmProb prob = probs[tile.prob];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave % WM;
  const int wn = wave / WM;

  acc_t accr[FM][FN], acci[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) { accr[i][j] = acc_t{0, 0, 0, 0}; acci[i][j] = acc_t{0, 0, 0, 0}; }

  // loader: the segment and K block whose elements are in ra / rb
  const V* A = nullptr;
  const V* B = nullptr;
  int64_t lda = 0, ldb = 0;
  int K = 0, k0 = 0;
  bool ta = false, tb = false;
  T ca = 1, cb = 1, sr = 1, sim = 0;
  int ls = tile.seg_lo;
  auto enter = [&](int si) {
    const GemmSeg sg = segs[si];
    A = static_cast<const V*>(arena_ptr(arenas, sg.a_off, 2 * sizeof(T)));
    B = static_cast<const V*>(arena_ptr(arenas, sg.b_off, 2 * sizeof(T)));
    lda = sg.lda;
    ldb = sg.ldb;
    K = sg.K;
    ta = (sg.flags & kTransA) != 0;
    tb = (sg.flags & kTransB) != 0;
    ca = (sg.flags & kConjA) ? T(-1) : T(1);
    cb = (sg.flags & kConjB) ? T(-1) : T(1);
    sr = static_cast<T>(sg.scale);
    sim = static_cast<T>(sg.scale_im);
    k0 = 0;
  };
  // next segment with K > 0 at or after si; false past the tile's segments
  auto seek = [&](int si) {
    while (si < tile.seg_hi && segs[si].K <= 0) ++si;
    ls = si;
    if (si >= tile.seg_hi) return false;
    enter(si);
    return true;
  };
  V ra[A_ELEMS], rb[B_ELEMS];
  auto fetch = [&]() {
#pragma unroll
    for (int e = 0; e < A_ELEMS; ++e) {
      const int idx = tid + e * NT;
      const int m = ta ? idx / TK : idx % TM, k = ta ? idx % TK : idx / TM;
      const int gm = tile.m0 + m, gk = k0 + k;
      const bool ok = gm < prob.M && gk < K;
      const int64_t o = ta ? (gk + int64_t(gm) * lda) : (gm + int64_t(gk) * lda);
      ra[e] = ok ? A[o] : V{0, 0};
    }
#pragma unroll
    for (int e = 0; e < B_ELEMS; ++e) {
      const int idx = tid + e * NT;
      const int n = tb ? idx % TN : idx / TK, k = tb ? idx / TN : idx % TK;
      const int gn = tile.n0 + n, gk = k0 + k;
      const bool ok = gn < prob.N && gk < K;
      const int64_t o = tb ? (gn + int64_t(gk) * ldb) : (gk + int64_t(gn) * ldb);
      rb[e] = ok ? B[o] : V{0, 0};
    }
  };

  bool have = seek(tile.seg_lo);
  if (have) fetch();
  while (have) {
    __syncthreads();  // the previous step's products are done with LDS
#pragma unroll
    for (int e = 0; e < A_ELEMS; ++e) {
      const int idx = tid + e * NT;
      const int m = ta ? idx / TK : idx % TM, k = ta ? idx % TK : idx / TM;
      const T xr = ra[e].x, xi = ca * ra[e].y;
      sAr[k][m] = xr * sr - xi * sim;
      sAi[k][m] = xr * sim + xi * sr;
    }
#pragma unroll
    for (int e = 0; e < B_ELEMS; ++e) {
      const int idx = tid + e * NT;
      const int n = tb ? idx % TN : idx / TK, k = tb ? idx / TN : idx % TK;
      sBr[k][n] = rb[e].x;
      sBi[k][n] = cb * rb[e].y;
    }
    __syncthreads();
    // issue the next step's loads before this step's products
    k0 += TK;
    have = k0 < K || seek(ls + 1);
    if (have) fetch();
#pragma unroll
    for (int kk = 0; kk < TK; kk += 4) {
      const int kr = kk + (lane >> 4);
      T ar[FM], ai[FM], nai[FM], br[FN], bi[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) {
        ar[i] = sAr[kr][(wm * FM + i) * 16 + (lane & 15)];
        ai[i] = sAi[kr][(wm * FM + i) * 16 + (lane & 15)];
        nai[i] = -ai[i];
      }
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        br[j] = sBr[kr][(wn * FN + j) * 16 + (lane & 15)];
        bi[j] = sBi[kr][(wn * FN + j) * 16 + (lane & 15)];
      }
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) {
          accr[i][j] = mfma_16x16x4(ar[i], br[j], accr[i][j]);
          accr[i][j] = mfma_16x16x4(nai[i], bi[j], accr[i][j]);
          acci[i][j] = mfma_16x16x4(ar[i], bi[j], acci[i][j]);
          acci[i][j] = mfma_16x16x4(ai[i], br[j], acci[i][j]);
        }
    }
  }

  typedef typename Cplx<TOut>::type VO;
  VO* C = static_cast<VO*>(const_cast<void*>(arena_ptr(arenas, prob.c_off, 2 * sizeof(TOut))));
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int gn = tile.n0 + (wn * FN + j) * 16 + (lane & 15);
      if (gn >= prob.N) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gm = tile.m0 + (wm * FM + i) * 16 + acc_row<T>(lane >> 4, r);
        if (gm >= prob.M) continue;
        VO* p = C + (gm + int64_t(gn) * prob.ldc);
        VO v{static_cast<TOut>(accr[i][j][r]), static_cast<TOut>(acci[i][j][r])};
        if (prob.accumulate) {
          const VO old = *p;
          v.x += old.x;
          v.y += old.y;
        }
        *p = v;
      }
    }
}

// ---------------------------------------------------------------------------
// Launch helpers.  The grid is split in chunks so that gridDim.x stays within
// limits for very large tile tables.
template <class T, class TOut, int TM, int TN, int TK, int WPE = 2, int PF = 2>
static hipError_t launch_real(const GemmLaunch& L, hipStream_t st) {
  typename PartT<TOut>::type* part = static_cast<typename PartT<TOut>::type*>(L.partial);
  // Persistent variants exist where the tile loop compiles without scratch:
  // FP64 128 x 64 tiles stage one K step ahead (two steps need the registers
  // the loop keeps), and the FP32 128 x 128 tiles and every BF16 tile only
  // with vector staging; other tables launch one workgroup per tile.
  constexpr int PFP = (sizeof(T) == 8 && TM == 128 && TN == 64) ? 1 : PF;
  // BF16 128 x 128 tiles fit two workgroups per CU with vector staging only;
  // element staging keeps one
  constexpr int WPEE = (sizeof(T) == 2 && TM == 128 && TN == 128) ? 1 : WPE;
  constexpr bool kElemPersist = !(sizeof(T) == 4 && TM == 128 && TN == 128) && sizeof(T) != 2;
  if (L.queue && (L.vec || kElemPersist)) {
    // persistent: about the resident workgroup count (WPE per CU)
    const int grid = L.ntiles < kPersistCUs * WPE ? L.ntiles : kPersistCUs * WPE;
    if (L.vec) {
      hipLaunchKernelGGL((grouped_gemm_kernel<T, TOut, TM, TN, TK, WPE, PFP, true, true>), dim3(grid), dim3(256), 0,
                         st, L.tiles, L.probs, L.segs, L.arenas, part, L.reduce, L.counters, 0, L.queue, L.ntiles);
      return hipGetLastError();
    }
    if constexpr (kElemPersist) {
      hipLaunchKernelGGL((grouped_gemm_kernel<T, TOut, TM, TN, TK, WPE, PFP, false, true>), dim3(grid), dim3(256), 0,
                         st, L.tiles, L.probs, L.segs, L.arenas, part, L.reduce, L.counters, 0, L.queue, L.ntiles);
      return hipGetLastError();
    }
  }
  const int chunk = 1 << 30;
  for (int base = 0; base < L.ntiles; base += chunk) {
    const int n = L.ntiles - base < chunk ? L.ntiles - base : chunk;
    if (L.vec) {
      hipLaunchKernelGGL((grouped_gemm_kernel<T, TOut, TM, TN, TK, WPE, PF, true>), dim3(n), dim3(256), 0, st,
                         L.tiles, L.probs, L.segs, L.arenas, part, L.reduce, L.counters, base, nullptr, n);
      continue;
    }
    hipLaunchKernelGGL((grouped_gemm_kernel<T, TOut, TM, TN, TK, WPEE, PF>), dim3(n), dim3(256), 0, st,
                       L.tiles, L.probs, L.segs, L.arenas, part, L.reduce, L.counters, base, nullptr, n);
  }
  return hipGetLastError();
}

// W x W waves per workgroup (a 32 x 32 tile on one wave, W = 1, measured
// slower: fewer resident waves for the same LDS)
template <int TM, int TN, int W>
static hipError_t launch_complex(GemmPrec prec, const GemmLaunch& L, hipStream_t st) {
  if (prec == GemmPrec::Z64)
    hipLaunchKernelGGL((grouped_zgemm_kernel<double, double, TM, TN, 16, W, W>), dim3(L.ntiles), dim3(64 * W * W), 0,
                       st, L.tiles, L.probs, L.segs, L.arenas, 0);
  else if (prec == GemmPrec::Z32)
    hipLaunchKernelGGL((grouped_zgemm_kernel<float, float, TM, TN, 16, W, W>), dim3(L.ntiles), dim3(64 * W * W), 0,
                       st, L.tiles, L.probs, L.segs, L.arenas, 0);
  else
    hipLaunchKernelGGL((grouped_zgemm_kernel<float, double, TM, TN, 16, W, W>), dim3(L.ntiles), dim3(64 * W * W), 0,
                       st, L.tiles, L.probs, L.segs, L.arenas, 0);
  return hipGetLastError();
}

hipError_t launch_grouped_gemm(GemmPrec prec, int tile_cfg, const GemmLaunch& L, hipStream_t st) {
  if (L.ntiles <= 0) return hipSuccess;
  if (L.nreduce > 0 && (L.partial == nullptr || L.reduce == nullptr || L.counters == nullptr))
    return hipErrorInvalidValue;
  switch (prec) {
    case GemmPrec::F64:
      if (tile_cfg == 3) return launch_real<double, double, 32, 32, 16>(L, st);
      if (tile_cfg == 1) return launch_real<double, double, 128, 128, 16, 1>(L, st);
      if (tile_cfg == 2) return launch_real<double, double, 128, 64, 16>(L, st);
      return launch_real<double, double, 64, 64, 16>(L, st);
    case GemmPrec::F32:
      if (tile_cfg == 3) return launch_real<float, float, 32, 32, 32>(L, st);
      if (tile_cfg == 1) return launch_real<float, float, 128, 128, 32, 2, 1>(L, st);
      if (tile_cfg == 2) return launch_real<float, float, 128, 64, 32>(L, st);
      return launch_real<float, float, 64, 64, 32>(L, st);
    case GemmPrec::F32_OUT64:
      if (tile_cfg == 3) return launch_real<float, double, 32, 32, 32>(L, st);
      if (tile_cfg == 1) return launch_real<float, double, 128, 128, 32, 2, 1>(L, st);
      if (tile_cfg == 2) return launch_real<float, double, 128, 64, 32>(L, st);
      return launch_real<float, double, 64, 64, 32>(L, st);
    case GemmPrec::BF16:
      if (tile_cfg == 3) return launch_real<bf16_t, bf16_t, 32, 32, 64, 2, 1>(L, st);
      if (tile_cfg == 1) return launch_real<bf16_t, bf16_t, 128, 128, 64, 2, 1>(L, st);
      if (tile_cfg == 2) return launch_real<bf16_t, bf16_t, 128, 64, 64, 2, 1>(L, st);
      return launch_real<bf16_t, bf16_t, 64, 64, 64, 2, 1>(L, st);
    case GemmPrec::BF16_OUT64:
      if (tile_cfg == 3) return launch_real<bf16_t, double, 32, 32, 64, 2, 1>(L, st);
      if (tile_cfg == 1) return launch_real<bf16_t, double, 128, 128, 64, 2, 1>(L, st);
      if (tile_cfg == 2) return launch_real<bf16_t, double, 128, 64, 64, 2, 1>(L, st);
      return launch_real<bf16_t, double, 64, 64, 64, 2, 1>(L, st);
    case GemmPrec::Z64:
    case GemmPrec::Z32:
    case GemmPrec::Z32_OUT64: {
      // no split reductions or large tiles for complex tables
      if (L.nreduce > 0 || !complex_tile_cfg(tile_cfg)) return hipErrorInvalidValue;
      if (tile_cfg == 3) return launch_complex<32, 32, 2>(prec, L, st);
      return launch_complex<64, 64, 2>(prec, L, st);
    }
  }
  return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
// Vector kernels (FP64).  Complex vectors are handled as interleaved doubles;
// `cplx` selects conjugate-linear dot products.

constexpr int kRedBlock = 256;
constexpr int kRedGrid = 1024;

// partial[g*ncol + c] = sum over this block's slice of conj(X[:,c]) . y
// dncol (optional, device): the column count is min(*dncol, ncol), read by
// every thread, so one launch serves steps of different width.
__global__ __launch_bounds__(kRedBlock) void gemv_t_partial(const double* __restrict__ X, int64_t ldx,
                                                            int ncol, const int* __restrict__ dncol,
                                                            const double* __restrict__ y,
                                                            int64_t n, int cplx,
                                                            double* __restrict__ partial) {
  __shared__ double red[kRedBlock / 64][2];
  if (dncol) ncol = min(*dncol, ncol);
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int c = 0; c < ncol; ++c) {
    const double* x = X + c * ldx * (cplx ? 2 : 1);
    double sr = 0, si = 0;
    if (!cplx) {
      for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += stride) sr += x[i] * y[i];
    } else {
      for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += stride) {
        const double xr = x[2 * i], xi = x[2 * i + 1], yr = y[2 * i], yi = y[2 * i + 1];
        sr += xr * yr + xi * yi;
        si += xr * yi - xi * yr;
      }
    }
    for (int off = 32; off > 0; off >>= 1) {
      sr += __shfl_down(sr, off);
      si += __shfl_down(si, off);
    }
    if ((threadIdx.x & 63) == 0) { red[threadIdx.x >> 6][0] = sr; red[threadIdx.x >> 6][1] = si; }
    __syncthreads();
    if (threadIdx.x == 0) {
      double tr = 0, ti = 0;
      for (int w = 0; w < kRedBlock / 64; ++w) { tr += red[w][0]; ti += red[w][1]; }
      partial[(int64_t(blockIdx.x) * ncol + c) * 2] = tr;
      partial[(int64_t(blockIdx.x) * ncol + c) * 2 + 1] = ti;
    }
    __syncthreads();
  }
}

// out[c] (complex pair) = sum_g partial[g, c]; out is scaled by `sign`.
__global__ void gemv_t_finish(const double* __restrict__ partial, int ngrid, int ncol,
                              const int* __restrict__ dncol, double sign, double* __restrict__ out) {
  const int c = blockIdx.x;
  if (dncol) ncol = min(*dncol, ncol);
  if (c >= ncol) return;
  __shared__ double red[64][2];
  double sr = 0, si = 0;
  for (int g = threadIdx.x; g < ngrid; g += blockDim.x) {
    sr += partial[(int64_t(g) * ncol + c) * 2];
    si += partial[(int64_t(g) * ncol + c) * 2 + 1];
  }
  red[threadIdx.x][0] = sr;
  red[threadIdx.x][1] = si;
  __syncthreads();
  if (threadIdx.x == 0) {
    double tr = 0, ti = 0;
    for (int w = 0; w < 64; ++w) { tr += red[w][0]; ti += red[w][1]; }
    out[2 * c] = sign * tr;
    out[2 * c + 1] = sign * ti;
  }
}

// y += sum_c coef[c] * X[:,c]   (coef complex pairs, device resident)
__global__ void gemv_n_update(const double* __restrict__ X, int64_t ldx, int ncol, const int* __restrict__ dncol,
                              const double* __restrict__ coef, double* __restrict__ y, int64_t n, int cplx) {
  if (dncol) ncol = min(*dncol, ncol);
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += stride) {
    if (!cplx) {
      double acc = y[i];
      for (int c = 0; c < ncol; ++c) acc += coef[2 * c] * X[i + c * ldx];
      y[i] = acc;
    } else {
      double ar = y[2 * i], ai = y[2 * i + 1];
      for (int c = 0; c < ncol; ++c) {
        const double xr = X[2 * (i + c * ldx)], xi = X[2 * (i + c * ldx) + 1];
        const double cr = coef[2 * c], ci = coef[2 * c + 1];
        ar += cr * xr - ci * xi;
        ai += cr * xi + ci * xr;
      }
      y[2 * i] = ar;
      y[2 * i + 1] = ai;
    }
  }
}

// ---- Lanczos steps without host synchronisation ----------------------------
// One step produces column m = s->m of the Krylov basis V (columns of ldv
// doubles, nd doubles used) from w = H x, x = column m-1:
//   alpha[m-1] = <x, w>;  w -= alpha x;  w -= beta[m-1] V[:, m-2]   (update)
//   w -= V[:, 0..m) (V^H w), twice                                  (gemv kernels)
//   beta[m] = |w|;  V[:, m] = x = w / beta[m];  m += 1              (finish, advance)
// The same expressions and rounding order as the host-driven step.  A step
// whose beta falls below tol sets s->stop and leaves m on that step; every
// later step is then a no-op.
__global__ void lanczos_update_kernel(LanczosDev* __restrict__ s, double* __restrict__ w,
                                      const double* __restrict__ x, const double* __restrict__ V, int64_t ldv,
                                      int64_t nd) {
  if (s->stop) return;
  const int m = s->m;
  const double a = s->dot[0];
  const double b = m >= 2 ? s->beta[m - 1] : 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) s->alpha[m - 1] = a;
  const double* v2 = m >= 2 ? V + int64_t(m - 2) * ldv : nullptr;
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < nd; i += stride) {
    double t = fma(-a, x[i], w[i]);
    if (v2) t = fma(-b, v2[i], t);
    w[i] = t;
  }
}

__global__ void lanczos_finish_kernel(LanczosDev* __restrict__ s, const double* __restrict__ w,
                                      double* __restrict__ x, double* __restrict__ V, int64_t ldv, int64_t nd,
                                      double tol) {
  if (s->stop) return;
  const int m = s->m;
  const double b = sqrt(fabs(s->nrm[0]));
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    s->beta[m] = b;
    if (b < tol) s->stop = 1;
  }
  if (b < tol) return;
  const double inv = 1.0 / b;
  double* col = V + int64_t(m) * ldv;
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < nd; i += stride) {
    const double t = inv * w[i];
    col[i] = t;
    x[i] = t;
  }
}

__global__ void lanczos_advance_kernel(LanczosDev* __restrict__ s) {
  if (threadIdx.x == 0 && !s->stop) s->m += 1;
}

// y = a*y + b*x with host scalars (real), one rounding of b*x + a*y (a*y is
// exact for a = 1, so the Lanczos update matches the device-step kernels)
__global__ void axpby_kernel(int64_t n, double a, double b, const double* __restrict__ x, double* __restrict__ y) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += stride)
    y[i] = x ? fma(b, x[i], a * y[i]) : a * y[i];
}

// y *= complex scalar (or y = x * s), interleaved
__global__ void zscal_kernel(int64_t n, double sr, double si, const double* __restrict__ x,
                             double* __restrict__ y) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += stride) {
    const double xr = x[2 * i], xi = x[2 * i + 1];
    y[2 * i] = sr * xr - si * xi;
    y[2 * i + 1] = sr * xi + si * xr;
  }
}

__global__ void f64_to_f32_kernel(int64_t n, const double* __restrict__ x, float* __restrict__ y) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += stride) y[i] = static_cast<float>(x[i]);
}

__global__ void f64_to_bf16_kernel(int64_t n, const double* __restrict__ x, bf16_t* __restrict__ y) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += stride)
    y[i] = static_cast<bf16_t>(x[i]);
}

__global__ void f32_to_f64_kernel(int64_t n, const float* __restrict__ x, double* __restrict__ y) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += stride) y[i] = static_cast<double>(x[i]);
}

__global__ void fill_kernel(int64_t n, double v, double* __restrict__ y) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += stride) y[i] = v;
}

// Strided 2-D copy with scale: dst[i + j*ldd] (+)= s * cs[j] * src[i*s_row + j*s_col]
// (element = `w` doubles; cs optional per-column scale)
__global__ void copy2d_kernel(int rows, int cols, const double* __restrict__ src, int64_t s_row, int64_t s_col,
                              double* __restrict__ dst, int64_t ldd, int w, double scale,
                              const double* __restrict__ cs, int accumulate) {
  const int64_t total = int64_t(rows) * cols;
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < total; t += stride) {
    const int i = static_cast<int>(t % rows);
    const int j = static_cast<int>(t / rows);
    const double f = cs ? scale * cs[j] : scale;
    const double* s = src + (i * s_row + j * s_col) * w;
    double* d = dst + (i + j * ldd) * w;
    for (int c = 0; c < w; ++c) d[c] = accumulate ? d[c] + f * s[c] : f * s[c];
  }
}

// out[v] = || A[v*s_vec + e*s_elem], e < len ||_2 for v < count: one workgroup
// per vector (column norms: s_elem = 1, s_vec = ld; row norms: s_elem = ld,
// s_vec = 1).
__global__ void __launch_bounds__(256) strided_norms_kernel(int count, int len, const double* __restrict__ a,
                                                            int64_t s_elem, int64_t s_vec, double* __restrict__ out) {
  __shared__ double red[256];
  const int v = blockIdx.x;
  if (v >= count) return;
  const double* p = a + v * s_vec;
  double acc = 0.0;
  for (int e = threadIdx.x; e < len; e += blockDim.x) {
    const double x = p[e * s_elem];
    acc += x * x;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[v] = sqrt(red[0]);
}

// A[v*s_vec + e*s_elem] *= f[v] for v < count, e < len.
__global__ void strided_scale_kernel(int count, int len, double* __restrict__ a, int64_t s_elem, int64_t s_vec,
                                     const double* __restrict__ f) {
  const int64_t total = int64_t(count) * len;
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < total; t += stride) {
    const int64_t v = t / len, e = t % len;
    a[v * s_vec + e * s_elem] *= f[v];
  }
}

// p[t * stride] = v  for t < n  (diagonal / strided fill)
__global__ void strided_fill_kernel(int64_t n, double v, double* __restrict__ p, int64_t stride) {
  const int64_t g = int64_t(gridDim.x) * blockDim.x;
  for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < n; t += g) p[t * stride] = v;
}

// Block scatter (enlarged-basis environment updates): each workgroup copies
// one column chunk of one operator block into its embedded position.
__global__ __launch_bounds__(256) void block_scatter_kernel(const ScatterTask* __restrict__ tasks,
                                                            const double* __restrict__ src,
                                                            double* __restrict__ dst, int w) {
  const ScatterTask t = tasks[blockIdx.x];
  if (t.kind == 1) {
    for (int i = threadIdx.x; i < t.rows; i += blockDim.x) {
      double* d = dst + (t.dst_off + int64_t(i) * (t.ldd + 1)) * w;
      d[0] = t.scale;
      if (w == 2) d[1] = 0.0;
    }
    return;
  }
  // the rectangle in column-major order, 256 elements per pass: (r, c) of
  // a thread's element advances by (256 mod rows, 256 / rows) with one
  // carry, so the loop divides only once (at the start)
  const int rows = t.rows, cols = t.cols;
  if (rows <= 0) return;
  const int dc = 256 / rows, dr = 256 - dc * rows;
  int c = static_cast<int>(threadIdx.x) / rows;
  int r = static_cast<int>(threadIdx.x) - c * rows;
  const double sc = t.scale;
  // U elements per thread in flight: their positions first, then the loads,
  // then the stores
  constexpr int U = 8;
  auto advance = [&]() {
    r += dr;
    c += dc;
    if (r >= rows) {
      r -= rows;
      ++c;
    }
  };
  if (w == 1) {
    const double* s0 = src + t.src_off;
    double* d0 = dst + t.dst_off;
    while (c < cols) {
      int64_t so[U], dof[U];
      bool ok[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        ok[u] = c < cols;
        so[u] = r + int64_t(c) * t.lds;
        dof[u] = r + int64_t(c) * t.ldd;
        advance();
      }
      double x[U];
#pragma unroll
      for (int u = 0; u < U; ++u) x[u] = ok[u] ? s0[so[u]] : 0.0;
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (ok[u]) d0[dof[u]] = sc * x[u];
    }
  } else {
    const double2* s0 = reinterpret_cast<const double2*>(src) + t.src_off;
    double2* d0 = reinterpret_cast<double2*>(dst) + t.dst_off;
    while (c < cols) {
      int64_t so[U], dof[U];
      bool ok[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        ok[u] = c < cols;
        so[u] = r + int64_t(c) * t.lds;
        dof[u] = r + int64_t(c) * t.ldd;
        advance();
      }
      double2 x[U];
#pragma unroll
      for (int u = 0; u < U; ++u) x[u] = ok[u] ? s0[so[u]] : double2{0.0, 0.0};
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (ok[u]) d0[dof[u]] = double2{sc * x[u].x, sc * x[u].y};
    }
  }
}

// Block copies into a padded layout, with conversion: for every column c of
// a task, dst(T)[t.dst_off + r + c*t.ldd] = src[t.src_off + r + c*t.lds] for
// r < t.rows and 0 for t.rows <= r < t.ldd (w doubles per element); one
// workgroup per task.  srcs (optional): per-task source bases, t.pad indexing
// them; otherwise src.
template <class TOut>
__global__ __launch_bounds__(256) void repack_kernel(const ScatterTask* __restrict__ tasks,
                                                     const double* __restrict__ src,
                                                     const double* const* __restrict__ srcs, TOut* __restrict__ dst,
                                                     int w) {
  const ScatterTask t = tasks[blockIdx.x];
  const double* base = srcs ? srcs[t.pad] : src;
  const int64_t total = t.ldd * t.cols;
  for (int64_t e = threadIdx.x; e < total; e += blockDim.x) {
    const int r = static_cast<int>(e % t.ldd);
    const int c = static_cast<int>(e / t.ldd);
    TOut* d = dst + (t.dst_off + r + int64_t(c) * t.ldd) * w;
    if (r < t.rows) {
      const double* s = base + (t.src_off + r + int64_t(c) * t.lds) * w;
      for (int k = 0; k < w; ++k) d[k] = static_cast<TOut>(s[k]);
    } else {
      for (int k = 0; k < w; ++k) d[k] = TOut(0);
    }
  }
}

hipError_t launch_repack(const ScatterTask* tasks, int ntasks, const double* src, const double* const* srcs,
                         void* dst, int w, int out_type, hipStream_t st) {
  if (ntasks <= 0) return hipSuccess;
  if (out_type == 0)
    hipLaunchKernelGGL(repack_kernel<double>, dim3(ntasks), dim3(256), 0, st, tasks, src, srcs,
                       static_cast<double*>(dst), w);
  else if (out_type == 1)
    hipLaunchKernelGGL(repack_kernel<float>, dim3(ntasks), dim3(256), 0, st, tasks, src, srcs,
                       static_cast<float*>(dst), w);
  else
    hipLaunchKernelGGL(repack_kernel<bf16_t>, dim3(ntasks), dim3(256), 0, st, tasks, src, srcs,
                       static_cast<bf16_t*>(dst), w);
  return hipGetLastError();
}

hipError_t launch_block_scatter(const ScatterTask* tasks, int ntasks, const double* src, double* dst, int w,
                                hipStream_t st) {
  if (ntasks <= 0) return hipSuccess;
  hipLaunchKernelGGL(block_scatter_kernel, dim3(ntasks), dim3(256), 0, st, tasks, src, dst, w);
  return hipGetLastError();
}

static int grid_for(int64_t n, int block = 256) {
  int64_t g = (n + block - 1) / block;
  if (g > 4096) g = 4096;
  if (g < 1) g = 1;
  return static_cast<int>(g);
}

// ---- planar complex layout ------------------------------------------------
template <class TOut>
__global__ void deinterleave_kernel(int64_t n, int64_t n4, const double* __restrict__ src, TOut* __restrict__ dst) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n4; i += stride) {
    const bool in = i < n;
    dst[i] = static_cast<TOut>(in ? src[2 * i] : 0.0);
    dst[n4 + i] = static_cast<TOut>(in ? src[2 * i + 1] : 0.0);
  }
}

__global__ void interleave_kernel(int64_t n, int64_t n4, const double* __restrict__ src, double* __restrict__ dst) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += stride) {
    dst[2 * i] = src[i];
    dst[2 * i + 1] = src[n4 + i];
  }
}

// ---- segmented vector kernels (batched solves) ----------------------------
// One workgroup per SegBlock; the columns are walked in order and each
// block's partial sums are reduced in a fixed order, so results do not
// depend on scheduling.
__global__ __launch_bounds__(kRedBlock) void seg_dot_kernel(const double* __restrict__ X, int64_t ldx, int ncol,
                                                            const double* __restrict__ y,
                                                            const SegBlock* __restrict__ blocks, int cplx,
                                                            double* __restrict__ partial) {
  __shared__ double red[kRedBlock / 64][2];
  const SegBlock b = blocks[blockIdx.x];
  const int w = cplx ? 2 : 1;
  for (int c = 0; c < ncol; ++c) {
    const double* x = X + c * ldx * w;
    double sr = 0, si = 0;
    if (!cplx) {
      for (int64_t i = b.begin + threadIdx.x; i < b.end; i += kRedBlock) sr = fma(x[i], y[i], sr);
    } else {
      for (int64_t i = b.begin + threadIdx.x; i < b.end; i += kRedBlock) {
        const double xr = x[2 * i], xi = x[2 * i + 1], yr = y[2 * i], yi = y[2 * i + 1];
        sr += xr * yr + xi * yi;
        si += xr * yi - xi * yr;
      }
    }
    for (int off = 32; off > 0; off >>= 1) {
      sr += __shfl_down(sr, off);
      si += __shfl_down(si, off);
    }
    if ((threadIdx.x & 63) == 0) {
      red[threadIdx.x >> 6][0] = sr;
      red[threadIdx.x >> 6][1] = si;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double tr = 0, ti = 0;
      for (int q = 0; q < kRedBlock / 64; ++q) {
        tr += red[q][0];
        ti += red[q][1];
      }
      partial[(int64_t(blockIdx.x) * ncol + c) * 2] = tr;
      partial[(int64_t(blockIdx.x) * ncol + c) * 2 + 1] = ti;
    }
    __syncthreads();
  }
}

__global__ void seg_reduce_kernel(const double* __restrict__ partial, const int* __restrict__ first, int ninst,
                                  int ncol, double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ninst * ncol) return;
  const int i = t / ncol, c = t % ncol;
  double sr = 0, si = 0;
  for (int b = first[i]; b < first[i + 1]; ++b) {
    sr += partial[(int64_t(b) * ncol + c) * 2];
    si += partial[(int64_t(b) * ncol + c) * 2 + 1];
  }
  out[int64_t(t) * 2] = sr;
  out[int64_t(t) * 2 + 1] = si;
}

__global__ __launch_bounds__(256) void seg_combine_kernel(const double* __restrict__ X, int64_t ldx, int ncol,
                                                          const double* __restrict__ coef,
                                                          const double* __restrict__ beta, double* __restrict__ y,
                                                          const SegBlock* __restrict__ blocks, int cplx) {
  const SegBlock b = blocks[blockIdx.x];
  const double* cf = coef + int64_t(b.inst) * ncol * 2;
  const double bt = beta[b.inst];
  if (!cplx) {
    for (int64_t i = b.begin + threadIdx.x; i < b.end; i += 256) {
      double acc = bt == 0.0 ? 0.0 : bt * y[i];
      for (int c = 0; c < ncol; ++c) acc = fma(cf[2 * c], X[c * ldx + i], acc);
      y[i] = acc;
    }
  } else {
    for (int64_t i = b.begin + threadIdx.x; i < b.end; i += 256) {
      double ar = bt == 0.0 ? 0.0 : bt * y[2 * i], ai = bt == 0.0 ? 0.0 : bt * y[2 * i + 1];
      for (int c = 0; c < ncol; ++c) {
        const double xr = X[(c * ldx + i) * 2], xi = X[(c * ldx + i) * 2 + 1];
        const double cr = cf[2 * c], ci = cf[2 * c + 1];
        ar = fma(cr, xr, fma(-ci, xi, ar));
        ai = fma(cr, xi, fma(ci, xr, ai));
      }
      y[2 * i] = ar;
      y[2 * i + 1] = ai;
    }
  }
}

hipError_t launch_deinterleave(int64_t n, int64_t n4, const double* src, void* dst, int out_type, hipStream_t st) {
  if (n4 <= 0) return hipSuccess;
  if (n > n4) return hipErrorInvalidValue;
  if (out_type == 0)
    hipLaunchKernelGGL(deinterleave_kernel<double>, dim3(grid_for(n4)), dim3(256), 0, st, n, n4, src,
                       static_cast<double*>(dst));
  else
    hipLaunchKernelGGL(deinterleave_kernel<float>, dim3(grid_for(n4)), dim3(256), 0, st, n, n4, src,
                       static_cast<float*>(dst));
  return hipGetLastError();
}

hipError_t launch_interleave(int64_t n, int64_t n4, const double* src, double* dst, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  if (n > n4) return hipErrorInvalidValue;
  hipLaunchKernelGGL(interleave_kernel, dim3(grid_for(n)), dim3(256), 0, st, n, n4, src, dst);
  return hipGetLastError();
}

hipError_t launch_seg_dot(const double* X, int64_t ldx, int ncol, const double* y, const SegBlock* blocks, int nblocks,
                          int cplx, double* partial, hipStream_t st) {
  if (nblocks <= 0 || ncol <= 0) return hipSuccess;
  hipLaunchKernelGGL(seg_dot_kernel, dim3(nblocks), dim3(kRedBlock), 0, st, X, ldx, ncol, y, blocks, cplx, partial);
  return hipGetLastError();
}

hipError_t launch_seg_reduce(const double* partial, const int* first, int ninst, int ncol, double* out,
                             hipStream_t st) {
  if (ninst <= 0 || ncol <= 0) return hipSuccess;
  const int n = ninst * ncol;
  hipLaunchKernelGGL(seg_reduce_kernel, dim3((n + 255) / 256), dim3(256), 0, st, partial, first, ninst, ncol, out);
  return hipGetLastError();
}

hipError_t launch_seg_combine(const double* X, int64_t ldx, int ncol, const double* coef, const double* beta,
                              double* y, const SegBlock* blocks, int nblocks, int cplx, hipStream_t st) {
  if (nblocks <= 0) return hipSuccess;
  hipLaunchKernelGGL(seg_combine_kernel, dim3(nblocks), dim3(256), 0, st, X, ldx, ncol, coef, beta, y, blocks, cplx);
  return hipGetLastError();
}

hipError_t launch_gemv_t(const double* X, int64_t ldx, int ncol, const double* y, int64_t n, int cplx,
                         double* partial, double sign, double* out, hipStream_t st) {
  if (ncol <= 0) return hipSuccess;
  if (ncol > 64) return hipErrorInvalidValue;  // partial holds 2 * kPartialGrid * 64 doubles
  int g = grid_for(n, kRedBlock);
  if (g > kRedGrid) g = kRedGrid;
  hipLaunchKernelGGL(gemv_t_partial, dim3(g), dim3(kRedBlock), 0, st, X, ldx, ncol, nullptr, y, n, cplx, partial);
  hipLaunchKernelGGL(gemv_t_finish, dim3(ncol), dim3(64), 0, st, partial, g, ncol, nullptr, sign, out);
  return hipGetLastError();
}

hipError_t launch_gemv_n_update(const double* X, int64_t ldx, int ncol, const double* coef, double* y, int64_t n,
                                int cplx, hipStream_t st) {
  if (ncol <= 0) return hipSuccess;
  hipLaunchKernelGGL(gemv_n_update, dim3(grid_for(n)), dim3(256), 0, st, X, ldx, ncol, nullptr, coef, y, n, cplx);
  return hipGetLastError();
}

hipError_t launch_gemv_t_dev(const double* X, int64_t ldx, int maxcol, const int* dncol, const double* y, int64_t n,
                             int cplx, double* partial, double sign, double* out, hipStream_t st) {
  if (maxcol <= 0) return hipSuccess;
  if (maxcol > 64 || dncol == nullptr) return hipErrorInvalidValue;
  int g = grid_for(n, kRedBlock);
  if (g > kRedGrid) g = kRedGrid;
  hipLaunchKernelGGL(gemv_t_partial, dim3(g), dim3(kRedBlock), 0, st, X, ldx, maxcol, dncol, y, n, cplx, partial);
  hipLaunchKernelGGL(gemv_t_finish, dim3(maxcol), dim3(64), 0, st, partial, g, maxcol, dncol, sign, out);
  return hipGetLastError();
}

hipError_t launch_gemv_n_update_dev(const double* X, int64_t ldx, int maxcol, const int* dncol, const double* coef,
                                    double* y, int64_t n, int cplx, hipStream_t st) {
  if (maxcol <= 0) return hipSuccess;
  if (dncol == nullptr) return hipErrorInvalidValue;
  hipLaunchKernelGGL(gemv_n_update, dim3(grid_for(n)), dim3(256), 0, st, X, ldx, maxcol, dncol, coef, y, n, cplx);
  return hipGetLastError();
}

hipError_t launch_lanczos_update(LanczosDev* s, double* w, const double* x, const double* V, int64_t ldv, int64_t nd,
                                 hipStream_t st) {
  hipLaunchKernelGGL(lanczos_update_kernel, dim3(grid_for(nd)), dim3(256), 0, st, s, w, x, V, ldv, nd);
  return hipGetLastError();
}

hipError_t launch_lanczos_finish(LanczosDev* s, const double* w, double* x, double* V, int64_t ldv, int64_t nd,
                                 double tol, hipStream_t st) {
  hipLaunchKernelGGL(lanczos_finish_kernel, dim3(grid_for(nd)), dim3(256), 0, st, s, w, x, V, ldv, nd, tol);
  return hipGetLastError();
}

hipError_t launch_lanczos_advance(LanczosDev* s, hipStream_t st) {
  hipLaunchKernelGGL(lanczos_advance_kernel, dim3(1), dim3(64), 0, st, s);
  return hipGetLastError();
}

hipError_t launch_axpby(int64_t n, double a, double b, const double* x, double* y, hipStream_t st) {
  hipLaunchKernelGGL(axpby_kernel, dim3(grid_for(n)), dim3(256), 0, st, n, a, b, x, y);
  return hipGetLastError();
}

hipError_t launch_zscal(int64_t n, double sr, double si, const double* x, double* y, hipStream_t st) {
  hipLaunchKernelGGL(zscal_kernel, dim3(grid_for(n)), dim3(256), 0, st, n, sr, si, x, y);
  return hipGetLastError();
}

hipError_t launch_f64_to_f32(int64_t n, const double* x, float* y, hipStream_t st) {
  hipLaunchKernelGGL(f64_to_f32_kernel, dim3(grid_for(n)), dim3(256), 0, st, n, x, y);
  return hipGetLastError();
}

hipError_t launch_f64_to_bf16(int64_t n, const double* x, uint16_t* y, hipStream_t st) {
  hipLaunchKernelGGL(f64_to_bf16_kernel, dim3(grid_for(n)), dim3(256), 0, st, n, x, reinterpret_cast<bf16_t*>(y));
  return hipGetLastError();
}

hipError_t launch_f32_to_f64(int64_t n, const float* x, double* y, hipStream_t st) {
  hipLaunchKernelGGL(f32_to_f64_kernel, dim3(grid_for(n)), dim3(256), 0, st, n, x, y);
  return hipGetLastError();
}

hipError_t launch_fill(int64_t n, double v, double* y, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_kernel, dim3(grid_for(n)), dim3(256), 0, st, n, v, y);
  return hipGetLastError();
}

hipError_t launch_copy2d(int rows, int cols, const double* src, int64_t s_row, int64_t s_col, double* dst,
                         int64_t ldd, int w, double scale, const double* colscale, int accumulate, hipStream_t st) {
  const int64_t total = int64_t(rows) * cols;
  if (total <= 0) return hipSuccess;
  hipLaunchKernelGGL(copy2d_kernel, dim3(grid_for(total)), dim3(256), 0, st, rows, cols, src, s_row, s_col, dst,
                     ldd, w, scale, colscale, accumulate);
  return hipGetLastError();
}

hipError_t launch_strided_norms(int count, int len, const double* a, int64_t s_elem, int64_t s_vec, double* out,
                                hipStream_t st) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(strided_norms_kernel, dim3(count), dim3(256), 0, st, count, len, a, s_elem, s_vec, out);
  return hipGetLastError();
}

hipError_t launch_strided_scale(int count, int len, double* a, int64_t s_elem, int64_t s_vec, const double* f,
                                hipStream_t st) {
  const int64_t total = int64_t(count) * len;
  if (total <= 0) return hipSuccess;
  hipLaunchKernelGGL(strided_scale_kernel, dim3(grid_for(total)), dim3(256), 0, st, count, len, a, s_elem, s_vec, f);
  return hipGetLastError();
}

hipError_t launch_strided_fill(int64_t n, double v, double* p, int64_t stride, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(strided_fill_kernel, dim3(grid_for(n)), dim3(256), 0, st, n, v, p, stride);
  return hipGetLastError();
}

}  // namespace sadmrg
