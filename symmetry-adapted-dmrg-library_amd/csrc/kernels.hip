// kernels.hip -- hand-written gfx950 kernels for the two-site DMRG inner loop.
//
//  * grouped_gemm_kernel : one launch executes a whole block table (see
//    grouped_gemm.hpp).  Each workgroup owns one TM x TN output tile of one
//    problem and walks that problem's K segments (all sector pairs / bond
//    terms that feed the tile), accumulating in MFMA registers.  No atomics:
//    every output element is produced by exactly one workgroup in a fixed
//    order, so results are bitwise reproducible run to run.
//      FP64: v_mfma_f64_16x16x4f64, FP32: v_mfma_f32_16x16x4f32.
//  * vector kernels used by Lanczos: deterministic two-stage reductions for
//    dot products / norms, axpy with a device-resident coefficient, block
//    Gram-Schmidt (V^T w and w -= V c), precision conversion.
#include <hip/hip_runtime.h>
#include <cstdint>
#include "grouped_gemm.hpp"
#include "kernels.hpp"

namespace sadmrg {

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

__device__ inline d4_t mfma_16x16x4(double a, double b, d4_t c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ inline f4_t mfma_16x16x4(float a, float b, f4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

template <class T> struct Acc4;
template <> struct Acc4<double> { typedef d4_t type; };
template <> struct Acc4<float> { typedef f4_t type; };

// Output row of accumulator element r held by lane group g = lane / 16 (the
// column is lane % 16).  The FP64 16x16x4 MFMA interleaves rows across lane
// groups (row = g + 4r); the FP32 16x16x4 MFMA gives each lane group four
// consecutive rows (row = 4g + r).
template <class T> __device__ inline int acc_row(int g, int r);
template <> __device__ inline int acc_row<double>(int g, int r) { return g + 4 * r; }
template <> __device__ inline int acc_row<float>(int g, int r) { return 4 * g + r; }

__device__ inline const void* arena_ptr(const ArenaTable& at, int64_t off, int elem_size) {
  const int a = static_cast<int>(static_cast<uint64_t>(off) >> kArenaShift);
  const int64_t o = off & ((int64_t(1) << kArenaShift) - 1);
  return static_cast<const char*>(at.base[a]) + o * elem_size;
}

// Cursor over the flattened (segment, k0) sequence of one problem.
struct KCursor {
  int seg, seg_end, k0;
};

// TM x TN tile, TK deep, 256 threads as WM x WN waves.
template <class T, class TOut, int TM, int TN, int TK, int WM, int WN>
__global__ __launch_bounds__(256) void grouped_gemm_kernel(const GemmTile* __restrict__ tiles,
                                                           const GemmProb* __restrict__ probs,
                                                           const GemmSeg* __restrict__ segs,
                                                           ArenaTable arenas, TOut* __restrict__ partial,
                                                           int tile_base) {
  static_assert(WM * WN == 4, "four waves per workgroup");
  constexpr int FM = TM / WM / 16;  // 16x16 fragments per wave along M
  constexpr int FN = TN / WN / 16;
  constexpr int PAD = 2;
  constexpr int A_ELEMS = TM * TK / 256;  // elements each thread stages per step
  constexpr int B_ELEMS = TN * TK / 256;
  typedef typename Acc4<T>::type acc_t;

  __shared__ T sA[TK][TM + PAD];
  __shared__ T sB[TK][TN + PAD];

  const GemmTile tile = tiles[tile_base + blockIdx.x];
  const GemmProb prob = probs[tile.prob];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave % WM;
  const int wn = wave / WM;

  acc_t acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

  T ra[A_ELEMS], rb[B_ELEMS];

  // Stage the (seg, k0) step into registers.
  auto load_step = [&](int si, int k0) {
    const GemmSeg sg = segs[si];
    const T* A = static_cast<const T*>(arena_ptr(arenas, sg.a_off, sizeof(T)));
    const T* B = static_cast<const T*>(arena_ptr(arenas, sg.b_off, sizeof(T)));
    const bool ta = sg.flags & kTransA;
    const bool tb = sg.flags & kTransB;
    const T scale = static_cast<T>(sg.scale);
#pragma unroll
    for (int e = 0; e < A_ELEMS; ++e) {
      int m, k;
      if (!ta) { const int idx = tid + e * 256; m = idx % TM; k = idx / TM; }
      else     { const int idx = tid + e * 256; k = idx % TK; m = idx / TK; }
      const int gm = tile.m0 + m, gk = k0 + k;
      T v = T(0);
      if (gm < prob.M && gk < sg.K) v = ta ? A[gk + int64_t(gm) * sg.lda] : A[gm + int64_t(gk) * sg.lda];
      ra[e] = v * scale;
    }
#pragma unroll
    for (int e = 0; e < B_ELEMS; ++e) {
      int n, k;
      if (tb) { const int idx = tid + e * 256; n = idx % TN; k = idx / TN; }
      else    { const int idx = tid + e * 256; k = idx % TK; n = idx / TK; }
      const int gn = tile.n0 + n, gk = k0 + k;
      T v = T(0);
      if (gn < prob.N && gk < sg.K) v = tb ? B[gn + int64_t(gk) * sg.ldb] : B[gk + int64_t(gn) * sg.ldb];
      rb[e] = v;
    }
  };
  auto store_step = [&](int si) {
    const int fl = segs[si].flags;
    const bool ta = fl & kTransA;
    const bool tb = fl & kTransB;
#pragma unroll
    for (int e = 0; e < A_ELEMS; ++e) {
      int m, k;
      if (!ta) { const int idx = tid + e * 256; m = idx % TM; k = idx / TM; }
      else     { const int idx = tid + e * 256; k = idx % TK; m = idx / TK; }
      sA[k][m] = ra[e];
    }
#pragma unroll
    for (int e = 0; e < B_ELEMS; ++e) {
      int n, k;
      if (tb) { const int idx = tid + e * 256; n = idx % TN; k = idx / TN; }
      else    { const int idx = tid + e * 256; k = idx % TK; n = idx / TK; }
      sB[k][n] = rb[e];
    }
  };

  KCursor cur{tile.seg_lo, tile.seg_hi, 0};
  // skip empty segments
  while (cur.seg < cur.seg_end && segs[cur.seg].K <= 0) ++cur.seg;
  if (cur.seg < cur.seg_end) load_step(cur.seg, 0);

  while (cur.seg < cur.seg_end) {
    const int this_seg = cur.seg;
    // advance cursor to the next step
    cur.k0 += TK;
    if (cur.k0 >= segs[cur.seg].K) {
      cur.k0 = 0;
      ++cur.seg;
      while (cur.seg < cur.seg_end && segs[cur.seg].K <= 0) ++cur.seg;
    }
    __syncthreads();
    store_step(this_seg);
    __syncthreads();
    if (cur.seg < cur.seg_end) load_step(cur.seg, cur.k0);  // prefetch overlaps the MFMAs
#pragma unroll
    for (int kk = 0; kk < TK; kk += 4) {
      const int kr = kk + (lane >> 4);
      T af[FM], bf[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) af[i] = sA[kr][(wm * FM + i) * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < FN; ++j) bf[j] = sB[kr][(wn * FN + j) * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = mfma_16x16x4(af[i], bf[j], acc[i][j]);
    }
  }

  if (tile.part >= 0) {
    // split reduction: dense partial tile, reduced later in slot order
    TOut* P = partial + int64_t(tile.part) * (TM * TN);
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        const int n = (wn * FN + j) * 16 + (lane & 15);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int m = (wm * FM + i) * 16 + acc_row<T>(lane >> 4, r);
          P[m + n * TM] = static_cast<TOut>(acc[i][j][r]);
        }
      }
    return;
  }
  TOut* C = static_cast<TOut*>(const_cast<void*>(arena_ptr(arenas, prob.c_off, sizeof(TOut))));
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
        TOut* p = C + gm + int64_t(gn) * prob.ldc;
        const TOut v = static_cast<TOut>(acc[i][j][r]);
        if (prob.accumulate) *p += v;
        else *p = v;
      }
    }
}

// Folds split-reduction partial tiles into C: one workgroup per output tile,
// slots summed in increasing order.
template <class TOut, int TM, int TN>
__global__ __launch_bounds__(256) void gemm_reduce_kernel(const GemmReduce* __restrict__ tasks,
                                                          const GemmProb* __restrict__ probs,
                                                          ArenaTable arenas, const TOut* __restrict__ partial) {
  const GemmReduce t = tasks[blockIdx.x];
  const GemmProb prob = probs[t.prob];
  TOut* C = static_cast<TOut*>(const_cast<void*>(arena_ptr(arenas, prob.c_off, sizeof(TOut))));
  const int rows = min(TM, prob.M - t.m0);
  const int cols = min(TN, prob.N - t.n0);
  for (int idx = threadIdx.x; idx < TM * TN; idx += 256) {
    const int m = idx % TM, n = idx / TM;
    if (m >= rows || n >= cols) continue;
    const TOut* P = partial + int64_t(t.part_begin) * (TM * TN) + idx;
    TOut v = P[0];
    for (int q = 1; q < t.nparts; ++q) v += P[int64_t(q) * (TM * TN)];
    TOut* p = C + (t.m0 + m) + int64_t(t.n0 + n) * prob.ldc;
    if (prob.accumulate) *p += v;
    else *p = v;
  }
}

// Complex grouped GEMM on interleaved (re, im) double data.  Four real
// MFMA products per step; conjugation flags select the signs.
template <int TM, int TN, int TK, int WM, int WN>
__global__ __launch_bounds__(256) void grouped_zgemm_kernel(const GemmTile* __restrict__ tiles,
                                                            const GemmProb* __restrict__ probs,
                                                            const GemmSeg* __restrict__ segs,
                                                            ArenaTable arenas, int tile_base) {
  constexpr int FM = TM / WM / 16;
  constexpr int FN = TN / WN / 16;
  constexpr int PAD = 2;
  constexpr int A_ELEMS = TM * TK / 256;
  constexpr int B_ELEMS = TN * TK / 256;
  __shared__ double sAr[TK][TM + PAD], sAi[TK][TM + PAD];
  __shared__ double sBr[TK][TN + PAD], sBi[TK][TN + PAD];

  const GemmTile tile = tiles[tile_base + blockIdx.x];
  const GemmProb prob = probs[tile.prob];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave % WM;
  const int wn = wave / WM;

  d4_t accr[FM][FN], acci[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) { accr[i][j] = d4_t{0, 0, 0, 0}; acci[i][j] = d4_t{0, 0, 0, 0}; }

  for (int si = tile.seg_lo; si < tile.seg_hi; ++si) {
    const GemmSeg sg = segs[si];
    const double* A = static_cast<const double*>(arena_ptr(arenas, sg.a_off, 16));
    const double* B = static_cast<const double*>(arena_ptr(arenas, sg.b_off, 16));
    const bool ta = sg.flags & kTransA, tb = sg.flags & kTransB;
    const double ca = (sg.flags & kConjA) ? -1.0 : 1.0;
    const double cb = (sg.flags & kConjB) ? -1.0 : 1.0;
    for (int k0 = 0; k0 < sg.K; k0 += TK) {
      __syncthreads();
#pragma unroll
      for (int e = 0; e < A_ELEMS; ++e) {
        int m, k;
        const int idx = tid + e * 256;
        if (!ta) { m = idx % TM; k = idx / TM; } else { k = idx % TK; m = idx / TK; }
        const int gm = tile.m0 + m, gk = k0 + k;
        double xr = 0, xi = 0;
        if (gm < prob.M && gk < sg.K) {
          const int64_t o = ta ? (gk + int64_t(gm) * sg.lda) : (gm + int64_t(gk) * sg.lda);
          xr = A[2 * o];
          xi = ca * A[2 * o + 1];
        }
        sAr[k][m] = xr * sg.scale - xi * sg.scale_im;
        sAi[k][m] = xr * sg.scale_im + xi * sg.scale;
      }
#pragma unroll
      for (int e = 0; e < B_ELEMS; ++e) {
        int n, k;
        const int idx = tid + e * 256;
        if (tb) { n = idx % TN; k = idx / TN; } else { k = idx % TK; n = idx / TK; }
        const int gn = tile.n0 + n, gk = k0 + k;
        double xr = 0, xi = 0;
        if (gn < prob.N && gk < sg.K) {
          const int64_t o = tb ? (gn + int64_t(gk) * sg.ldb) : (gk + int64_t(gn) * sg.ldb);
          xr = B[2 * o];
          xi = cb * B[2 * o + 1];
        }
        sBr[k][n] = xr;
        sBi[k][n] = xi;
      }
      __syncthreads();
#pragma unroll
      for (int kk = 0; kk < TK; kk += 4) {
        const int kr = kk + (lane >> 4);
        double ar[FM], ai[FM], nai[FM], br[FN], bi[FN];
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
  }

  double* C = static_cast<double*>(const_cast<void*>(arena_ptr(arenas, prob.c_off, 16)));
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int gn = tile.n0 + (wn * FN + j) * 16 + (lane & 15);
      if (gn >= prob.N) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gm = tile.m0 + (wm * FM + i) * 16 + acc_row<double>(lane >> 4, r);
        if (gm >= prob.M) continue;
        double* p = C + 2 * (gm + int64_t(gn) * prob.ldc);
        if (prob.accumulate) { p[0] += accr[i][j][r]; p[1] += acci[i][j][r]; }
        else { p[0] = accr[i][j][r]; p[1] = acci[i][j][r]; }
      }
    }
}

// ---------------------------------------------------------------------------
// Launch helpers.  The grid is split in chunks so that gridDim.x stays within
// limits for very large tile tables.
template <class T, class TOut, int TM, int TN, int TK, int WM, int WN>
static hipError_t launch_real(const GemmLaunch& L, hipStream_t st) {
  TOut* part = static_cast<TOut*>(L.partial);
  const int chunk = 1 << 30;
  for (int base = 0; base < L.ntiles; base += chunk) {
    const int n = L.ntiles - base < chunk ? L.ntiles - base : chunk;
    hipLaunchKernelGGL((grouped_gemm_kernel<T, TOut, TM, TN, TK, WM, WN>), dim3(n), dim3(256), 0, st,
                       L.tiles, L.probs, L.segs, L.arenas, part, base);
  }
  if (L.nreduce > 0)
    hipLaunchKernelGGL((gemm_reduce_kernel<TOut, TM, TN>), dim3(L.nreduce), dim3(256), 0, st, L.reduce, L.probs,
                       L.arenas, static_cast<const TOut*>(part));
  return hipGetLastError();
}

hipError_t launch_grouped_gemm(GemmPrec prec, int tile_cfg, const GemmLaunch& L, hipStream_t st) {
  if (L.ntiles <= 0) return hipSuccess;
  if (L.nreduce > 0 && L.partial == nullptr) return hipErrorInvalidValue;
  switch (prec) {
    case GemmPrec::F64:
      if (tile_cfg == 1) return launch_real<double, double, 128, 128, 16, 2, 2>(L, st);
      return launch_real<double, double, 64, 64, 16, 2, 2>(L, st);
    case GemmPrec::F32:
      if (tile_cfg == 1) return launch_real<float, float, 128, 128, 16, 2, 2>(L, st);
      return launch_real<float, float, 64, 64, 16, 2, 2>(L, st);
    case GemmPrec::F32_OUT64:
      if (tile_cfg == 1) return launch_real<float, double, 128, 128, 16, 2, 2>(L, st);
      return launch_real<float, double, 64, 64, 16, 2, 2>(L, st);
    case GemmPrec::Z64: {
      if (L.nreduce > 0 || tile_cfg != 0) return hipErrorInvalidValue;  // no split or large tiles for complex
      hipLaunchKernelGGL((grouped_zgemm_kernel<64, 64, 16, 2, 2>), dim3(L.ntiles), dim3(256), 0, st, L.tiles,
                         L.probs, L.segs, L.arenas, 0);
      return hipGetLastError();
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
__global__ __launch_bounds__(kRedBlock) void gemv_t_partial(const double* __restrict__ X, int64_t ldx,
                                                            int ncol, const double* __restrict__ y,
                                                            int64_t n, int cplx,
                                                            double* __restrict__ partial) {
  __shared__ double red[kRedBlock / 64][2];
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
__global__ void gemv_t_finish(const double* __restrict__ partial, int ngrid, int ncol, double sign,
                              double* __restrict__ out) {
  const int c = blockIdx.x;
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
__global__ void gemv_n_update(const double* __restrict__ X, int64_t ldx, int ncol,
                              const double* __restrict__ coef, double* __restrict__ y, int64_t n, int cplx) {
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

// y = a*y + b*x with host scalars (real)
__global__ void axpby_kernel(int64_t n, double a, double b, const double* __restrict__ x, double* __restrict__ y) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += stride)
    y[i] = a * y[i] + (x ? b * x[i] : 0.0);
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
  const int64_t total = int64_t(t.rows) * t.cols;
  for (int64_t e = threadIdx.x; e < total; e += blockDim.x) {
    const int r = static_cast<int>(e % t.rows);
    const int c = static_cast<int>(e / t.rows);
    const double* s = src + (t.src_off + r + int64_t(c) * t.lds) * w;
    double* d = dst + (t.dst_off + r + int64_t(c) * t.ldd) * w;
    for (int k = 0; k < w; ++k) d[k] = t.scale * s[k];
  }
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

hipError_t launch_gemv_t(const double* X, int64_t ldx, int ncol, const double* y, int64_t n, int cplx,
                         double* partial, double sign, double* out, hipStream_t st) {
  if (ncol <= 0) return hipSuccess;
  if (ncol > 64) return hipErrorInvalidValue;  // partial holds 2 * kPartialGrid * 64 doubles
  int g = grid_for(n, kRedBlock);
  if (g > kRedGrid) g = kRedGrid;
  hipLaunchKernelGGL(gemv_t_partial, dim3(g), dim3(kRedBlock), 0, st, X, ldx, ncol, y, n, cplx, partial);
  hipLaunchKernelGGL(gemv_t_finish, dim3(ncol), dim3(64), 0, st, partial, g, ncol, sign, out);
  return hipGetLastError();
}

hipError_t launch_gemv_n_update(const double* X, int64_t ldx, int ncol, const double* coef, double* y, int64_t n,
                                int cplx, hipStream_t st) {
  if (ncol <= 0) return hipSuccess;
  hipLaunchKernelGGL(gemv_n_update, dim3(grid_for(n)), dim3(256), 0, st, X, ldx, ncol, coef, y, n, cplx);
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

hipError_t launch_strided_fill(int64_t n, double v, double* p, int64_t stride, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(strided_fill_kernel, dim3(grid_for(n)), dim3(256), 0, st, n, v, p, stride);
  return hipGetLastError();
}

}  // namespace sadmrg
