// eigen.hip -- see eigen.hpp.
#include "eigen.hpp"
#include "kernels.hpp"
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>
#include <algorithm>
#include <atomic>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>

// Scratch of the allocation-free phase 1 (SymEigBatch::eigenvalues_noalloc):
// one copy per device, held until the process ends.  Host code uses the
// device address looked up by noalloc_scratch_setup, never this name.
static __device__ __align__(256) double scratch[sadmrg::SymEigBatch::kNoAllocBytes / 8];

namespace sadmrg {

namespace {

// ---------------------------------------------------------------------------
// 1. Tridiagonalisation, one persistent cooperative launch for every matrix.
//
// Matrix q is served by a group of nwg workgroups; workgroup w holds rows
// r = w + nwg * i (i < rows) of the full symmetric matrix in LDS.  Step k
// (k = 0 .. n-2) of the reduction:
//   row pass   every workgroup applies the rank-2 update of step k-1 to its
//              rows (A -= v w^T + w v^T on the trailing block), computes its
//              rows of p = tau_k A v_k, and reads its entries of column k+1;
//   publish    those entries go to the group's buffer as tagged words: each
//              double as two 64-bit words (one 32-bit half and the step tag
//              k+1), each stored with one write-through store, so a reader
//              sees either the old word or the whole new one;
//   gather     every workgroup polls the buffer until all entries of step k
//              carry tag k+1 (no atomics, no fence: the data is the signal);
//   redundant  every workgroup computes, identically (same data, same
//              operation order), alpha = v^T p, w = p - tau alpha v / 2,
//              column k+1 of the updated matrix, d[k+1] and the next
//              reflector (dlarfg: beta, tau, v).
// Workgroup 0 of the group writes d, e, tau and the reflector vectors
// (LAPACK layout: reflector i acts on rows i+1..n-1, v[i+1] = 1,
// v[i+2..n-1] stored in column i below the subdiagonal).  The buffers of
// even and odd steps are separate, so a workgroup that runs ahead cannot
// overwrite entries a slower one still reads.  A workgroup that waits longer
// than the wall-clock limit passed by the host (a quarter second; a step
// normally waits microseconds), or past kSpinLimit polls, sets the error flag,
// and every workgroup then ends.
constexpr int kTrdThreads = 512;
constexpr int kTrdWaves = kTrdThreads / 64;
constexpr int kLdsBytes = 160 * 1024;
constexpr int kTrdLdsDoubles = (kLdsBytes - 512) / 8;  // dynamic part; the rest is static shared
constexpr long long kSpinLimit = 1LL << 22;
constexpr int kCuMargin = 16;

struct TrdTask {
  double* a;
  double* d;
  double* e;
  double* tau;
  uint64_t* pub;  // 2 parities x n rows x 4 tagged words {p lo, p hi, col lo, col hi}
  int n, lda, wg0, nwg, rows;  // rows: LDS-held rows per workgroup
  int rreg;                     // register-held rows per workgroup (one per wave, waves 0 .. rreg-1)
  int ldr;        // LDS stride of rows and vectors: n rounded up to a multiple of 64
};

// DPP move of a double (both halves), lanes without a source reading zero.
template <int Ctrl, int RowMask = 0xf>
__device__ inline double dpp_move(double x) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), Ctrl, RowMask, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), Ctrl, RowMask, 0xf, true);
  return __hiloint2double(hi, lo);
}

// Sum over the 64 lanes of a wave, in a fixed order, with DPP row shifts and
// row broadcasts (no LDS crossbar); the result is returned to every lane.
// Every lane of the wave must take part.
__device__ inline double wave_sum(double x) {
  x += dpp_move<0x111>(x);        // row_shr:1
  x += dpp_move<0x112>(x);        // row_shr:2
  x += dpp_move<0x114>(x);        // row_shr:4
  x += dpp_move<0x118>(x);        // row_shr:8   (lane 15 of a row: the row's sum)
  x += dpp_move<0x142, 0xa>(x);   // row_bcast:15 (lane 31: rows 0-1, lane 63: rows 2-3)
  x += dpp_move<0x143, 0xc>(x);   // row_bcast:31 (lane 63: all rows)
  const int lo = __builtin_amdgcn_readlane(__double2loint(x), 63);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(x), 63);
  return __hiloint2double(hi, lo);
}

// Sum over the workgroup; every thread returns the same value (the wave sums
// are added in wave order).
__device__ inline double block_sum(double x, double* red) {
  x = wave_sum(x);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
  __syncthreads();
  double s = 0;
  for (int i = 0; i < kTrdWaves; ++i) s += red[i];
  return s;
}

// dlarfg on c[lo .. n): H (c[lo], c[lo+1..]) = (beta, 0).  c becomes v (v[lo] = 1);
// returns beta, *tau = tau.
__device__ double householder(double* c, int lo, int n, double* tau, double* red) {
  double s = 0;
  for (int j = lo + 1 + static_cast<int>(threadIdx.x); j < n; j += kTrdThreads) s += c[j] * c[j];
  const double sig = block_sum(s, red);
  const double x0 = c[lo];
  double beta, t, scale;
  if (sig == 0.0) {
    beta = x0;
    t = 0.0;
    scale = 0.0;
  } else {
    beta = -copysign(sqrt(fma(x0, x0, sig)), x0);
    t = (beta - x0) / beta;
    scale = 1.0 / (x0 - beta);
  }
  __syncthreads();  // c[lo] read by every thread
  for (int j = lo + 1 + static_cast<int>(threadIdx.x); j < n; j += kTrdThreads) c[j] *= scale;
  if (threadIdx.x == 0) c[lo] = 1.0;
  __syncthreads();
  *tau = t;
  return beta;
}

__device__ inline void put_tagged(uint64_t* p, double x, uint32_t tag) {
  const uint64_t b = __double_as_longlong(x);
  __hip_atomic_store(p, (b & 0xffffffffull) | (uint64_t(tag) << 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(p + 1, (b >> 32) | (uint64_t(tag) << 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Register-held rows (one per wave) fit next to the rest of the kernel's
// registers for rows of up to 64 * 24 entries.
constexpr int kTrdMaxRegNT = 24;
template <int NT>
constexpr bool kTrdRegRows = NT <= kTrdMaxRegNT;

// The instantiation (NT) that holds rows of n entries: 64 NT >= n.
inline int trd_template(int n) {
  const int nt = (n + 63) / 64;
  return nt <= 4 ? 4 : nt <= 8 ? 8 : nt <= 12 ? 12 : nt <= 16 ? 16 : nt <= 20 ? 20 : nt <= 24 ? 24 : 32;
}

template <int NT>
__global__ __launch_bounds__(kTrdThreads) void trd_kernel(const TrdTask* __restrict__ tasks, int ntask,
                                                          int* __restrict__ err, long long wait_ticks) {
  extern __shared__ double lds[];
  __shared__ double red[kTrdWaves];
  __shared__ int s_abort;
  int t = 0;
  while (t < ntask && static_cast<int>(blockIdx.x) >= tasks[t].wg0 + tasks[t].nwg) ++t;
  if (t >= ntask) return;
  const TrdTask T = tasks[t];
  const int n = T.n, W = T.nwg, w = static_cast<int>(blockIdx.x) - T.wg0, lda = T.lda, R = T.rows;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  double* A = T.a;
  if (n == 1) {
    if (w == 0 && tid == 0) T.d[0] = A[0];
    return;
  }
  if (tid == 0) s_abort = 0;
  // rows and vectors padded to ldr (a multiple of 64) entries, zero past n;
  // the row pass loads every slice unconditionally (indices clamped to the
  // padded row) and masks by value
  const int ldr = T.ldr;
  double* rows = lds;
  double* v = rows + static_cast<size_t>(R) * ldr;
  double* vp = v + ldr;
  double* wp = vp + ldr;
  // own rows, from the lower triangle: row r = column r below the diagonal,
  // and A(r, j) for j < r
  for (int i = 0; i < R; ++i) {
    const int r = w + W * i;
    for (int j = tid; j < ldr; j += kTrdThreads)
      rows[static_cast<size_t>(i) * ldr + j] =
          (r >= n || j >= n) ? 0.0 : (j >= r ? A[j + static_cast<size_t>(r) * lda] : A[r + static_cast<size_t>(j) * lda]);
  }
  // a register-held row for waves below rreg: row index R + wave, lane l
  // holding entries l + 64 m
  const int rrow = (kTrdRegRows<NT> && wv < T.rreg) ? w + W * (R + wv) : n;  // n: none
  double areg[kTrdRegRows<NT> ? NT : 1];
#pragma unroll
  for (int m = 0; m < (kTrdRegRows<NT> ? NT : 1); ++m) {
    const int j = lane + 64 * m;
    areg[m] = (rrow >= n || j >= n) ? 0.0
                                    : (j >= rrow ? A[j + static_cast<size_t>(rrow) * lda]
                                                 : A[rrow + static_cast<size_t>(j) * lda]);
  }
  for (int j = tid; j < 3 * ldr; j += kTrdThreads) v[j] = 0.0;
  __syncthreads();
  for (int j = 1 + tid; j < n; j += kTrdThreads) v[j] = A[j];
  __syncthreads();
  double tau;
  const double beta0 = householder(v, 1, n, &tau, red);
  if (w == 0 && tid == 0) {
    T.d[0] = A[0];
    T.e[0] = beta0;
    T.tau[0] = tau;
  }
  for (int k = 0; k <= n - 2; ++k) {
    uint64_t* pub = T.pub + static_cast<size_t>(k & 1) * 4 * n;
    const uint32_t tag = static_cast<uint32_t>(k + 1);
    // row pass: update k-1, p = tau A v, column k+1.  This lane's first NR
    // slices of the three vectors are read once per step and held in
    // registers; further slices (rows longer than 64 NR) are read from LDS
    // per row.  Each row is processed in chunks of kChunk slices: loaded,
    // updated, stored back.
    {
      constexpr int NR = NT <= 16 ? NT : (kTrdRegRows<NT> ? 8 : 12);
      constexpr int kChunk = 4;
      static_assert(NR % kChunk == 0 && NT % kChunk == 0, "chunks must tile the register-held and LDS slices");
      double vs[NR], vps[NR], wps[NR];
#pragma unroll
      for (int m = 0; m < NR; ++m) {
        const int j = lane + 64 * m;
        const int jc = min(j, ldr - 1);
        const bool on = j > k && j < n;
        const double a = v[jc], b = vp[jc], c = wp[jc];
        vs[m] = on ? a : 0.0;
        vps[m] = (on && k > 0) ? b : 0.0;
        wps[m] = (on && k > 0) ? c : 0.0;
      }
      for (int i = wv; i < R; i += kTrdWaves) {
        const int r = w + W * i;
        if (r >= n || r <= k) continue;
        double* row = rows + static_cast<size_t>(i) * ldr;
        const double vr = k > 0 ? vp[r] : 0.0, wr = k > 0 ? wp[r] : 0.0;
        double acc = 0.0, cn = 0.0;
        // one chunk of slices m0 .. m0+kChunk-1 (those below NT): load,
        // update, store; `vec` gives slice m's (v, vp, w) entries
        auto chunk = [&](int m0, auto vec) {
          double a[kChunk];
#pragma unroll
          for (int q = 0; q < kChunk; ++q)
            if (m0 + q < NT) a[q] = row[min(lane + 64 * (m0 + q), ldr - 1)];
#pragma unroll
          for (int q = 0; q < kChunk; ++q) {
            const int m = m0 + q;
            if (m < NT) {
              const int j = lane + 64 * m;
              double vsm, vpsm, wpsm;
              vec(q, j, vsm, vpsm, wpsm);
              // vpsm, wpsm, vsm are zero for j <= k and j >= n
              const double x = fma(-wr, vpsm, fma(-vr, wpsm, a[q]));
              acc = fma(x, vsm, acc);
              if (j == k + 1) cn = x;
              a[q] = x;
            }
          }
          if (k > 0) {
#pragma unroll
            for (int q = 0; q < kChunk; ++q) {
              const int j = lane + 64 * (m0 + q);
              if (m0 + q < NT && j > k && j < n) row[j] = a[q];
            }
          }
        };
        // register-held slices (fully unrolled: the array indices are constants)
#pragma unroll
        for (int m0 = 0; m0 < NR; m0 += kChunk)
          chunk(m0, [&](int q, int, double& x, double& y, double& z) {
            x = vs[m0 + q];
            y = vps[m0 + q];
            z = wps[m0 + q];
          });
        // further slices from LDS, one chunk at a time
#pragma unroll 1
        for (int m0 = NR; m0 < NT; m0 += kChunk)
          chunk(m0, [&](int, int j, double& x, double& y, double& z) {
            const int jc = min(j, ldr - 1);
            const bool on = j > k && j < n;
            const double a0 = v[jc], b0 = vp[jc], c0 = wp[jc];
            x = on ? a0 : 0.0;
            y = (on && k > 0) ? b0 : 0.0;
            z = (on && k > 0) ? c0 : 0.0;
          });
        acc = wave_sum(acc);
        if (lane == 0) put_tagged(pub + 4 * static_cast<size_t>(r), tau * acc, tag);
        if (lane == ((k + 1) & 63)) put_tagged(pub + 4 * static_cast<size_t>(r) + 2, cn, tag);
      }
      // the wave's register-held row (wave-uniform condition)
      if constexpr (kTrdRegRows<NT>) if (rrow < n && rrow > k) {
        const double vr = k > 0 ? vp[rrow] : 0.0, wr = k > 0 ? wp[rrow] : 0.0;
        double acc = 0.0, cn = 0.0;
#pragma unroll
        for (int m = 0; m < NT; ++m) {
          const int j = lane + 64 * m;
          double vsm, vpsm, wpsm;
          if (m < NR) {
            vsm = vs[m];
            vpsm = vps[m];
            wpsm = wps[m];
          } else {
            const int jc = min(j, ldr - 1);
            const bool on = j > k && j < n;
            const double a0 = v[jc], b0 = vp[jc], c0 = wp[jc];
            vsm = on ? a0 : 0.0;
            vpsm = (on && k > 0) ? b0 : 0.0;
            wpsm = (on && k > 0) ? c0 : 0.0;
          }
          // zero vector slices leave entries j <= k and j >= n as they are
          const double x = fma(-wr, vpsm, fma(-vr, wpsm, areg[m]));
          acc = fma(x, vsm, acc);
          if (j == k + 1) cn = x;
          areg[m] = x;
        }
        acc = wave_sum(acc);
        if (lane == 0) put_tagged(pub + 4 * static_cast<size_t>(rrow), tau * acc, tag);
        if (lane == ((k + 1) & 63)) put_tagged(pub + 4 * static_cast<size_t>(rrow) + 2, cn, tag);
      }
    }
    // reflector k to column k (every workgroup loaded its rows before it
    // published step 0, and workgroup 0 is past its own gather of step 0 only
    // after every workgroup published)
    if (w == 0 && k > 0)
      for (int j = k + 1 + tid; j < n; j += kTrdThreads) A[j + static_cast<size_t>(k - 1) * lda] = vp[j];
    // gather p and column k+1 into the slots of step k-1's vectors
    double* P = vp;
    double* C = wp;
    double s_alpha = 0.0;
    __syncthreads();  // the row pass read vp, wp
    {
      // every entry this thread gathers is requested at once; entries whose
      // tags are not yet current are requested again
      constexpr int kMaxRows = (2048 + kTrdThreads - 1) / kTrdThreads;
      uint64_t x[kMaxRows][4];
      unsigned pending = 0;
#pragma unroll
      for (int c = 0; c < kMaxRows; ++c) {
        const int j = k + 1 + tid + c * kTrdThreads;
        if (j < n) pending |= 1u << c;
      }
      long long spins = 0;
      const long long t0 = wall_clock64();
      while (pending) {
#pragma unroll
        for (int c = 0; c < kMaxRows; ++c)
          if (pending & (1u << c)) {
            const uint64_t* q = pub + 4 * static_cast<size_t>(k + 1 + tid + c * kTrdThreads);
#pragma unroll
            for (int u = 0; u < 4; ++u) x[c][u] = __hip_atomic_load(q + u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          }
#pragma unroll
        for (int c = 0; c < kMaxRows; ++c)
          if ((pending & (1u << c)) && (x[c][0] >> 32) == tag && (x[c][1] >> 32) == tag && (x[c][2] >> 32) == tag &&
              (x[c][3] >> 32) == tag) {
            const int j = k + 1 + tid + c * kTrdThreads;
            const double pj = __longlong_as_double(static_cast<long long>((x[c][0] & 0xffffffffull) | (x[c][1] << 32)));
            P[j] = pj;
            C[j] = __longlong_as_double(static_cast<long long>((x[c][2] & 0xffffffffull) | (x[c][3] << 32)));
            pending &= ~(1u << c);
          }
        if (!pending) break;
        __builtin_amdgcn_s_sleep(1);
        if (++spins > kSpinLimit || wall_clock64() - t0 > wait_ticks ||
            __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
          atomicExch(err, 1);
          s_abort = 1;
          break;
        }
      }
    }
    // alpha = v^T p, this thread's entries added in row order (not in the
    // order they arrived: every workgroup must form bitwise the same sums);
    // the block sum's barriers also publish the gathered entries
    for (int j = k + 1 + tid; j < n; j += kTrdThreads) s_alpha = fma(v[j], P[j], s_alpha);
    const double alpha = block_sum(s_alpha, red);
    if (s_abort) return;
    const double h = 0.5 * tau * alpha;
    const double vk1 = v[k + 1], wk1 = fma(-h, vk1, P[k + 1]);
    // w = p - h v, column k+1 of the updated matrix, and the squares the next
    // reflector needs, in one pass over this thread's entries
    // (entry k+1 of w is not stored: other threads read P[k+1] for wk1 in
    // this pass, and the next update uses w only past k+1)
    double sig = 0.0;
    for (int j = k + 1 + tid; j < n; j += kTrdThreads) {
      const double wj = fma(-h, v[j], P[j]);
      if (j > k + 1) P[j] = wj;
      const double cj = C[j] - v[j] * wk1 - wj * vk1;
      C[j] = cj;
      if (j > k + 2) sig = fma(cj, cj, sig);
    }
    sig = block_sum(sig, red);
    if (w == 0 && tid == 0) T.d[k + 1] = C[k + 1];
    if (k + 1 <= n - 2) {
      // dlarfg on C[k+2 ..): beta, tau, v (v[k+2] = 1)
      const int lo = k + 2;
      const double x0 = C[lo];
      double bn, tn, scale;
      if (sig == 0.0) {
        bn = x0;
        tn = 0.0;
        scale = 0.0;
      } else {
        bn = -copysign(sqrt(fma(x0, x0, sig)), x0);
        tn = (bn - x0) / bn;
        scale = 1.0 / (x0 - bn);
      }
      __syncthreads();  // every thread has read x0 = C[lo]
      for (int j = k + 1 + tid; j < n; j += kTrdThreads) {
        if (j > lo) C[j] *= scale;
        else if (j == lo) C[j] = 1.0;
      }
      if (w == 0 && tid == 0) {
        T.e[k + 1] = bn;
        T.tau[k + 1] = tn;
      }
      tau = tn;
    }
    double* vn = C;
    vp = v;
    wp = P;
    v = vn;
    __syncthreads();
  }
  // reflectors 0 .. n-3 were stored at steps 1 .. n-2; reflector n-2 has
  // length one and nothing to store
}

template <int NT>
hipError_t launch_trd(const TrdTask* tasks, int ntask, int* err, long long wait_ticks, int nwg, int lds,
                      hipStream_t st) {
  static bool attr = false;
  if (!attr) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(trd_kernel<NT>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, kTrdLdsDoubles * 8);
    if (e != hipSuccess) return e;
    attr = true;
  }
  void* args[] = {&tasks, &ntask, &err, &wait_ticks};
  return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(trd_kernel<NT>), dim3(nwg), dim3(kTrdThreads), args,
                                    static_cast<unsigned>(lds), st);
}

// ---------------------------------------------------------------------------
// 2. Eigenvalues of T by multisection on the Sturm count (the bisection of
// LAPACK dstebz, 17-way): 16 lanes per eigenvalue, lane l evaluating the
// count at the l-th of 16 interior points of the current interval, which
// shrinks to the sub-interval where the count passes the index, a factor 17
// per round (about 13 rounds to working precision: 208 counts per
// eigenvalue, where 64 lanes and 65-way cuts need 576 for nine rounds; the
// counts fill the machine, so their number sets the time, while 8 lanes and
// 9-way cuts make the 17 dependent rounds the bound).  A workgroup holds 16
// eigenvalues of one matrix and first copies the diagonal and the squared
// off-diagonal into LDS, so the count's per-step operands are LDS
// broadcasts instead of global loads.
constexpr int kBisOrder = 2048;  // LDS rows of the copy; no larger order reaches the solver
constexpr int kBisLanes = 16;                    // lanes (interior points) per eigenvalue
constexpr int kBisPerBlock = 4 * 64 / kBisLanes;  // eigenvalues per workgroup of four waves

struct BisTask {
  const double* d;
  const double* e;
  double* w;
  int n, off;  // off: the matrix's first workgroup
};

__device__ inline int sturm_count(const double* d, const double* e2, int n, double x, double pivmin) {
  double q = d[0] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  int c = q < 0.0;
  double dn = n > 1 ? d[1] : 0.0, en = n > 1 ? e2[0] : 0.0;
  for (int j = 1; j < n; ++j) {
    const double dj = dn, e2j = en;
    if (j + 1 < n) {  // operands of the next step, loaded ahead of the division chain
      dn = d[j + 1];
      en = e2[j];
    }
    q = dj - x - e2j / q;
    if (fabs(q) < pivmin) q = -pivmin;
    c += q < 0.0;
  }
  return c;
}

__global__ __launch_bounds__(256) void bisect_kernel(const BisTask* __restrict__ tasks, int ntask) {
  __shared__ double sD[kBisOrder], sE2[kBisOrder];
  const int blk = blockIdx.x;
  int t = 0;
  while (t + 1 < ntask && blk >= tasks[t + 1].off) ++t;
  const BisTask T = tasks[t];
  const int n = T.n;
  const double* d = T.d;
  const double* e = T.e;
  for (int j = threadIdx.x; j < n; j += 256) {
    sD[j] = d[j];
    sE2[j] = j + 1 < n ? e[j] * e[j] : 0.0;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int grp = lane / kBisLanes, sub = lane % kBisLanes;  // kBisLanes lanes per eigenvalue
  const int i0 = (blk - T.off) * kBisPerBlock + (threadIdx.x >> 6) * (64 / kBisLanes);
  if (i0 >= n) return;  // the whole wave is past the matrix
  const int i = i0 + grp;
  // Gershgorin interval and pivot threshold, lanes over the rows
  double lo = INFINITY, hi = -INFINITY, emax2 = 0.0;
  for (int j = lane; j < n; j += 64) {
    const double r = (j > 0 ? fabs(e[j - 1]) : 0.0) + (j + 1 < n ? fabs(e[j]) : 0.0);
    lo = fmin(lo, d[j] - r);
    hi = fmax(hi, d[j] + r);
    if (j + 1 < n) emax2 = fmax(emax2, e[j] * e[j]);
  }
  for (int o = 32; o > 0; o >>= 1) {
    lo = fmin(lo, __shfl_xor(lo, o, 64));
    hi = fmax(hi, __shfl_xor(hi, o, 64));
    emax2 = fmax(emax2, __shfl_xor(emax2, o, 64));
  }
  const double eps = DBL_EPSILON;
  const double tnorm = fmax(fabs(lo), fabs(hi));
  const double pivmin = DBL_MIN * fmax(1.0, emax2);
  const double fudge = 2.1 * eps * tnorm * n + 4.2 * pivmin;
  lo -= fudge;
  hi += fudge;
  const double abstol = eps * tnorm;
  // lo and hi are uniform over a group; a group stops when its interval has
  // converged, the wave when every group has
  bool done = i >= n;
  for (int it = 0; it < 64; ++it) {
    done = done || hi - lo <= fmax(abstol, 2.0 * eps * fmax(fabs(lo), fabs(hi)));
    if (__ballot(!done) == 0) break;
    const double step = (hi - lo) / (kBisLanes + 1);
    const double x = lo + step * (sub + 1);
    const bool above = !done && sturm_count(sD, sE2, n, x, pivmin) > i;
    const unsigned gm = static_cast<unsigned>(__ballot(above) >> (grp * kBisLanes)) & ((1u << kBisLanes) - 1u);
    const int l = gm ? __builtin_ctz(gm) : kBisLanes - 1;
    const double xl = __shfl(x, grp * kBisLanes + l, 64);
    const double xm = __shfl(x, grp * kBisLanes + (l > 0 ? l - 1 : 0), 64);
    if (!done) {
      if (gm == 0) {
        lo = xl;  // above the group's highest point
      } else {
        hi = xl;
        if (l > 0) lo = xm;
      }
    }
  }
  if (sub == 0 && i < n) T.w[i] = 0.5 * (lo + hi);
  static_assert(64 % kBisLanes == 0 && kBisLanes <= 32, "lane groups tile a wave");
}

// ---------------------------------------------------------------------------
// 3. Eigenvectors of T by inverse iteration (LAPACK dstein scheme).  A task
// is one lane: either one eigenvalue, or a cluster of eigenvalues that
// inverse iteration cannot separate (neighbours closer than kClusterGap * |T|,
// a few times the bisection's absolute accuracy; at most kMaxCluster
// members) run serially.  By default the members of a cluster are separate
// tasks and cluster_gs_kernel orthonormalises them afterwards.  For each
// member, from the largest: three solves of (T - lambda I) x = y from a
// seeded random start, each followed by Gram-Schmidt against the task's
// earlier members and normalisation.  For a task's first member the norm is
// summed by the back substitution and the scaling is applied as the next
// solve loads y, so no serial pass over the iterate remains besides the
// solves.  Each solve runs the LU factorisation with partial pivoting
// (dlagtf) fused with the forward elimination, then the back substitution.
//
// Eigenvalues further apart converge to vectors whose overlaps are about
// eps |T| / gap (at most ~0.03), which the orthonormalisation after this
// kernel removes; a vector's error is then a rotation among eigenvectors of
// close eigenvalues, which leaves the span of the kept set -- all the
// truncation uses -- unchanged to that order.  The split of a longer run of
// exactly degenerate eigenvalues into chunks gives chunks of random, hence
// well-conditioned, vectors of the same eigenspace.
//
// The recurrences are sequential along the matrix, so the parallelism is
// across lanes: all 64 lanes of a wave serve tasks of one matrix.  A
// lane's LU factors and iterate (4 n doubles) live in a global workspace
// interleaved by lane -- entry j of array a of lane l at (a n + j) 64 + l --
// so every access of the wave is one coalesced 512-byte line.  The forward
// sweep only stores; the back substitution loads entry j-1 before it uses
// entry j, so the load latency overlaps the division chain.
constexpr double kClusterGap = 32 * DBL_EPSILON;
constexpr int kMaxCluster = 8;

struct InvTask {
  const double* d;
  const double* e;
  const double* w;
  double* z;
  int n, ldz, c0, m;  // members c0 .. c0+m-1 (column c <-> w[n-1-c])
};

struct InvWave {
  int task0, ntask, n;
  int64_t ws;  // workspace offset (doubles) of the wave: 4 n x 64
};

__device__ inline double seeded_uniform(uint64_t x) {
  // splitmix64 -> [-1, 1)
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  x ^= x >> 31;
  return static_cast<double>(x >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

__global__ __launch_bounds__(64) void invit_kernel(const InvTask* __restrict__ tasks,
                                                   const InvWave* __restrict__ waves, double* __restrict__ work) {
  // the wave's matrix (one per wave): diagonal and off-diagonal in LDS
  __shared__ double sD[2048], sE[2048];
  const InvWave W = waves[blockIdx.x];
  const int lane = threadIdx.x;
  {
    const InvTask T0 = tasks[W.task0];
    for (int j = lane; j < T0.n; j += 64) {
      sD[j] = T0.d[j];
      sE[j] = j + 1 < T0.n ? T0.e[j] : 0.0;
    }
  }
  __syncthreads();
  if (lane >= W.ntask) return;
  const InvTask T = tasks[W.task0 + lane];
  const int n = T.n;
  double* ws = work + W.ws + lane;
  auto at = [&](int arr, int j) -> double& { return ws[(static_cast<size_t>(arr) * n + j) * 64]; };
  constexpr int U0 = 0, U1 = 1, U2 = 2, Y = 3;
  constexpr int kBlk = 8;  // entries loaded ahead of the recurrences
  const double tnorm = fmax(fabs(T.w[0]), fabs(T.w[n - 1]));
  const double tolp = fmax(DBL_EPSILON * tnorm, DBL_MIN);
  for (int a = 0; a < T.m; ++a) {
    const int c = T.c0 + a;
    const double lam = T.w[n - 1 - c];
    const uint64_t seed = (static_cast<uint64_t>(n) << 44) ^ (static_cast<uint64_t>(c) << 24);
    for (int j = 0; j < n; ++j) at(Y, j) = seeded_uniform(seed + j);
    double pend = 1.0;  // normalisation of y not yet applied (applied as y is loaded)
    for (int it = 0; it < 3; ++it) {
      // factorise T - lambda I = P L U and apply L^-1 P to y in one sweep;
      // y[j+1] of a block of steps is loaded before the block (step j
      // stores y[j] only, so the loaded entries are still the inputs)
      double aj = sD[0] - lam, bj = sE[0], cur = at(Y, 0) * pend;
      for (int j0 = 0; j0 + 1 < n; j0 += kBlk) {
        double yb[kBlk];
#pragma unroll
        for (int q = 0; q < kBlk; ++q) yb[q] = j0 + q + 1 < n ? at(Y, j0 + q + 1) * pend : 0.0;
#pragma unroll
        for (int q = 0; q < kBlk; ++q) {
          const int j = j0 + q;
          if (j + 1 >= n) break;
          const double cj = sE[j], an = sD[j + 1] - lam, bn = sE[j + 1];
          const double y1 = yb[q];
          double mm, top, bot;
          if (fabs(aj) >= fabs(cj)) {
            mm = aj != 0.0 ? cj / aj : 0.0;
            at(U0, j) = aj;
            at(U1, j) = bj;
            at(U2, j) = 0.0;
            top = cur;
            bot = y1;
            aj = an - mm * bj;
            bj = bn;
          } else {
            mm = aj / cj;
            at(U0, j) = cj;
            at(U1, j) = an;
            at(U2, j) = bn;
            top = y1;
            bot = cur;
            aj = bj - mm * an;
            bj = -mm * bn;
          }
          at(Y, j) = top;
          cur = bot - mm * top;
        }
      }
      at(U0, n - 1) = aj;
      at(Y, n - 1) = cur;
      // back substitution (tiny pivots replaced by +-eps |T|), rescaling when
      // the iterate grows large; a block of entries is loaded before it is
      // used (the rescaling touches only entries already solved); the
      // squared norm accumulates on the way
      double x1 = 0.0, x2 = 0.0, nrm = 0.0;
      for (int j1 = n - 1; j1 >= 0; j1 -= kBlk) {
        double ry[kBlk], r0[kBlk], r1[kBlk], r2[kBlk];
#pragma unroll
        for (int q = 0; q < kBlk; ++q) {
          const int j = j1 - q;
          ry[q] = j >= 0 ? at(Y, j) : 0.0;
          r0[q] = j >= 0 ? at(U0, j) : 1.0;
          r1[q] = (j >= 0 && j + 1 < n) ? at(U1, j) : 0.0;
          r2[q] = (j >= 0 && j + 2 < n) ? at(U2, j) : 0.0;
        }
#pragma unroll
        for (int q = 0; q < kBlk; ++q) {
          const int j = j1 - q;
          if (j < 0) break;
          double r = ry[q] - r1[q] * x1 - r2[q] * x2;
          double u = r0[q];
          if (fabs(u) < tolp) u = u < 0.0 ? -tolp : tolp;
          double x = r / u;
          if (fabs(x) > 1e150) {
            for (int i = j + 1; i < n; ++i) at(Y, i) *= 1e-150;
            x *= 1e-150;
            x1 *= 1e-150;
            x2 *= 1e-150;
            nrm *= 1e-300;
          }
          at(Y, j) = x;
          nrm = fma(x, x, nrm);
          x2 = x1;
          x1 = x;
        }
      }
      if (a > 0) {
        // Gram-Schmidt against the cluster's earlier members, then the norm
        for (int b = 0; b < a; ++b) {
          const double* zb = T.z + static_cast<size_t>(T.c0 + b) * T.ldz;
          double dt = 0.0;
          for (int j = 0; j < n; ++j) dt = fma(zb[j], at(Y, j), dt);
          for (int j = 0; j < n; ++j) at(Y, j) = fma(-dt, zb[j], at(Y, j));
        }
        nrm = 0.0;
        for (int j = 0; j < n; ++j) nrm = fma(at(Y, j), at(Y, j), nrm);
      }
      pend = 1.0 / sqrt(nrm);
    }
    double* zc = T.z + static_cast<size_t>(c) * T.ldz;
    for (int j0 = 0; j0 < n; j0 += kBlk) {
      double t[kBlk];
#pragma unroll
      for (int q = 0; q < kBlk; ++q) t[q] = j0 + q < n ? at(Y, j0 + q) : 0.0;
#pragma unroll
      for (int q = 0; q < kBlk; ++q)
        if (j0 + q < n) zc[j0 + q] = t[q] * pend;
    }
  }
}

// Orthonormalisation inside clusters: the members of a cluster run their
// inverse iterations in separate lanes (each from its own seeded start), so
// each converges into the cluster's eigenspace independently; afterwards the
// m member columns of the cluster are made orthonormal by two passes of
// modified Gram-Schmidt, one workgroup per cluster.  A cluster is a run of
// eigenvalues that inverse iteration cannot separate, so any orthonormal
// basis of its span is as good as another -- provided the members came out
// independent.  A member left with less than kGsKeep of its norm by the
// first pass (two iterates nearly parallel, so the remainder would carry
// their rounding errors magnified) flags the cluster, and the host runs that
// cluster again the serial way: one lane, each member orthogonalised against
// the earlier ones after every solve.
constexpr double kGsKeep = 1e-2;
bool g_cluster_serial = false;  // every cluster the serial way (process-wide)

struct ClusterGs {
  double* z;  // column c0 of the cluster; column j at z + j * ldz
  int n, ldz, m, pad;
};

__device__ inline double block_sum256(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  __syncthreads();  // red is free again
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  return red[0] + red[1] + red[2] + red[3];
}

__global__ __launch_bounds__(256) void cluster_gs_kernel(const ClusterGs* __restrict__ cl, int* __restrict__ dep) {
  __shared__ double red[4];
  const ClusterGs C = cl[blockIdx.x];
  int flag = 0;
  for (int pass = 0; pass < 2; ++pass)
    for (int a = 0; a < C.m; ++a) {
      double* za = C.z + static_cast<int64_t>(a) * C.ldz;
      double q0 = 0.0;
      if (pass == 0) {
        double q = 0.0;
        for (int j = threadIdx.x; j < C.n; j += 256) q = fma(za[j], za[j], q);
        q0 = block_sum256(q, red);
      }
      for (int b = 0; b < a; ++b) {
        const double* zb = C.z + static_cast<int64_t>(b) * C.ldz;
        double p = 0.0;
        for (int j = threadIdx.x; j < C.n; j += 256) p = fma(za[j], zb[j], p);
        const double dt = block_sum256(p, red);
        for (int j = threadIdx.x; j < C.n; j += 256) za[j] = fma(-dt, zb[j], za[j]);
        __syncthreads();
      }
      double q = 0.0;
      for (int j = threadIdx.x; j < C.n; j += 256) q = fma(za[j], za[j], q);
      const double nrm = block_sum256(q, red);
      if (pass == 0 && !(nrm > kGsKeep * kGsKeep * q0)) flag = 1;
      const double sc = nrm > 0.0 ? 1.0 / sqrt(nrm) : 0.0;
      for (int j = threadIdx.x; j < C.n; j += 256) za[j] *= sc;
      __syncthreads();
    }
  if (threadIdx.x == 0) dep[blockIdx.x] = flag;
}

// B = 1.5 I - 0.5 M on every k x k slot (M = Z^T Z): one Newton-Schulz step
// Z <- Z B towards the orthonormal polar factor of Z.
__global__ __launch_bounds__(256) void ns_coef_kernel(const double* __restrict__ M, double* __restrict__ B, int k,
                                                      int64_t total) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int64_t r = i % (int64_t(k) * k);
  const double diag = (r % k) == (r / k) ? 1.5 : 0.0;
  B[i] = diag - 0.5 * M[i];
}

// max over every slot's leading k_s x k_s block of |M - I|: blockIdx.y is
// the slot, blockIdx.x a range of kOrthoCols columns.  Non-negative doubles
// order like their bit patterns, so the workgroups' maxima meet in an integer
// atomic max on out[s] (cleared to zero before the launch).
constexpr int kOrthoCols = 16;
__global__ __launch_bounds__(256) void ortho_err_kernel(const double* __restrict__ M, const int* __restrict__ ks,
                                                        int k, unsigned long long* __restrict__ out) {
  const int s = blockIdx.y;
  const int kk = ks[s];
  const int c0 = blockIdx.x * kOrthoCols;
  if (c0 >= kk) return;
  const int c1 = min(kk, c0 + kOrthoCols);
  const double* Ms = M + static_cast<size_t>(s) * k * k;
  double m = 0.0;
  for (int c = c0; c < c1; ++c)
    for (int r = threadIdx.x; r < kk; r += 256)
      m = fmax(m, fabs(Ms[r + static_cast<size_t>(c) * k] - (r == c ? 1.0 : 0.0)));
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  __shared__ double red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0)
    atomicMax(out + s, static_cast<unsigned long long>(
                           __double_as_longlong(fmax(fmax(red[0], red[1]), fmax(red[2], red[3])))));
}

// ---------------------------------------------------------------------------
// 4. Back-transformation Z <- H_0 H_1 ... H_{n-2} Z, the reflectors applied
// from the last to the first.  A workgroup of kBtWaves waves serves up to
// two columns per wave, held in registers (lane l keeps rows l + 64 m,
// m < NT).  The reflectors pass through LDS in blocks of kBtBlkOf<NT>,
// double-buffered: the next block's global loads are issued before the
// current block is applied and stored after it, one barrier per block, so
// every reflector is read from memory once per workgroup.
constexpr int kBtWaves = 8;
constexpr int kBtThreads = 64 * kBtWaves;
// the longest rows (NT > 24) take one column per wave and blocks of two
// reflectors, so their registers fit
template <int NT>
constexpr int kBtBlkOf = NT > 24 ? 2 : 4;
template <int NT>
constexpr int kBtColsPerWave = NT > 24 ? 1 : 2;
template <int NT>
constexpr int kBtColsOf = kBtColsPerWave<NT> * kBtWaves;

struct BtTask {
  const double* a;
  const double* tau;
  double* z;
  int n, lda, ldz, k, blk0, nblk;
};

template <int NT>
__global__ __launch_bounds__(kBtThreads) void bt_kernel(const BtTask* __restrict__ tasks, int ntask) {
  extern __shared__ double sv[];  // 2 buffers x kBtBlkOf<NT> reflectors x 64 NT entries
  constexpr int kBtBlk = kBtBlkOf<NT>;
  __shared__ double stau[2][kBtBlk];
  int t = 0;
  while (t < ntask && static_cast<int>(blockIdx.x) >= tasks[t].blk0 + tasks[t].nblk) ++t;
  if (t >= ntask) return;
  const BtTask T = tasks[t];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  constexpr int L = 64 * NT;  // LDS stride of one reflector
  const int n = T.n;
  constexpr bool kTwo = kBtColsPerWave<NT> == 2;
  const int c0 = (static_cast<int>(blockIdx.x) - T.blk0) * kBtColsOf<NT> + kBtColsPerWave<NT> * wv;
  const bool one = c0 < T.k, two = kTwo && c0 + 1 < T.k;  // wave-uniform
  double z0[NT], z1[kTwo ? NT : 1];
  const double* Z0 = T.z + static_cast<size_t>(c0) * T.ldz;
  const double* Z1 = Z0 + T.ldz;
#pragma unroll
  for (int m = 0; m < NT; ++m) {
    const int j = lane + 64 * m;
    z0[m] = (one && j < n) ? Z0[j] : 0.0;
    if constexpr (kTwo) z1[m] = (two && j < n) ? Z1[j] : 0.0;
  }
  // block b holds reflectors hi(b), hi(b)-1, ..., at most kBtBlk, hi(b) = n-2-b kBtBlk
  const int nref = n - 1;
  const int nblocks = nref > 0 ? (nref + kBtBlk - 1) / kBtBlk : 0;
  constexpr int kPf = (kBtBlk * L + kBtThreads - 1) / kBtThreads;  // entries per thread per block
  double pf[kPf];
  auto fetch = [&](int b) {
    const int hi = n - 2 - b * kBtBlk;
#pragma unroll
    for (int p = 0; p < kPf; ++p) {
      const int e = tid + kBtThreads * p;
      const int r = e / L, j = e - r * L;
      const int i = hi - r;
      double x = 0.0;
      if (r < kBtBlk && i >= 0) x = j == i + 1 ? 1.0 : ((j > i + 1 && j < n) ? T.a[j + static_cast<size_t>(i) * T.lda] : 0.0);
      pf[p] = x;
    }
  };
  auto stash = [&](int b, int buf) {
    const int hi = n - 2 - b * kBtBlk;
#pragma unroll
    for (int p = 0; p < kPf; ++p) {
      const int e = tid + kBtThreads * p;
      if (e < kBtBlk * L) sv[buf * kBtBlk * L + e] = pf[p];
    }
    if (tid < kBtBlk) stau[buf][tid] = hi - tid >= 0 ? T.tau[hi - tid] : 0.0;
  };
  if (nblocks > 0) {
    fetch(0);
    stash(0, 0);
  }
  __syncthreads();
  for (int b = 0; b < nblocks; ++b) {
    const int buf = b & 1;
    if (b + 1 < nblocks) fetch(b + 1);
    if (one) {
      const double* vb = sv + buf * kBtBlk * L;
#pragma unroll 1
      for (int r = 0; r < kBtBlk; ++r) {
        const double ti = stau[buf][r];
        if (ti == 0.0) continue;  // also reflectors past the first
        const double* vv = vb + r * L;
        double s0 = 0.0, s1 = 0.0;
#pragma unroll
        for (int m = 0; m < NT; ++m) {
          const double x = vv[lane + 64 * m];
          s0 = fma(x, z0[m], s0);
          if constexpr (kTwo) s1 = fma(x, z1[m], s1);
        }
        s0 = wave_sum(s0) * ti;
        if constexpr (kTwo) s1 = wave_sum(s1) * ti;
#pragma unroll
        for (int m = 0; m < NT; ++m) {
          const double x = vv[lane + 64 * m];
          z0[m] = fma(-s0, x, z0[m]);
          if constexpr (kTwo) z1[m] = fma(-s1, x, z1[m]);
        }
      }
    }
    if (b + 1 < nblocks) stash(b + 1, buf ^ 1);
    __syncthreads();
  }
  if (!one) return;
  double* W0 = T.z + static_cast<size_t>(c0) * T.ldz;
  double* W1 = W0 + T.ldz;
#pragma unroll
  for (int m = 0; m < NT; ++m) {
    const int j = lane + 64 * m;
    if (j < n) {
      W0[j] = z0[m];
      if constexpr (kTwo)
        if (two) W1[j] = z1[m];
    }
  }
}

template <int NT>
hipError_t launch_bt(const BtTask* tasks, int ntask, int nblk, hipStream_t st) {
  constexpr int lds = 2 * kBtBlkOf<NT> * 64 * NT * 8;
  static bool attr = false;
  if (!attr) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(bt_kernel<NT>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    if (e != hipSuccess) return e;
    attr = true;
  }
  hipLaunchKernelGGL(bt_kernel<NT>, dim3(nblk), dim3(kBtThreads), lds, st, tasks, ntask);
  return hipGetLastError();
}

std::atomic<long long> g_cholqr_fallbacks{0};
EigProfile g_prof;

double secs_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

int device_cus() {
  static int cus = [] {
    int dev = 0;
    hipDeviceProp_t p;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&p, dev) != hipSuccess) return 0;
    return p.multiProcessorCount;
  }();
  return cus;
}

// Wall-clock ticks of the tridiagonalisation's longest wait (a quarter second).
long long wait_ticks() {
  static long long t = [] {
    int dev = 0, khz = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0)
      khz = 100000;
    return static_cast<long long>(khz) * 250;
  }();
  return t;
}

}  // namespace

long long SymEigBatch::cholqr_fallbacks() { return g_cholqr_fallbacks; }
EigProfile& SymEigBatch::profile() { return g_prof; }
void SymEigBatch::set_cluster_serial(bool serial) { g_cluster_serial = serial; }
bool SymEigBatch::cluster_serial() { return g_cluster_serial; }

int SymEigBatch::max_order() {
  // the largest order whose rows fit the LDS of at most (CUs - kCuMargin)
  // co-resident workgroups next to three vectors each, and at most 64 * 32
  // rows in the back-transformation's registers
  static const int m = [] {
    const int cus = device_cus() - kCuMargin;
    for (int n = 2048; n > 1; --n) {
      const int ldr = (n + 63) / 64 * 64;
      const int rmax = (kTrdLdsDoubles - 3 * ldr) / ldr;
      const int cap = rmax + (trd_template(n) > kTrdMaxRegNT ? 0 : kTrdWaves);
      if (rmax >= 1 && (n + cap - 1) / cap <= cus) return n;
    }
    return 1;
  }();
  return m;
}

namespace {

// Device memory of phase 1, one region: d, e, tau (3 n doubles per job, at
// offset doff[q] from its start), the publication words of the groups (8 n
// per job) and the error flag (both cleared by a fill kernel before the
// launches), then the task tables (one TrdTask and one BisTask per
// non-empty job).  Parts start on 256-byte boundaries.
struct Phase1Layout {
  std::vector<int64_t> doff;  // per job, in doubles; doff[nj] = total
  int64_t pubs = 0, err = 0, tasks = 0, bis = 0, total = 0;  // byte offsets / size
};

int64_t align256(int64_t b) { return (b + 255) / 256 * 256; }

Phase1Layout phase1_layout(const std::vector<EigJob>& jobs) {
  Phase1Layout L;
  const int nj = static_cast<int>(jobs.size());
  L.doff.assign(nj + 1, 0);
  int64_t nonempty = 0;
  for (int q = 0; q < nj; ++q) {
    require(jobs[q].n >= 0 && jobs[q].n <= SymEigBatch::max_order(), "eigensolver: matrix order out of range");
    require(jobs[q].n == 0 || (jobs[q].a && jobs[q].w && jobs[q].lda >= jobs[q].n), "eigensolver: bad job");
    L.doff[q + 1] = L.doff[q] + 3 * int64_t(jobs[q].n);
    nonempty += jobs[q].n > 0;
  }
  L.pubs = align256(L.doff[nj] * 8);
  L.err = L.pubs + (L.doff[nj] / 3) * 8 * 8;  // 8 words per row
  L.tasks = align256(L.err + 8);
  L.bis = align256(L.tasks + nonempty * int64_t(sizeof(TrdTask)));
  L.total = align256(L.bis + nonempty * int64_t(sizeof(BisTask)));
  return L;
}

// Phase 1 on the region `mem` laid out by L: tridiagonalisations in
// cooperative launches of at most (CUs - kCuMargin) workgroups, then
// bisection.  Synchronises the stream before returning.
void phase1(Context* ctx, const std::vector<EigJob>& jobs, const Phase1Layout& L, char* mem) {
  const int nj = static_cast<int>(jobs.size());
  auto t0 = std::chrono::steady_clock::now();
  double* base = reinterpret_cast<double*>(mem);
  uint64_t* pubs = reinterpret_cast<uint64_t*>(mem + L.pubs);
  int* err = reinterpret_cast<int*>(mem + L.err);
  TrdTask* dtasks = reinterpret_cast<TrdTask*>(mem + L.tasks);
  BisTask* dbis = reinterpret_cast<BisTask*>(mem + L.bis);
  // publication words and the error flag: zero bits
  SADMRG_HIP(launch_fill((L.err + 8 - L.pubs) / 8, 0.0, reinterpret_cast<double*>(pubs), ctx->stream()));
  // workgroups per matrix: as many rows per workgroup as LDS holds next to
  // three vectors, plus one register-held row per wave where the kernel
  // instantiation has them; rows spread evenly, LDS first
  struct Plan {
    int q, W, R, rreg;
    bool big;  // rows longer than the register-row instantiations take
  };
  std::vector<Plan> plans;
  for (int q = 0; q < nj; ++q) {
    const int n = jobs[q].n;
    if (n == 0) continue;
    const int ldr = (n + 63) / 64 * 64;
    const int rmax = (kTrdLdsDoubles - 3 * ldr) / ldr;
    require(rmax >= 1, "eigensolver: matrix too large for the LDS-resident reduction");
    const bool big = trd_template(n) > kTrdMaxRegNT;
    const int reg = big ? 0 : kTrdWaves;
    const int W = (n + rmax + reg - 1) / (rmax + reg);
    const int Rt = (n + W - 1) / W;
    const int R = std::min(Rt, rmax);
    plans.push_back(Plan{q, W, R, Rt - R, big});
  }
  // launches hold one kind of rows only (the instantiation decides whether
  // register rows exist); the widest matrices first
  std::stable_sort(plans.begin(), plans.end(),
                   [](const Plan& x, const Plan& y) { return x.big != y.big ? x.big : x.W > y.W; });
  // cooperative launches stay a margin below the CU count, so a grid never
  // depends on every CU being free at once
  const int cus = device_cus() - kCuMargin;
  require(cus > 0, "eigensolver: no device properties");
  const long long wait = wait_ticks();
  // every launch's tasks and the bisection tasks in one host table, one copy
  struct Launch {
    int t0, ntask, wg, lds, nt;
  };
  std::vector<Launch> launches;
  std::vector<char> host(L.total - L.tasks, 0);
  TrdTask* htasks = reinterpret_cast<TrdTask*>(host.data());
  BisTask* hbis = reinterpret_cast<BisTask*>(host.data() + (L.bis - L.tasks));
  // first-fit decreasing: each plan, widest first, joins the first launch of
  // its kind with room, or opens one
  std::vector<std::vector<size_t>> bins;
  std::vector<int> bin_wg;
  for (size_t pi = 0; pi < plans.size(); ++pi) {
    const Plan& P = plans[pi];
    require(P.W <= cus, "eigensolver: matrix needs more workgroups than the device holds");
    size_t b = 0;
    while (b < bins.size() && !(plans[bins[b][0]].big == P.big && bin_wg[b] + P.W <= cus)) ++b;
    if (b == bins.size()) {
      bins.emplace_back();
      bin_wg.push_back(0);
    }
    bins[b].push_back(pi);
    bin_wg[b] += P.W;
  }
  int nt_all = 0;
  for (const std::vector<size_t>& bin : bins) {
    Launch l{nt_all, 0, 0, 0, 0};
    for (size_t pi : bin) {
      const Plan& P = plans[pi];
      const EigJob& J = jobs[P.q];
      TrdTask t;
      t.a = J.a;
      t.d = base + L.doff[P.q];
      t.e = t.d + J.n;
      t.tau = t.e + J.n;
      t.pub = pubs + (L.doff[P.q] / 3) * 8;
      t.n = J.n;
      t.lda = J.lda;
      t.wg0 = l.wg;
      t.nwg = P.W;
      t.rows = P.R;
      t.rreg = P.rreg;
      t.ldr = (J.n + 63) / 64 * 64;
      std::memcpy(htasks + nt_all, &t, sizeof(TrdTask));
      ++nt_all;
      ++l.ntask;
      l.wg += P.W;
      l.lds = std::max(l.lds, (P.R + 3) * t.ldr * 8);
      l.nt = std::max(l.nt, (J.n + 63) / 64);
    }
    launches.push_back(l);
  }
  int total = 0, nb = 0;  // total: workgroups of the bisection (kBisPerBlock eigenvalues each)
  for (int q = 0; q < nj; ++q) {
    if (jobs[q].n == 0) continue;
    if (jobs[q].n > kBisOrder) fail(kErrInvalid, "eigensolver: matrix order beyond the bisection's LDS copy");
    BisTask b;
    b.d = base + L.doff[q];
    b.e = b.d + jobs[q].n;
    b.w = jobs[q].w;
    b.n = jobs[q].n;
    b.off = total;
    std::memcpy(hbis + nb, &b, sizeof(BisTask));
    ++nb;
    total += (jobs[q].n + kBisPerBlock - 1) / kBisPerBlock;
  }
  if (total == 0) return;
  SADMRG_HIP(hipMemcpyAsync(mem + L.tasks, host.data(), host.size(), hipMemcpyHostToDevice, ctx->stream()));
  for (const Launch& l : launches) {
    const TrdTask* tp = dtasks + l.t0;
    hipError_t le;
    if (l.nt <= 4) le = launch_trd<4>(tp, l.ntask, err, wait, l.wg, l.lds, ctx->stream());
    else if (l.nt <= 8) le = launch_trd<8>(tp, l.ntask, err, wait, l.wg, l.lds, ctx->stream());
    else if (l.nt <= 12) le = launch_trd<12>(tp, l.ntask, err, wait, l.wg, l.lds, ctx->stream());
    else if (l.nt <= 16) le = launch_trd<16>(tp, l.ntask, err, wait, l.wg, l.lds, ctx->stream());
    else if (l.nt <= 20) le = launch_trd<20>(tp, l.ntask, err, wait, l.wg, l.lds, ctx->stream());
    else if (l.nt <= 24) le = launch_trd<24>(tp, l.ntask, err, wait, l.wg, l.lds, ctx->stream());
    else le = launch_trd<32>(tp, l.ntask, err, wait, l.wg, l.lds, ctx->stream());
    SADMRG_HIP(le);
    g_prof.launches++;
  }
  int herr = 0;
  SADMRG_HIP(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, ctx->stream()));
  ctx->sync();  // also: the host task table may leave scope
  if (herr) fail(kErrNumerics, "eigensolver: a tridiagonalisation step did not receive its group's data");
  g_prof.trd += secs_since(t0);
  t0 = std::chrono::steady_clock::now();
  g_prof.matrices += nj;
  for (const EigJob& J : jobs) g_prof.max_order = std::max<int64_t>(g_prof.max_order, J.n);
  hipLaunchKernelGGL(bisect_kernel, dim3(total), dim3(256), 0, ctx->stream(), dbis, nb);
  SADMRG_HIP(hipGetLastError());
  ctx->sync();
  g_prof.bisect += secs_since(t0);
}

}  // namespace

namespace {
std::mutex g_scratch_mu;
std::vector<char*> g_scratch_ptr;  // per device: address of `scratch`, null until looked up
}  // namespace

void noalloc_scratch_setup(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (static_cast<int>(g_scratch_ptr.size()) <= device) g_scratch_ptr.resize(device + 1, nullptr);
  if (g_scratch_ptr[device]) return;
  void* p = nullptr;
  SADMRG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  require(p != nullptr, "eigensolver: no device address of the static scratch");
  g_scratch_ptr[device] = static_cast<char*>(p);
}

void SymEigBatch::eigenvalues_noalloc(Context* ctx, const std::vector<EigJob>& jobs) {
  const Phase1Layout L = phase1_layout(jobs);
  if (L.total > kNoAllocBytes)
    fail(kErrInvalid, "eigvalsh (no allocation): the batch needs " + std::to_string(L.total) +
                          " bytes of scratch, more than the " + std::to_string(kNoAllocBytes) +
                          " held per device (88 bytes per matrix row, 96 per matrix)");
  char* p = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (ctx->device() < static_cast<int>(g_scratch_ptr.size())) p = g_scratch_ptr[ctx->device()];
  }
  require(p != nullptr, "eigvalsh (no allocation): scratch of this device not set up");
  phase1(ctx, jobs, L, p);
}

void SymEigBatch::eigenvalues(const std::vector<EigJob>& jobs) {
  const Phase1Layout L = phase1_layout(jobs);
  doff_ = L.doff;
  trd_ = DevBuf(ctx_, std::max<int64_t>(L.total, 256));
  phase1(ctx_, jobs, L, static_cast<char*>(trd_.get()));
}

void SymEigBatch::top_vectors(const std::vector<EigJob>& jobs) {
  Context* ctx = ctx_;
  const int nj = static_cast<int>(jobs.size());
  require(static_cast<int>(doff_.size()) == nj + 1, "eigensolver: top_vectors before eigenvalues");
  double* base = trd_.as<double>();
  // the jobs asking for vectors, in one padded batch: slot s holds an
  // nmax x kmax block, rows >= n and columns >= k zero
  std::vector<int> sel;
  int nmax = 0, kmax = 0;
  for (int q = 0; q < nj; ++q) {
    const EigJob& J = jobs[q];
    if (J.n == 0 || J.k <= 0) continue;
    require(J.k <= J.n && J.z && J.ldz >= J.n, "eigensolver: bad eigenvector request");
    sel.push_back(q);
    nmax = std::max(nmax, J.n);
    kmax = std::max(kmax, J.k);
  }
  const int ns = static_cast<int>(sel.size());
  if (ns == 0) return;
  auto t0 = std::chrono::steady_clock::now();
  const int64_t slot = int64_t(nmax) * kmax;
  DevBuf zp(ctx, slot * ns * 8), gm(ctx, int64_t(kmax) * kmax * ns * 8), dinfo(ctx, sizeof(int) * 2 * ns);
  SADMRG_HIP(launch_fill(slot * ns, 0.0, zp.as<double>(), ctx->stream()));
  SADMRG_HIP(hipMemsetAsync(dinfo.get(), 0, sizeof(int) * 2 * ns, ctx->stream()));
  // clusters of eigenvalues closer than kClusterGap |T| (host copy of w)
  std::vector<InvTask> it;
  std::vector<ClusterGs> cgs;
  std::vector<InvTask> serial;  // the clusters of cgs as single serial tasks
  std::vector<std::vector<double>> hw(ns);
  for (int s = 0; s < ns; ++s) {
    const EigJob& J = jobs[sel[s]];
    hw[s].resize(J.n);
    SADMRG_HIP(hipMemcpyAsync(hw[s].data(), J.w, J.n * 8, hipMemcpyDeviceToHost, ctx->stream()));
  }
  ctx->sync();
  for (int s = 0; s < ns; ++s) {
    const EigJob& J = jobs[sel[s]];
    const std::vector<double>& w = hw[s];
    const double tn = std::max(std::fabs(w[0]), std::fabs(w[J.n - 1]));
    InvTask v;
    v.d = base + doff_[sel[s]];
    v.e = v.d + J.n;
    v.w = J.w;
    v.z = zp.as<double>() + slot * s;
    v.n = J.n;
    v.ldz = nmax;
    for (int c = 0; c < J.k;) {
      int m = 1;
      while (c + m < J.k && m < kMaxCluster && w[J.n - 1 - (c + m - 1)] - w[J.n - 1 - (c + m)] <= kClusterGap * tn)
        ++m;
      if (m > 1 && g_cluster_serial) {
        v.c0 = c;
        v.m = m;
        it.push_back(v);
        g_prof.max_cluster = std::max<int64_t>(g_prof.max_cluster, m);
        g_prof.vectors += m;
        g_prof.clusters++;
        c += m;
        continue;
      }
      // one lane per member: a cluster's members iterate in parallel and
      // are made orthonormal afterwards (cluster_gs_kernel)
      for (int a = 0; a < m; ++a) {
        v.c0 = c + a;
        v.m = 1;
        it.push_back(v);
      }
      if (m > 1) {
        cgs.push_back(ClusterGs{v.z + static_cast<int64_t>(c) * v.ldz, v.n, v.ldz, m, 0});
        serial.push_back(v);
        serial.back().c0 = c;
        serial.back().m = m;
      }
      g_prof.max_cluster = std::max<int64_t>(g_prof.max_cluster, m);
      g_prof.vectors += m;
      g_prof.clusters++;
      c += m;
    }
  }
  // waves of up to 64 lanes (tasks) of one matrix, each wave with a
  // workspace of 4 n x 64 doubles; the launch is left in flight
  auto run_invit = [&](const std::vector<InvTask>& tasks, std::vector<DevBuf>& keep) {
    std::vector<InvWave> wv;
    int64_t ws = 0;
    for (size_t t0 = 0; t0 < tasks.size();) {
      const int n = tasks[t0].n;
      int cnt = 0;
      while (t0 + cnt < tasks.size() && cnt < 64 && tasks[t0 + cnt].z == tasks[t0].z) ++cnt;
      wv.push_back(InvWave{static_cast<int>(t0), cnt, n, ws});
      ws += 4 * int64_t(n) * 64;
      t0 += cnt;
    }
    DevBuf dt(ctx, tasks.size() * sizeof(InvTask)), dw(ctx, wv.size() * sizeof(InvWave)), work(ctx, ws * 8);
    SADMRG_HIP(hipMemcpyAsync(dt.get(), tasks.data(), tasks.size() * sizeof(InvTask), hipMemcpyHostToDevice,
                              ctx->stream()));
    SADMRG_HIP(
        hipMemcpyAsync(dw.get(), wv.data(), wv.size() * sizeof(InvWave), hipMemcpyHostToDevice, ctx->stream()));
    hipLaunchKernelGGL(invit_kernel, dim3(wv.size()), dim3(64), 0, ctx->stream(), dt.as<InvTask>(), dw.as<InvWave>(),
                       work.as<double>());
    SADMRG_HIP(hipGetLastError());
    keep.push_back(std::move(dt));
    keep.push_back(std::move(dw));
    keep.push_back(std::move(work));
  };
  {
    std::vector<DevBuf> keep;
    const auto tk = std::chrono::steady_clock::now();
    run_invit(it, keep);
    if (!cgs.empty()) {
      const int nc = static_cast<int>(cgs.size());
      DevBuf dc(ctx, cgs.size() * sizeof(ClusterGs)), dd(ctx, sizeof(int) * nc);
      SADMRG_HIP(
          hipMemcpyAsync(dc.get(), cgs.data(), cgs.size() * sizeof(ClusterGs), hipMemcpyHostToDevice, ctx->stream()));
      hipLaunchKernelGGL(cluster_gs_kernel, dim3(nc), dim3(256), 0, ctx->stream(), dc.as<ClusterGs>(), dd.as<int>());
      SADMRG_HIP(hipGetLastError());
      std::vector<int> dep(nc);
      SADMRG_HIP(hipMemcpyAsync(dep.data(), dd.get(), sizeof(int) * nc, hipMemcpyDeviceToHost, ctx->stream()));
      ctx->sync();
      std::vector<InvTask> redo;
      for (int q = 0; q < nc; ++q)
        if (dep[q]) redo.push_back(serial[q]);
      if (!redo.empty()) {
        g_prof.serial_clusters += static_cast<int64_t>(redo.size());
        run_invit(redo, keep);
      }
    }
    ctx->sync();
    g_prof.invit_kernel += secs_since(tk);
  }
  g_prof.invit += secs_since(t0);
  t0 = std::chrono::steady_clock::now();
  // Newton-Schulz steps Z <- Z (3 I - Z^T Z) / 2 over the whole batch (the
  // zero padding columns stay zero) until every slot is orthonormal to
  // kOrthoTol; Cholesky QR twice if that does not happen within kNsMax steps
  rocblas_handle h = static_cast<rocblas_handle>(ctx->solver());
  const double one = 1.0, zero = 0.0;
  const int64_t sg = int64_t(kmax) * kmax;
  constexpr double kOrthoTol = 1e-13;
  constexpr int kNsMax = 6;
  DevBuf zq(ctx, slot * ns * 8), bm(ctx, sg * ns * 8), derr(ctx, ns * 8), dks(ctx, sizeof(int) * ns);
  {
    std::vector<int> hk(ns);
    for (int s = 0; s < ns; ++s) hk[s] = jobs[sel[s]].k;
    SADMRG_HIP(hipMemcpyAsync(dks.get(), hk.data(), sizeof(int) * ns, hipMemcpyHostToDevice, ctx->stream()));
  }
  double* zcur = zp.as<double>();
  double* znext = zq.as<double>();
  bool ortho = false;
  for (int step = 0; step <= kNsMax && !ortho; ++step) {
    if (rocblas_dgemm_strided_batched(h, rocblas_operation_transpose, rocblas_operation_none, kmax, kmax, nmax, &one,
                                      zcur, nmax, slot, zcur, nmax, slot, &zero, gm.as<double>(), kmax, sg, ns) !=
        rocblas_status_success)
      fail(kErrNumerics, "eigensolver: rocblas gemm failed");
    SADMRG_HIP(launch_fill(ns, 0.0, derr.as<double>(), ctx->stream()));
    hipLaunchKernelGGL(ortho_err_kernel, dim3((kmax + kOrthoCols - 1) / kOrthoCols, ns), dim3(256), 0, ctx->stream(),
                       gm.as<double>(), dks.as<int>(), kmax, derr.as<unsigned long long>());
    SADMRG_HIP(hipGetLastError());
    std::vector<double> herr(ns);
    SADMRG_HIP(hipMemcpyAsync(herr.data(), derr.get(), ns * 8, hipMemcpyDeviceToHost, ctx->stream()));
    ctx->sync();
    ortho = *std::max_element(herr.begin(), herr.end()) <= kOrthoTol;
    if (ortho || step == kNsMax) break;
    hipLaunchKernelGGL(ns_coef_kernel, dim3((sg * ns + 255) / 256), dim3(256), 0, ctx->stream(), gm.as<double>(),
                       bm.as<double>(), kmax, sg * ns);
    SADMRG_HIP(hipGetLastError());
    if (rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, nmax, kmax, kmax, &one, zcur,
                                      nmax, slot, bm.as<double>(), kmax, sg, &zero, znext, nmax, slot, ns) !=
        rocblas_status_success)
      fail(kErrNumerics, "eigensolver: rocblas gemm failed");
    std::swap(zcur, znext);
    g_prof.ns_steps++;
  }
  if (!ortho) {
    for (int pass = 0; pass < 2; ++pass) {
      if (rocblas_dsyrk_strided_batched(h, rocblas_fill_lower, rocblas_operation_transpose, kmax, nmax, &one, zcur,
                                        nmax, slot, &zero, gm.as<double>(), kmax, sg, ns) != rocblas_status_success)
        fail(kErrNumerics, "eigensolver: rocblas syrk failed");
      for (int s = 0; s < ns; ++s) {
        const int k = jobs[sel[s]].k;
        if (k < kmax)
          SADMRG_HIP(launch_strided_fill(kmax - k, 1.0, gm.as<double>() + sg * s + int64_t(k) * (kmax + 1),
                                         kmax + 1, ctx->stream()));
      }
      if (rocsolver_dpotrf_strided_batched(h, rocblas_fill_lower, kmax, gm.as<double>(), kmax, sg,
                                           dinfo.as<int>() + pass * ns, ns) != rocblas_status_success)
        fail(kErrNumerics, "eigensolver: rocsolver potrf failed");
      if (rocblas_dtrsm_strided_batched(h, rocblas_side_right, rocblas_fill_lower, rocblas_operation_transpose,
                                        rocblas_diagonal_non_unit, nmax, kmax, &one, gm.as<double>(), kmax, sg, zcur,
                                        nmax, slot, ns) != rocblas_status_success)
        fail(kErrNumerics, "eigensolver: rocblas trsm failed");
    }
  }
  g_cholqr_fallbacks += ortho ? 0 : 1;
  std::vector<int> hinfo(2 * ns);
  SADMRG_HIP(hipMemcpyAsync(hinfo.data(), dinfo.get(), sizeof(int) * hinfo.size(), hipMemcpyDeviceToHost,
                            ctx->stream()));
  // back-transformation on the padded slots, then the k columns to each job
  std::vector<BtTask> bts;
  int blk = 0;
  // columns per workgroup of the instantiation the launch takes (NT of nmax)
  const int bt_cols = (trd_template(nmax) > kTrdMaxRegNT ? 1 : 2) * kBtWaves;
  static_assert(kBtColsPerWave<24> == 2 && kBtColsPerWave<32> == 1, "columns per wave follow the row length");
  for (int s = 0; s < ns; ++s) {
    const EigJob& J = jobs[sel[s]];
    BtTask b;
    b.a = J.a;
    b.tau = base + doff_[sel[s]] + 2 * int64_t(J.n);
    b.z = zcur + slot * s;
    b.n = J.n;
    b.lda = J.lda;
    b.ldz = nmax;
    b.k = J.k;
    b.blk0 = blk;
    b.nblk = (J.k + bt_cols - 1) / bt_cols;
    bts.push_back(b);
    blk += b.nblk;
  }
  DevBuf dbt(ctx, bts.size() * sizeof(BtTask));
  SADMRG_HIP(hipMemcpyAsync(dbt.get(), bts.data(), bts.size() * sizeof(BtTask), hipMemcpyHostToDevice, ctx->stream()));
  ctx->sync();
  for (int x : hinfo)
    if (x != 0) fail(kErrNumerics, "eigensolver: eigenvectors not independent (Cholesky QR failed)");
  g_prof.ortho += secs_since(t0);
  t0 = std::chrono::steady_clock::now();
  const int nt = (nmax + 63) / 64;
  const BtTask* tp = dbt.as<BtTask>();
  const int ntask = static_cast<int>(bts.size());
  hipError_t e;
  if (nt <= 4) e = launch_bt<4>(tp, ntask, blk, ctx->stream());
  else if (nt <= 8) e = launch_bt<8>(tp, ntask, blk, ctx->stream());
  else if (nt <= 12) e = launch_bt<12>(tp, ntask, blk, ctx->stream());
  else if (nt <= 16) e = launch_bt<16>(tp, ntask, blk, ctx->stream());
  else if (nt <= 20) e = launch_bt<20>(tp, ntask, blk, ctx->stream());
  else if (nt <= 24) e = launch_bt<24>(tp, ntask, blk, ctx->stream());
  else e = launch_bt<32>(tp, ntask, blk, ctx->stream());
  SADMRG_HIP(e);
  for (int s = 0; s < ns; ++s) {
    const EigJob& J = jobs[sel[s]];
    SADMRG_HIP(launch_copy2d(J.n, J.k, zcur + slot * s, 1, nmax, J.z, J.ldz, 1, 1.0, nullptr, 0, ctx->stream()));
  }
  ctx->sync();
  g_prof.backt += secs_since(t0);
}

// ---- small-sector SVD: one-sided Jacobi in LDS ------------------------------
bool jacobi_svd_fits(int m, int n, int max_order) {
  const int q = std::min(m, n), p = std::max(m, n);
  const int qe = q + (q & 1);
  return q >= 1 && q <= std::min(max_order, kJacobiMaxOrder) &&
         int64_t(p) * qe + int64_t(qe) * qe <= kJacobiLdsWords;
}

__global__ __launch_bounds__(256) void jacobi_svd_kernel(const JacobiSvdJob* __restrict__ jobs, int max_sweeps,
                                                         int* __restrict__ flags) {
  constexpr int NT = 256;
  extern __shared__ double lds[];
  __shared__ int s_rot[2];
  __shared__ double s_sig[kJacobiMaxOrder + 8];
  __shared__ int s_rank[kJacobiMaxOrder + 8];
  const JacobiSvdJob J = jobs[blockIdx.x];
  const int tid = threadIdx.x;
  const bool tr = J.m < J.n;  // W = A^T
  const int p = tr ? J.n : J.m, q = tr ? J.m : J.n;
  const int qe = q + (q & 1);  // an odd count gets a zero column
  double* W = lds;             // p x qe, column-major
  double* V = lds + int64_t(p) * qe;  // qe x qe
  if (!tr) {
    for (int idx = tid; idx < p * q; idx += NT) {
      const int r = idx % p, c = idx / p;
      W[c * p + r] = J.a[r + int64_t(c) * J.lda];
    }
  } else {
    for (int idx = tid; idx < p * q; idx += NT) {
      const int i = idx % J.m, j = idx / J.m;  // A(i, j) = W(j, i), read along A's columns
      W[i * p + j] = J.a[i + int64_t(j) * J.lda];
    }
  }
  for (int r = tid; r < p; r += NT)
    if (qe > q) W[q * p + r] = 0.0;
  for (int idx = tid; idx < qe * qe; idx += NT) V[idx] = (idx % qe == idx / qe) ? 1.0 : 0.0;
  if (tid < 2) s_rot[tid] = 0;
  __syncthreads();

  const int npairs = qe / 2;
  int T = 64;
  while (T > 1 && npairs * T > NT) T >>= 1;
  const int pk = tid / T, pt = tid % T;
  const double tol = 2.220446049250313e-16 * sqrt(static_cast<double>(p));
  int converged = qe < 2 ? 1 : 0;
  for (int sweep = 0; sweep < max_sweeps && !converged; ++sweep) {
    for (int r = 0; r < qe - 1; ++r) {
      if (pk < npairs) {
        // round-robin pairing: qe - 1 rounds meet every pair once
        const int i = pk == 0 ? qe - 1 : (r + pk) % (qe - 1);
        const int j = pk == 0 ? r : (r + qe - 1 - pk) % (qe - 1);
        double* wi = W + int64_t(i) * p;
        double* wj = W + int64_t(j) * p;
        double al = 0.0, be = 0.0, ga = 0.0;
        for (int k = pt; k < p; k += T) {
          const double x = wi[k], y = wj[k];
          al = fma(x, x, al);
          be = fma(y, y, be);
          ga = fma(x, y, ga);
        }
        for (int o = T / 2; o > 0; o >>= 1) {
          al += __shfl_xor(al, o, 64);
          be += __shfl_xor(be, o, 64);
          ga += __shfl_xor(ga, o, 64);
        }
        if (i < q && j < q && al > 0.0 && be > 0.0 && fabs(ga) > tol * sqrt(al) * sqrt(be)) {
          // rotation that makes columns i and j orthogonal (the smaller root
          // of t^2 + 2 zeta t - 1 = 0)
          const double zeta = (be - al) / (2.0 * ga);
          const double t = copysign(1.0, zeta) / (fabs(zeta) + sqrt(fma(zeta, zeta, 1.0)));
          const double c = 1.0 / sqrt(fma(t, t, 1.0)), sn = c * t;
          for (int k = pt; k < p; k += T) {
            const double x = wi[k], y = wj[k];
            wi[k] = c * x - sn * y;
            wj[k] = sn * x + c * y;
          }
          double* vi = V + int64_t(i) * qe;
          double* vj = V + int64_t(j) * qe;
          for (int k = pt; k < qe; k += T) {
            const double x = vi[k], y = vj[k];
            vi[k] = c * x - sn * y;
            vj[k] = sn * x + c * y;
          }
          if (pt == 0) s_rot[sweep & 1] = 1;
        }
      }
      __syncthreads();
    }
    const int rot = s_rot[sweep & 1];
    __syncthreads();  // every thread has read the flag before it is cleared
    if (tid == 0) s_rot[sweep & 1] = 0;
    converged = rot == 0;
  }
  if (tid == 0) flags[blockIdx.x] = converged ? 0 : 1;

  // singular values and their descending order (ties by column index)
  for (int c = tid; c < q; c += NT) {
    double x = 0.0;
    for (int k = 0; k < p; ++k) x = fma(W[c * p + k], W[c * p + k], x);
    s_sig[c] = sqrt(x);
  }
  __syncthreads();
  for (int c = tid; c < q; c += NT) {
    int rank = 0;
    for (int k = 0; k < q; ++k) rank += (s_sig[k] > s_sig[c] || (s_sig[k] == s_sig[c] && k < c)) ? 1 : 0;
    s_rank[c] = rank;
    J.s[rank] = s_sig[c];
  }
  __syncthreads();
  double smax = 0.0;
  for (int k = 0; k < q; ++k) smax = fmax(smax, s_sig[k]);
  const bool all_zero = smax == 0.0;
  // W side: w_c / S_c (length p); rotation side: column c of V (length q)
  for (int idx = tid; idx < p * q; idx += NT) {
    const int k = idx % p, c = idx / p;
    const double sg = s_sig[c];
    double x = sg > 0.0 ? W[c * p + k] / sg : 0.0;
    if (all_zero && s_rank[c] == 0) x = k == 0 ? 1.0 : 0.0;
    const int rk = s_rank[c];
    if (!tr) J.u[k + int64_t(rk) * J.ldu] = x;    // U(:, rk), m = p rows
    else J.vt[rk + int64_t(k) * J.ldvt] = x;      // Vt(rk, :), n = p columns
  }
  for (int idx = tid; idx < q * q; idx += NT) {
    const int k = idx % q, c = idx / q;
    const int rk = s_rank[c];
    const double x = V[c * qe + k];
    if (!tr) J.vt[rk + int64_t(k) * J.ldvt] = x;  // Vt(rk, :), n = q columns
    else J.u[k + int64_t(rk) * J.ldu] = x;        // U(:, rk), m = q rows
  }
}

hipError_t launch_jacobi_svd(const JacobiSvdJob* jobs, int njobs, int lds_words, int max_sweeps, int* flags,
                             hipStream_t st) {
  if (njobs <= 0) return hipSuccess;
  if (lds_words <= 0 || lds_words > kJacobiLdsWords || max_sweeps < 1) return hipErrorInvalidValue;
  hipLaunchKernelGGL(jacobi_svd_kernel, dim3(njobs), dim3(256), size_t(lds_words) * sizeof(double), st, jobs,
                     max_sweeps, flags);
  return hipGetLastError();
}

}  // namespace sadmrg
