// eigen.hpp -- batched symmetric eigensolver of the truncation's Gram route.
//
// The Gram route of BlockTensor::svd (block_tensor_inst.cpp:501-637,
// dense_inst.cpp:447-541) needs, per sector of one truncation step, every
// eigenvalue of a symmetric Gram matrix (for the global selection) and the
// eigenvectors of the kept (largest) ones only.  The reference obtains them
// from gesvd (dense_inst.cpp:486).  Here the matrices of one step are solved
// together:
//
//   1. tridiagonalisation Q^T G Q = T by Householder reflectors (the LAPACK
//      dsytd2 'L' scheme), all matrices in one persistent cooperative launch:
//      the rows of each matrix stay in the LDS of a group of workgroups, one
//      group barrier per column;
//   2. every eigenvalue of T by Sturm-sequence multisection (one wave each);
//   3. eigenvectors of T for the k largest eigenvalues by inverse iteration
//      with partial-pivoting LU of T - lambda I (one thread per cluster of
//      eigenvalues closer than inverse iteration separates, Gram-Schmidt
//      inside a cluster), made orthonormal by Newton-Schulz steps (GEMMs);
//   4. back-transformation V = Q Z (the reflectors applied in reverse order).
//
// Phase 1 (steps 1-2) and phase 2 (steps 3-4) are separate calls, so the
// caller can choose k from the eigenvalues of every matrix in between.
#pragma once
#include <cstdint>
#include <vector>
#include <hip/hip_runtime.h>
#include "core.hpp"

namespace sadmrg {

// Process-wide phase timers and counters of the solver (wall seconds; every
// phase ends with a stream synchronisation, so the host clock times it).
struct EigProfile {
  double trd = 0, bisect = 0, invit = 0, invit_kernel = 0, ortho = 0, backt = 0;
  int64_t matrices = 0, launches = 0, vectors = 0, clusters = 0, max_cluster = 0, ns_steps = 0, max_order = 0;
  int64_t serial_clusters = 0;  // clusters whose members came out dependent and ran again serially
};

struct EigJob {
  double* a = nullptr;  // n x n, column-major, leading dimension lda; lower triangle read, then overwritten
  int n = 0, lda = 0;
  double* w = nullptr;  // n eigenvalues, ascending (device)
  int k = 0;            // phase 2: eigenvectors of the k largest eigenvalues
  double* z = nullptr;  // phase 2: n x k, leading dimension ldz; column c belongs to w[n - 1 - c]
  int ldz = 0;
};

class SymEigBatch {
 public:
  explicit SymEigBatch(Context* ctx) : ctx_(ctx) {}
  // Phase 1: tridiagonalises every job's matrix and writes all eigenvalues.
  void eigenvalues(const std::vector<EigJob>& jobs);
  // Phase 2 (after phase 1, same jobs, k and z set): eigenvectors.
  void top_vectors(const std::vector<EigJob>& jobs);
  // Phase 1 without allocating: the working memory is a static device array
  // of kNoAllocBytes per device (its address looked up when the device's
  // first context is created); batches needing more are rejected.  Calls
  // that overlap on one device share that array and race.
  static constexpr int64_t kNoAllocBytes = int64_t(2) << 20;
  static void eigenvalues_noalloc(Context* ctx, const std::vector<EigJob>& jobs);
  // Largest matrix order the tridiagonalisation holds in LDS.
  static int max_order();
  // Phase-2 batches whose Newton-Schulz orthonormalisation did not converge
  // and fell back to Cholesky QR (process-wide count).
  static long long cholqr_fallbacks();
  static EigProfile& profile();
  // Clusters of close eigenvalues: members iterate in parallel lanes and are
  // orthonormalised afterwards (false, the default), or each cluster runs in
  // one lane with Gram-Schmidt after every solve (true).  Process-wide.
  static void set_cluster_serial(bool serial);
  static bool cluster_serial();

 private:
  Context* ctx_;
  std::vector<int64_t> doff_;  // per job: offset of d, e, tau (3n doubles) in trd_
  DevBuf trd_;
};

// Small-sector SVD (one-sided Jacobi, Hestenes): A = U S V^T of a real
// m x n block (column-major, leading dimension lda), S descending,
// U m x mn (ldu), Vt mn x n (ldvt), mn = min(m, n), each matrix in the LDS
// of one workgroup.  The longer side is the column length: the q = mn
// columns of W = A (or A^T when m < n) are rotated in pairs, q/2 disjoint
// pairs per round in round-robin order, until a whole sweep needs no
// rotation (|w_i . w_j| <= sqrt(p) eps |w_i| |w_j| for every pair).  S_j =
// |w_j|; the W side's vectors are w_j / S_j, the other side accumulates the
// rotations.  Column orthogonality is enforced relative to each column's
// norm, so small singular values keep their relative accuracy (unlike the
// Gram matrix route and gesdd).  A column with S_j = 0 gets a zero vector
// (e_0 when the whole block is zero); the truncation never keeps such a
// column except that one.
struct JacobiSvdJob {
  const double* a = nullptr;
  double* s = nullptr;
  double* u = nullptr;
  double* vt = nullptr;
  int m = 0, n = 0, lda = 0, ldu = 0, ldvt = 0, pad = 0;
};
// Largest min(m, n) and LDS words (doubles) of a matrix the kernel takes.
constexpr int kJacobiMaxOrder = 88;
constexpr int kJacobiLdsWords = 8192;  // 64 KB
bool jacobi_svd_fits(int m, int n, int max_order);
// One workgroup per job; flags[j] = 0 converged, 1 not converged within
// max_sweeps (the outputs are then those of the last sweep).
hipError_t launch_jacobi_svd(const JacobiSvdJob* jobs, int njobs, int lds_words, int max_sweeps, int* flags,
                             hipStream_t st);

// Looks up the device address of the allocation-free scratch of `device`
// (the current device), once per device.
void noalloc_scratch_setup(int device);

}  // namespace sadmrg
