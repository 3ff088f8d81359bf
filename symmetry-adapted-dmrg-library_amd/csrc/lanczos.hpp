// lanczos.hpp -- device Lanczos ground-state solver.
//
// Mirrors Lanczos<Scalar,Sym>::lanczos1 (dmrg_inst.cpp:1286-1348): Krylov
// basis of at most mlanc vectors, full re-orthogonalisation, tridiagonal
// eigensolve after every step, early exit when beta < 1e-6, convergence check
// every 4th step with one extra matvec (stop when |<v,Hv>|/|Hv| is within 1e-8
// of 1).  Krylov vectors stay on the device in FP64; only the tridiagonal
// (<= 40 x 40) goes to the host.
//
// Mixed precision: the Krylov matvecs run with FP32 operands on FP32 MFMA and
// the convergence-check matvecs in FP64, so the stopping test sees the exact
// operator and the last check's H x gives the exact residual
// r = H x - (x^T H x) x of the Ritz vector.  If |r| exceeds refine_tol |E|
// the vector is refined by restarts: a few FP32 Krylov steps on r, then an
// FP64 matvec for the new residual.  The energy returned is the FP64
// Rayleigh quotient of the final vector.
#pragma once
#include <vector>
#include "heff.hpp"

namespace sadmrg {

struct LanczosOpts {
  int mlanc = 40;
  Prec prec = Prec::FP64;     // precision of the bulk matvecs
  bool refine = false;        // FP64 residual-correction restarts (FP32 bulk only)
  double refine_tol = 1e-7;   // stop when |r64| <= refine_tol * |E|
  int refine_krylov = 6;      // FP32 Krylov steps per restart
  int max_restarts = 6;
  int check_every = 4;        // convergence check after every check_every-th step (dmrg_inst.cpp:1334)
  // Krylov steps between checks run without host synchronisation (alpha,
  // beta and the step index stay on the device); with graph, one step is
  // captured in a HIP graph once per solve and replayed.  Same arithmetic
  // as the host-driven step.  Bases of more than 64 vectors use the
  // host-driven step.
  bool device_steps = true;
  bool graph = true;
  // Mixed precision with refinement only.  false: the solve makes no FP64
  // matvec.  After every check_every-th step the Ritz pair (theta, y) of the
  // FP32 Krylov recurrence is tested by its recurrence residual
  // beta_m |y_m| (no matvec); the solve stops when that is at most
  // refine_tol |theta| (the target the FP64 refinement would reach) and
  // returns theta, the Ritz value of the FP32-operand operator.  The vector
  // is that of an FP64 solve to within the FP32 operator error (about
  // 1e-7 |H| / gap); the returned energy carries the FP32 operator error.
  bool fp64_energy = true;
};

// Process-wide switches (default on): a solve uses device steps / graphs
// when both its options and these allow them.
void set_lanczos_mode(bool device_steps, bool graph);
void lanczos_mode(bool* device_steps, bool* graph);
// graph replays and captures since the last reset (process-wide)
void lanczos_graph_counts(long long* replays, long long* captures, bool reset);
// Nodes of the last captured Krylov step (every one a kernel node: checked at capture).
long long lanczos_graph_nodes();

struct LanczosOut {
  double energy = 0;
  int matvec64 = 0, matvec32 = 0;
  int steps = 0;
  int graph_launches = 0;   // Krylov steps replayed from a captured graph
  int restarts = 0;
  double residual64 = -1;     // |H x - E x| of the returned vector (FP64), if measured
  double ritz32 = 0;          // mixed precision: the last Ritz value of the FP32 Krylov recurrence
  bool energy64 = true;       // energy is an FP64 Rayleigh quotient (false: the FP32 Ritz value)
  std::vector<double> trace;  // Ritz values at each check
};

// Symmetric tridiagonal eigensolver (implicit QL with Wilkinson shifts).
// d[n] diagonal, e[n-1] off-diagonal; z (n x n, column-major) receives the
// eigenvectors; eigenvalues ascend in d.
void tridiag_eig(int n, std::vector<double>& d, std::vector<double> e, std::vector<double>& z);

// d_vec: start vector in the layout of H (n elements); on return the
// normalised Ritz vector.
LanczosOut lanczos_solve(HeffPlan& H, double* d_vec, const LanczosOpts& opt);

// Lanczos::lanczos1 of every instance of a merged plan (HeffPlan::merge) at
// once: one merged H_eff application per Krylov step serves every instance,
// and the vector operations (inner products, the three-term recurrence,
// re-orthogonalisation, normalisation, Ritz vectors) are segmented per
// instance, each instance with its own alpha/beta, basis size
// min(mlanc, n_i + 1), breakdown stop and convergence test (after every
// check_every-th step, |<v,Hv>| / |Hv| within 1e-8 of 1, the check matvec
// shared by the instances checking at that step).  Instances that finish
// stop updating; the merged matvec still covers them.  With refine: FP64
// checks, then FP64 residual-correction restarts of every instance whose
// |r| exceeds refine_tol |E| (batched over those instances, Krylov steps in
// the bulk precision: FP32 for the mixed path, FP64 for a refined FP64
// reference); energies are FP64 Rayleigh quotients.  seg: instance offsets in elements (n_inst + 1
// entries, seg[0] = 0, seg.back() = H.n()).  Host-driven steps; mlanc <= 64.
struct BatchedLanczosOut {
  std::vector<double> energy, residual64;
  std::vector<int> steps, restarts;
  int matvec64 = 0, matvec32 = 0;  // merged applications
};
BatchedLanczosOut lanczos_batched(HeffPlan& H, const std::vector<int64_t>& seg, double* d_vec,
                                  const LanczosOpts& opt);

}  // namespace sadmrg
