// dmrg.hpp -- two-site DMRG driver on the device engine.
//
// Keeps the reference's DMRG<Scalar,Sym> state and step order (dmrg.hpp:29-133,
// src/dmrg_inst.cpp): lattice maps hmap/coup/fll/frr, MPS tensors uu[ns],
// block Hamiltonians hh[ns], open-bond operators opr[ns][ns], iDMRG warm-up,
// three-pass finite sweep.  Every tensor is device resident; each step builds
// one H_eff plan whose Lanczos solve runs entirely on the device.
#pragma once
#include <chrono>
#include <memory>
#include <string>
#include <vector>
#include "btensor.hpp"
#include "heff.hpp"
#include "lanczos.hpp"

namespace sadmrg {

struct StepRecord {
  int il, ir;
  double energy;
  double trunc_err;   // sum of squared discarded singular values
  int kept;           // multiplets (SU(2)) / states (U(1)) kept
  int matvec64, matvec32;
  double t_env, t_lanczos, t_svd;  // seconds
  double heff_flops;  // algorithmic flops per H_eff application at this step
  int64_t ndata;      // reduced elements of the two-site wave function
  double ritz32;      // mixed precision: the last Ritz value of the FP32 Krylov recurrence
  int energy64;       // 1: energy is an FP64 Rayleigh quotient; 0: the FP32 Ritz value
  double t_plan = 0;  // seconds of t_lanczos spent building the step's H_eff plan
};

class Dmrg {
 public:
  // xx, yy, bondd, physpn, totspin, max_exci as DMRG(int,int,int,int,int,int)
  // (dmrg_inst.cpp:63-187)
  Dmrg(Context* ctx, SymKind sym, bool cplx, int xx, int yy, int bondd, int physpn, int totspin, int max_exci = 0);

  int ns() const { return ns_; }
  // Replace the lattice map and recompute fll/frr (dmrg_inst.cpp:163-176).
  void set_couplings(const std::vector<int>& hmap, const std::vector<double>& coup);
  // Channel weights (U(1) only; no reference counterpart): bond (j < k) with
  // coupling index c contributes coup[c] * sum_q w[c][q] O_j(q) O_k(q) over
  // the channel labels q = -2, 0, +2 of the spin operators, instead of
  // coup[c] * S_j.S_k.  w holds 6 doubles (re, im for q = -2, 0, +2) per
  // coupling index; an empty vector restores the isotropic exchange.  XXZ
  // with a z-axis Dzyaloshinskii-Moriya term: w = (Jxy - i D, Jz, Jxy + i D).
  void set_channel_weights(const std::vector<double>& w);
  bool has_channel_weights() const { return !chw_.empty(); }
  void set_max_dcut(int d) { max_dcut_ = d; }
  void set_lanczos(const LanczosOpts& o) { lopt_ = o; }
  // 1: one line per sweep step on stderr (bond, energy, kept states, stage
  // times), as the reference prints each step of DMRG::sweep to cout.
  void set_verbose(int v) { verbose_ = v; }
  // Mixed precision with refinement: which sweep steps compute the FP64
  // energy (the FP64 Rayleigh quotient and the residual-correction restarts).
  // 1 (default): every step.  k > 1: every k-th step of a sweep and its last
  // step.  0: the last step of a sweep only.  The other steps stop on the
  // FP32 recurrence residual (LanczosOpts::fp64_energy = false) and record
  // the FP32 Ritz value.  iDMRG steps always compute the FP64 energy.
  void set_fp64_energy_every(int k) { fp64_energy_every_ = k; }
  int fp64_energy_every() const { return fp64_energy_every_; }
  const LanczosOpts& lanczos_opts() const { return lopt_; }

  // Deep copy of another driver's lattice maps, MPS, environments, settings
  // and last Lanczos wave function (same symmetry, scalar and lattice size);
  // device-to-device.
  void copy_state_from(const Dmrg& o);

  void do_idmrg();
  void prepare_sweep();
  void sweep();
  double energy() const { return gs_enr_; }

  // step-level entry points (private in the reference; exposed for parity)
  void wavefunc_transformation(int i, int flag);
  void prepare_site_operator_from_left(int i, int flag);
  void prepare_site_operator_from_right(int i, int flag);
  void prepare_site_operator_from_left(int i);
  void prepare_site_operator_from_right(int i);
  void prepare_input_vector(int il, int ir, BTensor& vec);
  void prepare_input_vector_initial(int il, int ir, BTensor& vec);
  std::unique_ptr<HeffPlan> make_heff(int il, int ir, const BTensor& psi_layout,
                                      std::vector<std::unique_ptr<BTensor>>* keep);
  double lanczos_solve(int il, int ir, BTensor& vec);
  void svd(const BTensor& vec, int a, int b, double p1, double p2);

  // dE/dJ_c for every coupling index c (Hellmann-Feynman, <psi| dH/dJ_c |psi>)
  // at the two-site wave function of the last Lanczos solve, with the
  // environments of that step: bonds across the two sites are measured
  // directly; bonds inside a block through the adjoint of the block-
  // Hamiltonian recursion (lam_{i-1} = T_i^*(lam_i), starting from
  // lam = dE/dH_block) contracted with each bond's pair-up term.  *energy
  // receives <psi|H|psi> assembled from the same pieces.  No reference
  // counterpart.
  std::vector<double> coupling_gradient(double* energy);

  // Checkpoint files in the reference's format (dmrg_inst.cpp:947-1177):
  // uu-{ns}-{bondd}-{totspin}-{exci}-{i}.dat/.bin per site, ww-...bin,
  // enr-...dat, in directory dir ("" = the current directory).
  void save_mps(const std::string& dir, int exci) const;  // save_mps2
  void save_ww(const std::string& dir) const;
  void save_enr(const std::string& dir) const;
  bool read_mps(const std::string& dir, int read1);             // read_mps(int)
  bool read_mps(const std::string& dir, int read1, int read2);  // + read_ww, read_enr
  void read_ww(const std::string& dir, int read1);
  void read_enr(const std::string& dir, int read1);
  void read_enr(const std::string& dir, int read1, int read2);  // + stored states' energies
  // sweep() ends with save_mps / save_ww / save_enr into dir, as the
  // reference's does (dmrg_inst.cpp:940-942); off by default.
  void set_checkpoint(const std::string& dir, bool on) { ckpt_dir_ = dir; ckpt_on_ = on; }
  const std::vector<double>& ww() const { return wout_; }

  // Excited states (dmrg_inst.cpp:283-583).  The driver targets state
  // exci() and keeps the stored states 0..exci()-1 (orth), their overlap
  // environments with the current MPS (ovlp) and their energies; every
  // H_eff on two adjacent sites carries the penalty -E_j |phi_j><phi_j|,
  // phi_j = stored state j projected on the current two-site basis
  // (overlapvec).  The reference loads the stored states from files
  // (read_mps with read2); set_lower_state takes them from another driver
  // (no reference counterpart).
  int exci() const { return exci_; }
  void set_lower_state(int j, const Dmrg& o);
  double lower_energy(int j) const;
  BTensor& orth(int j, int i) { return orth_.at(j).at(i); }
  BTensor& ovlp(int j, int i) { return ovlp_.at(j).at(i); }
  BTensor& overlapvec(int j) { return overlapvec_.at(j); }

  BTensor& uu(int i) { return uu_[i]; }
  BTensor& hh(int i) { return hh_[i]; }
  BTensor& opr(int i, int j) { return opr_[i][j]; }
  const std::vector<double>& last_wout() const { return wout_; }
  std::vector<StepRecord>& records() { return rec_; }
  CgcTables* cgc() { return cgc_.get(); }
  const std::vector<int>& hmap() const { return hmap_; }
  const std::vector<double>& coup() const { return coup_; }

 private:
  void initialize_local_operators();
  void boundary_tensors();
  int hm(int i, int j) const { return hmap_[i * ns_ + j]; }
  // op += coup[c] * x, channel weighted when weights are set
  void add_coupled(BTensor& op, int c, const BTensor& x) const;
  std::string file_tag(int dim, int exci) const;
  void overlap_env_left(int i, int flag);
  void overlap_env_right(int i, int flag);
  void build_overlapvec(int il, int ir, const BTensor& t1, const BTensor& t2);

  Context* ctx_;
  SymKind sym_;
  bool cplx_;
  int lx_, ly_, ns_, nfree_ = 2, physpn_, totspin_, max_dcut_, bondd_;
  std::string ckpt_dir_;
  bool ckpt_on_ = false;
  std::vector<int> hmap_, fll_, frr_;
  std::vector<double> coup_;
  std::vector<BTensor> uu_, hh_;
  std::vector<std::vector<BTensor>> opr_;
  BTensor sigma_[2];
  std::unique_ptr<CgcTables> cgc_;
  LanczosOpts lopt_;
  int verbose_ = 0;
  int fp64_energy_every_ = 1;
  bool step_fp64_energy_ = true;  // policy for the next lanczos_solve
  double gs_enr_ = 0;
  BTensor last_vec_;  // Ritz vector of the last Lanczos solve
  int last_il_ = -1, last_ir_ = -1;
  std::vector<double> wout_;
  std::vector<StepRecord> rec_;
  StepRecord cur_{};
  int exci_ = 0;
  std::vector<std::vector<BTensor>> orth_, ovlp_;  // [state][site]
  std::vector<BTensor> overlapvec_;
  std::vector<double> lower_enr_;
  std::vector<dcmplex> chw_;  // 3 per coupling index (q = -2, 0, +2), empty: isotropic
};

}  // namespace sadmrg
