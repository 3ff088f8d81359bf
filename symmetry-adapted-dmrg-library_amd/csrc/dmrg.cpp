// dmrg.cpp -- see dmrg.hpp.
#include "dmrg.hpp"
#include "eigen.hpp"
#include <cstdio>
#include <cmath>
#include <iostream>

namespace sadmrg {

namespace {
using Clock = std::chrono::steady_clock;
double since(Clock::time_point t0) { return std::chrono::duration<double>(Clock::now() - t0).count(); }
}  // namespace

Dmrg::Dmrg(Context* ctx, SymKind sym, bool cplx, int xx, int yy, int bondd, int physpn, int totspin, int max_exci)
    : ctx_(ctx), sym_(sym), cplx_(cplx), lx_(xx), ly_(yy), ns_(xx * yy), physpn_(physpn),
      totspin_(totspin), max_dcut_(bondd), bondd_(bondd), exci_(max_exci) {
  require(ns_ >= 4, "DMRG: at least 4 sites are required");
  require(max_exci >= 0 && max_exci <= 64, "DMRG: max_exci must lie in [0, 64]");
  // the overlap environments of the centre pair reach three sites past it
  require(max_exci == 0 || ns_ >= 6, "DMRG: excited states need at least 6 sites");
  require(physpn == 1, "DMRG: only spin-1/2 sites (physpn = 1) are supported");
  hmap_.assign(ns_ * ns_, 0);
  fll_.assign(ns_ * ns_, 0);
  frr_.assign(ns_ * ns_, 0);
  // snake ordering, nearest neighbours, periodic in both directions
  // (dmrg_inst.cpp:117-144)
  for (int k = 0; k < 2; ++k)
    for (int i = 0; i < ns_; ++i) {
      const int x1 = i / ly_;
      const int y1 = (x1 % 2 == 0) ? i % ly_ : ly_ - 1 - (i % ly_);
      int x2, y2;
      if (k == 0) { x2 = (x1 + 1) % lx_; y2 = y1; }
      else { x2 = x1; y2 = (y1 + 1) % ly_; }
      const int j = (x2 % 2 == 0) ? x2 * ly_ + y2 : x2 * ly_ + (ly_ - 1 - y2);
      if (i != j) {
        hmap_[i * ns_ + j] = 1;
        hmap_[j * ns_ + i] = 1;
      }
    }
  coup_.assign(ns_ + 1, 0.0);
  coup_[1] = 1.0;
  set_couplings(hmap_, coup_);
  uu_.resize(ns_);
  hh_.resize(ns_);
  opr_.resize(ns_);
  for (auto& r : opr_) r.resize(ns_);
  if (sym_ == SymKind::SU2) cgc_ = std::make_unique<CgcTables>(physpn_);
  orth_.resize(exci_);
  ovlp_.resize(exci_);
  for (int j = 0; j < exci_; ++j) {
    orth_[j].resize(ns_);
    ovlp_[j].resize(ns_);
  }
  overlapvec_.resize(exci_);
  lower_enr_.assign(exci_, 0.0);
  initialize_local_operators();
}

void Dmrg::set_lower_state(int j, const Dmrg& o) {
  require(j >= 0 && j < exci_, "set_lower_state: state index out of range");
  require(o.ns_ == ns_ && o.sym_ == sym_ && o.cplx_ == cplx_ && o.totspin_ == totspin_,
          "set_lower_state: drivers differ in lattice size, symmetry, scalar type or sector");
  for (int i = 0; i < ns_; ++i) copy_tensor(orth_[j][i], o.uu_[i]);
  lower_enr_[j] = o.gs_enr_;
  ctx_->sync();
}

double Dmrg::lower_energy(int j) const {
  require(j >= 0 && j < exci_, "lower_energy: state index out of range");
  return lower_enr_[j];
}

void Dmrg::set_couplings(const std::vector<int>& hmap, const std::vector<double>& coup) {
  require(static_cast<int>(hmap.size()) == ns_ * ns_, "set_couplings: hmap must be ns x ns");
  require(chw_.empty() || chw_.size() == 3 * coup.size(),
          "set_couplings: the channel weights cover a different number of coupling indices; set them again");
  for (int v : hmap) require(v >= 0 && v < static_cast<int>(coup.size()), "set_couplings: coupling index out of range");
  hmap_ = hmap;
  coup_ = coup;
  for (int i = 0; i < ns_; ++i) {
    for (int j = 0; j < ns_; ++j) { fll_[i * ns_ + j] = 0; frr_[i * ns_ + j] = 0; }
    for (int j = 0; j < i; ++j)
      for (int k = i + 1; k < ns_; ++k)
        if (hm(j, k) != 0) fll_[i * ns_ + j] = 1;
    for (int j = i + 1; j < ns_; ++j)
      for (int k = 0; k < i; ++k)
        if (hm(j, k) != 0) frr_[i * ns_ + j] = 1;
  }
}

void Dmrg::set_channel_weights(const std::vector<double>& w) {
  if (w.empty()) {
    chw_.clear();
    return;
  }
  require(sym_ == SymKind::U1, "channel weights: U(1) drivers only (SU(2) fixes the isotropic exchange)");
  require(w.size() == 6 * coup_.size(), "channel weights: 6 doubles per coupling index are required");
  std::vector<dcmplex> c(3 * coup_.size());
  for (size_t i = 0; i < c.size(); ++i) {
    c[i] = dcmplex(w[2 * i], w[2 * i + 1]);
    require(cplx_ || c[i].imag() == 0.0, "channel weights: complex weights need a complex driver");
  }
  for (size_t k = 0; k < coup_.size(); ++k)  // Hermitian bond terms: w(-q) = conj(w(q))
    require(std::abs(c[3 * k] - std::conj(c[3 * k + 2])) <= 1e-14 * (1.0 + std::abs(c[3 * k])) &&
                c[3 * k + 1].imag() == 0.0,
            "channel weights: w(-2) must be the conjugate of w(+2) and w(0) real");
  chw_ = c;
}

void Dmrg::add_coupled(BTensor& op, int c, const BTensor& x) const {
  if (chw_.empty()) {
    op_axpy(op, coup_[c], 0.0, x);
    return;
  }
  const dcmplex w[3] = {chw_[3 * c], chw_[3 * c + 1], chw_[3 * c + 2]};
  op_axpy_channels(op, coup_[c], w, x, 1);
}

void Dmrg::copy_state_from(const Dmrg& o) {
  require(o.ns_ == ns_ && o.sym_ == sym_ && o.cplx_ == cplx_ && o.physpn_ == physpn_,
          "copy_state_from: drivers differ in lattice size, symmetry or scalar type");
  hmap_ = o.hmap_;
  fll_ = o.fll_;
  frr_ = o.frr_;
  coup_ = o.coup_;
  totspin_ = o.totspin_;
  max_dcut_ = o.max_dcut_;
  lopt_ = o.lopt_;
  fp64_energy_every_ = o.fp64_energy_every_;
  gs_enr_ = o.gs_enr_;
  wout_ = o.wout_;
  for (int i = 0; i < ns_; ++i) {
    copy_tensor(uu_[i], o.uu_[i]);
    copy_tensor(hh_[i], o.hh_[i]);
    for (int j = 0; j < ns_; ++j) copy_tensor(opr_[i][j], o.opr_[i][j]);
  }
  require(o.exci_ == exci_, "copy_state_from: drivers target different excited states");
  for (int j = 0; j < exci_; ++j) {
    for (int i = 0; i < ns_; ++i) {
      copy_tensor(orth_[j][i], o.orth_[j][i]);
      copy_tensor(ovlp_[j][i], o.ovlp_[j][i]);
    }
    copy_tensor(overlapvec_[j], o.overlapvec_[j]);
  }
  lower_enr_ = o.lower_enr_;
  chw_ = o.chw_;
  // the last Lanczos solve's wave function and bond (coupling_gradient)
  copy_tensor(last_vec_, o.last_vec_);
  last_il_ = o.last_il_;
  last_ir_ = o.last_ir_;
  ctx_->sync();
}

// dmrg_inst.cpp:222-260
void Dmrg::initialize_local_operators() {
  op_make_spinor_start(sigma_[0], ctx_, cplx_, sym_, physpn_);
  op_make_spinor_start(sigma_[1], ctx_, cplx_, sym_, physpn_);
  if (sym_ == SymKind::SU2) {
    op_scale(sigma_[0], std::sqrt(3.0) / 2.0);
    op_scale(sigma_[1], -std::sqrt(3.0) / 2.0);
  } else {
    const Bond b = sigma_[0].st.bonds[1].inverted();
    BTensor t1, t2;
    op_fuse_to_singlet(t1, ctx_, cplx_, sym_, b, b);
    op_contract(t2, sigma_[1], 1, t1, 1);
    op_shift(t2, 1, 0);
    sigma_[1] = std::move(t2);
  }
  // the environment updates read the site operators' values as scalar weights
  sigma_[0].keep_host();
  sigma_[1].keep_host();
}

void Dmrg::boundary_tensors() {
  if (sym_ == SymKind::SU2) {
    const Bond b0 = Bond::make(1, {0}, {1}), b1 = Bond::make(1, {physpn_}, {1});
    const Bond b2 = Bond::make(1, {physpn_}, {1}), b3 = Bond::make(1, {totspin_}, {1});
    op_fuse(uu_[0], ctx_, cplx_, sym_, b0, b1);
    op_fuse(uu_[ns_ - 1], ctx_, cplx_, sym_, b2, b3);
  } else {
    const Bond b0 = Bond::make(1, {0}, {1}), b1 = Bond::make(1, {-1, 1}, {1, 1});
    const Bond b2 = Bond::make(1, {-1, 1}, {1, 1}), b3 = Bond::make(-1, {totspin_}, {1});
    op_fuse(uu_[0], ctx_, cplx_, sym_, b0, b1);
    op_fuse(uu_[ns_ - 1], ctx_, cplx_, sym_, b2, b3);
  }
}

// ---- environment updates ---------------------------------------------------
// Overlap environments of the stored states (dmrg_inst.cpp:448-449, 493-500):
// ovlp[j][i] = <orth_j| uu> over sites 0..i; with flag 1 the two sites right
// of i get the 3-leg pieces prepare_input_vector combines into overlapvec:
// ovlp[j][i+1] = ovlp[j][i] . orth_j[i+1] and ovlp[j][i+2] = orth_j[i+2] .
// ovlp[j][i+3] (legs shifted).
void Dmrg::overlap_env_left(int i, int flag) {
  CgcTables* cg = cgc_.get();
  for (int j = 0; j < exci_; ++j) {
    auto& ov = ovlp_[j];
    if (i == 0) {
      op_env(ov[0], kOverlapInitial, orth_[j][0], uu_[0], nullptr, nullptr, 0, cg);
      continue;
    }
    require(!ov[i - 1].is_null(), "overlap environment: the left environment of the stored state is missing");
    op_env(ov[i], kOverlapTransformation, orth_[j][i], uu_[i], &ov[i - 1], nullptr, 0, cg);
    if (flag == 1) {
      require(i + 3 < ns_, "overlap environment: site out of range");
      op_contract(ov[i + 1], ov[i], 0, orth_[j][i + 1], 0);
      op_contract(ov[i + 2], orth_[j][i + 2], 1, ov[i + 3], 0);
      op_shift(ov[i + 2], 1, 0);
    }
  }
}

// mirror image of overlap_env_left (dmrg_inst.cpp:522-523, 567-574)
void Dmrg::overlap_env_right(int i, int flag) {
  CgcTables* cg = cgc_.get();
  for (int j = 0; j < exci_; ++j) {
    auto& ov = ovlp_[j];
    if (i == ns_ - 1) {
      op_env(ov[i], kOverlapInitial, orth_[j][i], uu_[i], nullptr, nullptr, 1, cg);
      continue;
    }
    require(!ov[i + 1].is_null(), "overlap environment: the right environment of the stored state is missing");
    op_env(ov[i], kOverlapTransformation, orth_[j][i], uu_[i], &ov[i + 1], nullptr, 1, cg);
    if (flag == 1) {
      require(i - 3 >= 0, "overlap environment: site out of range");
      op_contract(ov[i - 2], ov[i - 3], 0, orth_[j][i - 2], 0);
      op_contract(ov[i - 1], orth_[j][i - 1], 1, ov[i], 0);
      op_shift(ov[i - 1], 1, 0);
    }
  }
}

// dmrg_inst.cpp:440-511
void Dmrg::prepare_site_operator_from_left(int i, int flag) {
  const auto t0 = Clock::now();
  CgcTables* cg = cgc_.get();
  if (i == 0) {
    op_env(opr_[0][0], kOperatorInitial, uu_[0], uu_[0], &sigma_[0], nullptr, 0, cg);
    overlap_env_left(0, 0);
    hh_[0].clean();
    cur_.t_env += since(t0);
    return;
  }
  if (!hh_[i - 1].is_null()) op_env(hh_[i], kOverlapTransformation, uu_[i], uu_[i], &hh_[i - 1], nullptr, 0, cg);
  else hh_[i].clean();
  {
    std::vector<BTensor*> outs;
    std::vector<const BTensor*> ops;
    for (int k = 0; k < i; ++k)
      if (fll_[i * ns_ + k] == 1) {
        outs.push_back(&opr_[i][k]);
        ops.push_back(&opr_[i - 1][k]);
      }
    op_env_transform_many(outs, uu_[i], uu_[i], ops, 0, cg);
  }
  BTensor op;
  for (int k = 0; k < i; ++k)
    if (hm(i, k) > 0 && !opr_[i - 1][k].is_null()) add_coupled(op, hm(i, k), opr_[i - 1][k]);
  if (!op.is_null()) {
    BTensor ht;
    op_env(ht, kOperatorPairup, uu_[i], uu_[i], &op, &sigma_[1], 0, cg);
    op_axpy(hh_[i], 1.0, 0.0, ht);
  }
  op_env(opr_[i][i], kOperatorInitial, uu_[i], uu_[i], &sigma_[0], nullptr, 0, cg);
  overlap_env_left(i, flag);
  cur_.t_env += since(t0);
}

// dmrg_inst.cpp:515-583
void Dmrg::prepare_site_operator_from_right(int i, int flag) {
  const auto t0 = Clock::now();
  CgcTables* cg = cgc_.get();
  if (i == ns_ - 1) {
    op_env(opr_[i][i], kOperatorInitial, uu_[i], uu_[i], &sigma_[1], nullptr, 1, cg);
    overlap_env_right(i, 0);
    hh_[i].clean();
    cur_.t_env += since(t0);
    return;
  }
  if (!hh_[i + 1].is_null()) op_env(hh_[i], kOverlapTransformation, uu_[i], uu_[i], &hh_[i + 1], nullptr, 1, cg);
  else hh_[i].clean();
  {
    std::vector<BTensor*> outs;
    std::vector<const BTensor*> ops;
    for (int m0 = 0; m0 < ns_ - 1 - i; ++m0) {
      const int k = i + 1 + m0;
      if (frr_[i * ns_ + k] == 1) {
        outs.push_back(&opr_[i][k]);
        ops.push_back(&opr_[i + 1][k]);
      }
    }
    op_env_transform_many(outs, uu_[i], uu_[i], ops, 1, cg);
  }
  BTensor op;
  for (int m = 0; m < ns_ - 1 - i; ++m) {
    const int k = i + 1 + m;
    if (hm(i, k) > 0 && !opr_[i + 1][k].is_null()) add_coupled(op, hm(i, k), opr_[i + 1][k]);
  }
  if (!op.is_null()) {
    BTensor ht;
    op_env(ht, kOperatorPairup, uu_[i], uu_[i], &op, &sigma_[0], 1, cg);
    op_axpy(hh_[i], 1.0, 0.0, ht);
  }
  op_env(opr_[i][i], kOperatorInitial, uu_[i], uu_[i], &sigma_[1], nullptr, 1, cg);
  overlap_env_right(i, flag);
  cur_.t_env += since(t0);
}

// dmrg_inst.cpp:587-639 (enlarged basis: identity fusion of uu[i]'s legs)
void Dmrg::prepare_site_operator_from_left(int i) {
  const auto t0 = Clock::now();
  CgcTables* cg = cgc_.get();
  BTensor u1;
  op_fuse(u1, ctx_, cplx_, sym_, uu_[i].st.bonds[0], uu_[i].st.bonds[1]);
  if (!hh_[i - 1].is_null()) op_env(hh_[i], kOverlapTransformation, u1, u1, &hh_[i - 1], nullptr, 0, cg);
  else hh_[i].clean();
  {
    std::vector<BTensor*> outs;
    std::vector<const BTensor*> ops;
    for (int k = 0; k < i; ++k)
      if (fll_[i * ns_ + k] == 1) {
        outs.push_back(&opr_[i][k]);
        ops.push_back(&opr_[i - 1][k]);
      }
    op_env_transform_many(outs, u1, u1, ops, 0, cg);
  }
  BTensor op;
  for (int k = 0; k < i; ++k)
    if (hm(i, k) > 0 && !opr_[i - 1][k].is_null()) add_coupled(op, hm(i, k), opr_[i - 1][k]);
  if (!op.is_null()) {
    BTensor ht;
    op_env(ht, kOperatorPairup, u1, u1, &op, &sigma_[1], 0, cg);
    op_axpy(hh_[i], 1.0, 0.0, ht);
  }
  op_env(opr_[i][i], kOperatorInitial, u1, u1, &sigma_[0], nullptr, 0, cg);
  cur_.t_env += since(t0);
}

// dmrg_inst.cpp:643-695
void Dmrg::prepare_site_operator_from_right(int i) {
  const auto t0 = Clock::now();
  CgcTables* cg = cgc_.get();
  BTensor u1;
  op_fuse(u1, ctx_, cplx_, sym_, uu_[i].st.bonds[0], uu_[i].st.bonds[1]);
  if (!hh_[i + 1].is_null()) op_env(hh_[i], kOverlapTransformation, u1, u1, &hh_[i + 1], nullptr, 1, cg);
  else hh_[i].clean();
  {
    std::vector<BTensor*> outs;
    std::vector<const BTensor*> ops;
    for (int m0 = 0; m0 < ns_ - 1 - i; ++m0) {
      const int k = i + 1 + m0;
      if (frr_[i * ns_ + k] == 1) {
        outs.push_back(&opr_[i][k]);
        ops.push_back(&opr_[i + 1][k]);
      }
    }
    op_env_transform_many(outs, u1, u1, ops, 1, cg);
  }
  BTensor op;
  for (int m = 0; m < ns_ - 1 - i; ++m) {
    const int k = i + 1 + m;
    if (hm(i, k) > 0 && !opr_[i + 1][k].is_null()) add_coupled(op, hm(i, k), opr_[i + 1][k]);
  }
  if (!op.is_null()) {
    BTensor ht;
    op_env(ht, kOperatorPairup, u1, u1, &op, &sigma_[0], 1, cg);
    op_axpy(hh_[i], 1.0, 0.0, ht);
  }
  op_env(opr_[i][i], kOperatorInitial, u1, u1, &sigma_[1], nullptr, 1, cg);
  cur_.t_env += since(t0);
}

// ---- wave-function preparation ---------------------------------------------
// dmrg_inst.cpp:375-408
void Dmrg::prepare_input_vector_initial(int il, int ir, BTensor& vec) {
  BTensor t1, t2, left, rght, tmp, tmp3;
  op_fuse(t1, ctx_, cplx_, sym_, uu_[il].st.bonds[0], uu_[il].st.bonds[1]);
  op_fuse(t2, ctx_, cplx_, sym_, uu_[ir].st.bonds[0], uu_[ir].st.bonds[1]);
  op_env(left, kOverlapInitial, uu_[il], t1, nullptr, nullptr, 0, cgc_.get());
  op_env(rght, kOverlapInitial, uu_[ir], t2, nullptr, nullptr, 1, cgc_.get());
  op_fuse_to_singlet(tmp, ctx_, cplx_, sym_, left.st.bonds[0].inverted(), rght.st.bonds[0].inverted());
  op_contract(tmp3, left, 0, tmp, 0);
  op_contract(vec, tmp3, 1, rght, 0);
  op_makeup_input_vector(vec);
  if (il + 1 == ir) build_overlapvec(il, ir, t1, t2);
}

// overlapvec[j]: stored state j projected on the two-site basis of (il, ir)
// (dmrg_inst.cpp:399-407, 429-436)
void Dmrg::build_overlapvec(int il, int ir, const BTensor& t1, const BTensor& t2) {
  for (int j = 0; j < exci_; ++j) {
    require(!ovlp_[j][il].is_null() && !ovlp_[j][ir].is_null(),
            "overlap vector: the environments of the stored state are missing");
    BTensor left, rght;
    op_env(left, kOverlapInitial, ovlp_[j][il], t1, nullptr, nullptr, 0, cgc_.get());
    op_env(rght, kOverlapInitial, ovlp_[j][ir], t2, nullptr, nullptr, 1, cgc_.get());
    if (sym_ == SymKind::SU2) op_take_conjugate(left, 0);
    op_contract(overlapvec_[j], left, 0, rght, 0);
    op_makeup_input_vector(overlapvec_[j]);
  }
}

// dmrg_inst.cpp:412-437
void Dmrg::prepare_input_vector(int il, int ir, BTensor& vec) {
  const auto t0 = Clock::now();
  BTensor t1, t2, left, rght;
  op_fuse(t1, ctx_, cplx_, sym_, uu_[il].st.bonds[0], uu_[il].st.bonds[1]);
  op_fuse(t2, ctx_, cplx_, sym_, uu_[ir].st.bonds[0], uu_[ir].st.bonds[1]);
  op_env(left, kOverlapInitial, uu_[il], t1, nullptr, nullptr, 0, cgc_.get());
  op_env(rght, kOverlapInitial, uu_[ir], t2, nullptr, nullptr, 1, cgc_.get());
  if (sym_ == SymKind::SU2) op_take_conjugate(left, 0);
  op_contract(vec, left, 0, rght, 0);
  op_makeup_input_vector(vec);
  if (il + 1 == ir) build_overlapvec(il, ir, t1, t2);
  cur_.t_env += since(t0);
}

// dmrg_inst.cpp:858-870: the stored states' site tensors follow uu's layout
void Dmrg::wavefunc_transformation(int i, int flag) {
  op_vectran(uu_[i], flag != 0);
  for (int j = 0; j < exci_; ++j) op_vectran(orth_[j][i], flag != 0);
}

// ---- H_eff and Lanczos ----------------------------------------------------------
// Bond terms of DMRG::hamiltonian_vector_multiplication_idmrg (dmrg_inst.cpp:310-351):
// on the left half each left operator opr[il][j] pairs with the coupling-
// weighted sum of right operators; on the right half the roles swap.
std::unique_ptr<HeffPlan> Dmrg::make_heff(int il, int ir, const BTensor& psi,
                                          std::vector<std::unique_ptr<BTensor>>* keep) {
  std::vector<std::unique_ptr<BTensor>> local;
  auto& hold = keep ? *keep : local;
  std::vector<const BTensor*> o1, o2;
  const bool left_half = il < ns_ / 2;
  const int m1 = left_half ? il + 1 : ns_ - ir;
  const int m4 = left_half ? ns_ - ir : il + 1;
  for (int m0 = 0; m0 < m1; ++m0) {
    auto op = std::make_unique<BTensor>();
    if (left_half) {
      const int j = m0;
      for (int m = 0; m < m4; ++m) {
        const int k = ir + m;
        if (hm(j, k) > 0 && !opr_[ir][k].is_null()) add_coupled(*op, hm(j, k), opr_[ir][k]);
      }
      if (op->is_null()) continue;
      o1.push_back(&opr_[il][j]);
      o2.push_back(op.get());
    } else {
      const int k = ir + m0;
      for (int m = 0; m < m4; ++m) {
        const int j = m;
        if (hm(j, k) > 0 && !opr_[il][j].is_null()) add_coupled(*op, hm(j, k), opr_[il][j]);
      }
      if (op->is_null()) continue;
      o1.push_back(op.get());
      o2.push_back(&opr_[ir][k]);
    }
    hold.push_back(std::move(op));
  }
  auto plan = std::make_unique<HeffPlan>(ctx_, cgc_.get(), hh_[il], hh_[ir], o1, o2, psi);
  // orthogonality penalty on adjacent sites (dmrg_inst.cpp:352-363)
  if (il == ir - 1)
    for (int j = 0; j < exci_; ++j) {
      require(!overlapvec_[j].is_null(), "H_eff: the overlap vector of a stored state is missing");
      plan->add_projector(overlapvec_[j], -lower_enr_[j]);
    }
  return plan;
}

// dmrg_inst.cpp:264-279
double Dmrg::lanczos_solve(int il, int ir, BTensor& vec) {
  const auto t0 = Clock::now();
  int64_t ndata = 0;
  for (int b = 0; b < vec.nten(); ++b)
    if (vec.st.allowed(b)) ndata += vec.block_size(b);
  int mlanc = 40;
  if (mlanc > ndata) mlanc = static_cast<int>(ndata) + 1;
  if (mlanc < 2) mlanc = 2;
  for (int b = 0; b < vec.nten(); ++b)
    require(!vec.st.allowed(b) || vec.present(b), "lanczos: input vector must hold every allowed block");
  const auto tp = Clock::now();
  auto plan = make_heff(il, ir, vec, nullptr);
  cur_.t_plan += since(tp);
  LanczosOpts o = lopt_;
  o.mlanc = mlanc;
  o.fp64_energy = step_fp64_energy_;
  step_fp64_energy_ = true;
  const LanczosOut r = ::sadmrg::lanczos_solve(*plan, vec.dptr(), o);
  ctx_->sync();
  gs_enr_ = r.energy;
  cur_.ritz32 = r.ritz32;
  cur_.energy64 = r.energy64 ? 1 : 0;
  cur_.il = il;
  cur_.ir = ir;
  cur_.energy = r.energy;
  cur_.matvec64 = r.matvec64;
  cur_.matvec32 = r.matvec32;
  cur_.heff_flops = plan->stats().flops;
  cur_.ndata = ndata;
  copy_tensor(last_vec_, vec);
  last_il_ = il;
  last_ir_ = ir;
  cur_.t_lanczos += since(t0);
  return r.energy;
}

void Dmrg::svd(const BTensor& vec, int a, int b, double p1, double p2) {
  const auto t0 = Clock::now();
  const SvdStats s0 = svd_stats();
  const EigProfile e0 = SymEigBatch::profile();
  op_svd(vec, uu_[a], p1, uu_[b], p2, max_dcut_, wout_);
  const int n = static_cast<int>(wout_.size());
  const int dcut = n > max_dcut_ ? max_dcut_ : n;
  double te = 0;
  for (int k = dcut; k < n; ++k) te += wout_[k] * wout_[k];
  cur_.trunc_err = te;
  cur_.kept = dcut;
  cur_.t_svd += since(t0);
  rec_.push_back(cur_);
  if (verbose_ > 0) {
    std::fprintf(stderr,
                 "[sadmrg] step %d-%d E=%.12f kept=%d trunc=%.3e env=%.4fs lanczos=%.4fs (plan %.4fs) svd=%.4fs\n", a,
                 b, cur_.energy, cur_.kept, te, cur_.t_env, cur_.t_lanczos, cur_.t_plan, cur_.t_svd);
    if (verbose_ > 1) {
      const SvdStats& s1 = svd_stats();
      const EigProfile& e1 = SymEigBatch::profile();
      std::fprintf(stderr,
                   "[sadmrg]   svd: gram=%lld gesdd=%lld redo=%lld max_sector=%lld t_gram=%.4f t_gesdd=%.4f "
                   "t_vectors=%.4f | eig: matrices=%lld launches=%lld max_order=%lld trd=%.4f bisect=%.4f "
                   "invit=%.4f (kernel %.4f) ortho=%.4f backt=%.4f vectors=%lld clusters=%lld max_cluster=%lld ns_steps=%lld\n",
                   (long long)(s1.gram_sectors - s0.gram_sectors), (long long)(s1.svd_sectors - s0.svd_sectors),
                   (long long)(s1.gram_redo - s0.gram_redo), (long long)s1.max_sector, s1.t_gram - s0.t_gram,
                   s1.t_gesdd - s0.t_gesdd, s1.t_vectors - s0.t_vectors, (long long)(e1.matrices - e0.matrices),
                   (long long)(e1.launches - e0.launches), (long long)e1.max_order, e1.trd - e0.trd,
                   e1.bisect - e0.bisect, e1.invit - e0.invit, e1.invit_kernel - e0.invit_kernel, e1.ortho - e0.ortho,
                   e1.backt - e0.backt,
                   (long long)(e1.vectors - e0.vectors), (long long)(e1.clusters - e0.clusters),
                   (long long)e1.max_cluster, (long long)(e1.ns_steps - e0.ns_steps));
    }
    std::fflush(stderr);
  }
  cur_ = StepRecord{};
}

// ---- coupling gradient -------------------------------------------------------
namespace {

// dE/dH for E = <psi, H_L^T psi> (left, H[a, a'] -> sum_b psi[a, b] psi[a', b])
// or E = <psi, psi H_R> (right, H[b, b'] -> sum_a psi[a, b] psi[a, b']), laid
// out with the block Hamiltonian's structure.
void block_density(BTensor& lam, const BTensor& psi, const BTensor& h, bool left) {
  Context* ctx = psi.ctx;
  std::vector<std::vector<int>> per(h.nten());
  const int n0 = psi.st.bonds[0].nmom();
  const Bond& hb0 = h.st.bonds[0];
  const Bond& hb1 = h.st.bonds[1];
  for (int m = 0; m < h.nten(); ++m) {
    if (!h.st.allowed(m)) continue;
    const int i0 = m % hb0.nmom(), i1 = m / hb0.nmom();
    if (hb0.q[i0] != hb1.q[i1]) continue;
    for (int b = 0; b < psi.nten(); ++b) {
      if (!psi.present(b)) continue;
      const int r = b % n0, c = b / n0;
      const int q = left ? psi.st.bonds[0].q[r] : psi.st.bonds[1].q[c];
      if (q == hb0.q[i0]) per[m].push_back(b);
    }
  }
  std::vector<uint8_t> mask(h.nten());
  for (int m = 0; m < h.nten(); ++m) mask[m] = !per[m].empty();
  lam.clean();
  lam.layout(ctx, false, h.st, mask, kCgcIdent);
  GemmBatch g;
  for (int m = 0; m < h.nten(); ++m) {
    if (per[m].empty()) continue;
    int d[2];
    h.st.block_dims(m, d);
    const int p = g.add_problem(1, lam.off[m], d[0], d[0], d[1], false);
    for (int b : per[m]) {
      int pd[2];
      psi.st.block_dims(b, pd);
      Operand A, B;
      A.arena = 0;
      A.off = psi.off[b];
      A.ld = pd[0];
      B = A;
      if (left) {  // psi_b psi_b^T
        require(pd[0] == d[0], "block density: dimension mismatch");
        B.trans = true;
        g.add_seg(p, A, B, pd[1], 1.0);
      } else {     // psi_b^T psi_b
        require(pd[1] == d[0], "block density: dimension mismatch");
        A.trans = true;
        g.add_seg(p, A, B, pd[0], 1.0);
      }
    }
  }
  g.limit(0, psi.nelem);
  g.limit(1, lam.nelem);
  g.finalize(ctx, false);
  ArenaTable at{};
  at.base[0] = psi.dptr();
  at.base[1] = lam.dptr();
  g.run(ctx, GemmPrec::F64, at);
  ctx->sync();
}

}  // namespace

std::vector<double> Dmrg::coupling_gradient(double* energy) {
  require(last_il_ >= 0 && !last_vec_.is_null(), "coupling gradient: no Lanczos solve has run");
  require(exci_ == 0, "coupling gradient: ground-state drivers only (max_exci = 0)");
  require(chw_.empty(), "coupling gradient: isotropic couplings only (no channel weights)");
  require(!cplx_, "coupling gradient: real tensors only");
  const int il = last_il_, ir = last_ir_;
  const BTensor& psi = last_vec_;
  CgcTables* cg = cgc_.get();
  std::vector<double> grad(coup_.size(), 0.0);
  double e = 0;
  // bonds across the two sites: <psi| O1_j^T psi O2_k>
  for (int j = 0; j <= il; ++j)
    for (int k = ir; k < ns_; ++k) {
      if (hm(j, k) <= 0 || opr_[il][j].is_null() || opr_[ir][k].is_null()) continue;
      BTensor t;
      op_hvm(t, psi, opr_[il][j], opr_[ir][k], cg);
      if (t.is_null()) continue;
      const double v = op_dot(psi, t).real();
      grad[hm(j, k)] += v;
      e += coup_[hm(j, k)] * v;
    }
  // left block
  if (!hh_[il].is_null()) {
    BTensor lam;
    block_density(lam, psi, hh_[il], true);
    e += op_dot(lam, hh_[il]).real();
    for (int i = il; i >= 1; --i) {
      BTensor u1;
      if (i == il) op_fuse(u1, ctx_, cplx_, sym_, uu_[i].st.bonds[0], uu_[i].st.bonds[1]);
      const BTensor& basis = (i == il) ? u1 : uu_[i];
      for (int k = 0; k < i; ++k) {
        if (hm(i, k) <= 0 || opr_[i - 1][k].is_null()) continue;
        BTensor ht;
        op_env(ht, kOperatorPairup, basis, basis, &opr_[i - 1][k], &sigma_[1], 0, cg);
        if (ht.is_null()) continue;
        require(ht.st == lam.st, "coupling gradient: pair-up term and adjoint differ in structure");
        grad[hm(i, k)] += op_dot(lam, ht).real();
      }
      if (i - 1 < 1 || hh_[i - 1].is_null()) break;
      BTensor nl;
      op_env_adjoint(nl, lam, basis, basis, hh_[i - 1], 0);
      lam = std::move(nl);
    }
  }
  // right block
  if (!hh_[ir].is_null()) {
    BTensor lam;
    block_density(lam, psi, hh_[ir], false);
    e += op_dot(lam, hh_[ir]).real();
    for (int i = ir; i <= ns_ - 2; ++i) {
      BTensor u1;
      if (i == ir) op_fuse(u1, ctx_, cplx_, sym_, uu_[i].st.bonds[0], uu_[i].st.bonds[1]);
      const BTensor& basis = (i == ir) ? u1 : uu_[i];
      for (int k = i + 1; k < ns_; ++k) {
        if (hm(i, k) <= 0 || opr_[i + 1][k].is_null()) continue;
        BTensor ht;
        op_env(ht, kOperatorPairup, basis, basis, &opr_[i + 1][k], &sigma_[0], 1, cg);
        if (ht.is_null()) continue;
        require(ht.st == lam.st, "coupling gradient: pair-up term and adjoint differ in structure");
        grad[hm(i, k)] += op_dot(lam, ht).real();
      }
      if (i + 1 > ns_ - 2 || hh_[i + 1].is_null()) break;
      BTensor nl;
      op_env_adjoint(nl, lam, basis, basis, hh_[i + 1], 1);
      lam = std::move(nl);
    }
  }
  if (energy) *energy = e;
  return grad;
}

// ---- drivers ------------------------------------------------------------------
// dmrg_inst.cpp:699-804
void Dmrg::do_idmrg() {
  boundary_tensors();
  Bond pb1, pb2;
  if (sym_ == SymKind::SU2) { pb1 = Bond::make(1, {physpn_}, {1}); pb2 = pb1; }
  else { pb1 = Bond::make(1, {-1, 1}, {1, 1}); pb2 = pb1; }
  for (int i = 0; i <= ns_ / 2 - 2; ++i) {
    prepare_site_operator_from_left(i, 0);
    prepare_site_operator_from_right(ns_ - 1 - i, i < ns_ / 2 - 2 ? 0 : 1);
    const Bond b0 = uu_[i].st.bonds[2].inverted();
    const Bond b1 = uu_[ns_ - 1 - i].st.bonds[2].inverted();
    op_fuse(uu_[i + 1], ctx_, cplx_, sym_, b0, pb1);
    op_fuse(uu_[ns_ - 2 - i], ctx_, cplx_, sym_, pb2, b1);
    if (((i + 2) * 2 * physpn_ >= totspin_ && i >= 1) || i == ns_ / 2 - 2) {
      BTensor vec;
      prepare_input_vector_initial(i + 1, ns_ - 2 - i, vec);
      prepare_site_operator_from_left(i + 1);
      prepare_site_operator_from_right(ns_ - 2 - i);
      op_random_init(vec);
      op_normalize(vec);
      lanczos_solve(i + 1, ns_ - 2 - i, vec);
      if (i == ns_ / 2 - 2) svd(vec, i + 1, ns_ - 2 - i, 1.0, 0.0);
      else svd(vec, i + 1, ns_ - 2 - i, 0.0, 0.0);
    }
  }
}

// dmrg_inst.cpp:808-854
void Dmrg::prepare_sweep() {
  boundary_tensors();
  for (int i = 0; i < ns_ / 2 - 1; ++i) {
    prepare_site_operator_from_left(i, 0);
    prepare_site_operator_from_right(ns_ - 1 - i, i < ns_ / 2 - 2 ? 0 : 1);
  }
}

// dmrg_inst.cpp:874-943 (checkpoint files only when a directory is set)
void Dmrg::sweep() {
  BTensor vec;
  // FP64-energy policy of the mixed path (set_fp64_energy_every): step s of
  // nsteps computes the FP64 energy when every == 1, when s is a multiple of
  // every (every > 1), and at the last step
  const int nsteps = (ns_ / 2 - nfree_ - 1) + (ns_ - 2 * nfree_ - 2) + (ns_ - nfree_ - 1 - ns_ / 2);
  int step = 0;
  auto policy = [&]() {
    const int k = fp64_energy_every_;
    ++step;
    step_fp64_energy_ = k == 1 || step >= nsteps || (k > 1 && step % k == 0);
  };
  for (int i = ns_ / 2; i >= nfree_ + 2; --i) {
    wavefunc_transformation(i - 1, 1);
    prepare_site_operator_from_right(i, 1);
    prepare_input_vector(i - 2, i - 1, vec);
    prepare_site_operator_from_left(i - 2);
    prepare_site_operator_from_right(i - 1);
    policy();
    lanczos_solve(i - 2, i - 1, vec);
    if (i > nfree_ + 2) svd(vec, i - 2, i - 1, 1.0, 0.0);
    else svd(vec, i - 2, i - 1, 0.0, 1.0);
  }
  for (int i = nfree_; i < ns_ - nfree_ - 2; ++i) {
    wavefunc_transformation(i + 1, 0);
    prepare_site_operator_from_left(i, 1);
    prepare_input_vector(i + 1, i + 2, vec);
    prepare_site_operator_from_left(i + 1);
    prepare_site_operator_from_right(i + 2);
    policy();
    lanczos_solve(i + 1, i + 2, vec);
    if (i < ns_ - nfree_ - 3) svd(vec, i + 1, i + 2, 0.0, 1.0);
    else svd(vec, i + 1, i + 2, 1.0, 0.0);
  }
  for (int i = ns_ - nfree_ - 1; i >= ns_ / 2 + 1; --i) {
    wavefunc_transformation(i - 1, 1);
    prepare_site_operator_from_right(i, 1);
    prepare_input_vector(i - 2, i - 1, vec);
    prepare_site_operator_from_left(i - 2);
    prepare_site_operator_from_right(i - 1);
    policy();
    lanczos_solve(i - 2, i - 1, vec);
    svd(vec, i - 2, i - 1, 1.0, 0.0);
  }
  ctx_->sync();
  if (ckpt_on_) {
    save_mps(ckpt_dir_, exci_);
    save_ww(ckpt_dir_);
    save_enr(ckpt_dir_);
  }
}

}  // namespace sadmrg
