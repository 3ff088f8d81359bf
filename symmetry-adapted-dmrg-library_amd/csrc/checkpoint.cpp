// checkpoint.cpp -- MPS checkpoint files in the reference's format.
//
// DMRG::save_mps2 / read_mps / save_ww / read_ww / save_enr / read_enr
// (dmrg_inst.cpp:947-1177).  Per site i, with base = uu-{ns}-{bondd}-{totspin}-{exci}-:
//   {base}{i}.dat  text: "nbond\tlocspin\n", then per bond "dir\tnmom\n", the
//                  nmom labels (2J or charge) each followed by '\t', '\n', the
//                  nmom reduced dimensions likewise (dmrg_inst.cpp:965-975);
//   {base}{i}.bin  binary: ndata Scalars, then ncgc doubles (SU(2)), where
//                  ndata / ncgc sum the reduced / CGC sizes of every allowed
//                  block (SU2Struct::get_nelement, su2.cpp:535-548).  The
//                  present blocks come first, block after block in block
//                  order (BlockTensor::get_elems, block_tensor_inst.cpp:643-660);
//                  the rest is zero (dmrg_inst.cpp:981-987).
// ww-{ns}-{bondd}-{totspin}-{exci}.bin holds bondd*phdim doubles of the last
// truncation's singular values and enr-{...}.dat the energies, appended one
// per line with 16 significant digits.
//
// Reading follows BlockTensor::set(struct, data, cgc) (block_tensor_inst.cpp:
// 120-159): every allowed block of the structure is filled, in block order,
// from consecutive elements of the file.  The engine keeps SU(2) CGC blocks
// symbolically, so a CGC block read back must be the canonical one of its
// labels (Dense::make_cgc / make_identity / make_singlet,
// dense_inst.cpp:673-708) within 1e-12; any other block is rejected.
#include <cmath>
#include <cstdint>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>
#include <vector>
#include "dmrg.hpp"

namespace sadmrg {

namespace {

std::string path_in(const std::string& dir, const std::string& name) {
  if (dir.empty()) return name;
  return dir.back() == '/' ? dir + name : dir + "/" + name;
}

// Canonical CGC block of SU(2) block labels (column-major, m = 2i - j).
void canonical_cgc(int nb, const int* lab, uint8_t tag, std::vector<double>& out) {
  if (nb == 3) {
    const int j1 = lab[0], j2 = lab[1], j3 = lab[2];
    out.assign(size_t(j1 + 1) * (j2 + 1) * (j3 + 1), 0.0);
    for (int k = 0; k <= j3; ++k)
      for (int j = 0; j <= j2; ++j)
        for (int i = 0; i <= j1; ++i)
          out[i + size_t(j) * (j1 + 1) + size_t(k) * (j1 + 1) * (j2 + 1)] =
              clebsch_gordan(j1, 2 * i - j1, j2, 2 * j - j2, j3, 2 * k - j3);
    return;
  }
  require(nb == 2, "checkpoint: SU(2) tensors of 2 or 3 bonds only");
  const int jj = lab[0];
  out.assign(size_t(jj + 1) * (jj + 1), 0.0);
  if (tag == kCgcIdent) {
    for (int i = 0; i <= jj; ++i) out[i + size_t(i) * (jj + 1)] = 1.0;
  } else if (tag == kCgcSinglet) {
    for (int i = 0; i <= jj; ++i)
      for (int j = 0; j <= jj; ++j) {
        const int m1 = 2 * i - jj, m2 = 2 * j - jj;
        if (m1 + m2 != 0) continue;
        const double sgn = (((jj - m1) / 2) % 2 == 0) ? 1.0 : -1.0;
        out[i + size_t(j) * (jj + 1)] = sgn / std::sqrt(static_cast<double>(jj + 1));
      }
  } else {
    fail(kErrUnsupported, "checkpoint: 2-leg block with a CGC tensor other than identity or singlet");
  }
}

int64_t cgc_size(const Struct& st, int blk) {
  std::vector<int> lab(st.nbond());
  st.labels(blk, lab.data());
  int64_t n = 1;
  for (int l : lab) n *= l + 1;
  return n;
}

void write_site(const BTensor& t, const std::string& dat, const std::string& bin) {
  {
    std::ofstream f(dat);
    if (!f) fail(kErrInvalid, "checkpoint: cannot write " + dat);
    f << t.nbond() << "\t" << t.st.locspin << "\n";
    for (const Bond& b : t.st.bonds) {
      f << b.dir << "\t" << b.nmom() << "\n";
      for (int q : b.q) f << q << "\t";
      f << "\n";
      for (int d : b.dim) f << d << "\t";
      f << "\n";
    }
    if (!f) fail(kErrInvalid, "checkpoint: write failed: " + dat);
  }
  std::ofstream f(bin, std::ios::binary);
  if (!f) fail(kErrInvalid, "checkpoint: cannot write " + bin);
  if (t.is_null()) return;
  const int w = t.w();
  const bool su2 = t.sym() == SymKind::SU2;
  int64_t ndata = 0, ncgc = 0;
  for (int b = 0; b < t.nten(); ++b)
    if (t.st.allowed(b)) {
      ndata += t.block_size(b);
      if (su2) ncgc += cgc_size(t.st, b);
    }
  std::vector<unsigned char> present(t.nten());
  std::vector<double> data(std::max<int64_t>(ndata, 1) * w, 0.0);
  require(t.nelem <= ndata, "checkpoint: tensor holds blocks outside its structure");
  tensor_to_host(t, present.data(), data.data());
  f.write(reinterpret_cast<const char*>(data.data()), ndata * w * sizeof(double));
  if (su2) {
    std::vector<double> cg(std::max<int64_t>(ncgc, 1), 0.0), blk;
    std::vector<int> lab(t.nbond());
    int64_t off = 0;
    for (int b = 0; b < t.nten(); ++b) {
      if (!t.st.allowed(b) || !t.present(b)) continue;
      t.st.labels(b, lab.data());
      canonical_cgc(t.nbond(), lab.data(), t.tag.empty() ? kCgcCanon : t.tag[b], blk);
      std::copy(blk.begin(), blk.end(), cg.begin() + off);
      off += static_cast<int64_t>(blk.size());
    }
    f.write(reinterpret_cast<const char*>(cg.data()), ncgc * sizeof(double));
  }
  if (!f) fail(kErrInvalid, "checkpoint: write failed: " + bin);
}

// Returns false when either file cannot be opened (the reference's pass flag).
bool read_site(BTensor& t, Context* ctx, SymKind sym, bool cplx, const std::string& dat, const std::string& bin) {
  std::ifstream fd(dat);
  if (!fd) return false;
  std::vector<int> hdr;
  int nb = 0, ls = 0;
  if (!(fd >> nb >> ls) || nb < 1 || nb > 8) fail(kErrInvalid, "checkpoint: bad header in " + dat);
  hdr = {nb, ls};
  for (int j = 0; j < nb; ++j) {
    int dir = 0, nm = 0;
    if (!(fd >> dir >> nm) || nm < 1) fail(kErrInvalid, "checkpoint: bad bond in " + dat);
    std::vector<int> q(nm), d(nm);
    for (int k = 0; k < nm; ++k)
      if (!(fd >> q[k])) fail(kErrInvalid, "checkpoint: bad labels in " + dat);
    for (int k = 0; k < nm; ++k)
      if (!(fd >> d[k]) || d[k] < 0) fail(kErrInvalid, "checkpoint: bad dimensions in " + dat);
    hdr.push_back(dir);
    hdr.push_back(nm);
    hdr.insert(hdr.end(), q.begin(), q.end());
    hdr.insert(hdr.end(), d.begin(), d.end());
  }
  std::ifstream fb(bin, std::ios::binary);
  if (!fb) return false;
  const Struct st = parse_header(sym, hdr.data(), static_cast<int>(hdr.size()));
  const int w = cplx ? 2 : 1;
  int64_t ndata = 0, ncgc = 0;
  std::vector<unsigned char> present(st.nten(), 0);
  for (int b = 0; b < st.nten(); ++b)
    if (st.allowed(b)) {
      present[b] = 1;
      ndata += st.block_size(b);
      if (sym == SymKind::SU2) ncgc += cgc_size(st, b);
    }
  std::vector<double> data(std::max<int64_t>(ndata, 1) * w, 0.0);
  fb.read(reinterpret_cast<char*>(data.data()), ndata * w * sizeof(double));
  if (fb.gcount() != static_cast<std::streamsize>(ndata * w * sizeof(double)))
    fail(kErrInvalid, "checkpoint: " + bin + " is shorter than its structure");
  std::vector<uint8_t> tags(st.nten(), kCgcCanon);
  if (sym == SymKind::SU2) {
    std::vector<double> cg(std::max<int64_t>(ncgc, 1), 0.0), blk, alt;
    fb.read(reinterpret_cast<char*>(cg.data()), ncgc * sizeof(double));
    if (fb.gcount() != static_cast<std::streamsize>(ncgc * sizeof(double)))
      fail(kErrInvalid, "checkpoint: " + bin + " lacks the CGC blocks of its structure");
    std::vector<int> lab(nb);
    int64_t off = 0;
    auto same = [&](const std::vector<double>& ref) {
      for (size_t e = 0; e < ref.size(); ++e)
        if (std::fabs(cg[off + e] - ref[e]) > 1e-12) return false;
      return true;
    };
    for (int b = 0; b < st.nten(); ++b) {
      if (!present[b]) continue;
      st.labels(b, lab.data());
      if (nb == 2) {
        canonical_cgc(nb, lab.data(), kCgcIdent, blk);
        canonical_cgc(nb, lab.data(), kCgcSinglet, alt);
        if (same(blk)) tags[b] = kCgcIdent;
        else if (same(alt)) tags[b] = kCgcSinglet;
        else fail(kErrUnsupported, "checkpoint: " + bin + " block " + std::to_string(b) +
                                       " carries a CGC tensor other than identity or singlet");
      } else {
        canonical_cgc(nb, lab.data(), kCgcCanon, blk);
        if (!same(blk))
          fail(kErrUnsupported, "checkpoint: " + bin + " block " + std::to_string(b) +
                                    " carries a CGC tensor other than the canonical coupling");
      }
      off += static_cast<int64_t>(blk.size());
    }
  }
  tensor_from_host(t, ctx, sym, cplx, hdr.data(), static_cast<int>(hdr.size()), present.data(), data.data(), 0);
  if (sym == SymKind::SU2)
    for (int b = 0; b < st.nten(); ++b)
      if (present[b] && b < static_cast<int>(t.tag.size())) t.tag[b] = tags[b];
  return true;
}

}  // namespace

std::string Dmrg::file_tag(int dim, int exci) const {
  return std::to_string(ns_) + "-" + std::to_string(dim) + "-" + std::to_string(totspin_) + "-" +
         std::to_string(exci);
}

// dmrg_inst.cpp:947-994
void Dmrg::save_mps(const std::string& dir, int exci) const {
  const std::string base = "uu-" + file_tag(bondd_, exci) + "-";
  for (int i = 0; i < ns_; ++i)
    write_site(uu_[i], path_in(dir, base + std::to_string(i) + ".dat"),
               path_in(dir, base + std::to_string(i) + ".bin"));
}

// dmrg_inst.cpp:1146-1153: bondd * phdim doubles of ww[exci]
void Dmrg::save_ww(const std::string& dir) const {
  const std::string name = path_in(dir, "ww-" + file_tag(bondd_, exci_) + ".bin");
  std::ofstream f(name, std::ios::binary);
  if (!f) fail(kErrInvalid, "checkpoint: cannot write " + name);
  std::vector<double> w(size_t(bondd_) * (physpn_ + 1), 0.0);
  for (size_t k = 0; k < w.size() && k < wout_.size(); ++k) w[k] = wout_[k];
  f.write(reinterpret_cast<const char*>(w.data()), w.size() * sizeof(double));
}

// dmrg_inst.cpp:997-1003
void Dmrg::save_enr(const std::string& dir) const {
  const std::string name = path_in(dir, "enr-" + file_tag(bondd_, exci_) + ".dat");
  std::ofstream f(name, std::ios::app);
  if (!f) fail(kErrInvalid, "checkpoint: cannot write " + name);
  f << std::setprecision(16) << gs_enr_ << "\n";
}

// dmrg_inst.cpp:1031-1076
bool Dmrg::read_mps(const std::string& dir, int read1) {
  const std::string base = "uu-" + file_tag(read1, exci_) + "-";
  bool pass = true;
  for (int i = 0; i < ns_; ++i) {
    BTensor t;
    if (read_site(t, ctx_, sym_, cplx_, path_in(dir, base + std::to_string(i) + ".dat"),
                  path_in(dir, base + std::to_string(i) + ".bin")))
      uu_[i] = std::move(t);
    else
      pass = false;
  }
  ctx_->sync();
  return pass;
}

// dmrg_inst.cpp:1078-1141: the driver's own files (tag exci), then the
// stored states 0..exci-1 from the files of bond dimension read2.  The
// reference skips a stored state whose files are missing and later works on
// empty tensors; here that is an error.
bool Dmrg::read_mps(const std::string& dir, int read1, int read2) {
  const bool pass = read_mps(dir, read1);
  for (int idx = 0; idx < exci_; ++idx) {
    const std::string base = "uu-" + file_tag(read2, idx) + "-";
    for (int i = 0; i < ns_; ++i) {
      BTensor t;
      if (!read_site(t, ctx_, sym_, cplx_, path_in(dir, base + std::to_string(i) + ".dat"),
                     path_in(dir, base + std::to_string(i) + ".bin")))
        fail(kErrInvalid, "read_mps: stored state " + std::to_string(idx) + " is missing " + base +
                              std::to_string(i) + ".dat/.bin");
      orth_[idx][i] = std::move(t);
    }
  }
  ctx_->sync();
  read_ww(dir, read1);
  read_enr(dir, read1, read2);
  return pass;
}

// dmrg_inst.cpp:1156-1176: read1 doubles, or ones when the file is missing
void Dmrg::read_ww(const std::string& dir, int read1) {
  std::ifstream f(path_in(dir, "ww-" + file_tag(read1, exci_) + ".bin"), std::ios::binary);
  wout_.assign(std::max(read1, 0), 1.0);
  if (f) {
    std::fill(wout_.begin(), wout_.end(), 0.0);
    f.read(reinterpret_cast<char*>(wout_.data()), wout_.size() * sizeof(double));
  }
}

// dmrg_inst.cpp:1007-1027: the last value of the file
void Dmrg::read_enr(const std::string& dir, int read1) {
  std::ifstream f(path_in(dir, "enr-" + file_tag(read1, exci_) + ".dat"));
  double e;
  while (f >> e) gs_enr_ = e;
}

// the stored states' energies from the files of bond dimension read2 (the
// penalty weights; a missing file is an error)
void Dmrg::read_enr(const std::string& dir, int read1, int read2) {
  read_enr(dir, read1);
  for (int idx = 0; idx < exci_; ++idx) {
    const std::string name = path_in(dir, "enr-" + file_tag(read2, idx) + ".dat");
    std::ifstream f(name);
    if (!f) fail(kErrInvalid, "read_enr: missing " + name);
    double e;
    bool any = false;
    while (f >> e) {
      lower_enr_[idx] = e;
      any = true;
    }
    if (!any) fail(kErrInvalid, "read_enr: no energy in " + name);
  }
}

}  // namespace sadmrg
