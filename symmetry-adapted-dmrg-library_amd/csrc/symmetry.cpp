// symmetry.cpp -- see symmetry.hpp.
#include "symmetry.hpp"
#include <algorithm>
#include <cstdlib>
#include <sstream>
#include "core.hpp"

namespace sadmrg {

Bond Bond::make(int dir, std::vector<int> q, std::vector<int> dim) {
  require(q.size() == dim.size(), "bond: label/dim length mismatch");
  // pairwise exchange sort, ascending by label (src/symmetry/su2.cpp:22-29)
  const int n = static_cast<int>(q.size());
  for (int i = 0; i < n - 1; ++i)
    for (int j = i + 1; j < n; ++j)
      if (q[j] < q[i]) {
        std::swap(q[i], q[j]);
        std::swap(dim[i], dim[j]);
      }
  Bond b;
  b.dir = dir;
  b.q = std::move(q);
  b.dim = std::move(dim);
  return b;
}

int Bond::find(int label) const {
  for (int i = 0; i < nmom(); ++i)
    if (q[i] == label) return i;
  return -1;
}

int64_t Bond::total_dim() const {
  int64_t t = 0;
  for (int d : dim) t += d;
  return t;
}

std::string Bond::str() const {
  std::ostringstream os;
  os << "dir=" << dir << " [";
  for (int i = 0; i < nmom(); ++i) os << (i ? " " : "") << q[i] << ":" << dim[i];
  os << "]";
  return os.str();
}

Bond fuse_bonds(SymKind sym, const Bond& a, const Bond& b) {
  std::vector<int> lab, dim;
  auto add = [&](int c, int d) {
    for (size_t l = 0; l < lab.size(); ++l)
      if (lab[l] == c) { dim[l] += d; return; }
    lab.push_back(c);
    dim.push_back(d);
  };
  if (sym == SymKind::SU2) {
    // triangle decomposition; result direction -a.dir (src/symmetry/su2.cpp:182-217)
    require(a.dir == b.dir, "SU2 fuse: bond directions differ");
    for (int i = 0; i < a.nmom(); ++i)
      for (int j = 0; j < b.nmom(); ++j)
        for (int jk = std::abs(a.q[i] - b.q[j]); jk <= a.q[i] + b.q[j]; jk += 2) add(jk, a.dim[i] * b.dim[j]);
    return Bond::make(-a.dir, lab, dim);
  }
  // U(1): charge q1+q2 for equal directions, -q1+q2 otherwise; result -b.dir
  // (src/symmetry/u1.cpp:111-141)
  for (int j = 0; j < b.nmom(); ++j)
    for (int i = 0; i < a.nmom(); ++i) {
      const int c = (a.dir * b.dir == 1) ? a.q[i] + b.q[j] : -a.q[i] + b.q[j];
      add(c, a.dim[i] * b.dim[j]);
    }
  return Bond::make(-b.dir, lab, dim);
}

int Struct::nten() const {
  if (bonds.empty()) return 0;
  int n = 1;
  for (const auto& b : bonds) n *= b.nmom();
  return n;
}

void Struct::decode(int blk, int* idx) const {
  for (int i = 0; i < nbond(); ++i) {
    idx[i] = blk % bonds[i].nmom();
    blk /= bonds[i].nmom();
  }
}

int Struct::encode(const int* idx) const {
  int b = 0;
  for (int i = nbond() - 1; i >= 0; --i) b = b * bonds[i].nmom() + idx[i];
  return b;
}

void Struct::labels(int blk, int* lab) const {
  for (int i = 0; i < nbond(); ++i) {
    const int s = blk % bonds[i].nmom();
    blk /= bonds[i].nmom();
    lab[i] = bonds[i].q[s];
  }
}

int Struct::index_of_labels(const int* lab) const {
  int b = 0;
  for (int i = nbond() - 1; i >= 0; --i) {
    const int s = bonds[i].find(lab[i]);
    if (s < 0) return -1;
    b = b * bonds[i].nmom() + s;
  }
  return b;
}

bool Struct::allowed_labels(const int* lab) const {
  const int nb = nbond();
  if (nb <= 0) return false;
  if (sym == SymKind::U1) {
    // existence plus charge balance (src/symmetry/u1.cpp:367-378)
    int tot = 0;
    for (int i = 0; i < nb; ++i) {
      if (bonds[i].find(lab[i]) < 0) return false;
      tot += bonds[i].dir * lab[i];
    }
    return tot == locspin;
  }
  // SU(2): successive coupling reachability (src/symmetry/su2.cpp:483-507)
  if (nb == 1) return lab[0] == locspin;
  int maxj = 0;
  for (int i = 0; i < nb; ++i) maxj += lab[i];
  std::vector<unsigned char> reach(maxj + 1, 0), next(maxj + 1, 0);
  reach[lab[0]] = 1;
  for (int i = 1; i < nb; ++i) {
    std::fill(next.begin(), next.end(), 0);
    const int j = lab[i];
    for (int J = 0; J <= maxj; ++J) {
      if (!reach[J]) continue;
      for (int K = std::abs(J - j); K <= J + j; K += 2) next[K] = 1;
    }
    reach.swap(next);
  }
  return (locspin >= 0 && locspin <= maxj) ? reach[locspin] != 0 : false;
}

bool Struct::allowed(int blk) const {
  int lab[8];
  labels(blk, lab);
  return allowed_labels(lab);
}

void Struct::block_dims(int blk, int* d) const {
  for (int i = 0; i < nbond(); ++i) {
    const int s = blk % bonds[i].nmom();
    blk /= bonds[i].nmom();
    d[i] = bonds[i].dim[s];
  }
}

int64_t Struct::block_size(int blk) const {
  int d[8];
  block_dims(blk, d);
  int64_t n = 1;
  for (int i = 0; i < nbond(); ++i) n *= d[i];
  return n;
}

}  // namespace sadmrg
