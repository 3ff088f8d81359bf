// symmetry.hpp -- U(1) / SU(2) quantum-number bookkeeping (host, integer).
//
// Semantics follow the reference's symmetry classes:
//   * a bond holds sectors sorted ascending by label (2J or charge) with
//     their reduced dimensions and a direction (+1 in, -1 out)
//     (symmetry/su2.hpp:218-256, symmetry/u1.hpp:20-59);
//   * a block of an n-leg tensor is addressed by one sector index per bond,
//     flattened column-major with bond 0 fastest (src/symmetry/su2.cpp:432-450,
//     src/symmetry/u1.cpp:322-339);
//   * SU(2): a block is allowed when the legs' spins can couple to locspin
//     (src/symmetry/su2.cpp:483-507); U(1): sum(dir*q) == locspin
//     (src/symmetry/u1.cpp:348-378).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace sadmrg {

enum class SymKind : int { U1 = 0, SU2 = 1 };

struct Bond {
  int dir = 0;
  std::vector<int> q;    // sorted ascending
  std::vector<int> dim;  // reduced dimension per sector

  Bond() = default;
  // Builds a bond from unsorted sectors (sorted like src/symmetry/su2.cpp:22-29).
  static Bond make(int dir, std::vector<int> q, std::vector<int> dim);
  int nmom() const { return static_cast<int>(q.size()); }
  int find(int label) const;  // sector index or -1
  Bond inverted() const {
    Bond b = *this;
    b.dir = -b.dir;
    return b;
  }
  int64_t total_dim() const;
  bool operator==(const Bond& o) const { return dir == o.dir && q == o.q && dim == o.dim; }
  bool operator!=(const Bond& o) const { return !(*this == o); }
  std::string str() const;
};

// Fusion of two bonds (src/symmetry/su2.cpp:182-217; src/symmetry/u1.cpp:111-141).
Bond fuse_bonds(SymKind sym, const Bond& a, const Bond& b);

struct Struct {
  SymKind sym = SymKind::SU2;
  int locspin = 0;
  std::vector<Bond> bonds;

  int nbond() const { return static_cast<int>(bonds.size()); }
  int nten() const;
  // block index -> per-bond sector indices
  void decode(int blk, int* idx) const;
  int encode(const int* idx) const;
  // labels of block `blk` (2J or charge per leg)
  void labels(int blk, int* lab) const;
  // block index from labels; -1 if a label is missing on its bond
  int index_of_labels(const int* lab) const;
  // allowed-block rule applied to labels (check_angularmoments)
  bool allowed_labels(const int* lab) const;
  bool allowed(int blk) const;
  int64_t block_size(int blk) const;
  void block_dims(int blk, int* d) const;
  bool operator==(const Struct& o) const {
    return sym == o.sym && locspin == o.locspin && bonds == o.bonds;
  }
  bool operator!=(const Struct& o) const { return !(*this == o); }
};

}  // namespace sadmrg
