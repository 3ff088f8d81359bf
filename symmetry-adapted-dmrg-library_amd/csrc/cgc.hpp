// cgc.hpp -- SU(2) Clebsch-Gordan coefficients and recoupling factor tables.
//
// Every SU(2) block GEMM on the inner loop is scaled by one scalar factor
// read from these tables.  The factors are defined exactly as the reference
// defines them (src/cgc_tables.cpp:102-357): contract canonical CGC tensors
// in the pattern of the corresponding block operation and take the ratio of
// the result to the canonical CGC (or identity) tensor of the output block.
// Entries are computed on first use and memoised, with the same guard
// conditions as the reference's table loops, so every entry the reference
// tabulates (all indices < max_angm = 20) is reproduced and larger spins are
// available without a cap.
#pragma once
#include <cstdint>
#include <mutex>
#include <unordered_map>
#include <vector>

namespace sadmrg {

// <j1 m1; j2 m2 | j3 m3> with all arguments doubled (Racah's closed form).
double clebsch_gordan(int j1, int m1, int j2, int m2, int j3, int m3);

class CgcTables {
 public:
  explicit CgcTables(int physpn);
  int physpn() const { return physpn_; }
  int p2() const { return 2 * physpn_ + 1; }

  // fac_hamilt_vec[i][j][k]                       (cgc_tables.cpp:156-164)
  double hamilt_vec(int i, int j, int k);
  // fac_operator_onsite_{left,rght}[j][m][l][i][p0 + p1*P2]   (:190-227)
  double onsite(int side, int j, int m, int l, int i, int pidx);
  // fac_operator_transformation_{left,rght}[j][m][l][i][k][p0] (:257-297)
  double transformation(int side, int j, int m, int l, int i, int k, int p0);
  // fac_operator_pairup_{left,rght}[j][i][k][m][p0 + p1*P2]    (:323-357)
  double pairup(int side, int j, int i, int k, int m, int pidx);

 private:
  double lookup(uint64_t key, double (CgcTables::*fn)(const int*), const int* args);
  double eval_hamilt(const int* a);
  double eval_onsite(const int* a);
  double eval_transformation(const int* a);
  double eval_pairup(const int* a);
  int physpn_;
  std::mutex mu_;
  std::unordered_map<uint64_t, double> cache_;
};

}  // namespace sadmrg
