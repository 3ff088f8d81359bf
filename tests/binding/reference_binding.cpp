// The reference-side binding, compiled: the hot-path methods of the
// reference's BlockTensor<Scalar,Sym> (block_tensor.hpp:38-193) and of
// DMRG::lanczos_solve_eigenvector_idmrg (dmrg_inst.cpp:264-279) with their
// bodies routed into the engine's C ABI (include/sadmrg.h), as a maintainer
// of the reference would write them.  The class below keeps the reference's
// method names and argument meaning; its host metadata is reduced to the
// tensor header.  Built and run by tests/test_binding.py: without a GPU the
// program checks that the library reports SADMRG_ERR_NO_DEVICE; with one it
// runs the bound methods on a spin-1/2 SU(2) site and prints "binding ok".
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>
#include "sadmrg.h"

namespace binding {

[[noreturn]] void die(const char* where) {
  // the reference prints and exits on inconsistent input
  std::printf("%s: %s\n", where, sadmrg_last_error());
  std::exit(1);
}

sadmrg_ctx* g_ctx = nullptr;
sadmrg_cgc* g_cgc = nullptr;

template <typename Scalar, int Sym>
class BlockTensor {
 public:
  sadmrg_tensor* dev = nullptr;  // device twin of data_blocks
  static sadmrg_ctx* ctx() { return g_ctx; }
  static sadmrg_cgc* tables() { return g_cgc; }
  static constexpr int kCplx = sizeof(Scalar) == 16 ? SADMRG_COMPLEX : SADMRG_REAL;

  BlockTensor() {
    if (sadmrg_tensor_null(ctx(), Sym, kCplx, &dev) != SADMRG_OK) die("BlockTensor");
  }
  ~BlockTensor() { sadmrg_tensor_destroy(dev); }
  BlockTensor(const BlockTensor&) = delete;
  BlockTensor& operator=(const BlockTensor& o) {  // block_tensor.hpp:107
    if (sadmrg_tensor_copy(dev, o.dev) != SADMRG_OK) die("operator=");
    return *this;
  }
  // make_spinor_start (block_tensor.hpp:156)
  void make_spinor_start(int physpn) {
    if (sadmrg_make_spinor_start(dev, physpn) != SADMRG_OK) die("make_spinor_start");
  }
  // fuse(BondType, BondType) (block_tensor.hpp:129); bond = [dir, nmom, q..., dim...]
  void fuse(const std::vector<int>& b1, const std::vector<int>& b2) {
    if (sadmrg_fuse(dev, b1.data(), b2.data()) != SADMRG_OK) die("fuse");
  }
  // contract (block_tensor.hpp:142)
  void contract(BlockTensor& t1, int i1, BlockTensor& t2, int i2) {
    if (sadmrg_contract(dev, t1.dev, i1, t2.dev, i2) != SADMRG_OK) die("contract");
  }
  // hamiltonian_vector_multiplication (block_tensor.hpp:179)
  void hamiltonian_vector_multiplication(BlockTensor& vec, BlockTensor& op1, BlockTensor& op2) {
    if (sadmrg_hamiltonian_vector_multiplication(dev, vec.dev, op1.dev, op2.dev, tables()) != SADMRG_OK)
      die("hamiltonian_vector_multiplication");
  }
  // operator+= / -= (block_tensor.hpp:108-109)
  BlockTensor& operator+=(BlockTensor& x) {
    if (sadmrg_tensor_axpy(dev, 1.0, 0.0, x.dev) != SADMRG_OK) die("operator+=");
    return *this;
  }
  // operator*= (block_tensor.hpp:110)
  BlockTensor& operator*=(double a) {
    if (sadmrg_tensor_transform(dev, 4, a, 0.0) != SADMRG_OK) die("operator*=");
    return *this;
  }
  // inner_prod_c (block_tensor.hpp:117): real part for real Scalar
  double inner_prod(BlockTensor& y) {
    double z[2];
    if (sadmrg_tensor_dot(dev, y.dev, z) != SADMRG_OK) die("inner_prod_c");
    return z[0];
  }
  // normalize_vector (block_tensor.hpp:119)
  void normalize_vector() {
    if (sadmrg_tensor_transform(dev, 5, 0.0, 0.0) != SADMRG_OK) die("normalize_vector");
  }
  long long get_nelement() {
    long long n = 0;
    if (sadmrg_tensor_nelem(dev, &n) != SADMRG_OK) die("get_nelement");
    return n;
  }
};

// DMRG::lanczos_solve_eigenvector_idmrg (dmrg_inst.cpp:264-279): the
// H_eff plan of the step, the whole Lanczos solve on the device, release.
template <typename Scalar, int Sym>
double lanczos_solve_eigenvector(BlockTensor<Scalar, Sym>& hl, BlockTensor<Scalar, Sym>& hr,
                                 std::vector<BlockTensor<Scalar, Sym>*>& o1, std::vector<BlockTensor<Scalar, Sym>*>& o2,
                                 BlockTensor<Scalar, Sym>& vec, int mlanc) {
  std::vector<const sadmrg_tensor*> d1, d2;
  for (auto* t : o1) d1.push_back(t->dev);
  for (auto* t : o2) d2.push_back(t->dev);
  sadmrg_heff* h = nullptr;
  if (sadmrg_heff_create(BlockTensor<Scalar, Sym>::ctx(), BlockTensor<Scalar, Sym>::tables(), hl.dev, hr.dev,
                         static_cast<int>(d1.size()), d1.data(), d2.data(), vec.dev, &h) != SADMRG_OK)
    die("hamiltonian_vector_multiplication_idmrg");
  double e = 0;
  int counters[4];
  if (sadmrg_lanczos(h, vec.dev, mlanc, SADMRG_FP64, 0, 0.0, &e, counters, nullptr) != SADMRG_OK)
    die("lanczos1");
  sadmrg_heff_destroy(h);
  return e;
}

}  // namespace binding

int main() {
  using namespace binding;
  const int st = sadmrg_ctx_create(0, &g_ctx);
  if (st == SADMRG_ERR_NO_DEVICE) {
    std::printf("no device: %s\n", sadmrg_last_error());
    return 0;
  }
  if (st != SADMRG_OK) die("ctx");
  if (sadmrg_cgc_create(1, &g_cgc) != SADMRG_OK) die("cgc");
  using BT = BlockTensor<double, SADMRG_SU2>;
  {
    // the spin-1/2 site operator and its square through contract, + and *=
    BT s, s2, t;
    s.make_spinor_start(1);
    t = s;
    t *= 2.0;
    t += s;  // 3 s
    const double n1 = s.inner_prod(s), n3 = t.inner_prod(t);
    if (!(n1 > 0) || std::fabs(n3 - 9.0 * n1) > 1e-12 * n3) {
      std::printf("vector algebra mismatch: %g %g\n", n1, n3);
      return 1;
    }
    t.normalize_vector();
    if (std::fabs(t.inner_prod(t) - 1.0) > 1e-12) {
      std::printf("normalize_vector mismatch\n");
      return 1;
    }
    if (s.get_nelement() <= 0) return 1;
  }
  sadmrg_cgc_destroy(g_cgc);
  sadmrg_ctx_destroy(g_ctx);
  std::printf("binding ok\n");
  return 0;
}
