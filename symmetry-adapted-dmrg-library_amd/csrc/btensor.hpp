// btensor.hpp -- device-resident block-sparse symmetric tensor.
//
// BTensor is the runtime-typed core behind BlockTensor<Scalar,Sym>
// (block_tensor.hpp:38-193 of the reference).  Sector metadata lives on the
// host; the reduced-element data of all present blocks lives in one device
// allocation, block after block in block-index order, each block column-major.
// A block that is allowed but absent is "null", exactly like a Dense with
// nbond == 0 in the reference (dense_inst.cpp:755); null blocks own no memory.
//
// Operations mirror the reference's BlockTensor methods: the host enumerates
// the same loop nests with the same matching predicates, fac tables and skip
// thresholds, and lowers the per-block GEMMs into grouped block tables
// (gemm_plan.hpp) executed by one or two device launches.
#pragma once
#include <array>
#include <cstdint>
#include <random>
#include <unordered_map>
#include <vector>
#include "cgc.hpp"
#include "core.hpp"
#include "gemm_plan.hpp"
#include "symmetry.hpp"

namespace sadmrg {

// Largest number of legs of a block tensor.
constexpr int kMaxLeg = 8;

// Symbolic CGC block for SU(2) 2-leg tensors (3-leg tensors always carry the
// canonical <j0 j1 | j2> block in the paths used here).
enum CgcTag : uint8_t { kCgcIdent = 0, kCgcSinglet = 1, kCgcOther = 2, kCgcCanon = 3 };

struct BTensor {
  Context* ctx = nullptr;
  bool cplx = false;
  Struct st;                   // st.bonds empty => null tensor
  std::vector<int64_t> off;    // per block element offset, -1 = null block
  std::vector<uint8_t> tag;    // per block CGC tag (SU(2))
  DevBuf buf;                  // nelem elements of (cplx ? 16 : 8) bytes
  int64_t nelem = 0;
  bool fuse_id = false;        // built by op_fuse: 0/1 identity-embedding blocks
  // Host copy of the data (nelem * w doubles) kept for small constant tensors
  // that are read as scalar weights (site operators); empty when not kept.
  // layout() and clean() drop it; whoever writes the data afterwards must
  // drop or refresh it (keep_host()).
  std::vector<double> host;

  BTensor() = default;
  BTensor(BTensor&&) = default;
  BTensor& operator=(BTensor&&) = default;
  BTensor(const BTensor&) = delete;
  BTensor& operator=(const BTensor&) = delete;

  bool is_null() const { return st.bonds.empty(); }
  SymKind sym() const { return st.sym; }
  int nbond() const { return st.nbond(); }
  int nten() const { return st.nten(); }
  bool present(int b) const { return b >= 0 && b < static_cast<int>(off.size()) && off[b] >= 0; }
  int w() const { return cplx ? 2 : 1; }
  double* dptr() const { return buf.as<double>(); }
  double* block_ptr(int b) const { return dptr() + off[b] * w(); }
  int64_t block_size(int b) const { return st.block_size(b); }

  // Lays out the blocks flagged in mask (zero-filled device memory).
  void layout(Context* c, bool complex_, const Struct& s, const std::vector<uint8_t>& mask, uint8_t tag0);
  void clean();
  // Reads the data back into `host` (one synchronisation).
  void keep_host();
  // The data on the host: `host` when kept, otherwise read back into tmp.
  const double* host_values(std::vector<double>& tmp) const;
};

// The present, allowed blocks of a tensor with their labels, in block order
// (the order in which the reference's loops visit them), and indexes of
// those blocks by one or two labels.  Block loop nests (environment
// updates, H_eff plans) walk only the matching blocks, in the same order as
// the reference's full loop nests.
struct BlockList {
  std::vector<int> b;
  std::vector<std::array<int, 3>> lab;
  explicit BlockList(const BTensor& t) {
    int l[kMaxLeg];
    for (int x = 0; x < t.nten(); ++x) {
      if (!t.present(x) || !t.st.allowed(x)) continue;
      t.st.labels(x, l);
      b.push_back(x);
      lab.push_back({l[0], t.nbond() > 1 ? l[1] : 0, t.nbond() > 2 ? l[2] : 0});
    }
  }
  int size() const { return static_cast<int>(b.size()); }
};

struct LabelIndex {
  std::unordered_map<int64_t, std::vector<int>> m;
  static int64_t key(int a, int c) { return (static_cast<int64_t>(a) << 32) ^ static_cast<uint32_t>(c); }
  // positions p of list (ascending) keyed by (lab[p][i0], lab[p][i1]); i1 < 0: one label
  LabelIndex(const BlockList& list, int i0, int i1) {
    for (int p = 0; p < list.size(); ++p) m[key(list.lab[p][i0], i1 < 0 ? 0 : list.lab[p][i1])].push_back(p);
  }
  const std::vector<int>& get(int a, int c = 0) const {
    static const std::vector<int> none;
    auto it = m.find(key(a, c));
    return it == m.end() ? none : it->second;
  }
};

// ---- exchange with host (the C ABI tensor format) ------------------------
// header = [nbond, locspin, {dir, nmom, q..., dim...} per bond]
Struct parse_header(SymKind sym, const int* header, int len);
std::vector<int> make_header(const Struct& s);
void tensor_from_host(BTensor& t, Context* ctx, SymKind sym, bool cplx, const int* header, int len,
                      const unsigned char* present, const double* data, int cgc2_kind);
// data must hold nelem * w doubles
void tensor_to_host(const BTensor& t, unsigned char* present, double* data);
void copy_tensor(BTensor& dst, const BTensor& src);

// ---- reference-method mirrors --------------------------------------------
// BlockTensor::hamiltonian_vector_multiplication (block_tensor_inst.cpp:2170-2318)
void op_hvm(BTensor& out, const BTensor& vec, const BTensor& op1, const BTensor& op2, CgcTables* cgc);
// BlockTensor::contract (block_tensor_inst.cpp:1046-1126)
void op_contract(BTensor& out, const BTensor& t1, int i1, const BTensor& t2, int i2);

enum EnvKind : int {
  kOverlapInitial = 0,         // block_tensor_inst.cpp:1661-1788
  kOverlapTransformation = 1,  // :1790-1951
  kOperatorInitial = 2,        // :1282-1469
  kOperatorTransformation = 3, // :1471-1658
  kOperatorPairup = 4,         // :1953-2167
};
void op_env(BTensor& out, int kind, const BTensor& uu, const BTensor& vv, const BTensor* op1, const BTensor* op2,
            int flag, CgcTables* cgc);
// operator_transformation of several operators in one basis: outs[q] =
// T(ops[q]).  Operators with the same block layout (structure and block
// offsets) share one block table, built once and executed per operator with
// one host synchronisation per group; the results equal op_env's.
void op_env_transform_many(const std::vector<BTensor*>& outs, const BTensor& uu, const BTensor& vv,
                           const std::vector<const BTensor*>& ops, int flag, CgcTables* cgc);

// Adjoint of the overlap_transformation update (no reference counterpart;
// used for coupling gradients).  With T(X) = overlap_transformation(uu, vv,
// X, flag), out = T*(lam) has the structure of `like` and satisfies
// <lam, T(X)> = <out, X> for every X of that structure.  Real tensors.
void op_env_adjoint(BTensor& out, const BTensor& lam, const BTensor& uu, const BTensor& vv, const BTensor& like,
                    int flag);

// BlockTensor::svd (block_tensor_inst.cpp:501-637).  tu/tv are overwritten;
// wout receives kept values (sector order) followed by discarded values.
// Sweep-step truncations ((p1, p2) = (1, 0) or (0, 1), real data) decompose
// sectors whose larger side has at least gram_min() rows through the Gram
// matrix (eigenvectors of psi^T psi or psi psi^T) when that provably leaves
// the kept set unchanged, and by gesdd otherwise; gram_min() <= 0 turns the
// Gram route off.  Process-wide, default 256.
void op_svd(const BTensor& vec, BTensor& tu, double p1, BTensor& tv, double p2, int max_dcut,
            std::vector<double>& wout, std::vector<int>* kept_per_sector = nullptr);
void set_gram_min(int n);
int gram_min();
// Eigensolver of the Gram route: 1 (default) the hand-written batched solver
// (eigen.hpp; unpadded, eigenvectors of the kept values only), 0 rocSOLVER's
// strided-batched syevd on matrices padded to the batch's largest order.
void set_svd_eigensolver(int v);
int svd_eigensolver();
// Process-wide truncation counters: op_svd calls, sectors decomposed through
// the Gram route / by gesdd, Gram sectors redone by gesdd, wall seconds; the
// seconds spent in the Gram decompositions, in gesdd and in the eigen
// route's vectors, and the largest sector (max of m, n).
struct SvdStats {
  int64_t calls = 0, gram_sectors = 0, svd_sectors = 0, gram_redo = 0, max_sector = 0;
  int64_t jacobi_sectors = 0, jacobi_redo = 0;
  double seconds = 0, t_gram = 0, t_gesdd = 0, t_vectors = 0, t_jacobi = 0;
};
// In sweep steps, sectors of real blocks with min(m, n) <= q (and a Jacobi
// working set that fits 64 KB of LDS) are decomposed by the one-sided Jacobi
// kernel (launch_jacobi_svd); 0 turns it off.  Process-wide, default 64.
void set_svd_jacobi_max(int q);
int svd_jacobi_max();
SvdStats& svd_stats();

// BlockTensor::fuse(BondType, BondType) (block_tensor_inst.cpp:708-743)
void op_fuse(BTensor& out, Context* ctx, bool cplx, SymKind sym, const Bond& b1, const Bond& b2);
// BlockTensor::fuse_to_singlet (block_tensor_inst.cpp:780-816)
void op_fuse_to_singlet(BTensor& out, Context* ctx, bool cplx, SymKind sym, const Bond& b1, const Bond& b2);
// BlockTensor::left2right_vectran / right2left_vectran (:429-497)
void op_vectran(BTensor& t, bool left_to_right);
// BlockTensor::shift (:913-948)
void op_shift(BTensor& t, int i0, int i1);
// BlockTensor::take_conjugate(int) (:684-701): bond direction flip + CGC
void op_take_conjugate(BTensor& t, int bond);
// BlockTensor::makeup_input_vector (:1231-1257)
void op_makeup_input_vector(BTensor& t);
// BlockTensor::make_spinor_start (:1145-1217)
void op_make_spinor_start(BTensor& t, Context* ctx, bool cplx, SymKind sym, int physpn);

// ---- vector algebra (block_tensor_inst.cpp:197-376) ----------------------
void op_scale(BTensor& t, double re, double im = 0.0);            // *=, /=
void op_axpy(BTensor& y, double re, double im, const BTensor& x); // += (1,0), -= (-1,0)
dcmplex op_dot(const BTensor& x, const BTensor& y);               // inner_prod_c
double op_normalize(BTensor& t);                                  // normalize_vector
// y += s * w[c] * x with w[c] chosen per block by the label c of bond `bond`
// (U(1) channel labels -2, 0, +2 -> w[0], w[1], w[2]).  Complex weights need
// complex tensors.  No reference counterpart (the reference's couplings are
// one real number per bond).
void op_axpy_channels(BTensor& y, double s, const dcmplex (&w)[3], const BTensor& x, int bond);
void op_fill(BTensor& t, double v);                               // operator=(double)

// Global generator of the reference (dense_inst.cpp:23-24): mt19937_64(42),
// uniform(-1, 1), consumed element by element in block order.
std::mt19937_64& global_rng();
void reset_global_rng(uint64_t seed);
void op_random_init(BTensor& t);  // random_initialize_tensor (:1221-1229)

}  // namespace sadmrg
