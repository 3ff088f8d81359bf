// ref_harness.cpp -- TEST INFRASTRUCTURE ONLY (part of the oracle).
//
// Exposes the reference library's own C++ classes (compiled from the
// reference sources where they lie, see oracle/Makefile) through a small
// extern "C" surface so that pytest and bench.py's CPU-baseline leg can
// drive them with ctypes.  Nothing in the product links or loads this file.
//
// The reference's drivers (main/main_dmrg_su2_real.cpp:32-35) define the
// globals max_dcut / psize / myrank / g_cgc_tables; this harness defines
// them instead.  DMRG<..> keeps hh/opr/uu/hmap/coup private (dmrg.hpp:36-94);
// the harness reaches them by compiling this translation unit with the
// access specifier relaxed (no reference source is modified or copied).
//
// Tensor transfer format (shared with the product's C ABI):
//   header = [nbond, locspin, {dir, nmom, q[0..nmom-1], dim[0..nmom-1]} x nbond]
//   present[b] = 1 if block b (column-major block index over bonds) holds data
//   data = concatenation, in block order, of every present block, each block
//          column-major (bond 0 fastest); complex data is interleaved re,im.
#include <complex>
#include <cstring>
#include <cmath>
#include <cstdlib>
#include <iostream>
#include <fstream>
#include <sstream>
#include <vector>
#include <random>
#include <string>
#include <algorithm>
#include <random>
#include <chrono>
#include <omp.h>
#include <unistd.h>

#define private public
#include "dmrg.hpp"
#include "cgc_tables.hpp"
#undef private

int        max_dcut     = 64;
int        psize        = 1;
int        myrank       = 0;
CGCTables* g_cgc_tables = nullptr;

extern "C" int MKL_Set_Threading_Layer(int);
extern "C" void MKL_Set_Num_Threads(int);

namespace {

std::string g_err;
std::streambuf* g_saved_cout = nullptr;
std::ostringstream g_sink;

template <class S> struct ScalarIO;
template <> struct ScalarIO<double> {
  static constexpr int width = 1;
  static double get(const double* p) { return p[0]; }
  static void put(double v, double* p) { p[0] = v; }
};
template <> struct ScalarIO<dcmplex> {
  static constexpr int width = 2;
  static dcmplex get(const double* p) { return dcmplex(p[0], p[1]); }
  static void put(dcmplex v, double* p) { p[0] = v.real(); p[1] = v.imag(); }
};

// Runtime-typed handle around one of the four BlockTensor instantiations.
struct RefTensor {
  int sym;   // 0 = U1, 1 = SU2
  int cplx;  // 0 = double, 1 = complex<double>
  void* t;
};

template <class S, class Y> BlockTensor<S, Y>* as(RefTensor* h) {
  return static_cast<BlockTensor<S, Y>*>(h->t);
}

template <class S, class Y> RefTensor* wrap(BlockTensor<S, Y>* t) {
  RefTensor* h = new RefTensor;
  h->sym = std::is_same<Y, SU2Symmetry>::value ? 1 : 0;
  h->cplx = std::is_same<S, dcmplex>::value ? 1 : 0;
  h->t = t;
  return h;
}

template <class Y>
void parse_bonds(const int* hdr, int& nb, int& ls, std::vector<typename Y::BondType>& bonds) {
  nb = hdr[0];
  ls = hdr[1];
  bonds.resize(nb);
  int p = 2;
  for (int i = 0; i < nb; ++i) {
    int dir = hdr[p], nm = hdr[p + 1];
    bonds[i].set(dir, nm, hdr + p + 2, hdr + p + 2 + nm);
    p += 2 + 2 * nm;
  }
}

template <class S, class Y>
BlockTensor<S, Y>* make_tensor(const int* hdr, const unsigned char* present, const double* data,
                               int cgc2_kind) {
  int nb, ls;
  std::vector<typename Y::BondType> bonds;
  parse_bonds<Y>(hdr, nb, ls, bonds);
  typename Y::StructType st;
  st.set_struct(nb, ls, bonds.data());
  auto* t = new BlockTensor<S, Y>();
  // Every allowed block gets storage first (block_tensor_inst.cpp:121-159),
  // then absent blocks are reset to the null state.
  std::vector<int> angm(nb), bdim(nb), cdim(nb);
  long long tot = 0;
  for (int b = 0; b < st.get_nten(); ++b) {
    bool ok;
    if constexpr (Y::has_cgc) ok = st.get_tensor_argument(b, angm.data(), bdim.data(), cdim.data());
    else ok = st.get_tensor_argument(b, angm.data(), bdim.data());
    if (!ok) continue;
    long long sz = 1;
    for (int i = 0; i < nb; ++i) sz *= bdim[i];
    tot += sz;
  }
  std::vector<S> buf(std::max<long long>(tot, 1));
  long long off = 0, src = 0;
  for (int b = 0; b < st.get_nten(); ++b) {
    bool ok;
    if constexpr (Y::has_cgc) ok = st.get_tensor_argument(b, angm.data(), bdim.data(), cdim.data());
    else ok = st.get_tensor_argument(b, angm.data(), bdim.data());
    if (!ok) continue;
    long long sz = 1;
    for (int i = 0; i < nb; ++i) sz *= bdim[i];
    if (present[b]) {
      for (long long e = 0; e < sz; ++e)
        buf[off + e] = ScalarIO<S>::get(data + (src + e) * ScalarIO<S>::width);
      src += sz;
    }
    off += sz;
  }
  t->set(st, buf.data());
  for (int b = 0; b < t->get_nten(); ++b) {
    bool ok;
    if constexpr (Y::has_cgc) ok = st.get_tensor_argument(b, angm.data(), bdim.data(), cdim.data());
    else ok = st.get_tensor_argument(b, angm.data(), bdim.data());
    if (!ok) continue;
    if (!present[b]) {
      t->get_data_block(b)->clean();
      if constexpr (Y::has_cgc) t->get_cgc_block(b)->clean();
    } else if constexpr (Y::has_cgc) {
      if (nb == 2 && cgc2_kind == 1) t->get_cgc_block(b)->make_singlet(angm[0]);
    }
  }
  return t;
}

template <class S, class Y>
int tensor_header(BlockTensor<S, Y>* t, int* hdr, int cap) {
  std::vector<int> h;
  h.push_back(t->get_nbond());
  h.push_back(t->get_locspin());
  for (int i = 0; i < t->get_nbond(); ++i) {
    int nm = t->get_nmoment(i);
    h.push_back(t->get_bonddir(i));
    h.push_back(nm);
    for (int j = 0; j < nm; ++j) h.push_back(t->get_angularmoment(i, j));
    for (int j = 0; j < nm; ++j) h.push_back(t->get_bonddim(i, j));
  }
  if (hdr) {
    if ((int)h.size() > cap) return -1;
    std::copy(h.begin(), h.end(), hdr);
  }
  return (int)h.size();
}

template <class S, class Y>
long long tensor_get(BlockTensor<S, Y>* t, unsigned char* present, double* data, int* cgc_kind) {
  int nb = t->get_nbond();
  if (nb == 0) return 0;
  std::vector<int> angm(nb), bdim(nb), cdim(nb);
  long long off = 0;
  for (int b = 0; b < t->get_nten(); ++b) {
    bool ok = t->get_tensor_argument(b, angm.data(), bdim.data(), cdim.data());
    if (present) present[b] = 0;
    if (cgc_kind) cgc_kind[b] = -1;
    if (!ok || t->is_null(b)) continue;
    Dense<S>* d = t->get_data_block(b);
    long long sz = 1;
    for (int i = 0; i < nb; ++i) sz *= bdim[i];
    if (sz != d->get_nelem()) return -2;
    if (present) present[b] = 1;
    if (data)
      for (long long e = 0; e < sz; ++e)
        ScalarIO<S>::put(d->get_elem((int)e), data + (off + e) * ScalarIO<S>::width);
    if constexpr (Y::has_cgc) {
      if (cgc_kind && nb == 2) {
        Dense<double>* c = t->get_cgc_block(b);
        Dense<double> id, sg;
        id.make_identity(angm[0]);
        sg.make_singlet(angm[0]);
        cgc_kind[b] = 9;
        if (!c->check_null() && c->get_nelem() == id.get_nelem()) {
          bool is_id = true, is_sg = true;
          for (int e = 0; e < id.get_nelem(); ++e) {
            if (std::fabs(c->get_elem(e) - id.get_elem(e)) > 1e-12) is_id = false;
            if (std::fabs(c->get_elem(e) - sg.get_elem(e)) > 1e-12) is_sg = false;
          }
          cgc_kind[b] = is_id ? 0 : (is_sg ? 1 : 9);
        }
      }
    }
    off += sz;
  }
  return off;
}

#define DISPATCH(h, ...)                                                         \
  do {                                                                           \
    if ((h)->sym == 1 && (h)->cplx == 0) { using S = double; using Y = SU2Symmetry; __VA_ARGS__; } \
    else if ((h)->sym == 1 && (h)->cplx == 1) { using S = dcmplex; using Y = SU2Symmetry; __VA_ARGS__; } \
    else if ((h)->sym == 0 && (h)->cplx == 0) { using S = double; using Y = U1Symmetry; __VA_ARGS__; } \
    else { using S = dcmplex; using Y = U1Symmetry; __VA_ARGS__; }              \
  } while (0)

struct RefDMRG {
  int sym, cplx;
  void* d;
};

template <class S, class Y> DMRG<S, Y>* asd(RefDMRG* h) { return static_cast<DMRG<S, Y>*>(h->d); }

}  // namespace

extern "C" {

const char* ref_last_error() { return g_err.c_str(); }

void ref_quiet(int on) {
  if (on && !g_saved_cout) {
    g_saved_cout = std::cout.rdbuf(g_sink.rdbuf());
  } else if (!on && g_saved_cout) {
    std::cout.rdbuf(g_saved_cout);
    g_saved_cout = nullptr;
  }
  g_sink.str("");
}

// Globals normally set by main/main_dmrg_*.cpp:32-35 and :188-197.
// MKL threading before the first MKL call of the process: layer 1
// sequential (the default: the reference's own OpenMP over bond terms is the
// parallelism), 3 GNU OpenMP with `threads` BLAS/LAPACK threads.
namespace {
bool g_blas_chosen = false;
}
void ref_set_blas_threading(int layer, int threads) {
  MKL_Set_Threading_Layer(layer);
  if (threads > 0) MKL_Set_Num_Threads(threads);
  g_blas_chosen = true;
}

int ref_init(int physpn, int max_angm, int max_dcut_, int psize_) {
  if (!g_blas_chosen) MKL_Set_Threading_Layer(1);
  max_dcut = max_dcut_;
  psize = psize_ > 0 ? psize_ : 1;
  omp_set_num_threads(psize);
  if (physpn > 0) {
    delete g_cgc_tables;
    g_cgc_tables = new CGCTables();
    makeup_clebsch_gordan_coefficient_tensors(physpn, max_angm, *g_cgc_tables);
  }
  return 0;
}

void ref_set_max_dcut(int v) { max_dcut = v; }
void ref_set_psize(int v) { psize = v > 0 ? v : 1; omp_set_num_threads(psize); }

double ref_cg(int j1, int m1, int j2, int m2, int j3, int m3) { return CGCTable::cg(j1, m1, j2, m2, j3, m3); }

// Table layouts follow cgc_tables.hpp:18-49 flattened C-order over the
// pointer levels: hamilt_vec[i][j][k]; onsite/pairup[i][j][k][l][p];
// transformation[i][j][k][l][m][p].
int ref_fac_hamilt_vec(double* out) {
  if (!g_cgc_tables) return -1;
  int M = g_cgc_tables->max_angm;
  for (int i = 0; i < M; ++i)
    for (int j = 0; j < M; ++j)
      for (int k = 0; k < M; ++k) out[(i * M + j) * M + k] = g_cgc_tables->fac_hamilt_vec[i][j][k];
  return 0;
}

int ref_fac_table5(int which, double* out) {
  if (!g_cgc_tables) return -1;
  int M = g_cgc_tables->max_angm, P = g_cgc_tables->physpnsqr;
  double***** t = nullptr;
  switch (which) {
    case 0: t = g_cgc_tables->fac_operator_onsite_left; break;
    case 1: t = g_cgc_tables->fac_operator_onsite_rght; break;
    case 2: t = g_cgc_tables->fac_operator_pairup_left; break;
    case 3: t = g_cgc_tables->fac_operator_pairup_rght; break;
    default: return -2;
  }
  long long o = 0;
  for (int i = 0; i < M; ++i)
    for (int j = 0; j < M; ++j)
      for (int k = 0; k < M; ++k)
        for (int l = 0; l < M; ++l)
          for (int p = 0; p < P; ++p) out[o++] = t[i][j][k][l][p];
  return 0;
}

int ref_fac_transformation(int which, double* out) {
  if (!g_cgc_tables) return -1;
  int M = g_cgc_tables->max_angm, P = g_cgc_tables->physpn2;
  double****** t = which == 0 ? g_cgc_tables->fac_operator_transformation_left
                              : g_cgc_tables->fac_operator_transformation_rght;
  long long o = 0;
  for (int i = 0; i < M; ++i)
    for (int j = 0; j < M; ++j)
      for (int k = 0; k < M; ++k)
        for (int l = 0; l < M; ++l)
          for (int m = 0; m < M; ++m)
            for (int p = 0; p < P; ++p) out[o++] = t[i][j][k][l][m][p];
  return 0;
}

// ---- tensors ---------------------------------------------------------------
void* ref_tensor_new(int sym, int cplx, const int* hdr, const unsigned char* present,
                     const double* data, int cgc2_kind) {
  RefTensor tmp{sym, cplx, nullptr};
  RefTensor* out = nullptr;
  DISPATCH(&tmp, out = wrap(make_tensor<S, Y>(hdr, present, data, cgc2_kind)));
  return out;
}

void ref_tensor_free(void* h) {
  if (!h) return;
  RefTensor* r = static_cast<RefTensor*>(h);
  DISPATCH(r, delete as<S, Y>(r));
  delete r;
}

int ref_tensor_header(void* h, int* hdr, int cap) {
  RefTensor* r = static_cast<RefTensor*>(h);
  int n = 0;
  DISPATCH(r, n = tensor_header(as<S, Y>(r), hdr, cap));
  return n;
}

int ref_tensor_nten(void* h) {
  RefTensor* r = static_cast<RefTensor*>(h);
  int n = 0;
  DISPATCH(r, n = as<S, Y>(r)->get_nten());
  return n;
}

long long ref_tensor_get(void* h, unsigned char* present, double* data, int* cgc_kind) {
  RefTensor* r = static_cast<RefTensor*>(h);
  long long n = 0;
  DISPATCH(r, n = tensor_get(as<S, Y>(r), present, data, cgc_kind));
  return n;
}

void* ref_tensor_copy(void* h) {
  RefTensor* r = static_cast<RefTensor*>(h);
  RefTensor* out = nullptr;
  DISPATCH(r, out = wrap(as<S, Y>(r)->is_null() ? new BlockTensor<S, Y>() : new BlockTensor<S, Y>(*as<S, Y>(r))));
  return out;
}

// BlockTensor::hamiltonian_vector_multiplication (block_tensor.hpp:179)
void* ref_hvm(void* vec, void* op1, void* op2) {
  RefTensor* v = static_cast<RefTensor*>(vec);
  RefTensor* out = nullptr;
  DISPATCH(v, {
    auto* t = new BlockTensor<S, Y>();
    t->hamiltonian_vector_multiplication(*as<S, Y>(v), *as<S, Y>((RefTensor*)op1), *as<S, Y>((RefTensor*)op2));
    out = wrap(t);
  });
  return out;
}

// BlockTensor::contract (block_tensor.hpp:142)
void* ref_contract(void* t1, int i1, void* t2, int i2) {
  RefTensor* a = static_cast<RefTensor*>(t1);
  RefTensor* out = nullptr;
  DISPATCH(a, {
    auto* t = new BlockTensor<S, Y>();
    t->contract(*as<S, Y>(a), i1, *as<S, Y>((RefTensor*)t2), i2);
    out = wrap(t);
  });
  return out;
}

// Environment updates (block_tensor.hpp:163-176).  kind: 0 overlap_initial,
// 1 overlap_transformation, 2 operator_initial, 3 operator_transformation,
// 4 operator_pairup.
void* ref_env(int kind, void* uu, void* vv, void* op1, void* op2, int flag) {
  RefTensor* u = static_cast<RefTensor*>(uu);
  RefTensor* out = nullptr;
  DISPATCH(u, {
    auto* t = new BlockTensor<S, Y>();
    auto& U = *as<S, Y>(u);
    auto& V = *as<S, Y>((RefTensor*)vv);
    switch (kind) {
      case 0: t->overlap_initial(U, V, flag); break;
      case 1: t->overlap_transformation(U, V, *as<S, Y>((RefTensor*)op1), flag); break;
      case 2: t->operator_initial(U, V, *as<S, Y>((RefTensor*)op1), flag); break;
      case 3: t->operator_transformation(U, V, *as<S, Y>((RefTensor*)op1), flag); break;
      case 4: t->operator_pairup(U, V, *as<S, Y>((RefTensor*)op1), *as<S, Y>((RefTensor*)op2), flag); break;
    }
    out = wrap(t);
  });
  return out;
}

// BlockTensor::svd (block_tensor.hpp:135).  tu/tv are updated in place.
int ref_svd(void* vec, void* tu, double p1, void* tv, double p2, double* wout) {
  RefTensor* v = static_cast<RefTensor*>(vec);
  DISPATCH(v, as<S, Y>(v)->svd(*as<S, Y>((RefTensor*)tu), p1, *as<S, Y>((RefTensor*)tv), p2, wout));
  return 0;
}

// BlockTensor::fuse(BondType, BondType) (block_tensor.hpp:129)
void* ref_fuse(int sym, int cplx, const int* b1, const int* b2) {
  RefTensor tmp{sym, cplx, nullptr};
  RefTensor* out = nullptr;
  DISPATCH(&tmp, {
    int nb, ls;
    std::vector<typename Y::BondType> x, y;
    parse_bonds<Y>(b1, nb, ls, x);
    parse_bonds<Y>(b2, nb, ls, y);
    auto* t = new BlockTensor<S, Y>();
    t->fuse(x[0], y[0]);
    out = wrap(t);
  });
  return out;
}

int ref_tensor_op(void* h, int op, double a, double b) {
  RefTensor* r = static_cast<RefTensor*>(h);
  DISPATCH(r, {
    auto& T = *as<S, Y>(r);
    switch (op) {
      case 0: T.left2right_vectran(); break;
      case 1: T.right2left_vectran(); break;
      case 2: T.take_conjugate((int)a); break;
      case 3: T.makeup_input_vector(); break;
      case 4: T *= a; break;
      case 5: T.normalize_vector(); break;
      case 6: T.initialize_input_vector(); break;
      case 7: T.shift((int)a, (int)b); break;
      case 8: T = a; break;
      case 9: T *= dcmplex(a, b); break;
    }
  });
  return 0;
}

int ref_tensor_binop(void* h, int op, void* o) {
  RefTensor* r = static_cast<RefTensor*>(h);
  DISPATCH(r, {
    auto& T = *as<S, Y>(r);
    auto& O = *as<S, Y>((RefTensor*)o);
    if (op == 0) T += O;
    else if (op == 1) T -= O;
    else T = O;
  });
  return 0;
}

int ref_inner_prod_c(void* x, void* y, double* out2) {
  RefTensor* r = static_cast<RefTensor*>(x);
  DISPATCH(r, {
    S v = as<S, Y>(r)->inner_prod_c(*as<S, Y>((RefTensor*)y));
    dcmplex c(v);
    out2[0] = c.real();
    out2[1] = c.imag();
  });
  return 0;
}

// ---- DMRG driver -----------------------------------------------------------
void* ref_dmrg_new_exci(int sym, int cplx, int ly, int lx, int bondd, int physpn, int totspin, int exci) {
  RefDMRG* h = new RefDMRG{sym, cplx, nullptr};
  DISPATCH(h, {
    auto* d = new DMRG<S, Y>(ly, lx, bondd, physpn, totspin, exci);
    // BlockTensor::svd writes kept + discarded values (block_tensor_inst.cpp:589-600)
    // into ww[exci], which the constructor sizes to bondd (dmrg_inst.cpp:82-85),
    // and save_ww reads bondd*phdim of it; give it room for both.
    for (int e = 0; e <= d->max_exci; ++e) {
      delete[] d->ww[e];
      d->ww[e] = new double[4 * bondd + 4096]();
    }
    h->d = d;
  });
  return h;
}

// DMRG(ly, lx, bondd, physpn, totspin, max_exci = 0)
void* ref_dmrg_new(int sym, int cplx, int ly, int lx, int bondd, int physpn, int totspin) {
  return ref_dmrg_new_exci(sym, cplx, ly, lx, bondd, physpn, totspin, 0);
}

void ref_dmrg_free(void* p) {
  RefDMRG* h = static_cast<RefDMRG*>(p);
  DISPATCH(h, delete asd<S, Y>(h));
  delete h;
}

int ref_dmrg_ns(void* p) {
  RefDMRG* h = static_cast<RefDMRG*>(p);
  int n = 0;
  DISPATCH(h, n = asd<S, Y>(h)->ns);
  return n;
}

// Replace the hard-coded lattice map (dmrg_inst.cpp:117-160) and recompute
// fll/frr with the constructor's own rule (dmrg_inst.cpp:163-176).
int ref_dmrg_set_couplings(void* p, const int* hmap, const double* coup, int ncoup) {
  RefDMRG* h = static_cast<RefDMRG*>(p);
  DISPATCH(h, {
    auto* d = asd<S, Y>(h);
    int ns = d->ns;
    if (ncoup > ns + 1) return -1;
    for (int i = 0; i < ns; ++i)
      for (int j = 0; j < ns; ++j) d->hmap[i][j] = hmap[i * ns + j];
    for (int i = 0; i <= ns; ++i) d->coup[i] = i < ncoup ? coup[i] : 0.0;
    for (int i = 0; i < ns; ++i) {
      for (int j = 0; j < ns; ++j) { d->fll[i][j] = 0; d->frr[i][j] = 0; }
      for (int j = 0; j < i; ++j)
        for (int k = i + 1; k < ns; ++k)
          if (d->hmap[j][k] != 0) d->fll[i][j] = 1;
      for (int j = i + 1; j < ns; ++j)
        for (int k = 0; k < i; ++k)
          if (d->hmap[j][k] != 0) d->frr[i][j] = 1;
    }
  });
  return 0;
}

int ref_dmrg_get_couplings(void* p, int* hmap, double* coup) {
  RefDMRG* h = static_cast<RefDMRG*>(p);
  DISPATCH(h, {
    auto* d = asd<S, Y>(h);
    int ns = d->ns;
    for (int i = 0; i < ns; ++i)
      for (int j = 0; j < ns; ++j) hmap[i * ns + j] = d->hmap[i][j];
    for (int i = 0; i <= ns; ++i) coup[i] = d->coup[i];
  });
  return 0;
}

int ref_dmrg_idmrg(void* p) {
  RefDMRG* h = static_cast<RefDMRG*>(p);
  DISPATCH(h, asd<S, Y>(h)->do_idmrg());
  return 0;
}

int ref_dmrg_sweep(void* p) {
  RefDMRG* h = static_cast<RefDMRG*>(p);
  DISPATCH(h, asd<S, Y>(h)->sweep());
  return 0;
}

double ref_dmrg_energy(void* p) {
  RefDMRG* h = static_cast<RefDMRG*>(p);
  double e = 0;
  DISPATCH(h, e = asd<S, Y>(h)->get_enr());
  return e;
}

// which: 0 uu[i], 1 hh[i], 2 opr[i][j]; for stored state j: 3 orth[j][i],
// 4 ovlp[j][i], 5 overlapvec[j] (dmrg.hpp:64-67).
extern "C++" {
template <class S, class Y>
static BlockTensor<S, Y>* dmrg_member(DMRG<S, Y>* d, int which, int i, int j) {
  switch (which) {
    case 0: return &d->uu[i];
    case 1: return &d->hh[i];
    case 2: return &d->opr[i][j];
    case 3: return &d->orth[j][i];
    case 4: return &d->ovlp[j][i];
    case 6: return &d->sigma[i];  // local spin operators (dmrg_inst.cpp:222-260)
    default: return &d->overlapvec[j];
  }
}
}

// Returns a deep copy.
void* ref_dmrg_tensor(void* p, int which, int i, int j) {
  RefDMRG* h = static_cast<RefDMRG*>(p);
  RefTensor* out = nullptr;
  DISPATCH(h, {
    auto* d = asd<S, Y>(h);
    BlockTensor<S, Y>* src = dmrg_member(d, which, i, j);
    // copying a null tensor exits inside SU2Struct::set_su2struct (nbond < 1)
    out = wrap(src->is_null() ? new BlockTensor<S, Y>() : new BlockTensor<S, Y>(*src));
  });
  return out;
}

int ref_dmrg_set_tensor(void* p, int which, int i, int j, void* t) {
  RefDMRG* h = static_cast<RefDMRG*>(p);
  DISPATCH(h, {
    auto* d = asd<S, Y>(h);
    BlockTensor<S, Y>* dst = dmrg_member(d, which, i, j);
    *dst = *as<S, Y>((RefTensor*)t);
  });
  return 0;
}

// gs_enr[j] (dmrg.hpp:77): the energy of stored state j, the penalty weight
double ref_dmrg_state_energy(void* p, int j) {
  RefDMRG* h = static_cast<RefDMRG*>(p);
  double e = 0;
  DISPATCH(h, e = asd<S, Y>(h)->gs_enr[j]);
  return e;
}

// Individual sweep-step methods (dmrg_inst.cpp:874-943 calls them in order).
// code: 0 wavefunc_transformation(a, b); 1 prepare_site_operator_from_left(a, b);
// 2 prepare_site_operator_from_right(a, b); 3 prepare_site_operator_from_left(a);
// 4 prepare_site_operator_from_right(a); 5 prepare_sweep().
int ref_dmrg_call(void* p, int code, int a, int b) {
  RefDMRG* h = static_cast<RefDMRG*>(p);
  DISPATCH(h, {
    auto* d = asd<S, Y>(h);
    switch (code) {
      case 0: d->wavefunc_transformation(a, b); break;
      case 1: d->prepare_site_operator_from_left(a, b); break;
      case 2: d->prepare_site_operator_from_right(a, b); break;
      case 3: d->prepare_site_operator_from_left(a); break;
      case 4: d->prepare_site_operator_from_right(a); break;
      case 5: d->prepare_sweep(); break;
    }
  });
  return 0;
}

void* ref_dmrg_input_vector(void* p, int il, int ir) {
  RefDMRG* h = static_cast<RefDMRG*>(p);
  RefTensor* out = nullptr;
  DISPATCH(h, {
    auto* t = new BlockTensor<S, Y>();
    asd<S, Y>(h)->prepare_input_vector(il, ir, *t);
    out = wrap(t);
  });
  return out;
}

// DMRG::hamiltonian_vector_multiplication_idmrg (dmrg_inst.cpp:283-371)
void* ref_dmrg_heff(void* p, int il, int ir, void* vin) {
  RefDMRG* h = static_cast<RefDMRG*>(p);
  RefTensor* out = nullptr;
  DISPATCH(h, {
    auto* t = new BlockTensor<S, Y>();
    asd<S, Y>(h)->hamiltonian_vector_multiplication_idmrg(il, ir, *as<S, Y>((RefTensor*)vin), *t);
    out = wrap(t);
  });
  return out;
}

// DMRG::lanczos_solve_eigenvector_idmrg (dmrg_inst.cpp:264-279); vec updated in place.
double ref_dmrg_lanczos(void* p, int il, int ir, void* vec) {
  RefDMRG* h = static_cast<RefDMRG*>(p);
  double e = 0;
  DISPATCH(h, {
    auto* d = asd<S, Y>(h);
    d->lanczos_solve_eigenvector_idmrg(il, ir, *as<S, Y>((RefTensor*)vec));
    e = d->gs_enr[d->exci];
  });
  return e;
}

// vec.svd(uu[a], p1, uu[b], p2, ww[exci]) as in dmrg_inst.cpp:894-935; the
// singular values are copied to wout (capacity cap).
int ref_dmrg_svd(void* p, void* vec, int a, int b, double p1, double p2, double* wout, int cap) {
  RefDMRG* h = static_cast<RefDMRG*>(p);
  DISPATCH(h, {
    auto* d = asd<S, Y>(h);
    std::vector<double> w(std::max(cap, 4 * d->bondd + 4096), 0.0);
    as<S, Y>((RefTensor*)vec)->svd(d->uu[a], p1, d->uu[b], p2, w.data());
    for (int k = 0; k < cap; ++k) wout[k] = w[k];
  });
  return 0;
}

// Times `reps` applications of hamiltonian_vector_multiplication_idmrg at
// (il, ir) and returns seconds per application.
double ref_dmrg_time_heff(void* p, int il, int ir, void* vin, int reps) {
  RefDMRG* h = static_cast<RefDMRG*>(p);
  double sec = 0;
  DISPATCH(h, {
    auto* d = asd<S, Y>(h);
    auto& v = *as<S, Y>((RefTensor*)vin);
    BlockTensor<S, Y> out;
    auto t0 = std::chrono::steady_clock::now();
    for (int r = 0; r < reps; ++r) d->hamiltonian_vector_multiplication_idmrg(il, ir, v, out);
    auto t1 = std::chrono::steady_clock::now();
    sec = std::chrono::duration<double>(t1 - t0).count() / reps;
  });
  return sec;
}

// Number of uniform(-1, 1) draws the reference's global generator
// (dense_inst.cpp:23-24) has produced so far: draws one more value through
// Dense::random_init and locates it in a fresh mt19937_64(42) stream.
// Returns -1 if not found within maxn draws.  Consumes one draw.
long long ref_rng_position(long long maxn) {
  int one = 1;
  Dense<double> t;
  t.alloc(1, &one);
  t.random_init();
  const double v = t.getptr()[0];
  std::mt19937_64 g(42);
  std::uniform_real_distribution<double> dist(-1.0, 1.0);
  for (long long k = 0; k < maxn; ++k)
    if (dist(g) == v) return k;
  return -1;
}

// The reference's checkpoint I/O writes and reads relative paths in the
// current directory; these run it with the working directory set to dir.
// DMRG::save_mps2(exci), save_ww(), save_enr() (dmrg_inst.cpp:940-942, 947-1003,
// 1146-1153).
int ref_dmrg_save(void* p, const char* dir, int exci) {
  RefDMRG* h = static_cast<RefDMRG*>(p);
  char cwd[4096];
  if (!getcwd(cwd, sizeof(cwd)) || chdir(dir) != 0) return -1;
  DISPATCH(h, {
    auto* d = asd<S, Y>(h);
    d->save_mps2(exci);
    d->save_ww();
    d->save_enr();
  });
  return chdir(cwd) == 0 ? 0 : -1;
}

// DMRG::read_mps(read1, read2) (dmrg_inst.cpp:1078-1141); returns its pass flag.
int ref_dmrg_read(void* p, const char* dir, int read1, int read2) {
  RefDMRG* h = static_cast<RefDMRG*>(p);
  char cwd[4096];
  if (!getcwd(cwd, sizeof(cwd)) || chdir(dir) != 0) return -1;
  bool pass = false;
  DISPATCH(h, pass = asd<S, Y>(h)->read_mps(read1, read2));
  if (chdir(cwd) != 0) return -1;
  return pass ? 1 : 0;
}

// First cap entries of ww[exci] (the singular values of the last truncation).
int ref_dmrg_ww(void* p, double* out, int cap) {
  RefDMRG* h = static_cast<RefDMRG*>(p);
  DISPATCH(h, {
    auto* d = asd<S, Y>(h);
    for (int k = 0; k < cap; ++k) out[k] = d->ww[d->exci][k];
  });
  return 0;
}

}  // extern "C"
