// cgc.cpp -- see cgc.hpp.  Host-side, evaluated once per distinct entry.
#include "cgc.hpp"
#include <cmath>
#include <cstdlib>
#include "core.hpp"

namespace sadmrg {

namespace {

double fact(int n) {
  double f = 1.0;
  for (int i = 2; i <= n; ++i) f *= i;
  return f;
}

// Small column-major host tensor used only to evaluate table entries.
struct HT {
  std::vector<int> d;
  std::vector<double> v;
  int64_t size() const {
    int64_t n = 1;
    for (int x : d) n *= x;
    return n;
  }
};

HT make_cgc(int j1, int j2, int j3) {
  HT t;
  t.d = {j1 + 1, j2 + 1, j3 + 1};
  t.v.assign(t.size(), 0.0);
  for (int i = 0; i <= j1; ++i)
    for (int j = 0; j <= j2; ++j)
      for (int k = 0; k <= j3; ++k)
        t.v[i + j * (j1 + 1) + k * (j1 + 1) * (j2 + 1)] =
            clebsch_gordan(j1, 2 * i - j1, j2, 2 * j - j2, j3, 2 * k - j3);
  return t;
}

// 2-leg spin singlet: S(p, a) = (-1)^(j - p) / sqrt(j+1) when m_p + m_a = 0.
HT make_singlet(int j) {
  HT t;
  t.d = {j + 1, j + 1};
  t.v.assign(t.size(), 0.0);
  for (int p = 0; p <= j; ++p) {
    const int a = j - p;
    const double s = ((j - p) % 2 == 0) ? 1.0 : -1.0;
    t.v[p + a * (j + 1)] = s / std::sqrt(static_cast<double>(j + 1));
  }
  return t;
}

HT make_identity(int j) {
  HT t;
  t.d = {j + 1, j + 1};
  t.v.assign(t.size(), 0.0);
  for (int p = 0; p <= j; ++p) t.v[p + p * (j + 1)] = 1.0;
  return t;
}

inline double at3(const HT& t, int a, int b, int c) { return t.v[a + t.d[0] * (b + t.d[1] * c)]; }

// R(c,y,c') = sum t1(a,b,c) t3(a,y,z) t2(z,b,c')   [pattern "A"]
// R(c,y,c') = sum t1(a,b,c) t3(b,y,z) t2(a,z,c')   [pattern "B"]
HT trilinear(const HT& t1, const HT& t2, const HT& t3, bool pattern_b) {
  HT r;
  r.d = {t1.d[2], t3.d[1], t2.d[2]};
  r.v.assign(r.size(), 0.0);
  for (int c = 0; c < t1.d[2]; ++c)
    for (int y = 0; y < t3.d[1]; ++y)
      for (int cp = 0; cp < t2.d[2]; ++cp) {
        double s = 0;
        for (int a = 0; a < t1.d[0]; ++a)
          for (int b = 0; b < t1.d[1]; ++b)
            for (int z = 0; z < t3.d[2]; ++z) {
              if (!pattern_b) {
                if (a >= t3.d[0] || z >= t2.d[0] || b >= t2.d[1]) continue;
                s += at3(t1, a, b, c) * at3(t3, a, y, z) * at3(t2, z, b, cp);
              } else {
                if (b >= t3.d[0] || a >= t2.d[0] || z >= t2.d[1]) continue;
                s += at3(t1, a, b, c) * at3(t3, b, y, z) * at3(t2, a, z, cp);
              }
            }
        r.v[c + r.d[0] * (y + r.d[1] * cp)] = s;
      }
  return r;
}

// pairup flag 0: R(c,c') = sum uu(a,b,c) op1(a,y,z) op2(b,y,r) vv(z,r,c')
// pairup flag 1: R(c,c') = sum uu(a,b,c) op1(b,y,z) op2(a,y,r) vv(r,z,c')
HT pairup_contract(const HT& uu, const HT& vv, const HT& op1, const HT& op2, int flag) {
  HT r;
  r.d = {uu.d[2], vv.d[2]};
  r.v.assign(r.size(), 0.0);
  for (int c = 0; c < uu.d[2]; ++c)
    for (int cp = 0; cp < vv.d[2]; ++cp) {
      double s = 0;
      for (int a = 0; a < uu.d[0]; ++a)
        for (int b = 0; b < uu.d[1]; ++b) {
          const double u = at3(uu, a, b, c);
          if (u == 0.0) continue;
          for (int y = 0; y < op1.d[1]; ++y)
            for (int z = 0; z < op1.d[2]; ++z)
              for (int rr = 0; rr < op2.d[2]; ++rr) {
                if (flag == 0) {
                  if (a >= op1.d[0] || b >= op2.d[0] || y >= op2.d[1] || z >= vv.d[0] || rr >= vv.d[1]) continue;
                  s += u * at3(op1, a, y, z) * at3(op2, b, y, rr) * at3(vv, z, rr, cp);
                } else {
                  if (b >= op1.d[0] || a >= op2.d[0] || y >= op2.d[1] || rr >= vv.d[0] || z >= vv.d[1]) continue;
                  s += u * at3(op1, b, y, z) * at3(op2, a, y, rr) * at3(vv, rr, z, cp);
                }
              }
        }
      r.v[c + r.d[0] * cp] = s;
    }
  return r;
}

// Proportionality factor of r against o, with the reference's rules
// (src/dense_inst.cpp:800-844): shape mismatch -> entry left at 0; r all
// below 1e-12 -> 0; otherwise r/o at the first position of max |o|.
double ratio(const HT& r, const HT& o) {
  if (r.d != o.d) return 0.0;
  const double tol = 1e-12;
  bool zero = true;
  for (double x : r.v)
    if (std::fabs(x) > tol) { zero = false; break; }
  if (zero) return 0.0;
  int64_t pos = 0;
  double mx = 0;
  for (int64_t i = 0; i < static_cast<int64_t>(o.v.size()); ++i)
    if (std::fabs(o.v[i]) > mx) { mx = std::fabs(o.v[i]); pos = i; }
  if (mx <= tol) fail(kErrNumerics, "CGC ratio against a zero tensor");
  return r.v[pos] / o.v[pos];
}

inline bool in_range(int x, int lo, int hi) { return x >= lo && x <= hi; }

uint64_t make_key(int table, const int* a, int n) {
  uint64_t k = static_cast<uint64_t>(table);
  for (int i = 0; i < n; ++i) k = k * 131u + static_cast<uint64_t>(a[i] + 7);
  return k;
}

}  // namespace

double clebsch_gordan(int j1, int m1, int j2, int m2, int j3, int m3) {
  if (m1 + m2 != m3) return 0.0;
  if (j3 < std::abs(j1 - j2) || j3 > j1 + j2) return 0.0;
  if ((j1 + j2 + j3) % 2 != 0) return 0.0;
  if (std::abs(m1) > j1 || std::abs(m2) > j2 || std::abs(m3) > j3) return 0.0;
  const double pre = std::sqrt((j3 + 1) * fact((j3 + j1 - j2) / 2) * fact((j3 - j1 + j2) / 2) *
                               fact((j1 + j2 - j3) / 2) / fact((j1 + j2 + j3) / 2 + 1));
  const double f1 = fact((j1 + m1) / 2) * fact((j1 - m1) / 2);
  const double f2 = fact((j2 + m2) / 2) * fact((j2 - m2) / 2);
  const double f3 = fact((j3 + m3) / 2) * fact((j3 - m3) / 2);
  const double mf = std::sqrt(f1 * f2 * f3);
  const int a1 = (j1 + j2 - j3) / 2, a2 = (j1 - m1) / 2, a3 = (j2 + m2) / 2;
  const int b1 = (j3 - j2 + m1) / 2, b2 = (j3 - j1 - m2) / 2;
  double sum = 0;
  for (int k = 0; k <= (j1 + j2 + j3) / 2; ++k) {
    if (a1 - k < 0 || a2 - k < 0 || a3 - k < 0 || b1 + k < 0 || b2 + k < 0) continue;
    const double den = fact(k) * fact(a1 - k) * fact(a2 - k) * fact(a3 - k) * fact(b1 + k) * fact(b2 + k);
    sum += ((k % 2) ? -1.0 : 1.0) / den;
  }
  return pre * mf * sum;
}

CgcTables::CgcTables(int physpn) : physpn_(physpn) {
  require(physpn >= 1, "CGC tables need physpn >= 1");
}

double CgcTables::lookup(uint64_t key, double (CgcTables::*fn)(const int*), const int* args) {
  {
    std::lock_guard<std::mutex> g(mu_);
    auto it = cache_.find(key);
    if (it != cache_.end()) return it->second;
  }
  const double v = (this->*fn)(args);
  std::lock_guard<std::mutex> g(mu_);
  cache_[key] = v;
  return v;
}

double CgcTables::hamilt_vec(int i, int j, int k) {
  const int a[3] = {i, j, k};
  return lookup(make_key(1, a, 3), &CgcTables::eval_hamilt, a);
}

double CgcTables::onsite(int side, int j, int m, int l, int i, int pidx) {
  const int a[6] = {side, j, m, l, i, pidx};
  return lookup(make_key(2, a, 6), &CgcTables::eval_onsite, a);
}

double CgcTables::transformation(int side, int j, int m, int l, int i, int k, int p0) {
  const int a[7] = {side, j, m, l, i, k, p0};
  return lookup(make_key(3, a, 7), &CgcTables::eval_transformation, a);
}

double CgcTables::pairup(int side, int j, int i, int k, int m, int pidx) {
  const int a[6] = {side, j, i, k, m, pidx};
  return lookup(make_key(4, a, 6), &CgcTables::eval_pairup, a);
}

// fac = sqrt(j+1) * < C_ijk , C_ijk x S_i x S_j x S_k >   (cgc_tables.cpp:156-163)
double CgcTables::eval_hamilt(const int* a) {
  const int i = a[0], j = a[1], k = a[2];
  if (i < 0 || j < 0 || k < 0) return 0.0;
  if (!(k >= std::abs(i - j) && k <= i + j && (i % 2 == k % 2))) return 0.0;
  const HT c = make_cgc(i, j, k);
  const HT si = make_singlet(i), sj = make_singlet(j), sk = make_singlet(k);
  double s = 0;
  for (int p = 0; p <= i; ++p)
    for (int q = 0; q <= j; ++q)
      for (int r = 0; r <= k; ++r) {
        // each singlet has one nonzero per row: S(p, i-p)
        const double t = at3(c, i - p, j - q, k - r) * si.v[p + (i - p) * (i + 1)] *
                         sj.v[q + (j - q) * (j + 1)] * sk.v[r + (k - r) * (k + 1)];
        s += at3(c, p, q, r) * t;
      }
  return s * std::sqrt(static_cast<double>(j + 1));
}

double CgcTables::eval_onsite(const int* a) {
  const int side = a[0], j = a[1], m = a[2], l = a[3], i = a[4], pidx = a[5];
  const int P2 = p2();
  if (pidx < 0 || pidx >= P2 * P2 || m < 0 || m > 6 || (m % 2) != 0 || i < 0 || j < 0 || l < 0) return 0.0;
  const int p0 = pidx % P2, p1 = pidx / P2, k = i;
  if (!in_range(j, std::abs(i - p0), i + p0)) return 0.0;
  if (!in_range(l, std::abs(k - p1), k + p1)) return 0.0;
  if (!(in_range(l, std::abs(j - m), j + m) && (j % 2 == l % 2))) return 0.0;
  if (!in_range(p1, std::abs(p0 - m), p0 + m)) return 0.0;
  const HT t3 = make_cgc(p0, m, p1);
  const HT target = make_cgc(j, m, l);
  if (side == 0) {
    // contract_dmrg_operator_initial(C[i][p0][j], C[k][p1][l], C[p0][m][p1], 0)
    return ratio(trilinear(make_cgc(i, p0, j), make_cgc(k, p1, l), t3, true), target);
  }
  // contract_dmrg_operator_initial(C[p0][i][j], C[p1][k][l], C[p0][m][p1], 1)
  return ratio(trilinear(make_cgc(p0, i, j), make_cgc(p1, k, l), t3, false), target);
}

double CgcTables::eval_transformation(const int* a) {
  const int side = a[0], j = a[1], m = a[2], l = a[3], i = a[4], k = a[5], p0 = a[6];
  const int P2 = p2();
  if (p0 < 0 || p0 >= P2 || m < 0 || m > 6 || (m % 2) != 0 || i < 0 || j < 0 || k < 0 || l < 0) return 0.0;
  const int p1 = p0;
  if (!in_range(j, std::abs(i - p0), i + p0)) return 0.0;
  if (!in_range(l, std::abs(k - p1), k + p1)) return 0.0;
  if (!(in_range(k, std::abs(i - m), i + m) && (k % 2 == i % 2) && in_range(l, std::abs(j - m), j + m) &&
        (j % 2 == l % 2)))
    return 0.0;
  const HT t3 = make_cgc(i, m, k);
  const HT target = make_cgc(j, m, l);
  if (side == 0)  // contract_dmrg_operator_transformation(C[i][p0][j], C[k][p1][l], C[i][m][k], 0)
    return ratio(trilinear(make_cgc(i, p0, j), make_cgc(k, p1, l), t3, false), target);
  // contract_dmrg_operator_transformation(C[p0][i][j], C[p1][k][l], C[i][m][k], 1)
  return ratio(trilinear(make_cgc(p0, i, j), make_cgc(p1, k, l), t3, true), target);
}

double CgcTables::eval_pairup(const int* a) {
  const int side = a[0], j = a[1], i = a[2], k = a[3], m = a[4], pidx = a[5];
  const int P2 = p2();
  if (pidx < 0 || pidx >= P2 * P2 || m < 0 || m > 6 || (m % 2) != 0 || i < 0 || j < 0 || k < 0) return 0.0;
  const int p0 = pidx % P2, p1 = pidx / P2, l = j;
  if (!in_range(j, std::abs(i - p0), i + p0)) return 0.0;
  if (!in_range(l, std::abs(k - p1), k + p1)) return 0.0;
  if (!(in_range(k, std::abs(i - m), i + m) && in_range(p1, std::abs(p0 - m), p0 + m))) return 0.0;
  // op2(a, p, c) = sqrt(m+1) * sum_b C[p0][m][p1](a, b, c) * S_m(p, b)
  const HT cpm = make_cgc(p0, m, p1);
  const HT sm = make_singlet(m);
  HT op2;
  op2.d = {p0 + 1, m + 1, p1 + 1};
  op2.v.assign(op2.size(), 0.0);
  for (int aa = 0; aa <= p0; ++aa)
    for (int p = 0; p <= m; ++p)
      for (int c = 0; c <= p1; ++c) {
        double s = 0;
        for (int b = 0; b <= m; ++b) s += at3(cpm, aa, b, c) * sm.v[p + b * (m + 1)];
        op2.v[aa + (p0 + 1) * (p + (m + 1) * c)] = s * std::sqrt(static_cast<double>(m + 1));
      }
  const HT op1 = make_cgc(i, m, k);
  const HT target = make_identity(j);
  if (side == 0)
    return ratio(pairup_contract(make_cgc(i, p0, j), make_cgc(k, p1, l), op1, op2, 0), target);
  return ratio(pairup_contract(make_cgc(p0, i, j), make_cgc(p1, k, l), op1, op2, 1), target);
}

}  // namespace sadmrg
