// lanczos.cpp -- see lanczos.hpp.
#include "lanczos.hpp"
#include <algorithm>
#include <atomic>
#include <cmath>
#include <limits>
#include "kernels.hpp"

namespace sadmrg {

// Implicit QL iteration for a symmetric tridiagonal matrix (the EISPACK tql2
// scheme: Wilkinson shift, chasing with plane rotations accumulated into z).
void tridiag_eig(int n, std::vector<double>& d, std::vector<double> e, std::vector<double>& z) {
  z.assign(static_cast<size_t>(n) * n, 0.0);
  for (int i = 0; i < n; ++i) z[i + i * n] = 1.0;
  e.resize(n, 0.0);
  if (n > 0) e[n - 1] = 0.0;
  const double eps = std::numeric_limits<double>::epsilon();
  for (int l = 0; l < n; ++l) {
    int iter = 0;
    int m;
    do {
      for (m = l; m < n - 1; ++m) {
        const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= eps * dd) break;
      }
      if (m != l) {
        if (++iter > 200) fail(kErrNumerics, "tridiagonal eigensolver did not converge");
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double r = std::hypot(g, 1.0);
        g = d[m] - d[l] + e[l] / (g + (g >= 0 ? std::fabs(r) : -std::fabs(r)));
        double s = 1.0, c = 1.0, p = 0.0;
        int i;
        bool underflow = false;
        for (i = m - 1; i >= l; --i) {
          double f = s * e[i];
          const double b = c * e[i];
          r = std::hypot(f, g);
          e[i + 1] = r;
          if (r == 0.0) {
            d[i + 1] -= p;
            e[m] = 0.0;
            underflow = true;
            break;
          }
          s = f / r;
          c = g / r;
          g = d[i + 1] - p;
          r = (d[i] - g) * s + 2.0 * c * b;
          p = s * r;
          d[i + 1] = g + p;
          g = c * r - b;
          for (int k = 0; k < n; ++k) {
            f = z[k + (i + 1) * n];
            z[k + (i + 1) * n] = s * z[k + i * n] + c * f;
            z[k + i * n] = c * z[k + i * n] - s * f;
          }
        }
        if (underflow) continue;
        d[l] -= p;
        e[l] = g;
        e[m] = 0.0;
      }
    } while (m != l);
  }
  // ascending order, eigenvectors with a non-negative leading component
  std::vector<int> idx(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return d[a] < d[b]; });
  std::vector<double> d2(n), z2(static_cast<size_t>(n) * n);
  for (int j = 0; j < n; ++j) {
    d2[j] = d[idx[j]];
    const double sg = z[0 + idx[j] * n] < 0 ? -1.0 : 1.0;
    for (int k = 0; k < n; ++k) z2[k + j * n] = sg * z[k + idx[j] * n];
  }
  d = d2;
  z = z2;
}

namespace {
std::atomic<bool> g_device_steps{true}, g_graph{true};
std::atomic<long long> g_replays{0}, g_captures{0}, g_graph_nodes{0};
}  // namespace

long long lanczos_graph_nodes() { return g_graph_nodes; }

void lanczos_graph_counts(long long* replays, long long* captures, bool reset) {
  if (replays) *replays = g_replays;
  if (captures) *captures = g_captures;
  if (reset) {
    g_replays = 0;
    g_captures = 0;
  }
}

void set_lanczos_mode(bool device_steps, bool graph) {
  g_device_steps = device_steps;
  g_graph = graph;
}

void lanczos_mode(bool* device_steps, bool* graph) {
  if (device_steps) *device_steps = g_device_steps;
  if (graph) *graph = g_graph;
}

namespace {

// One Krylov step captured in a HIP graph, destroyed with the solve.
struct StepGraph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~StepGraph() {
    if (exec) hipGraphExecDestroy(exec);
    if (graph) hipGraphDestroy(graph);
  }
};

void check_kernel_only(hipGraph_t g) {
  size_t nn = 0;
  SADMRG_HIP(hipGraphGetNodes(g, nullptr, &nn));
  std::vector<hipGraphNode_t> nodes(nn);
  if (nn) SADMRG_HIP(hipGraphGetNodes(g, nodes.data(), &nn));
  for (hipGraphNode_t nd : nodes) {
    hipGraphNodeType t;
    SADMRG_HIP(hipGraphNodeGetType(nd, &t));
    if (t != hipGraphNodeTypeKernel)
      fail(kErrHip, "lanczos: the captured Krylov step holds a non-kernel node (type " + std::to_string(int(t)) + ")");
  }
  g_graph_nodes = static_cast<long long>(nn);
}

struct Dev {
  Context* ctx;
  int64_t n;
  bool cx;
  int w() const { return cx ? 2 : 1; }

  dcmplex dot(const double* x, const double* y) {
    SADMRG_HIP(launch_gemv_t(x, 0, 1, y, n, cx ? 1 : 0, ctx->partial(), 1.0, ctx->dscal(), ctx->stream()));
    double h[2];
    SADMRG_HIP(hipMemcpyAsync(h, ctx->dscal(), sizeof(h), hipMemcpyDeviceToHost, ctx->stream()));
    ctx->sync();
    return dcmplex(h[0], h[1]);
  }
  double nrm(const double* x) { return std::sqrt(std::fabs(dot(x, x).real())); }
  void scal(double a, double* x) { SADMRG_HIP(launch_axpby(n * w(), a, 0.0, nullptr, x, ctx->stream())); }
  // y = a*y + b*x (real scalars act componentwise on complex data)
  void axpby(double a, double b, const double* x, double* y) {
    SADMRG_HIP(launch_axpby(n * w(), a, b, x, y, ctx->stream()));
  }
  void copy(const double* x, double* y) {
    SADMRG_HIP(hipMemcpyAsync(y, x, n * w() * 8, hipMemcpyDeviceToDevice, ctx->stream()));
  }
  // w -= V[:, 0..m) (V^H w), twice (classical Gram-Schmidt with re-orthogonalisation).
  // The reduction kernel takes at most kOrthPanel columns, so longer bases are
  // processed panel by panel (block Gram-Schmidt over the panels).
  static constexpr int kOrthPanel = 64;
  void orth(const double* V, int m, double* x) {
    if (m <= 0) return;
    for (int pass = 0; pass < 2; ++pass) {
      for (int c0 = 0; c0 < m; c0 += kOrthPanel) {
        const int mc = std::min(kOrthPanel, m - c0);
        const double* Vp = V + static_cast<int64_t>(c0) * n * w();
        SADMRG_HIP(launch_gemv_t(Vp, n, mc, x, n, cx ? 1 : 0, ctx->partial(), -1.0, ctx->dscal(), ctx->stream()));
        SADMRG_HIP(launch_gemv_n_update(Vp, n, mc, ctx->dscal(), x, n, cx ? 1 : 0, ctx->stream()));
      }
    }
  }
  // y = sum_j c_j V[:, j]
  void combine(const double* V, const std::vector<double>& c, double* y) {
    const int m = static_cast<int>(c.size());
    std::vector<double> h(2 * m, 0.0);
    for (int j = 0; j < m; ++j) h[2 * j] = c[j];
    SADMRG_HIP(hipMemcpyAsync(ctx->dscal(), h.data(), h.size() * 8, hipMemcpyHostToDevice, ctx->stream()));
    SADMRG_HIP(launch_fill(n * w(), 0.0, y, ctx->stream()));
    SADMRG_HIP(launch_gemv_n_update(V, n, m, ctx->dscal(), y, n, cx ? 1 : 0, ctx->stream()));
    ctx->sync();  // h leaves scope
  }
};

}  // namespace

LanczosOut lanczos_solve(HeffPlan& H, double* d_vec, const LanczosOpts& opt) {
  Context* ctx = H.ctx();
  Dev D{ctx, H.n(), H.complex()};
  const int64_t nw = D.n * D.w();
  const int ml = std::max(2, opt.mlanc);
  require(ml + 1 <= 2048, "lanczos: basis too large");
  require(Context::kMaxDscal >= 2 * (ml + 1), "lanczos: device scalar area too small");
  LanczosOut out;
  DevBuf Vb(ctx, std::max<int64_t>(nw, 1) * 8 * (ml + 1));
  DevBuf hvb(ctx, std::max<int64_t>(nw, 1) * 8);
  double* V = Vb.as<double>();
  auto col = [&](int j) { return V + static_cast<int64_t>(j) * nw; };
  double* hv = hvb.as<double>();
  const Prec bulk = opt.prec;
  auto matvec = [&](const double* x, double* y, Prec p) {
    H.apply(x, y, p);
    if (p == Prec::FP64) out.matvec64++;
    else out.matvec32++;
  };

  std::vector<double> aal(ml, 0.0), nnl(ml + 1, 0.0), eig, z;
  auto diatridiag = [&](int m) {
    std::vector<double> d(aal.begin(), aal.begin() + m), e(std::max(m - 1, 0));
    for (int i = 0; i + 1 < m; ++i) e[i] = nnl[i + 1];
    tridiag_eig(m, d, e, z);
    eig = d;
  };
  bool hv_valid = false;  // hv = H d_vec, from the last convergence check
  auto ritz = [&](int mc) {
    std::vector<double> c(mc);
    for (int j = 0; j < mc; ++j) c[j] = z[j];  // eigenvector 0, leading dimension mc
    D.combine(V, c, d_vec);
    hv_valid = false;
  };
  // check_eigenvector (dmrg_inst.cpp:1243-1254): returns |<v,Hv>| / |Hv|.
  // With refinement the check matvec is FP64: the stopping test then sees the
  // exact operator, and hv = H v is reused as the first residual below.
  const Prec check_prec = (bulk != Prec::FP64 && opt.refine) ? Prec::FP64 : bulk;
  // FP32 Ritz pairs judged by their recurrence residual, no FP64 matvec
  const bool est_only = bulk != Prec::FP64 && opt.refine && !opt.fp64_energy;
  out.energy64 = !est_only;
  // |r| = beta_m |y_m| of the lowest Ritz pair of the m-step recurrence
  // (needs diatridiag(m) and nnl[m])
  auto est_converged = [&](int m) {
    const double th = eig[0], rr = nnl[m] * std::fabs(z[m - 1]);
    out.ritz32 = th;
    return rr <= opt.refine_tol * std::fabs(th);
  };
  auto check = [&]() {
    const double nor = D.nrm(d_vec);
    D.scal(1.0 / nor, d_vec);
    matvec(d_vec, hv, check_prec);
    hv_valid = true;
    const double prod = D.nrm(hv);
    const dcmplex ov = D.dot(d_vec, hv);
    return std::abs(ov) / prod;
  };

  // initialize_lanczos (dmrg_inst.cpp:1201-1212)
  D.copy(d_vec, col(0));
  {
    const double nor = D.nrm(col(0));
    if (nor > 1e-12) D.scal(1.0 / nor, col(0));
  }
  const double tol = 1e-6;
  bool done = false;
  matvec(col(0), col(1), bulk);
  aal[0] = D.dot(col(1), col(0)).real();
  D.axpby(1.0, -aal[0], col(0), col(1));
  nnl[1] = D.nrm(col(1));
  D.scal(1.0 / nnl[1], col(1));
  out.steps = 1;
  if (ml == 2) {
    if (nnl[1] < tol) {
      diatridiag(1);
      ritz(1);
    } else {
      matvec(col(1), col(2), bulk);
      aal[1] = D.dot(col(2), col(1)).real();
      diatridiag(2);
      ritz(2);
    }
    done = true;
  }
  const bool dev_steps = !done && opt.device_steps && g_device_steps && ml <= kLanczosDevMax;
  if (dev_steps) {
    // Steps m = 2 .. ml-1 on the device: x = column m-1 (kept in xin), w = H x.
    DevBuf stb(ctx, sizeof(LanczosDev));
    DevBuf xb(ctx, std::max<int64_t>(nw, 1) * 8), wb(ctx, std::max<int64_t>(nw, 1) * 8);
    LanczosDev* st = stb.as<LanczosDev>();
    double* xin = xb.as<double>();
    double* w = wb.as<double>();
    LanczosDev hs{};
    hs.m = 2;
    hs.alpha[0] = aal[0];
    hs.beta[1] = nnl[1];
    SADMRG_HIP(hipMemcpyAsync(st, &hs, sizeof(hs), hipMemcpyHostToDevice, ctx->stream()));
    D.copy(col(1), xin);
    ctx->sync();  // hs is reused as the read-back buffer
    const int cx = D.cx ? 1 : 0;
    auto step = [&]() {
      H.apply(xin, w, bulk);
      SADMRG_HIP(launch_gemv_t(xin, 0, 1, w, D.n, cx, ctx->partial(), 1.0, st->dot, ctx->stream()));
      SADMRG_HIP(launch_lanczos_update(st, w, xin, V, nw, nw, ctx->stream()));
      for (int pass = 0; pass < 2; ++pass) {
        SADMRG_HIP(launch_gemv_t_dev(V, D.n, ml, &st->m, w, D.n, cx, ctx->partial(), -1.0, ctx->dscal(),
                                     ctx->stream()));
        SADMRG_HIP(launch_gemv_n_update_dev(V, D.n, ml, &st->m, ctx->dscal(), w, D.n, cx, ctx->stream()));
      }
      SADMRG_HIP(launch_gemv_t(w, 0, 1, w, D.n, cx, ctx->partial(), 1.0, st->nrm, ctx->stream()));
      SADMRG_HIP(launch_lanczos_finish(st, w, xin, V, nw, nw, tol, ctx->stream()));
      SADMRG_HIP(launch_lanczos_advance(st, ctx->stream()));
    };
    StepGraph sg;
    if (opt.graph && g_graph && ml > 3) {
      SADMRG_HIP(hipStreamBeginCapture(ctx->stream(), hipStreamCaptureModeThreadLocal));
      try {
        step();
      } catch (...) {
        hipGraph_t g = nullptr;
        hipStreamEndCapture(ctx->stream(), &g);
        if (g) hipGraphDestroy(g);
        throw;
      }
      SADMRG_HIP(hipStreamEndCapture(ctx->stream(), &sg.graph));
      // The step must hold kernel nodes only: a captured copy or fill would
      // keep the source pointer or value of capture time, and a host-side
      // source could be gone by the replay.  Every kernel argument points
      // into the plan (operator packs, block and tile tables, scratch), the
      // context (reduction partials, device scalars) or this solve's buffers
      // (V, the step state, x, w), all of which outlive sg.
      check_kernel_only(sg.graph);
      SADMRG_HIP(hipGraphInstantiate(&sg.exec, sg.graph, nullptr, nullptr, 0));
      g_captures++;
    }
    // host copies of alpha / beta and the stop state
    auto fetch = [&]() {
      SADMRG_HIP(hipMemcpyAsync(&hs, st, sizeof(hs), hipMemcpyDeviceToHost, ctx->stream()));
      ctx->sync();
      const int top = hs.stop ? hs.m : hs.m - 1;  // last step that completed
      for (int i = 1; i <= top && i < ml; ++i) aal[i] = hs.alpha[i];
      for (int i = 2; i <= top && i <= ml; ++i) nnl[i] = hs.beta[i];
    };
    for (int m = 2; !done && m < ml; ++m) {
      if (sg.exec) {
        SADMRG_HIP(hipGraphLaunch(sg.exec, ctx->stream()));
        out.graph_launches++;
        g_replays++;
      } else {
        step();
      }
      if (bulk == Prec::FP64) out.matvec64++;
      else out.matvec32++;
      out.steps = m;
      if (m % opt.check_every != 0 && m != ml - 1) continue;
      fetch();
      if (hs.stop) {
        // breakdown at step k: the host-driven loop stops there
        const int k = hs.m;
        const int wasted = m - k;
        if (bulk == Prec::FP64) out.matvec64 -= wasted;
        else out.matvec32 -= wasted;
        out.steps = k;
        diatridiag(k);
        ritz(k);
        out.ritz32 = eig[0];
        if (!est_only) check();
        done = true;
        break;
      }
      if (m % opt.check_every != 0) continue;
      diatridiag(m);
      if (est_only) {
        out.trace.push_back(eig[0]);
        if (est_converged(m)) {
          ritz(m);
          done = true;
          break;
        }
        continue;
      }
      ritz(m);
      out.ritz32 = eig[0];
      const double ov = check();
      out.trace.push_back(eig[0]);
      if (std::fabs(ov - 1.0) < 1e-8) {
        done = true;
        break;
      }
    }
    ctx->sync();  // the step buffers leave scope
  }
  for (int m = 2; !done && !dev_steps && m < ml; ++m) {
    matvec(col(m - 1), col(m), bulk);
    aal[m - 1] = D.dot(col(m), col(m - 1)).real();
    D.axpby(1.0, -aal[m - 1], col(m - 1), col(m));
    D.axpby(1.0, -nnl[m - 1], col(m - 2), col(m));
    D.orth(V, m, col(m));
    nnl[m] = D.nrm(col(m));
    D.scal(1.0 / nnl[m], col(m));
    out.steps = m;
    diatridiag(m);
    if (nnl[m] < tol) {
      ritz(m);
      out.ritz32 = eig[0];
      if (!est_only) check();
      done = true;
      break;
    }
    if (m % opt.check_every == 0 && est_only) {
      out.trace.push_back(eig[0]);
      if (est_converged(m)) {
        ritz(m);
        done = true;
        break;
      }
    } else if (m % opt.check_every == 0) {
      ritz(m);
      out.ritz32 = eig[0];
      const double ov = check();
      out.trace.push_back(eig[0]);
      if (std::fabs(ov - 1.0) < 1e-8) {
        done = true;
        break;
      }
    }
  }
  if (!done) {
    matvec(col(ml - 1), col(ml), bulk);
    aal[ml - 1] = D.dot(col(ml), col(ml - 1)).real();
    diatridiag(ml);
    ritz(ml);
    out.ritz32 = eig[0];
    if (!est_only) check();
    out.steps = ml;
  }
  out.energy = eig[0];

  if (bulk != Prec::FP64 && opt.refine && !est_only) {
    // FP64 residual-correction restarts on the Ritz vector in d_vec.
    const int k = std::max(1, std::min(opt.refine_krylov, ml - 1));
    double* x = d_vec;
    for (int r = 0; r <= opt.max_restarts; ++r) {
      D.copy(x, col(0));
      const double nor = D.nrm(col(0));
      D.scal(1.0 / nor, col(0));
      if (r == 0 && hv_valid) D.copy(hv, col(1));  // the last check's FP64 H v (x already normalised)
      else matvec(col(0), col(1), Prec::FP64);
      const double a0 = D.dot(col(1), col(0)).real();
      D.axpby(1.0, -a0, col(0), col(1));
      const double b1 = D.nrm(col(1));
      out.energy = a0;
      out.residual64 = b1;
      D.copy(col(0), x);
      if (b1 <= opt.refine_tol * std::fabs(a0) || r == opt.max_restarts) break;
      out.restarts++;
      std::vector<double> ta{a0}, tb{0.0, b1};
      D.scal(1.0 / b1, col(1));
      int mk = 2;
      for (int i = 1; i <= k; ++i) {
        matvec(col(i), col(i + 1), bulk);
        const double ai = D.dot(col(i + 1), col(i)).real();
        ta.push_back(ai);
        if (i == k) break;
        D.axpby(1.0, -ai, col(i), col(i + 1));
        D.axpby(1.0, -tb[i], col(i - 1), col(i + 1));
        D.orth(V, i + 1, col(i + 1));
        const double bi = D.nrm(col(i + 1));
        if (bi < 1e-12 * std::fabs(a0)) break;
        D.scal(1.0 / bi, col(i + 1));
        tb.push_back(bi);
        mk = i + 2;
      }
      const int msz = static_cast<int>(ta.size());
      mk = msz;
      std::vector<double> d(ta.begin(), ta.end()), e(std::max(mk - 1, 0));
      for (int i = 0; i + 1 < mk; ++i) e[i] = tb[i + 1];
      tridiag_eig(mk, d, e, z);
      std::vector<double> c(mk);
      for (int j = 0; j < mk; ++j) c[j] = z[j];
      D.combine(V, c, x);
    }
  }
  return out;
}


// ---- batched solves over the instances of a merged plan -------------------
namespace {

// Segmented vector operations on the instances' vectors stored back to back.
struct SegOps {
  static constexpr int kMaxCol = 64;
  static constexpr int64_t kChunk = 16384;  // elements per workgroup block
  Context* ctx;
  bool cx;
  int ninst;
  int64_t n;
  int nb = 0;
  DevBuf blocks, first, partial, out, coef, beta;
  SegOps(Context* c, bool cx_, const std::vector<int64_t>& seg) : ctx(c), cx(cx_) {
    ninst = static_cast<int>(seg.size()) - 1;
    n = seg.back();
    std::vector<SegBlock> hb;
    std::vector<int> hf(ninst + 1, 0);
    for (int i = 0; i < ninst; ++i) {
      hf[i] = static_cast<int>(hb.size());
      for (int64_t b = seg[i]; b < seg[i + 1]; b += kChunk)
        hb.push_back(SegBlock{b, std::min(seg[i + 1], b + kChunk), i, 0});
    }
    hf[ninst] = static_cast<int>(hb.size());
    nb = static_cast<int>(hb.size());
    blocks = DevBuf(ctx, std::max<size_t>(hb.size(), 1) * sizeof(SegBlock));
    first = DevBuf(ctx, hf.size() * sizeof(int));
    if (nb) ctx->upload(blocks.get(), hb.data(), hb.size() * sizeof(SegBlock));
    ctx->upload(first.get(), hf.data(), hf.size() * sizeof(int));
    partial = DevBuf(ctx, std::max(nb, 1) * kMaxCol * 2 * sizeof(double));
    out = DevBuf(ctx, ninst * kMaxCol * 2 * sizeof(double));
    coef = DevBuf(ctx, ninst * kMaxCol * 2 * sizeof(double));
    beta = DevBuf(ctx, ninst * sizeof(double));
  }
  // conj(X[:, c]) . y per instance: result[i * ncol + c]
  std::vector<dcmplex> dots(const double* X, int64_t ldx, int ncol, const double* y) {
    require(ncol >= 1 && ncol <= kMaxCol, "batched lanczos: too many columns");
    SADMRG_HIP(launch_seg_dot(X, ldx, ncol, y, blocks.as<SegBlock>(), nb, cx ? 1 : 0, partial.as<double>(),
                              ctx->stream()));
    SADMRG_HIP(launch_seg_reduce(partial.as<double>(), first.as<int>(), ninst, ncol, out.as<double>(), ctx->stream()));
    std::vector<double> h(static_cast<size_t>(ninst) * ncol * 2);
    SADMRG_HIP(hipMemcpyAsync(h.data(), out.get(), h.size() * 8, hipMemcpyDeviceToHost, ctx->stream()));
    ctx->sync();
    std::vector<dcmplex> r(static_cast<size_t>(ninst) * ncol);
    for (size_t k = 0; k < r.size(); ++k) r[k] = dcmplex(h[2 * k], h[2 * k + 1]);
    return r;
  }
  std::vector<double> norms(const double* y) {
    const std::vector<dcmplex> d = dots(y, 0, 1, y);
    std::vector<double> r(ninst);
    for (int i = 0; i < ninst; ++i) r[i] = std::sqrt(std::fabs(d[i].real()));
    return r;
  }
  // y = b[i] y + sum_c c[i * ncol + c] X[:, c] on instance i's elements
  void combine(const double* X, int64_t ldx, int ncol, const std::vector<dcmplex>& c, const std::vector<double>& b,
               double* y) {
    require(ncol >= 0 && ncol <= kMaxCol, "batched lanczos: too many columns");
    if (ncol > 0) {
      std::vector<double> h(static_cast<size_t>(ninst) * ncol * 2);
      for (size_t k = 0; k < c.size(); ++k) {
        h[2 * k] = c[k].real();
        h[2 * k + 1] = c[k].imag();
      }
      ctx->upload(coef.get(), h.data(), h.size() * 8);
    }
    ctx->upload(beta.get(), b.data(), b.size() * 8);
    SADMRG_HIP(launch_seg_combine(X, ldx, ncol, coef.as<double>(), beta.as<double>(), y, blocks.as<SegBlock>(), nb,
                                  cx ? 1 : 0, ctx->stream()));
  }
  void scale(const std::vector<double>& a, double* y) { combine(nullptr, 0, 0, {}, a, y); }
};

}  // namespace

BatchedLanczosOut lanczos_batched(HeffPlan& H, const std::vector<int64_t>& seg, double* d_vec, const LanczosOpts& opt) {
  Context* ctx = H.ctx();
  const bool cx = H.complex();
  const int w = cx ? 2 : 1;
  require(seg.size() >= 2 && seg.front() == 0 && seg.back() == H.n(), "batched lanczos: bad instance offsets");
  for (size_t i = 0; i + 1 < seg.size(); ++i) require(seg[i + 1] > seg[i], "batched lanczos: empty instance");
  const int ni = static_cast<int>(seg.size()) - 1;
  const int64_t nw = H.n() * w;
  std::vector<int> ml(ni);
  int ML = 2;
  for (int i = 0; i < ni; ++i) {
    ml[i] = static_cast<int>(std::max<int64_t>(2, std::min<int64_t>(opt.mlanc, seg[i + 1] - seg[i] + 1)));
    ML = std::max(ML, ml[i]);
  }
  require(ML + 1 <= SegOps::kMaxCol, "batched lanczos: mlanc must be at most 63");
  SegOps S(ctx, cx, seg);
  BatchedLanczosOut out;
  out.energy.assign(ni, 0.0);
  out.residual64.assign(ni, -1.0);
  out.steps.assign(ni, 0);
  out.restarts.assign(ni, 0);
  DevBuf Vb(ctx, std::max<int64_t>(nw, 1) * 8 * (ML + 1)), hvb(ctx, std::max<int64_t>(nw, 1) * 8);
  double* V = Vb.as<double>();
  double* hv = hvb.as<double>();
  auto col = [&](int j) { return V + static_cast<int64_t>(j) * nw; };
  const Prec bulk = opt.prec;
  // refinement restarts (FP64 residuals, bulk-precision Krylov steps) in
  // either precision; FP32 bulk checks in FP64
  const bool mixed = opt.refine;
  const Prec check_prec = opt.refine ? Prec::FP64 : bulk;
  auto matvec = [&](const double* x, double* y, Prec p) {
    H.apply(x, y, p);
    if (p == Prec::FP64) out.matvec64++;
    else out.matvec32++;
  };
  const double tol = 1e-6;
  std::vector<std::vector<double>> aal(ni, std::vector<double>(ML + 1, 0.0)), nnl(ni, std::vector<double>(ML + 2, 0.0));
  std::vector<uint8_t> active(ni, 1);
  const std::vector<double> ones(ni, 1.0), zeros(ni, 0.0);
  std::vector<dcmplex> zc;

  // start vectors (initialize_lanczos, dmrg_inst.cpp:1201-1212)
  SADMRG_HIP(hipMemcpyAsync(col(0), d_vec, nw * 8, hipMemcpyDeviceToDevice, ctx->stream()));
  {
    std::vector<double> nor = S.norms(col(0)), a(ni);
    for (int i = 0; i < ni; ++i) a[i] = nor[i] > 1e-12 ? 1.0 / nor[i] : 1.0;
    S.scale(a, col(0));
  }
  // x (d_vec) holds each instance's current Ritz vector; tridiagonal
  // eigenvectors per instance for the Ritz combinations
  auto tridiag = [&](int i, int m, std::vector<double>& d, std::vector<double>& z) {
    d.assign(aal[i].begin(), aal[i].begin() + m);
    std::vector<double> e(std::max(m - 1, 0));
    for (int k = 0; k + 1 < m; ++k) e[k] = nnl[i][k + 1];
    tridiag_eig(m, d, e, z);
  };
  // Ritz vectors of the instances in `set` (k[i] basis vectors), the
  // others' x left as is; then, if check, normalisation and one shared
  // check matvec: returns |<x,Hx>| / |Hx| per instance
  auto ritz_and_check = [&](const std::vector<int>& set, const std::vector<int>& kk,
                            const std::vector<std::vector<double>>& zz, bool check) {
    int K = 1;
    for (int i : set) K = std::max(K, kk[i]);
    std::vector<dcmplex> c(static_cast<size_t>(ni) * K, dcmplex(0, 0));
    std::vector<double> b(ni, 1.0);
    for (int i : set) {
      b[i] = 0.0;
      for (int j = 0; j < kk[i]; ++j) c[static_cast<size_t>(i) * K + j] = zz[i][j];
    }
    S.combine(V, H.n(), K, c, b, d_vec);
    std::vector<double> ov(ni, 0.0);
    if (!check) return ov;
    std::vector<double> nor = S.norms(d_vec), a(ni, 1.0);
    for (int i : set) a[i] = nor[i] > 0 ? 1.0 / nor[i] : 1.0;
    S.scale(a, d_vec);
    matvec(d_vec, hv, check_prec);
    const std::vector<double> hn = S.norms(hv);
    const std::vector<dcmplex> d = S.dots(d_vec, 0, 1, hv);
    for (int i : set) ov[i] = hn[i] > 0 ? std::abs(d[i]) / hn[i] : 1.0;
    return ov;
  };
  std::vector<int> kk(ni, 0);
  std::vector<std::vector<double>> zz(ni);
  std::vector<double> dtmp, ztmp;
  auto finish = [&](int i, int k, std::vector<int>& set) {
    tridiag(i, k, dtmp, ztmp);
    out.energy[i] = dtmp[0];
    zz[i].assign(ztmp.begin(), ztmp.begin() + k);  // eigenvector 0 (leading dimension k)
    kk[i] = k;
    set.push_back(i);
  };

  // step 1
  matvec(col(0), col(1), bulk);
  {
    const std::vector<dcmplex> a = S.dots(col(0), 0, 1, col(1));
    std::vector<dcmplex> c(ni);
    for (int i = 0; i < ni; ++i) {
      aal[i][0] = a[i].real();
      c[i] = -aal[i][0];
    }
    S.combine(col(0), 0, 1, c, ones, col(1));
    const std::vector<double> b = S.norms(col(1));
    std::vector<double> sc(ni);
    std::vector<int> fin;
    for (int i = 0; i < ni; ++i) {
      nnl[i][1] = b[i];
      out.steps[i] = 1;
      sc[i] = b[i] >= tol ? 1.0 / b[i] : 0.0;
      if (b[i] < tol) {  // Krylov space exhausted after one vector
        finish(i, 1, fin);
        active[i] = 0;
      }
    }
    S.scale(sc, col(1));
    if (!fin.empty()) ritz_and_check(fin, kk, zz, !mixed ? true : true);
  }
  for (int m = 2; m <= ML; ++m) {
    bool any = false;
    for (int i = 0; i < ni; ++i) any = any || active[i];
    if (!any) break;
    matvec(col(m - 1), col(m), bulk);
    const std::vector<dcmplex> a = S.dots(col(m - 1), 0, 1, col(m));
    // instances at their last basis vector (m == ml): the reference's final
    // step takes alpha and stops (dmrg_inst.cpp:1341-1347)
    std::vector<int> fin, chk;
    std::vector<uint8_t> step(ni, 0);
    for (int i = 0; i < ni; ++i) {
      if (!active[i]) continue;
      aal[i][m - 1] = a[i].real();
      if (m == ml[i]) {
        out.steps[i] = m;
        finish(i, m, fin);
        active[i] = 0;
      } else {
        step[i] = 1;
      }
    }
    // three-term recurrence, two re-orthogonalisation passes and the
    // normalisation for the stepping instances; every other instance's
    // column m is zeroed
    {
      std::vector<dcmplex> c(static_cast<size_t>(ni) * 2, dcmplex(0, 0));
      std::vector<double> b(ni, 0.0);
      for (int i = 0; i < ni; ++i)
        if (step[i]) {
          c[2 * i] = -nnl[i][m - 1];
          c[2 * i + 1] = -aal[i][m - 1];
          b[i] = 1.0;
        }
      S.combine(col(m - 2), nw / w, 2, c, b, col(m));
    }
    for (int pass = 0; pass < 2; ++pass) {
      std::vector<dcmplex> c = S.dots(V, H.n(), m, col(m));
      for (int i = 0; i < ni; ++i)
        for (int j = 0; j < m; ++j) c[static_cast<size_t>(i) * m + j] = step[i] ? -c[static_cast<size_t>(i) * m + j] : 0.0;
      S.combine(V, H.n(), m, c, ones, col(m));
    }
    const std::vector<double> bn = S.norms(col(m));
    std::vector<double> sc(ni, 0.0);
    for (int i = 0; i < ni; ++i) {
      if (!step[i]) continue;
      nnl[i][m] = bn[i];
      out.steps[i] = m;
      if (bn[i] < tol) {  // breakdown: stop at this step (dmrg_inst.cpp:1322-1328)
        finish(i, m, fin);
        active[i] = 0;
        continue;
      }
      sc[i] = 1.0 / bn[i];
      if (m % opt.check_every == 0) finish(i, m, chk);
    }
    S.scale(sc, col(m));
    std::vector<int> set = fin;
    set.insert(set.end(), chk.begin(), chk.end());
    if (set.empty()) continue;
    const std::vector<double> ov = ritz_and_check(set, kk, zz, true);
    for (int i : chk)
      if (std::fabs(ov[i] - 1.0) < 1e-8) active[i] = 0;
  }
  ctx->sync();
  if (!mixed) return out;

  // FP64 residual-correction restarts (batched over the instances that need one)
  const int k = std::max(1, std::min(opt.refine_krylov, ML - 1));
  auto residual = [&](const std::vector<int>& set, std::vector<double>& a0, std::vector<double>& b1) {
    // x normalised; hv = H x (FP64); col(1) = hv - a0 x for the instances in set
    std::vector<double> nor = S.norms(d_vec), a(ni, 1.0);
    for (int i : set) a[i] = nor[i] > 0 ? 1.0 / nor[i] : 1.0;
    S.scale(a, d_vec);
    matvec(d_vec, hv, Prec::FP64);
    const std::vector<dcmplex> d = S.dots(d_vec, 0, 1, hv);
    SADMRG_HIP(hipMemcpyAsync(col(1), hv, nw * 8, hipMemcpyDeviceToDevice, ctx->stream()));
    std::vector<dcmplex> c(ni, dcmplex(0, 0));
    std::vector<double> b(ni, 0.0);
    for (int i : set) {
      a0[i] = d[i].real();
      c[i] = -a0[i];
      b[i] = 1.0;
    }
    S.combine(d_vec, 0, 1, c, b, col(1));
    const std::vector<double> bn = S.norms(col(1));
    for (int i : set) {
      b1[i] = bn[i];
      out.energy[i] = a0[i];
      out.residual64[i] = b1[i];
    }
  };
  std::vector<int> all(ni);
  for (int i = 0; i < ni; ++i) all[i] = i;
  std::vector<double> a0(ni, 0.0), b1(ni, 0.0);
  residual(all, a0, b1);
  for (int r = 0; r < opt.max_restarts; ++r) {
    std::vector<int> need;
    for (int i = 0; i < ni; ++i)
      if (b1[i] > opt.refine_tol * std::fabs(a0[i])) need.push_back(i);
    if (need.empty()) break;
    std::vector<uint8_t> on(ni, 0);
    for (int i : need) {
      on[i] = 1;
      out.restarts[i]++;
    }
    // col(0) = x, col(1) = r / |r| on the instances refined; the others zero
    SADMRG_HIP(hipMemcpyAsync(col(0), d_vec, nw * 8, hipMemcpyDeviceToDevice, ctx->stream()));
    {
      std::vector<double> z0(ni, 0.0), s1(ni, 0.0);
      for (int i : need) {
        z0[i] = 1.0;
        s1[i] = 1.0 / b1[i];
      }
      S.scale(z0, col(0));
      S.scale(s1, col(1));
    }
    std::vector<std::vector<double>> ta(ni), tb(ni);
    std::vector<uint8_t> live = on;
    for (int i : need) {
      ta[i] = {a0[i]};
      tb[i] = {0.0, b1[i]};
    }
    for (int j = 1; j <= k; ++j) {
      matvec(col(j), col(j + 1), bulk);
      const std::vector<dcmplex> a = S.dots(col(j), 0, 1, col(j + 1));
      for (int i : need)
        if (live[i]) ta[i].push_back(a[i].real());
      if (j == k) break;
      std::vector<dcmplex> c(static_cast<size_t>(ni) * 2, dcmplex(0, 0));
      std::vector<double> b(ni, 0.0);
      for (int i : need)
        if (live[i]) {
          c[2 * i] = -tb[i][j];
          c[2 * i + 1] = -ta[i][j];
          b[i] = 1.0;
        }
      S.combine(col(j - 1), nw / w, 2, c, b, col(j + 1));
      for (int pass = 0; pass < 2; ++pass) {
        std::vector<dcmplex> cc = S.dots(V, H.n(), j + 1, col(j + 1));
        for (int i = 0; i < ni; ++i)
          for (int q = 0; q <= j; ++q)
            cc[static_cast<size_t>(i) * (j + 1) + q] = live[i] ? -cc[static_cast<size_t>(i) * (j + 1) + q] : 0.0;
        S.combine(V, H.n(), j + 1, cc, ones, col(j + 1));
      }
      const std::vector<double> bn = S.norms(col(j + 1));
      std::vector<double> sc(ni, 0.0);
      for (int i : need) {
        if (!live[i]) continue;
        if (bn[i] < 1e-12 * std::fabs(a0[i])) {
          live[i] = 0;
          continue;
        }
        sc[i] = 1.0 / bn[i];
        tb[i].push_back(bn[i]);
      }
      S.scale(sc, col(j + 1));
    }
    // Ritz vector of each refined instance in its small Krylov space
    std::vector<int> kk2(ni, 0);
    std::vector<std::vector<double>> z2(ni);
    for (int i : need) {
      const int mk = static_cast<int>(ta[i].size());
      std::vector<double> d(ta[i].begin(), ta[i].end()), e(std::max(mk - 1, 0)), z;
      for (int q = 0; q + 1 < mk; ++q) e[q] = tb[i][q + 1];
      tridiag_eig(mk, d, e, z);
      z2[i].assign(z.begin(), z.begin() + mk);
      kk2[i] = mk;
    }
    ritz_and_check(need, kk2, z2, false);
    residual(need, a0, b1);
  }
  ctx->sync();
  return out;
}

}  // namespace sadmrg
