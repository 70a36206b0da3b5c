// minim.f:13-250 on a context: local minimisation of the energy by the
// variables that are not fixed, with the derivatives from the device
// (engine_gradient), and a batched form that minimises many conformations
// side by side, one device launch per round of function calls.
#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>

#include "engine.hpp"
#include "minimize.hpp"
#include "minim.hpp"

namespace smmp {

namespace {

constexpr double kPiM = 3.141592653589793238462643;
constexpr double kPi2M = 2.0 * kPiM;

// The variables minim.f works on: those not fixed (minim.f:101-110).
std::vector<int> free_variables(const State& s) {
  std::vector<int> fv;
  const int nvt = s.nvars();
  for (int i = 1; i <= nvt; ++i)
    if (!s.fxvr(i)) fv.push_back(i - 1);
  return fv;
}

void check_minim(const State& s, int imin) {
  if (s.ntlml < 1) throw Error("minim> no molecule initialised");
  if (imin != 1 && imin != 2) throw Error("minim> imin must be 1 (quasi-Newton) or 2 (conjugate gradients)");
  if (s.ireg != 0) throw Error("minim> regularisation (ireg != 0, opereg.f) is not provided");
  gradient_supported(s);
}

// One minimiser of either kind behind one interface.
struct Minimiser {
  int imin = 1;
  QuasiNewton qn;
  ConjugateGradient cg;
  void start(int n, const double* x, double f, const double* g, double acc, int maxit) {
    if (imin == 1) {
      std::vector<double> scl(n, 1.0);  // minim.f:106 scale 1 for every variable
      qn.start(n, x, f, g, scl.data(), acc, maxit);
    } else {
      cg.start(n, x, f, g, acc, maxit);
    }
  }
  bool pending() const { return imin == 1 ? qn.pending() : cg.pending(); }
  const double* point() const { return imin == 1 ? qn.point() : cg.point(); }
  void supply(double f, const double* g) { imin == 1 ? qn.supply(f, g) : cg.supply(f, g); }
  const std::vector<double>& x() const { return imin == 1 ? qn.x() : cg.x(); }
  double f() const { return imin == 1 ? qn.f() : cg.f(); }
  int nfun() const { return imin == 1 ? qn.nfun() : cg.nfun(); }
};

double wrap(double v) { return std::fabs(v) > kPiM ? v - std::copysign(kPi2M, v) : v; }

// fn(c) for c in [0, n): on up to 16 host threads when there is enough work
// (the quasi-Newton factor updates cost O(nvr^2) per conformation and round),
// each thread a contiguous range.
template <typename F>
void parallel_for(int n, F&& fn) {
  const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
  const int nt = std::min({16, hw, n / 32});
  if (nt <= 1) {
    for (int c = 0; c < n; ++c) fn(c);
    return;
  }
  std::vector<std::thread> th;
  th.reserve(nt);
  for (int t = 0; t < nt; ++t)
    th.emplace_back([&, t] {
      for (int c = (int)((int64_t)n * t / nt); c < (int)((int64_t)n * (t + 1) / nt); ++c) fn(c);
    });
  for (auto& x : th) x.join();
}

}  // namespace

int minim(Engine*& e, State& s, int imin, int maxit, double eps) {
  check_minim(s, imin);
  const int nvt = s.nvars();
  const std::vector<int> fv = free_variables(s);
  const int n = (int)fv.size();
  std::vector<double> vl(s.vlvr.data(), s.vlvr.data() + nvt), ge(kGradEnergies), gd(nvt);
  auto evaluate = [&](const double* xfree) {  // move() of minim.f: variables, gradient()
    for (int k = 0; k < n; ++k) vl[fv[k]] = xfree[k];
    engine_gradient(e, s, 1, vl.data(), ge.data(), gd.data());
    for (int k = 0; k < nvt; ++k) s.vlvr(k + 1) = vl[k];
    s.eysm = ge[0];
    s.eyel = ge[1];
    s.eyvw = ge[2];
    s.eyhb = ge[3];
    s.eyvr = ge[4];
    s.eysl = s.itysol < 0 ? ge[0] - (ge[1] + ge[2] + ge[3] + ge[4]) : 0.0;
    for (int k = 0; k < nvt; ++k) s.gdeyvr(k + 1) = gd[k];
  };
  std::vector<double> x(n), g(n);
  for (int k = 0; k < n; ++k) x[k] = vl[fv[k]];
  evaluate(x.data());  // energy and gradient at the start
  for (int k = 0; k < n; ++k) g[k] = gd[fv[k]];
  Minimiser m;
  m.imin = imin;
  m.start(n, x.data(), ge[0], g.data(), eps, maxit);
  while (m.pending()) {
    evaluate(m.point());
    for (int k = 0; k < n; ++k) g[k] = gd[fv[k]];
    m.supply(ge[0], g.data());
  }
  // final energies at the variables the last call left (minim.f:189), then
  // the best point's variables wrapped into [-pi, pi] (minim.f:218-232)
  s.eysm = engine_energy(e, s);
  for (int k = 0; k < n; ++k) s.vlvr(fv[k] + 1) = wrap(m.x()[k]);
  return m.nfun();
}

void minim_batch(Engine*& e, State& s, int imin, int nconf, double* vlvr, int maxit, double eps, double* fout,
                 int32_t* nfun) {
  check_minim(s, imin);
  if (nconf <= 0) return;
  const int nvt = s.nvars();
  const std::vector<int> fv = free_variables(s);
  const int n = (int)fv.size();
  std::vector<Minimiser> ms(nconf);
  std::vector<double> ge((size_t)kGradEnergies * nconf), gd((size_t)nvt * nconf);
  engine_gradient(e, s, nconf, vlvr, ge.data(), gd.data());
  parallel_for(nconf, [&](int c) {
    std::vector<double> x(n), g(n);
    for (int k = 0; k < n; ++k) {
      x[k] = vlvr[(size_t)c * nvt + fv[k]];
      g[k] = gd[(size_t)c * nvt + fv[k]];
    }
    ms[c].imin = imin;
    ms[c].start(n, x.data(), ge[(size_t)kGradEnergies * c], g.data(), eps, maxit);
  });
  // rounds: every minimisation still running asks for one point
  std::vector<int> active;
  std::vector<double> pts;
  for (;;) {
    active.clear();
    for (int c = 0; c < nconf; ++c)
      if (ms[c].pending()) active.push_back(c);
    if (active.empty()) break;
    const int na = (int)active.size();
    pts.resize((size_t)na * nvt);
    parallel_for(na, [&](int a) {
      double* p = pts.data() + (size_t)a * nvt;
      for (int k = 0; k < nvt; ++k) p[k] = vlvr[(size_t)active[a] * nvt + k];  // fixed variables
      const double* q = ms[active[a]].point();
      for (int k = 0; k < n; ++k) p[fv[k]] = q[k];
    });
    engine_gradient(e, s, na, pts.data(), ge.data(), gd.data());
    parallel_for(na, [&](int a) {
      std::vector<double> g(n);
      for (int k = 0; k < n; ++k) g[k] = gd[(size_t)a * nvt + fv[k]];
      ms[active[a]].supply(ge[(size_t)kGradEnergies * a], g.data());
    });
  }
  for (int c = 0; c < nconf; ++c) {
    for (int k = 0; k < n; ++k) vlvr[(size_t)c * nvt + fv[k]] = wrap(ms[c].x()[k]);
    if (fout) fout[c] = ms[c].f();
    if (nfun) nfun[c] = ms[c].nfun();
  }
}

}  // namespace smmp
