// Local minimisers of minim.f: quasi-Newton with BFGS updates of a factored
// Hessian (minqsn.f:12-430 with mc11a/mc11e) and conjugate gradients with
// Powell restarts (mincjg.f:12-250).
//
// Both run in reverse communication: after start(), while pending() the
// caller evaluates the function and its gradient at point() and hands them
// over with supply().  That way many independent minimisations can share
// one batched device evaluation per round, while each follows exactly the
// sequence of steps the reference's routine takes with its MOVE callback.
#pragma once
#include <cstdint>
#include <vector>

namespace smmp {

class QuasiNewton {
 public:
  // x, f, g: variables, function value and gradient at the start; scal: per
  // variable scale (initial step and its lower bound, initial Hessian
  // diagonal); acur: fractional precision of the variables; maxfun: most
  // function calls.
  void start(int n, const double* x, double f, const double* g, const double* scal, double acur, int maxfun);
  bool pending() const { return state_ == kEvalLine || state_ == kEvalFinal; }
  const double* point() const { return state_ == kEvalLine ? xb_.data() : x_.data(); }
  void supply(double f, const double* g);
  // Best point so far (the reference's X, F, G on return), calls used.
  const std::vector<double>& x() const { return x_; }
  double f() const { return f_; }
  const std::vector<double>& g() const { return g_; }
  int nfun() const { return nfun_; }
  bool rank_deficient() const { return rank_deficient_; }

 private:
  enum State { kIdle, kEvalLine, kEvalFinal, kDone };
  void restart();      // label 1: line searches from the best point so far
  void line_search();  // label 2: new direction and first step
  void step_or_stop(); // label 3: next trial point, or the call limit
  void after_trial();  // labels 3 (after MOVE) and 4
  void finish();       // label 5
  int n_ = 0, maxfun_ = 0, nfun_ = 0, itr_ = 0, isfv_ = 0;
  double f_ = 0, fa_ = 0, fb_ = 0, dff_ = 0, dga_ = 0, dgb_ = 0, fmin_ = 0, gmin_ = 0;
  double stmin_ = 0, stepub_ = 0, steplb_ = 0, step_ = 0, c_ = 0, acur_ = 0;
  bool rank_deficient_ = false;
  State state_ = kIdle;
  std::vector<double> x_, g_, scal_, h_, d_, w_, xa_, ga_, xb_, gb_;
};

class ConjugateGradient {
 public:
  void start(int n, const double* x, double f, const double* g, double acur, int maxfun);
  bool pending() const { return state_ == kEvalLine || state_ == kEvalFinal; }
  const double* point() const { return x_.data(); }
  void supply(double f, const double* g);
  const std::vector<double>& x() const { return x_; }
  double f() const { return f_; }
  const std::vector<double>& g() const { return g_; }
  int nfun() const { return nfun_; }
  int ier() const { return ier_; }  // 0 normal end, 1 call limit, 2 step too small, 3 uphill, 4 no reduction

 private:
  enum State { kIdle, kEvalLine, kEvalFinal, kDone };
  void iteration();   // label 1
  void trial();       // label 2
  void after_trial(); // after MOVE in label 2
  void next_step();   // label 4
  void limit_or_step();  // label 3
  void end_search();  // label 6
  void new_direction();  // label 5
  int n_ = 0, maxfun_ = 0, nfun_ = 0, ier_ = 0, iter_ = 0, iterfm_ = 0, iterrs_ = 0, nfopt_ = 0, nfbeg_ = 0,
      iretry_ = 0;
  double f_ = 0, acur_ = 0, gsqrd_ = 0, fmin_ = 0, gsq2_ = 0, gnew_ = 0, dfpr_ = 0, stmin_ = 0, fuit_ = 0,
         gdit_ = 0, gdmi_ = 0, sbound_ = 0, stepch_ = 0, step_ = 0, fch_ = 0, ddspln_ = 0, gamden_ = 0;
  State state_ = kIdle;
  std::vector<double> x_, g_, d_, xa_, ga_, dt_, yt_, gt_;
};

// mc11a of minqsn.f: a := a + sig z z^T for a positive definite matrix held
// as its factors L D L^T (packed by rows of the upper triangle, D on the
// diagonal), keeping positive definiteness; ir: rank in and out; mk = 1:
// w holds the forward-substituted z from the preceding mc11e; eps: see
// minqsn.f.  z is overwritten; w may be z when sig > 0.
void mc11a(double* a, int n, double* z, double sig, double* w, int& ir, int mk, double eps);
// mc11e: solves (L D L^T) z_out = z_in in place; w receives L^-1 z_in.
void mc11e(const double* a, int n, double* z, double* w, int ir);

}  // namespace smmp
