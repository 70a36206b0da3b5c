#include "minimize.hpp"

#include <algorithm>
#include <cmath>

namespace smmp {

namespace {
// minqsn.f:72-78 parameters
constexpr double kEps1 = 0.1;   // sufficient reduction of the function along a search
constexpr double kEps2 = 0.7;   // accuracy of a line search
constexpr double kTiny = 1e-32;
// Default-real (single precision) literals of minqsn.f, widened to double
// where they meet double operands: .01 (initial Hessian) and .1 (step cut).
constexpr double kHess0 = (double)0.01f;
constexpr double kCut = (double)0.1f;
// mincjg.f:27-31 parameters
constexpr double kAmf = 10.0;
constexpr int kMxfcon = 2;
constexpr int kMaxlin = 5;
constexpr double kTol = 1e-7;
constexpr double kEpsCg = 0.7;
}  // namespace

// ------------------------------------------------------------ mc11a/mc11e

void mc11a(double* a0, int n, double* z0, double sig, double* w0, int& ir, int mk, double eps) {
  // 1-based views of the packed matrix and the vectors, as in minqsn.f
  auto a = [&](int i) -> double& { return a0[i - 1]; };
  auto z = [&](int i) -> double& { return z0[i - 1]; };
  auto w = [&](int i) -> double& { return w0[i - 1]; };
  if (n <= 1) {
    a(1) = a(1) + sig * z(1) * z(1);
    ir = 1;
    if (a(1) > 0.0) return;
    a(1) = 0.0;
    ir = 0;
    return;
  }
  if (sig == 0.0 || ir == 0) return;
  const int np = n + 1;
  int ij = 1, mm = 0;
  double ti = 0.0, tim = 0.0;
  bool positive_path = sig >= 0.0;
  if (sig < 0.0) {
    // the rank-one term removes curvature: the ratios of the new diagonal
    // come from the solve with z, accumulated backwards
    ti = 1.0 / sig;
    if (mk == 0) {
      for (int i = 1; i <= n; ++i) w(i) = z(i);
      for (int i = 1; i <= n; ++i) {
        if (a(ij) > 0.0) {
          const double v = w(i);
          ti = ti + v * v / a(ij);
          for (int j = i + 1; j <= n; ++j) {
            ++ij;
            w(j) = w(j) - v * a(ij);
          }
          ++ij;
        } else {
          w(i) = 0.0;
          ij = ij + np - i;
        }
      }
    } else {
      for (int i = 1; i <= n; ++i) {
        if (a(ij) != 0.0) ti = ti + w(i) * w(i) / a(ij);
        ij = ij + np - i;
      }
    }
    if (ir <= 0) {
      ti = 0.0;
      ir = -ir - 1;
    } else if (ti > 0.0) {
      if (eps != 0.0) {
        ti = eps / sig;
      } else {
        ir = ir - 1;
        ti = 0.0;
      }
    } else if (mk <= 1) {
      positive_path = true;
    }
    if (!positive_path) {
      mm = 1;
      tim = ti;
      for (int i = 1; i <= n; ++i) {
        const int j = np - i;
        ij = ij - i;
        if (a(ij) != 0.0) tim = ti - w(j) * w(j) / a(ij);
        w(j) = ti;
        ti = tim;
      }
    }
  }
  if (positive_path) {
    mm = 0;
    tim = 1.0 / sig;
  }
  ij = 1;
  for (int i = 1; i <= n; ++i) {
    const double v = z(i);
    if (a(ij) > 0.0) {
      const double al = v / a(ij);
      ti = mm <= 0 ? tim + v * al : w(i);
      const double r = ti / tim;
      a(ij) = a(ij) * r;
      if (r == 0.0 || i == n) break;
      const double b = al / ti;
      if (r > 4.0) {
        const double gm = tim / ti;
        for (int j = i + 1; j <= n; ++j) {
          ++ij;
          const double y = a(ij);
          a(ij) = b * z(j) + y * gm;
          z(j) = z(j) - v * y;
        }
      } else {
        for (int j = i + 1; j <= n; ++j) {
          ++ij;
          z(j) = z(j) - v * a(ij);
          a(ij) = a(ij) + b * z(j);
        }
      }
      tim = ti;
      ++ij;
    } else if (ir > 0 || sig < 0.0 || v == 0.0) {
      ti = tim;
      ij = ij + np - i;
    } else {
      // a zero pivot takes the new direction: the rank grows by one
      ir = 1 - ir;
      a(ij) = v * v / tim;
      if (i == n) return;
      for (int j = i + 1; j <= n; ++j) {
        ++ij;
        a(ij) = z(j) / v;
      }
      return;
    }
  }
  if (ir < 0) ir = -ir;
}

void mc11e(const double* a0, int n, double* z0, double* w0, int ir) {
  auto a = [&](int i) { return a0[i - 1]; };
  auto z = [&](int i) -> double& { return z0[i - 1]; };
  auto w = [&](int i) -> double& { return w0[i - 1]; };
  if (ir < n) return;  // rank deficient: z unchanged
  w(1) = z(1);
  if (n <= 1) {
    z(1) = z(1) / a(1);
    return;
  }
  int ij = 0;
  for (int i = 2; i <= n; ++i) {  // forward substitution with the unit lower factor
    ij = i;
    double v = z(i);
    for (int j = 1; j <= i - 1; ++j) {
      v = v - a(ij) * z(j);
      ij = ij + n - j;
    }
    z(i) = v;
    w(i) = v;
  }
  z(n) = z(n) / a(ij);
  const int np = n + 1;
  for (int nip = 2; nip <= n; ++nip) {  // diagonal, then back substitution
    const int i = np - nip;
    int ii = ij - nip;
    double v = z(i) / a(ii);
    ij = ii;
    for (int j = i + 1; j <= n; ++j) {
      ++ii;
      v = v - a(ii) * z(j);
    }
    z(i) = v;
  }
}

// ------------------------------------------------------------ quasi-Newton

void QuasiNewton::start(int n, const double* x, double f, const double* g, const double* scal, double acur,
                        int maxfun) {
  n_ = n;
  maxfun_ = maxfun;
  acur_ = acur;
  x_.assign(x, x + n);
  g_.assign(g, g + n);
  scal_.assign(scal, scal + n);
  f_ = f;
  nfun_ = 0;
  itr_ = 0;
  dff_ = 0.0;
  rank_deficient_ = false;
  d_.assign(n, 0.0);
  w_.assign(n, 0.0);
  xa_.assign(n, 0.0);
  ga_.assign(n, 0.0);
  xb_.assign(n, 0.0);
  gb_.assign(n, 0.0);
  // initial Hessian: diagonal from the largest scaled gradient component
  double c = 0.0;
  for (int i = 0; i < n; ++i) c = std::max(c, std::fabs(g_[i] * scal_[i]));
  if (c <= 0.0) c = 1.0;
  h_.assign((size_t)n * (n + 1) / 2, 0.0);
  for (int i = 0, j = 0; i < n; ++i) {
    h_[j] = kHess0 * c / (scal_[i] * scal_[i]);
    j += n - i;
  }
  restart();
}

void QuasiNewton::restart() {
  isfv_ = 1;
  fa_ = f_;
  xa_ = x_;
  ga_ = g_;
  line_search();
}

void QuasiNewton::line_search() {
  ++itr_;
  for (int i = 0; i < n_; ++i) d_[i] = -ga_[i];
  mc11e(h_.data(), n_, d_.data(), w_.data(), n_);
  double c = 0.0;
  dga_ = 0.0;
  for (int i = 0; i < n_; ++i) {
    const double di = d_[i];
    c = std::max(c, std::fabs(di / scal_[i]));
    dga_ = dga_ + ga_[i] * di;  // directional derivative
  }
  c = 1.0 / c;
  if (dga_ >= 0.0) {  // uphill
    finish();
    return;
  }
  fmin_ = fa_;
  gmin_ = dga_;
  stmin_ = 0.0;
  stepub_ = 0.0;
  steplb_ = acur_ * c;
  step_ = dff_ > 0.0 ? std::min(1.0, (dff_ + dff_) / (-dga_)) : std::min(1.0, c);
  step_or_stop();
}

void QuasiNewton::step_or_stop() {
  if (nfun_ >= maxfun_) {
    state_ = kDone;
    return;
  }
  c_ = stmin_ + step_;
  for (int i = 0; i < n_; ++i) xb_[i] = xa_[i] + c_ * d_[i];
  ++nfun_;
  state_ = kEvalLine;
}

void QuasiNewton::supply(double f, const double* g) {
  if (state_ == kEvalLine) {
    fb_ = f;
    gb_.assign(g, g + n_);
    after_trial();
  } else if (state_ == kEvalFinal) {
    f_ = f;
    g_.assign(g, g + n_);
    state_ = kDone;
  }
}

void QuasiNewton::after_trial() {
  isfv_ = std::min(2, isfv_);
  if (fb_ <= f_) {
    bool keep = true;
    if (fb_ == f_) {  // a tie goes to the smaller scaled gradient
      double gl1 = 0.0, gl2 = 0.0;
      for (int i = 0; i < n_; ++i) {
        const double si = scal_[i] * scal_[i];
        gl1 = gl1 + si * g_[i] * g_[i];
        gl2 = gl2 + si * gb_[i] * gb_[i];
      }
      keep = gl2 < gl1;
    }
    if (keep) {  // smallest value so far
      f_ = fb_;
      x_ = xb_;
      g_ = gb_;
      isfv_ = 3;
    }
  }
  dgb_ = 0.0;
  for (int i = 0; i < n_; ++i) dgb_ = dgb_ + gb_[i] * d_[i];
  if (fb_ - fa_ <= kEps1 * c_ * dga_) {  // sufficient reduction
    stmin_ = c_;
    fmin_ = fb_;
    gmin_ = dgb_;
    stepub_ = stepub_ - step_;
    step_ = stepub_ > 0.0 ? 0.5 * stepub_ : 9.0 * stmin_;  // extrapolation
    double c = dga_ + 3.0 * dgb_ - 4.0 * (fb_ - fa_) / stmin_;
    if (c > 0.0) step_ = std::min(step_, stmin_ * std::max(1.0, -dgb_ / c));
    if (dgb_ < kEps2 * dga_) {  // line minimum not yet accurate enough
      step_or_stop();
      return;
    }
    isfv_ = 4 - isfv_;
    if (stmin_ + step_ > steplb_) {  // search done: BFGS update of the factors
      for (int i = 0; i < n_; ++i) {
        xa_[i] = xb_[i];
        xb_[i] = ga_[i];
        d_[i] = gb_[i] - ga_[i];
        ga_[i] = gb_[i];
      }
      int ir = -n_;
      double sig = 1.0 / dga_;
      mc11a(h_.data(), n_, xb_.data(), sig, w_.data(), ir, 1, 0.0);
      ir = -ir;
      sig = 1.0 / (stmin_ * (dgb_ - dga_));
      mc11a(h_.data(), n_, d_.data(), sig, d_.data(), ir, 0, 0.0);
      if (ir == n_) {
        dff_ = fa_ - fb_;
        fa_ = fb_;
        line_search();
      } else {  // rank of the new matrix deficient
        rank_deficient_ = true;
        state_ = kDone;
      }
      return;
    }
  } else if (step_ > steplb_) {  // insufficient reduction: cubic interpolation
    stepub_ = step_;
    double c = gmin_ + dgb_ - 3.0 * (fb_ - fmin_) / step_;
    c = c + gmin_ - std::sqrt(c * c - gmin_ * dgb_);
    if (std::fabs(c) < kTiny)
      step_ = kCut * step_;
    else
      step_ = step_ * std::max(0.1, gmin_ / c);
    step_or_stop();
    return;
  }
  finish();
}

void QuasiNewton::finish() {
  if (isfv_ >= 2) {  // restart from the best point so far
    restart();
    return;
  }
  ++nfun_;
  state_ = kEvalFinal;  // the function at the best point, then done
}

// ------------------------------------------------------------ conjugate gradients

void ConjugateGradient::start(int n, const double* x, double f, const double* g, double acur, int maxfun) {
  n_ = n;
  maxfun_ = maxfun;
  acur_ = acur;
  x_.assign(x, x + n);
  g_.assign(g, g + n);
  f_ = f;
  d_.assign(n, 0.0);
  xa_.assign(n, 0.0);
  ga_.assign(n, 0.0);
  dt_.assign(n, 0.0);
  yt_.assign(n, 0.0);
  gt_.assign(n, 0.0);
  ier_ = 0;
  iter_ = 0;
  iterfm_ = 0;
  iterrs_ = 0;
  nfun_ = 0;
  nfopt_ = 0;
  gsqrd_ = 0.0;
  fmin_ = f;
  gamden_ = 0.0;
  for (int i = 0; i < n; ++i) {
    const double t = g_[i];
    ga_[i] = t;
    d_[i] = -t;
    gsqrd_ = gsqrd_ + t * t;
    xa_[i] = x_[i];
  }
  gsq2_ = 0.2 * gsqrd_;
  gnew_ = -gsqrd_;
  dfpr_ = kAmf;
  stmin_ = kAmf / gsqrd_;
  iteration();
}

void ConjugateGradient::iteration() {
  ++iter_;
  fuit_ = f_;
  gdit_ = 0.0;
  for (int i = 0; i < n_; ++i) {
    const double t = g_[i];
    gt_[i] = t;
    gdit_ = gdit_ + d_[i] * t;
  }
  if (gdit_ >= 0.0) {  // uphill
    ier_ = 3;
    end_search();
    return;
  }
  gdmi_ = gdit_;
  sbound_ = -1.0;
  nfbeg_ = nfun_;
  iretry_ = -1;
  stepch_ = std::min(stmin_, std::fabs(dfpr_ / gdit_));
  stmin_ = 0.0;
  trial();
}

void ConjugateGradient::trial() {
  step_ = stmin_ + stepch_;
  double wo = 0.0;
  for (int i = 0; i < n_; ++i) {
    const double t = stepch_ * d_[i];
    wo = std::max(wo, std::fabs(t));
    x_[i] = xa_[i] + t;
  }
  if (wo > kTol) {
    ++nfun_;
    state_ = kEvalLine;
    return;
  }
  // the step is effectively zero
  if (nfun_ > nfbeg_ + 1 || std::fabs(gdmi_ / gdit_) > kEpsCg) ier_ = 2;
  end_search();
}

void ConjugateGradient::supply(double f, const double* g) {
  if (state_ != kEvalLine && state_ != kEvalFinal) return;
  f_ = f;
  g_.assign(g, g + n_);
  if (state_ == kEvalFinal) {
    state_ = kDone;
    return;
  }
  after_trial();
}

void ConjugateGradient::after_trial() {
  gnew_ = 0.0;
  double sum = 0.0;
  for (int i = 0; i < n_; ++i) {
    const double t = g_[i];
    gnew_ = gnew_ + d_[i] * t;
    sum = sum + t * t;
  }
  fch_ = f_ - fmin_;
  if (fch_ <= 0.0) {
    if (fch_ < 0.0 || gnew_ / gdmi_ >= -1.0) {  // new best point
      fmin_ = f_;
      gsqrd_ = sum;
      gsq2_ = 0.2 * gsqrd_;
      nfopt_ = nfun_;
      xa_ = x_;
      ga_ = g_;
    }
    if (sum <= acur_) {  // normal end
      state_ = kDone;
      return;
    }
  }
  if (nfun_ == maxfun_) {
    ier_ = 1;
    state_ = kDone;
    return;
  }
  const double wo = (fch_ + fch_) / stepch_ - gnew_ - gdmi_;
  ddspln_ = (gnew_ - gdmi_) / stepch_;
  if (nfun_ <= nfopt_) {
    if (gdmi_ * gnew_ <= 0.0) sbound_ = stmin_;
    stmin_ = step_;
    gdmi_ = gnew_;
    stepch_ = -stepch_;
  } else {
    sbound_ = step_;
  }
  if (fch_ != 0.0) ddspln_ = ddspln_ + (wo + wo) / stepch_;
  if (gdmi_ == 0.0) {
    end_search();
    return;
  }
  if (nfun_ <= nfbeg_ + 1) {
    next_step();
    return;
  }
  if (std::fabs(gdmi_ / gdit_) <= kEpsCg) {
    end_search();
    return;
  }
  limit_or_step();
}

void ConjugateGradient::limit_or_step() {
  if (nfun_ >= nfopt_ + kMaxlin) {
    ier_ = 2;
    end_search();
    return;
  }
  next_step();
}

void ConjugateGradient::next_step() {
  if (sbound_ < -0.5)
    stepch_ = 9.0 * stmin_;
  else
    stepch_ = 0.5 * (sbound_ - stmin_);
  const double gspln = gdmi_ + stepch_ * ddspln_;
  if (gdmi_ * gspln < 0.0) stepch_ = stepch_ * gdmi_ / (gdmi_ - gspln);
  trial();
}

void ConjugateGradient::end_search() {
  if (nfun_ != nfopt_) {  // back to the best point of the search
    f_ = fmin_;
    x_ = xa_;
    g_ = ga_;
  }
  if (ier_ == 0) {
    new_direction();
    return;
  }
  ++nfun_;
  state_ = kEvalFinal;
}

void ConjugateGradient::new_direction() {
  double sum = 0.0;
  for (int i = 0; i < n_; ++i) sum = sum + g_[i] * gt_[i];
  const double beta = (gsqrd_ - sum) / (gdmi_ - gdit_);
  if (std::fabs(beta * gdmi_) > gsq2_) {
    ++iretry_;
    if (iretry_ <= 0) {
      limit_or_step();
      return;
    }
  }
  if (f_ < fuit_) iterfm_ = iter_;
  if (iter_ >= iterfm_ + kMxfcon) {  // line searches no longer reduce the function
    ier_ = 4;
    state_ = kDone;
    return;
  }
  dfpr_ = stmin_ * gdit_;
  if (iretry_ > 0) {  // steepest descent
    for (int i = 0; i < n_; ++i) d_[i] = -g_[i];
    iterrs_ = 0;
  } else {
    if (iterrs_ != 0 && (iter_ - iterrs_) < (n_ - 1) && std::fabs(sum) < gsq2_) {
      double gama = 0.0;
      sum = 0.0;
      for (int i = 0; i < n_; ++i) {
        const double t = g_[i];
        gama = gama + t * yt_[i];
        sum = sum + t * dt_[i];
      }
      gama = gama / gamden_;
      if (std::fabs(beta * gdmi_ + gama * sum) < gsq2_) {  // three-term recurrence (Powell restart)
        for (int i = 0; i < n_; ++i) d_[i] = -g_[i] + beta * d_[i] + gama * dt_[i];
        iteration();
        return;
      }
    }
    gamden_ = gdmi_ - gdit_;  // new restart direction
    for (int i = 0; i < n_; ++i) {
      const double t = g_[i], di = d_[i];
      dt_[i] = di;
      yt_[i] = t - gt_[i];
      d_[i] = -t + beta * di;
    }
    iterrs_ = iter_;
  }
  iteration();
}

}  // namespace smmp
