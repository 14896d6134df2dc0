// Optimizers of the training loop on the device, FP64: one state machine per element network
// (TOptimizer_next, optimizers.F90:184-221), one block per network and one launch per training
// iteration, acting on the resident parameters and the gradient of the last training call.
//
//  * steepest descent (steepdesc.F90:125-204): x -= lr g, scaled to the maximal displacement;
//  * line minimiser (linemin.F90): bracketing by golden-ratio growth, then cubic, secant or
//    bisection steps on the directional derivative; steps capped at the maximal displacement;
//  * conjugate gradients (conjgrad.F90): Polak-Ribiere directions, one line minimisation per
//    direction, first step 5 |g|;
//  * L-BFGS (lbfgs.F90:283-488): two-loop recursion over the stored (s, y) pairs, then a line
//    minimisation with first step 1, or the plain quasi-Newton step;
//  * FIRE (fire.F90, Bitzek et al., PRL 97, 170201 (2006)).
//
// Every thread of a block evaluates the same scalar control flow from block reductions that sum
// in a fixed order (so the step is reproducible); vector updates are spread over the threads.
// Thread 0 writes the scalar state back at the end of the launch.  Products and sums are not
// contracted into FMAs: the line minimiser's cubic step (b^2 - 4ac and the like) can cancel,
// and it must round as the reference's separate operations do.
#include <cfloat>
#include <cmath>

#include "../../include/fnet_hip.h"
#include "fnet_internal.hpp"

#pragma clang fp contract(off)

namespace fnet {
namespace {

constexpr int kOptThreads = 256;
constexpr double kEps = DBL_EPSILON;
constexpr double kGoldenMeanP1 = 1.6180339887498949;  // (1 + sqrt(5)) / 2, constants.F90

struct OptVecs {
  double* x0;    // line minimiser: start point, then the minimum
  double* d0;    // line minimiser: unit direction, then the gradient at the minimum
  double* u;     // CG: current point
  double* gg;    // CG: -gradient of the last line minimum
  double* hh;    // CG: search direction
  double* xold;  // L-BFGS
  double* gold;
  double* dir;
  double* q;
  double* v;     // FIRE velocities
  double* s;     // L-BFGS pairs [mem][n]
  double* y;
  double* rho;   // [mem]
  double* alpha; // [mem]
};

__device__ OptVecs vecs_of(double* base, long long n, int mem) {
  OptVecs v;
  double* p = base;
  v.x0 = p; p += n;
  v.d0 = p; p += n;
  v.u = p; p += n;
  v.gg = p; p += n;
  v.hh = p; p += n;
  v.xold = p; p += n;
  v.gold = p; p += n;
  v.dir = p; p += n;
  v.q = p; p += n;
  v.v = p; p += n;
  v.s = p; p += (long long)mem * n;
  v.y = p; p += (long long)mem * n;
  v.rho = p; p += mem;
  v.alpha = p;
  return v;
}

// block reductions: thread partials over k = tid, tid + 256, ... in order, then a fixed tree
__device__ double block_sum(double v, double* red) {
  const int t = threadIdx.x;
  red[t] = v;
  __syncthreads();
  for (int st = kOptThreads / 2; st > 0; st >>= 1) {
    if (t < st) red[t] += red[t + st];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

__device__ double block_max(double v, double* red) {
  const int t = threadIdx.x;
  red[t] = v;
  __syncthreads();
  for (int st = kOptThreads / 2; st > 0; st >>= 1) {
    if (t < st) red[t] = fmax(red[t], red[t + st]);
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

struct Ctx {
  long long n;
  double* red;
  __device__ double dot(const double* a, const double* b) const {
    double s = 0.0;
    for (long long k = threadIdx.x; k < n; k += kOptThreads) s += a[k] * b[k];
    return block_sum(s, red);
  }
  __device__ double maxabs(const double* a) const {
    double m = 0.0;
    for (long long k = threadIdx.x; k < n; k += kOptThreads) m = fmax(m, fabs(a[k]));
    return block_max(m, red);
  }
  __device__ void copy(double* d, const double* s) const {
    for (long long k = threadIdx.x; k < n; k += kOptThreads) d[k] = s[k];
    __syncthreads();
  }
};

// ---- line minimiser (linemin.F90) ----
__device__ void lm_reset(OptState& st, const OptVecs& v, const Ctx& c, const double* x0,
                         const double* d, double first_step) {
  st.lm_state = 1;
  st.lm_it = 0;
  st.lm_xx[0] = st.lm_xx[1] = 0.0;
  st.lm_dx[0] = st.lm_dx[1] = 0.0;
  st.lm_xcur = 0.0;
  st.lm_first = first_step;
  st.lm_conv = 0;
  const double nrm = sqrt(c.dot(d, d));
  for (long long k = threadIdx.x; k < c.n; k += kOptThreads) {
    v.x0[k] = x0[k];
    v.d0[k] = d[k] / nrm;
  }
  __syncthreads();
  st.lm_maxx = st.max_disp / c.maxabs(v.d0);
}

// One line-minimiser iteration at the point whose gradient is g; the next point goes to x.
// Returns 1 when the minimum is reached (x = the minimum, d0 = its gradient).
__device__ int lm_next(OptState& st, const OptVecs& v, const Ctx& c, double fx, const double* g,
                       double* x) {
  if (st.lm_conv) {
    c.copy(x, v.x0);
    return 1;
  }
  st.lm_it += 1;
  double dcur = c.dot(g, v.d0);
  if ((fabs(dcur) < st.lm_tol && st.lm_state == 3) || st.lm_it > st.lm_maxiter) {
    for (long long k = threadIdx.x; k < c.n; k += kOptThreads) {
      v.x0[k] = v.x0[k] + st.lm_xcur * v.d0[k];
      v.d0[k] = g[k];
      x[k] = v.x0[k];
    }
    __syncthreads();
    st.lm_xx[0] = fx;
    st.lm_xx[1] = st.lm_xcur;
    st.lm_conv = 1;
    return 1;
  }
  int state = st.lm_state;
  double xnew = 0.0;
  int nxt = state;
  if (state == 1) {
    if (dcur > 0.0) {
      for (long long k = threadIdx.x; k < c.n; k += kOptThreads) v.d0[k] = -v.d0[k];
      __syncthreads();
      dcur = -dcur;
    }
    st.lm_xx[0] = st.lm_xcur;
    st.lm_dx[0] = dcur;
    xnew = st.lm_first;
    nxt = 2;
  }
  if (state == 2) {
    if (dcur < 0.0) {
      xnew = st.lm_xcur + kGoldenMeanP1 * (st.lm_xcur - st.lm_xx[0]);
      st.lm_xx[0] = st.lm_xcur;
      st.lm_dx[0] = dcur;
      nxt = 2;
    } else {
      state = 3;
    }
  }
  if (state == 3) {
    bool done = false;
    nxt = 3;
    const double x1 = st.lm_xx[0], x2 = st.lm_xx[1], d1 = st.lm_dx[0], d2 = st.lm_dx[1];
    const double xc = st.lm_xcur;
    if (st.lm_it > 2 && fabs(x1 - x2) > kEps) {
      const double qq = (xc - x1) / (x1 - x2);
      double tmp = dcur - d1;
      double cc = qq * qq * (d2 - d1);
      double aa = qq * tmp + cc;
      const double bb = 2.0 * qq * tmp + cc + tmp;
      cc = (qq + 1.0) * dcur;
      tmp = bb * bb - 4 * aa * cc;
      if (tmp >= 0.0) {
        tmp = sqrt(tmp);
        aa = bb > 0.0 ? bb + tmp : bb - tmp;
        if (fabs(aa) > kEps) {
          const double xn = xc - (xc - x1) * 2.0 * cc / aa;
          if (((x1 < xn) == (xn < x2)) && ((dcur < 0.0) == (xn > xc))) {
            if (xn > xc) {
              st.lm_xx[0] = xc;
              st.lm_dx[0] = dcur;
            } else {
              st.lm_xx[1] = xc;
              st.lm_dx[1] = dcur;
            }
            xnew = xn;
            done = true;
          }
        }
      }
    }
    if (!done) {
      if (dcur < 0.0) {
        st.lm_xx[0] = xc;
        st.lm_dx[0] = dcur;
      } else {
        st.lm_xx[1] = xc;
        st.lm_dx[1] = dcur;
      }
      const double tmp = st.lm_dx[0] - st.lm_dx[1];
      if (fabs(tmp) > kEps) {
        const double xn = (st.lm_xx[1] * st.lm_dx[0] - st.lm_xx[0] * st.lm_dx[1]) / tmp;
        if ((st.lm_xx[0] < xn) == (xn < st.lm_xx[1])) {
          xnew = xn;
          done = true;
        }
      }
    }
    if (!done) xnew = 0.5 * (st.lm_xx[0] + st.lm_xx[1]);
  }
  const double step = xnew - st.lm_xcur;
  if (fabs(step) <= st.lm_maxx) st.lm_xcur = xnew;
  else st.lm_xcur = st.lm_xcur + copysign(st.lm_maxx, step);
  st.lm_state = nxt;
  for (long long k = threadIdx.x; k < c.n; k += kOptThreads) x[k] = v.x0[k] + st.lm_xcur * v.d0[k];
  __syncthreads();
  return 0;
}

// ---- steepest descent (steepdesc.F90:160-204) ----
__device__ int sd_next(OptState& st, const Ctx& c, const double* g, double* x) {
  const double gmax = c.maxabs(g);
  if (st.converged || gmax < st.tol) {
    st.converged = 1;
    return 1;
  }
  const double m = st.lr * gmax;  // max |-lr g|
  const double f = m <= st.max_disp ? 1.0 : st.max_disp / m;
  for (long long k = threadIdx.x; k < c.n; k += kOptThreads) {
    const double step = -1.0 * st.lr * g[k];
    x[k] = f == 1.0 ? x[k] + step : x[k] + f * step;
  }
  __syncthreads();
  return 0;
}

// ---- conjugate gradients (conjgrad.F90) ----
__device__ int cg_next(OptState& st, const OptVecs& v, const Ctx& c, double fx, const double* g,
                       double* x) {
  if (st.converged) {
    c.copy(x, v.u);
    return 1;
  }
  if (st.cg_state == 1) {
    if (c.maxabs(g) < st.tol) {
      st.converged = 1;
      st.cg_state = 2;
      c.copy(x, v.u);
      return 1;
    }
    for (long long k = threadIdx.x; k < c.n; k += kOptThreads) {
      v.gg[k] = -g[k];
      v.hh[k] = -g[k];
    }
    __syncthreads();
    lm_reset(st, v, c, v.u, v.hh, 5.0 * sqrt(c.dot(v.gg, v.gg)));
    st.cg_state = 2;
  }
  int conv_line = 1;
  while (!st.converged && conv_line) {
    conv_line = lm_next(st, v, c, fx, g, v.u);
    if (conv_line) {
      const double* xi = v.d0;  // gradient at the line minimum
      if (c.maxabs(xi) < st.tol) {
        st.converged = 1;
      } else {
        const double gg2 = c.dot(v.gg, v.gg);
        double dg2 = 0.0;
        {
          double s = 0.0;
          for (long long k = threadIdx.x; k < c.n; k += kOptThreads) s += (xi[k] + v.gg[k]) * xi[k];
          dg2 = block_sum(s, c.red);
        }
        if (gg2 < kEps) {
          st.converged = 1;
        } else {
          for (long long k = threadIdx.x; k < c.n; k += kOptThreads) {
            v.gg[k] = -xi[k];
            v.hh[k] = v.gg[k] + (dg2 / gg2) * v.hh[k];
            v.u[k] = v.x0[k];
          }
          __syncthreads();
          lm_reset(st, v, c, v.u, v.hh, st.lm_xx[1]);
        }
      }
    }
  }
  if (st.converged) c.copy(v.u, v.x0);
  c.copy(x, v.u);
  return st.converged;
}

// ---- L-BFGS (lbfgs.F90:283-488) ----
__device__ void lbfgs_direction(OptState& st, const OptVecs& v, const Ctx& c, const double* x,
                                const double* g) {
  const long long n = c.n;
  const int mem = st.mem;
  if (st.lb_it == 0) {
    for (long long k = threadIdx.x; k < n; k += kOptThreads) {
      v.gold[k] = g[k];
      v.dir[k] = -g[k];
    }
    __syncthreads();
    st.lb_it = 1;
    return;
  }
  const int nw = (st.lb_it - 1) % mem;
  double* sn = v.s + (long long)nw * n;
  double* yn = v.y + (long long)nw * n;
  for (long long k = threadIdx.x; k < n; k += kOptThreads) {
    sn[k] = x[k] - v.xold[k];
    yn[k] = g[k] - v.gold[k];
    v.xold[k] = x[k];
    v.gold[k] = g[k];
  }
  __syncthreads();
  const double beta = c.dot(yn, sn);
  v.rho[nw] = 1.0 / fmax(fabs(beta), kEps);  // every thread writes the same value
  __syncthreads();
  const double diag = st.lb_it == 1 ? 1.0 : c.dot(sn, yn) / fmax(c.dot(yn, yn), kEps);
  // pairs stored so far; pairs never written have rho = s = y = 0 and change nothing
  const int stored = st.lb_it < mem ? st.lb_it : mem;
  c.copy(v.q, g);
  for (int m = 0; m < stored; ++m) {  // newest to oldest
    const int i = ((nw - m) % mem + mem) % mem;
    const double a = v.rho[i] * c.dot(v.s + (long long)i * n, v.q);
    const double* yi = v.y + (long long)i * n;
    for (long long k = threadIdx.x; k < n; k += kOptThreads) v.q[k] = v.q[k] - a * yi[k];
    if (threadIdx.x == 0) v.alpha[i] = a;
    __syncthreads();
  }
  for (long long k = threadIdx.x; k < n; k += kOptThreads) v.dir[k] = diag * v.q[k];
  __syncthreads();
  const int oldest = stored < mem ? 0 : (nw + 1) % mem;
  for (int m = 0; m < stored; ++m) {  // oldest to newest
    const int i = (oldest + m) % mem;
    const double bb = v.rho[i] * c.dot(v.y + (long long)i * n, v.dir);
    const double* si = v.s + (long long)i * n;
    const double f = v.alpha[i] - bb;
    for (long long k = threadIdx.x; k < n; k += kOptThreads) v.dir[k] = v.dir[k] + si[k] * f;
    __syncthreads();
  }
  for (long long k = threadIdx.x; k < n; k += kOptThreads) v.dir[k] = -v.dir[k];
  __syncthreads();
  st.lb_it += 1;
}

__device__ int lbfgs_next(OptState& st, const OptVecs& v, const Ctx& c, double fx,
                          const double* g, double* x) {
  if (c.maxabs(g) < st.tol) return 1;
  const double* gtmp = g;
  if (st.line_search && st.lb_it > 0) {
    const int conv = lm_next(st, v, c, fx, g, x);
    if (!conv) return 0;
    gtmp = v.d0;
  }
  lbfgs_direction(st, v, c, x, gtmp);
  if (st.line_search) {
    lm_reset(st, v, c, x, v.dir, 1.0);
    (void)lm_next(st, v, c, fx, g, x);
  } else {
    // the quasi-Newton step itself (no displacement cap: isQNDisp is false, lbfgs.F90)
    for (long long k = threadIdx.x; k < c.n; k += kOptThreads) x[k] = x[k] + v.dir[k];
    __syncthreads();
  }
  return 0;
}

// ---- FIRE (fire.F90) ----
__device__ int fire_next(OptState& st, const OptVecs& v, const Ctx& c, const double* g,
                         double* x) {
  const int conv = c.maxabs(g) < st.tol;
  if (!conv) {
    const double p = -c.dot(v.v, g);
    const double vn = sqrt(c.dot(v.v, v.v)), gn = sqrt(c.dot(g, g));
    for (long long k = threadIdx.x; k < c.n; k += kOptThreads)
      v.v[k] = (1.0 - st.f_a) * v.v[k] - st.f_a * g[k] * vn / gn;
    __syncthreads();
    if (p < 0.0) {
      for (long long k = threadIdx.x; k < c.n; k += kOptThreads) v.v[k] = 0.0;
      __syncthreads();
      st.f_reset = st.f_it;
      st.f_dt *= 0.5;    // f_dec
      st.f_a = 0.1;      // a_start
    } else if (st.f_it - st.f_reset > 5) {  // n_min
      st.f_dt = fmin(st.f_dt * 1.1, st.max_disp);  // f_inc, dt_max
      st.f_a *= 0.99;    // f_alpha
    }
    for (long long k = threadIdx.x; k < c.n; k += kOptThreads) {
      v.v[k] = v.v[k] - st.f_dt * g[k];
      x[k] = x[k] + v.v[k] * st.f_dt;
    }
    __syncthreads();
  }
  st.f_it += 1;
  return conv;
}

__global__ void __launch_bounds__(kOptThreads)
    k_opt_step(OptState* __restrict__ states, double* __restrict__ params,
               const double* __restrict__ grad, double* __restrict__ vecs, long long n,
               long long vstride, double fx, int* __restrict__ conv_out) {
  __shared__ double red[kOptThreads];
  const int sp = blockIdx.x;
  OptState st = states[sp];
  const OptVecs v = vecs_of(vecs + (long long)sp * vstride, n, st.mem);
  Ctx c;
  c.n = n;
  c.red = red;
  double* x = params + (long long)sp * n;
  const double* g = grad + (long long)sp * n;
  int conv = 0;
  switch (st.kind) {
    case FNET_OPT_SD: conv = sd_next(st, c, g, x); break;
    case FNET_OPT_CG: conv = cg_next(st, v, c, fx, g, x); break;
    case FNET_OPT_LBFGS: conv = lbfgs_next(st, v, c, fx, g, x); break;
    default: conv = fire_next(st, v, c, g, x); break;
  }
  if (threadIdx.x == 0) {
    states[sp] = st;
    conv_out[sp] = conv;
  }
}

// start state: x (the resident parameters) copied into the vectors that need it
__global__ void __launch_bounds__(kOptThreads)
    k_opt_init(const double* __restrict__ params, double* __restrict__ vecs, long long n,
               long long vstride, int mem) {
  const int sp = blockIdx.x;
  const OptVecs v = vecs_of(vecs + (long long)sp * vstride, n, mem);
  const double* x = params + (long long)sp * n;
  for (long long k = threadIdx.x; k < n; k += kOptThreads) {
    v.u[k] = x[k];
    v.xold[k] = x[k];
    v.gold[k] = 0.0;
    v.v[k] = 0.0;
  }
  for (long long k = threadIdx.x; k < (long long)mem * n; k += kOptThreads) {
    v.s[k] = 0.0;
    v.y[k] = 0.0;
  }
  for (int k = threadIdx.x; k < mem; k += kOptThreads) {
    v.rho[k] = 0.0;
    v.alpha[k] = 0.0;
  }
}

}  // namespace

long long opt_vec_doubles(long long n, int mem) { return 10 * n + 2 * (long long)mem * n + 2 * (long long)mem; }

hipError_t launch_opt_init(const double* params, double* vecs, long long n, long long vstride,
                           int mem, int nspecies, hipStream_t s) {
  k_opt_init<<<nspecies, kOptThreads, 0, s>>>(params, vecs, n, vstride, mem);
  return hipGetLastError();
}

hipError_t launch_opt_step(OptState* states, double* params, const double* grad, double* vecs,
                           long long n, long long vstride, double fx, int* conv, int nspecies,
                           hipStream_t s) {
  k_opt_step<<<nspecies, kOptThreads, 0, s>>>(states, params, grad, vecs, n, vstride, fx, conv);
  return hipGetLastError();
}

}  // namespace fnet
