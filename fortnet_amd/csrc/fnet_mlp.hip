// Element networks on gfx950: all atoms of one element form tiles; one block runs a tile through
// every layer with the activations in LDS (network.F90:146-180 fprop, :314-347 iPredict), then
// back to the input for the force adjoint (fdevi, network.F90:183-244, taken in reverse mode:
// nOut reverse sweeps give the same Jacobian rows), or back through the weights for the
// training gradient (bprop, network.F90:248-296, with the sysTrain normalisation of
// bpnn.F90:692-702).  Per-structure energies and weight gradients are reduced in FP64.
#include <algorithm>

#include "fnet_internal.hpp"

namespace fnet {
namespace {

template <typename T>
__device__ __forceinline__ T act_f(int id, T x) {
  switch (id) {
    case 7: return exp(-(x * x));                                   // gaussian
    case 4: return log((T)1 + exp(x));                              // softplus
    case 5: return (sqrt(x * x + (T)1) - (T)1) / (T)2 + x;          // bent identity
    case 2: return x > (T)0 ? x : (T)0;                             // relu
    case 3: return ((T)0.01 * x > x) ? (T)0.01 * x : x;             // leaky relu
    case 1: return (T)1 / ((T)1 + exp(-x));                         // sigmoid
    case 8: return x > (T)0 ? (T)1 : (T)0;                          // heaviside
    case 6: return atan(x);                                         // atan
    case 9: return x;                                               // linear
    default: return tanh(x);                                        // tanh
  }
}

template <typename T>
__device__ __forceinline__ T act_d(int id, T x) {
  switch (id) {
    case 7: return (T)-2 * x * exp(-(x * x));
    case 4: return (T)1 / ((T)1 + exp(-x));
    case 5: return x / ((T)2 * sqrt(x * x + (T)1)) + (T)1;
    case 2: return x >= (T)0 ? (T)1 : (T)0;
    case 3: return x >= (T)0 ? (T)1 : (T)0.01;
    case 1: {
      const T s = (T)1 / ((T)1 + exp(-x));
      return s * ((T)1 - s);
    }
    case 8: return (T)0;
    case 6: return (T)1 / (x * x + (T)1);
    case 9: return (T)1;
    default: {
      const T t = tanh(x);
      return (T)1 - t * t;
    }
  }
}

struct MlpArgs {
  NetShape sh;
  const void* params; // [nspecies][nparams] T
  const double* G;    // [N][nG]
  const double* ext;
  int next;
  const int* ext_in;  // [next_in]
  int nG, next_in, nout;
  const double* zmean;
  const double* zinv;
  const int* perm;
  const int* tile_sp;
  const int* tile_start;
  const int* tile_cnt;
  int first_tile;
  int TA;
  double* Eatom;      // [N][nout]
  double* cG;         // [N][nout][nG]
  const double* lossgrad; // [N][nout]
  double* grad;       // [nspecies][nparams]
};

// LDS per tile: input x [TA][d0], activations of every layer [TA][d_l], f'(z) of hidden layers,
// and two delta buffers [TA][dmax].
template <typename T>
__device__ __forceinline__ void load_input(const MlpArgs& a, T* x, int cnt, int start) {
  const int d0 = a.sh.dims[0];
  for (int idx = threadIdx.x; idx < a.TA * d0; idx += blockDim.x) {
    const int at = idx % a.TA, f = idx / a.TA;
    T v = 0;
    if (at < cnt) {
      const long long atom = a.perm[start + at];
      if (f < a.nG) {
        v = (T)((a.G[atom * a.nG + f] - a.zmean[f]) * a.zinv[f]);
      } else {
        v = (T)a.ext[atom * a.next + a.ext_in[f - a.nG]];
      }
    }
    x[at * d0 + f] = v;
  }
}

// fprop of one tile.  acts[l] = [TA][d_l] (acts[0] = x), fp[l] = f'(z_l) for hidden layers.
template <typename T>
__device__ void forward(const MlpArgs& a, const T* P, int act, T** acts, T** fp) {
  const int nl = a.sh.nl;
  for (int l = 1; l < nl; ++l) {
    const int din = a.sh.dims[l - 1], dout = a.sh.dims[l];
    const T* W = P + a.sh.woff[l - 1];
    const T* bb = P + a.sh.boff[l];
    const T* in = acts[l - 1];
    T* out = acts[l];
    const bool hidden = l < nl - 1;
    for (int idx = threadIdx.x; idx < a.TA * dout; idx += blockDim.x) {
      const int at = idx % a.TA, o = idx / a.TA;
      const T* w = W + (long long)o * din;
      const T* xi = in + at * din;
      T z = 0;
      for (int k = 0; k < din; ++k) z += w[k] * xi[k];
      z += bb[o];
      if (hidden) {
        out[at * dout + o] = act_f(act, z);
        fp[l][at * dout + o] = act_d(act, z);
      } else {
        out[at * dout + o] = z;
      }
    }
    __syncthreads();
  }
}

template <typename T>
__device__ __forceinline__ size_t mlp_lds_floats(const NetShape& sh, int TA) {
  size_t s = 0;
  int dmax = 0;
  for (int l = 0; l < sh.nl; ++l) {
    s += sh.dims[l];
    if (l > 0 && l < sh.nl - 1) s += sh.dims[l];
    if (sh.dims[l] > dmax) dmax = sh.dims[l];
  }
  s += 2 * dmax;
  return s * TA;
}

template <typename T, bool GRADIN>
__global__ void __launch_bounds__(256) k_mlp_fwd(MlpArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int tile = a.first_tile + blockIdx.x;
  const int sp = a.tile_sp[tile], start = a.tile_start[tile], cnt = a.tile_cnt[tile];
  const int nl = a.sh.nl;
  const T* P = (const T*)a.params + (long long)sp * a.sh.nparams;
  const int act = a.sh.act[sp];
  T* acts[kMaxLayers];
  T* fp[kMaxLayers];
  T* p = (T*)smem;
  int dmax = 0;
  for (int l = 0; l < nl; ++l) {
    acts[l] = p;
    p += a.TA * a.sh.dims[l];
    if (a.sh.dims[l] > dmax) dmax = a.sh.dims[l];
  }
  for (int l = 0; l < nl; ++l) {
    fp[l] = nullptr;
    if (l > 0 && l < nl - 1) {
      fp[l] = p;
      p += a.TA * a.sh.dims[l];
    }
  }
  T* del0 = p;
  T* del1 = p + a.TA * dmax;

  load_input(a, acts[0], cnt, start);
  __syncthreads();
  forward(a, P, act, acts, fp);
  const int nout = a.sh.dims[nl - 1];
  for (int idx = threadIdx.x; idx < cnt * nout; idx += blockDim.x) {
    const int at = idx % cnt, o = idx / cnt;
    a.Eatom[(long long)a.perm[start + at] * nout + o] = (double)acts[nl - 1][at * nout + o];
  }
  if (!GRADIN) return;
  for (int t = 0; t < nout; ++t) {
    // reverse sweep for output t: delta = e_t at the (linear) output layer
    T* dcur = del0;
    T* dnext = del1;
    for (int idx = threadIdx.x; idx < a.TA * nout; idx += blockDim.x) {
      const int at = idx % a.TA, o = idx / a.TA;
      dcur[at * nout + o] = (o == t) ? (T)1 : (T)0;
    }
    __syncthreads();
    for (int l = nl - 1; l >= 1; --l) {
      const int din = a.sh.dims[l - 1], dout = a.sh.dims[l];
      const T* W = P + a.sh.woff[l - 1];
      for (int idx = threadIdx.x; idx < a.TA * din; idx += blockDim.x) {
        const int at = idx % a.TA, k = idx / a.TA;
        T s = 0;
        for (int o = 0; o < dout; ++o) s += W[(long long)o * din + k] * dcur[at * dout + o];
        if (l - 1 >= 1) s *= fp[l - 1][at * din + k];
        dnext[at * din + k] = s;
      }
      __syncthreads();
      T* tmp = dcur;
      dcur = dnext;
      dnext = tmp;
    }
    // dcur = dOut_t/dx  [TA][d0]; only the ACSF part feeds the force adjoint
    for (int idx = threadIdx.x; idx < cnt * a.nG; idx += blockDim.x) {
      const int at = idx % cnt, f = idx / cnt;
      const long long atom = a.perm[start + at];
      a.cG[(atom * a.nout + t) * a.nG + f] = (double)dcur[at * a.sh.dims[0] + f] * a.zinv[f];
    }
    __syncthreads();
  }
}

// Weight gradient of one tile (bprop), summed over the tile's atoms in T and written as the
// tile's FP64 partial vector (a.grad + tile * nparams); k_grad_reduce adds the partials of each
// species in tile order.  Entries that receive no gradient (the input layer's biases and the
// dummy output weights, network.F90:397-427) are written as zero.
template <typename T>
__global__ void __launch_bounds__(256) k_mlp_grad(MlpArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int tile = a.first_tile + blockIdx.x;
  const int sp = a.tile_sp[tile], start = a.tile_start[tile], cnt = a.tile_cnt[tile];
  const int nl = a.sh.nl;
  const T* P = (const T*)a.params + (long long)sp * a.sh.nparams;
  double* g = a.grad + (long long)tile * a.sh.nparams;
  for (int k = threadIdx.x; k < a.sh.dims[nl - 1]; k += blockDim.x) g[a.sh.woff[nl - 1] + k] = 0.0;
  for (int k = threadIdx.x; k < a.sh.dims[0]; k += blockDim.x) g[a.sh.boff[0] + k] = 0.0;
  const int act = a.sh.act[sp];
  T* acts[kMaxLayers];
  T* fp[kMaxLayers];
  T* p = (T*)smem;
  int dmax = 0;
  for (int l = 0; l < nl; ++l) {
    acts[l] = p;
    p += a.TA * a.sh.dims[l];
    if (a.sh.dims[l] > dmax) dmax = a.sh.dims[l];
  }
  for (int l = 0; l < nl; ++l) {
    fp[l] = nullptr;
    if (l > 0 && l < nl - 1) {
      fp[l] = p;
      p += a.TA * a.sh.dims[l];
    }
  }
  T* dcur = p;
  T* dnext = p + a.TA * dmax;
  load_input(a, acts[0], cnt, start);
  __syncthreads();
  forward(a, P, act, acts, fp);
  const int nout = a.sh.dims[nl - 1];
  for (int idx = threadIdx.x; idx < a.TA * nout; idx += blockDim.x) {
    const int at = idx % a.TA, o = idx / a.TA;
    dcur[at * nout + o] =
        (at < cnt) ? (T)a.lossgrad[(long long)a.perm[start + at] * nout + o] : (T)0;
  }
  __syncthreads();
  for (int l = nl - 1; l >= 1; --l) {
    const int din = a.sh.dims[l - 1], dout = a.sh.dims[l];
    // db(l) and dW(l-1) = a_{l-1} delta_l^T summed over the tile
    for (int o = threadIdx.x; o < dout; o += blockDim.x) {
      T s = 0;
      for (int at = 0; at < cnt; ++at) s += dcur[at * dout + o];
      g[a.sh.boff[l] + o] = (double)s;
    }
    for (int idx = threadIdx.x; idx < din * dout; idx += blockDim.x) {
      const int k = idx % din, o = idx / din;
      T s = 0;
      for (int at = 0; at < cnt; ++at) s += acts[l - 1][at * din + k] * dcur[at * dout + o];
      g[a.sh.woff[l - 1] + (long long)o * din + k] = (double)s;
    }
    if (l == 1) break;
    const T* W = P + a.sh.woff[l - 1];
    for (int idx = threadIdx.x; idx < a.TA * din; idx += blockDim.x) {
      const int at = idx % a.TA, k = idx / a.TA;
      T s = 0;
      for (int o = 0; o < dout; ++o) s += W[(long long)o * din + k] * dcur[at * dout + o];
      dnext[at * din + k] = s * fp[l - 1][at * din + k];
    }
    __syncthreads();
    T* tmp = dcur;
    dcur = dnext;
    dnext = tmp;
  }
}

// grad[sp][k] = sum over the tiles of species sp, in tile order, of the tiles' partials, in two
// fixed-order levels: chunks of kGradChunk tiles (one thread per (chunk, species, k)), then the
// chunks in order.
constexpr int kGradChunk = 64;

// Partial t belongs to tile t / nsub (the MFMA path writes nsub FP32 partials per tile; their
// values are FP32 sums, so FP32 storage loses nothing, and they are added in FP64 from here on).
template <typename P>
__global__ void k_grad_chunks(const P* __restrict__ gpart, const int* __restrict__ tile_sp,
                              int nparts, int nsub, long long nparams, int nspecies,
                              double* __restrict__ gchunk) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nparams * nspecies) return;
  const int c = blockIdx.y;
  const int sp = (int)(q / nparams);
  const long long k = q - (long long)sp * nparams;
  const int t0 = c * kGradChunk, t1 = min(nparts, t0 + kGradChunk);
  double v = 0.0;
  for (int t = t0; t < t1; ++t)
    if (tile_sp[t / nsub] == sp) v += (double)gpart[(long long)t * nparams + k];
  gchunk[(long long)c * nparams * nspecies + q] = v;
}

__global__ void k_grad_reduce(const double* __restrict__ gchunk, int nchunks, long long n,
                              double* __restrict__ grad) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  double v = 0.0;
  for (int c = 0; c < nchunks; ++c) v += gchunk[(long long)c * n + q];
  grad[q] = v;
}

// Per-structure sums of the atomic outputs (predictBatch sums, bpnn.F90:1040-1047;
// forces.F90:381-385) in FP64, in two passes with a fixed order: each block sums one chunk of
// kEnergyChunk atoms of one structure (thread t takes atoms t, t + 256, ... in order; the four
// wave sums combine in a fixed order), then one thread per (structure, output) adds its chunks'
// partials in chunk order.  The result does not depend on scheduling, and a large cell is spread
// over many blocks instead of one.
__global__ void __launch_bounds__(256) k_energy_partial(const double* __restrict__ Eatom,
                                                       const long long* __restrict__ off,
                                                       const int* __restrict__ chunk_struct,
                                                       const int* __restrict__ chunk_first,
                                                       int c0, int nout,
                                                       double* __restrict__ Epart) {
  __shared__ double ws[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c = c0 + blockIdx.x;
  const int s = chunk_struct[c];
  const long long a0 = off[s] + (long long)(c - chunk_first[s]) * kEnergyChunk;
  const long long a1 = min(a0 + (long long)kEnergyChunk, off[s + 1]);
  for (int t = 0; t < nout; ++t) {
    double v = 0.0;
    for (long long i = a0 + threadIdx.x; i < a1; i += 256) v += Eatom[i * nout + t];
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    if (lane == 0) ws[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) Epart[(long long)c * nout + t] = (ws[0] + ws[1]) + (ws[2] + ws[3]);
    __syncthreads();
  }
}

__global__ void k_energy_final(const double* __restrict__ Epart,
                               const int* __restrict__ chunk_first, int s0, int s1, int nout,
                               double* __restrict__ Estruct) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long s = s0 + q / nout;
  const int t = (int)(q % nout);
  if (s >= s1) return;
  double v = 0.0;
  for (int c = chunk_first[s]; c < chunk_first[s + 1]; ++c) v += Epart[(long long)c * nout + t];
  Estruct[s * nout + t] = v;
}

// Loss-gradient functions of loss.F90:217-281, elementwise on (prediction, target):
// mse 2 (p - t); mae and rms (p - t) / |p - t| (the reference writes rms as (p - t) / sqrt((p - t)^2));
// mape 100 (p - t) / (t^2 |p / t - 1|).
__device__ __forceinline__ double loss_grad(int loss, double p, double t) {
  const double d = p - t;
  switch (loss) {
    case 1: return d / fabs(d);
    case 2: return 100.0 * d / (t * t * fabs(p / t - 1.0));
    case 3: return d / sqrt(d * d);
    default: return 2.0 * d;
  }
}

// Loss gradient of every atomic output with the sysTrain scaling (bpnn.F90:677-700) and the
// datapoint weight / normalisation of updateGradients + update (bpnn.F90:445-448,766).
__global__ void k_lossgrad(const double* __restrict__ Eatom, const double* __restrict__ Estruct,
                           const int* __restrict__ atom_struct, const long long* __restrict__ off,
                           long long N, int nout, int nglob, const double* __restrict__ gt,
                           const double* __restrict__ at, const double* __restrict__ aw,
                           const double* __restrict__ dpw, double inv_wsum, int loss,
                           double* __restrict__ lg) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int s = atom_struct[i];
  const double natom = (double)(off[s + 1] - off[s]);
  const double sc = (aw ? aw[i] : 1.0) / natom * (dpw ? dpw[s] : 1.0) * inv_wsum;
  const int natomic = nout - nglob;
  for (int t = 0; t < nout; ++t) {
    double v;
    if (t < nglob) v = loss_grad(loss, Estruct[(long long)s * nout + t], gt[(long long)s * nglob + t]);
    else v = loss_grad(loss, Eatom[i * nout + t], at[i * natomic + (t - nglob)]);
    lg[i * nout + t] = v * sc;
  }
}

// Per-item loss of loss.F90:285-366 over the targets of one item (a datapoint's global targets
// or one atom's atomic targets): mean (t - p)^2, mean |t - p|, 100 mean |(t - p) / t|, or the
// root of the mean square.
__device__ __forceinline__ double item_loss(int loss, double sum_sq, double sum_abs,
                                            double sum_rel, int n) {
  switch (loss) {
    case 1: return sum_abs / n;
    case 2: return 100.0 * sum_rel / n;
    case 3: return sqrt(sum_sq / n);
    default: return sum_sq / n;
  }
}

// Loss of the batch (maLoss / mapLoss / msLoss / rmsLoss, loss.F90:370-700) in FP64: global
// targets weighted by w_s * sum(atomic weights of s), atomic targets by w_s * atomic weight; the
// result is divided by the summed weights.  Three fixed-order levels: k_loss_chunks sums each
// 2048-atom chunk of the energy reduction (one block: thread t takes atoms t, t + 256, ... in
// order, then a fixed tree) into (atomic loss, atomic weight, summed atomic weights);
// k_loss_terms adds a structure's chunks in order and its global term; k_loss_sum adds the
// structures (thread t takes t, t + 256, ..., then a fixed tree).
__global__ void __launch_bounds__(256) k_loss_chunks(const double* __restrict__ Eatom,
                                                     const long long* __restrict__ off,
                                                     const int* __restrict__ chunk_struct,
                                                     const int* __restrict__ chunk_first,
                                                     int nout, int nglob,
                                                     const double* __restrict__ at,
                                                     const double* __restrict__ aw,
                                                     const double* __restrict__ dpw, int loss,
                                                     double* __restrict__ part) {
  __shared__ double r0[256], r1[256], r2[256];
  const int c = blockIdx.x;
  const int s = chunk_struct[c];
  const long long a0 = off[s] + (long long)(c - chunk_first[s]) * kEnergyChunk;
  const long long a1 = min(a0 + (long long)kEnergyChunk, off[s + 1]);
  const double w = dpw ? dpw[s] : 1.0;
  const int natomic = nout - nglob;
  double l = 0.0, nv = 0.0, saw = 0.0;
  for (long long i = a0 + threadIdx.x; i < a1; i += blockDim.x) {
    const double awi = aw ? aw[i] : 1.0;
    saw += awi;
    if (natomic > 0) {
      double q = 0.0, a = 0.0, r = 0.0;
      for (int o = 0; o < natomic; ++o) {
        const double tv = at[i * natomic + o];
        const double d = tv - Eatom[i * nout + nglob + o];
        q += d * d;
        a += fabs(d);
        r += fabs(d / tv);
      }
      l += w * awi * item_loss(loss, q, a, r, natomic);
      nv += w * awi;
    }
  }
  r0[threadIdx.x] = l;
  r1[threadIdx.x] = nv;
  r2[threadIdx.x] = saw;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
      r0[threadIdx.x] += r0[threadIdx.x + st];
      r1[threadIdx.x] += r1[threadIdx.x + st];
      r2[threadIdx.x] += r2[threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[3 * (long long)c] = r0[0];
    part[3 * (long long)c + 1] = r1[0];
    part[3 * (long long)c + 2] = r2[0];
  }
}

__global__ void k_loss_terms(const double* __restrict__ Estruct,
                             const int* __restrict__ chunk_first, int S, int nout, int nglob,
                             const double* __restrict__ gt, const double* __restrict__ dpw,
                             int loss, const double* __restrict__ part,
                             double* __restrict__ terms) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  double l = 0.0, nv = 0.0, saw = 0.0;
  for (int c = chunk_first[s]; c < chunk_first[s + 1]; ++c) {
    l += part[3 * (long long)c];
    nv += part[3 * (long long)c + 1];
    saw += part[3 * (long long)c + 2];
  }
  if (nglob > 0) {
    const double w = dpw ? dpw[s] : 1.0;
    double q = 0.0, a = 0.0, r = 0.0;
    for (int o = 0; o < nglob; ++o) {
      const double tv = gt[(long long)s * nglob + o];
      const double d = tv - Estruct[(long long)s * nout + o];
      q += d * d;
      a += fabs(d);
      r += fabs(d / tv);
    }
    l += w * saw * item_loss(loss, q, a, r, nglob);
    nv += w * saw;
  }
  terms[2 * (long long)s] = l;
  terms[2 * (long long)s + 1] = nv;
}

__global__ void __launch_bounds__(256) k_loss_sum(const double* __restrict__ terms, int S,
                                                  double* __restrict__ out) {
  __shared__ double rl[256], rn[256];
  double l = 0.0, nv = 0.0;
  for (int s = threadIdx.x; s < S; s += blockDim.x) {
    l += terms[2 * (long long)s];
    nv += terms[2 * (long long)s + 1];
  }
  rl[threadIdx.x] = l;
  rn[threadIdx.x] = nv;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
      rl[threadIdx.x] += rl[threadIdx.x + st];
      rn[threadIdx.x] += rn[threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = rl[0] / rn[0];
}

// Elastic-net term of the weight gradients (TWeightDerivs_elasticNetRegularization,
// nestedtypes.F90:336-370, applied in TBpnn_update before the division by the summed datapoint
// weights, bpnn.F90:750-766): g_w += scale ((1 - alpha) w + alpha sign(w)) on the weights part
// of every species' vector, sign(0) = 0; scale = strength / nWeights / sum(w_s).
__global__ void k_regularize(const double* __restrict__ params, double* __restrict__ grad,
                             long long nparams, long long nweights, int nspecies, double scale,
                             double alpha) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nweights * nspecies) return;
  const long long sp = k / nweights, q = k - sp * nweights;
  const double w = params[sp * nparams + q];
  const double sg = w > 0.0 ? 1.0 : (w < 0.0 ? -1.0 : 0.0);
  grad[sp * nparams + q] += scale * ((1.0 - alpha) * w + alpha * sg);
}

// Steepest descent of one species' parameter vector (steepdesc.F90:160-204), FP64.
__global__ void __launch_bounds__(256) k_sd_step(double* __restrict__ x,
                                                 const double* __restrict__ g, long long n,
                                                 double lr, double maxdisp, double tol,
                                                 double* __restrict__ gmax_out) {
  __shared__ double red[256];
  double* xs = x + (long long)blockIdx.x * n;
  const double* gs = g + (long long)blockIdx.x * n;
  double m = 0.0;
  for (long long k = threadIdx.x; k < n; k += blockDim.x) m = fmax(m, fabs(gs[k]));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + st]);
    __syncthreads();
  }
  const double gmax = red[0];
  const double stepmax = lr * gmax;  // max |-lr * g|
  const double f = (stepmax <= maxdisp) ? 1.0 : maxdisp / stepmax;
  if (threadIdx.x == 0) gmax_out[blockIdx.x] = gmax;
  if (gmax < tol) return;  // converged: xNew = xOld
  for (long long k = threadIdx.x; k < n; k += blockDim.x) {
    const double step = -1.0 * lr * gs[k];
    xs[k] = (f == 1.0) ? xs[k] + step : xs[k] + f * step;
  }
}

__global__ void k_to_float(const double* __restrict__ in, float* __restrict__ out, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (float)in[i];
}

template <typename T>
size_t host_lds(const NetShape& sh, int TA) {
  size_t s = 0;
  int dmax = 0;
  for (int l = 0; l < sh.nl; ++l) {
    s += sh.dims[l];
    if (l > 0 && l < sh.nl - 1) s += sh.dims[l];
    if (sh.dims[l] > dmax) dmax = sh.dims[l];
  }
  s += 2 * dmax;
  return s * TA * sizeof(T);
}

template <typename T>
MlpArgs make_args(fnet_batch* b, const DevModel& m, int first_tile) {
  MlpArgs a;
  a.sh = m.shape;
  a.params = std::is_same<T, double>::value ? (const void*)m.params : (const void*)m.params_f;
  a.G = (const double*)b->G;
  a.ext = b->ext;
  a.next = b->next;
  a.ext_in = m.ext_in;
  a.nG = m.nG;
  a.next_in = m.next_in;
  a.nout = m.nout;
  a.zmean = m.zmean;
  a.zinv = m.zinv;
  a.perm = b->perm;
  a.tile_sp = b->tile_sp;
  a.tile_start = b->tile_start;
  a.tile_cnt = b->tile_cnt;
  a.first_tile = first_tile;
  a.TA = b->tile_atoms;
  a.Eatom = b->Eatom;
  a.cG = (double*)b->cG;
  a.lossgrad = b->lossgrad;
  a.grad = nullptr;
  return a;
}

}  // namespace

int grad_chunks(int ntiles) { return std::max(1, (ntiles + kGradChunk - 1) / kGradChunk); }

int mlp_tile_atoms(const NetShape& sh, int fp64) {
  int TA = 64;
  while (TA > 1 && (fp64 ? host_lds<double>(sh, TA) : host_lds<float>(sh, TA)) > 64 * 1024) TA >>= 1;
  return TA;
}

hipError_t launch_mlp_fwd(fnet_batch* b, const DevModel& m, int first_tile, int ntiles, int fp64,
                          int want_grad_in, hipStream_t s) {
  if (ntiles <= 0) return hipSuccess;
  if (fp64) {
    MlpArgs a = make_args<double>(b, m, first_tile);
    const size_t lds = host_lds<double>(m.shape, a.TA);
    if (want_grad_in) k_mlp_fwd<double, true><<<ntiles, 256, lds, s>>>(a);
    else k_mlp_fwd<double, false><<<ntiles, 256, lds, s>>>(a);
  } else {
    MlpArgs a = make_args<float>(b, m, first_tile);
    const size_t lds = host_lds<float>(m.shape, a.TA);
    if (want_grad_in) k_mlp_fwd<float, true><<<ntiles, 256, lds, s>>>(a);
    else k_mlp_fwd<float, false><<<ntiles, 256, lds, s>>>(a);
  }
  return hipGetLastError();
}

hipError_t launch_energy_reduce(fnet_batch* b, const DevModel& m, int s0, int s1, hipStream_t s) {
  const int n = s1 - s0;
  if (n <= 0) return hipSuccess;
  const int c0 = b->h_chunk_first[s0], nc = b->h_chunk_first[s1] - c0;
  if (nc > 0)
    k_energy_partial<<<nc, 256, 0, s>>>(b->Eatom, b->off, b->chunk_struct, b->chunk_first, c0,
                                        m.nout, b->Epart);
  const long long q = (long long)n * m.nout;
  k_energy_final<<<(unsigned)((q + 255) / 256), 256, 0, s>>>(b->Epart, b->chunk_first, s0, s1,
                                                            m.nout, b->Estruct);
  return hipGetLastError();
}

hipError_t launch_lossgrad(fnet_batch* b, const DevModel& m, int n_global, const double* gt,
                           const double* at, const double* aw, const double* dpw, double inv_wsum,
                           int loss, hipStream_t s) {
  if (b->N == 0) return hipSuccess;
  k_lossgrad<<<(unsigned)((b->N + 255) / 256), 256, 0, s>>>(b->Eatom, b->Estruct, b->atom_struct,
                                                            b->off, b->N, m.nout, n_global, gt, at,
                                                            aw, dpw, inv_wsum, loss, b->lossgrad);
  return hipGetLastError();
}

hipError_t launch_loss(fnet_batch* b, const DevModel& m, int n_global, const double* gt,
                       const double* at, const double* aw, const double* dpw, int loss,
                       double* out, double* terms, double* part, hipStream_t s) {
  if (b->nchunks > 0)
    k_loss_chunks<<<b->nchunks, 256, 0, s>>>(b->Eatom, b->off, b->chunk_struct, b->chunk_first,
                                             m.nout, n_global, at, aw, dpw, loss, part);
  if (b->S > 0)
    k_loss_terms<<<(b->S + 255) / 256, 256, 0, s>>>(b->Estruct, b->chunk_first, b->S, m.nout,
                                                    n_global, gt, dpw, loss, part, terms);
  k_loss_sum<<<1, 256, 0, s>>>(terms, b->S, out);
  return hipGetLastError();
}

hipError_t launch_regularize(const DevModel& m, double* grad, double scale, double alpha,
                             hipStream_t s) {
  long long nw = 0;
  for (int l = 0; l + 1 < m.shape.nl; ++l) nw += (long long)m.shape.dims[l] * m.shape.dims[l + 1];
  nw += m.shape.dims[m.shape.nl - 1];  // the dummy output layer ww(d_n, 1)
  const long long n = nw * m.nspecies;
  if (n <= 0) return hipSuccess;
  k_regularize<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(m.params, grad, m.shape.nparams, nw,
                                                           m.nspecies, scale / (double)nw, alpha);
  return hipGetLastError();
}

int grad_parts(const fnet_batch* b, const DevModel& m, int fp64) {
  const int nsub = fp64 || !m.wpack ? 0 : mlp_grad_subtiles(m.shape, b->mtile_waves);
  return nsub ? b->nmtiles * nsub : b->ntiles;
}

hipError_t launch_mlp_grad(fnet_batch* b, const DevModel& m, int fp64, double* grad,
                           hipStream_t s) {
  const long long np = m.shape.nparams * m.nspecies;
  const int nsub = fp64 || !m.wpack ? 0 : mlp_grad_subtiles(m.shape, b->mtile_waves);
  const int nparts = nsub ? b->nmtiles * nsub : b->ntiles;
  if (nsub) {
    const hipError_t e = launch_mlp_grad_mfma(b, m, s);
    if (e != hipSuccess) return e;
  } else if (b->ntiles > 0) {
    if (fp64) {
      MlpArgs a = make_args<double>(b, m, 0);
      a.grad = b->gpart;
      k_mlp_grad<double><<<b->ntiles, 256, host_lds<double>(m.shape, a.TA), s>>>(a);
    } else {
      MlpArgs a = make_args<float>(b, m, 0);
      a.grad = b->gpart;
      k_mlp_grad<float><<<b->ntiles, 256, host_lds<float>(m.shape, a.TA), s>>>(a);
    }
  }
  const int nchunks = grad_chunks(nparts);
  double* gchunk = b->gpart + (long long)std::max(nparts, 1) * m.shape.nparams;
  if (nparts > 0) {
    const dim3 grid((unsigned)((np + 255) / 256), (unsigned)nchunks);
    if (nsub)
      k_grad_chunks<float><<<grid, 256, 0, s>>>((const float*)b->gpart, b->mtile_sp, nparts, nsub,
                                                m.shape.nparams, m.nspecies, gchunk);
    else
      k_grad_chunks<double><<<grid, 256, 0, s>>>(b->gpart, b->tile_sp, nparts, 1,
                                                 m.shape.nparams, m.nspecies, gchunk);
  }
  k_grad_reduce<<<(unsigned)((np + 255) / 256), 256, 0, s>>>(gchunk, nparts > 0 ? nchunks : 0, np,
                                                             grad);
  return hipGetLastError();
}

hipError_t launch_sd_step(double* params, const double* grad, long long n, double lr,
                          double max_disp, double tol, double* scratch, int nspecies,
                          hipStream_t s) {
  k_sd_step<<<nspecies, 256, 0, s>>>(params, grad, n, lr, max_disp, tol, scratch);
  return hipGetLastError();
}

hipError_t launch_to_float(const double* in, float* out, long long n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  k_to_float<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(in, out, n);
  return hipGetLastError();
}

}  // namespace fnet
