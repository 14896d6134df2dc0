// FP32 element networks on the matrix cores of gfx950 (v_mfma_f32_16x16x4_f32).
//
// Same mathematics as the generic network kernel (fprop, network.F90:146-180 / iPredict :314-347;
// the input gradient of fdevi, network.F90:183-244, in reverse mode, one sweep per output), laid
// out for MFMA: each wave owns 16 atoms of one element and runs them through every layer as
//   Z^T [d_out x 16] = W^T [d_out x d_in] . X^T [d_in x 16]
// in 16 x 16 output blocks with K = 4 per instruction.  W^T fragments come from a copy of the
// FP32 weights packed at model upload so that each instruction's A operand is one coalesced
// 256-byte load (L1/L2-resident, shared by all waves); X^T fragments come from the wave's
// activation rows in LDS, 4 K-steps per ds_read_b128 (the K order inside a 16-wide chunk is
// permuted to k = 16 c + 4 g + s, identically in the packed weights, so lane group g reads four
// consecutive features).  The accumulator fragment (rows = outputs 4 g + r, column = atom) is
// activated in registers, f'(z) of every hidden layer is kept in LDS in fragment order, and the
// activations go back to LDS as the next layer's rows.  The reverse sweep runs the transposed
// products delta_l^T = W_l . delta_{l+1}^T on separately packed fragments.  Products and sums are
// FP32 (as the generic FP32 kernel); per-atom outputs and input gradients are written in FP64.
//
// BF16 modes: the products whose input and output are both hidden layers (the hidden-to-hidden
// weights, forward and reverse) run on v_mfma_f32_16x16x32_bf16 with FP32 accumulation, 32 K per
// instruction.  Each lane takes 8 consecutive features of its atom's row (k = 32 c + 8 g + e) and
// the weights are packed with the same k order, so the products pair up whatever order the
// instruction uses inside its K chunk.  Split mode (FNET_BF16) writes each operand as
// hi + lo, hi = bf16(x) and lo = bf16(x - hi), and accumulates hi.hi + hi.lo + lo.hi: about 16
// mantissa bits, enough for the 1e-5 eV/atom and 1e-4 eV/A gates; single mode (FNET_BF16_FAST)
// takes hi.hi only (8 bits; its errors are reported, not gated).  The input layer (z-scored
// features, which can be ill-conditioned) and the output layer stay FP32.
#include <algorithm>
#include <cstring>

#include "../../include/fnet_hip.h"
#include "fnet_internal.hpp"

namespace fnet {
namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// every layer width is padded to a multiple of 32: output blocks are computed in pairs
__host__ __device__ __forceinline__ int r32(int d) { return (d + 31) & ~31; }

__device__ __forceinline__ float actf(int id, float x) {
  switch (id) {
    case 7: return expf(-(x * x));                                // gaussian
    case 4: return logf(1.0f + expf(x));                          // softplus
    case 5: return (sqrtf(x * x + 1.0f) - 1.0f) / 2.0f + x;       // bent identity
    case 2: return x > 0.0f ? x : 0.0f;                           // relu
    case 3: return (0.01f * x > x) ? 0.01f * x : x;               // leaky relu
    case 1: return 1.0f / (1.0f + expf(-x));                      // sigmoid
    case 8: return x > 0.0f ? 1.0f : 0.0f;                        // heaviside
    case 6: return atanf(x);                                      // atan
    case 9: return x;                                             // linear
    default: return tanhf(x);                                     // tanh
  }
}

__device__ __forceinline__ float actd(int id, float x) {
  switch (id) {
    case 7: return -2.0f * x * expf(-(x * x));
    case 4: return 1.0f / (1.0f + expf(-x));
    case 5: return x / (2.0f * sqrtf(x * x + 1.0f)) + 1.0f;
    case 2: return x >= 0.0f ? 1.0f : 0.0f;
    case 3: return x >= 0.0f ? 1.0f : 0.01f;
    case 1: {
      const float s = 1.0f / (1.0f + expf(-x));
      return s * (1.0f - s);
    }
    case 8: return 0.0f;
    case 6: return 1.0f / (x * x + 1.0f);
    case 9: return 1.0f;
    default: {
      const float t = tanhf(x);
      return 1.0f - t * t;
    }
  }
}

// f(x) and f'(x) together: tanh and sigmoid share their transcendental (the same values as actf
// and actd, bit for bit), the others are evaluated one after the other.
__device__ __forceinline__ void actfd(int id, float x, float& f, float& d) {
  if (id == 0) {
    const float t = tanhf(x);
    f = t;
    d = 1.0f - t * t;
  } else if (id == 1) {
    const float s = 1.0f / (1.0f + expf(-x));
    f = s;
    d = s * (1.0f - s);
  } else {
    f = actf(id, x);
    d = actd(id, x);
  }
}

struct MfmaArgs {
  NetShape sh;
  const float* wpack;
  long long pk_stride;
  long long fw[kMaxLayers], rw[kMaxLayers], pb[kMaxLayers];
  // BF16 fragments of the hidden-to-hidden layers (hi, then lo at + bsz): forward and reverse
  const bf8* bpack;
  long long bk_stride;  // per species, in bf8 units
  long long bfw[kMaxLayers], brw[kMaxLayers], bsz[kMaxLayers];
  int fpoff[kMaxLayers];  // offset (floats) of f'(z) of hidden activation layer m in the wave's store
  const float* params_f;  // [nspecies][nparams] FP32 (last-layer columns of the reverse sweep)
  const double* G;        // [N][nG]
  const double* ext;
  int next;
  const int* ext_in;
  int nG, nout;
  const double* zmean;
  const double* zinv;
  const int* perm;
  const int* tile_sp;
  const int* tile_start;
  const int* tile_cnt;
  int first_tile;
  int ldx;        // row stride (floats) of the wave's activation rows
  int wave_floats;
  double* Eatom;  // [N][nout]
  double* cG;     // [N][nout][nG]
};

// Row stride and per-wave LDS (floats) for a network shape: 16 activation rows of the widest
// layer (+4 floats so that rows start on different banks) and the fragment slots of every hidden
// layer, which hold z, then f'(z), then the reverse-sweep delta of that layer.
void wave_lds(const NetShape& sh, int& ldx, int& wave_floats, int* fpoff) {
  int dmax = 0;
  for (int l = 0; l < sh.nl; ++l) dmax = std::max(dmax, r32(sh.dims[l]));
  ldx = dmax + 4;
  int fp = 0;
  for (int m = 0; m < kMaxLayers; ++m) fpoff[m] = 0;
  for (int m = 1; m < sh.nl - 1; ++m) {
    fpoff[m] = fp;
    fp += r32(sh.dims[m]) * 16;
  }
  wave_floats = 16 * ldx + fp;
}

// Two 16 x 16 output blocks (ob0, ob0 + 1) of Z^T = A . X^T over Kc 16-wide K chunks: A fragments
// from the packed operand Af (block stride Kc * 256 floats), X^T fragments from the wave's rows.
__device__ __forceinline__ void gemm_pair(const float* __restrict__ Af, int ob0, int Kc,
                                          const float* X, int ldx, int lane, int j, int g,
                                          f4& c0, f4& c1) {
  const float* A0 = Af + (long long)ob0 * Kc * 256 + lane;
  const float* A1 = A0 + (long long)Kc * 256;
  float n0[4], n1[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    n0[s] = A0[s * 64];
    n1[s] = A1[s * 64];
  }
#pragma unroll 1
  for (int c = 0; c < Kc; ++c) {
    const f4 xb = *(const f4*)(X + j * ldx + 16 * c + 4 * g);
    float u0[4], u1[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      u0[s] = n0[s];
      u1[s] = n1[s];
    }
    if (c + 1 < Kc) {
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        n0[s] = A0[((c + 1) * 4 + s) * 64];
        n1[s] = A1[((c + 1) * 4 + s) * 64];
      }
    }
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      c0 = __builtin_amdgcn_mfma_f32_16x16x4f32(u0[s], xb[s], c0, 0, 0, 0);
      c1 = __builtin_amdgcn_mfma_f32_16x16x4f32(u1[s], xb[s], c1, 0, 0, 0);
    }
  }
}

// hi and lo BF16 parts of 8 FP32 values
__device__ __forceinline__ void split8(const f4& x0, const f4& x1, bf8& h, bf8& l) {
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const __bf16 a = (__bf16)x0[e], b = (__bf16)x1[e];
    h[e] = a;
    h[e + 4] = b;
    l[e] = (__bf16)(x0[e] - (float)a);
    l[e + 4] = (__bf16)(x1[e] - (float)b);
  }
}

// gemm_pair on the BF16 matrix cores: Kc32 chunks of 32; A fragments Ah (hi) and Ah + bsz (lo),
// block stride Kc32 * 64 bf8; BF = 3 split products, 1 the hi products only.
template <int BF>
__device__ __forceinline__ void gemm_pair_bf(const bf8* __restrict__ Ah, long long bsz, int ob0,
                                             int Kc32, const float* X, int ldx, int lane, int j,
                                             int g, f4& c0, f4& c1) {
  const bf8* A0 = Ah + (long long)ob0 * Kc32 * 64 + lane;
  const bf8* A1 = A0 + (long long)Kc32 * 64;
  bf8 na0 = A0[0], na1 = A1[0], nl0, nl1;
  if (BF == 3) {
    nl0 = A0[bsz];
    nl1 = A1[bsz];
  }
#pragma unroll 1
  for (int c = 0; c < Kc32; ++c) {
    const float* xr = X + j * ldx + 32 * c + 8 * g;
    const f4 x0 = *(const f4*)xr, x1 = *(const f4*)(xr + 4);
    bf8 xh, xl;
    split8(x0, x1, xh, xl);
    const bf8 a0 = na0, a1 = na1;
    bf8 l0, l1;
    if (BF == 3) {
      l0 = nl0;
      l1 = nl1;
    }
    if (c + 1 < Kc32) {
      na0 = A0[(c + 1) * 64];
      na1 = A1[(c + 1) * 64];
      if (BF == 3) {
        nl0 = A0[bsz + (c + 1) * 64];
        nl1 = A1[bsz + (c + 1) * 64];
      }
    }
    c0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, xh, c0, 0, 0, 0);
    c1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, xh, c1, 0, 0, 0);
    if (BF == 3) {
      c0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, xl, c0, 0, 0, 0);
      c1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, xl, c1, 0, 0, 0);
      c0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(l0, xh, c0, 0, 0, 0);
      c1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(l1, xh, c1, 0, 0, 0);
    }
  }
}

// layer l (ww(d_l, d_{l+1})) maps a hidden layer to a hidden layer
__device__ __forceinline__ bool hidden_hidden(int l, int nl) { return l >= 1 && l + 1 <= nl - 2; }

template <bool GRADIN, int BF>
__global__ void __launch_bounds__(256) k_mlp_mfma(MfmaArgs a) {
  extern __shared__ __align__(16) float smf[];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int g = lane >> 4, j = lane & 15;
  const int tile = a.first_tile + blockIdx.x;
  const int sp = a.tile_sp[tile], start = a.tile_start[tile], cnt = a.tile_cnt[tile];
  const int a0 = 16 * w;
  if (a0 >= cnt) return;  // whole wave
  const int nv = min(16, cnt - a0);
  const int nl = a.sh.nl;
  const int act = a.sh.act[sp];
  const int ldx = a.ldx;
  float* X = smf + (size_t)w * a.wave_floats;
  float* FP = X + 16 * ldx;
  const float* WP = a.wpack + (long long)sp * a.pk_stride;
  const bool valid = j < nv;
  const long long atom = valid ? (long long)a.perm[start + a0 + j] : 0;

  // input rows: standardised features (and external features), zero beyond d0 and nv
  const int d0 = a.sh.dims[0], K0 = r32(d0);
  for (int idx = lane; idx < 16 * K0; idx += 64) {
    const int r = idx / K0, k = idx - r * K0;
    float v = 0.0f;
    if (r < nv && k < d0) {
      const long long at = a.perm[start + a0 + r];
      v = k < a.nG ? (float)((a.G[at * a.nG + k] - a.zmean[k]) * a.zinv[k])
                   : (float)a.ext[at * a.next + a.ext_in[k - a.nG]];
    }
    X[r * ldx + k] = v;
  }
  wave_sync();

  for (int l = 0; l < nl - 1; ++l) {
    const int Kc = r32(a.sh.dims[l]) >> 4, Nb = r32(a.sh.dims[l + 1]) >> 4;
    const float* Wf = WP + a.fw[l];
    const float* Bp = WP + a.pb[l];
    const bool hidden = l < nl - 2;
    float* zs = FP + a.fpoff[l + 1];
    const bool bf = BF && hidden_hidden(l, nl);
    for (int ob0 = 0; ob0 < Nb; ob0 += 2) {
      f4 c0 = *(const f4*)(Bp + 16 * ob0 + 4 * g);
      f4 c1 = *(const f4*)(Bp + 16 * ob0 + 16 + 4 * g);
      if (bf)
        gemm_pair_bf<BF>(a.bpack + (long long)sp * a.bk_stride + a.bfw[l], a.bsz[l], ob0, Kc >> 1,
                         X, ldx, lane, j, g, c0, c1);
      else
        gemm_pair(Wf, ob0, Kc, X, ldx, lane, j, g, c0, c1);
      if (hidden) {
        *(f4*)(zs + (ob0 * 64 + lane) * 4) = c0;
        *(f4*)(zs + ((ob0 + 1) * 64 + lane) * 4) = c1;
      } else if (valid) {
        // linear output layer (nout <= 16): lane group g holds outputs 4 g + r of atom j
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int o = 4 * g + r;
          if (o < a.nout) a.Eatom[atom * a.nout + o] = (double)c0[r];
        }
      }
    }
    wave_sync();  // every lane has read this layer's rows before they are overwritten
    if (hidden) {
      // activation: z in the fragment slots -> next rows y = f(z), slots <- f'(z); four at a
      // time, so that their LDS reads and transcendentals overlap (the kernel's occupancy is
      // set by its LDS, not by these registers)
#pragma unroll 4
      for (int q = 0; q < 4 * Nb; ++q) {
        const int ob = q >> 2, r = q & 3;
        float* slot = zs + (ob * 64 + lane) * 4 + r;
        const float z = *slot;
        float f, d;
        actfd(act, z, f, d);
        X[j * ldx + 16 * ob + 4 * g + r] = f;
        *slot = d;
      }
      wave_sync();
    }
  }
  if (!GRADIN) return;

  // reverse sweep of the single output (input gradient = row of the Jacobian)
  double* out = a.cG + atom * a.nG;
  int m = nl - 2;
  {
    // delta of activation layer m from the one-hot output: column 0 of ww(m), times f'(z_m)
    const int dm = a.sh.dims[m], Kb = r32(dm) >> 4;
    const float* col = a.params_f + (long long)sp * a.sh.nparams + a.sh.woff[m];
    if (m >= 1) {
      float* sl = FP + a.fpoff[m];
#pragma unroll 4
      for (int q = 0; q < 4 * Kb; ++q) {
        const int ib = q >> 2, r = q & 3, k = 16 * ib + 4 * g + r;
        float* slot = sl + (ib * 64 + lane) * 4 + r;
        const float d = (k < dm ? col[k] : 0.0f) * *slot;
        *slot = d;
        X[j * ldx + k] = d;
      }
      wave_sync();
    } else if (valid) {
      for (int k = 4 * g; k < a.nG; k += 16)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (k + r < a.nG) out[k + r] = (double)col[k + r] * a.zinv[k + r];
    }
  }
  // delta_{m-1}^T = W_{m-1} . delta_m^T (rows in X), times f'(z_{m-1}) for hidden layers
  for (; m >= 1; --m) {
    const int Mb = r32(a.sh.dims[m - 1]) >> 4, Kc = r32(a.sh.dims[m]) >> 4;
    const float* Rf = WP + a.rw[m - 1];
    float* sl = FP + a.fpoff[m - 1];
    const bool bf = BF && hidden_hidden(m - 1, nl);
    for (int ib0 = 0; ib0 < Mb; ib0 += 2) {
      f4 c0 = {0.0f, 0.0f, 0.0f, 0.0f}, c1 = {0.0f, 0.0f, 0.0f, 0.0f};
      if (bf)
        gemm_pair_bf<BF>(a.bpack + (long long)sp * a.bk_stride + a.brw[m - 1], a.bsz[m - 1], ib0,
                         Kc >> 1, X, ldx, lane, j, g, c0, c1);
      else
        gemm_pair(Rf, ib0, Kc, X, ldx, lane, j, g, c0, c1);
      if (m - 1 >= 1) {
        f4* s0 = (f4*)(sl + (ib0 * 64 + lane) * 4);
        f4* s1 = (f4*)(sl + ((ib0 + 1) * 64 + lane) * 4);
        *s0 = c0 * *s0;
        *s1 = c1 * *s1;
      } else if (valid) {
        // input layer: the ACSF part of d out / d x feeds the force adjoint (FP64, z-score factor)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int f0 = 16 * ib0 + 4 * g + r, f1 = f0 + 16;
          if (f0 < a.nG) out[f0] = (double)c0[r] * a.zinv[f0];
          if (f1 < a.nG) out[f1] = (double)c1[r] * a.zinv[f1];
        }
      }
    }
    wave_sync();
    if (m - 1 >= 1) {
      // delta_{m-1} fragments -> rows for the next product
#pragma unroll 4
      for (int q = 0; q < 4 * Mb; ++q) {
        const int ib = q >> 2, r = q & 3;
        X[j * ldx + 16 * ib + 4 * g + r] = sl[(ib * 64 + lane) * 4 + r];
      }
      wave_sync();
    }
  }
}

// ---- weight gradient (bprop, network.F90:248-296, with sysTrain's scaling, bpnn.F90:692-702) ----
// One block = nw waves = 16 nw atoms of one element (a tile of the MFMA layout, or a fixed part of
// one).  Forward: every wave runs its 16 atoms through the hidden layers as above, keeping each
// layer's pre-activation rows z_l for all atoms of the block in LDS (R[l]; R[0] holds the input
// rows) and its own current activations in Xw.  Backward, top layer first, with delta_out = the
// loss gradient of each atom (nout = 1):
//  * the output layer (rank one) on the vector unit: dW[0][k] = sum_at delta_out a_l[at][k],
//    delta_l = W[0][k] delta_out f'(z_l);
//  * every other layer l: dW_l[o][k] = sum_at delta_{l+1}[at][o] a_l[at][k] on the matrix cores
//    with the block's atoms as the K dimension (A = delta rows, B = a_l = f(z_l) rows; a wave
//    takes a k-block with up to 8 o-blocks, so each element is activated once per 8 o-blocks), the bias
//    gradient as a column sum in atom order, then delta_l = (W_l^T delta_{l+1}) f'(z_l) with the
//    reverse fragments, written over z_l.
// Products and sums over the block's atoms are FP32 (column sums over the atoms in four fixed
// quarters, then the quarters in order); the block writes its FP32 partial vector
// (gpart + block * nparams), which k_grad_chunks / k_grad_reduce add in FP64 in a fixed order.
struct GradArgs {
  const double* lossgrad;  // [N][nout]
  float* gpart;            // [blocks][nparams]
  int nsub;                // blocks per tile
  int roff[kMaxLayers];    // LDS offsets (floats) of R[l], l = 0 .. nl - 2
  int ld[kMaxLayers];      // their row strides
  int doff;                // delta_out [16 nw]
  int woff;                // Xw of wave 0 (16 rows of stride ldw per wave); after the forward,
                           // scratch of the column sums (4 r32(widest layer) floats)
  int ldw;
};

// out[k] (k < n) = sum over the block's TA atoms of v(at, k) = wt[at] M[at][k] (ACT: f(M)), in
// four quarters of the atoms, each in order, then ((q0 + q1) + (q2 + q3)).  Ends synchronised.
template <bool WT, bool ACT>
__device__ __forceinline__ void col_sums(const float* M, int ld, int n, int TA, const float* wt,
                                         int act, float* red, float* out_f) {
  const int qa = TA >> 2;
  for (int idx = threadIdx.x; idx < 4 * n; idx += blockDim.x) {
    const int q = idx / n, k = idx - q * n;
    float s = 0.0f;
    for (int at = q * qa; at < (q + 1) * qa; ++at) {
      float v = M[at * ld + k];
      if (ACT) v = actf(act, v);
      s += WT ? wt[at] * v : v;
    }
    red[idx] = s;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < n; k += blockDim.x)
    out_f[k] = (red[k] + red[n + k]) + (red[2 * n + k] + red[3 * n + k]);
  __syncthreads();
}

template <int BF>
__global__ void __launch_bounds__(256) k_mlp_grad_mfma(MfmaArgs a, GradArgs ga) {
  extern __shared__ __align__(16) float smf[];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int g = lane >> 4, j = lane & 15;
  const int TA = 16 * nw;
  const int t = a.first_tile + (int)blockIdx.x / ga.nsub, sub = (int)blockIdx.x % ga.nsub;
  const int sp = a.tile_sp[t];
  const int start = a.tile_start[t] + sub * TA;
  const int cnt = min(TA, a.tile_cnt[t] - sub * TA);  // may be <= 0: the partial is then zero
  const int nl = a.sh.nl;
  const int act = a.sh.act[sp];
  const float* WP = a.wpack + (long long)sp * a.pk_stride;
  float* gp = ga.gpart + (long long)blockIdx.x * a.sh.nparams;
  float* dout = smf + ga.doff;
  float* Xw = smf + ga.woff + w * 16 * ga.ldw;
  float* red = smf + ga.woff;
  const int a0 = 16 * w;

  // entries without gradient: the dummy output weights and the input layer's biases
  for (int k = threadIdx.x; k < a.sh.dims[nl - 1]; k += blockDim.x) gp[a.sh.woff[nl - 1] + k] = 0.0f;
  for (int k = threadIdx.x; k < a.sh.dims[0]; k += blockDim.x) gp[a.sh.boff[0] + k] = 0.0f;

  // input rows (own atoms) and delta_out
  {
    float* X0 = smf + ga.roff[0] + a0 * ga.ld[0];
    const int d0 = a.sh.dims[0], K0 = r32(d0);
    for (int idx = lane; idx < 16 * K0; idx += 64) {
      const int r = idx / K0, k = idx - r * K0;
      float v = 0.0f;
      if (a0 + r < cnt && k < d0) {
        const long long at = a.perm[start + a0 + r];
        v = k < a.nG ? (float)((a.G[at * a.nG + k] - a.zmean[k]) * a.zinv[k])
                     : (float)a.ext[at * a.next + a.ext_in[k - a.nG]];
      }
      X0[r * ga.ld[0] + k] = v;
    }
    if (lane < 16)
      dout[a0 + lane] =
          a0 + lane < cnt ? (float)ga.lossgrad[(long long)a.perm[start + a0 + lane] * a.nout] : 0.0f;
  }
  wave_sync();

  // forward through the hidden layers: z_{l+1} rows into R[l+1], f(z_{l+1}) into Xw
  for (int l = 0; l + 2 < nl; ++l) {
    const int Kc = r32(a.sh.dims[l]) >> 4, Nb = r32(a.sh.dims[l + 1]) >> 4;
    const float* Xin = l == 0 ? smf + ga.roff[0] + a0 * ga.ld[0] : Xw;
    const int ldin = l == 0 ? ga.ld[0] : ga.ldw;
    const int ldz = ga.ld[l + 1];
    float* Z = smf + ga.roff[l + 1] + a0 * ldz;
    const float* Bp = WP + a.pb[l];
    const bool bf = BF && hidden_hidden(l, nl);
    for (int ob0 = 0; ob0 < Nb; ob0 += 2) {
      f4 c0 = *(const f4*)(Bp + 16 * ob0 + 4 * g);
      f4 c1 = *(const f4*)(Bp + 16 * ob0 + 16 + 4 * g);
      if (bf)
        gemm_pair_bf<BF>(a.bpack + (long long)sp * a.bk_stride + a.bfw[l], a.bsz[l], ob0, Kc >> 1,
                         Xin, ldin, lane, j, g, c0, c1);
      else
        gemm_pair(WP + a.fw[l], ob0, Kc, Xin, ldin, lane, j, g, c0, c1);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        Z[j * ldz + 16 * ob0 + 4 * g + r] = c0[r];
        Z[j * ldz + 16 * ob0 + 16 + 4 * g + r] = c1[r];
      }
    }
    wave_sync();
    const int W = 16 * Nb;
    for (int idx = lane; idx < 16 * W; idx += 64) {
      const int r = idx / W, k = idx - r * W;
      Xw[r * ga.ldw + k] = actf(act, Z[r * ldz + k]);
    }
    wave_sync();
  }
  __syncthreads();  // the backward reads every wave's rows

  // output layer (one output): dW[0][k], the bias, and delta of the last hidden layer
  {
    const int l = nl - 2, dl = a.sh.dims[l], ldl = ga.ld[l];
    float* Z = smf + ga.roff[l];
    if (l == 0) col_sums<true, false>(Z, ldl, dl, TA, dout, act, red, gp + a.sh.woff[l]);
    else col_sums<true, true>(Z, ldl, dl, TA, dout, act, red, gp + a.sh.woff[l]);
    col_sums<false, false>(dout, 1, 1, TA, nullptr, act, red, gp + a.sh.boff[nl - 1]);
    if (l >= 1) {
      const float* col = a.params_f + (long long)sp * a.sh.nparams + a.sh.woff[l];
      for (int idx = threadIdx.x; idx < TA * dl; idx += blockDim.x) {
        const int at = idx / dl, k = idx - at * dl;
        float* zp = Z + at * ldl + k;
        *zp = col[k] * dout[at] * actd(act, *zp);
      }
      __syncthreads();
    }
  }

  // the other layers, top down: R[l + 1] holds delta_{l+1}, R[l] holds z_l (x for l = 0)
  for (int l = nl - 3; l >= 0; --l) {
    const int dl = a.sh.dims[l], dn = a.sh.dims[l + 1];
    const int Kb = r32(dl) >> 4, Ob = r32(dn) >> 4;
    const int ldl = ga.ld[l], ldn = ga.ld[l + 1];
    float* Zl = smf + ga.roff[l];
    const float* Dn = smf + ga.roff[l + 1];
    // work items (k-block, group of up to 8 o-blocks); groups narrower than 8 when there are
    // fewer k-blocks than waves (the input layer), so that every wave has one
    const int gob = Kb >= nw ? 8 : max(1, min(8, (Ob * Kb) / nw));
    const int ngr = (Ob + gob - 1) / gob;
    for (int it = w; it < Kb * ngr; it += nw) {
      const int kb = it / ngr;
      {
        const int ob0 = (it - kb * ngr) * gob, ob1 = min(Ob, ob0 + gob);
        f4 c[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) c[q] = f4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll 1
        for (int s4 = 0; s4 < TA; s4 += 4) {
          const int at = s4 + g;
          const float z = Zl[at * ldl + 16 * kb + j];
          const float bv = l == 0 ? z : actf(act, z);
#pragma unroll
          for (int q = 0; q < 8; ++q)
            if (ob0 + q < ob1)
              c[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(Dn[at * ldn + 16 * (ob0 + q) + j], bv,
                                                          c[q], 0, 0, 0);
        }
        const int k = 16 * kb + j;
#pragma unroll
        for (int q = 0; q < 8; ++q)
          if (ob0 + q < ob1)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int o = 16 * (ob0 + q) + 4 * g + r;
              if (o < dn && k < dl) gp[a.sh.woff[l] + (long long)o * dl + k] = c[q][r];
            }
      }
    }
    col_sums<false, false>(Dn, ldn, dn, TA, nullptr, act, red, gp + a.sh.boff[l + 1]);
    if (l == 0) break;
    // delta_l = (W_l^T delta_{l+1}) f'(z_l), own atoms, over z_l
    const int Mb = Kb, Kc = Ob;
    const bool bf = BF && hidden_hidden(l, nl);
    float* zr = Zl + (a0 + j) * ldl;
    for (int ib0 = 0; ib0 < Mb; ib0 += 2) {
      f4 c0 = {0.0f, 0.0f, 0.0f, 0.0f}, c1 = {0.0f, 0.0f, 0.0f, 0.0f};
      if (bf)
        gemm_pair_bf<BF>(a.bpack + (long long)sp * a.bk_stride + a.brw[l], a.bsz[l], ib0, Kc >> 1,
                         Dn + a0 * ldn, ldn, lane, j, g, c0, c1);
      else
        gemm_pair(WP + a.rw[l], ib0, Kc, Dn + a0 * ldn, ldn, lane, j, g, c0, c1);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int k0 = 16 * ib0 + 4 * g + r;
        zr[k0] = c0[r] * actd(act, zr[k0]);
        zr[k0 + 16] = c1[r] * actd(act, zr[k0 + 16]);
      }
    }
    __syncthreads();
  }
}

}  // namespace

int mlp_mfma_waves(const NetShape& sh) {
  // single-output networks (the energy), layer widths up to kMaxDim
  if (sh.nl < 2 || sh.dims[sh.nl - 1] != 1) return 0;
  for (int l = 0; l < sh.nl; ++l)
    if (sh.dims[l] < 1 || sh.dims[l] > kMaxDim) return 0;
  int ldx, wf, fpoff[kMaxLayers];
  wave_lds(sh, ldx, wf, fpoff);
  int waves = 4;
  while (waves > 1 && (size_t)waves * wf * sizeof(float) > 64 * 1024) waves >>= 1;
  if ((size_t)waves * wf * sizeof(float) > 64 * 1024) return 0;
  return waves;
}

// Packs the FP32 weights of every species into MFMA A-operand fragments (see the header comment),
// every layer width padded to a multiple of 32.  Forward fragment of layer l: index
// ((ob * Kc + c) * 4 + s) * 64 + lane, lane = 16 g + i, holds ww_l(k = 16 c + 4 g + s,
// o = 16 ob + i); reverse fragment: ((ib * Oc + c) * 4 + s) * 64 + lane holds
// ww_l(k = 16 ib + i, o = 16 c + 4 g + s); zero outside the layer's dimensions.
void pack_mfma_weights(const NetShape& sh, int nspecies, const double* params,
                       std::vector<float>& out, long long& stride, long long* fw, long long* rw,
                       long long* pb) {
  long long o = 0;
  for (int l = 0; l + 1 < sh.nl; ++l) {
    const long long sz = (long long)r32(sh.dims[l]) * r32(sh.dims[l + 1]);
    fw[l] = o;
    o += sz;
    rw[l] = o;
    o += sz;
    pb[l] = o;
    o += r32(sh.dims[l + 1]);
  }
  stride = (o + 63) & ~63LL;
  out.assign((size_t)stride * nspecies, 0.0f);
  for (int sp = 0; sp < nspecies; ++sp) {
    const double* P = params + (long long)sp * sh.nparams;
    float* D = out.data() + (long long)sp * stride;
    for (int l = 0; l + 1 < sh.nl; ++l) {
      const int din = sh.dims[l], dout = sh.dims[l + 1];
      const int Kc = r32(din) / 16, Oc = r32(dout) / 16;
      auto ww = [&](int k, int oo) -> float {
        return (k < din && oo < dout) ? (float)P[sh.woff[l] + (long long)oo * din + k] : 0.0f;
      };
      for (int ob = 0; ob < Oc; ++ob)
        for (int c = 0; c < Kc; ++c)
          for (int s = 0; s < 4; ++s)
            for (int lane = 0; lane < 64; ++lane) {
              const int g = lane / 16, i = lane % 16;
              D[fw[l] + ((long long)(ob * Kc + c) * 4 + s) * 64 + lane] =
                  ww(16 * c + 4 * g + s, 16 * ob + i);
            }
      for (int ib = 0; ib < Kc; ++ib)
        for (int c = 0; c < Oc; ++c)
          for (int s = 0; s < 4; ++s)
            for (int lane = 0; lane < 64; ++lane) {
              const int g = lane / 16, i = lane % 16;
              D[rw[l] + ((long long)(ib * Oc + c) * 4 + s) * 64 + lane] =
                  ww(16 * ib + i, 16 * c + 4 * g + s);
            }
      for (int q = 0; q < dout; ++q) D[pb[l] + q] = (float)P[sh.boff[l + 1] + q];
    }
  }
}

namespace {
// float -> bf16 bits, round to nearest even (finite values)
uint16_t to_bf16(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
float from_bf16(uint16_t h) {
  const uint32_t u = (uint32_t)h << 16;
  float x;
  std::memcpy(&x, &u, 4);
  return x;
}
}  // namespace

// BF16 fragments (hi, then lo = bf16(w - hi)) of the hidden-to-hidden layers for
// v_mfma_f32_16x16x32_bf16.  Forward fragment of layer l: element
// ((ob * Kc32 + c) * 64 + lane) * 8 + e, lane = 16 g + i, holds ww_l(k = 32 c + 8 g + e,
// o = 16 ob + i); reverse: ((ib * Oc32 + c) * 64 + lane) * 8 + e holds
// ww_l(k = 16 ib + i, o = 32 c + 8 g + e).  Offsets and sizes are in units of 8 values.
void pack_bf16_weights(const NetShape& sh, int nspecies, const double* params,
                       std::vector<uint16_t>& out, long long& stride, long long* bfw,
                       long long* brw, long long* bsz) {
  long long o = 0;
  for (int l = 0; l < kMaxLayers; ++l) bfw[l] = brw[l] = bsz[l] = 0;
  for (int l = 1; l + 2 < sh.nl; ++l) {
    const long long sz = (long long)r32(sh.dims[l]) * r32(sh.dims[l + 1]) / 8;
    bsz[l] = sz;
    bfw[l] = o;
    o += 2 * sz;
    brw[l] = o;
    o += 2 * sz;
  }
  stride = o;
  out.assign((size_t)(stride * 8 * nspecies), 0);
  for (int sp = 0; sp < nspecies; ++sp) {
    const double* P = params + (long long)sp * sh.nparams;
    uint16_t* D = out.data() + (long long)sp * stride * 8;
    for (int l = 1; l + 2 < sh.nl; ++l) {
      const int din = sh.dims[l], dout = sh.dims[l + 1];
      const int Kc32 = r32(din) / 32, Oc32 = r32(dout) / 32;
      const int Kb = r32(din) / 16, Ob = r32(dout) / 16;
      const long long sz8 = bsz[l] * 8;
      auto put = [&](long long base, long long idx, float w) {
        const uint16_t h = to_bf16(w);
        D[base + idx] = h;
        D[base + sz8 + idx] = to_bf16(w - from_bf16(h));
      };
      auto ww = [&](int k, int oo) -> float {
        return (k < din && oo < dout) ? (float)P[sh.woff[l] + (long long)oo * din + k] : 0.0f;
      };
      for (int ob = 0; ob < Ob; ++ob)
        for (int c = 0; c < Kc32; ++c)
          for (int lane = 0; lane < 64; ++lane)
            for (int e = 0; e < 8; ++e) {
              const int g = lane / 16, i = lane % 16;
              put(bfw[l] * 8, (((long long)ob * Kc32 + c) * 64 + lane) * 8 + e,
                  ww(32 * c + 8 * g + e, 16 * ob + i));
            }
      for (int ib = 0; ib < Kb; ++ib)
        for (int c = 0; c < Oc32; ++c)
          for (int lane = 0; lane < 64; ++lane)
            for (int e = 0; e < 8; ++e) {
              const int g = lane / 16, i = lane % 16;
              put(brw[l] * 8, (((long long)ib * Oc32 + c) * 64 + lane) * 8 + e,
                  ww(16 * ib + i, 32 * c + 8 * g + e));
            }
    }
  }
}

namespace {
// Waves per weight-gradient block (the largest divisor of the tile's waves whose LDS fits 64 KiB),
// with the LDS layout of k_mlp_grad_mfma; 0 when the network has no MFMA tiles.
int mlp_grad_waves(const NetShape& sh, int tile_waves, GradArgs* ga, size_t* bytes) {
  if (tile_waves <= 0) return 0;
  for (int nw = tile_waves; nw >= 1; nw >>= 1) {
    if (tile_waves % nw) continue;
    const int TA = 16 * nw;
    GradArgs q{};
    int f = 0;
    for (int l = 0; l + 1 < sh.nl; ++l) {
      q.ld[l] = r32(sh.dims[l]) + 4;
      q.roff[l] = f;
      f += TA * q.ld[l];
    }
    q.doff = f;
    f += TA;
    q.ldw = 0;
    int wmax = 0;
    for (int l = 0; l < sh.nl; ++l) wmax = std::max(wmax, r32(sh.dims[l]));
    for (int l = 1; l + 1 < sh.nl; ++l) q.ldw = std::max(q.ldw, r32(sh.dims[l]) + 4);
    q.woff = f;
    f += std::max(nw * 16 * q.ldw, 4 * wmax);
    const size_t b = (size_t)f * sizeof(float);
    if (b <= 64 * 1024) {
      q.nsub = tile_waves / nw;
      if (ga) *ga = q;
      if (bytes) *bytes = b;
      return nw;
    }
  }
  return 0;
}
}  // namespace

// Weight-gradient blocks per MFMA tile (0: the network has no MFMA path)
int mlp_grad_subtiles(const NetShape& sh, int tile_waves) {
  GradArgs ga;
  return mlp_grad_waves(sh, tile_waves, &ga, nullptr) ? ga.nsub : 0;
}

hipError_t launch_mlp_grad_mfma(fnet_batch* b, const DevModel& m, hipStream_t s) {
  GradArgs ga;
  size_t lds = 0;
  const int nw = mlp_grad_waves(m.shape, b->mtile_waves, &ga, &lds);
  if (nw == 0) return hipErrorInvalidConfiguration;
  if (b->nmtiles <= 0) return hipSuccess;
  MfmaArgs a{};
  a.sh = m.shape;
  a.wpack = m.wpack;
  a.pk_stride = m.pk_stride;
  for (int l = 0; l < kMaxLayers; ++l) {
    a.fw[l] = m.pk_fw[l];
    a.rw[l] = m.pk_rw[l];
    a.pb[l] = m.pk_b[l];
    a.bfw[l] = m.bk_fw[l];
    a.brw[l] = m.bk_rw[l];
    a.bsz[l] = m.bk_sz[l];
  }
  a.params_f = m.params_f;
  a.G = (const double*)b->G;
  a.ext = b->ext;
  a.next = b->next;
  a.ext_in = m.ext_in;
  a.nG = m.nG;
  a.nout = m.nout;
  a.zmean = m.zmean;
  a.zinv = m.zinv;
  a.perm = b->perm;
  a.tile_sp = b->mtile_sp;
  a.tile_start = b->mtile_start;
  a.tile_cnt = b->mtile_cnt;
  a.first_tile = 0;
  a.bpack = (const bf8*)m.bpack;
  a.bk_stride = m.bk_stride;
  ga.lossgrad = b->lossgrad;
  ga.gpart = (float*)b->gpart;
  const int bf = (b->precision == FNET_BF16 && m.bpack) ? 3
               : (b->precision == FNET_BF16_FAST && m.bpack) ? 1 : 0;
  const dim3 grid((unsigned)((long long)b->nmtiles * ga.nsub)), block(64 * nw);
  if (bf == 3) k_mlp_grad_mfma<3><<<grid, block, lds, s>>>(a, ga);
  else if (bf == 1) k_mlp_grad_mfma<1><<<grid, block, lds, s>>>(a, ga);
  else k_mlp_grad_mfma<0><<<grid, block, lds, s>>>(a, ga);
  return hipGetLastError();
}

hipError_t launch_mlp_mfma(fnet_batch* b, const DevModel& m, int first_tile, int ntiles,
                           int want_grad_in, hipStream_t s) {
  if (ntiles <= 0) return hipSuccess;
  MfmaArgs a;
  a.sh = m.shape;
  a.wpack = m.wpack;
  a.pk_stride = m.pk_stride;
  for (int l = 0; l < kMaxLayers; ++l) {
    a.fw[l] = m.pk_fw[l];
    a.rw[l] = m.pk_rw[l];
    a.pb[l] = m.pk_b[l];
  }
  wave_lds(m.shape, a.ldx, a.wave_floats, a.fpoff);
  a.params_f = m.params_f;
  a.G = (const double*)b->G;
  a.ext = b->ext;
  a.next = b->next;
  a.ext_in = m.ext_in;
  a.nG = m.nG;
  a.nout = m.nout;
  a.zmean = m.zmean;
  a.zinv = m.zinv;
  a.perm = b->perm;
  a.tile_sp = b->mtile_sp;
  a.tile_start = b->mtile_start;
  a.tile_cnt = b->mtile_cnt;
  a.first_tile = first_tile;
  a.Eatom = b->Eatom;
  a.cG = (double*)b->cG;
  a.bpack = (const bf8*)m.bpack;
  a.bk_stride = m.bk_stride;
  for (int l = 0; l < kMaxLayers; ++l) {
    a.bfw[l] = m.bk_fw[l];
    a.brw[l] = m.bk_rw[l];
    a.bsz[l] = m.bk_sz[l];
  }
  const int bf = (b->precision == FNET_BF16 && m.bpack) ? 3
               : (b->precision == FNET_BF16_FAST && m.bpack) ? 1 : 0;
  const int waves = b->mtile_waves;
  const size_t lds = (size_t)waves * a.wave_floats * sizeof(float);
  const dim3 grid(ntiles), block(64 * waves);
  if (want_grad_in) {
    if (bf == 3) k_mlp_mfma<true, 3><<<grid, block, lds, s>>>(a);
    else if (bf == 1) k_mlp_mfma<true, 1><<<grid, block, lds, s>>>(a);
    else k_mlp_mfma<true, 0><<<grid, block, lds, s>>>(a);
  } else {
    if (bf == 3) k_mlp_mfma<false, 3><<<grid, block, lds, s>>>(a);
    else if (bf == 1) k_mlp_mfma<false, 1><<<grid, block, lds, s>>>(a);
    else k_mlp_mfma<false, 0><<<grid, block, lds, s>>>(a);
  }
  return hipGetLastError();
}

}  // namespace fnet
