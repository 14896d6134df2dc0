// ACSF values and their adjoint (force) pass on gfx950.
//
// One wavefront per centre atom.  The centre's CSR row of the neighbour list (image displacement
// and distance per entry) is staged once into the wave's LDS slice with coalesced loads, with
// the cutoff function of the list cutoff and per-element index lists built by ballots, so every
// species-resolved list of buildGFunctionNeighborlists (acsf.F90:943-1066) is a slice of LDS
// instead of a rebuilt neighbour list.
//
//  * radial G1/G2/G3 (acsf.F90:1287-1373): lanes = (entry stride group, function), partial sums
//    combined through LDS;
//  * angular G4/G5 (acsf.F90:1377-1492): functions sharing (type, cutoff, eta, species pair,
//    atom id) form a group; lanes stride over the group's neighbour pairs once and update every
//    member.  Pairs are enumerated by offset (k = j + d): the diagonal, then each unordered pair
//    once with weight 2 when both lists are the same (the ordered double sum with its j == k
//    terms, acsf.F90:1420-1432), or the full list1 x list2 product otherwise.  With this order
//    the 64 lanes of one step touch 64 different entries, which keeps the adjoint's LDS
//    accumulation free of same-address collisions;
//  * the adjoint pass computes F = -dE/dR directly from c = dE/dG (the network's input gradient
//    times the z-score factor): the derivative of every function with respect to each entry's
//    position (gFuncGrad, acsf.F90:1496-1670) is weighted by c and summed per entry, then scattered
//    to the neighbour's force (-) and the centre's force (+), as forces.F90:400-413 contracts the
//    dense array, but without forming it and with periodic images summed exactly.
//
// Arithmetic is T (double or float); sums over neighbours and pairs, the values G and the forces
// are accumulated in FP64 in both modes.
#include "fnet_internal.hpp"

namespace fnet {
namespace {

template <typename T> struct Vec4;
template <> struct Vec4<float> { using type = float4; };
template <> struct Vec4<double> { using type = double4; };

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ unsigned long long lanemask_lt(int lane) {
  return (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
}

__device__ __forceinline__ float fexp(float x) { return __expf(x); }
__device__ __forceinline__ double fexp(double x) { return exp(x); }
__device__ __forceinline__ float fcos(float x) { return __cosf(x); }
__device__ __forceinline__ double fcos(double x) { return cos(x); }
__device__ __forceinline__ float fsin(float x) { return __sinf(x); }
__device__ __forceinline__ double fsin(double x) { return sin(x); }
__device__ __forceinline__ float fpow(float x, float y) { return powf(x, y); }
__device__ __forceinline__ double fpow(double x, double y) { return pow(x, y); }

// x^n for a wave-uniform integer n >= 0, else x^y
template <typename T>
__device__ __forceinline__ T xpow(T x, int n, T y) {
  if (n < 0) return fpow(x, y);
  T r = (T)1;
  while (n) {
    if (n & 1) r *= x;
    x *= x;
    n >>= 1;
  }
  return r;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

struct AcsfArgs {
  const int* rowptr;
  const int* nj;
  const void* nd;
  const int* zcode;
  const double* ext;
  int next;
  const RadialFn* radial;
  int nR;
  const AngGroup* groups;
  int nAG;
  int nZ;
  int nG;
  int nout;
  double rclist;
  long long a0, a1;
  int maxnb;
  double* G;      // [N][nG] (forward)
  const double* cG; // [N][nout][nG] (adjoint)
  double* F;      // [N][nout][3] (adjoint)
};

// Bytes of one wave's LDS slice (host and device agree on this layout).  Per entry: the
// displacement and distance d, the pair-loop factors pa = (1/r, exp(-eta r^2) fc, atom-id
// prefactor, radial derivative), the list-cutoff fc and fc', the neighbour index and element
// code, the adjoint accumulator (adjoint only) and the per-element index lists.
template <typename T, bool ADJ>
__host__ __device__ size_t slice_bytes(int M, int nZ) {
  size_t b = 2 * sizeof(typename Vec4<T>::type) * M;  // d, pa
  b += 2 * sizeof(T) * M;                             // fc, dfc
  if (ADJ) b += 3 * sizeof(T) * M;                    // gs
  b += 2 * sizeof(int) * M;                           // j, zc
  b += 64 * sizeof(double);                           // radial partial sums
  b += sizeof(int) * kMaxZ;                           // nz
  b += sizeof(unsigned short) * (size_t)nZ * M;       // idx
  return (b + 31) & ~size_t(31);
}

template <typename T>
struct Slice {
  using V = typename Vec4<T>::type;
  V* d;
  V* pa;
  T *fc, *dfc;
  T* gs;
  int *j, *zc;
  double* red;
  int* nz;
  unsigned short* idx;
};

template <typename T, bool ADJ>
__device__ Slice<T> make_slice(unsigned char* p, int M) {
  using V = typename Vec4<T>::type;
  Slice<T> s;
  s.d = (V*)p;
  p += sizeof(V) * M;
  s.pa = (V*)p;
  p += sizeof(V) * M;
  s.red = (double*)p;
  p += 64 * sizeof(double);
  s.fc = (T*)p;
  p += sizeof(T) * M;
  s.dfc = (T*)p;
  p += sizeof(T) * M;
  s.gs = nullptr;
  if (ADJ) {
    s.gs = (T*)p;
    p += 3 * sizeof(T) * M;
  }
  s.j = (int*)p;
  p += sizeof(int) * M;
  s.zc = (int*)p;
  p += sizeof(int) * M;
  s.nz = (int*)p;
  p += sizeof(int) * kMaxZ;
  s.idx = (unsigned short*)p;
  return s;
}

__device__ __forceinline__ double ext_q(const double* ext, int next, int atomid, long long atom) {
  return atomid > 0 ? ext[atom * next + (atomid - 1)] : 1.0;
}

// Pair enumeration of one group.  Same lists: the diagonal (d = 0, weight 1), then every
// unordered pair once as (j, j + d mod L), d = 1 .. L/2, weight 2.  Different lists: every pair of
// list1 x list2 once; the longer list runs over the lanes and k = j + d mod (shorter length).
struct PairIter {
  const unsigned short* la;
  const unsigned short* lb;
  int LA, LB, d, j, dmax;
  bool same, even, done;

  __device__ __forceinline__ void init(int lane, bool s, int L1, int L2, const unsigned short* l1,
                                       const unsigned short* l2) {
    same = s;
    if (same) {
      LA = LB = L1;
      la = lb = l1;
    } else {
      const bool sw = L2 > L1;
      LA = sw ? L2 : L1;
      LB = sw ? L1 : L2;
      la = sw ? l2 : l1;
      lb = sw ? l1 : l2;
    }
    dmax = (LA - 1) / 2;
    even = (LA & 1) == 0;
    d = 0;
    j = lane;
    done = (LA <= 0 || LB <= 0);
    if (!done)
      while (j >= LA) {
        j -= LA;
        ++d;
      }
  }

  __device__ __forceinline__ bool next(int& ej, int& ek, bool& diag, int& w) {
    if (done) return false;
    if (same) {
      if (d > dmax && !(even && d == dmax + 1 && j < (LA >> 1))) {
        done = true;
        return false;
      }
    } else if (d >= LB) {
      done = true;
      return false;
    }
    int k = j + d;
    while (k >= LB) k -= LB;
    ej = la ? (int)la[j] : j;
    ek = lb ? (int)lb[k] : k;
    diag = same && d == 0;
    w = (same && d != 0) ? 2 : 1;
    j += 64;
    while (j >= LA) {
      j -= LA;
      ++d;
    }
    return true;
  }
};

template <typename T, bool ADJ, int MQ>
__global__ void __launch_bounds__(512) k_acsf(AcsfArgs a) {
  using V = typename Vec4<T>::type;
  extern __shared__ __align__(32) unsigned char smem[];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const long long i = a.a0 + (long long)blockIdx.x * (blockDim.x >> 6) + wave;
  if (i >= a.a1) return;
  const int M = a.maxnb;
  Slice<T> sl = make_slice<T, ADJ>(smem + slice_bytes<T, ADJ>(M, a.nZ) * wave, M);

  const int b = a.rowptr[i];
  const int n = a.rowptr[i + 1] - b;
  const V* nd = (const V*)a.nd + b;
  const T pi = (T)kPi;
  const T rcl = (T)a.rclist;

  // ---- stage the centre's entries, list-cutoff fc, per-element index lists ----
  int running[kMaxZ];
#pragma unroll
  for (int c = 0; c < kMaxZ; ++c) running[c] = 0;
  for (int e0 = 0; e0 < n; e0 += 64) {
    const int e = e0 + lane;
    int zc = -1;
    if (e < n) {
      const int jj = a.nj[b + e];
      const V d = nd[e];
      sl.d[e] = d;
      sl.j[e] = jj;
      zc = a.zcode[jj];
      sl.zc[e] = zc;
      const T x = pi * d.w / rcl;
      sl.fc[e] = (T)0.5 * (fcos(x) + (T)1);
      sl.dfc[e] = -(T)0.5 * (pi / rcl) * fsin(x);
    }
#pragma unroll
    for (int c = 0; c < kMaxZ; ++c) {
      if (c < a.nZ) {
        const unsigned long long m = __ballot(e < n && zc == c);
        if (e < n && zc == c) sl.idx[c * M + running[c] + __popcll(m & lanemask_lt(lane))] = e;
        running[c] += __popcll(m);
      }
    }
  }
#pragma unroll
  for (int c = 0; c < kMaxZ; ++c)
    if (c == lane) sl.nz[c] = running[c];
  wave_sync();

  const int nouts = ADJ ? a.nout : 1;
  for (int t = 0; t < nouts; ++t) {
    const double* cG = ADJ ? a.cG + (i * a.nout + t) * a.nG : nullptr;

    // ---- radial functions ----
    if (!ADJ) {
      for (int f0 = 0; f0 < a.nR; f0 += 64) {
        const int nf = min(64, a.nR - f0);
        const int ng = 64 / nf;
        const int fl = lane % nf, gl = lane / nf;
        const bool act = gl < ng;
        const RadialFn& fn = a.radial[f0 + fl];
        const bool shared = (fn.rc == a.rclist) && fn.atomid == 0;
        const T rc = (T)fn.rc, eta = (T)fn.eta, rs = (T)fn.rs, kap = (T)fn.kappa;
        const T qi = (T)ext_q(a.ext, a.next, fn.atomid, i);
        double acc = 0.0;
        if (act) {
          for (int e = gl; e < n; e += ng) {
            const T r = sl.d[e].w;
            const int zc = sl.zc[e];
            if ((fn.zc < 0 || fn.zc == zc) && r <= rc) {
              T fcv;
              if (shared) {
                fcv = sl.fc[e];
              } else {
                const T qj = (T)ext_q(a.ext, a.next, fn.atomid, sl.j[e]);
                fcv = (T)0.5 * qi * qj * (fcos(pi * r / rc) + (T)1);
              }
              T term;
              if (fn.type == 1) term = fcv;
              else if (fn.type == 2) term = fexp(-eta * ((r - rs) * (r - rs))) * fcv;
              else term = fcos(kap * r) * fcv;
              acc += (double)term;
            }
          }
        }
        sl.red[lane] = acc;
        wave_sync();
        if (lane < nf) {
          double s = 0.0;
          for (int g = 0; g < ng; ++g) s += sl.red[g * nf + lane];
          a.G[i * a.nG + fn.col] = s;
        }
        wave_sync();
      }
    } else {
      for (int e0 = 0; e0 < n; e0 += 64) {
        const int e = e0 + lane;
        if (e >= n) break;
        const V d = sl.d[e];
        const int zc = sl.zc[e];
        const T r = d.w;
        T s = 0;
        for (int f = 0; f < a.nR; ++f) {
          const RadialFn& fn = a.radial[f];
          const T cf = (T)cG[fn.col];
          if (cf == (T)0) continue;
          if (!(fn.zc < 0 || fn.zc == zc) || !(r <= (T)fn.rc)) continue;
          T fc, dfc;
          if (fn.rc == a.rclist && fn.atomid == 0) {
            fc = sl.fc[e];
            dfc = sl.dfc[e];
          } else {
            const T rc = (T)fn.rc;
            const T qq = (T)(ext_q(a.ext, a.next, fn.atomid, i) *
                             ext_q(a.ext, a.next, fn.atomid, sl.j[e]));
            const T x = pi * r / rc;
            fc = (T)0.5 * qq * (fcos(x) + (T)1);
            dfc = -(T)0.5 * qq * (pi / rc) * fsin(x);
          }
          T g;
          if (fn.type == 1) {
            g = dfc;
          } else if (fn.type == 2) {
            const T eta = (T)fn.eta, dr = r - (T)fn.rs;
            g = (dfc - fc * (T)2 * eta * dr) * fexp(-eta * dr * dr);
          } else {
            const T kap = (T)fn.kappa;
            g = fcos(kap * r) * dfc - fsin(kap * r) * fc * kap;
          }
          s += cf * g;
        }
        const T sr = s / r;
        sl.gs[3 * e + 0] = sr * d.x;
        sl.gs[3 * e + 1] = sr * d.y;
        sl.gs[3 * e + 2] = sr * d.z;
      }
      wave_sync();
    }

    // ---- angular groups ----
    for (int g = 0; g < a.nAG; ++g) {
      const AngGroup& grp = a.groups[g];
      const int nm = grp.nmem;
      T A[MQ];
      if (ADJ) {
        bool any = false;
#pragma unroll
        for (int q = 0; q < MQ; ++q) {
          A[q] = (q < nm) ? (T)(cG[grp.col[q]] * grp.scale[q]) : (T)0;
          any = any || (A[q] != (T)0);
        }
        if (!any) continue;
      }
      const T rc = (T)grp.rc, eta = (T)grp.eta;
      const bool g4 = grp.type == 4;
      const bool same = grp.same != 0;
      const bool shared = (grp.rc == a.rclist) && grp.atomid == 0;
      const unsigned short* l1 = grp.zc1 >= 0 ? sl.idx + grp.zc1 * M : nullptr;
      const unsigned short* l2 = grp.zc2 >= 0 ? sl.idx + grp.zc2 * M : nullptr;
      const int L1 = grp.zc1 >= 0 ? sl.nz[grp.zc1] : n;
      const int L2 = grp.zc2 >= 0 ? sl.nz[grp.zc2] : n;
      const T qi = (T)ext_q(a.ext, a.next, grp.atomid, i);
      // per-entry radial factor exp(-eta r^2) fc(r) of this group (and its r-derivative)
      for (int e0 = 0; e0 < n; e0 += 64) {
        const int e = e0 + lane;
        if (e >= n) break;
        const T r = sl.d[e].w;
        T qj = 1, E = 0, Ep = 0;
        if (r <= rc) {
          T fc, dfc;
          if (shared) {
            fc = sl.fc[e];
            dfc = sl.dfc[e];
          } else {
            qj = (T)ext_q(a.ext, a.next, grp.atomid, sl.j[e]);
            const T x = pi * r / rc;
            fc = (T)0.5 * qi * qj * (fcos(x) + (T)1);
            dfc = -(T)0.5 * qi * qj * (pi / rc) * fsin(x);
          }
          const T ex = fexp(-eta * r * r);
          E = ex * fc;
          Ep = ex * (dfc - (T)2 * eta * r * fc);
        }
        V pa;
        pa.x = (T)1 / r;
        pa.y = E;
        pa.z = qj;
        pa.w = Ep;
        sl.pa[e] = pa;
      }
      wave_sync();

      int xint[MQ];
      T xi[MQ], lam[MQ];
#pragma unroll
      for (int q = 0; q < MQ; ++q) {
        xint[q] = (q < nm) ? grp.xint[q] : 0;
        xi[q] = (q < nm) ? (T)grp.xi[q] : (T)0;
        lam[q] = (q < nm) ? (T)grp.lam[q] : (T)0;
      }

      if (!ADJ) {
        double acc[MQ];
#pragma unroll
        for (int q = 0; q < MQ; ++q) acc[q] = 0.0;
        PairIter it;
        it.init(lane, same, L1, L2, l1, l2);
        int ej, ek, wi;
        bool diag;
        while (it.next(ej, ek, diag, wi)) {
          const T w = (T)wi;
          const V dj = sl.d[ej], dk = sl.d[ek];
          const V pj = sl.pa[ej], pk = sl.pa[ek];
          T c = (dj.x * dk.x + dj.y * dk.y + dj.z * dk.z) * pj.x * pk.x;
          c = fmin(fmax(c, (T)-1), (T)1);
          T base = w * pj.y * pk.y;
          if (g4) {
            const T rx = dj.x - dk.x, ry = dj.y - dk.y, rz = dj.z - dk.z;
            const T r2 = rx * rx + ry * ry + rz * rz;
            const T rjk = sqrt(r2);
            if (rjk <= rc) {
              const T fcjk = (T)0.5 * pj.z * pk.z * (fcos(pi * rjk / rc) + (T)1);
              base *= fexp(-eta * r2) * fcjk;
            } else {
              base = 0;
            }
          }
#pragma unroll
          for (int q = 0; q < MQ; ++q)
            if (q < nm) acc[q] += (double)(base * xpow((T)1 + lam[q] * c, xint[q], xi[q]));
        }
#pragma unroll
        for (int q = 0; q < MQ; ++q) {
          if (q < nm) {
            const double s = wave_sum(acc[q]);
            if (lane == 0) a.G[i * a.nG + grp.col[q]] = s * grp.scale[q];
          }
        }
      } else {
        // diagonal coefficient: P = 1 + lambda at c = 1 (acsf.F90:1580-1585)
        T S0d = 0;
#pragma unroll
        for (int q = 0; q < MQ; ++q)
          if (q < nm) S0d += A[q] * xpow((T)1 + lam[q], xint[q], xi[q]);
        PairIter it;
        it.init(lane, same, L1, L2, l1, l2);
        int ej, ek, wi;
        bool diag;
        while (it.next(ej, ek, diag, wi)) {
          const T w = (T)wi;
          const V dj = sl.d[ej];
          const V pj = sl.pa[ej];
          const T irj = pj.x;
          const T Ej = pj.y, Epj = pj.w;
          if (diag) {
            const T H = g4 ? pj.z * pj.z : (T)1;
            const T gsc = S0d * (T)2 * Ej * Epj * H * irj;
            atomicAdd(&sl.gs[3 * ej + 0], gsc * dj.x);
            atomicAdd(&sl.gs[3 * ej + 1], gsc * dj.y);
            atomicAdd(&sl.gs[3 * ej + 2], gsc * dj.z);
            continue;
          }
          const V dk = sl.d[ek];
          const V pk = sl.pa[ek];
          const T irk = pk.x;
          const T Ek = pk.y, Epk = pk.w;
          T H = 1, Hp = 0, rjk = 1;
          T rx = 0, ry = 0, rz = 0;
          if (g4) {
            rx = dj.x - dk.x;
            ry = dj.y - dk.y;
            rz = dj.z - dk.z;
            const T r2 = rx * rx + ry * ry + rz * rz;
            rjk = sqrt(r2);
            if (!(rjk <= rc)) continue;
            const T qq = pj.z * pk.z;
            const T x = pi * rjk / rc;
            const T ex = fexp(-eta * r2);
            const T fc = (T)0.5 * qq * (fcos(x) + (T)1);
            const T dfc = -(T)0.5 * qq * (pi / rc) * fsin(x);
            H = ex * fc;
            Hp = ex * (dfc - (T)2 * eta * rjk * fc);
          }
          const T ujx = dj.x * irj, ujy = dj.y * irj, ujz = dj.z * irj;
          const T ukx = dk.x * irk, uky = dk.y * irk, ukz = dk.z * irk;
          const T c = fmin(fmax(ujx * ukx + ujy * uky + ujz * ukz, (T)-1), (T)1);
          T S0 = 0, S1 = 0;
#pragma unroll
          for (int q = 0; q < MQ; ++q) {
            if (q < nm) {
              const T P = (T)1 + lam[q] * c;
              T pw, pm1;
              if (xint[q] == 0) {
                pw = (T)1;
                pm1 = (T)0;
              } else if (xint[q] > 0) {
                pm1 = xpow(P, xint[q] - 1, (T)0);
                pw = pm1 * P;
              } else {
                pw = fpow(P, xi[q]);
                pm1 = fpow(P, xi[q] - (T)1);
              }
              S0 += A[q] * pw;
              S1 += A[q] * xi[q] * lam[q] * pm1;
            }
          }
          const T B = Ej * Ek * H;
          const T cj = w * S1 * B * irj, ck = w * S1 * B * irk;
          const T aj = w * S0 * Epj * Ek * H, ak = w * S0 * Ej * Epk * H;
          T gjx = cj * (ukx - c * ujx) + aj * ujx;
          T gjy = cj * (uky - c * ujy) + aj * ujy;
          T gjz = cj * (ukz - c * ujz) + aj * ujz;
          T gkx = ck * (ujx - c * ukx) + ak * ukx;
          T gky = ck * (ujy - c * uky) + ak * uky;
          T gkz = ck * (ujz - c * ukz) + ak * ukz;
          if (g4) {
            const T h = w * S0 * Ej * Ek * Hp / rjk;
            gjx += h * rx;
            gjy += h * ry;
            gjz += h * rz;
            gkx -= h * rx;
            gky -= h * ry;
            gkz -= h * rz;
          }
          atomicAdd(&sl.gs[3 * ej + 0], gjx);
          atomicAdd(&sl.gs[3 * ej + 1], gjy);
          atomicAdd(&sl.gs[3 * ej + 2], gjz);
          atomicAdd(&sl.gs[3 * ek + 0], gkx);
          atomicAdd(&sl.gs[3 * ek + 1], gky);
          atomicAdd(&sl.gs[3 * ek + 2], gkz);
        }
      }
      wave_sync();
    }

    // ---- scatter of the adjoint to the forces ----
    if (ADJ) {
      double* F = a.F;
      double sx = 0, sy = 0, sz = 0;
      for (int e0 = 0; e0 < n; e0 += 64) {
        const int e = e0 + lane;
        if (e >= n) break;
        const double gx = (double)sl.gs[3 * e], gy = (double)sl.gs[3 * e + 1],
                     gz = (double)sl.gs[3 * e + 2];
        const long long jj = sl.j[e];
        unsafeAtomicAdd(&F[(jj * a.nout + t) * 3 + 0], -gx);
        unsafeAtomicAdd(&F[(jj * a.nout + t) * 3 + 1], -gy);
        unsafeAtomicAdd(&F[(jj * a.nout + t) * 3 + 2], -gz);
        sx += gx;
        sy += gy;
        sz += gz;
      }
      sx = wave_sum(sx);
      sy = wave_sum(sy);
      sz = wave_sum(sz);
      if (lane == 0) {
        unsafeAtomicAdd(&F[(i * a.nout + t) * 3 + 0], sx);
        unsafeAtomicAdd(&F[(i * a.nout + t) * 3 + 1], sy);
        unsafeAtomicAdd(&F[(i * a.nout + t) * 3 + 2], sz);
      }
      wave_sync();
    }
  }
}

template <typename T, bool ADJ, int MQ>
hipError_t launch_mq(fnet_batch* b, const DevModel& m, long long a0, long long a1,
                     hipStream_t s) {
  const long long n = a1 - a0;
  if (n <= 0) return hipSuccess;
  const int M = ((b->max_nb + 31) / 32) * 32 + 32;
  const size_t sb = slice_bytes<T, ADJ>(M, m.nZ);
  int wpb = 8;
  while (wpb > 1 && sb * wpb > 64 * 1024) wpb >>= 1;
  if (sb * wpb > 64 * 1024) return hipErrorInvalidConfiguration;
  AcsfArgs a;
  a.rowptr = b->rowptr;
  a.nj = b->nj;
  a.nd = b->nd;
  a.zcode = b->zcode;
  a.ext = b->ext;
  a.next = b->next;
  a.radial = m.radial;
  a.nR = m.nR;
  a.groups = m.groups;
  a.nAG = m.nAG;
  a.nZ = m.nZ;
  a.nG = m.nG;
  a.nout = m.nout;
  a.rclist = m.rcmax;
  a.a0 = a0;
  a.a1 = a1;
  a.maxnb = M;
  a.G = (double*)b->G;
  a.cG = (const double*)b->cG;
  a.F = (double*)b->F;
  const unsigned grid = (unsigned)((n + wpb - 1) / wpb);
  k_acsf<T, ADJ, MQ><<<grid, 64 * wpb, sb * wpb, s>>>(a);
  return hipGetLastError();
}

template <typename T, bool ADJ>
hipError_t launch(fnet_batch* b, const DevModel& m, long long a0, long long a1, hipStream_t s) {
  if (m.maxmem <= 4) return launch_mq<T, ADJ, 4>(b, m, a0, a1, s);
  if (m.maxmem <= 8) return launch_mq<T, ADJ, 8>(b, m, a0, a1, s);
  return launch_mq<T, ADJ, kMaxMembers>(b, m, a0, a1, s);
}

}  // namespace

size_t lds_bytes_acsf(int max_nb, int nZ, int fp64, int adj) {
  const int M = ((max_nb + 31) / 32) * 32 + 32;
  if (fp64) return adj ? slice_bytes<double, true>(M, nZ) : slice_bytes<double, false>(M, nZ);
  return adj ? slice_bytes<float, true>(M, nZ) : slice_bytes<float, false>(M, nZ);
}

hipError_t launch_acsf_fwd(fnet_batch* b, const DevModel& m, long long a0, long long a1, int fp64,
                           hipStream_t s) {
  return fp64 ? launch<double, false>(b, m, a0, a1, s) : launch<float, false>(b, m, a0, a1, s);
}

hipError_t launch_acsf_adj(fnet_batch* b, const DevModel& m, long long a0, long long a1, int fp64,
                           hipStream_t s) {
  return fp64 ? launch<double, true>(b, m, a0, a1, s) : launch<float, true>(b, m, a0, a1, s);
}

}  // namespace fnet
