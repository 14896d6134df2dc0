// ACSF values and their adjoint (force) pass on gfx950.
//
// One wavefront per centre atom.  The centre's CSR row of the neighbour list (image displacement
// and distance per entry) is staged once into the wave's LDS slice with coalesced loads, and
// per-element index lists are built with ballots, so every species-resolved list of
// buildGFunctionNeighborlists (acsf.F90:943-1066) is a slice of LDS instead of a rebuilt
// neighbour list.
//
//  * radial G1/G2/G3 (acsf.F90:1287-1373): lanes = functions, loop over the staged entries;
//  * angular G4/G5 (acsf.F90:1377-1492): functions that share (type, cutoff, eta, species pair,
//    atom id) form a group; lanes stride over the neighbour pairs of the group once and update
//    every member.  When both lists are the same (unresolved, or Z1 == Z2) the ordered double
//    sum with its j == k diagonal is taken as 2 * (j < k) + diagonal;
//  * the adjoint pass computes F = -dE/dR directly from c = dE/dG (the network's input gradient
//    times the z-score factor): per entry, the derivative of every function is weighted by c and
//    summed into one 3-vector per neighbour entry (LDS atomics), which is then scattered to the
//    neighbour's force (-) and the centre's force (+).  This is gFuncGrad (acsf.F90:1496-1670)
//    contracted with the Jacobian as in forces.F90:400-413, without the dense dG/dr array, and
//    with periodic images summed exactly.
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

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
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
  void* G;        // [N][nG] (forward)
  const void* cG; // [N][nout][nG] (adjoint)
  void* F;        // [N][nout][3] (adjoint)
};

// LDS slice of one wave.
template <typename T, bool ADJ>
struct Slice {
  using V = typename Vec4<T>::type;
  V* d;          // [M] displacement (x,y,z) and distance (w)
  int* j;        // [M] global neighbour index
  int* zc;       // [M] element code
  T* E;          // [M] exp(-eta r^2) fc(r) of the current group
  T* Ep;         // [M] its r-derivative (adjoint)
  T* q;          // [M] atom-id prefactor of the current group
  T* gs;         // [3M] accumulated dE/d(entry position) (adjoint)
  unsigned short* idx; // [nZ][M]
  int* nz;       // [nZ]
};

template <typename T, bool ADJ>
__device__ size_t slice_bytes(int M, int nZ) {
  size_t b = sizeof(typename Vec4<T>::type) * M + 2 * sizeof(int) * M + 2 * sizeof(T) * M;
  if (ADJ) b += sizeof(T) * M + 3 * sizeof(T) * M;
  b += sizeof(unsigned short) * (size_t)nZ * M + sizeof(int) * kMaxZ;
  return (b + 31) & ~size_t(31);
}

template <typename T, bool ADJ>
__device__ Slice<T, ADJ> make_slice(unsigned char* base, int M, int nZ) {
  using V = typename Vec4<T>::type;
  Slice<T, ADJ> s;
  unsigned char* p = base;
  s.d = (V*)p;
  p += sizeof(V) * M;
  s.E = (T*)p;
  p += sizeof(T) * M;
  s.q = (T*)p;
  p += sizeof(T) * M;
  if (ADJ) {
    s.Ep = (T*)p;
    p += sizeof(T) * M;
    s.gs = (T*)p;
    p += 3 * sizeof(T) * M;
  } else {
    s.Ep = nullptr;
    s.gs = nullptr;
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

// Enumerates the pairs of one group: same lists -> j <= k (triangle with diagonal), otherwise the
// full rectangle list1 x list2.  Calls fn(e_j, e_k, diag).
template <typename Fn>
__device__ __forceinline__ void for_pairs(int lane, bool same, int L1, int L2,
                                          const unsigned short* l1, const unsigned short* l2,
                                          Fn&& fn) {
  if (same) {
    if (L1 <= 0) return;
    int j = 0, k = lane;
    while (j < L1 && k >= L1) {
      k = k - L1 + j + 1;
      ++j;
    }
    while (j < L1) {
      const int ej = l1 ? l1[j] : j;
      const int ek = l1 ? l1[k] : k;
      fn(ej, ek, j == k);
      k += 64;
      while (j < L1 && k >= L1) {
        k = k - L1 + j + 1;
        ++j;
      }
    }
  } else {
    if (L1 <= 0 || L2 <= 0) return;
    int j = 0, k = lane;
    while (j < L1 && k >= L2) {
      k -= L2;
      ++j;
    }
    while (j < L1) {
      const int ej = l1 ? l1[j] : j;
      const int ek = l2 ? l2[k] : k;
      fn(ej, ek, false);
      k += 64;
      while (j < L1 && k >= L2) {
        k -= L2;
        ++j;
      }
    }
  }
}

template <typename T, bool ADJ>
__global__ void __launch_bounds__(256) k_acsf(AcsfArgs a) {
  using V = typename Vec4<T>::type;
  extern __shared__ __align__(32) unsigned char smem[];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const long long i = a.a0 + (long long)blockIdx.x * (blockDim.x >> 6) + wave;
  if (i >= a.a1) return;
  const int M = a.maxnb;
  const size_t sb = slice_bytes<T, ADJ>(M, a.nZ);
  Slice<T, ADJ> sl = make_slice<T, ADJ>(smem + sb * wave, M, a.nZ);

  const int b = a.rowptr[i];
  const int n = a.rowptr[i + 1] - b;
  const V* nd = (const V*)a.nd + b;

  // ---- stage the centre's entries and build per-element index lists ----
  if (lane < kMaxZ) sl.nz[lane] = 0;
  int running[kMaxZ];
#pragma unroll
  for (int c = 0; c < kMaxZ; ++c) running[c] = 0;
  for (int e0 = 0; e0 < n; e0 += 64) {
    const int e = e0 + lane;
    int zc = -1;
    if (e < n) {
      const int jj = a.nj[b + e];
      sl.d[e] = nd[e];
      sl.j[e] = jj;
      zc = a.zcode[jj];
      sl.zc[e] = zc;
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
  if (lane < a.nZ) {
#pragma unroll
    for (int c = 0; c < kMaxZ; ++c)
      if (c == lane) sl.nz[c] = running[c];
  }
  wave_sync();

  const T pi = (T)kPi;
  const int nouts = ADJ ? a.nout : 1;
  for (int t = 0; t < nouts; ++t) {
    const T* cG = ADJ ? (const T*)a.cG + (i * a.nout + t) * a.nG : nullptr;

    // ---- radial functions ----
    if (!ADJ) {
      for (int f0 = 0; f0 < a.nR; f0 += 64) {
        const int f = f0 + lane;
        const bool act = f < a.nR;
        RadialFn fn;
        if (act) fn = a.radial[f];
        else {
          fn.col = 0; fn.type = 0; fn.zc = -1; fn.atomid = 0; fn.rc = 1; fn.eta = fn.rs = fn.kappa = 0;
        }
        const T rc = (T)fn.rc, eta = (T)fn.eta, rs = (T)fn.rs, kap = (T)fn.kappa;
        const T qi = (T)ext_q(a.ext, a.next, fn.atomid, i);
        T acc = 0;
        for (int e = 0; e < n; ++e) {
          const V d = sl.d[e];
          const int zc = sl.zc[e];
          if (act && (fn.zc < 0 || fn.zc == zc) && d.w <= rc) {
            const T qj = fn.atomid > 0 ? (T)ext_q(a.ext, a.next, fn.atomid, sl.j[e]) : (T)1;
            const T fc = (T)0.5 * qi * qj * (fcos(pi * d.w / rc) + (T)1);
            if (fn.type == 1) acc += fc;
            else if (fn.type == 2) acc += fexp(-eta * ((d.w - rs) * (d.w - rs))) * fc;
            else acc += fcos(kap * d.w) * fc;
          }
        }
        if (act) ((T*)a.G)[i * a.nG + fn.col] = acc;
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
          const T cf = cG[fn.col];
          if (cf == (T)0) continue;
          if (!(fn.zc < 0 || fn.zc == zc) || !(r <= (T)fn.rc)) continue;
          const T rc = (T)fn.rc;
          const T qq = (T)(ext_q(a.ext, a.next, fn.atomid, i) *
                           (fn.atomid > 0 ? ext_q(a.ext, a.next, fn.atomid, sl.j[e]) : 1.0));
          const T x = pi * r / rc;
          const T fc = (T)0.5 * qq * (fcos(x) + (T)1);
          const T dfc = -(T)0.5 * qq * (pi / rc) * fsin(x);
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
      T A[kMaxMembers];
      bool any = !ADJ;
      if (ADJ) {
#pragma unroll
        for (int q = 0; q < kMaxMembers; ++q) {
          A[q] = (q < nm) ? cG[grp.col[q]] * (T)grp.scale[q] : (T)0;
          any = any || (A[q] != (T)0);
        }
        if (!any) continue;
      }
      const T rc = (T)grp.rc, eta = (T)grp.eta;
      const bool g4 = grp.type == 4;
      const bool same = grp.same != 0;
      const unsigned short* l1 = grp.zc1 >= 0 ? sl.idx + grp.zc1 * M : nullptr;
      const unsigned short* l2 = grp.zc2 >= 0 ? sl.idx + grp.zc2 * M : nullptr;
      const int L1 = grp.zc1 >= 0 ? sl.nz[grp.zc1] : n;
      const int L2 = same ? L1 : (grp.zc2 >= 0 ? sl.nz[grp.zc2] : n);
      const T qi = (T)ext_q(a.ext, a.next, grp.atomid, i);
      // per-entry radial factors of this group
      for (int e0 = 0; e0 < n; e0 += 64) {
        const int e = e0 + lane;
        if (e >= n) break;
        const T r = sl.d[e].w;
        const T qj = grp.atomid > 0 ? (T)ext_q(a.ext, a.next, grp.atomid, sl.j[e]) : (T)1;
        sl.q[e] = qj;
        T E = 0, Ep = 0;
        if (r <= rc) {
          const T x = pi * r / rc;
          const T ex = fexp(-eta * r * r);
          const T fc = (T)0.5 * qi * qj * (fcos(x) + (T)1);
          E = ex * fc;
          if (ADJ) {
            const T dfc = -(T)0.5 * qi * qj * (pi / rc) * fsin(x);
            Ep = ex * (dfc - (T)2 * eta * r * fc);
          }
        }
        sl.E[e] = E;
        if (ADJ) sl.Ep[e] = Ep;
      }
      wave_sync();

      int xint[kMaxMembers];
      T xi[kMaxMembers], lam[kMaxMembers];
#pragma unroll
      for (int q = 0; q < kMaxMembers; ++q) {
        xint[q] = (q < nm) ? grp.xint[q] : 0;
        xi[q] = (q < nm) ? (T)grp.xi[q] : (T)0;
        lam[q] = (q < nm) ? (T)grp.lam[q] : (T)0;
      }

      if (!ADJ) {
        T acc[kMaxMembers];
#pragma unroll
        for (int q = 0; q < kMaxMembers; ++q) acc[q] = 0;
        for_pairs(lane, same, L1, L2, l1, l2, [&](int ej, int ek, bool diag) {
          const V dj = sl.d[ej], dk = sl.d[ek];
          const T dot = dj.x * dk.x + dj.y * dk.y + dj.z * dk.z;
          const T c = dot / (dj.w * dk.w + (T)1e-13);
          T base = sl.E[ej] * sl.E[ek] * ((same && !diag) ? (T)2 : (T)1);
          if (g4) {
            const T rx = dj.x - dk.x, ry = dj.y - dk.y, rz = dj.z - dk.z;
            const T r2 = rx * rx + ry * ry + rz * rz;
            const T rjk = sqrt(r2);
            if (rjk <= rc) {
              const T fcjk = (T)0.5 * sl.q[ej] * sl.q[ek] * (fcos(pi * rjk / rc) + (T)1);
              base *= fexp(-eta * r2) * fcjk;
            } else {
              base = 0;
            }
          }
#pragma unroll
          for (int q = 0; q < kMaxMembers; ++q)
            if (q < nm) acc[q] += base * xpow((T)1 + lam[q] * c, xint[q], xi[q]);
        });
#pragma unroll
        for (int q = 0; q < kMaxMembers; ++q) {
          if (q < nm) {
            const T s = wave_sum(acc[q]);
            if (lane == 0) ((T*)a.G)[i * a.nG + grp.col[q]] = s * (T)grp.scale[q];
          }
        }
      } else {
        // diagonal coefficient: P = 1 + lambda at c = 1 (acsf.F90:1580-1585)
        T S0d = 0;
#pragma unroll
        for (int q = 0; q < kMaxMembers; ++q)
          if (q < nm) S0d += A[q] * xpow((T)1 + lam[q], xint[q], xi[q]);
        for_pairs(lane, same, L1, L2, l1, l2, [&](int ej, int ek, bool diag) {
          const V dj = sl.d[ej];
          const T rj = dj.w;
          const T Ej = sl.E[ej], Epj = sl.Ep[ej];
          if (diag) {
            const T H = g4 ? sl.q[ej] * sl.q[ej] : (T)1;
            const T gsc = S0d * (T)2 * Ej * Epj * H / rj;
            atomicAdd(&sl.gs[3 * ej + 0], gsc * dj.x);
            atomicAdd(&sl.gs[3 * ej + 1], gsc * dj.y);
            atomicAdd(&sl.gs[3 * ej + 2], gsc * dj.z);
            return;
          }
          const V dk = sl.d[ek];
          const T rk = dk.w;
          const T Ek = sl.E[ek], Epk = sl.Ep[ek];
          T H = 1, Hp = 0, rjk = 0;
          T rx = 0, ry = 0, rz = 0;
          if (g4) {
            rx = dj.x - dk.x;
            ry = dj.y - dk.y;
            rz = dj.z - dk.z;
            const T r2 = rx * rx + ry * ry + rz * rz;
            rjk = sqrt(r2);
            if (!(rjk <= rc)) return;
            const T qq = sl.q[ej] * sl.q[ek];
            const T x = pi * rjk / rc;
            const T ex = fexp(-eta * r2);
            const T fc = (T)0.5 * qq * (fcos(x) + (T)1);
            const T dfc = -(T)0.5 * qq * (pi / rc) * fsin(x);
            H = ex * fc;
            Hp = ex * (dfc - (T)2 * eta * rjk * fc);
          }
          const T irj = (T)1 / rj, irk = (T)1 / rk;
          const T ujx = dj.x * irj, ujy = dj.y * irj, ujz = dj.z * irj;
          const T ukx = dk.x * irk, uky = dk.y * irk, ukz = dk.z * irk;
          const T c = ujx * ukx + ujy * uky + ujz * ukz;
          T S0 = 0, S1 = 0;
#pragma unroll
          for (int q = 0; q < kMaxMembers; ++q) {
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
          const T w = same ? (T)2 : (T)1;
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
        });
      }
      wave_sync();
    }

    // ---- scatter of the adjoint to the forces ----
    if (ADJ) {
      T* F = (T*)a.F;
      T sx = 0, sy = 0, sz = 0;
      for (int e0 = 0; e0 < n; e0 += 64) {
        const int e = e0 + lane;
        if (e >= n) break;
        const T gx = sl.gs[3 * e], gy = sl.gs[3 * e + 1], gz = sl.gs[3 * e + 2];
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

template <typename T, bool ADJ>
size_t host_slice_bytes(int M, int nZ) {
  using V = typename Vec4<T>::type;
  size_t b = sizeof(V) * M + 2 * sizeof(int) * M + 2 * sizeof(T) * M;
  if (ADJ) b += sizeof(T) * M + 3 * sizeof(T) * M;
  b += sizeof(unsigned short) * (size_t)nZ * M + sizeof(int) * kMaxZ;
  return (b + 31) & ~size_t(31);
}

template <typename T, bool ADJ>
hipError_t launch(fnet_batch* b, const DevModel& m, long long a0, long long a1, hipStream_t s) {
  const long long n = a1 - a0;
  if (n <= 0) return hipSuccess;
  const int M = ((b->max_nb + 63) / 64) * 64 + 64;
  const size_t sb = host_slice_bytes<T, ADJ>(M, m.nZ);
  int wpb = 4;
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
  a.G = b->G;
  a.cG = b->cG;
  a.F = b->F;
  const unsigned grid = (unsigned)((n + wpb - 1) / wpb);
  k_acsf<T, ADJ><<<grid, 64 * wpb, sb * wpb, s>>>(a);
  return hipGetLastError();
}

}  // namespace

size_t lds_bytes_acsf(int max_nb, int nZ, int fp64, int adj) {
  const int M = ((max_nb + 63) / 64) * 64 + 64;
  if (fp64) return adj ? host_slice_bytes<double, true>(M, nZ) : host_slice_bytes<double, false>(M, nZ);
  return adj ? host_slice_bytes<float, true>(M, nZ) : host_slice_bytes<float, false>(M, nZ);
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
