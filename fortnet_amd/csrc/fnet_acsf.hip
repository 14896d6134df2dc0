// ACSF values and their adjoint (force) pass on gfx950.
//
// One wavefront per centre atom.  The centre's CSR row of the neighbour list (image displacement
// and distance per entry) is staged once into the wave's LDS slice with coalesced loads, ordered
// by element (a counting pass and a placing pass, both with ballots), so every species-resolved
// neighbour list of buildGFunctionNeighborlists (acsf.F90:943-1066) is a contiguous range of the
// slice instead of a rebuilt neighbour list.
//
//  * radial G1/G2/G3 (acsf.F90:1287-1373): lanes = (entry stride group, function), partial sums
//    combined through LDS;
//  * angular G4/G5 (acsf.F90:1377-1492): functions sharing (type, cutoff, eta, species pair,
//    atom id) form a group; lanes stride over the group's neighbour pairs once and update every
//    member.  Pairs are enumerated by offset (k = j + d): the diagonal, then each unordered pair
//    once with weight 2 when both lists are the same (the ordered double sum with its j == k
//    terms, acsf.F90:1420-1432), or the full list1 x list2 product otherwise.  With this order
//    the 64 lanes of one step touch 64 different entries, which keeps the adjoint's LDS
//    accumulation free of same-address collisions, and the walk needs no division per step.
//    When every group's exponents are powers of two (up to 16) with at most two lambda values,
//    the kernel is built for that form: the powers come from one squaring chain per lambda, the
//    same straight-line code for every member (a per-member branch on the exponent cost 25 % of
//    the adjoint's time on the water boxes); other models use the per-member loop;
//  * the adjoint pass computes F = -dE/dR directly from c = dE/dG (the network's input gradient
//    times the z-score factor): the derivative of every function with respect to each entry's
//    position (gFuncGrad, acsf.F90:1496-1670) is weighted by c and summed per entry, then scattered
//    to the neighbour's force (-) and the centre's force (+), as forces.F90:400-413 contracts the
//    dense array, but without forming it and with periodic images summed exactly.
//
// Arithmetic is T (double or float); sums over neighbours and pairs, the values G and the forces
// are accumulated in FP64 in both modes.  FP64 evaluates the elementary functions with libm in the
// reference's operation order; FP32 uses the hardware instructions (v_exp_f32, v_cos_f32,
// v_sin_f32, v_sqrt_f32, v_rcp_f32: about 1 ulp, cos/sin to ~1e-6 absolute on [0, pi]).
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
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double fsqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double frcp(double x) { return 1.0 / x; }
__device__ __forceinline__ float fdiv(float x, float y) { return x * __builtin_amdgcn_rcpf(y); }
// r = sqrt(r2) and 1 / r (FP32: one v_rsq_f32)
__device__ __forceinline__ void dist_inv(float r2, float& r, float& ir) {
  ir = __builtin_amdgcn_rsqf(r2);
  r = r2 * ir;
}
__device__ __forceinline__ void dist_inv(double r2, double& r, double& ir) {
  r = sqrt(r2);
  ir = 1.0 / r;
}
__device__ __forceinline__ double fdiv(double x, double y) { return x / y; }
// argument pi r / rc of the cosine cutoff (FP32: with the wave-uniform pi / rc hoisted)
__device__ __forceinline__ float cutarg(float r, float pi, float rc, float pirc) {
  return r * pirc;
}
__device__ __forceinline__ double cutarg(double r, double pi, double rc, double pirc) {
  return pi * r / rc;
}

// A wave-uniform value moved to scalar registers (values converted from FP64 on the vector
// unit would otherwise hold one vector register each through the pair loops).
__device__ __forceinline__ float uniform(float x) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}
__device__ __forceinline__ double uniform(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffll));
  const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

// Wave sums of NA per-lane values at once (NA a power of two, at most 64), by halving exchange:
// at the step of lane bit m (32, 16, ...) a lane keeps the half of its values that bit m of its
// index selects and adds the partner's copy of that half, so after log2(NA) steps it holds one
// value, index idx, summed over the lanes that differ in those bits; the remaining bits are
// summed as in wave_sum.  NA sums take NA - 1 + log2(64 / NA) exchanges instead of 6 NA, and
// every addition pairs the same partial sums as wave_sum's butterfly, so each total is
// bit-identical to wave_sum of that value.  Lanes whose low log2(64 / NA) bits are zero hold the
// NA totals once each.
template <int NA, int H, int M>
__device__ __forceinline__ void scatter_step(double (&v)[NA], int lane, int& idx) {
  if constexpr (H >= 1) {
    const bool up = (lane & M) != 0;
#pragma unroll
    for (int q = 0; q < H; ++q) {
      const double keep = up ? v[q + H] : v[q];
      const double send = up ? v[q] : v[q + H];
      v[q] = keep + __shfl_xor(send, M, 64);
    }
    if (up) idx += H;
    scatter_step<NA, H / 2, M / 2>(v, lane, idx);
  }
}

template <int NA>
__device__ __forceinline__ double wave_sum_scatter(double (&v)[NA], int lane, int& idx) {
  static_assert(NA >= 1 && NA <= 64 && (NA & (NA - 1)) == 0, "NA: a power of two <= 64");
  idx = 0;
  scatter_step<NA, NA / 2, 32>(v, lane, idx);
  double s = v[0];
#pragma unroll
  for (int m = 32 / NA; m >= 1; m >>= 1) s += __shfl_xor(s, m, 64);
  return s;
}

__host__ __device__ constexpr int pow2_at_least(int n) { return n <= 1 ? 1 : 2 * pow2_at_least((n + 1) / 2); }

__device__ __forceinline__ int wave_sum_i(int v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

// exclusive prefix sum over the 64 lanes
__device__ __forceinline__ int wave_excl_scan_i(int v, int lane) {
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  return x - v;
}

// cos(theta) of an off-diagonal pair in the G4/G5 values, clamped to [-1, 1].  FP64 follows the
// reference's value formula dot / (r_j r_k + 1e-13) (theta, acsf.F90:1159-1176; its
// cos(acos(x)) round trip moves x by about an ulp): the regulariser lowers every cosine by a
// relative ~1e-15, always in the same direction, and z-scores with sigma ~1e-7 turn that shift
// into force changes of ~1e-9, so it is kept.  FP32: dot (1 / r_j) (1 / r_k).
template <typename V>
__device__ __forceinline__ float value_cos(const V& dj, const V& dk, float irj, float irk) {
  const float c = (dj.x * dk.x + dj.y * dk.y + dj.z * dk.z) * irj * irk;
  return fminf(fmaxf(c, -1.0f), 1.0f);
}
template <typename V>
__device__ __forceinline__ double value_cos(const V& dj, const V& dk, double irj, double irk) {
  const double c = (dj.x * dk.x + dj.y * dk.y + dj.z * dk.z) / (dj.w * dk.w + 1e-13);
  return fmin(fmax(c, -1.0), 1.0);
}

// cos(theta) of the j == k term of a value: 1 in FP32; in FP64 the reference's
// dot(d, d) / (r^2 + 1e-13), just below 1 (the gradient code uses 1 exactly, acsf.F90:1580-1585).
template <typename V>
__device__ __forceinline__ float diag_cos(const V& d, float) { return 1.0f; }
template <typename V>
__device__ __forceinline__ double diag_cos(const V& d, double) {
  return (d.x * d.x + d.y * d.y + d.z * d.z) / (d.w * d.w + 1e-13);
}

struct AcsfArgs {
  const int* rowptr;  // CSR rows, or (cap > 0) row i at i * cap with cnt[i] entries
  const int* cnt;
  int cap;
  const int* nj;
  const void* nd;
  const int* zcode;
  const double* ext;
  int next;
  const RadialFn* radial;
  int nR;
  int rbal;  // forward: radial lanes in proportion to the functions' element ranges
  const AngGroup* groups;
  int nAG;
  int nZ;
  int nG;
  int nout;
  double rclist;
  long long a0, a1;
  int maxnb;
  int persist;      // unused by the kernels (the persistent grid is an instance of its own)
  int zone;         // the model has one element slot and every atom takes it
  double* G;        // [N][nG] (forward)
  const double* cG; // [N][nout][nG] (adjoint)
  double* F;        // [N][nout][3] (adjoint): the centre's own sum of its entries' gradients
  double* gvec;     // [entries][nout][3] (adjoint): gradient of E with respect to each entry's
                    // image position, stored at the index of the entry's mirror (the entry of
                    // the neighbour's row that points back), or at its own index without one;
                    // FP64 values from the FP64 kernels, FP32 values from the FP32 ones
  const int* mir;   // [entries] mirror entries (-1: none)
};

// Bytes of one wave's LDS slice (host and device agree on this layout).  Per entry, in element
// order: the displacement and distance d, the pair-loop factors pa = (1/r, exp(-eta r^2) fc,
// atom-id prefactor, radial derivative / r), then the radial partial sums and the list-cutoff fc
// (forward only) or the FP64 adjoint accumulator (adjoint only), the neighbour index and element
// code, and the start and length of each element's range.  The slice is kept small: the number of
// resident waves per CU, and with it the adjoint's speed, follows the LDS it takes (a slice 1.5x
// larger made the adjoint 1.3x slower on the water boxes).  With paired groups the slice also
// holds pb, the partner group's pa.
template <typename T, bool ADJ>
__host__ __device__ size_t slice_bytes(int M, int pairs = 0) {
  size_t b = 2 * sizeof(typename Vec4<T>::type) * M;  // d, pa
  if (pairs) b += sizeof(typename Vec4<T>::type) * M;  // pb
  if (!ADJ) b += 64 * sizeof(double) + sizeof(T) * M;  // radial partial sums, fc
  if (ADJ) b += 3 * sizeof(double) * M;               // gs
  b += 2 * sizeof(int) * M;                           // j, zc
  b += 2 * sizeof(int) * (kMaxZ + 1);                 // st, nz
  return (b + 31) & ~size_t(31);
}

// Entries per slice: the largest row, rounded up to a multiple of 8 (keeps every array aligned).
__host__ __device__ inline int slice_entries(int max_nb) {
  return max_nb < 8 ? 8 : ((max_nb + 7) / 8) * 8;
}

// LDS-typed pointer: its index arithmetic stays 32-bit (a generic pointer's is 64-bit)
typedef __attribute__((address_space(3))) double lds_f64;

// The model's tables and the dE/dG rows do not change during a launch and are read through the
// constant address space: wave-uniform reads of them then become scalar loads, which the compiler
// does not emit for generic pointers in a kernel that also writes global memory (as vector loads,
// the chains slot -> column -> dE/dG kept the Si adjoints waiting on memory ~60% of the time).
typedef const __attribute__((address_space(4))) AngGroup CGroup;
typedef const __attribute__((address_space(4))) RadialFn CRadial;
typedef const __attribute__((address_space(4))) double cdouble;

__device__ __forceinline__ void lds_add(lds_f64* p, double v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}

template <typename T>
struct Slice {
  using V = typename Vec4<T>::type;
  V* d;
  V* pa;
  V* pb;  // adjoint with paired groups: the partner group's pa
  T* fc;  // forward only
  lds_f64* gs;  // FP64: ds_add_f64 measured faster than ds_add_f32 here, and exact to more bits
  int *j, *zc;
  double* red;
  int *st, *nz;
};

template <typename T, bool ADJ>
__device__ Slice<T> make_slice(unsigned char* p, int M, int pairs) {
  using V = typename Vec4<T>::type;
  Slice<T> s;
  s.d = (V*)p;
  p += sizeof(V) * M;
  s.pa = (V*)p;
  p += sizeof(V) * M;
  s.pb = nullptr;
  if (pairs) {
    s.pb = (V*)p;
    p += sizeof(V) * M;
  }
  s.red = nullptr;
  s.fc = nullptr;
  if (!ADJ) {
    s.red = (double*)p;
    p += 64 * sizeof(double);
    s.fc = (T*)p;
    p += sizeof(T) * M;
  }
  s.gs = nullptr;
  if (ADJ) {
    s.gs = (lds_f64*)p;
    p += 3 * sizeof(double) * M;
  }
  s.j = (int*)p;
  p += sizeof(int) * M;
  s.zc = (int*)p;
  p += sizeof(int) * M;
  s.st = (int*)p;
  p += sizeof(int) * (kMaxZ + 1);
  s.nz = (int*)p;
  return s;
}

__device__ __forceinline__ double ext_q(const double* ext, int next, int atomid, long long atom) {
  return atomid > 0 ? ext[atom * next + (atomid - 1)] : 1.0;
}

// Off-diagonal pair walk of one group over the element ranges [A0, A0 + LA) and [B0, B0 + LB).
// Step index t = 64 s + lane + t0 maps to (j, d) = (t mod LA, t div LA) and k = (j + d) mod LB;
// j and k advance by wave-uniform increments, so a step needs no division.  Same lists (SAME):
// t0 = LA skips the diagonal d = 0, which the callers treat on their own; every unordered pair
// then comes once (weight 2): d = 1 .. (LA - 1) / 2, plus half of d = LA / 2 when LA is even.
// Different lists: the longer list is indexed by j, every pair once (weight 1).
template <bool SAME>
struct PairWalk {
  int A0, B0, LA, LB;
  int j, k, t, tend;
  int rs, kstep, kwrap;

  __device__ __forceinline__ void init(int lane, int a0, int l1, int b0, int l2) {
    if (!SAME && l2 > l1) {
      A0 = b0;
      LA = l2;
      B0 = a0;
      LB = l1;
    } else {
      A0 = a0;
      LA = l1;
      B0 = b0;
      LB = l2;
    }
    j = k = rs = kstep = kwrap = 0;
    t = lane;
    tend = 0;
    if (LA <= 0 || LB <= 0) return;
    int d;
    if (SAME) {
      tend = LA * ((LA - 1) / 2 + 1) + (((LA & 1) == 0) ? (LA >> 1) : 0);
      t = lane + LA;
      j = t % LA;
      d = t / LA;
    } else {
      tend = LA * LB;
      j = lane % LA;
      d = lane / LA;
    }
    k = (j + d) % LB;
    const int qs = 64 / LA;
    rs = 64 % LA;
    kstep = (rs + qs) % LB;
    kwrap = ((1 - LA) % LB + LB) % LB;
  }

  __device__ __forceinline__ bool next(int& ej, int& ek) {
    if (t >= tend) return false;
    ej = A0 + j;
    ek = B0 + k;
    t += 64;
    j += rs;
    k += kstep;
    k = (k >= LB) ? k - LB : k;
    if (j >= LA) {
      j -= LA;
      k += kwrap;
      k = (k >= LB) ? k - LB : k;
    }
    return true;
  }
};

// The walk of PairWalk with two offsets per step: lane pairs (j, d) with d = dmin + 2e and d + 1,
// k = (j + d) mod LB, so one step covers two pairs that share j (its loads, and in the adjoint
// its accumulator update) and the 64 lanes still touch 64 consecutive entries.  v1 / v2 say
// whether each pair exists (the last offset of a row may not); an invalid pair's k is still a
// valid entry index.
template <bool SAME>
struct PairWalk2 {
  // absolute entry indices J (list A) and K1 (list B) with their ends; Jhalf: entries below it
  // also take the half offset d = LA / 2 of an even same-list walk
  int J, Jend, Jhalf, K1, Kend, LA, LB;
  int d1, u, uend;
  int rs, dstep, kstep, kwrap, dfull;

  __device__ __forceinline__ void init(int lane, int a0, int l1, int b0, int l2) {
    int A0, B0;
    if (!SAME && l2 > l1) {
      A0 = b0;
      LA = l2;
      B0 = a0;
      LB = l1;
    } else {
      A0 = a0;
      LA = l1;
      B0 = b0;
      LB = l2;
    }
    J = Jend = Jhalf = K1 = Kend = d1 = rs = dstep = kstep = kwrap = dfull = 0;
    u = lane;
    uend = 0;
    if (LA <= 0 || LB <= 0) return;
    int dtot, dhalf = 0;
    if (SAME) {
      dfull = (LA - 1) / 2;
      dhalf = (LA & 1) == 0 ? LA / 2 : 0;
      dtot = dfull + (dhalf ? 1 : 0);
    } else {
      dfull = LB - 1;
      dtot = LB;
    }
    uend = LA * ((dtot + 1) / 2);
    const int j = lane % LA, e = lane / LA;
    d1 = (SAME ? 1 : 0) + 2 * e;
    J = A0 + j;
    Jend = A0 + LA;
    Jhalf = A0 + dhalf;
    K1 = B0 + (j + d1) % LB;
    Kend = B0 + LB;
    const int qs = 64 / LA;
    rs = 64 % LA;
    dstep = 2 * qs;
    kstep = (rs + 2 * qs) % LB;
    kwrap = ((2 - LA) % LB + LB) % LB;
  }

  __device__ __forceinline__ bool next(int& ej, int& ek1, int& ek2, bool& v1, bool& v2) {
    if (u >= uend) return false;
    // the pair (j, d) exists for d <= dlim: d <= dfull, or d = LA / 2 for j < LA / 2
    const int dlim = SAME ? dfull + (J < Jhalf ? 1 : 0) : dfull;
    v1 = d1 <= dlim;
    v2 = d1 < dlim;
    int K2 = K1 + 1;
    K2 = K2 >= Kend ? K2 - LB : K2;
    ej = J;
    ek1 = K1;
    ek2 = K2;
    u += 64;
    J += rs;
    d1 += dstep;
    K1 += kstep;
    K1 = K1 >= Kend ? K1 - LB : K1;
    if (J >= Jend) {
      J -= LA;
      d1 += 2;
      K1 += kwrap;
      K1 = K1 >= Kend ? K1 - LB : K1;
    }
    return true;
  }
};

// x^n, n >= 0 wave-uniform: the common exponents of ACSF sets as straight-line products (the
// switch is on a scalar, so only the taken product costs vector instructions), others by square
// and multiply
template <typename T>
__device__ __forceinline__ T powi(T x, int n) {
  switch (n) {
    case 0: return (T)1;
    case 1: return x;
    case 2: return x * x;
    case 3: return x * x * x;
    case 4: {
      const T y = x * x;
      return y * y;
    }
    case 7: {
      const T y = x * x;
      const T z = y * y;
      return z * y * x;
    }
    case 8: {
      const T y = x * x;
      const T z = y * y;
      return z * z;
    }
    case 15: {
      const T y = x * x;
      const T z = y * y;
      const T u = z * z;
      return u * z * y * x;
    }
    case 16: {
      const T y = x * x;
      const T z = y * y;
      const T u = z * z;
      return u * u;
    }
    default: {
      T r = (T)1;
      while (n) {
        if (n & 1) r *= x;
        x *= x;
        n >>= 1;
      }
      return r;
    }
  }
}

// Members of one angular group (wave-uniform).  Padding members (q >= nmem) have lam = 0,
// xint = 0 and A = 0, so they can be evaluated unconditionally and contribute nothing.
template <typename T, int MQ>
struct Members {
  T lam[MQ], xi[MQ];
  int xint[MQ];
  T pdiag[MQ];  // (1 + lam)^xi: the j == k terms, where cos(theta) = 1
  T A[MQ];      // adjoint: dE/dG_q times 2^(1 - xi_q)
  T AXL[MQ];    // adjoint: A xi lam
};

// Forward: off-diagonal pairs of one group, FP64 partial sums per member in acc.
template <typename T, int MQ, bool SAME, bool G4, bool IPOW>
__device__ __forceinline__ void pairs_fwd(const Slice<T>& sl, int lane, int A0, int LA, int B0,
                                          int LB, const Members<T, MQ>& mb, T rc, T eta, T pi,
                                          T pirc, double (&acc)[MQ]) {
  using V = typename Vec4<T>::type;
  PairWalk<SAME> it;
  it.init(lane, A0, LA, B0, LB);
  int ej, ek;
  while (it.next(ej, ek)) {
    const V dj = sl.d[ej], dk = sl.d[ek];
    const V pj = sl.pa[ej], pk = sl.pa[ek];
    const T c = value_cos(dj, dk, pj.x, pk.x);
    T base = (SAME ? (T)2 : (T)1) * pj.y * pk.y;
    if (G4) {
      const T rx = dj.x - dk.x, ry = dj.y - dk.y, rz = dj.z - dk.z;
      const T r2 = rx * rx + ry * ry + rz * rz;
      const T rjk = fsqrt(r2);
      const T fcjk = (T)0.5 * pj.z * pk.z * (fcos(cutarg(rjk, pi, rc, pirc)) + (T)1);
      const T f = fexp(-eta * r2) * fcjk;
      base = (rjk <= rc) ? base * f : (T)0;
    }
#pragma unroll
    for (int q = 0; q < MQ; ++q) {
      const T P = (T)1 + mb.lam[q] * c;
      const T pw = IPOW ? powi(P, mb.xint[q]) : fpow(P, mb.xi[q]);
      acc[q] += (double)(base * pw);
    }
  }
}

// Per-pair factors of the adjoint: rr = 1 / (r_j r_k), cos(theta), and for G4 the r_jk factor
// H = exp(-eta r_jk^2) fc(r_jk) with Hq = H' / r_jk (H = 1, Hq = 0 for G5).  False for G4 pairs
// whose r_jk lies beyond the cutoff (no contribution).
template <typename T, bool G4>
__device__ __forceinline__ bool adj_geom(const typename Vec4<T>::type& dj,
                                         const typename Vec4<T>::type& dk,
                                         const typename Vec4<T>::type& pj,
                                         const typename Vec4<T>::type& pk, T rc, T eta, T pi,
                                         T pirc, T& rr, T& c, T& H, T& Hq) {
  rr = pj.x * pk.x;
  c = fmin(fmax((dj.x * dk.x + dj.y * dk.y + dj.z * dk.z) * rr, (T)-1), (T)1);
  H = 1;
  Hq = 0;
  if (G4) {
    const T rx = dj.x - dk.x, ry = dj.y - dk.y, rz = dj.z - dk.z;
    const T r2 = rx * rx + ry * ry + rz * rz;
    T rjk, irjk;
    dist_inv(r2, rjk, irjk);
    if (!(rjk <= rc)) return false;
    const T qq = pj.z * pk.z;
    const T x = cutarg(rjk, pi, rc, pirc);
    const T ex = fexp(-eta * r2);
    const T fc = (T)0.5 * qq * (fcos(x) + (T)1);
    const T dfc = -(T)0.5 * qq * pirc * fsin(x);
    H = ex * fc;
    Hq = ex * (dfc - (T)2 * eta * rjk * fc) * irjk;
  }
  return true;
}

// Wave-uniform constants of a group's r_jk factor exp(-eta r^2) fc(r).  FP32: -eta log2(e) and
// pi / (2 pi rc) prepared for v_exp_f32 (2^x) and v_cos_f32 / v_sin_f32 (argument in turns);
// FP64: libm in the reference's operation order.
template <typename T> struct JkConst;
template <> struct JkConst<float> {
  float rc, eta, pirc, el2, prev;
  __device__ __forceinline__ void set(float rc_, float eta_, float pirc_) {
    rc = rc_;
    eta = eta_;
    pirc = pirc_;
    el2 = uniform(-eta_ * 1.44269504088896341f);
    prev = uniform(pirc_ * 0.159154943091895336f);
  }
  __device__ __forceinline__ float ex(float r2) const { return __builtin_amdgcn_exp2f(r2 * el2); }
  __device__ __forceinline__ void cs(float r, float& c, float& sn) const {
    const float x = r * prev;
    c = __builtin_amdgcn_cosf(x);
    sn = __builtin_amdgcn_sinf(x);
  }
};
template <> struct JkConst<double> {
  double rc, eta, pirc;
  __device__ __forceinline__ void set(double rc_, double eta_, double pirc_) {
    rc = rc_;
    eta = eta_;
    pirc = pirc_;
  }
  __device__ __forceinline__ double ex(double r2) const { return exp(-eta * r2); }
  __device__ __forceinline__ void cs(double r, double& c, double& sn) const {
    const double x = kPi * r / rc;
    c = cos(x);
    sn = sin(x);
  }
};

// adj_geom without branches: a pair that does not exist (v false) or lies beyond the cutoff
// gets H = Hq = 0 (and so no contribution) by selection, which also keeps the values of a
// degenerate invalid pair (r_jk = 0) out of the result.
template <typename T, bool G4>
__device__ __forceinline__ void adj_geom_sel(const typename Vec4<T>::type& dj,
                                             const typename Vec4<T>::type& dk,
                                             const typename Vec4<T>::type& pj,
                                             const typename Vec4<T>::type& pk,
                                             const JkConst<T>& K, bool v, T& rr, T& c, T& H,
                                             T& Hq) {
  rr = pj.x * pk.x;
  c = fmin(fmax((dj.x * dk.x + dj.y * dk.y + dj.z * dk.z) * rr, (T)-1), (T)1);
  H = v ? (T)1 : (T)0;
  Hq = 0;
  if (G4) {
    const T rx = dj.x - dk.x, ry = dj.y - dk.y, rz = dj.z - dk.z;
    const T r2 = rx * rx + ry * ry + rz * rz;
    T rjk, irjk;
    dist_inv(r2, rjk, irjk);
    const T qq = pj.z * pk.z;
    T cx, sx;
    K.cs(rjk, cx, sx);
    const T ex = K.ex(r2);
    const T fc = (T)0.5 * qq * (cx + (T)1);
    const T dfc = -(T)0.5 * qq * K.pirc * sx;
    const bool ok = v && (rjk <= K.rc);
    H = ok ? ex * fc : (T)0;
    Hq = ok ? ex * (dfc - (T)2 * K.eta * rjk * fc) * irjk : (T)0;
  }
}

// Adds the gradient of one pair to the per-entry accumulators.  S0 = sum over members of
// dE/dG (1 + lambda c)^xi and S1 = its derivative in c, both times the pair weight.  With the
// per-entry factors E = exp(-eta r^2) fc(r) (pa.y) and E' / r (pa.w), the gradients are
//   g_j = Pj d_j + Q d_k,  g_k = Pk d_k + Q d_j,
//   Pj = S0 H E_k E'_j / r_j - c t / r_j^2 + h,  Q = t / (r_j r_k) - h,
// with t = S1 E_j E_k H and h = S0 E_j E_k H' / r_jk (the derivative of
// S0 * E_j E_k H through c, r_j, r_k and r_jk; gFuncGrad, acsf.F90:1496-1670).
template <typename T, bool G4>
__device__ __forceinline__ void adj_coef(const typename Vec4<T>::type& pj,
                                         const typename Vec4<T>::type& pk, T rr, T c, T H, T Hq,
                                         T S0, T S1, T& Pj, T& Pk, T& Q) {
  const T EE = pj.y * pk.y;
  const T t = S1 * EE * H;
  const T U = S0 * H;
  const T h = G4 ? S0 * EE * Hq : (T)0;
  Q = t * rr - h;
  const T ct = c * t;
  Pj = U * pk.y * pj.w + h - ct * (pj.x * pj.x);
  Pk = U * pj.y * pk.w + h - ct * (pk.x * pk.x);
}

template <typename T, bool G4>
__device__ __forceinline__ void adj_scatter(const Slice<T>& sl, int ej, int ek,
                                            const typename Vec4<T>::type& dj,
                                            const typename Vec4<T>::type& dk,
                                            const typename Vec4<T>::type& pj,
                                            const typename Vec4<T>::type& pk, T rr, T c, T H, T Hq,
                                            T S0, T S1) {
  T Pj, Pk, Q;
  adj_coef<T, G4>(pj, pk, rr, c, H, Hq, S0, S1, Pj, Pk, Q);
  lds_f64* gj = sl.gs + __umul24(ej, 3);
  lds_f64* gk = sl.gs + __umul24(ek, 3);
  lds_add(gj + 0, (double)(Pj * dj.x + Q * dk.x));
  lds_add(gj + 1, (double)(Pj * dj.y + Q * dk.y));
  lds_add(gj + 2, (double)(Pj * dj.z + Q * dk.z));
  lds_add(gk + 0, (double)(Pk * dk.x + Q * dj.x));
  lds_add(gk + 1, (double)(Pk * dk.y + Q * dj.y));
  lds_add(gk + 2, (double)(Pk * dk.z + Q * dj.z));
}

// Adjoint: off-diagonal pairs of one group, gradients accumulated per entry in sl.gs.
template <typename T, int MQ, bool SAME, bool G4, bool IPOW>
__device__ __forceinline__ void pairs_adj(const Slice<T>& sl, int lane, int A0, int LA, int B0,
                                          int LB, const Members<T, MQ>& mb, T rc, T eta, T pi,
                                          T pirc) {
  using V = typename Vec4<T>::type;
  PairWalk<SAME> it;
  it.init(lane, A0, LA, B0, LB);
  const T w = SAME ? (T)2 : (T)1;
  int ej, ek;
  while (it.next(ej, ek)) {
    const V dj = sl.d[ej], dk = sl.d[ek];
    const V pj = sl.pa[ej], pk = sl.pa[ek];
    T rr, c, H, Hq;
    if (!adj_geom<T, G4>(dj, dk, pj, pk, rc, eta, pi, pirc, rr, c, H, Hq)) continue;
    T S0 = 0, S1 = 0;
#pragma unroll
    for (int q = 0; q < MQ; ++q) {
      const T P = (T)1 + mb.lam[q] * c;
      T pw, pm1;
      if (IPOW) {
        const int n = mb.xint[q];
        pm1 = n > 0 ? powi(P, n - 1) : (T)0;
        pw = n > 0 ? pm1 * P : (T)1;
      } else {
        pw = fpow(P, mb.xi[q]);
        pm1 = fpow(P, mb.xi[q] - (T)1);
      }
      S0 += mb.A[q] * pw;
      S1 += mb.AXL[q] * pm1;
    }
    adj_scatter<T, G4>(sl, ej, ek, dj, dk, pj, pk, rr, c, H, Hq, w * S0, w * S1);
  }
}

template <typename T, int MQ, bool SAME, bool G4>
__device__ __forceinline__ void pairs_fwd_p(bool ipow, const Slice<T>& sl, int lane, int A0,
                                            int LA, int B0, int LB, const Members<T, MQ>& mb,
                                            T rc, T eta, T pi, T pirc, double (&acc)[MQ]) {
  if (ipow) pairs_fwd<T, MQ, SAME, G4, true>(sl, lane, A0, LA, B0, LB, mb, rc, eta, pi, pirc, acc);
  else pairs_fwd<T, MQ, SAME, G4, false>(sl, lane, A0, LA, B0, LB, mb, rc, eta, pi, pirc, acc);
}

template <typename T, int MQ, bool SAME, bool G4>
__device__ __forceinline__ void pairs_adj_p(bool ipow, const Slice<T>& sl, int lane, int A0,
                                            int LA, int B0, int LB, const Members<T, MQ>& mb,
                                            T rc, T eta, T pi, T pirc) {
  if (ipow) pairs_adj<T, MQ, SAME, G4, true>(sl, lane, A0, LA, B0, LB, mb, rc, eta, pi, pirc);
  else pairs_adj<T, MQ, SAME, G4, false>(sl, lane, A0, LA, B0, LB, mb, rc, eta, pi, pirc);
}

// ---- exponent-ladder groups ----
// Every member has xi = 2^m (m <= MX) and one of two lambda values; slot (i, m) stands for
// (lambda_i, 2^m).  Per pair and lambda, P = 1 + lambda cos(theta) is squared MX times, which gives
// P^(2^m) for every slot and, as the running product, P^(2^m - 1) for the derivative: the same
// straight-line code for every slot, with no per-member branch on the exponent.

// Forward: off-diagonal pairs, two per step (PairWalk2), FP64 sums per slot; the two pairs'
// terms are added in T before the conversion.
template <typename T, int MX, int MASK, bool SAME, bool G4>
__device__ __forceinline__ void pairs_fwd_lad(const Slice<T>& sl, int lane, int A0, int LA, int B0,
                                              int LB, T lam0, T lam1, const JkConst<T>& K,
                                              T (&acc)[2][MX + 1]) {
  using V = typename Vec4<T>::type;
  PairWalk2<SAME> it;
  it.init(lane, A0, LA, B0, LB);
  int ej, ek1, ek2;
  bool v1, v2;
  // radial-and-r_jk factor of one pair (0 when it does not exist or r_jk > rc) and its cos(theta)
  auto pair_base = [&](const V& dj, const V& pj, const V& dk, const V& pk, bool v, T& c) {
    c = value_cos(dj, dk, pj.x, pk.x);
    T base = (SAME ? (T)2 : (T)1) * pj.y * pk.y;
    bool in = v;
    if (G4) {
      const T rx = dj.x - dk.x, ry = dj.y - dk.y, rz = dj.z - dk.z;
      const T r2 = rx * rx + ry * ry + rz * rz;
      const T rjk = fsqrt(r2);
      T cx, sx;
      K.cs(rjk, cx, sx);
      const T fcjk = (T)0.5 * pj.z * pk.z * (cx + (T)1);
      base *= K.ex(r2) * fcjk;
      in = in && (rjk <= K.rc);
    }
    return in ? base : (T)0;
  };
  // G4: the slot sums go through partial sums that are added to the lane's sums every 16 steps
  // (water boxes: forward 1.13 -> 1.09 ms when the lane's sums were FP64; for the cheaper G5
  // pairs the extra registers cost more, so they add every step)
  constexpr int FLUSH = G4 ? 16 : 1;
  T part[2][MX + 1];
#pragma unroll
  for (int l = 0; l < 2; ++l)
#pragma unroll
    for (int m = 0; m <= MX; ++m) part[l][m] = (T)0;
  auto flush = [&]() {
#pragma unroll
    for (int l = 0; l < 2; ++l)
#pragma unroll
      for (int m = 0; m <= MX; ++m) {
        if (MASK & (1 << m)) acc[l][m] += part[l][m];
        part[l][m] = (T)0;
      }
  };
  int np = 0;
  while (it.next(ej, ek1, ek2, v1, v2)) {
    const V dj = sl.d[ej], pj = sl.pa[ej];
    const V dk1 = sl.d[ek1], pk1 = sl.pa[ek1];
    const V dk2 = sl.d[ek2], pk2 = sl.pa[ek2];
    T c1, c2;
    const T b1 = pair_base(dj, pj, dk1, pk1, v1, c1);
    const T b2 = pair_base(dj, pj, dk2, pk2, v2, c2);
#pragma unroll
    for (int l = 0; l < 2; ++l) {
      const T lam = l ? lam1 : lam0;
      T p1 = (T)1 + lam * c1, p2 = (T)1 + lam * c2;
      if (MASK & 1) part[l][0] += b1 * p1 + b2 * p2;
#pragma unroll
      for (int m = 1; m <= MX; ++m) {
        p1 = p1 * p1;
        p2 = p2 * p2;
        if (MASK & (1 << m)) part[l][m] += b1 * p1 + b2 * p2;
      }
    }
    if (++np == FLUSH) {
      np = 0;
      flush();
    }
  }
  flush();
}

// Adjoint: off-diagonal pairs, two per step (PairWalk2) sharing j, whose gradient is summed in
// registers and added once.  c0 = w dE/dG 2^(1 - xi) of the slot's member (0 for an empty slot,
// w the pair weight), c1 = c0 xi lambda; c1const = the c1 of the xi = 1 slots, whose
// P^(xi - 1) is 1.
template <typename T, int MX, int MASK, bool SAME, bool G4>
__device__ __forceinline__ void pairs_adj_lad(const Slice<T>& sl, int lane, int A0, int LA, int B0,
                                              int LB, T lam0, T lam1, const T (&c0)[2][MX + 1],
                                              const T (&c1)[2][MX + 1], T c1const,
                                              const JkConst<T>& K) {
  using V = typename Vec4<T>::type;
  PairWalk2<SAME> it;
  it.init(lane, A0, LA, B0, LB);
  int ej, ek1, ek2;
  bool v1, v2;
  // one pair: adds its k-side gradient to the accumulator of ek, returns the j-side in g
  auto pair = [&](const V& dj, const V& pj, int ek, bool v, T& gx, T& gy, T& gz) {
    const V dk = sl.d[ek], pk = sl.pa[ek];
    T rr, c, H, Hq;
    adj_geom_sel<T, G4>(dj, dk, pj, pk, K, v, rr, c, H, Hq);
    T S0 = 0, S1 = c1const;
#pragma unroll
    for (int l = 0; l < 2; ++l) {
      T p = (T)1 + (l ? lam1 : lam0) * c;
      T pm = (T)1;
      if (MASK & 1) S0 += c0[l][0] * p;
#pragma unroll
      for (int m = 1; m <= MX; ++m) {
        pm = pm * p;
        p = p * p;
        if (MASK & (1 << m)) {
          S0 += c0[l][m] * p;
          S1 += c1[l][m] * pm;
        }
      }
    }
    T Pj, Pk, Q;
    adj_coef<T, G4>(pj, pk, rr, c, H, Hq, S0, S1, Pj, Pk, Q);
    gx += Pj * dj.x + Q * dk.x;
    gy += Pj * dj.y + Q * dk.y;
    gz += Pj * dj.z + Q * dk.z;
    lds_f64* gk = sl.gs + __umul24(ek, 3);
    lds_add(gk + 0, (double)(Pk * dk.x + Q * dj.x));
    lds_add(gk + 1, (double)(Pk * dk.y + Q * dj.y));
    lds_add(gk + 2, (double)(Pk * dk.z + Q * dj.z));
  };
  while (it.next(ej, ek1, ek2, v1, v2)) {
    const V dj = sl.d[ej], pj = sl.pa[ej];
    T gx = 0, gy = 0, gz = 0;
    pair(dj, pj, ek1, v1, gx, gy, gz);
    pair(dj, pj, ek2, v2, gx, gy, gz);
    lds_f64* gj = sl.gs + __umul24(ej, 3);
    lds_add(gj + 0, (double)gx);
    lds_add(gj + 1, (double)gy);
    lds_add(gj + 2, (double)gz);
  }
}

// One ladder group of one centre: values (forward) or the per-entry adjoint accumulation.  MASK:
// the exponent classes m (xi = 2^m) some member of the group has; the others are not evaluated.
template <typename T, bool ADJ, int MX, int MASK>
__device__ void group_lad(CGroup& grp, const Slice<T>& sl, int lane, long long i, int A0,
                          int L1, int B0, int L2, T rc, T eta, T pi, T pirc, cdouble* cG,
                          double* G, int nG) {
  using V = typename Vec4<T>::type;
  const bool g4 = grp.type == 4;
  const bool same = grp.same != 0;
  T lam0 = (T)grp.lamv[0], lam1 = (T)grp.lamv[1];
  if (ADJ) {
    lam0 = uniform(lam0);
    lam1 = uniform(lam1);
    rc = uniform(rc);
    eta = uniform(eta);
    pirc = uniform(pirc);
  }
  JkConst<T> K;
  K.set(rc, eta, pirc);
  if (!ADJ) {
    // each lane's slot sums in T (FP32 in the FP32 modes: a lane adds a few dozen pair terms on
    // typical rows), the wave's sum of them in FP64
    T acc[2][MX + 1];
#pragma unroll
    for (int l = 0; l < 2; ++l)
#pragma unroll
      for (int m = 0; m <= MX; ++m) acc[l][m] = (T)0;
    if (same) {
      // j == k terms: cos(theta) = 1, r_jk = 0, fc(0) = 1 (acsf.F90:1420-1432)
      for (int e = lane; e < L1; e += 64) {
        const V pj = sl.pa[A0 + e];
        T base = pj.y * pj.y;
        if (g4) base *= pj.z * pj.z;
        const T cjj = diag_cos(sl.d[A0 + e], (T)0);
#pragma unroll
        for (int l = 0; l < 2; ++l) {
          T p = (T)1 + (l ? lam1 : lam0) * cjj;
#pragma unroll
          for (int m = 0; m <= MX; ++m) {
            if (m > 0) p = p * p;
            acc[l][m] += base * p;
          }
        }
      }
      if (g4) pairs_fwd_lad<T, MX, MASK, true, true>(sl, lane, A0, L1, B0, L2, lam0, lam1, K, acc);
      else pairs_fwd_lad<T, MX, MASK, true, false>(sl, lane, A0, L1, B0, L2, lam0, lam1, K, acc);
    } else {
      if (g4) pairs_fwd_lad<T, MX, MASK, false, true>(sl, lane, A0, L1, B0, L2, lam0, lam1, K, acc);
      else pairs_fwd_lad<T, MX, MASK, false, false>(sl, lane, A0, L1, B0, L2, lam0, lam1, K, acc);
    }
    // the 2 (MX + 1) slot sums in one exchange; slot (l, m) has flat index l (MX + 1) + m
    constexpr int NA = pow2_at_least(2 * (MX + 1));
    double v[NA];
#pragma unroll
    for (int k = 0; k < NA; ++k)
      v[k] = k < 2 * (MX + 1) ? (double)acc[k / (MX + 1)][k % (MX + 1)] : 0.0;
    int k;
    const double s = wave_sum_scatter<NA>(v, lane, k);
    if ((lane & (64 / NA - 1)) == 0 && k < 2 * (MX + 1)) {
      const int l = k / (MX + 1), m = k % (MX + 1);
      const int q = grp.smem[5 * l + m];
      if (q >= 0) G[i * nG + grp.col[q]] = s * grp.scale[q];
    }
  } else {
    T c0[2][MX + 1], c1[2][MX + 1];
    T S0d = 0;
    const T w = same ? (T)2 : (T)1;  // unordered pairs of one list count twice
#pragma unroll
    for (int l = 0; l < 2; ++l)
#pragma unroll
      for (int m = 0; m <= MX; ++m) {
        const int q = grp.smem[5 * l + m];
        const T v = q >= 0 ? (T)(cG[grp.col[q]] * grp.scale[q]) : (T)0;
        c0[l][m] = uniform(w * v);
        c1[l][m] = uniform(w * v * (T)(1 << m) * (l ? lam1 : lam0));
        S0d += v * (T)grp.lpdiag[5 * l + m];
      }
    S0d = uniform(S0d);
    const T c1const = uniform(c1[0][0] + c1[1][0]);
    if (same) {
      // j == k terms: only the radial factors depend on r_j; P = 1 + lambda at c = 1
      // (acsf.F90:1580-1585).  Each lane owns its entries here, so the update is plain.
      for (int e = lane; e < L1; e += 64) {
        const int ej = A0 + e;
        const V dj = sl.d[ej];
        const V pj = sl.pa[ej];
        const T H = g4 ? pj.z * pj.z : (T)1;
        const T gsc = S0d * (T)2 * pj.y * pj.w * H;
        sl.gs[3 * ej + 0] += gsc * dj.x;
        sl.gs[3 * ej + 1] += gsc * dj.y;
        sl.gs[3 * ej + 2] += gsc * dj.z;
      }
      if (g4) pairs_adj_lad<T, MX, MASK, true, true>(sl, lane, A0, L1, B0, L2, lam0, lam1, c0, c1, c1const, K);
      else pairs_adj_lad<T, MX, MASK, true, false>(sl, lane, A0, L1, B0, L2, lam0, lam1, c0, c1, c1const, K);
    } else {
      if (g4) pairs_adj_lad<T, MX, MASK, false, true>(sl, lane, A0, L1, B0, L2, lam0, lam1, c0, c1, c1const, K);
      else pairs_adj_lad<T, MX, MASK, false, false>(sl, lane, A0, L1, B0, L2, lam0, lam1, c0, c1, c1const, K);
    }
  }
}

// Values of two G5 ladder groups over the same lists with the same cutoff and lambdas but another
// eta, in one walk: the loads of the entries' geometry, the cosine and the ladder powers are
// shared; each group has its own radial product (pa.y for the first group, pb.y for the partner)
// and its own slot sums, which get exactly the terms, in the same order, that pairs_fwd_lad adds
// for the group alone.
template <typename T, int MX, int MASK, bool SAME>
__device__ __forceinline__ void pairs_fwd_lad2(const Slice<T>& sl, int lane, int A0, int LA,
                                               int B0, int LB, T lam0, T lam1,
                                               T (&acc)[2][2][MX + 1]) {
  using V = typename Vec4<T>::type;
  PairWalk2<SAME> it;
  it.init(lane, A0, LA, B0, LB);
  int ej, ek1, ek2;
  bool v1, v2;
  const T w = SAME ? (T)2 : (T)1;
  while (it.next(ej, ek1, ek2, v1, v2)) {
    const V dj = sl.d[ej], pj = sl.pa[ej];
    const T qj = sl.pb[ej].y;
    const V dk1 = sl.d[ek1], pk1 = sl.pa[ek1];
    const V dk2 = sl.d[ek2], pk2 = sl.pa[ek2];
    const T qk1 = sl.pb[ek1].y, qk2 = sl.pb[ek2].y;
    const T c1 = value_cos(dj, dk1, pj.x, pk1.x);
    const T c2 = value_cos(dj, dk2, pj.x, pk2.x);
    const T a1 = v1 ? w * pj.y * pk1.y : (T)0, a2 = v2 ? w * pj.y * pk2.y : (T)0;
    const T b1 = v1 ? w * qj * qk1 : (T)0, b2 = v2 ? w * qj * qk2 : (T)0;
#pragma unroll
    for (int l = 0; l < 2; ++l) {
      const T lam = l ? lam1 : lam0;
      T p1 = (T)1 + lam * c1, p2 = (T)1 + lam * c2;
      if (MASK & 1) {
        acc[0][l][0] += a1 * p1 + a2 * p2;
        acc[1][l][0] += b1 * p1 + b2 * p2;
      }
#pragma unroll
      for (int m = 1; m <= MX; ++m) {
        p1 = p1 * p1;
        p2 = p2 * p2;
        if (MASK & (1 << m)) {
          acc[0][l][m] += a1 * p1 + a2 * p2;
          acc[1][l][m] += b1 * p1 + b2 * p2;
        }
      }
    }
  }
}

// Values of a G5 ladder group (ga) and its partner (gb) of another eta; sl.pa holds ga's
// per-entry factors, sl.pb gb's.
template <typename T, int MX, int MASK>
__device__ void group_lad_pair_fwd(CGroup& ga, CGroup& gb, const Slice<T>& sl, int lane,
                                   long long i, int A0, int L1, int B0, int L2, double* G,
                                   int nG) {
  using V = typename Vec4<T>::type;
  const T lam0 = (T)ga.lamv[0], lam1 = (T)ga.lamv[1];
  T acc[2][2][MX + 1];  // each lane's slot sums in T, as in group_lad
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int l = 0; l < 2; ++l)
#pragma unroll
      for (int m = 0; m <= MX; ++m) acc[h][l][m] = (T)0;
  if (ga.same) {
    // j == k terms of both groups: cos(theta) = 1 (acsf.F90:1420-1432)
    for (int e = lane; e < L1; e += 64) {
      const T ea = sl.pa[A0 + e].y, eb = sl.pb[A0 + e].y;
      const T ba = ea * ea, bb = eb * eb;
      const T cjj = diag_cos(sl.d[A0 + e], (T)0);
#pragma unroll
      for (int l = 0; l < 2; ++l) {
        T p = (T)1 + (l ? lam1 : lam0) * cjj;
#pragma unroll
        for (int m = 0; m <= MX; ++m) {
          if (m > 0) p = p * p;
          acc[0][l][m] += ba * p;
          acc[1][l][m] += bb * p;
        }
      }
    }
    pairs_fwd_lad2<T, MX, MASK, true>(sl, lane, A0, L1, B0, L2, lam0, lam1, acc);
  } else {
    pairs_fwd_lad2<T, MX, MASK, false>(sl, lane, A0, L1, B0, L2, lam0, lam1, acc);
  }
  // the slot sums of each group in one exchange (group by group: fewer values live at once);
  // slot (l, m) has flat index l (MX + 1) + m
  constexpr int NS = 2 * (MX + 1), NA = pow2_at_least(NS);
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    double v[NA];
#pragma unroll
    for (int k = 0; k < NA; ++k) v[k] = k < NS ? (double)acc[h][k / (MX + 1)][k % (MX + 1)] : 0.0;
    int k;
    const double s = wave_sum_scatter<NA>(v, lane, k);
    if ((lane & (64 / NA - 1)) == 0 && k < NS) {
      const int l = k / (MX + 1), m = k % (MX + 1);
      CGroup& gr = h ? gb : ga;
      const int q = gr.smem[5 * l + m];
      if (q >= 0) G[i * nG + gr.col[q]] = s * gr.scale[q];
    }
  }
}

// Adjoint of two G5 ladder groups over the same lists with the same cutoff, lambdas and exponent
// slots but another eta, in one walk.  Per pair, the cosine, the ladder powers, the loads of the
// shared entry data and the gradient vectors with their accumulator updates are shared; each
// group adds its own S0, S1 and radial factors (pa: the first group's E and E' / r, pb: the
// partner's), so the coefficients of adj_coef become sums over the two groups:
//   t = S1a Ea_j Ea_k + S1b Eb_j Eb_k,  Q = t / (r_j r_k),
//   Pj = S0a Ea_k Ea'_j / r_j + S0b Eb_k Eb'_j / r_j - c t / r_j^2.
template <typename T, int MX, int MASK, bool SAME>
__device__ __forceinline__ void pairs_adj_lad2(const Slice<T>& sl, int lane, int A0, int LA,
                                               int B0, int LB, T lam0, T lam1,
                                               const T (&c0)[2][2][MX + 1],
                                               const T (&c1)[2][2][MX + 1],
                                               const T (&c1const)[2]) {
  using V = typename Vec4<T>::type;
  PairWalk2<SAME> it;
  it.init(lane, A0, LA, B0, LB);
  int ej, ek1, ek2;
  bool v1, v2;
  auto pair = [&](const V& dj, const V& pj, const V& qj, int ek, bool v, T& gx, T& gy, T& gz) {
    const V dk = sl.d[ek], pk = sl.pa[ek], qk = sl.pb[ek];
    const T rr = pj.x * pk.x;
    const T c = fmin(fmax((dj.x * dk.x + dj.y * dk.y + dj.z * dk.z) * rr, (T)-1), (T)1);
    T S0a = 0, S1a = c1const[0], S0b = 0, S1b = c1const[1];
#pragma unroll
    for (int l = 0; l < 2; ++l) {
      T p = (T)1 + (l ? lam1 : lam0) * c;
      T pm = (T)1;
      if (MASK & 1) {
        S0a += c0[0][l][0] * p;
        S0b += c0[1][l][0] * p;
      }
#pragma unroll
      for (int m = 1; m <= MX; ++m) {
        pm = pm * p;
        p = p * p;
        if (MASK & (1 << m)) {
          S0a += c0[0][l][m] * p;
          S1a += c1[0][l][m] * pm;
          S0b += c0[1][l][m] * p;
          S1b += c1[1][l][m] * pm;
        }
      }
    }
    // a pair that does not exist adds nothing
    S0a = v ? S0a : (T)0;
    S1a = v ? S1a : (T)0;
    S0b = v ? S0b : (T)0;
    S1b = v ? S1b : (T)0;
    const T t = S1a * (pj.y * pk.y) + S1b * (qj.y * qk.y);
    const T Q = t * rr;
    const T ct = c * t;
    const T Pj = S0a * pk.y * pj.w + S0b * qk.y * qj.w - ct * (pj.x * pj.x);
    const T Pk = S0a * pj.y * pk.w + S0b * qj.y * qk.w - ct * (pk.x * pk.x);
    gx += Pj * dj.x + Q * dk.x;
    gy += Pj * dj.y + Q * dk.y;
    gz += Pj * dj.z + Q * dk.z;
    lds_f64* gk = sl.gs + __umul24(ek, 3);
    lds_add(gk + 0, (double)(Pk * dk.x + Q * dj.x));
    lds_add(gk + 1, (double)(Pk * dk.y + Q * dj.y));
    lds_add(gk + 2, (double)(Pk * dk.z + Q * dj.z));
  };
  while (it.next(ej, ek1, ek2, v1, v2)) {
    const V dj = sl.d[ej], pj = sl.pa[ej], qj = sl.pb[ej];
    T gx = 0, gy = 0, gz = 0;
    pair(dj, pj, qj, ek1, v1, gx, gy, gz);
    pair(dj, pj, qj, ek2, v2, gx, gy, gz);
    lds_f64* gj = sl.gs + __umul24(ej, 3);
    lds_add(gj + 0, (double)gx);
    lds_add(gj + 1, (double)gy);
    lds_add(gj + 2, (double)gz);
  }
}

// Adjoint of a G5 ladder group (ga) and its partner (gb) of another eta; sl.pa holds ga's
// per-entry factors, sl.pb gb's.  The coefficients are those of group_lad, one set per group.
template <typename T, int MX, int MASK>
__device__ void group_lad_pair(CGroup& ga, CGroup& gb, const Slice<T>& sl,
                               int lane, int A0, int L1, int B0, int L2, cdouble* cG) {
  using V = typename Vec4<T>::type;
  const bool same = ga.same != 0;
  const T lam0 = uniform((T)ga.lamv[0]), lam1 = uniform((T)ga.lamv[1]);
  const T w = same ? (T)2 : (T)1;  // unordered pairs of one list count twice
  T c0[2][2][MX + 1], c1[2][2][MX + 1], c1const[2], S0d[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    CGroup& gr = h ? gb : ga;
    T sd = 0;
#pragma unroll
    for (int l = 0; l < 2; ++l)
#pragma unroll
      for (int m = 0; m <= MX; ++m) {
        const int q = gr.smem[5 * l + m];
        const T v = q >= 0 ? (T)(cG[gr.col[q]] * gr.scale[q]) : (T)0;
        c0[h][l][m] = uniform(w * v);
        c1[h][l][m] = uniform(w * v * (T)(1 << m) * (l ? lam1 : lam0));
        sd += v * (T)gr.lpdiag[5 * l + m];
      }
    S0d[h] = uniform(sd);
    c1const[h] = uniform(c1[h][0][0] + c1[h][1][0]);
  }
  if (same) {
    // j == k terms of both groups (acsf.F90:1580-1585); each lane owns its entries here
    for (int e = lane; e < L1; e += 64) {
      const int ej = A0 + e;
      const V dj = sl.d[ej], pj = sl.pa[ej], qj = sl.pb[ej];
      const T gsc = (T)2 * (S0d[0] * pj.y * pj.w + S0d[1] * qj.y * qj.w);
      sl.gs[3 * ej + 0] += gsc * dj.x;
      sl.gs[3 * ej + 1] += gsc * dj.y;
      sl.gs[3 * ej + 2] += gsc * dj.z;
    }
    pairs_adj_lad2<T, MX, MASK, true>(sl, lane, A0, L1, B0, L2, lam0, lam1, c0, c1, c1const);
  } else {
    pairs_adj_lad2<T, MX, MASK, false>(sl, lane, A0, L1, B0, L2, lam0, lam1, c0, c1, c1const);
  }
}

// One ladder group, or (pg set) a G5 group with its partner.
template <typename T, bool ADJ, int MX, int MASK>
__device__ __forceinline__ void lad_one(CGroup& grp, CGroup* pg,
                                        const Slice<T>& sl, int lane, long long i, int A0, int L1,
                                        int B0, int L2, T rc, T eta, T pi, T pirc,
                                        cdouble* cG, double* G, int nG) {
  if (pg) {
    if constexpr (ADJ) group_lad_pair<T, MX, MASK>(grp, *pg, sl, lane, A0, L1, B0, L2, cG);
    else group_lad_pair_fwd<T, MX, MASK>(grp, *pg, sl, lane, i, A0, L1, B0, L2, G, nG);
    return;
  }
  group_lad<T, ADJ, MX, MASK>(grp, sl, lane, i, A0, L1, B0, L2, rc, eta, pi, pirc, cG, G, nG);
}

// Forward value of one radial function (G1/G2/G3, acsf.F90:1287-1373) summed over the staged
// entries e0, e0 + step, ... below e1, with the terms evaluated in U and summed in FP64.  U is the
// kernel's T, or FP64 in an FP32 kernel whose model standardises some radial feature by a sigma
// far below its mean (R64, DevModel::radial64): the z-score divides G - mean by that sigma, and the FP32
// sums' relative error (~4e-7 over a few dozen neighbours) would reach the network input, and
// through it the forces, amplified by |mean| / sigma.  The FP32 distance is then widened exactly,
// and the cutoff value is recomputed in FP64 instead of read from the FP32 slice.
template <typename U, typename T>
__device__ __forceinline__ double radial_sum(const AcsfArgs& a, const Slice<T>& sl,
                                             const RadialFn& fn, long long i, int e0, int e1,
                                             int step) {
  const bool shared = std::is_same<U, T>::value && fn.rc == a.rclist && fn.atomid == 0;
  const U pi = (U)kPi;
  const U rc = (U)fn.rc, eta = (U)fn.eta, rs = (U)fn.rs, kap = (U)fn.kappa;
  const U pirc = pi / rc;
  const U qi = (U)ext_q(a.ext, a.next, fn.atomid, i);
  double acc = 0.0;
  for (int e = e0; e < e1; e += step) {
    const U r = (U)sl.d[e].w;
    if (r <= rc) {
      U fcv;
      if (shared) {
        fcv = (U)sl.fc[e];
      } else {
        const U qj = (U)ext_q(a.ext, a.next, fn.atomid, sl.j[e]);
        fcv = (U)0.5 * qi * qj * (fcos(cutarg(r, pi, rc, pirc)) + (U)1);
      }
      U term;
      if (fn.type == 1) term = fcv;
      else if (fn.type == 2) term = fexp(-eta * ((r - rs) * (r - rs))) * fcv;
      else term = fcos(kap * r) * fcv;
      acc += (double)term;
    }
  }
  return acc;
}

// LAD: every angular group has the exponent-ladder form (AngGroup::lad > 0); MQ is then unused.
// PAIRS (ladder groups only): paired groups are evaluated together (own instance, so
// that the kernels of models without pairs keep their register budget).  R64 (FP32 forward
// only): radial terms in FP64 (radial_sum), also an instance of its own.
template <typename T, bool ADJ, int MQ, bool LAD, bool PAIRS, bool R64 = false>
__device__ __forceinline__ void acsf_centre(const AcsfArgs& a, const long long i) {
  using V = typename Vec4<T>::type;
  extern __shared__ __align__(32) unsigned char smem[];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int M = a.maxnb;
  Slice<T> sl = make_slice<T, ADJ>(smem + slice_bytes<T, ADJ>(M, PAIRS) * wave, M, PAIRS);

  const long long b = a.cap ? i * a.cap : (long long)a.rowptr[i];
  // a row longer than the slice (a list searched with a capacity from the previous step) is cut
  // to it; the host sees the longest row and recomputes such a step
  const int n = min(a.cap ? a.cnt[i] : (int)(a.rowptr[i + 1] - b), a.maxnb);
  const V* nd = (const V*)a.nd + b;
  // neighbour atom of staged entry e (the adjoint keeps the row index of the entry instead)
  auto nbr = [&](int e) -> long long { return ADJ ? (long long)a.nj[b + sl.j[e]] : sl.j[e]; };
  const T pi = (T)kPi;
  const T rcl = (T)a.rclist;
  const T pircl = pi / rcl;
  const int nZ = a.nZ;  // element slots 0 .. nZ - 1; slot nZ: elements no function names

  // ---- stage the centre's entries in element order with the list-cutoff fc ----
  // (one element slot that every atom takes: the entries keep their row order, and the element
  // codes need not be read)
  auto stage = [&](int pos, int e, int jj, int key) {
    V d = nd[e];
    d.w = entry_r(d.w);  // the list holds the squared distance
    sl.d[pos] = d;
    sl.j[pos] = ADJ ? e : jj;  // adjoint: the entry's index in its row, for the output
    sl.zc[pos] = key < nZ ? key : -1;
    if (!ADJ) sl.fc[pos] = (T)0.5 * (fcos(cutarg(d.w, pi, rcl, pircl)) + (T)1);
  };
  if (a.zone) {
    if (lane < 2) {
      sl.st[lane] = lane ? n : 0;
      sl.nz[lane] = lane ? 0 : n;
    }
    for (int e = lane; e < n; e += 64) stage(e, e, ADJ ? 0 : a.nj[b + e], 0);
  } else {
    int cnt[kMaxZ + 1];
#pragma unroll
    for (int c = 0; c <= kMaxZ; ++c) cnt[c] = 0;
    for (int e0 = 0; e0 < n; e0 += 64) {
      const int e = e0 + lane;
      int key = -1;
      if (e < n) {
        const int zc = a.zcode[a.nj[b + e]];
        key = zc >= 0 ? zc : nZ;
      }
#pragma unroll
      for (int c = 0; c <= kMaxZ; ++c)
        if (c <= nZ) cnt[c] += __popcll(__ballot(key == c));
    }
    int start[kMaxZ + 1];
    {
      int s = 0;
#pragma unroll
      for (int c = 0; c <= kMaxZ; ++c) {
        start[c] = s;
        if (c <= nZ) s += cnt[c];
      }
    }
#pragma unroll
    for (int c = 0; c <= kMaxZ; ++c)
      if (c == lane && c <= nZ) {
        sl.st[c] = start[c];
        sl.nz[c] = cnt[c];
      }
    int placed[kMaxZ + 1];
#pragma unroll
    for (int c = 0; c <= kMaxZ; ++c) placed[c] = 0;
    for (int e0 = 0; e0 < n; e0 += 64) {
      const int e = e0 + lane;
      int key = -1, jj = 0;
      if (e < n) {
        jj = a.nj[b + e];
        const int zc = a.zcode[jj];
        key = zc >= 0 ? zc : nZ;
      }
      int pos = 0;
#pragma unroll
      for (int c = 0; c <= kMaxZ; ++c) {
        if (c <= nZ) {
          const unsigned long long m = __ballot(key == c);
          if (key == c) pos = start[c] + placed[c] + __popcll(m & lanemask_lt(lane));
          placed[c] += __popcll(m);
        }
      }
      if (e < n) stage(pos, e, jj, key);
    }
  }
  wave_sync();

  const int nouts = ADJ ? a.nout : 1;
  for (int t = 0; t < nouts; ++t) {
    cdouble* cG = ADJ ? (cdouble*)a.cG + (i * a.nout + t) * a.nG : nullptr;

    // ---- radial functions ----
    if (!ADJ) {
      // Lanes are (function, share of its element range) pairs.  When the functions fit one
      // wave and some are species-resolved, each gets lanes in proportion to the length of its
      // range (on the water boxes the H functions walk twice as many entries as the O ones:
      // forward 1.07 -> 1.04 ms); else 64 / nf each (the set-up costs ~1 % where the ranges are
      // all the same).
      for (int f0 = 0; f0 < a.nR; f0 += 64) {
        const int nf = min(64, a.nR - f0);
        int nl, fst, fl, gl, ng;  // lanes and first lane of function f0 + lane; a lane's
        bool act;                 // function, share and stride
        if (a.rbal && a.nR <= 64) {
          int L = 0;
          if (lane < nf) {
            const int zc = a.radial[lane].zc;
            L = zc >= 0 ? sl.nz[zc] : n;
          }
          const int W = wave_sum_i(L);
          const int nb = lane < nf ? max(1, (64 * L) / max(W, 1)) : 0;
          nl = wave_sum_i(nb) <= 64 ? nb : (lane < nf ? 64 / nf : 0);
          // every function has at least one lane, so the first lanes fst are distinct and < 64:
          // flag them, and a lane's function is the number of flags up to it
          fst = wave_excl_scan_i(nl, lane);
          const int used = __shfl(fst + nl, nf - 1, 64);
          int* own = reinterpret_cast<int*>(sl.red);
          own[lane] = 0;
          wave_sync();
          if (lane < nf) own[fst] = 1;
          wave_sync();
          const unsigned long long fm = __ballot(own[lane] != 0);
          wave_sync();
          const unsigned long long le = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
          fl = __popcll(fm & le) - 1;
          gl = lane - __shfl(fst, fl, 64);
          ng = __shfl(nl, fl, 64);
          act = lane < used;
        } else {
          const int u = 64 / nf;
          nl = lane < nf ? u : 0;
          fst = lane * u;
          fl = min(lane / u, nf - 1);
          gl = lane - fl * u;
          ng = u;
          act = lane < nf * u;
        }
        const RadialFn& fn = a.radial[f0 + fl];
        double acc = 0.0;
        if (act) {
          // a species-resolved function sums over its element's range of the staged entries
          const int e0 = fn.zc >= 0 ? sl.st[fn.zc] : 0;
          const int e1 = fn.zc >= 0 ? e0 + sl.nz[fn.zc] : n;
          if constexpr (R64) acc = radial_sum<double>(a, sl, fn, i, e0 + gl, e1, ng);
          else acc = radial_sum<T>(a, sl, fn, i, e0 + gl, e1, ng);
        }
        sl.red[lane] = acc;
        wave_sync();
        if (lane < nf) {
          double s = 0.0;
          for (int g = 0; g < nl; ++g) s += sl.red[fst + g];
          a.G[i * a.nG + a.radial[f0 + lane].col] = s;
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
        // list-cutoff fc and fc' of the entry (recomputed rather than kept in LDS)
        const T xl = cutarg(r, pi, rcl, pircl);
        const T fcl = (T)0.5 * (fcos(xl) + (T)1), dfcl = -(T)0.5 * pircl * fsin(xl);
        T s = 0;
        for (int f = 0; f < a.nR; ++f) {
          // the whole record, held in scalar registers from here on (the empty asm takes every
          // field as an operand): its loads issue together instead of one after each branch
          CRadial& fr = ((CRadial*)a.radial)[f];
          RadialFn fn;
          fn.col = fr.col;
          fn.type = fr.type;
          fn.zc = fr.zc;
          fn.atomid = fr.atomid;
          fn.rc = fr.rc;
          fn.eta = fr.eta;
          fn.rs = fr.rs;
          fn.kappa = fr.kappa;
          __asm__ volatile("" ::"s"(fn.col), "s"(fn.type), "s"(fn.zc), "s"(fn.atomid), "s"(fn.rc),
                           "s"(fn.eta), "s"(fn.rs), "s"(fn.kappa));
          const T cf = (T)cG[fn.col];
          if (cf == (T)0) continue;
          if (!(fn.zc < 0 || fn.zc == zc) || !(r <= (T)fn.rc)) continue;
          T fc, dfc;
          if (fn.rc == a.rclist && fn.atomid == 0) {
            fc = fcl;
            dfc = dfcl;
          } else {
            const T rc = (T)fn.rc;
            const T pirc = pi / rc;
            const T qq = (T)(ext_q(a.ext, a.next, fn.atomid, i) *
                             ext_q(a.ext, a.next, fn.atomid, nbr(e)));
            const T x = cutarg(r, pi, rc, pirc);
            fc = (T)0.5 * qq * (fcos(x) + (T)1);
            dfc = -(T)0.5 * qq * pirc * fsin(x);
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
        const T sr = fdiv(s, r);
        sl.gs[3 * e + 0] = sr * d.x;
        sl.gs[3 * e + 1] = sr * d.y;
        sl.gs[3 * e + 2] = sr * d.z;
      }
      wave_sync();
    }

    // ---- angular groups ----
    for (int g = 0; g < a.nAG; ++g) {
      CGroup& grp = ((CGroup*)a.groups)[g];
      const int nm = grp.nmem;
      if constexpr (PAIRS && !ADJ) {
        if (grp.follower) continue;  // evaluated together with its partner
      }
      T A[MQ];
      if (ADJ) {
        // a group whose dE/dG are all zero adds nothing to the forces
        bool any = false;
        if (LAD) {
          if (PAIRS && grp.follower) continue;  // evaluated together with its partner
          for (int q = 0; q < nm; ++q) any = any || (cG[grp.col[q]] != 0.0);
          if (PAIRS && grp.partner >= 0) {
            CGroup& pg = ((CGroup*)a.groups)[grp.partner];
            for (int q = 0; q < pg.nmem; ++q) any = any || (cG[pg.col[q]] != 0.0);
          }
        } else {
#pragma unroll
          for (int q = 0; q < MQ; ++q) {
            A[q] = (q < nm) ? (T)(cG[grp.col[q]] * grp.scale[q]) : (T)0;
            any = any || (A[q] != (T)0);
          }
        }
        if (!any) continue;
      }
      const T rc = (T)grp.rc, eta = (T)grp.eta;
      const T pirc = pi / rc;
      const bool g4 = grp.type == 4;
      const bool same = grp.same != 0;
      const bool shared = (grp.rc == a.rclist) && grp.atomid == 0;
      const int A0 = grp.zc1 >= 0 ? sl.st[grp.zc1] : 0;
      const int L1 = grp.zc1 >= 0 ? sl.nz[grp.zc1] : n;
      const int B0 = grp.zc2 >= 0 ? sl.st[grp.zc2] : 0;
      const int L2 = grp.zc2 >= 0 ? sl.nz[grp.zc2] : n;
      const T qi = (T)ext_q(a.ext, a.next, grp.atomid, i);
      // adjoint of a group evaluated with its partner: the partner's factors go to sl.pb
      CGroup* pg = (PAIRS && grp.partner >= 0) ? &((CGroup*)a.groups)[grp.partner] : nullptr;
      const T eta2 = pg ? (T)pg->eta : (T)0;
      // per-entry radial factor E = exp(-eta r^2) fc(r) of this group and E'(r) / r, over the
      // range the group's two element lists span
      const int ea = min(A0, B0), eb = max(A0 + L1, B0 + L2);
      for (int e0 = ea; e0 < eb; e0 += 64) {
        const int e = e0 + lane;
        if (e >= eb) break;
        const T r = sl.d[e].w;
        T qj = 1, E = 0, Ep = 0, E2 = 0, Ep2 = 0;
        if (r <= rc) {
          T fc, dfc;
          if (shared) {
            if (ADJ) {
              const T x = cutarg(r, pi, rcl, pircl);
              fc = (T)0.5 * (fcos(x) + (T)1);
              dfc = -(T)0.5 * pircl * fsin(x);
            } else {
              fc = sl.fc[e];
              dfc = 0;
            }
          } else {
            qj = (T)ext_q(a.ext, a.next, grp.atomid, nbr(e));
            const T x = cutarg(r, pi, rc, pirc);
            fc = (T)0.5 * qi * qj * (fcos(x) + (T)1);
            dfc = -(T)0.5 * qi * qj * pirc * fsin(x);
          }
          const T ex = fexp(-eta * r * r);
          E = ex * fc;
          if (ADJ) Ep = ex * (dfc - (T)2 * eta * r * fc);
          if (PAIRS && pg) {
            const T ex2 = fexp(-eta2 * r * r);
            E2 = ex2 * fc;
            if (ADJ) Ep2 = ex2 * (dfc - (T)2 * eta2 * r * fc);
          }
        }
        V pa;
        pa.x = frcp(r);
        pa.y = E;
        pa.z = qj;
        pa.w = Ep * pa.x;
        sl.pa[e] = pa;
        if (PAIRS && pg) {
          V pb;
          pb.x = pa.x;
          pb.y = E2;
          pb.z = qj;
          pb.w = Ep2 * pa.x;
          sl.pb[e] = pb;
        }
      }
      wave_sync();

      if constexpr (LAD) {
        switch (grp.lad) {
          // exponent sets {1, 4} and {1, 2, 4, 16} have their own instances, others the full one
          case 1: lad_one<T, ADJ, 1, 0x3>(grp, pg, sl, lane, i, A0, L1, B0, L2, rc, eta, pi, pirc, cG, a.G, a.nG); break;
          case 2:
            if (grp.lmask == 0x5) lad_one<T, ADJ, 2, 0x5>(grp, pg, sl, lane, i, A0, L1, B0, L2, rc, eta, pi, pirc, cG, a.G, a.nG);
            else lad_one<T, ADJ, 2, 0x7>(grp, pg, sl, lane, i, A0, L1, B0, L2, rc, eta, pi, pirc, cG, a.G, a.nG);
            break;
          case 3: lad_one<T, ADJ, 3, 0xf>(grp, pg, sl, lane, i, A0, L1, B0, L2, rc, eta, pi, pirc, cG, a.G, a.nG); break;
          default:
            if (grp.lmask == 0x17) lad_one<T, ADJ, 4, 0x17>(grp, pg, sl, lane, i, A0, L1, B0, L2, rc, eta, pi, pirc, cG, a.G, a.nG);
            else lad_one<T, ADJ, 4, 0x1f>(grp, pg, sl, lane, i, A0, L1, B0, L2, rc, eta, pi, pirc, cG, a.G, a.nG);
            break;
        }
        wave_sync();
        continue;
      } else {
        Members<T, MQ> mb;
        bool ipow = true;
#pragma unroll
        for (int q = 0; q < MQ; ++q) {
          const bool real = q < nm;
          mb.xint[q] = real ? grp.xint[q] : 0;
          mb.xi[q] = real ? (T)grp.xi[q] : (T)0;
          mb.lam[q] = real ? (T)grp.lam[q] : (T)0;
          ipow = ipow && mb.xint[q] >= 0;
          mb.pdiag[q] = real ? (T)grp.pdiag[q] : (T)0;
          mb.A[q] = ADJ ? A[q] : (T)0;
          mb.AXL[q] = ADJ ? A[q] * mb.xi[q] * mb.lam[q] : (T)0;
        }

        if (!ADJ) {
          double acc[MQ];
#pragma unroll
          for (int q = 0; q < MQ; ++q) acc[q] = 0.0;
          if (same) {
            // j == k terms: cos(theta) = 1, r_jk = 0, fc(0) = 1 (acsf.F90:1420-1432)
            for (int e = lane; e < L1; e += 64) {
              const V pj = sl.pa[A0 + e];
              T base = pj.y * pj.y;
              if (g4) base *= pj.z * pj.z;
              const T cjj = diag_cos(sl.d[A0 + e], (T)0);
#pragma unroll
              for (int q = 0; q < MQ; ++q) {
                const T P = (T)1 + mb.lam[q] * cjj;
                const T pw = ipow ? powi(P, mb.xint[q]) : fpow(P, mb.xi[q]);
                acc[q] += (double)(base * (q < nm ? pw : (T)0));
              }
            }
            if (g4) pairs_fwd_p<T, MQ, true, true>(ipow, sl, lane, A0, L1, B0, L2, mb, rc, eta, pi, pirc, acc);
            else pairs_fwd_p<T, MQ, true, false>(ipow, sl, lane, A0, L1, B0, L2, mb, rc, eta, pi, pirc, acc);
          } else {
            if (g4) pairs_fwd_p<T, MQ, false, true>(ipow, sl, lane, A0, L1, B0, L2, mb, rc, eta, pi, pirc, acc);
            else pairs_fwd_p<T, MQ, false, false>(ipow, sl, lane, A0, L1, B0, L2, mb, rc, eta, pi, pirc, acc);
          }
#pragma unroll
          for (int q = 0; q < MQ; ++q) {
            if (q < nm) {
              const double s = wave_sum(acc[q]);
              if (lane == 0) a.G[i * a.nG + grp.col[q]] = s * grp.scale[q];
            }
          }
        } else {
          if (same) {
            // j == k terms: only the radial factors depend on r_j; P = 1 + lambda at c = 1
            // (acsf.F90:1580-1585).  Each lane owns its entries here, so the update is plain.
            T S0d = 0;
#pragma unroll
            for (int q = 0; q < MQ; ++q) S0d += mb.A[q] * mb.pdiag[q];
            for (int e = lane; e < L1; e += 64) {
              const int ej = A0 + e;
              const V dj = sl.d[ej];
              const V pj = sl.pa[ej];
              const T H = g4 ? pj.z * pj.z : (T)1;
              const T gsc = S0d * (T)2 * pj.y * pj.w * H;
              sl.gs[3 * ej + 0] += gsc * dj.x;
              sl.gs[3 * ej + 1] += gsc * dj.y;
              sl.gs[3 * ej + 2] += gsc * dj.z;
            }
            if (g4) pairs_adj_p<T, MQ, true, true>(ipow, sl, lane, A0, L1, B0, L2, mb, rc, eta, pi, pirc);
            else pairs_adj_p<T, MQ, true, false>(ipow, sl, lane, A0, L1, B0, L2, mb, rc, eta, pi, pirc);
          } else {
            if (g4) pairs_adj_p<T, MQ, false, true>(ipow, sl, lane, A0, L1, B0, L2, mb, rc, eta, pi, pirc);
            else pairs_adj_p<T, MQ, false, false>(ipow, sl, lane, A0, L1, B0, L2, mb, rc, eta, pi, pirc);
          }
        }
        wave_sync();
      }
    }

    // ---- output of the adjoint: per-entry gradients and the centre's sum ----
    // The neighbour side (-g of every entry whose neighbour is an atom) is gathered per atom
    // afterwards (k_force_gather), in a fixed order and without atomics: g goes to the slot of
    // the entry's mirror, in the neighbour's own row, so that the gather reads its row
    // contiguously instead of one scattered 24-byte record per entry.  The centre side (+sum of
    // the centre's entries) is written here.
    if (ADJ) {
      double sx = 0, sy = 0, sz = 0;
      for (int e0 = 0; e0 < n; e0 += 64) {
        const int e = e0 + lane;
        if (e >= n) break;
        const double gx = (double)sl.gs[3 * e], gy = (double)sl.gs[3 * e + 1],
                     gz = (double)sl.gs[3 * e + 2];
        const long long own = b + sl.j[e];
        const int me = a.mir[own];
        const long long slot = ((me >= 0 ? (long long)me : own) * a.nout + t) * 3;
        if constexpr (std::is_same<T, double>::value) {
          double* gv = a.gvec + slot;
          gv[0] = gx;
          gv[1] = gy;
          gv[2] = gz;
        } else {
          // FP32 kernels store the entry gradients in FP32 (half the bytes of the scattered
          // stores and of the gather); the centre's own sum below keeps the FP64 accumulators
          float* gv = (float*)a.gvec + slot;
          gv[0] = (float)gx;
          gv[1] = (float)gy;
          gv[2] = (float)gz;
        }
        sx += gx;
        sy += gy;
        sz += gz;
      }
      double v[4] = {sx, sy, sz, 0.0};
      int c;
      const double sc = wave_sum_scatter<4>(v, lane, c);
      if ((lane & 15) == 0 && c < 3) a.F[(i * a.nout + t) * 3 + c] = sc;
      wave_sync();
    }
  }
}

// One wave per centre.  PERSIST: a grid of as many blocks as stay resident takes the centres in
// strides of its wave count, restaging the slice for each (an instance of its own: the loop
// around the body costs the one-centre kernels registers).
template <typename T, bool ADJ, int MQ, bool LAD, bool PAIRS, bool R64 = false,
          bool PERSIST = false>
__device__ __forceinline__ void acsf_wave(const AcsfArgs& a) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int wpb = blockDim.x >> 6;
  long long i = a.a0 + (long long)blockIdx.x * wpb + wave;
  if constexpr (!PERSIST) {
    if (i < a.a1) acsf_centre<T, ADJ, MQ, LAD, PAIRS, R64>(a, i);
  } else {
    const long long stride = (long long)gridDim.x * wpb;
    for (; i < a.a1; i += stride) {
      acsf_centre<T, ADJ, MQ, LAD, PAIRS, R64>(a, i);
      wave_sync();
    }
  }
}

template <typename T, bool ADJ, int MQ, bool LAD, bool PAIRS>
__global__ void __launch_bounds__(512) k_acsf(AcsfArgs a) {
  acsf_wave<T, ADJ, MQ, LAD, PAIRS>(a);
}

// The FP32 exponent-ladder kernels, at least 5 waves per SIMD (the bound once kept them from the
// compiler's 4; with the forward's lane sums in FP32 they need 68-69 registers and run 7).
template <bool ADJ>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(5)))
    k_acsf_lad32(AcsfArgs a) {
  acsf_wave<float, ADJ, 1, true, false>(a);
}

// persistent grid of the same (the loop needs more than 96 registers: 4 waves per SIMD)
template <bool ADJ>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4)))
    k_acsf_lad32_persist(AcsfArgs a) {
  acsf_wave<float, ADJ, 1, true, false, false, true>(a);
}

// FP32 ladder adjoint with paired groups (72 registers: 7 waves per SIMD; the persistent
// instance needs 87)
template <bool PERSIST = false>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4)))
    k_acsf_lad32_pairs(AcsfArgs a) {
  acsf_wave<float, true, 1, true, true, false, PERSIST>(a);
}

// FP32 ladder forward with paired groups, bounded to 64 registers (8 waves per SIMD), which its
// FP32 slot sums per lane meet without spilling (66 unbounded, 7 waves)
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(8)))
    k_acsf_lad32_fpairs(AcsfArgs a) {
  acsf_wave<float, false, 1, true, true>(a);
}

// FP32 forward with the radial terms in FP64 (models with an ill-conditioned radial z-score)
template <int MQ, bool LAD>
__global__ void __launch_bounds__(512) k_acsf_fwd32_r64(AcsfArgs a) {
  acsf_wave<float, false, MQ, LAD, false, true>(a);
}

template <typename T, bool ADJ, int MQ, bool LAD>
hipError_t launch_mq(fnet_batch* b, const DevModel& m, long long a0, long long a1,
                     hipStream_t s) {
  const long long n = a1 - a0;
  if (n <= 0) return hipSuccess;
  const Tune& tn = tune_of(b);
  const int M = slice_entries(b->max_nb);
  const size_t lmax = (size_t)b->ctx->lds_block;
  // paired ladder groups share one walk when the larger slice still fits a block (the forward
  // with an FP64 radial part and the persistent forward keep the unpaired kernels)
  const int pairs = (LAD && m.has_pairs && slice_bytes<T, ADJ>(M, 1) <= lmax &&
                     (ADJ || !(std::is_same<T, float>::value && (m.radial64 || tn.persist))))
                        ? 1 : 0;
  const size_t sb = slice_bytes<T, ADJ>(M, pairs);
  // waves per block: the tuned value, else 1 (adjoint) / 2 (forward) (water boxes: adjoint
  // 4.75 / 4.83 / 5.9 ms with 1 / 2 / 4, forward 2.23 / 2.25 ms with 2 / 4)
  int wpb = ADJ ? (tn.adj_wpb > 0 ? tn.adj_wpb : 1) : (tn.fwd_wpb > 0 ? tn.fwd_wpb : 2);
  // more than one wave per block while their slices fit 64 KiB; a longer row's slice takes a
  // block of its own, up to the device's limit
  while (wpb > 1 && sb * wpb > 64 * 1024) wpb >>= 1;
  if (sb * wpb > lmax) return hipErrorInvalidConfiguration;
  AcsfArgs a;
  a.rowptr = b->rowptr;
  a.cnt = b->cnt;
  a.cap = b->nl_padded ? b->nb_cap : 0;
  a.nj = b->nl_padded ? b->pnj : b->nj;
  a.nd = b->nl_padded ? b->pnd : b->nd;
  a.zcode = b->zcode;
  a.ext = b->ext;
  a.next = b->next;
  a.radial = m.radial;
  a.nR = m.nR;
  a.rbal = tn.rbal >= 0 ? tn.rbal : m.radial_resolved;
  a.groups = m.groups;
  a.nAG = m.nAG;
  a.nZ = m.nZ;
  a.nG = m.nG;
  a.nout = m.nout;
  a.rclist = m.rcmax;
  a.a0 = a0;
  a.a1 = a1;
  a.maxnb = M;
  a.persist = tn.persist ? 1 : 0;
  a.zone = b->zone;
  a.G = (double*)b->G;
  a.cG = (const double*)b->cG;
  a.F = (double*)b->F;
  a.gvec = b->gvec;
  a.mir = b->mir;
  const unsigned grid = (unsigned)((n + wpb - 1) / wpb);
  const size_t lds = sb * wpb;
  // persistent grids exist for the FP32 exponent-ladder kernels (the batched FP32 / BF16 path);
  // the other kernels run one wave per centre
  auto go = [&](void (*kern)(AcsfArgs), bool persistent) -> hipError_t {
    unsigned g = grid;
    if (persistent) {
      // as many blocks as stay resident on the device at once
      int per_cu = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64 * wpb, lds) !=
          hipSuccess)
        per_cu = 1;
      const long long full = (long long)std::max(per_cu, 1) * std::max(b->ctx->n_cu, 1);
      g = (unsigned)std::min<long long>(grid, full);
    }
    hipLaunchKernelGGL(kern, dim3(g), dim3(64 * wpb), lds, s, a);
    return hipGetLastError();
  };
  const bool ps = a.persist != 0;
  if constexpr (ADJ && LAD) {
    if (pairs) {
      if constexpr (std::is_same<T, float>::value)
        return ps ? go(k_acsf_lad32_pairs<true>, true) : go(k_acsf_lad32_pairs<false>, false);
      else return go(k_acsf<T, true, 1, true, true>, false);
    }
  }
  if constexpr (!ADJ && LAD) {
    if (pairs) {
      if constexpr (std::is_same<T, float>::value) return go(k_acsf_lad32_fpairs, false);
      else return go(k_acsf<T, false, 1, true, true>, false);
    }
  }
  if constexpr (!ADJ && std::is_same<T, float>::value) {
    if (m.radial64) return go(k_acsf_fwd32_r64<MQ, LAD>, false);
  }
  if constexpr (LAD && std::is_same<T, float>::value)
    return ps ? go(k_acsf_lad32_persist<ADJ>, true) : go(k_acsf_lad32<ADJ>, false);
  else return go(k_acsf<T, ADJ, MQ, LAD, false>, false);
}

template <typename T, bool ADJ>
hipError_t launch(fnet_batch* b, const DevModel& m, long long a0, long long a1, hipStream_t s) {
  if (m.all_lad) return launch_mq<T, ADJ, 1, true>(b, m, a0, a1, s);
  if (m.maxmem <= 4) return launch_mq<T, ADJ, 4, false>(b, m, a0, a1, s);
  if (m.maxmem <= 8) return launch_mq<T, ADJ, 8, false>(b, m, a0, a1, s);
  if (m.maxmem <= 12) return launch_mq<T, ADJ, 12, false>(b, m, a0, a1, s);
  return launch_mq<T, ADJ, kMaxMembers, false>(b, m, a0, a1, s);
}

}  // namespace

int acsf_slice_entries(int max_nb) { return slice_entries(max_nb); }

size_t lds_bytes_acsf(int max_nb, int nZ, int fp64, int adj) {
  (void)nZ;
  const int M = slice_entries(max_nb);
  if (fp64) return adj ? slice_bytes<double, true>(M) : slice_bytes<double, false>(M);
  return adj ? slice_bytes<float, true>(M) : slice_bytes<float, false>(M);
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
