// Periodic neighbour lists for gfx950: one wavefront per centre atom, candidates tested 64 at a
// time in the reference's entry order and compacted with a ballot prefix count, so the CSR rows
// come out in exactly the order of dynneighlist.F90:233-320 without any sort.
//
// The distance test is the reference's FP64 arithmetic with FMA contraction disabled:
//   neighCoords = coords0(:,j) + latVecs . n          (dynneighlist.F90:291, latpointiter order)
//   dist2 = sum((coordsAtom1 - neighCoords)**2) <= cutoff**2   (dynneighlist.F90:292-294)
// so the set of (centre, neighbour, image) entries is bit-identical to the reference's.
//
// Whole lattice images are skipped when no atom of the structure can reach the centre from
// them: |row_a(inv L) . dr| <= rc |row_a(inv L)| bounds every fractional offset of a neighbour,
// so an image n with n_a outside [s_a(i) - smax_a - w_a, s_a(i) - smin_a + w_a] holds none; and
// an image whose translated Cartesian bounding box of the atoms lies farther than rc from the
// centre holds none either (on cubic cells this drops the corner images of the first test's
// box of images, ~10 % of the candidates on the Si cells).  The exact test still decides every
// candidate that is looked at.
#include <algorithm>

#include "fnet_internal.hpp"

namespace fnet {
namespace {

// q / d for 0 <= q, d < 2^31 through the FP64 reciprocal rd = 1 / d: the estimate is within one of
// the quotient, and one correction step makes it exact (no integer division sequence)
__device__ __forceinline__ int udiv(int q, int d, double rd) {
  int a = (int)((double)q * rd);
  const long long r = (long long)q - (long long)a * d;
  a += r >= d ? 1 : (r < 0 ? -1 : 0);
  return a;
}

__device__ __forceinline__ unsigned long long lanemask_lt(int lane) {
  return (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
}

// Inverse lattice and image ranges, mirroring simplealgebra.F90:41-62,90-125 and
// latpointiter.F90:226-253 (extension 1, dynneighlist.F90:223-224).
__device__ void lattice_setup(const double* L, double rc, StructInfo& si) {
#pragma clang fp contract(off)
#define O(r, c) L[((c) - 1) * 3 + ((r) - 1)]
  double det = O(1, 1) * (O(2, 2) * O(3, 3) - O(3, 2) * O(2, 3));
  det = det - O(1, 2) * (O(2, 1) * O(3, 3) - O(3, 1) * O(2, 3));
  det = det + O(1, 3) * (O(2, 1) * O(3, 2) - O(3, 1) * O(2, 2));
  double* I = si.inv;
#define IV(r, c) I[((c) - 1) * 3 + ((r) - 1)]
  IV(1, 1) = -O(2, 3) * O(3, 2) + O(2, 2) * O(3, 3);
  IV(2, 1) = O(2, 3) * O(3, 1) - O(2, 1) * O(3, 3);
  IV(3, 1) = -O(2, 2) * O(3, 1) + O(2, 1) * O(3, 2);
  IV(1, 2) = O(1, 3) * O(3, 2) - O(1, 2) * O(3, 3);
  IV(2, 2) = -O(1, 3) * O(3, 1) + O(1, 1) * O(3, 3);
  IV(3, 2) = O(1, 2) * O(3, 1) - O(1, 1) * O(3, 2);
  IV(1, 3) = -O(1, 3) * O(2, 2) + O(1, 2) * O(2, 3);
  IV(2, 3) = O(1, 3) * O(2, 1) - O(1, 1) * O(2, 3);
  IV(3, 3) = -O(1, 2) * O(2, 1) + O(1, 1) * O(2, 2);
#undef IV
#undef O
  for (int q = 0; q < 9; ++q) I[q] = I[q] / det;
  for (int a = 0; a < 3; ++a) {
    double s = 0.0;
    for (int c = 0; c < 3; ++c) {
      double v = I[c * 3 + a];
      s = s + v * v;
    }
    double nrm = sqrt(s);
    int t = (int)floor(rc * nrm);
    si.lo[a] = -(t + 1);
    si.hi[a] = t + 1;
    si.w[a] = rc * nrm;
  }
}

// One block per structure: lattice set-up by thread 0, extent of fractional coordinates by all.
// The extent only serves the image pruning of the direct search; structures searched with cell
// lists skip it.
__global__ void __launch_bounds__(256) k_struct_prep(const double* __restrict__ xyz,
                                                     const double* __restrict__ lat,
                                                     const int* __restrict__ periodic,
                                                     const long long* __restrict__ off, double rc,
                                                     const int4* __restrict__ cellinfo,
                                                     StructInfo* __restrict__ sinfo) {
  __shared__ StructInfo si;
  __shared__ double red[2][6][256];
  const int s = blockIdx.x;
  const int tid = threadIdx.x;
  if (tid == 0) {
    si.periodic = periodic[s];
    if (si.periodic) {
      lattice_setup(lat + 9 * s, rc, si);
    } else {
      for (int q = 0; q < 9; ++q) si.inv[q] = 0.0;
      for (int a = 0; a < 3; ++a) {
        si.lo[a] = 0;
        si.hi[a] = -1;
        si.w[a] = 0.0;
      }
    }
  }
  __syncthreads();
  // fractional extent (a = 0..2) and Cartesian bounding box (a = 3..5)
  double mn[6], mx[6];
  for (int a = 0; a < 6; ++a) {
    mn[a] = 1e300;
    mx[a] = -1e300;
  }
  if (si.periodic && !cellinfo[s].w) {
    const long long a0 = off[s], a1 = off[s + 1];
    for (long long i = a0 + tid; i < a1; i += blockDim.x) {
      const double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
      for (int a = 0; a < 3; ++a) {
        double f = si.inv[0 * 3 + a] * x + si.inv[1 * 3 + a] * y + si.inv[2 * 3 + a] * z;
        mn[a] = fmin(mn[a], f);
        mx[a] = fmax(mx[a], f);
      }
      mn[3] = fmin(mn[3], x);
      mx[3] = fmax(mx[3], x);
      mn[4] = fmin(mn[4], y);
      mx[4] = fmax(mx[4], y);
      mn[5] = fmin(mn[5], z);
      mx[5] = fmax(mx[5], z);
    }
  }
  for (int a = 0; a < 6; ++a) {
    red[0][a][tid] = mn[a];
    red[1][a][tid] = mx[a];
  }
  __syncthreads();
  for (int st = blockDim.x / 2; st > 0; st >>= 1) {
    if (tid < st) {
      for (int a = 0; a < 6; ++a) {
        red[0][a][tid] = fmin(red[0][a][tid], red[0][a][tid + st]);
        red[1][a][tid] = fmax(red[1][a][tid], red[1][a][tid + st]);
      }
    }
    __syncthreads();
  }
  if (tid == 0) {
    for (int a = 0; a < 3; ++a) {
      si.smin[a] = red[0][a][0];
      si.smax[a] = red[1][a][0];
      si.bmin[a] = red[0][a + 3][0];
      si.bmax[a] = red[1][a + 3][0];
    }
    sinfo[s] = si;
  }
}

// Candidate test in the reference's arithmetic.  Returns the squared distance through d2.
__device__ __forceinline__ bool near_test(double ci0, double ci1, double ci2, const double* X,
                                          int j, double c0, double c1, double c2, double rc2,
                                          double& nd0, double& nd1, double& nd2, double& d2) {
#pragma clang fp contract(off)
  const double nc0 = X[3 * j] + c0, nc1 = X[3 * j + 1] + c1, nc2 = X[3 * j + 2] + c2;
  const double e0 = ci0 - nc0, e1 = ci1 - nc1, e2 = ci2 - nc2;
  double s = e0 * e0;
  s = s + e1 * e1;
  s = s + e2 * e2;
  nd0 = -e0;
  nd1 = -e1;
  nd2 = -e2;
  d2 = s;
  return s <= rc2;
}

// MODE: kCount writes the row lengths; kFill writes the rows at their CSR offsets; kPadded
// writes row i at i * cap (entries past cap dropped) and its full length, in one pass.
enum { kCount = 0, kFill = 1, kPadded = 2 };

// Bounded to 80 registers (6 waves per SIMD instead of the compiler's 82 and 5), which it meets
// without spilling: 8192 Si cells 2.74 -> 2.55 ms; at 64 registers (8 waves) it spills and takes
// 3.38 ms.
template <int MODE, typename V>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6)))
    k_neigh(const double* __restrict__ xyz, const long long* __restrict__ off,
            const int* __restrict__ atom_struct, const StructInfo* __restrict__ sinfo,
            const double* __restrict__ lat, double rc, long long a0, long long a1,
            int* __restrict__ cnt, const int* __restrict__ rowptr,
            int* __restrict__ nj, int* __restrict__ nimg, V* __restrict__ nd,
            const int4* __restrict__ cellinfo, int cap, int* __restrict__ mir) {
  // nimg receives each entry's key (image + 1) n + (j - i) mod n; padded rows also set each
  // entry's mirror slot to -1 (no mirror) when mir is given, for the mirror search to fill in
#pragma clang fp contract(off)
  constexpr bool WRITE = MODE != kCount;
  const int lane = threadIdx.x & 63;
  const long long i = a0 + ((long long)blockIdx.x * blockDim.x + threadIdx.x) / 64;
  if (i >= a1) return;
  const int s = atom_struct[i];
  if (cellinfo[s].w) return;  // structure handled by the cell-list kernels
  const long long o = off[s];
  const int n = (int)(off[s + 1] - o);
  const int li = (int)(i - o);
  const double* X = xyz + 3 * o;
  const double ci0 = X[3 * li], ci1 = X[3 * li + 1], ci2 = X[3 * li + 2];
  const double rc2 = rc * rc;
  const long long base = MODE == kFill ? rowptr[i] : (MODE == kPadded ? i * cap : 0);
  int count = 0;

  // key of an entry: (image + 1) n + (j - i) mod n, which increases along the row (the image
  // index is key / n - 1); the force gather's mirror search reads it
  auto emit = [&](bool hit, int j, int key, double nd0, double nd1, double nd2, double d2) {
    const unsigned long long mask = __ballot(hit);
    const int q = count + __popcll(mask & lanemask_lt(lane));
    if (WRITE && hit && (MODE != kPadded || q < cap)) {
      const long long pos = base + q;
      nj[pos] = (int)(o + j);
      nimg[pos] = key;
      V v;
      v.x = nd0;
      v.y = nd1;
      v.z = nd2;
      entry_d2(v, d2);
      nd[pos] = v;
      if (MODE == kPadded && mir) mir[pos] = -1;
    }
    count += __popcll(mask);
  };

  // central cell: j = i+1 .. n-1, 0 .. i-1
  for (int t0 = 1; t0 < n; t0 += 64) {
    const int t = t0 + lane;
    int j = li + t;
    if (j >= n) j -= n;
    double nd0 = 0, nd1 = 0, nd2 = 0, d2 = 0;
    bool hit = false;
    if (t < n) hit = near_test(ci0, ci1, ci2, X, j, 0.0, 0.0, 0.0, rc2, nd0, nd1, nd2, d2);
    emit(hit, j, t, nd0, nd1, nd2, d2);
  }
  const StructInfo& si = sinfo[s];
  if (si.periodic) {
    // Lattice images in the reference's lexicographic order (third index fastest, origin
    // skipped), 64 at a time: the ones that can hold a neighbour are compacted in order into an
    // LDS table with their translation, then the (image, atom) pairs of the table are tested as
    // one list, 64 per step, so small cells (n << 64) still fill the wave.
    __shared__ double4 tab[4][64];  // per wave: translation L . n (x, y, z), image index (w)
    double4* tw = tab[(threadIdx.x >> 6) & 3];
    const double* L = lat + 9 * s;
    double sf[3];
    for (int a = 0; a < 3; ++a)
      sf[a] = si.inv[0 * 3 + a] * ci0 + si.inv[1 * 3 + a] * ci1 + si.inv[2 * 3 + a] * ci2;
    double wlo[3], whi[3];
    for (int a = 0; a < 3; ++a) {
      const double m = 1e-8 * (1.0 + fabs(sf[a]) + fabs(si.smin[a]) + fabs(si.smax[a]) + si.w[a]);
      wlo[a] = sf[a] - si.smax[a] - si.w[a] - m;
      whi[a] = sf[a] - si.smin[a] + si.w[a] + m;
    }
    const int m1 = si.hi[0] - si.lo[0] + 1, m2 = si.hi[1] - si.lo[1] + 1,
              m3 = si.hi[2] - si.lo[2] + 1;
    const int nq = m1 * m2 * m3;
    const int m23 = m2 * m3;
    const double rm23 = 1.0 / (double)m23, rm3 = 1.0 / (double)m3;
    const int qorig = ((0 - si.lo[0]) * m2 + (0 - si.lo[1])) * m3 + (0 - si.lo[2]);
    const int qs = 64 / n, rs = 64 % n;
    for (int q0 = 0; q0 < nq; q0 += 64) {
      const int q = q0 + lane;
      bool keep = false;
      double4 tv = make_double4(0.0, 0.0, 0.0, 0.0);
      if (q < nq && q != qorig) {
        // q = ((p1 - lo0) m2 + (p2 - lo1)) m3 + (p3 - lo2)
        const int a1 = udiv(q, m23, rm23), rem = q - a1 * m23, a2 = udiv(rem, m3, rm3);
        const int p1 = si.lo[0] + a1, p2 = si.lo[1] + a2, p3 = si.lo[2] + (rem - a2 * m3);
        keep = !(p1 < wlo[0] || p1 > whi[0] || p2 < wlo[1] || p2 > whi[1] || p3 < wlo[2] ||
                 p3 > whi[2]);
        const double pv0 = (double)p1, pv1 = (double)p2, pv2 = (double)p3;
        double cv[3];
        for (int r = 0; r < 3; ++r) {
          double acc = 0.0;
          acc = acc + L[0 * 3 + r] * pv0;
          acc = acc + L[1 * 3 + r] * pv1;
          acc = acc + L[2 * 3 + r] * pv2;
          cv[r] = acc;
        }
        // distance from the centre to the image's box of atoms; the margin covers the
        // rounding of this estimate against the exact test's
        const double ci[3] = {ci0, ci1, ci2};
        double q2 = 0.0;
        for (int r = 0; r < 3; ++r) {
          const double g = fmax(0.0, fmax(si.bmin[r] + cv[r] - ci[r], ci[r] - si.bmax[r] - cv[r]));
          q2 += g * g;
        }
        keep = keep && q2 <= rc2 * (1.0 + 1e-9) + 1e-12;
        tv = make_double4(cv[0], cv[1], cv[2], (double)(q > qorig ? q - 1 : q));
      }
      const unsigned long long km = __ballot(keep);
      if (keep) tw[__popcll(km & lanemask_lt(lane))] = tv;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      // flat index u = slot n + t over the kept images, t = 0 .. n - 1 from the centre on
      const int tot = __popcll(km) * n;
      int slot = lane / n, t = lane % n;
      for (int u0 = 0; u0 < tot; u0 += 64) {
        double nd0 = 0, nd1 = 0, nd2 = 0, d2 = 0;
        bool hit = false;
        int j = 0, img = 0;
        if (u0 + lane < tot) {
          const double4 c = tw[slot];
          j = li + t;
          if (j >= n) j -= n;
          img = (int)c.w;
          hit = near_test(ci0, ci1, ci2, X, j, c.x, c.y, c.z, rc2, nd0, nd1, nd2, d2);
        }
        emit(hit, j, (img + 1) * n + t, nd0, nd1, nd2, d2);
        slot += qs;
        t += rs;
        if (t >= n) {
          t -= n;
          slot += 1;
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
  }
  if (MODE != kFill && lane == 0) cnt[i] = count;
}

// Copies the padded rows (cap entries per atom) to their CSR offsets.
template <typename V>
__global__ void __launch_bounds__(256)
    k_compact_rows(long long n, int cap, const int* __restrict__ cnt,
                   const int* __restrict__ rowptr, const int* __restrict__ pnj,
                   const int* __restrict__ pnimg, const V* __restrict__ pnd,
                   int* __restrict__ nj, int* __restrict__ nimg, V* __restrict__ nd) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long i = t / cap;
  const int e = (int)(t - i * cap);
  if (i >= n || e >= cnt[i]) return;
  const long long src = t, dst = (long long)rowptr[i] + e;
  nj[dst] = pnj[src];
  nimg[dst] = pnimg[src];
  nd[dst] = pnd[src];
}

// ---- exclusive scan of the per-atom counts -------------------------------------------------
constexpr int kScanBlock = 1024;

__global__ void __launch_bounds__(kScanBlock) k_scan_block(const int* __restrict__ in,
                                                          int* __restrict__ out, long long n,
                                                          int* __restrict__ totals,
                                                          int* __restrict__ bmax) {
  __shared__ int buf[kScanBlock];
  __shared__ int wmax[kScanBlock / 64];
  const long long g = (long long)blockIdx.x * kScanBlock + threadIdx.x;
  const int v = (g < n) ? in[g] : 0;
  if (bmax) {
    int m = v;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) m = max(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
  }
  buf[threadIdx.x] = v;
  __syncthreads();
  if (bmax && threadIdx.x == 0) {
    int m = wmax[0];
    for (int w = 1; w < kScanBlock / 64; ++w) m = max(m, wmax[w]);
    bmax[blockIdx.x] = m;
  }
  for (int st = 1; st < kScanBlock; st <<= 1) {
    int add = (threadIdx.x >= st) ? buf[threadIdx.x - st] : 0;
    __syncthreads();
    buf[threadIdx.x] += add;
    __syncthreads();
  }
  if (g < n) out[g] = buf[threadIdx.x] - v;
  if (threadIdx.x == kScanBlock - 1) totals[blockIdx.x] = buf[threadIdx.x];
}

__global__ void __launch_bounds__(kScanBlock) k_scan_totals(int* __restrict__ totals,
                                                           long long nb, int* __restrict__ last,
                                                           const int* __restrict__ bmax,
                                                           int* __restrict__ maxout) {
  __shared__ int buf[kScanBlock];
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  if (maxout && threadIdx.x == 0) {
    int m = 0;
    for (long long b = 0; b < nb; ++b) m = max(m, bmax[b]);
    *maxout = m;
  }
  __syncthreads();
  for (long long b0 = 0; b0 < nb; b0 += kScanBlock) {
    const long long g = b0 + threadIdx.x;
    const int v = (g < nb) ? totals[g] : 0;
    buf[threadIdx.x] = v;
    __syncthreads();
    for (int st = 1; st < kScanBlock; st <<= 1) {
      int add = (threadIdx.x >= st) ? buf[threadIdx.x - st] : 0;
      __syncthreads();
      buf[threadIdx.x] += add;
      __syncthreads();
    }
    if (g < nb) totals[g] = carry + buf[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == 0) carry += buf[kScanBlock - 1];
    __syncthreads();
  }
  if (threadIdx.x == 0) *last = carry;
}

__global__ void k_scan_add(int* __restrict__ out, long long n, const int* __restrict__ totals) {
  const long long g = (long long)blockIdx.x * kScanBlock + threadIdx.x;
  if (g < n) out[g] += totals[blockIdx.x];
}

// Longest row of the list: block maxima combined with atomicMax into *out (zeroed before).
__global__ void __launch_bounds__(256) k_row_max(const int* __restrict__ cnt, long long n,
                                                 int* __restrict__ out) {
  __shared__ int wm[4];
  int m = 0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    m = max(m, cnt[i]);
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) m = max(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) atomicMax(out, max(max(wm[0], wm[1]), max(wm[2], wm[3])));
}

}  // namespace

hipError_t launch_row_max(const int* cnt, long long n, int* out, hipStream_t s) {
  FNET_HIP_TRY(launch_zero_int(out, 1, s));
  if (n <= 0) return hipSuccess;
  const long long blocks = std::min<long long>(1024, (n + 255) / 256);
  k_row_max<<<(unsigned)blocks, 256, 0, s>>>(cnt, n, out);
  return hipGetLastError();
}

// scratch of launch_scan: block totals (nb + 1) and block maxima (nb)
long long scan_tmp_size(long long n) { return 2 * ((n + kScanBlock - 1) / kScanBlock) + 1; }

// Exclusive scan of cnt[0..n) into rowptr[0..n] (rowptr[n] = total); with maxout, also the
// largest count (the reduction rides on the scan's own pass over the counts).
hipError_t launch_scan(int* cnt, int* rowptr, long long n, int* tmp, hipStream_t s,
                       int* maxout) {
  const long long nb = (n + kScanBlock - 1) / kScanBlock;
  if (nb == 0) {
    FNET_HIP_TRY(hipMemsetAsync(rowptr, 0, sizeof(int), s));
    if (maxout) FNET_HIP_TRY(hipMemsetAsync(maxout, 0, sizeof(int), s));
    return hipSuccess;
  }
  int* bmax = maxout ? tmp + nb + 1 : nullptr;
  k_scan_block<<<(unsigned)nb, kScanBlock, 0, s>>>(cnt, rowptr, n, tmp, bmax);
  k_scan_totals<<<1, kScanBlock, 0, s>>>(tmp, nb, rowptr + n, bmax, maxout);
  k_scan_add<<<(unsigned)nb, kScanBlock, 0, s>>>(rowptr, n, tmp);
  return hipGetLastError();
}

hipError_t launch_struct_prep(fnet_batch* b, double rc, hipStream_t s) {
  if (b->S == 0) return hipSuccess;
  k_struct_prep<<<b->S, 256, 0, s>>>(b->xyz, b->lat, b->periodic, b->off, rc, b->cellinfo,
                                     b->sinfo);
  return hipGetLastError();
}

hipError_t launch_neigh_count(fnet_batch* b, double rc, long long a0, long long a1,
                              hipStream_t s) {
  const long long n = a1 - a0;
  if (n <= 0) return hipSuccess;
  const unsigned grid = (unsigned)((n + 3) / 4);
  k_neigh<kCount, double4><<<grid, 256, 0, s>>>(b->xyz, b->off, b->atom_struct, b->sinfo,
                                                 b->lat, rc, a0, a1, b->cnt, nullptr, nullptr,
                                                 nullptr, nullptr, b->cellinfo, 0, nullptr);
  return hipGetLastError();
}

hipError_t launch_neigh_fill(fnet_batch* b, double rc, long long a0, long long a1, int fp64,
                             hipStream_t s) {
  const long long n = a1 - a0;
  if (n <= 0) return hipSuccess;
  const unsigned grid = (unsigned)((n + 3) / 4);
  if (fp64) {
    k_neigh<kFill, double4><<<grid, 256, 0, s>>>(b->xyz, b->off, b->atom_struct, b->sinfo,
                                                 b->lat, rc, a0, a1, nullptr, b->rowptr, b->nj,
                                                 b->nimg, (double4*)b->nd, b->cellinfo, 0,
                                                 nullptr);
  } else {
    k_neigh<kFill, float4><<<grid, 256, 0, s>>>(b->xyz, b->off, b->atom_struct, b->sinfo,
                                                b->lat, rc, a0, a1, nullptr, b->rowptr, b->nj,
                                                b->nimg, (float4*)b->nd, b->cellinfo, 0,
                                                nullptr);
  }
  return hipGetLastError();
}

hipError_t launch_neigh_pad(fnet_batch* b, double rc, int fp64, hipStream_t s, int* mir) {
  const long long n = b->N;
  if (n <= 0) return hipSuccess;
  const unsigned grid = (unsigned)((n + 3) / 4);
  if (fp64) {
    k_neigh<kPadded, double4><<<grid, 256, 0, s>>>(b->xyz, b->off, b->atom_struct, b->sinfo,
                                                   b->lat, rc, 0, n, b->cnt, nullptr, b->pnj,
                                                   b->pnimg, (double4*)b->pnd, b->cellinfo,
                                                   b->nb_cap, mir);
  } else {
    k_neigh<kPadded, float4><<<grid, 256, 0, s>>>(b->xyz, b->off, b->atom_struct, b->sinfo,
                                                  b->lat, rc, 0, n, b->cnt, nullptr, b->pnj,
                                                  b->pnimg, (float4*)b->pnd, b->cellinfo,
                                                  b->nb_cap, mir);
  }
  return hipGetLastError();
}

hipError_t launch_compact_rows(fnet_batch* b, int fp64, hipStream_t s) {
  const long long tot = (long long)b->N * b->nb_cap;
  if (tot <= 0) return hipSuccess;
  const unsigned grid = (unsigned)((tot + 255) / 256);
  if (fp64) {
    k_compact_rows<double4><<<grid, 256, 0, s>>>(b->N, b->nb_cap, b->cnt, b->rowptr, b->pnj,
                                                 b->pnimg, (const double4*)b->pnd, b->nj,
                                                 b->nimg, (double4*)b->nd);
  } else {
    k_compact_rows<float4><<<grid, 256, 0, s>>>(b->N, b->nb_cap, b->cnt, b->rowptr, b->pnj,
                                                b->pnimg, (const float4*)b->pnd, b->nj, b->nimg,
                                                (float4*)b->nd);
  }
  return hipGetLastError();
}

}  // namespace fnet
