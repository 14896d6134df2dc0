// Cell-list neighbour search for large periodic structures, with the reference's entry order.
//
// Atoms are binned by their wrapped fractional coordinates into nb_a >= 3 bins per axis, each at
// least rc wide perpendicular to its faces (1/nb_a >= rc |row_a(inv L)|), so every neighbour
// image of a centre lies in one of the 27 bins around it and each (neighbour, image) pair is met
// once.  The lattice image of a candidate follows from the bins: n = m_i - m_j + sigma, with m the
// integer part of the fractional coordinates and sigma the periodic wrap of the stencil bin.
// Candidates whose image lies outside the reference's image ranges (latpointiter.F90:226-253) are
// dropped, as the reference never visits them; the rest pass the reference's FP64 distance test
// (dynneighlist.F90:291-294).  The hits of one centre are then sorted in LDS by (lattice image in
// the reference's lexicographic order, (j - i) mod n), which is the order in which
// TNeighIterator_getNextNeighbours (dynneighlist.F90:233-320) emits them.
#include "fnet_internal.hpp"

namespace fnet {
namespace {

__device__ __forceinline__ unsigned long long lanemask_lt(int lane) {
  return (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ void k_cell_bin(const double* __restrict__ xyz, const int* __restrict__ atom_struct,
                           const StructInfo* __restrict__ sinfo, const int4* __restrict__ cellinfo,
                           const int* __restrict__ cell_off, long long a0, long long a1,
                           int* __restrict__ atom_cell, int4* __restrict__ atom_m,
                           int* __restrict__ cell_count) {
  const long long i = a0 + (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a1) return;
  const int s = atom_struct[i];
  const int4 ci = cellinfo[s];
  if (!ci.w) return;
  const StructInfo& si = sinfo[s];
  const double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
  const int nb[3] = {ci.x, ci.y, ci.z};
  int m[3], bn[3];
  for (int a = 0; a < 3; ++a) {
    const double f = si.inv[0 * 3 + a] * x + si.inv[1 * 3 + a] * y + si.inv[2 * 3 + a] * z;
    const double fl = floor(f);
    m[a] = (int)fl;
    int b = (int)((f - fl) * nb[a]);
    bn[a] = b < 0 ? 0 : (b >= nb[a] ? nb[a] - 1 : b);
  }
  const int local = (bn[0] * nb[1] + bn[1]) * nb[2] + bn[2];
  const int cell = cell_off[s] + local;
  atom_cell[i] = cell;
  atom_m[i] = make_int4(m[0], m[1], m[2], local);
  atomicAdd(&cell_count[cell], 1);
}

__global__ void k_cell_scatter(const int* __restrict__ atom_struct,
                               const int4* __restrict__ cellinfo, long long a0, long long a1,
                               const int* __restrict__ atom_cell, const int* __restrict__ cell_start,
                               int* __restrict__ cell_cursor, int* __restrict__ cell_atoms) {
  const long long i = a0 + (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a1) return;
  if (!cellinfo[atom_struct[i]].w) return;
  const int c = atom_cell[i];
  cell_atoms[cell_start[c] + atomicAdd(&cell_cursor[c], 1)] = (int)i;
}

struct CellArgs {
  const double* xyz;
  const long long* off;
  const int* atom_struct;
  const StructInfo* sinfo;
  const int4* cellinfo;
  const int* cell_off;
  const int* cell_start;
  const int* cell_atoms;
  const int4* atom_m;
  const double* lat;
  double rc;
  long long a0, a1;
  int K;  // sort capacity per wave (power of two >= largest row, or >= cap)
  int cap;  // padded rows: entries per atom
  int* cnt;
  const int* rowptr;
  int* nj;
  int* nimg;
  void* nd;
  int* mir;  // padded rows: each entry's mirror slot set to -1 when given
};

// MODE: 0 writes the row lengths; 1 writes the sorted rows at their CSR offsets; 2 writes row i
// at i * cap (entries past cap dropped) and its full length, in one pass.
template <int MODE, typename V>
__global__ void __launch_bounds__(256) k_neigh_cells(CellArgs a) {
#pragma clang fp contract(off)
  constexpr bool WRITE = MODE != 0;
  extern __shared__ __align__(32) unsigned char smem[];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const long long i = a.a0 + ((long long)blockIdx.x * blockDim.x + threadIdx.x) / 64;
  if (i >= a.a1) return;
  const int s = a.atom_struct[i];
  const int4 ci = a.cellinfo[s];
  if (!ci.w) return;
  const int K = a.K;
  // LDS slice of this wave (fill pass only)
  unsigned char* base = smem + (size_t)wave * K * (sizeof(unsigned long long) + 3 * sizeof(int) + sizeof(V));
  unsigned long long* key = (unsigned long long*)base;
  V* pnd = (V*)(key + K);
  int* slot = (int*)(pnd + K);
  int* pj = slot + K;
  int* pimg = pj + K;

  const StructInfo& si = a.sinfo[s];
  const long long o = a.off[s];
  const int n = (int)(a.off[s + 1] - o);
  const int li = (int)(i - o);
  const double* X = a.xyz + 3 * o;
  const double* L = a.lat + 9 * s;
  const double ci0 = X[3 * li], ci1 = X[3 * li + 1], ci2 = X[3 * li + 2];
  const double rc2 = a.rc * a.rc;
  const int4 mi = a.atom_m[i];
  const int nb0 = ci.x, nb1 = ci.y, nb2 = ci.z;
  const int b0 = mi.w / (nb1 * nb2), b1 = (mi.w / nb2) % nb1, b2 = mi.w % nb2;
  const int N1 = si.hi[1] - si.lo[1] + 1, N2 = si.hi[2] - si.lo[2] + 1;
  const int origin = ((0 - si.lo[0]) * N1 + (0 - si.lo[1])) * N2 + (0 - si.lo[2]);
  // The 27 stencil bins, one per lane 0..26: first candidate in cell_atoms, count, and the
  // periodic wrap (s0, s1, s2) packed as 9 (s0 + 1) + 3 (s1 + 1) + (s2 + 1).  Their candidates are
  // then taken 64 at a time as one list (a bin holds ~10 atoms, a fraction of a wave); the order
  // does not matter, the fill pass sorts each row.
  int bc0 = 0, bcc = 0, bsg = 0;
  if (lane < 27) {
    const int d0 = lane / 9 - 1, d1 = (lane / 3) % 3 - 1, d2 = lane % 3 - 1;
    int B0 = b0 + d0, B1 = b1 + d1, B2 = b2 + d2;
    const int s0 = B0 < 0 ? -1 : (B0 >= nb0 ? 1 : 0);
    const int s1 = B1 < 0 ? -1 : (B1 >= nb1 ? 1 : 0);
    const int s2 = B2 < 0 ? -1 : (B2 >= nb2 ? 1 : 0);
    B0 -= s0 * nb0;
    B1 -= s1 * nb1;
    B2 -= s2 * nb2;
    const int cell = a.cell_off[s] + (B0 * nb1 + B1) * nb2 + B2;
    bc0 = a.cell_start[cell];
    bcc = a.cell_start[cell + 1] - bc0;
    bsg = (s0 + 1) * 9 + (s1 + 1) * 3 + (s2 + 1);
  }
  int incl = bcc;
#pragma unroll
  for (int d = 1; d < 32; d <<= 1) {
    const int t = __shfl_up(incl, d, 64);
    if (lane >= d) incl += t;
  }
  const int bex = incl - bcc;  // first flat index of the lane's bin
  const int total = __shfl(incl, 26, 64);
  int count = 0;
  for (int p0 = 0; p0 < total; p0 += 64) {
    const int p = p0 + lane;
    // bin of candidate p: the last bin whose first flat index is <= p (empty bins share their
    // successor's index and are passed over)
    int bb = 0;
#pragma unroll
    for (int st = 16; st > 0; st >>= 1) {
      const int nbb = bb + st;
      const int e = __shfl(bex, nbb < 27 ? nbb : 26, 64);
      if (nbb < 27 && e <= p) bb = nbb;
    }
    const int c0 = __shfl(bc0, bb, 64), e0b = __shfl(bex, bb, 64), sg = __shfl(bsg, bb, 64);
    const int s0 = sg / 9 - 1, s1 = (sg / 3) % 3 - 1, s2 = sg % 3 - 1;
    bool hit = false;
    int lj = 0, img = -1;
    double nd0 = 0, nd1 = 0, nd2 = 0, dd = 0;
    unsigned long long k = 0;
    if (p < total) {
      const int j = a.cell_atoms[c0 + (p - e0b)];
      lj = (int)(j - o);
      const int4 mj = a.atom_m[j];
      const int n0 = mi.x - mj.x + s0, n1 = mi.y - mj.y + s1, n2 = mi.z - mj.z + s2;
      const bool inr = n0 >= si.lo[0] && n0 <= si.hi[0] && n1 >= si.lo[1] &&
                       n1 <= si.hi[1] && n2 >= si.lo[2] && n2 <= si.hi[2];
      const bool central = (n0 == 0 && n1 == 0 && n2 == 0);
      if (inr && !(central && lj == li)) {
        const double pv0 = (double)n0, pv1 = (double)n1, pv2 = (double)n2;
        double cv[3];
        for (int r = 0; r < 3; ++r) {
          double acc = 0.0;
          acc = acc + L[0 * 3 + r] * pv0;
          acc = acc + L[1 * 3 + r] * pv1;
          acc = acc + L[2 * 3 + r] * pv2;
          cv[r] = acc;
        }
        const double nc0 = X[3 * lj] + cv[0], nc1 = X[3 * lj + 1] + cv[1],
                     nc2 = X[3 * lj + 2] + cv[2];
        const double e0 = ci0 - nc0, e1 = ci1 - nc1, e2 = ci2 - nc2;
        double sq = e0 * e0;
        sq = sq + e1 * e1;
        sq = sq + e2 * e2;
        hit = sq <= rc2;
        nd0 = -e0;
        nd1 = -e1;
        nd2 = -e2;
        dd = sq;
        int lex = -1;
        if (!central) {
          lex = ((n0 - si.lo[0]) * N1 + (n1 - si.lo[1])) * N2 + (n2 - si.lo[2]);
          if (lex > origin) lex -= 1;
        }
        img = lex;
        int t = lj - li;
        if (t < 0) t += n;
        k = (unsigned long long)(lex + 1) * (unsigned long long)n + (unsigned long long)t;
      }
    }
    const unsigned long long mask = __ballot(hit);
    if (WRITE && hit) {
      const int q = count + __popcll(mask & lanemask_lt(lane));
      if (q < K) {
        key[q] = k;
        slot[q] = q;
        pj[q] = (int)(o + lj);
        V v;
        v.x = nd0;
        v.y = nd1;
        v.z = nd2;
        entry_d2(v, dd);
        pnd[q] = v;
      }
    }
    count += __popcll(mask);
  }
  if (!WRITE) {
    if (lane == 0) a.cnt[i] = count;
    return;
  }
  const int m = count < K ? count : K;
  for (int e = m + lane; e < K; e += 64) {
    key[e] = ~0ull;
    slot[e] = e;
  }
  wave_sync();
  // bitonic sort of (key, slot) over K entries
  for (int kk = 2; kk <= K; kk <<= 1) {
    for (int jj = kk >> 1; jj > 0; jj >>= 1) {
      for (int e = lane; e < K; e += 64) {
        const int x = e ^ jj;
        if (x > e) {
          const bool up = (e & kk) == 0;
          const unsigned long long ke = key[e], kx = key[x];
          if ((ke > kx) == up) {
            key[e] = kx;
            key[x] = ke;
            const int t = slot[e];
            slot[e] = slot[x];
            slot[x] = t;
          }
        }
      }
      wave_sync();
    }
  }
  const long long rb = MODE == 2 ? i * a.cap : (long long)a.rowptr[i];
  const int mw = MODE == 2 ? (m < a.cap ? m : a.cap) : m;
  if (MODE == 2 && lane == 0) a.cnt[i] = count;
  V* out = (V*)a.nd;
  for (int e = lane; e < mw; e += 64) {
    const int q = slot[e];
    a.nj[rb + e] = pj[q];
    a.nimg[rb + e] = (int)key[e];  // the entry's key (image + 1) n + (j - i) mod n
    out[rb + e] = pnd[q];
    if (MODE == 2 && a.mir) a.mir[rb + e] = -1;
  }
}

template <int MODE, typename V>
hipError_t launch_cells_t(fnet_batch* b, double rc, long long a0, long long a1, hipStream_t s,
                          int* mir = nullptr) {
  constexpr bool WRITE = MODE != 0;
  const long long n = a1 - a0;
  if (n <= 0) return hipSuccess;
  CellArgs a;
  a.xyz = b->xyz;
  a.off = b->off;
  a.atom_struct = b->atom_struct;
  a.sinfo = b->sinfo;
  a.cellinfo = b->cellinfo;
  a.cell_off = b->cell_off;
  a.cell_start = b->cell_start;
  a.cell_atoms = b->cell_atoms;
  a.atom_m = b->atom_m;
  a.lat = b->lat;
  a.rc = rc;
  a.a0 = a0;
  a.a1 = a1;
  const int K = cells_sort_capacity(MODE == 2 ? b->nb_cap : b->max_nb);
  a.K = K;
  a.cap = b->nb_cap;
  a.cnt = b->cnt;
  a.rowptr = b->rowptr;
  a.nj = MODE == 2 ? b->pnj : b->nj;
  a.nimg = MODE == 2 ? b->pnimg : b->nimg;
  a.nd = MODE == 2 ? b->pnd : b->nd;
  a.mir = MODE == 2 ? mir : nullptr;
  size_t lds = 0;
  int wpb = 4;
  if (WRITE) {
    const size_t per = (size_t)K * (sizeof(unsigned long long) + 3 * sizeof(int) + sizeof(V));
    while (wpb > 1 && per * wpb > 64 * 1024) wpb >>= 1;
    if (per * wpb > (size_t)b->ctx->lds_block) return hipErrorInvalidConfiguration;
    lds = per * wpb;
  }
  const unsigned grid = (unsigned)((n + wpb - 1) / wpb);
  k_neigh_cells<MODE, V><<<grid, 64 * wpb, lds, s>>>(a);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_cell_bin(fnet_batch* b, hipStream_t s) {
  if (b->ncells == 0 || b->N == 0) return hipSuccess;
  // fill kernels rather than memsets: this sequence is also captured into a HIP graph
  FNET_HIP_TRY(launch_zero_int(b->cell_count, (int)b->ncells, s));
  const unsigned grid = (unsigned)((b->N + 255) / 256);
  k_cell_bin<<<grid, 256, 0, s>>>(b->xyz, b->atom_struct, b->sinfo, b->cellinfo, b->cell_off, 0,
                                  b->N, b->atom_cell, b->atom_m, b->cell_count);
  FNET_HIP_TRY(hipGetLastError());
  FNET_HIP_TRY(launch_scan(b->cell_count, b->cell_start, b->ncells, b->cell_scan_tmp, s, nullptr));
  FNET_HIP_TRY(launch_zero_int(b->cell_count, (int)b->ncells, s));
  k_cell_scatter<<<grid, 256, 0, s>>>(b->atom_struct, b->cellinfo, 0, b->N, b->atom_cell,
                                      b->cell_start, b->cell_count, b->cell_atoms);
  return hipGetLastError();
}

int cells_sort_capacity(int need) {
  int K = 64;
  while (K < need) K <<= 1;
  return K;
}

size_t cells_lds_bytes(int need, int fp64) {
  const size_t v = fp64 ? sizeof(double4) : sizeof(float4);
  return (size_t)cells_sort_capacity(need) * (sizeof(unsigned long long) + 3 * sizeof(int) + v);
}

hipError_t launch_cells_count(fnet_batch* b, double rc, long long a0, long long a1,
                              hipStream_t s) {
  if (b->ncells == 0) return hipSuccess;
  return launch_cells_t<0, double4>(b, rc, a0, a1, s);
}

hipError_t launch_cells_fill(fnet_batch* b, double rc, long long a0, long long a1, int fp64,
                             hipStream_t s) {
  if (b->ncells == 0) return hipSuccess;
  return fp64 ? launch_cells_t<1, double4>(b, rc, a0, a1, s)
              : launch_cells_t<1, float4>(b, rc, a0, a1, s);
}

hipError_t launch_cells_pad(fnet_batch* b, double rc, int fp64, hipStream_t s, int* mir) {
  if (b->ncells == 0) return hipSuccess;
  return fp64 ? launch_cells_t<2, double4>(b, rc, 0, b->N, s, mir)
              : launch_cells_t<2, float4>(b, rc, 0, b->N, s, mir);
}

}  // namespace fnet
