// cil_hip.hip — MI355X (gfx950) implementation of the overdamped colony Euler step behind
// the C ABI of include/cil_hip.h.
//
// Step (one Euler(ts) of CIL.cu:1013-1020), kernels in launch order:
//   k_update     per cell: lazy intra-cell force (CIL.cu:30-100) + sorted pair force ->
//                fused overdamped update (CIL.cu:902-939); when no division is configured
//                it also emits the particle's bin key and arrival rank (counting sort pass 1)
//   [division]   k_div_* flag / prefix-sum / append in parent order (CIL.cu:469-619)
//   [k_keys]     bin key + rank, when division moved particles
//   [k_noise]    Wiener deviates per cell (CIL.cu:793-818), reference noise mode only
//   k_scan_*     exclusive prefix sum of bin counts
//   k_scatter    particle -> slot in bin order
//   k_binsort    one wavefront per bucket (element type x spatial bin): bitonic sort of the
//                bucket's particle ids (deterministic order), gather of bin-local coordinates
//   k_pair<T>    one thread per sorted particle: full-shell gather over the per-type-pair
//                stencil, walked as contiguous z-runs; canonical (lower-UID) orientation per
//                pair so asymmetric terms match the reference
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "kernels.hpp"

using namespace cil;

#define WAVE 64
#define GUARD_SLOTS 4

// Neighbour-list mode gates the sort on the device rebuild flag (S.rebuild is null outside
// list mode): steps that keep the list skip the whole re-sort with no host round trip.
__device__ inline bool sort_skipped(const Dev& S) { return S.rebuild != nullptr && *rebuild_flag(S, step_ts(S)) == 0u; }

// Gate of the bucket scan: closed when list mode skips this pass's sort (same test as
// sort_skipped, for kernels that do not take the Dev).
struct SortGate {
  const unsigned* flags;  // null: always open
  const int* ts;
  int off;
};
__device__ inline bool gate_closed(const SortGate& g) { return g.flags && g.flags[(*g.ts + g.off) & 1] == 0u; }

#define NUM_XCD 8
// Workgroups are dealt to the 8 XCDs round-robin (workgroup b runs on XCD b % 8) and each XCD
// has its own L2.  Kernels whose consecutive workgroups work on neighbouring bins renumber
// them so that each XCD processes one contiguous range: a bin's neighbour data is then
// fetched into one L2 instead of eight.  A bijection of [0, nblk).
__device__ inline int xcd_block(int b, int nblk) {
  const int q = nblk / NUM_XCD, r = nblk % NUM_XCD, x = b % NUM_XCD;
  return x * q + (x < r ? x : r) + b / NUM_XCD;
}
// The same idea where work per workgroup is uneven along the index (type-major buckets put
// all of one element type first): the grid is cut into 8*XCD_CHUNKS contiguous chunks dealt
// round-robin to the XCDs, so every XCD gets a share of each region while a chunk's
// workgroups still share one L2.  The tail beyond whole chunks keeps its own numbering.
#define XCD_CHUNKS 16
__device__ inline int xcd_block_chunked(int b, int nblk) {
  const int C = nblk / (NUM_XCD * XCD_CHUNKS);
  if (C == 0 || b >= C * NUM_XCD * XCD_CHUNKS) return b;
  const int t = b / NUM_XCD, x = b % NUM_XCD;
  return ((t / C) * NUM_XCD + x) * C + t % C;
}

// ======================================================================================
// Bin keys
// ======================================================================================
__device__ inline void k_keys_body(const StepParams* __restrict__ Pp, const Dev& S, int bid) {
  const StepParams& P = *Pp;
  if (sort_skipped(S)) return;
  int p = bid * blockDim.x + threadIdx.x;
  if (p >= S.P) return;
  int s, c;
  int cslot = p / P.nE;
  if (!cell_active(P, S.ncells, cslot, &s, &c)) {
    S.key[p] = -1;
    return;
  }
  double x[3];
  bool finite = true;
  for (int d = 0; d < P.box.dim; ++d) {
    x[d] = S.pos[(size_t)d * S.P + p];
    finite = finite && isfinite(x[d]);
  }
  if (!finite) {
    atomicOr(S.err, (unsigned)ERR_NONFINITE);
    for (int d = 0; d < P.box.dim; ++d) x[d] = 0.0;
  }
  int b = bucket_key(P, p, s, x);
  S.key[p] = b;
  S.rank[p] = atomicAdd(&S.count[b], 1);
}
__global__ void k_keys(const StepParams* __restrict__ Pp, Dev S) { k_keys_body(Pp, S, (int)blockIdx.x); }

// ======================================================================================
// Exclusive prefix sum (2 small kernels): 4096 items per block of 256 threads.
// ======================================================================================
#define SCAN_ITEMS 16
#define SCAN_BLOCK 256
#define SCAN_TILE (SCAN_ITEMS * SCAN_BLOCK)

__device__ inline int block_exclusive_scan(int v, int* lds, int* total) {
  int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
  int x = v;
#pragma unroll
  for (int o = 1; o < WAVE; o <<= 1) {
    int y = __shfl_up(x, o, WAVE);
    if (lane >= o) x += y;
  }
  if (lane == WAVE - 1) lds[wid] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int k = 0; k < (int)(blockDim.x / WAVE); ++k) {
      int t = lds[k];
      lds[k] = acc;
      acc += t;
    }
    lds[WAVE] = acc;
  }
  __syncthreads();
  int res = x - v + lds[wid];
  *total = lds[WAVE];
  __syncthreads();
  return res;
}

__device__ inline void k_scan_local_body(const int* __restrict__ in, int* __restrict__ out, int* __restrict__ bsum, int n, int bid) {
  __shared__ int lds[WAVE + 1];
  int base = bid * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
  int v[SCAN_ITEMS];
  int sum = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    v[k] = (base + k < n) ? in[base + k] : 0;
    sum += v[k];
  }
  int total;
  int ex = block_exclusive_scan(sum, lds, &total);
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    if (base + k < n) out[base + k] = ex;
    ex += v[k];
  }
  if (threadIdx.x == 0) bsum[bid] = total;
}
__global__ void k_scan_local(const int* __restrict__ in, int* __restrict__ out, int* __restrict__ bsum, int n,
                             SortGate gate) {
  if (gate_closed(gate)) return;
  k_scan_local_body(in, out, bsum, n, (int)blockIdx.x);
}

// Sum of v over the workgroup (every thread gets it).
__device__ inline int block_sum(int v, int* lds) {
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
  if ((threadIdx.x & (WAVE - 1)) == 0) lds[threadIdx.x / WAVE] = v;
  __syncthreads();
  int s = 0;
  for (int k = 0; k < (int)(blockDim.x / WAVE); ++k) s += lds[k];
  return s;
}

// Second pass: each 256-item block lies inside one scan tile t and adds the sum of the tile
// totals before t (at most a few hundred, summed by the block itself); the block holding the
// last item writes the grand total.  With n == 0 block 0 writes the total (0).
__device__ inline void k_scan_add_body(int* __restrict__ out, const int* __restrict__ bsum, int n,
                                       int* __restrict__ total_out, int bid) {
  __shared__ int lds[WAVE];
  const int i0 = bid * (int)blockDim.x;
  const int t = i0 / SCAN_TILE;
  int acc = 0;
  for (int q = (int)threadIdx.x; q < t; q += (int)blockDim.x) acc += bsum[q];
  acc = block_sum(acc, lds);
  const int i = i0 + (int)threadIdx.x;
  if (i < n) out[i] += acc;
  if (total_out && (n > 0 ? i == n - 1 : (bid == 0 && threadIdx.x == 0))) *total_out = acc + bsum[t];
}
__global__ void __launch_bounds__(256) k_scan_add(int* __restrict__ out, const int* __restrict__ bsum, int n,
                                                  int* __restrict__ total_out, SortGate gate) {
  if (gate_closed(gate)) return;
  k_scan_add_body(out, bsum, n, total_out, (int)blockIdx.x);
}

// ======================================================================================
// Scatter and per-bin sort + gather
// ======================================================================================
__device__ inline void k_scatter_body(const Dev& S, int bid) {
  if (sort_skipped(S)) return;
  int p = bid * blockDim.x + threadIdx.x;
  if (p >= S.P) return;
  int b = S.key[p];
  if (b < 0) return;
  S.idx[S.start[b] + S.rank[p]] = p;
}
__global__ void k_scatter(Dev S) { k_scatter_body(S, (int)blockIdx.x); }

// Orders this wave's LDS accesses across lanes (wave-scope fence + wave barrier).
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

#define BINSORT_MAX 1024
#define BINSORT_WAVES 4
#define BINSORT_GRID 2048  // workgroups of the single-colony bin sort (8 per CU)

// Coordinates of linear bin index b (x-major, z fastest in 3D).
__device__ inline void bin_coords(const StepParams& P, int b, int* bc) {
  if (P.box.dim == 3) {
    bc[2] = b % P.ns[2];
    int t = b / P.ns[2];
    bc[1] = t % P.ns[1];
    bc[0] = t / P.ns[1];
  } else {
    bc[1] = b % P.ns[1];
    bc[0] = b / P.ns[1];
    bc[2] = 0;
  }
}

// Packed bin coordinates of a sorted slot (sbin): 3D x << 21 | y << 10 | z (x, y < 2048,
// z < 1024), 2D x << 16 | y (x, y < 65536); limits checked when the bins are laid out.
__device__ inline int pack_bin(const StepParams& P, const int* bc) {
  return P.box.dim == 3 ? (int)(((unsigned)bc[0] << 21) | ((unsigned)bc[1] << 10) | (unsigned)bc[2])
                        : (int)(((unsigned)bc[0] << 16) | (unsigned)bc[1]);
}
__device__ inline void unpack_bin(const StepParams& P, int s, int* bc) {
  const unsigned u = (unsigned)s;
  if (P.box.dim == 3) {
    bc[0] = (int)(u >> 21);
    bc[1] = (int)((u >> 10) & 2047u);
    bc[2] = (int)(u & 1023u);
  } else {
    bc[0] = (int)(u >> 16);
    bc[1] = (int)(u & 0xffffu);
    bc[2] = 0;
  }
}

template <typename T>
__device__ inline void write_sorted(const StepParams& P, const Dev& S, const int* bc, int k, int p);

template <>
__device__ inline void write_sorted<float>(const StepParams& P, const Dev& S, const int* bc, int k, int p) {
  float4 r;
  r.x = (float)(S.pos[p] - bc[0] * P.w[0]);
  r.y = (float)(S.pos[(size_t)S.P + p] - bc[1] * P.w[1]);
  r.z = P.box.dim == 3 ? (float)(S.pos[2 * (size_t)S.P + p] - bc[2] * P.w[2]) : 0.0f;
  r.w = (float)S.sigma[p];
  ((float4*)S.spos)[k] = r;
}
template <>
__device__ inline void write_sorted<double>(const StepParams& P, const Dev& S, const int* bc, int k, int p) {
  (void)bc;
  double4 r;
  r.x = S.pos[p];
  r.y = S.pos[(size_t)S.P + p];
  r.z = P.box.dim == 3 ? S.pos[2 * (size_t)S.P + p] : 0.0;
  r.w = S.sigma[p];
  ((double4*)S.spos)[k] = r;
}

// Outputs of sorted slot k holding particle p of the bin with coordinates bc (packed:
// pb): index maps, metadata and the sorted coordinates.
template <typename T>
__device__ inline void emit_sorted(const StepParams& P, const Dev& S, const int* bc, int pb, int k, int p) {
  S.idx[k] = p;
  const int cslot = p / P.nE, e = p - cslot * P.nE;
  const int sp = species_of_cell(P, cslot);
  const int t = sp * P.nE + e;
  S.smeta[k] = p * 8 + t;
  S.sbin[k] = pb;
  S.inv[p] = k;
  if (S.sigma[p] > P.sigma_bound[t]) atomicOr(S.err, (unsigned)ERR_CUTOFF);
  write_sorted<T>(P, S, bc, k, p);
}

// A bucket of more than BINSORT_SMALL particles, sorted by the whole wave in registers
// (bitonic across lanes) or in its LDS region L.
template <typename T>
__device__ inline void binsort_bucket_wave(const StepParams& P, const Dev& S, int u, int lane, int* L) {
  const int s0 = S.start[u], n = S.start[u + 1] - s0;
  int bc[3];
  bin_coords(P, u % P.nbins, bc);
  const int pb = pack_bin(P, bc);
  if (n <= WAVE) {
    int v = lane < n ? S.idx[s0 + lane] : 0x7fffffff;
#pragma unroll
    for (int k = 2; k <= WAVE; k <<= 1) {
#pragma unroll
      for (int j = k >> 1; j > 0; j >>= 1) {
        int o = __shfl_xor(v, j, WAVE);
        bool up = ((lane & k) == 0);
        bool lower = ((lane & j) == 0);
        int mn = v < o ? v : o, mx = v < o ? o : v;
        v = (lower == up) ? mn : mx;
      }
    }
    if (lane < n) emit_sorted<T>(P, S, bc, pb, s0 + lane, v);
    return;
  }
  if (n > BINSORT_MAX) {
    if (lane == 0) atomicOr(S.err, (unsigned)ERR_BIN_OVERFLOW);
    return;
  }
  // One wave sorts its own LDS region; only wave-local ordering is needed between stages
  // (no workgroup barrier: the other waves of the block take other paths).
  int np2 = 1;
  while (np2 < n) np2 <<= 1;
  wave_lds_sync();  // the region's previous bucket is fully read
  for (int i = lane; i < np2; i += WAVE) L[i] = i < n ? S.idx[s0 + i] : 0x7fffffff;
  wave_lds_sync();
  for (int k = 2; k <= np2; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = lane; i < np2; i += WAVE) {
        int ixj = i ^ j;
        if (ixj > i) {
          int a = L[i], c = L[ixj];
          bool up = ((i & k) == 0);
          if ((a > c) == up) {
            L[i] = c;
            L[ixj] = a;
          }
        }
      }
      wave_lds_sync();
    }
  }
  wave_lds_sync();
  for (int i = lane; i < n; i += WAVE) emit_sorted<T>(P, S, bc, pb, s0 + i, L[i]);
}

// Compare-exchange of Batcher's odd-even merge sort network for 8 keys (19 comparators).
#define BINSORT_SMALL 8
__device__ inline void sort8(int* v) {
  constexpr int net[19][2] = {{0, 1}, {2, 3}, {4, 5}, {6, 7}, {0, 2}, {1, 3}, {4, 6}, {5, 7}, {1, 2}, {5, 6},
                              {0, 4}, {1, 5}, {2, 6}, {3, 7}, {2, 4}, {3, 5}, {1, 2}, {3, 4}, {5, 6}};
#pragma unroll
  for (int c = 0; c < 19; ++c) {
    const int a = v[net[c][0]], d = v[net[c][1]];
    v[net[c][0]] = a < d ? a : d;
    v[net[c][1]] = a < d ? d : a;
  }
}

// Wave w of nw takes a contiguous run of 64-bucket groups.  In a group, lane l sorts bucket
// base + l by itself when it holds at most BINSORT_SMALL particles (the common case: a bin
// holds one or two); larger buckets are then sorted one at a time by the whole wave.  Every
// bucket ends in ascending particle index, so the order is the same on every run.
template <typename T>
__device__ inline void k_binsort_body(const StepParams* __restrict__ Pp, const Dev& S, int bid, int nblk) {
  const StepParams& P = *Pp;
  if (sort_skipped(S)) return;
  __shared__ int buf[BINSORT_WAVES][BINSORT_MAX];
  const int wid = threadIdx.x / WAVE, lane = threadIdx.x & (WAVE - 1);
  const int groups = (P.nbuckets + WAVE - 1) / WAVE;
  const int nw = nblk * BINSORT_WAVES, w = bid * BINSORT_WAVES + wid;
  const int per = (groups + nw - 1) / nw;
  const int g1 = (w + 1) * per < groups ? (w + 1) * per : groups;
  for (int g = w * per; g < g1; ++g) {
    const int u = g * WAVE + lane;
    int n = 0, s0 = 0;
    if (u < P.nbuckets) {
      s0 = S.start[u];
      n = S.start[u + 1] - s0;
      S.count[u] = 0;  // reset for the next counting pass
    }
    if (n > 0 && n <= BINSORT_SMALL) {
      int v[BINSORT_SMALL];
#pragma unroll
      for (int i = 0; i < BINSORT_SMALL; ++i) v[i] = i < n ? S.idx[s0 + i] : 0x7fffffff;
      sort8(v);
      int bc[3];
      bin_coords(P, u % P.nbins, bc);
      const int pb = pack_bin(P, bc);
#pragma unroll
      for (int i = 0; i < BINSORT_SMALL; ++i)
        if (i < n) emit_sorted<T>(P, S, bc, pb, s0 + i, v[i]);
    }
    unsigned long long big = __ballot(n > BINSORT_SMALL);
    while (big) {
      const int l = __ffsll(big) - 1;
      big &= big - 1;
      binsort_bucket_wave<T>(P, S, __shfl(u, l, WAVE), lane, buf[wid]);
    }
  }
}
template <typename T>
__global__ void __launch_bounds__(WAVE* BINSORT_WAVES) k_binsort(const StepParams* __restrict__ Pp, Dev S) {
  k_binsort_body<T>(Pp, S, xcd_block(blockIdx.x, gridDim.x), gridDim.x);
}

// ======================================================================================
// Persistent bin sort (neighbour-list mode): keys, bucket scan, scatter and per-bucket sort in
// one launch, over the same bodies as the separate kernels.  It works only on rebuild passes
// and otherwise returns at once, so a step that keeps its list pays one launch.
// The grid is one 256-thread workgroup per CU, all resident (a CU holds several once the
// stream's previous kernel has drained), and the phases are separated by grid barriers.  A
// barrier's wait is bounded: after SYNC_SPIN_LIMIT polls it records ERR_SYNC and the kernel
// runs to its end without waiting; the host then reports the error and resets the counters.
// ======================================================================================
#ifndef SORT_ALL_CU_PER_WG
#define SORT_ALL_CU_PER_WG 1  // CUs per workgroup of k_sort_all (each workgroup passes a grid barrier with one fence pair)
#endif
#define SYNC_SPIN_LIMIT (1u << 18)  // about 0.3 s of polling; a phase takes well under 1 ms
// Barrier counters, each on its own 128-byte line: arrivals of the workgroups b with
// b % NUM_XCD == x at line x, arrivals of whole groups at line NUM_XCD; departures likewise
// from line NUM_XCD + 1.  Arrivals on one word serialise, so each workgroup adds to its group's
// word and only the last of a group adds to the shared one: NUM_XCD arrivals per phase there
// instead of one per workgroup.  All words are zero between launches.
constexpr int SYNC_LINE = 32;
constexpr int SYNC_WORDS = 2 * (NUM_XCD + 1) * SYNC_LINE;
__device__ inline unsigned sync_group_size(unsigned G, unsigned x) { return x < G ? (G - x + NUM_XCD - 1) / NUM_XCD : 0u; }
// Phase `phase` (1, 2, ...) of a grid barrier over gridDim.x resident workgroups.
__device__ inline bool grid_barrier(unsigned* sync, unsigned phase, unsigned* err) {
  __shared__ int ok_s;
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned G = gridDim.x, x = blockIdx.x % NUM_XCD;
    const unsigned groups = G < (unsigned)NUM_XCD ? G : (unsigned)NUM_XCD;
    unsigned* all = sync + NUM_XCD * SYNC_LINE;
    int ok = 1;
    __threadfence();  // this workgroup's writes of the phase, visible device-wide
    if (atomicAdd(sync + x * SYNC_LINE, 1u) == phase * sync_group_size(G, x) - 1u) {
      __threadfence();  // the group's writes, passed on with its arrival
      atomicAdd(all, 1u);
    }
    for (unsigned n = 0; __hip_atomic_load(all, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < phase * groups; ++n) {
      if (n >= SYNC_SPIN_LIMIT) {
        ok = 0;
        atomicOr(err, (unsigned)ERR_SYNC);
        break;
      }
      __builtin_amdgcn_s_sleep(8);
    }
    __threadfence();  // and the other workgroups' writes visible here
    ok_s = ok;
  }
  __syncthreads();
  return ok_s != 0;
}
// After the last phase: the last workgroup to leave zeroes every counter for the next launch.
__device__ inline void grid_depart(unsigned* sync) {
  if (threadIdx.x != 0) return;
  const unsigned G = gridDim.x, x = blockIdx.x % NUM_XCD;
  const unsigned groups = G < (unsigned)NUM_XCD ? G : (unsigned)NUM_XCD;
  unsigned* dep = sync + (NUM_XCD + 1) * SYNC_LINE;
  if (atomicAdd(dep + x * SYNC_LINE, 1u) == sync_group_size(G, x) - 1u &&
      atomicAdd(dep + NUM_XCD * SYNC_LINE, 1u) == groups - 1u)
    for (int w = 0; w < 2 * (NUM_XCD + 1); ++w) sync[w * SYNC_LINE] = 0u;
}

// sync: SYNC_WORDS barrier counters (zero between launches).
template <typename T>
__global__ void __launch_bounds__(256) k_sort_all(const StepParams* __restrict__ Pp, Dev S, unsigned* sync) {
  if (sort_skipped(S)) return;
  const StepParams& P = *Pp;
  const unsigned G = gridDim.x;
  const int np = (S.P + 255) / 256;
  for (int vb = blockIdx.x; vb < np; vb += (int)G) k_keys_body(Pp, S, vb);
  bool ok = grid_barrier(sync, 1, S.err);
  const int nt = P.nbuckets > 0 ? (P.nbuckets + SCAN_TILE - 1) / SCAN_TILE : 1;
  if (ok)
    for (int vb = blockIdx.x; vb < nt; vb += (int)G) k_scan_local_body(S.count, S.start, S.blocksum, P.nbuckets, vb);
  ok = ok && grid_barrier(sync, 2, S.err);
  const int na = P.nbuckets > 0 ? (P.nbuckets + 255) / 256 : 1;
  if (ok)
    for (int vb = blockIdx.x; vb < na; vb += (int)G)
      k_scan_add_body(S.start, S.blocksum, P.nbuckets, S.start + P.nbuckets, vb);
  ok = ok && grid_barrier(sync, 3, S.err);
  if (ok)
    for (int vb = blockIdx.x; vb < np; vb += (int)G) k_scatter_body(S, vb);
  ok = ok && grid_barrier(sync, 4, S.err);
  if (ok) k_binsort_body<T>(Pp, S, (int)blockIdx.x, (int)G);
  __syncthreads();
  if (ok) grid_depart(sync);
}

// ======================================================================================
// Pair forces.  Thread per sorted particle: a wavefront holds 64 consecutive particles of
// one element type, i.e. a short run of bins along z.  For each source type the stencil
// (reach per type pair) is walked as (x, y) columns; the z-run of a column is one contiguous
// range of the sorted array, walked in a single loop that only tracks bin boundaries for
// the periodic shift.  Law parameters of the (target, source) type pair sit in registers;
// every candidate gets a conservative cut-off test, then the exact reference predicate.
// ======================================================================================

// Law of one (target p, source q) element-type pair with both orientations' asymmetric
// terms: the canonical "i" of a pair is its lower particle index (lower UID).
struct PairLaw {
  int law;
  double rcut2, eps, xi, e0, ew, hzk;
  double ewell_p, ewell_q;  // adhesion constant of p's / q's element type
  double ea_p, ea_q;        // EA modulus of p's / q's element type
};

__device__ inline PairLaw load_law(const StepParams& P, int tp, int tq) {
  const LawEntry& A = P.law[tp][tq];
  const LawEntry& B = P.law[tq][tp];
  PairLaw L;
  L.law = A.law;
  L.rcut2 = A.rcut2;
  L.eps = A.eps;
  L.xi = A.xi;
  L.e0 = A.e0;
  L.ew = A.ew;
  L.hzk = A.hzk;
  L.ewell_p = A.ewell_i;
  L.ewell_q = B.ewell_i;
  L.ea_p = A.ea_ei;
  L.ea_q = B.ea_ei;
  return L;
}

// Random direction of an exactly coincident 2D contact pair (MD.cu:321-331), canonical i.
__device__ inline void coincident_dir(const StepParams& P, int pidx_i, int pidx_j, uint32_t ts, uint32_t stream,
                                      double* rij) {
  double u[2];
  int ci = pidx_i / P.nE, cj = pidx_j / P.nE;
  int ei = pidx_i - ci * P.nE, ej = pidx_j - cj * P.nE;
  int si_ = species_of_cell(P, ci), sj_ = species_of_cell(P, cj);
  uint32_t ki = (cell_key(si_, ci - P.base[si_]) << 1) | ei, kj = (cell_key(sj_, cj - P.base[sj_]) << 1) | ej;
  philox_uniform2(P.seed, ki, ts, stream, kj, u);
  double alpha = u[0] * kTwoPi;
  rij[0] = sin(alpha);
  rij[1] = cos(alpha);
}

// FP64 pair evaluation from p's side with the reference's operation order.
// d = minimg(x_q - x_p) exactly as BoundingBox::distance; adds p's share to F and pr.
__device__ inline void pair64(const StepParams& P, const PairLaw& E, int pidx_p, int spec_p, double sp_, int pidx_q,
                              int spec_q, double sq_, double d[3], double r2, uint32_t ts, uint32_t st_pair, double F[3],
                              double* pr) {
  const int D = P.box.dim;
  const bool p_is_i = pidx_p < pidx_q;
  double si = p_is_i ? sp_ : sq_, sj = p_is_i ? sq_ : sp_;
  double f0;
  bool contact = false;
  double temp = 0.0;
  if (E.law == CIL_LAW_LJ) {
    double sij2 = sq(0.5 * (si + sj));
    double rc2 = sij2 * E.rcut2;
    if (!(r2 <= rc2)) return;
    double f = lj_f0(sij2, r2) - (P.force_shifted ? lj_f0(sij2, rc2) * sqrt(rc2 / r2) : 0.0);
    f0 = f * E.eps;
  } else if (E.law == CIL_LAW_SC) {
    double sigCore = 0.5 * (si + sj);
    double rc2 = E.rcut2 * sq(sigCore + E.xi);
    if (!(r2 > E.rcut2 * sq(sigCore - E.xi) && r2 <= rc2)) return;
    f0 = sc_f0(sigCore, r2, E.xi, E.e0, E.ew) -
         (P.force_shifted ? sc_f0(sigCore, rc2, E.xi, E.e0, E.ew) * sqrt(rc2 / r2) : 0.0);
  } else if (E.law == CIL_LAW_HZ || E.law == CIL_LAW_EA) {
    double r = sqrt(r2);
    if (!(r <= (si + sj) / 2)) return;
    contact = true;
    if (D == 2 && d[0] == 0.0 && d[1] == 0.0) {
      double rij[2];
      coincident_dir(P, p_is_i ? pidx_p : pidx_q, p_is_i ? pidx_q : pidx_p, ts, st_pair, rij);
      d[0] = p_is_i ? rij[0] : -rij[0];
      d[1] = p_is_i ? rij[1] : -rij[1];
    }
    double hh = (si + sj) / 2 - r;
    double h = hh >= 0.0 ? hh : 0.0;
    if (E.law == CIL_LAW_HZ) {
      double Ri = si / 2, Rj = sj / 2;
      f0 = -(pow(h, 1.5) / (E.hzk * (sqrt((1 / Ri) + (1 / Rj)))));
    } else {
      f0 = -(h * (p_is_i ? E.ea_p : E.ea_q));
    }
    double A = D == 3 ? contact_area(r, si, sj) : contact_length(si, sj, h);
    f0 += (p_is_i ? E.ewell_p : E.ewell_q) * A;
    temp = (A == 0 ? 0 : (4 * kPi * sq(si / 2) / 3));
    temp = (D == 3 ? temp : A);
  } else {
    return;
  }
  double s = (f0 < 0.0 || spec_p == spec_q) ? 1.0 : P.cross_adh[spec_p];
  double fs = s == 1.0 ? f0 : s * f0;
  for (int k = 0; k < D; ++k) F[k] += fs * d[k];
  if (contact && P.pressure && temp != 0.0) *pr -= fs / temp;
}

// Exact FP64 cut-off decision for an LJ pair whose FP32 distance is in the guard band.
__device__ inline bool lj_inside_exact(const StepParams& P, const Dev& S, int pidx_p, int pidx_q, double rcut2) {
  double xp[3], xq[3], rij[3];
  bool p_is_i = pidx_p < pidx_q;
  int a = p_is_i ? pidx_p : pidx_q, b = p_is_i ? pidx_q : pidx_p;
  for (int d = 0; d < P.box.dim; ++d) {
    xp[d] = S.pos[(size_t)d * S.P + a];
    xq[d] = S.pos[(size_t)d * S.P + b];
  }
  double r2 = box_dist(P.box, rij, xp, xq);
  double sij2 = sq(0.5 * (S.sigma[a] + S.sigma[b]));
  return r2 <= sij2 * rcut2;
}

// FP32 pair evaluation (fast mode).  Cut-off decisions of the discontinuous LJ law that fall
// inside a 1e-5 relative guard band are re-decided in FP64 from the canonical positions.
__device__ inline void pair32(const StepParams& P, const Dev& S, const PairLaw& E, int pidx_p, int spec_p, float sp_,
                              int pidx_q, int spec_q, float sq_, float d[3], float r2, uint32_t ts, float F[3],
                              float* pr) {
  const int D = P.box.dim;
  const bool p_is_i = pidx_p < pidx_q;
  float f0;
  bool contact = false;
  float temp = 0.0f;
  if (E.law == CIL_LAW_LJ) {
    float sh = 0.5f * (sp_ + sq_);
    float sij2 = sh * sh;
    float rc2 = sij2 * (float)E.rcut2;
    if (r2 > rc2 * 1.00001f) return;
    if (r2 > rc2 * 0.99999f && !lj_inside_exact(P, S, pidx_p, pidx_q, E.rcut2)) return;
    float ir2 = 1.0f / r2;
    float s2 = sij2 * ir2;
    float s6 = s2 * s2 * s2;
    float f = (s6 - 2.0f * s6 * s6) * ir2;
    if (P.force_shifted) f -= lj_f0(sij2, rc2) * sqrtf(rc2 * ir2);
    f0 = f * (float)E.eps;
  } else if (E.law == CIL_LAW_SC) {
    float xi = (float)E.xi, rcut2 = (float)E.rcut2;
    float sigCore = 0.5f * (sp_ + sq_);
    float up = sigCore + xi, lo = sigCore - xi;
    float rc2 = rcut2 * up * up;
    if (!(r2 > rcut2 * lo * lo && r2 <= rc2)) return;
    f0 = sc_f0(sigCore, r2, xi, (float)E.e0, (float)E.ew) -
         (P.force_shifted ? sc_f0(sigCore, rc2, xi, (float)E.e0, (float)E.ew) * sqrtf(rc2 / r2) : 0.0f);
  } else if (E.law == CIL_LAW_HZ || E.law == CIL_LAW_EA) {
    float si = p_is_i ? sp_ : sq_, sj = p_is_i ? sq_ : sp_;
    float r = sqrtf(r2);
    if (!(r <= (si + sj) * 0.5f)) return;
    contact = true;
    if (D == 2 && d[0] == 0.0f && d[1] == 0.0f) {
      double rij[2];
      coincident_dir(P, p_is_i ? pidx_p : pidx_q, p_is_i ? pidx_q : pidx_p, ts, force_stream(S, ST_PAIR), rij);
      d[0] = (float)(p_is_i ? rij[0] : -rij[0]);
      d[1] = (float)(p_is_i ? rij[1] : -rij[1]);
    }
    float hh = (si + sj) * 0.5f - r;
    float h = hh >= 0.0f ? hh : 0.0f;
    if (E.law == CIL_LAW_HZ) {
      float Ri = si * 0.5f, Rj = sj * 0.5f;
      f0 = -(h * sqrtf(h) / ((float)E.hzk * sqrtf(1.0f / Ri + 1.0f / Rj)));
    } else {
      f0 = -(h * (float)(p_is_i ? E.ea_p : E.ea_q));
    }
    float A;
    float ra = si * 0.5f, rb = sj * 0.5f;
    if (D == 3) {
      float u = r2 - rb * rb + ra * ra;
      float t = sqrtf(4.f * r2 * ra * ra - u * u) / (2.f * r);
      A = t > 0.f ? (float)kPi * t * t : 0.f;
    } else {
      float q = ra - (2.f * rb * h - h * h) / (2.f * (ra + rb - h));
      float t = 2.f * sqrtf(ra * ra - q * q);
      A = isnan(t) ? (si <= sj ? si : sj) : t;
    }
    f0 += (float)(p_is_i ? E.ewell_p : E.ewell_q) * A;
    temp = (A == 0.f ? 0.f : (4.f * (float)kPi * ra * ra / 3.f));
    temp = (D == 3 ? temp : A);
  } else {
    return;
  }
  float s = (f0 < 0.0f || spec_p == spec_q) ? 1.0f : (float)P.cross_adh[spec_p];
  float fs = s * f0;
  for (int k = 0; k < D; ++k) F[k] = fmaf(fs, d[k], F[k]);
  if (contact && P.pressure && temp != 0.0f) *pr -= fs / temp;
}

// Bucket of (element type t, spatial bin coordinates).
__device__ inline int bucket_of(const StepParams& P, int t, int x, int y, int z) {
  int b = P.box.dim == 3 ? (x * P.ns[1] + y) * P.ns[2] + z : x * P.ns[1] + y;
  return t * P.nbins + b;
}

// Up to two z-segments of the bins bz+olo .. bz+ohi (olo <= ohi, span <= ns in z), split at
// the periodic seam; 0 segments when a non-periodic box clips the range away.
__device__ inline int z_segments_range(const StepParams& P, int bz, int olo, int ohi, int* z0, int* z1, int* off) {
  if (P.box.dim != 3) {
    z0[0] = 0;
    z1[0] = 0;
    off[0] = 0;
    return 1;
  }
  const int nz = P.ns[2];
  int lo = bz + olo, hi = bz + ohi;
  if (!P.box.pbc[2]) {
    lo = lo < 0 ? 0 : lo;
    hi = hi > nz - 1 ? nz - 1 : hi;
    if (lo > hi) return 0;
    z0[0] = lo; z1[0] = hi; off[0] = lo - bz;
    return 1;
  }
  if (hi < 0) {
    z0[0] = lo + nz; z1[0] = hi + nz; off[0] = lo - bz;
    return 1;
  }
  if (lo > nz - 1) {
    z0[0] = lo - nz; z1[0] = hi - nz; off[0] = lo - bz;
    return 1;
  }
  if (lo < 0) {
    z0[0] = lo + nz; z1[0] = nz - 1; off[0] = lo - bz;
    z0[1] = 0; z1[1] = hi; off[1] = -bz;
    return 2;
  }
  if (hi > nz - 1) {
    z0[0] = lo; z1[0] = nz - 1; off[0] = lo - bz;
    z0[1] = 0; z1[1] = hi - nz; off[1] = nz - bz;
    return 2;
  }
  z0[0] = lo; z1[0] = hi; off[0] = lo - bz;
  return 1;
}

// Up to two z-segments of the stencil column [bz-R, bz+R] (split at the periodic seam).
// Segment k covers bins z0[k]..z1[k] whose unwrapped offset from bz starts at off[k].
__device__ inline int z_segments(const StepParams& P, int bz, int R, int* z0, int* z1, int* off) {
  if (P.box.dim != 3) {
    z0[0] = 0;
    z1[0] = 0;
    off[0] = 0;
    return 1;
  }
  const int nz = P.ns[2];
  int lo = bz - R, hi = bz + R;
  if (!P.box.pbc[2]) {
    lo = lo < 0 ? 0 : lo;
    hi = hi > nz - 1 ? nz - 1 : hi;
    z0[0] = lo;
    z1[0] = hi;
    off[0] = lo - bz;
    return 1;
  }
  if (lo < 0) {
    z0[0] = lo + nz; z1[0] = nz - 1; off[0] = lo - bz;
    z0[1] = 0; z1[1] = hi; off[1] = -bz;
    return 2;
  }
  if (hi > nz - 1) {
    z0[0] = lo; z1[0] = nz - 1; off[0] = lo - bz;
    z0[1] = 0; z1[1] = hi - nz; off[1] = nz - bz;
    return 2;
  }
  z0[0] = lo; z1[0] = hi; off[0] = lo - bz;
  return 1;
}

template <typename T>
__device__ inline void k_pair_body(const StepParams* __restrict__ Pp, const Dev& S, int bid) {
  const StepParams& P = *Pp;
  const int D = P.box.dim;
  const int total = S.start[P.nbuckets];
  const int k = bid * blockDim.x + threadIdx.x;
  if (k >= total) return;
  const uint32_t ts = (uint32_t)step_ts(S);
  const int meta = S.smeta[k];
  const int pidx = meta >> 3, tp = meta & 7;
  const int cshift = P.nE - 1;  // cell = particle >> cshift (nE is 1 or 2)
  const int cell_p = pidx >> cshift;
  const int spec_p = species_of_cell(P, cell_p);
  int bc[3];
  unpack_bin(P, S.sbin[k], bc);
  T xp[4];
  if constexpr (sizeof(T) == 4) {
    float4 r = ((const float4*)S.spos)[k];
    xp[0] = r.x; xp[1] = r.y; xp[2] = r.z; xp[3] = r.w;
  } else {
    double4 r = ((const double4*)S.spos)[k];
    xp[0] = r.x; xp[1] = r.y; xp[2] = r.z; xp[3] = r.w;
  }
  T F[3] = {0, 0, 0};
  T pr = 0;
  for (int tq = 0; tq < P.ntypes; ++tq) {
    const int Rx = P.reach_tt[tp][tq][0];
    if (Rx < 0) continue;
    const int Ry = P.reach_tt[tp][tq][1], Rz = D == 3 ? P.reach_tt[tp][tq][2] : 0;
    const PairLaw E = load_law(P, tp, tq);
    const T rca = (T)P.rca_tt[tp][tq], rcb = (T)P.rcb_tt[tp][tq], margin = (T)(sizeof(T) == 4 ? 1.0001 : 1.000001);
    int z0[2], z1[2], zo[2];
    const int nseg = z_segments(P, bc[2], Rz, z0, z1, zo);
    for (int ox = -Rx; ox <= Rx; ++ox) {
      int nx = bc[0] + ox;
      if (P.box.pbc[0]) nx = nx < 0 ? nx + P.ns[0] : (nx >= P.ns[0] ? nx - P.ns[0] : nx);
      else if (nx < 0 || nx >= P.ns[0]) continue;
      for (int oy = -Ry; oy <= Ry; ++oy) {
        int ny = bc[1] + oy;
        if (P.box.pbc[1]) ny = ny < 0 ? ny + P.ns[1] : (ny >= P.ns[1] ? ny - P.ns[1] : ny);
        else if (ny < 0 || ny >= P.ns[1]) continue;
        for (int sg = 0; sg < nseg; ++sg) {
          const int u0 = bucket_of(P, tq, nx, ny, z0[sg]);
          const int u1 = bucket_of(P, tq, nx, ny, z1[sg]);
          const int j0 = S.start[u0], j1 = S.start[u1 + 1];
          int u = u0, jend = S.start[u0 + 1];
          int oz = zo[sg];
          // p's position relative to the source bin's frame (FP32: bin-local coordinates)
          T qx = 0, qy = 0, qz = 0;
          if constexpr (sizeof(T) == 4) {
            qx = xp[0] - (T)(ox * P.w[0]);
            qy = xp[1] - (T)(oy * P.w[1]);
            qz = xp[2] - (T)(oz * P.w[2]);
          }
          for (int j = j0; j < j1; ++j) {
            while (j >= jend) {  // crossed into the next z-bin of the run
              ++u;
              ++oz;
              jend = S.start[u + 1];
              if constexpr (sizeof(T) == 4) qz -= (T)P.w[2];
            }
            const int mq = S.smeta[j];
            const int qidx = mq >> 3;
            if constexpr (sizeof(T) == 4) {
              const float4 r = ((const float4*)S.spos)[j];
              float d[3] = {r.x - qx, r.y - qy, D == 3 ? r.z - qz : 0.0f};
              if (P.reach[0] == 0 && P.box.pbc[0]) d[0] -= (float)P.box.L[0] * rintf(d[0] * (float)P.box.invL[0]);
              if (P.reach[1] == 0 && P.box.pbc[1]) d[1] -= (float)P.box.L[1] * rintf(d[1] * (float)P.box.invL[1]);
              if (D == 3 && P.reach[2] == 0 && P.box.pbc[2])
                d[2] -= (float)P.box.L[2] * rintf(d[2] * (float)P.box.invL[2]);
              const float r2 = fmaf(d[0], d[0], fmaf(d[1], d[1], d[2] * d[2]));
              const float rc = fmaf(rca, r.w + xp[3], rcb);
              if (r2 > rc * rc * margin || (qidx >> cshift) == cell_p) continue;
              float Fl[3] = {F[0], F[1], F[2]};
              float prl = pr;
              pair32(P, S, E, pidx, spec_p, xp[3], qidx, species_of_cell(P, qidx >> cshift), r.w, d, r2, ts, Fl, &prl);
              F[0] = Fl[0]; F[1] = Fl[1]; F[2] = Fl[2];
              pr = prl;
            } else {
              const double4 r = ((const double4*)S.spos)[j];
              double xq[3] = {r.x, r.y, r.z};
              double d[3] = {0, 0, 0};
              double r2 = 0.0;
              for (int dd = 0; dd < D; ++dd) {
                double x = xq[dd] - (double)xp[dd];
                if (P.box.pbc[dd]) x -= (double)nint(x / P.box.L[dd]) * P.box.L[dd];
                d[dd] = x;
                r2 += sq(x);
              }
              const double rc = rca * (r.w + (double)xp[3]) + rcb;
              if (r2 > rc * rc * margin || (qidx >> cshift) == cell_p) continue;
              double Fl[3] = {(double)F[0], (double)F[1], (double)F[2]};
              double prl = pr;
              pair64(P, E, pidx, spec_p, (double)xp[3], qidx, species_of_cell(P, qidx >> cshift), r.w, d, r2, ts,
                     force_stream(S, ST_PAIR), Fl, &prl);
              F[0] = Fl[0]; F[1] = Fl[1]; F[2] = Fl[2];
              pr = prl;
            }
          }
        }
      }
    }
  }
  T* fp = (T*)S.fpair;
  for (int dd = 0; dd < D; ++dd) fp[(size_t)dd * S.P + k] = F[dd];
  ((T*)S.ppair)[k] = pr;
}
template <typename T>
__global__ void __launch_bounds__(256) k_pair(const StepParams* __restrict__ Pp, Dev S) { k_pair_body<T>(Pp, S, xcd_block_chunked(blockIdx.x, gridDim.x)); }

// Lean FP32 pair kernel (fast mode), specialised on the dimension.  Waves are laid out per
// element type (the sorted array is type-major), so the target type, its species, the law
// of each (target, source) type pair, the stencil reach and the cross-adhesion choice are
// wave-uniform (scalar registers).  Same traversal and pair semantics as k_pair<float>.
// Single-disc cells: the same-cell exclusion is j == k, so the inner loop reads only the
// 16-byte sorted record of each candidate.
// Law coefficients of one (target, source) element-type pair for the FP32 kernels.
struct Law32 {
  int law;
  bool shifted;
  float rcut2, eps, xi, ixi3, e0, ew, hzk, ewell_p, ewell_q, ea_p, ea_q;
};
__device__ inline Law32 law32(const StepParams& P, int tp, int tq) {
  const LawEntry& A = P.law[tp][tq];
  const LawEntry& B = P.law[tq][tp];
  Law32 L;
  L.law = A.law;
  L.shifted = P.force_shifted != 0;
  L.rcut2 = (float)A.rcut2;
  L.eps = (float)A.eps;
  L.xi = (float)A.xi;
  L.ixi3 = 6.0f / (L.xi * L.xi * L.xi);
  L.e0 = (float)A.e0;
  L.ew = (float)A.ew;
  L.hzk = (float)A.hzk;
  L.ewell_p = (float)A.ewell_i;
  L.ewell_q = (float)B.ewell_i;
  L.ea_p = (float)A.ea_ei;
  L.ea_q = (float)B.ea_ei;
  return L;
}

// Hardware reciprocal and square root (v_rcp_f32, v_sqrt_f32, 1 ulp) for the FP32 laws: the
// FP32 force bar (1e-5 relative to the force scale) is far above their error, and each
// replaces a ~10-instruction correctly rounded sequence.
__device__ __forceinline__ float rcp32(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float sqrt32(float x) { return __builtin_amdgcn_sqrtf(x); }

// Law class of pair_term32: LC_ANY dispatches on the (wave-uniform) law at run time; a fixed
// class compiles one law only (LC_CONTACT: Hertz or elastic adhesion, told apart at run time).
constexpr int LC_ANY = -1, LC_CONTACT = 100;

// One pair term of the FP32 kernels (LJParams / SCParams::elemForce, MD.cuh:89-110,190-207;
// contact laws with the device semantics of MD.cu:317-362, canonical i = lower index).
// Returns false when the pair adds nothing: outside the law's range, or an LJ pair inside the
// 1e-5 cut-off guard band, which is recorded in gslot for the exact FP64 decision.  For a
// coincident 2D contact (dx, dy) become the reference's random direction.
template <int D, int LC = LC_ANY>
__device__ __forceinline__ bool pair_term32(const StepParams& P, const Dev& S, const Law32& L, int k, int j, int pidx,
                                            float sp_, float sq_, float r2, uint32_t ts, float& dx, float& dy,
                                            int& nguard, float& f0, bool& contact, float& temp) {
  const float ssum = sq_ + sp_;
  const int law = LC == LC_ANY ? L.law : LC;
  contact = false;
  temp = 0.0f;
  if (law == CIL_LAW_LJ) {  // LJParams::elemForce (MD.cuh:89-110)
    const float sh = 0.5f * ssum;
    const float sij2 = sh * sh;
    const float rc2 = sij2 * L.rcut2;
    if (r2 > rc2 * 1.00001f) return false;
    if (r2 > rc2 * 0.99999f) {  // decided exactly in FP64 by guard_force after the pair loop
      if (nguard < GUARD_SLOTS) S.gslot[(size_t)k * GUARD_SLOTS + nguard] = j;
      ++nguard;
      return false;
    }
    const float ir2 = rcp32(r2);
    const float s2 = sij2 * ir2;
    const float s6 = s2 * s2 * s2;
    float f = (s6 - 2.0f * s6 * s6) * ir2;
    if (L.shifted) f -= lj_f0(sij2, rc2) * sqrt32(rc2 * ir2);
    f0 = f * L.eps;
  } else if (law == CIL_LAW_SC) {  // SCParams::elemForce (MD.cuh:190-207)
    const float sigCore = 0.5f * ssum;
    const float up = sigCore + L.xi, lo = sigCore - L.xi;
    const float rc2 = L.rcut2 * up * up;
    if (!(r2 > L.rcut2 * lo * lo && r2 <= rc2)) return false;
    const float rr = sqrt32(r2);
    const float irr = rcp32(rr);
    const float y = rr - sigCore;
    if (y <= -L.xi) f0 = 0.0f;
    else if (y < 0.0f) f0 = -(L.e0 + L.ew) * ((-y) * (y + L.xi) * L.ixi3) * irr;
    else if (y < L.xi) f0 = L.ew * (y * (L.xi - y) * L.ixi3) * irr;
    else f0 = 0.0f;
    if (L.shifted) f0 -= sc_f0(sigCore, rc2, L.xi, L.e0, L.ew) * (sqrt32(rc2) * irr);
  } else {  // contact laws (device semantics MD.cu:317-362), canonical i = lower index
    // range test first (symmetric in the two discs), then the partner's index
    const float rr = sqrt32(r2);
    if (!(rr <= ssum * 0.5f)) return false;
    const int qidx = S.smeta[j] >> 3;
    const bool p_is_i = pidx < qidx;
    const float si = p_is_i ? sp_ : sq_, sj = p_is_i ? sq_ : sp_;
    contact = true;
    if (D == 2 && dx == 0.0f && dy == 0.0f) {
      double rij[2];
      coincident_dir(P, p_is_i ? pidx : qidx, p_is_i ? qidx : pidx, ts, force_stream(S, ST_PAIR), rij);
      dx = (float)(p_is_i ? rij[0] : -rij[0]);
      dy = (float)(p_is_i ? rij[1] : -rij[1]);
    }
    const float hh = (si + sj) * 0.5f - rr;
    const float h = hh >= 0.0f ? hh : 0.0f;
    const float ra = si * 0.5f, rb = sj * 0.5f;
    if (L.law == CIL_LAW_HZ) f0 = -(h * sqrtf(h) / (L.hzk * sqrtf(1.0f / ra + 1.0f / rb)));
    else f0 = -(h * (p_is_i ? L.ea_p : L.ea_q));
    float Acon;
    if (D == 3) {
      const float uu = r2 - rb * rb + ra * ra;
      const float t = sqrtf(4.f * r2 * ra * ra - uu * uu) / (2.f * rr);
      Acon = t > 0.f ? (float)kPi * t * t : 0.f;
    } else {
      const float q = ra - (2.f * rb * h - h * h) / (2.f * (ra + rb - h));
      const float t = 2.f * sqrtf(ra * ra - q * q);
      Acon = isnan(t) ? (si <= sj ? si : sj) : t;
    }
    f0 += (p_is_i ? L.ewell_p : L.ewell_q) * Acon;
    temp = (Acon == 0.f ? 0.f : (4.f * (float)kPi * ra * ra / 3.f));
    temp = (D == 3 ? temp : Acon);
  }
  return true;
}

// FP32 LJ pairs of sorted particle k that fell inside the 1e-5 cut-off guard band: the
// exact FP64 cut-off decision (the reference's predicate) for each, and the FP32 force of
// those inside, summed in gslot order into F.  Called by the FP32 pair kernels for their own
// particle after its pair loop (positions are stable during the pass), so no separate pass
// is launched; F's sum is then added to the pair sum, as one FP32 addition per component.
__device__ inline void guard_force(const StepParams& P, const Dev& S, int k, int n, float F[3]) {
  const int D = P.box.dim;
  const int meta = S.smeta[k];
  const int pidx = meta >> 3, tp = meta & 7;
  for (int g = 0; g < n && g < GUARD_SLOTS; ++g) {
    const int j = S.gslot[(size_t)k * GUARD_SLOTS + g];
    const int mq = S.smeta[j];
    const int qidx = mq >> 3, tq = mq & 7;
    const LawEntry& A = P.law[tp][tq];
    if (!lj_inside_exact(P, S, pidx, qidx, A.rcut2)) continue;
    double xp[3], xq[3], d64[3];
    for (int d = 0; d < D; ++d) {
      xp[d] = S.pos[(size_t)d * S.P + pidx];
      xq[d] = S.pos[(size_t)d * S.P + qidx];
    }
    box_dist(P.box, d64, xp, xq);
    float dx[3] = {(float)d64[0], (float)d64[1], D == 3 ? (float)d64[2] : 0.0f};
    const float r2 = fmaf(dx[0], dx[0], fmaf(dx[1], dx[1], dx[2] * dx[2]));
    const float sh = 0.5f * (((const float4*)S.spos)[k].w + ((const float4*)S.spos)[j].w);
    const float sij2 = sh * sh;
    const float rc2 = sij2 * (float)A.rcut2;
    const float ir2 = 1.0f / r2;
    const float s2 = sij2 * ir2;
    const float s6 = s2 * s2 * s2;
    float f = (s6 - 2.0f * s6 * s6) * ir2;
    if (P.force_shifted) f -= lj_f0(sij2, rc2) * sqrtf(rc2 * ir2);
    float f0 = f * (float)A.eps;
    const int spec_p = tp / P.nE, spec_q = tq / P.nE;
    const float fs = (f0 < 0.0f || spec_p == spec_q) ? f0 : (float)P.cross_adh[spec_p] * f0;
    for (int d = 0; d < D; ++d) F[d] = fmaf(fs, dx[d], F[d]);
  }
}

template <int D>
__device__ inline void k_pair32_body(const StepParams* __restrict__ Pp, const Dev& S, int bid) {
  const StepParams& P = *Pp;
  const float4* __restrict__ spos = (const float4*)S.spos;
  const int* __restrict__ smeta = S.smeta;
  const int* __restrict__ start = S.start;
  const int lane = threadIdx.x & (WAVE - 1);
  const int w = __builtin_amdgcn_readfirstlane(bid * (blockDim.x / WAVE) + threadIdx.x / WAVE);
  // wave -> (type, chunk of that type's sorted range)
  int tp = -1, k0 = 0, k1 = 0, acc = 0;
  for (int t = 0; t < P.ntypes; ++t) {
    const int T0 = start[t * P.nbins], T1 = start[(t + 1) * P.nbins];
    const int nw = (T1 - T0 + WAVE - 1) / WAVE;
    if (tp < 0 && w < acc + nw) {
      tp = t;
      k0 = T0 + (w - acc) * WAVE;
      k1 = T1;
    }
    acc += nw;
  }
  if (tp < 0) return;
  tp = __builtin_amdgcn_readfirstlane(tp);
  const int k = k0 + lane;
  if (k >= k1) return;
  const uint32_t ts = (uint32_t)step_ts(S);
  const int nE = P.nE;
  const int cshift = nE - 1;
  const int spec_p = tp / nE;
  const float ca_p = (float)P.cross_adh[spec_p];
  const int pidx = smeta[k] >> 3;
  const int cell_p = pidx >> cshift;
  int bc[3];
  unpack_bin(P, S.sbin[k], bc);
  const float4 me = spos[k];
  const float wx = (float)P.w[0], wy = (float)P.w[1], wz = (float)P.w[2];
  const bool mix = P.reach[0] == 0 && P.box.pbc[0], miy = P.reach[1] == 0 && P.box.pbc[1];
  const bool miz = D == 3 && P.reach[2] == 0 && P.box.pbc[2];
  float Fx = 0.f, Fy = 0.f, Fz = 0.f, pr = 0.f;
  int nguard = 0;
  for (int tq = 0; tq < P.ntypes; ++tq) {
    const int Rx = P.reach_tt[tp][tq][0];
    if (Rx < 0) continue;
    const int Ry = P.reach_tt[tp][tq][1], Rz = D == 3 ? P.reach_tt[tp][tq][2] : 0;
    const Law32 Lw = law32(P, tp, tq);
    const float rca = (float)P.rca_tt[tp][tq], rcb = (float)P.rcb_tt[tp][tq];
    const float cadh = (tq / nE == spec_p) ? 1.0f : ca_p;  // scale for f0 >= 0 (MD.cu:345-360)
    const bool same_species_cells = nE == 2 && (tq / nE == spec_p);
    const bool self_bucket = tq == tp;
    // This particle's largest cut-off against type tq (diameter bound of tq), with a margin
    // far above FP32 rounding: bins beyond it cannot hold a partner.  Columns are culled by
    // their x/y gap and each z-run is clipped to the bins its remaining radius reaches.
    const float rcp = fmaf(rca, me.w + (float)P.sigma_bound[tq], rcb) * 1.001f + 1e-6f;
    const float rcp2 = rcp * rcp;
    int z0[2], z1[2], zo[2];
    for (int ox = -Rx; ox <= Rx; ++ox) {
      int nx = bc[0] + ox;
      if (P.box.pbc[0]) nx = nx < 0 ? nx + P.ns[0] : (nx >= P.ns[0] ? nx - P.ns[0] : nx);
      else if (nx < 0 || nx >= P.ns[0]) continue;
      const float qx = me.x - (float)ox * wx;
      const float gx = mix ? 0.0f : (ox > 0 ? (float)ox * wx - me.x : (ox < 0 ? me.x - (float)(ox + 1) * wx : 0.0f));
      const float gx2 = gx > 0.0f ? gx * gx : 0.0f;
      if (gx2 > rcp2) continue;
      for (int oy = -Ry; oy <= Ry; ++oy) {
        int ny = bc[1] + oy;
        if (P.box.pbc[1]) ny = ny < 0 ? ny + P.ns[1] : (ny >= P.ns[1] ? ny - P.ns[1] : ny);
        else if (ny < 0 || ny >= P.ns[1]) continue;
        const float qy = me.y - (float)oy * wy;
        const float gy = miy ? 0.0f : (oy > 0 ? (float)oy * wy - me.y : (oy < 0 ? me.y - (float)(oy + 1) * wy : 0.0f));
        const float rem = rcp2 - gx2 - (gy > 0.0f ? gy * gy : 0.0f);
        if (rem < 0.0f) continue;
        int zlo = -Rz, zhi = Rz;
        if (D == 3 && !miz) {
          const float sr = sqrtf(rem);
          const int a = (int)floorf((me.z + sr) / wz), b = (int)ceilf((me.z - sr) / wz) - 1;
          zhi = a < Rz ? a : Rz;
          zlo = b > -Rz ? b : -Rz;
          if (zlo > zhi) continue;
        }
        const int nseg = z_segments_range(P, bc[2], zlo, zhi, z0, z1, zo);
        for (int sg = 0; sg < nseg; ++sg) {
          const int u0 = bucket_of(P, tq, nx, ny, z0[sg]);
          const int u1 = bucket_of(P, tq, nx, ny, z1[sg]);
          const int j1 = start[u1 + 1];
          int u = u0, jend = start[u0 + 1];
          float qz = me.z - (float)zo[sg] * wz;
          for (int j = start[u0]; j < j1; ++j) {
            while (j >= jend) {
              ++u;
              jend = start[u + 1];
              qz -= wz;
            }
            const float4 r = spos[j];
            float dx = r.x - qx, dy = r.y - qy, dz = D == 3 ? r.z - qz : 0.0f;
            if (mix) dx -= (float)P.box.L[0] * rintf(dx * (float)P.box.invL[0]);
            if (miy) dy -= (float)P.box.L[1] * rintf(dy * (float)P.box.invL[1]);
            if (miz) dz -= (float)P.box.L[2] * rintf(dz * (float)P.box.invL[2]);
            const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
            const float ssum = r.w + me.w;
            const float rc = fmaf(rca, ssum, rcb);
            if (r2 > rc * rc * 1.0001f) continue;
            if (self_bucket && j == k) continue;
            if (same_species_cells && ((smeta[j] >> 3) >> cshift) == cell_p) continue;
            float f0, temp;
            bool contact;
            if (!pair_term32<D>(P, S, Lw, k, j, pidx, me.w, r.w, r2, ts, dx, dy, nguard, f0, contact, temp)) continue;
            const float fs = f0 < 0.0f ? f0 : cadh * f0;
            Fx = fmaf(fs, dx, Fx);
            Fy = fmaf(fs, dy, Fy);
            if (D == 3) Fz = fmaf(fs, dz, Fz);
            if (contact && temp != 0.0f) pr -= fs / temp;
          }
        }
      }
    }
  }
  if (nguard > 0) {
    float g[3] = {0.f, 0.f, 0.f};
    guard_force(P, S, k, nguard, g);
    Fx += g[0];
    Fy += g[1];
    if (D == 3) Fz += g[2];
  }
  float* fp = (float*)S.fpair;
  fp[k] = Fx;
  fp[(size_t)S.P + k] = Fy;
  if (D == 3) fp[2 * (size_t)S.P + k] = Fz;
  ((float*)S.ppair)[k] = P.pressure ? pr : 0.0f;
  if (nguard > GUARD_SLOTS) atomicOr(S.err, (unsigned)ERR_GUARD);
}
template <int D>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6)))
k_pair32(const StepParams* __restrict__ Pp, Dev S) { k_pair32_body<D>(Pp, S, xcd_block_chunked(blockIdx.x, gridDim.x)); }

// LJ pairs whose FP32 distance fell within 1e-5 of the cut-off: exact FP64 decision
// (LJParams::elemForce predicate r^2 <= RCUT^2 sigma_ij^2 on the canonical positions), then
// the FP32 force added by the thread that owns the target particle (deterministic).
// ======================================================================================
// Neighbour-list mode (FP32).  A list entry is (sorted index << 10) | bin offset code; the
// bins are sized for cut-off + skin, so the stencil reach is at most 4 and the offsets
// (-4..4 per dimension) fit 10 bits.  Entries of one particle are grouped by source type.
// ======================================================================================
__device__ inline int list_code(int ox, int oy, int oz) { return ((ox + 4) * 9 + (oy + 4)) * 9 + (oz + 4); }

// Wave -> (element type, sorted index) of the list kernels.  Workgroups are dispatched to the
// XCDs round-robin (workgroup b runs on XCD b % 8), so XCD x takes the x-th contiguous eighth
// of every element type's waves: each XCD walks one compact region of space, whose neighbour
// positions stay in its own L2, and gets an equal share of each type (types differ in list
// length).  Within an XCD the waves run type by type.  The grid holds at least
// ceil((sum over types of ceil(waves_t / 8)) / waves per block) workgroups per XCD; waves
// beyond an XCD's share return false.
__device__ inline bool wave_type_slot(const StepParams& P, const Dev& S, int* tp_out, int* k_out) {
  const int lane = threadIdx.x & (WAVE - 1);
  const int x = blockIdx.x % NUM_XCD;
  const int v = __builtin_amdgcn_readfirstlane((int)(blockIdx.x / NUM_XCD) * (int)(blockDim.x / WAVE) +
                                               (int)(threadIdx.x / WAVE));
  int tp = -1, k0 = 0, k1 = 0, acc = 0;
  for (int t = 0; t < P.ntypes; ++t) {
    const int T0 = S.start[t * P.nbins], T1 = S.start[(t + 1) * P.nbins];
    const int nw = (T1 - T0 + WAVE - 1) / WAVE;
    const int lo = (int)(((long)x * nw) / NUM_XCD), hi = (int)(((long)(x + 1) * nw) / NUM_XCD);
    if (tp < 0 && v < acc + (hi - lo)) {
      tp = t;
      k0 = T0 + (lo + v - acc) * WAVE;
      k1 = T1;
    }
    acc += hi - lo;
  }
  if (tp < 0) return false;
  *tp_out = __builtin_amdgcn_readfirstlane(tp);
  *k_out = k0 + lane;
  return k0 + lane < k1;
}

// Wave-wide minimum / maximum over all 64 lanes (every lane must be active).
__device__ __forceinline__ int wave_min_all(int v) {
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, WAVE));
  return v;
}
__device__ __forceinline__ int wave_max_all(int v) {
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, WAVE));
  return v;
}
// Candidate window of the list build: sorted positions staged per wave in LDS.
constexpr int BUILD_WIN = 160;

// List build of sorted particle k (element type tp), run by the list pair kernel when this
// pass rebuilds: every pair within its own cut-off plus the skin, found with the culled
// stencil walk of k_pair32; records the build positions.  Only k's own list is written, so
// the pair loop of the same lane can read it right after.  Entries are stored in groups of
// four (one 16-byte store per group), each source type starting on a group boundary.
// The lanes of a wave (consecutive particles, mostly one z-run of one column) walk
// overlapping candidate ranges in each stencil column: when all 64 lanes hold a particle and
// the union of their ranges fits BUILD_WIN, the wave loads it once, coalesced, into its LDS
// window `win` and the lanes read their candidates there instead of gathering each from
// memory.  The candidates, their order and the arithmetic are the same either way.
template <int D>
__device__ inline void build_list_of(const StepParams& P, const Dev& S, int tp, int k, float4* win) {
  const float4* __restrict__ spos = (const float4*)S.spos;
  const int* __restrict__ start = S.start;
  const int nE = P.nE, cshift = nE - 1, spec_p = tp / nE;
  const int pidx = S.smeta[k] >> 3, cell_p = pidx >> cshift;
  for (int d = 0; d < D; ++d) S.xref[(size_t)d * S.P + pidx] = (float)S.pos[(size_t)d * S.P + pidx];
  S.xref[(size_t)3 * S.P + pidx] = (float)S.sigma[pidx];
  int bc[3];
  unpack_bin(P, S.sbin[k], bc);
  const float4 me = spos[k];
  if (S.spos0) reinterpret_cast<float4*>(S.spos0)[k] = me;
  const float wx = (float)P.w[0], wy = (float)P.w[1], wz = (float)P.w[2];
  const bool mix = P.reach[0] == 0 && P.box.pbc[0], miy = P.reach[1] == 0 && P.box.pbc[1];
  const bool miz = D == 3 && P.reach[2] == 0 && P.box.pbc[2];
  const float skin = P.skin;
  const size_t NP = (size_t)S.P;
  const int lane = threadIdx.x & (WAVE - 1);
  const bool full = __ballot(1) == ~0ull;  // all lanes present: the window can be shared
  int4* __restrict__ nl4 = (int4*)S.nl;
  const int ngroups = P.list_max >> 2;
  int slot = 0;  // next list slot; each source type starts on a group boundary
  int b0 = 0, b1 = 0, b2 = 0, b3 = 0;  // the group being filled
  bool over = false;
  for (int tq = 0; tq < P.ntypes; ++tq) {
    int m = 0;
    const int Rx = P.reach_tt[tp][tq][0];
    if (Rx >= 0) {
      const int Ry = P.reach_tt[tp][tq][1], Rz = D == 3 ? P.reach_tt[tp][tq][2] : 0;
      const float rca = (float)P.rca_tt[tp][tq], rcb = (float)P.rcb_tt[tp][tq];
      const bool same_species_cells = nE == 2 && (tq / nE == spec_p);
      const bool self_bucket = tq == tp;
      const float rcp = (fmaf(rca, me.w + (float)P.sigma_bound[tq], rcb) + skin) * 1.001f + 1e-6f;
      const float rcp2 = rcp * rcp;
      int z0[2], z1[2], zo[2];
      // the (ox, oy) loops are wave-uniform; a lane whose column is culled or outside the
      // box walks no segment but still takes part in the window's reduction and load
      for (int ox = -Rx; ox <= Rx; ++ox) {
        int nx = bc[0] + ox;
        bool okx = true;
        if (P.box.pbc[0]) nx = nx < 0 ? nx + P.ns[0] : (nx >= P.ns[0] ? nx - P.ns[0] : nx);
        else okx = nx >= 0 && nx < P.ns[0];
        const float qx = me.x - (float)ox * wx;
        const float gx = mix ? 0.0f : (ox > 0 ? (float)ox * wx - me.x : (ox < 0 ? me.x - (float)(ox + 1) * wx : 0.0f));
        const float gx2 = gx > 0.0f ? gx * gx : 0.0f;
        okx = okx && !(gx2 > rcp2);
        for (int oy = -Ry; oy <= Ry; ++oy) {
          int ny = bc[1] + oy;
          bool ok = okx;
          if (P.box.pbc[1]) ny = ny < 0 ? ny + P.ns[1] : (ny >= P.ns[1] ? ny - P.ns[1] : ny);
          else ok = ok && ny >= 0 && ny < P.ns[1];
          const float qy = me.y - (float)oy * wy;
          const float gy = miy ? 0.0f : (oy > 0 ? (float)oy * wy - me.y : (oy < 0 ? me.y - (float)(oy + 1) * wy : 0.0f));
          const float rem = rcp2 - gx2 - (gy > 0.0f ? gy * gy : 0.0f);
          ok = ok && !(rem < 0.0f);
          int zlo = -Rz, zhi = Rz;
          if (D == 3 && !miz && ok) {
            const float sr = sqrtf(rem);
            const int a = (int)floorf((me.z + sr) / wz), b = (int)ceilf((me.z - sr) / wz) - 1;
            zhi = a < Rz ? a : Rz;
            zlo = b > -Rz ? b : -Rz;
            ok = zlo <= zhi;
          }
          const int nseg = ok ? z_segments_range(P, bc[2], zlo, zhi, z0, z1, zo) : 0;
          int u0s[2] = {0, 0}, j0s[2] = {0, 0}, j1s[2] = {0, 0};
          int lo = 0x7fffffff, hi = (int)0x80000000;
          for (int sg = 0; sg < nseg; ++sg) {
            u0s[sg] = bucket_of(P, tq, nx, ny, z0[sg]);
            j0s[sg] = start[u0s[sg]];
            j1s[sg] = start[bucket_of(P, tq, nx, ny, z1[sg]) + 1];
            lo = j0s[sg] < lo ? j0s[sg] : lo;
            hi = j1s[sg] > hi ? j1s[sg] : hi;
          }
          bool staged = false;
          int wlo = 0;
          if (full) {
            wlo = wave_min_all(lo);
            const int whi = wave_max_all(hi);
            staged = whi > wlo && whi - wlo <= BUILD_WIN;
            if (staged) {
              for (int c = lane; c < whi - wlo; c += WAVE) win[c] = spos[wlo + c];
              wave_lds_sync();
            }
          }
          // one compiled walk per candidate source, so that each issues a single kind of load
          auto walk = [&](auto from_win) {
            constexpr bool FROM_WIN = decltype(from_win)::value;
            for (int sg = 0; sg < nseg; ++sg) {
              const int j1 = j1s[sg];
              int u = u0s[sg], jend = start[u + 1];
              float qz = me.z - (float)zo[sg] * wz;
              for (int j = j0s[sg]; j < j1; ++j) {
                while (j >= jend) {
                  ++u;
                  jend = start[u + 1];
                  qz -= wz;
                }
                const float4 r = FROM_WIN ? win[j - wlo] : spos[j];
                float dx = r.x - qx, dy = r.y - qy, dz = D == 3 ? r.z - qz : 0.0f;
                if (mix) dx -= (float)P.box.L[0] * rintf(dx * (float)P.box.invL[0]);
                if (miy) dy -= (float)P.box.L[1] * rintf(dy * (float)P.box.invL[1]);
                if (miz) dz -= (float)P.box.L[2] * rintf(dz * (float)P.box.invL[2]);
                const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                const float rl = (fmaf(rca, r.w + me.w, rcb) + skin) * 1.001f;
                if (r2 > rl * rl) continue;
                if (self_bucket && j == k) continue;
                if (same_species_cells && ((S.smeta[j] >> 3) >> cshift) == cell_p) continue;
                const int oz = D == 3 ? zo[sg] + (u - u0s[sg]) : 0;
                const int e = (int)(((unsigned)j << 10) | (unsigned)list_code(mix ? 0 : ox, miy ? 0 : oy, miz ? 0 : oz));
                const int s = slot & 3;
                b0 = s == 0 ? e : b0;
                b1 = s == 1 ? e : b1;
                b2 = s == 2 ? e : b2;
                b3 = s == 3 ? e : b3;
                ++slot;
                ++m;
                if (s == 3) {  // group complete: one 16-byte store
                  const int g = (slot >> 2) - 1;
                  if (g < ngroups) nl4[(size_t)g * NP + k] = make_int4(b0, b1, b2, b3);
                  else over = true;
                }
              }
            }
          };
          if (staged) walk(std::true_type{});
          else walk(std::false_type{});
          if (staged) wave_lds_sync();  // every lane done with the window before it is rewritten
        }
      }
    }
    if (slot & 3) {  // partial last group of this source type (its unused slots are never read)
      const int g = slot >> 2;
      if (g < ngroups) nl4[(size_t)g * NP + k] = make_int4(b0, b1, b2, b3);
      else over = true;
      slot = (slot + 3) & ~3;
    }
    S.nlt[(size_t)tq * NP + k] = m;
  }
  if (over) {  // capacity error: an empty list keeps the pair loop inside the allocation
    for (int tq = 0; tq < P.ntypes; ++tq) S.nlt[(size_t)tq * NP + k] = 0;
    atomicOr(S.err, (unsigned)ERR_LIST);
  }
}

// Bin-offset table of the list kernels, in LDS: entry code -> ((float)ox * wx, (float)oy * wy,
// (float)oz * wz), the products k_pair32 forms for the same bin offsets.  Filled by the whole
// block before any wave leaves.
constexpr int LIST_CODES = 729;
constexpr int LIST_BLOCK = 256;
#ifndef LIST_WAVES_PER_EU
#define LIST_WAVES_PER_EU 6
#endif
__device__ inline void list_offsets_fill(const StepParams& P, float4* tab) {
  const float wx = (float)P.w[0], wy = (float)P.w[1], wz = (float)P.w[2];
  for (int c = threadIdx.x; c < LIST_CODES; c += blockDim.x) {
    const int ox = c / 81 - 4, oy = (c / 9) % 9 - 4, oz = c % 9 - 4;
    tab[c] = make_float4((float)ox * wx, (float)oy * wy, (float)oz * wz, 0.0f);
  }
  __syncthreads();
}

// Branch-free LJ and soft-core terms of the list kernel: the same operations and roundings as
// pair_term32, with out-of-range pairs giving f0 = 0 by selection (their fmaf then leaves the
// sums unchanged), so every lane runs one instruction stream.  lj_term32 sets *guard for a
// pair inside the 1e-5 cut-off guard band (f0 = 0; decided in FP64 later).
// SH: the law is force-shifted (compile-time, so the loop carries no run-time test).  The
// guard band is |r2 - rc2| <= 1e-5 rc2; pairs below it interact, pairs above do not.
template <bool SH>
__device__ __forceinline__ float lj_term32(const Law32& L, float ssum, float r2, bool* guard) {
  const float sh = 0.5f * ssum;
  const float sij2 = sh * sh;
  const float rc2 = sij2 * L.rcut2;
  const float dr = r2 - rc2, tol = rc2 * 1e-5f;
  const bool in = dr < -tol;
  *guard = fabsf(dr) <= tol;
  const float ir2 = rcp32(r2);
  const float s2 = sij2 * ir2;
  const float s6 = s2 * s2 * s2;
  float f = (s6 - 2.0f * s6 * s6) * ir2;
  if (SH) f -= lj_f0(sij2, rc2) * sqrt32(rc2 * ir2);
  return in ? f * L.eps : 0.0f;
}
template <bool SH>
__device__ __forceinline__ float sc_term32(const Law32& L, float ssum, float r2) {
  const float sigCore = 0.5f * ssum;
  const float up = sigCore + L.xi, lo = sigCore - L.xi;
  const float rc2 = L.rcut2 * up * up;
  const bool valid = r2 > L.rcut2 * lo * lo && r2 <= rc2;
  const float rr = sqrt32(r2);
  const float irr = rcp32(rr);
  const float y = rr - sigCore;
  const bool neg = y < 0.0f;
  const float a = neg ? -y : y, b = neg ? (y + L.xi) : (L.xi - y);
  const float coef = neg ? -(L.e0 + L.ew) : L.ew;
  const float f = coef * (a * b * L.ixi3) * irr;
  float f0 = (valid && y > -L.xi && y < L.xi) ? f : 0.0f;
  if (SH) f0 = valid ? f0 - sc_f0(sigCore, rc2, L.xi, L.e0, L.ew) * (sqrt32(rc2) * irr) : 0.0f;
  return f0;
}

// A group of four list entries, loaded non-temporally: the list (about 150 MB per pass at 1M
// particles) is read once per pass, and streaming it past the caches keeps the update's
// arrays resident in the last-level cache (update 45 -> 37 us, this kernel 5 us slower).
__device__ __forceinline__ int4 list_group(const char* __restrict__ nlb, unsigned off) {
  typedef int v4i __attribute__((ext_vector_type(4)));
  const v4i v = __builtin_nontemporal_load((const v4i*)(nlb + off));
  return make_int4(v.x, v.y, v.z, v.w);
}

// One source-type segment of a particle's list: n entries in groups of four from byte offset
// `off` of the list (groups `stride` bytes apart), in list order (the summation order is the
// list order).  LC is the segment's law class, MI whether any dimension uses the single-bin
// minimum image.
// The inner list a refresh pass writes for one particle (two-level lists): the entries of the
// (outer) list within cut-off + inner skin, in list order, in groups of four; each source type
// starts on a group boundary, as in the list build.
struct InnerList {
  int4* nl4;
  size_t NP;
  int k, ngroups;
  float rca, rcb, skin;  // the list radius of the current source type: (rca (s_i + s_j) + rcb + skin) 1.001
  int slot, m, b0, b1, b2, b3;
  bool over;
  __device__ __forceinline__ void append(int e) {
    const int s = slot & 3;
    b0 = s == 0 ? e : b0;
    b1 = s == 1 ? e : b1;
    b2 = s == 2 ? e : b2;
    b3 = s == 3 ? e : b3;
    ++slot;
    ++m;
    if (s == 3) {
      const int g = (slot >> 2) - 1;
      if (g < ngroups) nl4[(size_t)g * NP + k] = make_int4(b0, b1, b2, b3);
      else over = true;
    }
  }
  // close the current source type: its partial last group, and its entry count
  __device__ __forceinline__ void end_type(int* count) {
    if (slot & 3) {
      const int g = slot >> 2;
      if (g < ngroups) nl4[(size_t)g * NP + k] = make_int4(b0, b1, b2, b3);
      else over = true;
      slot = (slot + 3) & ~3;
    }
    *count = m;
    m = 0;
  }
};

template <int D, int LC, bool MI, bool SH, bool RF>
__device__ __forceinline__ void list_segment(const StepParams& P, const Dev& S, const Law32& Lw, const float4* tab,
                                             const float4& me, int k, int pidx, uint32_t ts, float cadh, int n,
                                             const char* __restrict__ nlb, unsigned& off, unsigned stride, float& Fx,
                                             float& Fy, float& Fz, float& pr, int& nguard, InnerList& il) {
  const float4* __restrict__ spos = (const float4*)S.spos;
  auto term = [&](unsigned e, const float4& r, const float4& o) {
    float dx = r.x - (me.x - o.x);
    float dy = r.y - (me.y - o.y);
    float dz = D == 3 ? r.z - (me.z - o.z) : 0.0f;
    if (MI) {
      if (P.reach[0] == 0 && P.box.pbc[0]) dx -= (float)P.box.L[0] * rintf(dx * (float)P.box.invL[0]);
      if (P.reach[1] == 0 && P.box.pbc[1]) dy -= (float)P.box.L[1] * rintf(dy * (float)P.box.invL[1]);
      if (D == 3 && P.reach[2] == 0 && P.box.pbc[2]) dz -= (float)P.box.L[2] * rintf(dz * (float)P.box.invL[2]);
    }
    const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    if (RF) {  // refresh pass: keep the entry in the inner list when within its radius
      const float rl = (fmaf(il.rca, r.w + me.w, il.rcb) + il.skin) * 1.001f;
      if (!(r2 > rl * rl)) il.append((int)e);
    }
    float f0;
    if (LC == CIL_LAW_LJ) {
      bool guard;
      f0 = lj_term32<SH>(Lw, r.w + me.w, r2, &guard);
      if (guard) {  // rare: recorded for the exact FP64 decision after the loop
        if (nguard < GUARD_SLOTS) S.gslot[(size_t)k * GUARD_SLOTS + nguard] = (int)(e >> 10);
        ++nguard;
      }
    } else if (LC == CIL_LAW_SC) {
      f0 = sc_term32<SH>(Lw, r.w + me.w, r2);
    } else {
      float temp;
      bool contact;
      if (!pair_term32<D, LC>(P, S, Lw, k, (int)(e >> 10), pidx, me.w, r.w, r2, ts, dx, dy, nguard, f0, contact,
                              temp))
        return;
      const float fs = f0 < 0.0f ? f0 : cadh * f0;
      Fx = fmaf(fs, dx, Fx);
      Fy = fmaf(fs, dy, Fy);
      if (D == 3) Fz = fmaf(fs, dz, Fz);
      if (contact && temp != 0.0f) pr -= fs / temp;
      return;
    }
    const float fs = f0 < 0.0f ? f0 : cadh * f0;
    Fx = fmaf(fs, dx, Fx);
    Fy = fmaf(fs, dy, Fy);
    if (D == 3) Fz = fmaf(fs, dz, Fz);
  };
  int i = 0;
  // one group per round: its four positions are loaded before any is used, so that four
  // independent gathers are in flight per lane
  for (; i + 4 <= n; i += 4) {
    const int4 g = list_group(nlb, off);
    off += stride;
    const unsigned e[4] = {(unsigned)g.x, (unsigned)g.y, (unsigned)g.z, (unsigned)g.w};
    if (RF) {  // refresh pass: two gathers in flight, leaving registers for the inner list's state
#pragma unroll
      for (int h = 0; h < 4; h += 2) {
        const float4 r0 = spos[e[h] >> 10], r1 = spos[e[h + 1] >> 10];
        term(e[h], r0, tab[e[h] & 1023u]);
        term(e[h + 1], r1, tab[e[h + 1] & 1023u]);
      }
    } else {
      float4 r[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) r[u] = spos[e[u] >> 10];
      float4 o[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) o[u] = tab[e[u] & 1023u];
#pragma unroll
      for (int u = 0; u < 4; ++u) term(e[u], r[u], o[u]);
    }
  }
  if (i < n) {  // last, partial group: 1 to 3 entries
    const int4 g = list_group(nlb, off);
    off += stride;
    const unsigned e[3] = {(unsigned)g.x, (unsigned)g.y, (unsigned)g.z};
    term(e[0], spos[e[0] >> 10], tab[e[0] & 1023u]);
    if (i + 1 < n) term(e[1], spos[e[1] >> 10], tab[e[1] & 1023u]);
    if (i + 2 < n) term(e[2], spos[e[2] >> 10], tab[e[2] & 1023u]);
  }
}

// Pair forces over the neighbour list (every step), after rebuilding the particle's list when
// this pass's rebuild flag is set.  Same law code, exclusions, guard band and per-type wave
// layout as k_pair32.  List groups are addressed with 32-bit byte offsets
// (list_max * P * 4 < 2^32, checked by the host).
template <int D>
__device__ inline void k_pair_list_body(const StepParams* __restrict__ Pp, const Dev& S) {
  const StepParams& P = *Pp;
  __shared__ float4 tab[LIST_CODES];
  __shared__ float4 bwin[LIST_BLOCK / WAVE][BUILD_WIN];  // per-wave candidate windows of the build
  list_offsets_fill(P, tab);
  const bool rebuild = *rebuild_flag(S, step_ts(S)) != 0u;
  // a re-sort whose grid barrier timed out left the sorted arrays half built: no list is built
  // from them (the host reports the error at its next check)
  if (rebuild && (__hip_atomic_load(S.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & ERR_SYNC)) return;
  if (rebuild && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(S.rebuild + 2, 1u);  // list epoch
  int tp, k;
  if (!wave_type_slot(P, S, &tp, &k)) return;
  if (rebuild) build_list_of<D>(P, S, tp, k, bwin[threadIdx.x / WAVE]);
  const float4* __restrict__ spos = (const float4*)S.spos;
  const uint32_t ts = (uint32_t)step_ts(S);
  const int nE = P.nE, spec_p = tp / nE;
  const float ca_p = (float)P.cross_adh[spec_p];
  const int pidx = S.smeta[k] >> 3;
  const float4 me = spos[k];
  const bool mi = (P.reach[0] == 0 && P.box.pbc[0]) || (P.reach[1] == 0 && P.box.pbc[1]) ||
                  (D == 3 && P.reach[2] == 0 && P.box.pbc[2]);
  const size_t NP = (size_t)S.P;
  const unsigned stride = 16u * (unsigned)S.P;
  unsigned off = 16u * (unsigned)k;
  float Fx = 0.f, Fy = 0.f, Fz = 0.f, pr = 0.f;
  int nguard = 0;
  // two-level lists: a refresh pass evaluates the outer list and writes the inner one; other
  // passes evaluate the inner list.  The two loops are compiled apart, so that the inner list's
  // append state is live only in the refresh pass.
  const bool two = S.nl_in != nullptr;
  const bool refresh = two && (rebuild || *refresh_flag(S, step_ts(S)) != 0u);
  auto run_types = [&](auto rfc) {
    constexpr bool RF = decltype(rfc)::value;
    const bool inner = two && !RF;
    const char* __restrict__ nlb = inner ? (const char*)S.nl_in : (const char*)S.nl;
    const int* __restrict__ nlt = inner ? S.nlt_in : S.nlt;
    InnerList il{(int4*)S.nl_in, NP, k, P.list_max >> 2, 0.f, 0.f, P.skin_in, 0, 0, 0, 0, 0, 0, false};
    for (int tq = 0; tq < P.ntypes; ++tq) {
      const int n = nlt[(size_t)tq * NP + k];
      if (n > 0) {
        const Law32 Lw = law32(P, tp, tq);
        const float cadh = (tq / nE == spec_p) ? 1.0f : ca_p;  // scale for f0 >= 0 (MD.cu:345-360)
        if (RF) {
          il.rca = (float)P.rca_tt[tp][tq];
          il.rcb = (float)P.rcb_tt[tp][tq];
        }
#define CIL_SEG(LC, MIV, SHV)                                                                                 \
  list_segment<D, LC, MIV, SHV, RF>(P, S, Lw, tab, me, k, pidx, ts, cadh, n, nlb, off, stride, Fx, Fy, Fz, pr, \
                                    nguard, il)
        if (Lw.law == CIL_LAW_LJ) {
          if (Lw.shifted) {
            if (mi) CIL_SEG(CIL_LAW_LJ, true, true); else CIL_SEG(CIL_LAW_LJ, false, true);
          } else {
            if (mi) CIL_SEG(CIL_LAW_LJ, true, false); else CIL_SEG(CIL_LAW_LJ, false, false);
          }
        } else if (Lw.law == CIL_LAW_SC) {
          if (Lw.shifted) {
            if (mi) CIL_SEG(CIL_LAW_SC, true, true); else CIL_SEG(CIL_LAW_SC, false, true);
          } else {
            if (mi) CIL_SEG(CIL_LAW_SC, true, false); else CIL_SEG(CIL_LAW_SC, false, false);
          }
        } else {
          if (mi) CIL_SEG(LC_CONTACT, true, false); else CIL_SEG(LC_CONTACT, false, false);
        }
#undef CIL_SEG
      }
      if (RF) il.end_type(S.nlt_in + (size_t)tq * NP + k);
    }
    if (RF) {
      if (il.over) {  // capacity error: an empty inner list keeps later passes inside the allocation
        for (int tq = 0; tq < P.ntypes; ++tq) S.nlt_in[(size_t)tq * NP + k] = 0;
        atomicOr(S.err, (unsigned)ERR_LIST);
      }
      for (int d = 0; d < D; ++d) S.xref_in[(size_t)d * NP + pidx] = (float)S.pos[(size_t)d * NP + pidx];
    }
  };
  if (refresh) run_types(std::true_type{});
  else run_types(std::false_type{});
  if (nguard > 0) {
    float g[3] = {0.f, 0.f, 0.f};
    guard_force(P, S, k, nguard, g);
    Fx += g[0];
    Fy += g[1];
    if (D == 3) Fz += g[2];
  }
  float* fp = (float*)S.fpair;
  fp[k] = Fx;
  fp[NP + k] = Fy;
  if (D == 3) fp[2 * NP + k] = Fz;
  ((float*)S.ppair)[k] = P.pressure ? pr : 0.0f;
  if (nguard > GUARD_SLOTS) atomicOr(S.err, (unsigned)ERR_GUARD);
}
template <int D>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(LIST_WAVES_PER_EU)))
k_pair_list(const StepParams* __restrict__ Pp, Dev S) {
  k_pair_list_body<D>(Pp, S);
}

// Both rebuild flags (either step parity) to v.
// Rebuild flags of both parities = v, inner-list refresh flags of both parities = vin.
__global__ void k_set_flags(unsigned* flag, unsigned v, unsigned vin) {
  if (threadIdx.x < 2 && blockIdx.x == 0) flag[threadIdx.x] = v;
  else if (threadIdx.x < 4 && blockIdx.x == 0) flag[4 + (threadIdx.x - 2)] = vin;
}



// ======================================================================================
// Noise draws (updateWienerProcess, CIL.cu:793-818), reference noise modes only.
// ======================================================================================
__device__ inline void k_noise_body(const StepParams* __restrict__ Pp, const Dev& S, int bid) {
  const StepParams& P = *Pp;
  int cslot = bid * blockDim.x + threadIdx.x;
  if (cslot >= S.C) return;
  int s, c;
  if (!cell_active(P, S.ncells, cslot, &s, &c)) return;
  const int D = P.box.dim;
  const int ts = step_ts(S);
  const cil_species& sp = P.sp[s];
  const double* blk = nullptr;
  long off_passive = 0, off_active = 0;
  if (P.rng == CIL_RNG_EXTERNAL) {
    int bi = ts - S.ext_ts0;
    if (bi < 0 || bi >= S.ext_nsteps) {
      atomicOr(S.err, (unsigned)ERR_NOISE);
      return;
    }
    blk = S.ext + (long)bi * S.ext_per_step;
    for (int k = 0; k < P.nspec; ++k) {
      if (P.noise == CIL_NOISE_CELL_WIENER && P.sp[k].d_noise > 0.0) {
        if (k < s) off_passive += (long)S.ncells[k] * D;
        off_active += (long)S.ncells[k] * D;
      }
    }
  }
  if (P.noise == CIL_NOISE_CELL_WIENER) {
    for (int d = 0; d < D; ++d) {
      double w = 0.0;
      if (sp.d_noise > 0.0) {
        double z;
        long at = off_passive + (long)c * D + d;
        if (blk && at >= S.ext_per_step) {
          atomicOr(S.err, (unsigned)ERR_NOISE);
          return;
        }
        if (blk) z = blk[at];
        else {
          double zz[2];
          philox_normal2(P.seed, cell_key(s, c), (uint32_t)ts, force_stream(S, ST_WIENER), (uint32_t)(d >> 1), zz);
          z = zz[d & 1];
        }
        w = sp.d_noise * z;
      }
      S.wiener[(size_t)cslot * 3 + d] = w;
    }
  }
  if (P.active_noise && P.nE == 2) {
    long ar = off_active, aa = off_active;
    if (blk) {
      for (int k = 0; k < P.nspec; ++k) {
        if (P.sp[k].d_noise_radial > 0.0) {
          if (k < s) ar += S.ncells[k];
          aa += S.ncells[k];
        }
      }
      for (int k = 0; k < s; ++k)
        if (P.sp[k].d_noise_angular > 0.0) aa += S.ncells[k];
    }
    double zz[2] = {0, 0};
    if (blk && (ar + c >= S.ext_per_step || aa + c >= S.ext_per_step)) {
      atomicOr(S.err, (unsigned)ERR_NOISE);
      return;
    }
    if (!blk) philox_normal2(P.seed, cell_key(s, c), (uint32_t)ts, force_stream(S, ST_ACTIVE), 0, zz);
    double wr = 0, wa = 0;
    if (sp.d_noise_radial > 0.0) wr = sp.d_noise_radial * (blk ? blk[ar + c] : zz[0]);
    if (sp.d_noise_angular > 0.0) wa = sp.d_noise_angular * (blk ? blk[aa + c] : zz[1]);
    S.awiener[(size_t)cslot * 2] = wr;
    S.awiener[(size_t)cslot * 2 + 1] = wa;
  }
}
__global__ void k_noise(const StepParams* __restrict__ Pp, Dev S) { k_noise_body(Pp, S, (int)blockIdx.x); }

// ======================================================================================
// Forces of one cell at the last force evaluation: intra (lazy) + pair (sorted).
// ======================================================================================
template <typename T, int DIM, int NE>
__device__ inline void cell_total_force(const StepParams& P, const Dev& S, int s, int c, int cslot, bool materialized,
                                        double pos[2][3], double vel[2][3], const double* sig, double F[2][3],
                                        double* pc, bool* pos1_changed) {
  constexpr int D = DIM, nE = NE;
  *pos1_changed = false;
  if (materialized) {
    for (int e = 0; e < nE; ++e)
      for (int d = 0; d < D; ++d) F[e][d] = S.frc[(size_t)d * S.P + (size_t)cslot * nE + e];
    *pc = S.pressure[cslot];
    return;
  }
  intra_cell<DIM, NE>(P, S, s, c, cslot, step_ts(S), pos, vel, sig, F, pos1_changed);
  const T* fp = (const T*)S.fpair;
  double pr = 0.0;
  for (int e = 0; e < nE; ++e) {
    int k = S.inv[(size_t)cslot * nE + e];
    for (int d = 0; d < D; ++d) F[e][d] += (double)fp[(size_t)d * S.P + k];
    if (P.pressure) pr += (double)((const T*)S.ppair)[k];
  }
  *pc = pr;
}

// ======================================================================================
// Fused update (updateVelocityPositions, CIL.cu:902-939), one thread per cell slot.
// ======================================================================================

// Neighbour-list bookkeeping of particle p moved to xn (list mode): the skin test against the
// build position raises `next`, the flag of the force pass that follows, when it moved more
// than skin / 2; otherwise the bin-local FP32 coordinates of its sorted slot (bins of the
// last sort) are refreshed.  A set flag means that pass re-sorts anyway (and, right after an
// upload, that no sorted slot exists yet), so the refresh is skipped then.  Every lane of the
// wave that holds a particle calls it (one atomic per wave).
template <int D>
__device__ inline void list_track(const StepParams& P, const Dev& S, size_t p, const double* xn, double sig,
                                  unsigned* next) {
  bool finite = true;
  float dr2 = 0.0f;
  for (int d = 0; d < D; ++d) {
    finite = finite && isfinite(xn[d]);
    float dd = (float)(xn[d] - (double)S.xref[(size_t)d * S.P + p]);
    if (P.box.pbc[d]) dd -= (float)P.box.L[d] * rintf(dd * (float)P.box.invL[d]);
    dr2 = fmaf(dd, dd, dr2);
  }
  if (!finite) atomicOr(S.err, (unsigned)ERR_NONFINITE);
  const bool moved = !(dr2 <= 0.25f * P.skin * P.skin);
  // one atomic per wave (its first moved lane), none once the flag is up
  const unsigned long long mv = __ballot(moved);
  if (mv != 0ull && (int)(threadIdx.x & (WAVE - 1)) == __ffsll(mv) - 1 && *next == 0u) atomicOr(next, 1u);
  if (S.xref_in) {  // two-level lists: the inner list's own skin test, against its last refresh
    float di2 = 0.0f;
    for (int d = 0; d < D; ++d) {
      float dd = (float)(xn[d] - (double)S.xref_in[(size_t)d * S.P + p]);
      if (P.box.pbc[d]) dd -= (float)P.box.L[d] * rintf(dd * (float)P.box.invL[d]);
      di2 = fmaf(dd, dd, di2);
    }
    const unsigned long long mi = __ballot(!(di2 <= 0.25f * P.skin_in * P.skin_in));
    unsigned* next_in = next + 4;  // the refresh flag of the same pass
    if (mi != 0ull && (int)(threadIdx.x & (WAVE - 1)) == __ffsll(mi) - 1 && *next_in == 0u) atomicOr(next_in, 1u);
  }
  if (!moved && *next == 0u) {
    const int k = S.inv[p];
    int bc[3];
    unpack_bin(P, S.sbin[k], bc);
    // the whole float4 slot in one aligned store (sigma rewritten with the value the sort
    // stored), not a partial 12-byte write
    float l4[4] = {0.f, 0.f, 0.f, (float)sig};
    for (int d = 0; d < D; ++d) {
      double l = xn[d] - bc[d] * P.w[d];
      if (P.box.pbc[d]) l -= P.box.L[d] * rint(l * P.box.invL[d]);
      l4[d] = (float)l;
    }
    ((float4*)S.spos)[k] = make_float4(l4[0], l4[1], l4[2], l4[3]);
  }
}

template <typename T, int DIM, int NE>
__device__ inline void k_update_body(const StepParams* __restrict__ Pp, const Dev& S, int materialized, int emit_keys, int bid) {
  const StepParams& P = *Pp;
  // list mode: the flag of the previous force pass (same step parity as this update) has been
  // used; the skin test below sets the other one, for this step's force pass
  if (emit_keys == 2 && bid == 0 && threadIdx.x == 0) {
    *rebuild_flag(S, step_ts(S)) = 0u;
    if (S.xref_in) *refresh_flag(S, step_ts(S)) = 0u;
  }
  int cslot = bid * blockDim.x + threadIdx.x;
  if (cslot >= S.C) return;
  int s, c;
  if (!cell_active(P, S.ncells, cslot, &s, &c)) {
    if (emit_keys)
      for (int e = 0; e < P.nE; ++e) S.key[(size_t)cslot * P.nE + e] = -1;
    return;
  }
  constexpr int D = DIM, nE = NE;
  const cil_species& sp = P.sp[s];
  // the intra force reads the old velocity only for motility (one disc) or quadratic drag
  // (two discs); otherwise it is overwritten without being read
  const bool need_vel = nE == 1 ? sp.motility[0] != 0.0 : sp.quad_drag != 0;
  double pos[2][3], vel[2][3], sig[2], F[2][3];
  for (int e = 0; e < nE; ++e) {
    size_t p = (size_t)cslot * nE + e;
    sig[e] = S.sigma[p];
    for (int d = 0; d < D; ++d) {
      pos[e][d] = S.pos[(size_t)d * S.P + p];
      vel[e][d] = need_vel ? S.vel[(size_t)d * S.P + p] : 0.0;
    }
  }
  double pc;
  bool pos1_changed;
  cell_total_force<T, DIM, NE>(P, S, s, c, cslot, materialized != 0, pos, vel, sig, F, &pc, &pos1_changed);
  if (!materialized && P.pressure) S.pressure[cslot] = pc;
  const int tsu = step_ts(S);  // step index before the increment (FDT noise key)
  double area = sigma2area(sig[0]);
  double pND = pc / sp.cycle_p0;
  for (int e = 0; e < nE; ++e) {
    size_t p = (size_t)cslot * nE + e;
    double temp;
    if (S.heun) {  // updatePositions with the half step (CIL.cu:849-866; hdt = 0.5 dt, CIL.cuh:208)
      temp = (0.5 * P.dt) * (1.0 / sp.zeta[e]);
    } else {
      double pterm = sp.zeta_p[e] == 0.0 ? 0.0 : sp.zeta_p[e] * pow(pND, sp.zeta_p_beta[e]);
      temp = P.dt * 1 / ((sp.zeta[e] + pterm) * area);
    }
    double mu = temp / P.dt;
    double z[4] = {0, 0, 0, 0};
    bool fdt = !S.heun && P.noise == CIL_NOISE_FDT && sp.kT > 0.0;
    if (fdt) {
      uint32_t id = (cell_key(s, c) << 1) | (uint32_t)e;
      if (sizeof(T) == 4) {  // FP32 force mode: FP32 Box-Muller on the same draws
        philox_normal2_f32(P.seed, id, (uint32_t)tsu, ST_FDT, 0, z);
        if (D == 3) philox_normal2_f32(P.seed, id, (uint32_t)tsu, ST_FDT, 1, z + 2);
      } else {
        philox_normal2(P.seed, id, (uint32_t)tsu, ST_FDT, 0, z);
        if (D == 3) philox_normal2(P.seed, id, (uint32_t)tsu, ST_FDT, 1, z + 2);
      }
    }
    double amp = fdt ? sqrt(2.0 * sp.kT * temp) : 0.0;
    double xn[3];
    if (fdt || (!S.heun && P.max_disp > 0.0)) {
      double dx[3] = {0, 0, 0}, n2 = 0.0;
      for (int d = 0; d < D; ++d) {
        dx[d] = fdt ? temp * F[e][d] + amp * z[d] : temp * F[e][d];
        n2 += dx[d] * dx[d];
      }
      if (P.max_disp > 0.0 && n2 > P.max_disp * P.max_disp) {  // relaxation extension
        double f = P.max_disp / sqrt(n2);
        for (int d = 0; d < D; ++d) dx[d] *= f;
      }
      for (int d = 0; d < D; ++d) xn[d] = box_update_dr1(P.box, d, pos[e][d], dx[d]);
    } else {
      for (int d = 0; d < D; ++d) xn[d] = box_update1(P.box, d, pos[e][d], F[e][d], temp);
    }
    for (int d = 0; d < D; ++d) {
      S.pos[(size_t)d * S.P + p] = xn[d];
      // Heun leaves the velocities alone, except for the zero-velocity kick of a single disc
      // (MD.cu:609-614), which the reference writes into the velocity array
      if (!S.heun) S.vel[(size_t)d * S.P + p] = F[e][d] * mu;
      else if (nE == 1 && need_vel) S.vel[(size_t)d * S.P + p] = vel[e][d];
    }
    if (emit_keys == 1) {
      bool finite = true;
      for (int d = 0; d < D; ++d) finite = finite && isfinite(xn[d]);
      if (!finite) {
        atomicOr(S.err, (unsigned)ERR_NONFINITE);
        for (int d = 0; d < D; ++d) xn[d] = 0.0;
      }
      int b = bucket_key(P, (int)p, s, xn);
      S.key[p] = b;
      S.rank[p] = atomicAdd(&S.count[b], 1);
    } else if (emit_keys == 2) {
      list_track<D>(P, S, p, xn, sig[e], next_pass_flag(S, tsu));
    }
  }
}
// Register budget: six waves per SIMD for single-disc cells (no spills); two-disc cells keep
// the compiler's own allocation.
template <typename T, int DIM, int NE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NE == 1 ? 6 : 1)))
k_update(const StepParams* __restrict__ Pp, Dev S, int materialized,
                                                int emit_keys) {
  k_update_body<T, DIM, NE>(Pp, S, materialized, emit_keys, (int)blockIdx.x);
}

// Forces at the current state into the canonical FP64 frc array (for download / parity).
template <typename T, int DIM, int NE>
__global__ void k_materialize(const StepParams* __restrict__ Pp, Dev S) {
  const StepParams& P = *Pp;
  int cslot = blockIdx.x * blockDim.x + threadIdx.x;
  if (cslot >= S.C) return;
  int s, c;
  if (!cell_active(P, S.ncells, cslot, &s, &c)) return;
  constexpr int D = DIM, nE = NE;
  double pos[2][3], vel[2][3], sig[2], F[2][3], pc;
  for (int e = 0; e < nE; ++e) {
    size_t p = (size_t)cslot * nE + e;
    sig[e] = S.sigma[p];
    for (int d = 0; d < D; ++d) {
      pos[e][d] = S.pos[(size_t)d * S.P + p];
      vel[e][d] = S.vel[(size_t)d * S.P + p];
    }
  }
  bool pos1_changed;
  cell_total_force<T, DIM, NE>(P, S, s, c, cslot, false, pos, vel, sig, F, &pc, &pos1_changed);
  for (int e = 0; e < nE; ++e)
    for (int d = 0; d < D; ++d) S.frc[(size_t)d * S.P + (size_t)cslot * nE + e] = F[e][d];
  if (pos1_changed)
    for (int d = 0; d < D; ++d) S.pos[(size_t)d * S.P + (size_t)cslot * nE + 1] = pos[1][d];
  if (nE == 1 && P.sp[s].motility[0] != 0.0)
    for (int d = 0; d < D; ++d) S.vel[(size_t)d * S.P + (size_t)cslot] = vel[0][d];
  S.pressure[cslot] = P.pressure ? pc : 0.0;
}

// ======================================================================================
// Pair-set hook: exact FP64 law predicate over the stencil, emits UID pairs (i<j).
// ======================================================================================
__device__ inline int slot_to_uid_particle(const StepParams& P, const int* ncells, int pidx) {
  int cslot = pidx / P.nE, e = pidx - cslot * P.nE;
  int s = species_of_cell(P, cslot);
  int cum = 0;
  for (int k = 0; k < s; ++k) cum += ncells[k];
  return (cum + cslot - P.base[s]) * P.nE + e;
}

__global__ void k_pairs_list(const StepParams* __restrict__ Pp, Dev S, int reach_x, int reach_y, int reach_z,
                             int* __restrict__ out, unsigned long long* __restrict__ counter, long cap) {
  const StepParams& P = *Pp;
  const int D = P.box.dim;
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  int total = S.start[P.nbuckets];
  if (k >= total) return;
  int meta = S.smeta[k];
  int pidx = meta >> 3, tp = meta & 7;
  int cell_p = pidx / P.nE;
  double xp[3];
  for (int d = 0; d < D; ++d) xp[d] = S.pos[(size_t)d * S.P + pidx];
  double sp_ = S.sigma[pidx];
  int bc[3];
  unpack_bin(P, S.sbin[k], bc);
  int reach[3] = {reach_x, reach_y, D == 3 ? reach_z : 0};
  for (int tq = 0; tq < P.ntypes; ++tq) {
    const LawEntry& E = P.law[tp][tq];
    if (E.law == CIL_LAW_NONE) continue;
    for (int ox = -reach[0]; ox <= reach[0]; ++ox)
      for (int oy = -reach[1]; oy <= reach[1]; ++oy)
        for (int oz = -reach[2]; oz <= reach[2]; ++oz) {
          int o[3] = {ox, oy, oz}, nc[3] = {0, 0, 0};
          bool ok = true;
          for (int d = 0; d < D; ++d) {
            int v = bc[d] + o[d];
            if (P.box.pbc[d]) v = ((v % P.ns[d]) + P.ns[d]) % P.ns[d];
            else if (v < 0 || v >= P.ns[d]) ok = false;
            nc[d] = v;
          }
          if (!ok) continue;
          // skip duplicate visits of a bin when the stencil wraps around a small periodic grid
          bool dup = false;
          for (int d = 0; d < D && !dup; ++d) {
            if (P.box.pbc[d] && 2 * reach[d] + 1 > P.ns[d]) {
              for (int t = -reach[d]; t < o[d]; ++t)
                if (((bc[d] + t) % P.ns[d] + P.ns[d]) % P.ns[d] == nc[d]) dup = true;
            }
          }
          if (dup) continue;
          int nb = bucket_of(P, tq, nc[0], nc[1], nc[2]);
          for (int j = S.start[nb]; j < S.start[nb + 1]; ++j) {
            int qidx = S.smeta[j] >> 3;
            if (qidx <= pidx) continue;
            if (qidx / P.nE == cell_p) continue;
            double xq[3], rij[3];
            for (int d = 0; d < D; ++d) xq[d] = S.pos[(size_t)d * S.P + qidx];
            double r2 = box_dist(P.box, rij, xp, xq);
            double sq_ = S.sigma[qidx];
            bool hit = false;
            if (E.law == CIL_LAW_LJ) {
              double sij2 = sq(0.5 * (sp_ + sq_));
              hit = r2 <= sij2 * E.rcut2;
            } else if (E.law == CIL_LAW_SC) {
              double sigCore = 0.5 * (sp_ + sq_);
              hit = r2 > E.rcut2 * sq(sigCore - E.xi) && r2 <= E.rcut2 * sq(sigCore + E.xi);
            } else {
              hit = sqrt(r2) <= (sp_ + sq_) / 2;
            }
            if (hit) {
              unsigned long long slot = atomicAdd(counter, 1ull);
              if ((long)slot < cap) {
                out[2 * slot] = slot_to_uid_particle(P, S.ncells, pidx);
                out[2 * slot + 1] = slot_to_uid_particle(P, S.ncells, qidx);
              }
            }
          }
        }
  }
}

#include "cil_division.inc"
#include "cil_dsfmt.inc"
#include "cil_surrogate.inc"
#include "cil_context.inc"
#include "cil_batch.inc"
#include "cil_surrogate_host.inc"
#include "cil_noalloc.inc"
#include "cil_config.inc"
#include "cil_output.inc"
#include "cil_h5.inc"
