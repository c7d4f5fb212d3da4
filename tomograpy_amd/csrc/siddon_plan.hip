// Brick-privatised backprojection.
//
// A scatter of d * w into the cube with one global atomic per Siddon step is
// bound by atomic throughput.  Here the cube is cut into bricks that fit in
// LDS.  A one-time, value-independent plan traces every ray once and records,
// at each brick the ray enters, the traversal state there (the voxel's index
// inside the brick, the three distances D) and the number of steps the ray
// takes in the brick (the Sun obstacle included), grouped by brick.  A
// backprojection then runs one workgroup per chunk of a brick's records: each
// lane restores a record, takes its steps with exactly the operations of the
// continuous traversal (so every length d is the one the reference computes),
// adds d * w into the LDS brick, and the workgroup flushes the brick with one
// coalesced atomic per touched voxel.
//
// Plan passes: (1) count records per brick (LDS histogram per workgroup,
// flushed with one global atomic per touched brick); counts are copied to the
// host, which computes brick offsets and chunk boundaries; (2) fill: each
// workgroup counts again, reserves a contiguous range per brick with one
// global atomic, and writes its records (SoA) into it.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "siddon_common.h"
#include "siddon_host.h"

namespace tg {

template <typename T>
int make_args(const tg_call *c, int t0, int t1, KArgs<T> &a, int tile_j = 0);

namespace {

constexpr int kPlanBlock = 256;
// LDS one backprojection workgroup may allocate (gfx950: 160 KiB per CU)
constexpr int kPlanLdsBytes = 160 * 1024;
constexpr int kHistMax = 8192;  // bricks x bins histogrammed in LDS; above: global atomics
constexpr int kBpjBlock = 1024;
// chunks (kBpjBlock-lane workgroups) one dispatch holds: 2^32 - 1 work-items
constexpr unsigned long long kMaxChunksPerLaunch = 0xffffffffull / kBpjBlock;
static_assert(kBpjBlock == 1024, "the plan sort keys hold a 10-bit pass position");
constexpr size_t kPlanQueueHeader = 256;
// Default pixels per row of the ray tiles traced by the plan build: one image
// row of 256 pixels per workgroup, so that the 64 records a wavefront of the
// backprojection walks together come from 64 adjacent pixels of a row --
// nearly parallel rays with alike walks through the brick (C2: 8.2 ms against
// 9.7 ms with 8 x 32 tiles and 15.1 ms with 32 x 8).
constexpr int kPlanTraceTile = 256;

struct BrickGeom {
  int32_t b[3];       // brick edge per axis
  int32_t lg[3];      // log2 of the edges
  int32_t nb[3];      // bricks per axis
  int32_t nbricks;    // nb0 * nb1 * nb2 * nt
  int32_t nt;
  int32_t chunk;      // records per chunk
  int32_t shuffle;    // records of a brick stored in a permuted order
  int32_t sort;       // records of each workgroup pass sorted by their step counts
  int32_t spread;     // low bits of the pass position reversed to order ties
  int32_t canon;      // order 3: each brick's records in ray order (view, row, column),
                      // cells unpadded, for the deterministic backprojection
  int32_t ks[2];      // brick key shifts of axes 0 and 1 (see brick_key)
  int32_t nkeys;      // brick keys per time bin: nb0 << ks0
  int32_t keyed;      // nkeys * nt <= kHistMax: per-workgroup counts in LDS, indexed by key
  int32_t direct;     // fill: LDS cursors hold record positions (records x max(4, sizeof(T))
                      // bytes < 2^32, no shuffle)
};

// A record's ray id is the ray's linear index (t n1 + i) n2 + j: the index of
// its ray-table entry, and of its pixel in a view-major image stack.
__device__ __forceinline__ uint32_t ray_id(uint32_t n1, uint32_t n2, int t, int i, int j) {
  return ((uint32_t)t * n1 + (uint32_t)i) * n2 + (uint32_t)j;
}

__device__ __forceinline__ void ray_pixel(uint32_t n1, uint32_t n2, uint32_t rid, int &t, int &i,
                                          int &j) {
  const uint32_t per_view = n1 * n2;
  t = (int)(rid / per_view);
  const uint32_t rem = rid % per_view;
  i = (int)(rem / n2);
  j = (int)(rem % n2);
}

// Per-ray constants of C:750-774 (|p| = |CDELT / u| and the step direction
// of each axis), written by the plan build with the arithmetic of ray_setup
// so that the backprojection need not recompute the unit vector and three
// divisions for every record.  up = (up0 + 1) | (up1 + 1) << 2 | (up2 + 1) << 4.
template <typename T>
struct RayEntry {
  T pabs[3];
  int32_t up;  // step direction (-1, 0, 1) per axis: 2-bit two's complement at bits 0, 2, 4
};

__host__ __device__ __forceinline__ int32_t pack_up(const int up[3]) {
  return (up[0] & 3) | ((up[1] & 3) << 2) | ((up[2] & 3) << 4);
}
// Axis k's direction: one signed bitfield extract.
__host__ __device__ __forceinline__ int up_axis(int32_t up, int k) {
  return (int32_t)((uint32_t)up << (30 - 2 * k)) >> 30;
}

// Records, structure of arrays: ray id, lq = (voxel index inside the brick) |
// (steps in the brick) << 16, and the distances D at the brick entry.
template <typename T>
struct RecPtrs {
  uint32_t *rid;
  uint32_t *lq;
  T *d0, *d1, *d2;
  RayEntry<T> *ray;  // indexed by (t * n1 + i) * n2 + j
};


// LDS layout of a brick's accumulators: axis 2 fastest, each axis-2 row
// padded by one cell and each axis-1 plane by one more.  With power-of-two
// pitches, voxels that differ only along axis 0 or 1 share an LDS bank, and
// the neighbouring rays of a wavefront hit exactly such voxels together.
// A canonical plan (order 3) keeps the cells unpadded: its backprojection sums
// each cell serially, so banks do not matter, and LDS is short.
__host__ __device__ __forceinline__ int brick_pitch1(const BrickGeom &g) {
  return g.canon ? g.b[2] : g.b[2] + 1;
}
__host__ __device__ __forceinline__ int brick_pitch0(const BrickGeom &g) {
  return g.canon ? g.b[1] * g.b[2] : g.b[1] * (g.b[2] + 1) + 1;
}
__host__ __device__ __forceinline__ long long brick_lds_cells(const BrickGeom &g) {
  return (long long)g.b[0] * brick_pitch0(g);
}

// The LDS cell of voxel iv inside its brick: what a record stores in the low
// 16 bits of lq.  (24-bit multiplies: a brick's cells fit the LDS, under 2^15.)
__device__ __forceinline__ uint32_t brick_local(const BrickGeom &g, const int iv[3]) {
  return (uint32_t)(__mul24(iv[0] & (g.b[0] - 1), brick_pitch0(g)) +
                    __mul24(iv[1] & (g.b[1] - 1), brick_pitch1(g)) + (iv[2] & (g.b[2] - 1)));
}

// base[i] = v with the byte offset formed in 32 bits (the caller ensures
// i * sizeof(V) < 2^32): the store takes the scalar-base + 32-bit vector-offset
// form, with no 64-bit address arithmetic per store.
template <typename V>
__device__ __forceinline__ void store_at(V *base, uint32_t i, V v) {
  *reinterpret_cast<V *>(reinterpret_cast<char *>(base) + i * (uint32_t)sizeof(V)) = v;
}

// A brick's key: its coordinates packed as (bx << ks0) | (by << ks1) | bz,
// ks1 = ceil(log2 nb2) and ks0 = ks1 + ceil(log2 nb1), plus bin x nkeys for
// time bin `bin`.  Shifts and ors only, so a trace forms the key of its voxel
// at every step, and the per-workgroup counts in LDS are indexed by it: keys
// number the bricks of a bin with gaps where nb1 or nb2 is not a power of two.
__device__ __forceinline__ uint32_t brick_key(const BrickGeom &g, const int iv[3]) {
  return ((uint32_t)(iv[0] >> g.lg[0]) << g.ks[0]) | ((uint32_t)(iv[1] >> g.lg[1]) << g.ks[1]) |
         (uint32_t)(iv[2] >> g.lg[2]);
}
constexpr uint32_t kNoBrick = 0xffffffffu;  // no key: keys stay under 2^31

// Dynamic LDS of the count and fill passes: one word per brick key and bin.
inline size_t hist_lds_bytes(const BrickGeom &g) {
  return g.keyed ? 4 * (size_t)g.nkeys * (size_t)g.nt : 0;
}

// The brick index (axis 2 fastest, then axes 1 and 0, then bins) of a key of
// bin `bin`, and the key of a brick index.
__host__ __device__ __forceinline__ int brick_of_key(const BrickGeom &g, uint32_t key, int bin) {
  const uint32_t k = key - (uint32_t)bin * (uint32_t)g.nkeys;
  const int bx = (int)(k >> g.ks[0]);
  const int by = (int)((k >> g.ks[1]) & ((1u << (g.ks[0] - g.ks[1])) - 1u));
  const int bz = (int)(k & ((1u << g.ks[1]) - 1u));
  return ((bx * g.nb[1] + by) * g.nb[2] + bz) + bin * (g.nb[0] * g.nb[1] * g.nb[2]);
}
__host__ __device__ __forceinline__ uint32_t key_of_brick(const BrickGeom &g, int b) {
  const int nsp = g.nb[0] * g.nb[1] * g.nb[2];
  const int bin = b / nsp, r = b - bin * nsp;
  const int bz = r % g.nb[2], by = (r / g.nb[2]) % g.nb[1], bx = r / (g.nb[2] * g.nb[1]);
  return (((uint32_t)bx << g.ks[0]) | ((uint32_t)by << g.ks[1]) | (uint32_t)bz) +
         (uint32_t)bin * (uint32_t)g.nkeys;
}

// Traces the ray of (tile, lane); calls emit(key, bin, ray, id, first) at the
// first step in every brick (key: brick_key + bin x nkeys) and close(steps)
// when the ray leaves that brick
// (or the map, or meets the Sun) with the number of steps taken in it.  Same
// traversal as the projector kernels.  With per_ray (not null) it also writes
// the ray's total step count there, indexed (t * n1 + i) * n2 + j as
// tg_count_steps writes it (0 for a masked ray or one that misses the map).
template <typename T, bool SUN, class F, class G>
__device__ __forceinline__ void trace_bricks(const KArgs<T> &a, const BrickGeom &g, uint32_t tile,
                                             F &&emit, G &&close, uint32_t *per_ray = nullptr) {
  int i, j, t;
  if (!lane_pixel(a, tile, i, j, t)) return;
  uint32_t *out = per_ray ? per_ray + ((size_t)t * a.n1 + i) * a.n2 + j : nullptr;
  if (out) *out = 0u;
  if (a.pixmask && a.pixmask[i * a.ms[0] + j * a.ms[1] + (int64_t)t * a.ms[2]] != 0) return;
  const ViewHdr<T> &h = a.views[t];
  T u0[3];
  pixel_unit_vector(a, h, t, i, j, u0);
  Ray<T> r;
  if (!ray_setup(a, h.M, u0, r)) return;
  if (!ray_state_ok(r)) {
    flag_status(a.status, TG_EGEOMETRY);
    return;
  }
  int bin = 0;
  if (a.kind == TG_BPJT) {
    bin = a.tindex ? a.tindex[t] : t;
    if (bin < 0 || bin >= a.nt) {
      flag_status(a.status, TG_EINVAL);
      return;
    }
  }
  const uint32_t rid = ray_id((uint32_t)a.n1, (uint32_t)a.n2, t, i, j);
  const uint32_t kboff = (uint32_t)bin * (uint32_t)g.nkeys;
  uint32_t last = kNoBrick;
  uint32_t seg = 0;
  if (a.ray_steps) {
    // the ray's step count from the count pass, which walked this same
    // traversal with the map and obstacle tests: the walk needs neither
    const uint32_t n = a.ray_steps[rid];
    for (uint32_t k = 0; k < n; k++) {
      const uint32_t key = brick_key(g, r.iv);
      if (key != last) {
        if (last != kNoBrick) close(seg);
        emit(key + kboff, bin, r, rid, last == kNoBrick);
        last = key;
        seg = 0;
      }
      bool sx, sy, sz;
      step_min(r, sx, sy, sz);
      seg++;
    }
    if (last != kNoBrick) close(seg);
    if (out) *out = n;
    return;
  }
  const int guard0 = (int)(a.n[0] + a.n[1] + a.n[2]) + 2;
  int guard = guard0;  // also counts the steps: guard0 - guard of them
  while (in_map(a, r) && (!SUN || outside_sun(h.M, u0, r.ac))) {
    if (--guard < 0) {
      flag_status(a.status, TG_EGEOMETRY);
      break;
    }
    const uint32_t key = brick_key(g, r.iv);
    if (key != last) {
      if (last != kNoBrick) close(seg);
      emit(key + kboff, bin, r, rid, last == kNoBrick);
      last = key;
      seg = 0;
    }
    bool sx, sy, sz;
    step_min(r, sx, sy, sz);
    seg++;
  }
  if (last != kNoBrick) close(seg);
  if (out) *out = (uint32_t)(guard0 - (guard < 0 ? 0 : guard));
}

// Per-workgroup record counts the count pass hands to the fill pass, so that
// the fill traces every ray once: for workgroup w of the plan's trace grid,
// a header word (the number of bricks its rays enter, or kPairsOverflow) and
// up to kWgPairs (brick, records) pairs.
constexpr int kWgPairs = 255;
constexpr uint32_t kPairsOverflow = 0xffffffffu;
constexpr size_t kWgPairBytes = 8 * (size_t)(kWgPairs + 1);

// pairs: the workgroup's kWgPairBytes slot of per-workgroup pairs (or null);
// per_ray: per-ray step counts out (or null); rowmat (a canonical plan, which
// needs g.keyed): this launch's rows of the tile x brick count matrix, the
// workgroup writing its tile's.  Dynamic LDS: hist_lds_bytes(g), one word per
// brick key (none without g.keyed), so that a geometry with few bricks leaves
// the CU's LDS to more workgroups.
template <typename T, bool SUN>
__global__ __launch_bounds__(kPlanBlock) void k_plan_count(KArgs<T> a, BrickGeom g,
                                                           uint32_t *counts, uint32_t *pairs,
                                                           uint32_t *per_ray, uint32_t *rowmat) {
  extern __shared__ uint32_t hist[];
  __shared__ uint32_t npairs;
  const bool lds = g.keyed;
  const int nkeys = g.nkeys * g.nt;
  if (lds) {
    for (int k = threadIdx.x; k < nkeys; k += blockDim.x) hist[k] = 0;
    if (threadIdx.x == 0) npairs = 0;
    __syncthreads();
  }
  const uint32_t tile = xcd_remap(blockIdx.x, gridDim.x);
  if (pairs) pairs += (size_t)blockIdx.x * (kWgPairBytes / 4);
  trace_bricks<T, SUN>(
      a, g, tile,
      [&](uint32_t key, int bin, const Ray<T> &, uint32_t, bool) {
        if (lds) atomicAdd(&hist[key], 1u);
        else atomicAdd(&counts[brick_of_key(g, key, bin)], 1u);
      },
      [](uint32_t) {}, per_ray);
  if (lds) {
    __syncthreads();
    for (int k = threadIdx.x; k < nkeys; k += blockDim.x) {
      const uint32_t h = hist[k];
      if (!h) continue;
      const int b = brick_of_key(g, (uint32_t)k, g.nt > 1 ? k / g.nkeys : 0);
      atomicAdd(&counts[b], h);
      if (pairs) {
        const uint32_t q = atomicAdd(&npairs, 1u);
        if (q < (uint32_t)kWgPairs) {
          pairs[2 + 2 * q] = (uint32_t)b;
          pairs[3 + 2 * q] = h;
        }
      }
    }
    __syncthreads();
    if (pairs && threadIdx.x == 0) pairs[0] = npairs <= (uint32_t)kWgPairs ? npairs : kPairsOverflow;
    if (rowmat) {
      // every brick's count, through its key (keys past a bin's bricks along
      // axes 1 and 2 name no brick)
      uint32_t *row = rowmat + (size_t)tile * (size_t)g.nbricks;
      for (int k = threadIdx.x; k < nkeys; k += blockDim.x) {
        const int bin = g.nt > 1 ? k / g.nkeys : 0;
        const uint32_t c = (uint32_t)(k - bin * g.nkeys);
        if (((c >> g.ks[1]) & ((1u << (g.ks[0] - g.ks[1])) - 1u)) >= (uint32_t)g.nb[1] ||
            (c & ((1u << g.ks[1]) - 1u)) >= (uint32_t)g.nb[2])
          continue;
        row[brick_of_key(g, (uint32_t)k, bin)] = hist[k];
      }
    }
  } else if (pairs && threadIdx.x == 0) {
    pairs[0] = kPairsOverflow;
  }
}

// ---- canonical plans (order 3): every brick's records in ray order ----------
// The count pass leaves a matrix of record counts, one row per trace tile
// (256 pixels of one image row: tiles in (view, row, column) order) and one
// column per brick.  An exclusive scan down each column gives every tile the
// position of its first record in each brick, and inside a tile a lane's
// record goes after those of the lanes (pixels) before it that enter the same
// brick: per-brick lane masks in LDS.  The brick's records are then in the
// order of their ray ids, the order in which the reference's single-threaded
// run adds into the brick's voxels (C:489-494 with nthread = 1).
constexpr int kDetMaxBricks = 2048;  // 256-bit lane masks of all bricks in LDS (64 KiB)
constexpr int kDetSegRows = 256;     // rows per segment of the column scan
constexpr int kDetMinSlots = 256;    // fewest contribution slots of one deterministic batch
// The canonical plan's count is the plan count with the tile's row of the
// count matrix; its per-workgroup pairs (the bricks a tile's rays enter, up to
// kWgPairs; a tile of C2's geometry enters up to about 180 bricks of
// 32 x 16 x 16) name the bricks whose lane masks the fill builds.

// Segment sums of the columns: block (s, cb) sums rows [s * 256, (s + 1) * 256)
// of columns [cb * 256, cb * 256 + 256) into seg[s * nb + col].
__global__ __launch_bounds__(256) void k_det_colsum(const uint32_t *__restrict__ mat, int64_t rows,
                                                    int nb, uint32_t *__restrict__ seg) {
  const int col = blockIdx.y * 256 + threadIdx.x;
  if (col >= nb) return;
  const int64_t r0 = (int64_t)blockIdx.x * kDetSegRows;
  const int64_t r1 = r0 + kDetSegRows < rows ? r0 + kDetSegRows : rows;
  uint32_t sum = 0;
  for (int64_t r = r0; r < r1; r++) sum += mat[r * nb + col];
  seg[(int64_t)blockIdx.x * nb + col] = sum;
}

// Exclusive scan of each column's segment sums (one thread per column).
__global__ __launch_bounds__(256) void k_det_segscan(uint32_t *__restrict__ seg, int64_t nseg,
                                                     int nb) {
  const int col = blockIdx.x * 256 + threadIdx.x;
  if (col >= nb) return;
  uint32_t run = 0;
  for (int64_t q = 0; q < nseg; q++) {
    const uint32_t v = seg[q * nb + col];
    seg[q * nb + col] = run;
    run += v;
  }
}

// mat[r][col] = records of brick col from the tiles before r (exclusive scan).
__global__ __launch_bounds__(256) void k_det_colapply(uint32_t *__restrict__ mat, int64_t rows,
                                                      int nb, const uint32_t *__restrict__ seg) {
  const int col = blockIdx.y * 256 + threadIdx.x;
  if (col >= nb) return;
  const int64_t r0 = (int64_t)blockIdx.x * kDetSegRows;
  const int64_t r1 = r0 + kDetSegRows < rows ? r0 + kDetSegRows : rows;
  uint32_t run = seg[(int64_t)blockIdx.x * nb + col];
  for (int64_t r = r0; r < r1; r++) {
    const uint32_t v = mat[r * nb + col];
    mat[r * nb + col] = run;
    run += v;
  }
}

// Fill of a canonical plan, for the tiles whose count pass could not list
// their bricks: each brick's lane mask from a first trace, then a trace that
// writes each record at the brick's offset + the tile's scanned count + the
// number of earlier lanes in the mask.  With g.direct the first two are summed
// per brick into LDS before the trace (record positions are under 2^32), so a
// record's position takes no global load.
template <typename T, bool SUN>
__global__ __launch_bounds__(kPlanBlock) void k_plan_fill_canon(KArgs<T> a, BrickGeom g,
                                                                const unsigned long long *offsets,
                                                                const uint32_t *rowmat,
                                                                const uint32_t *pairs,
                                                                RecPtrs<T> rec) {
  __shared__ __align__(16) uint32_t mask[kDetMaxBricks * (kPlanBlock / 32)];
  __shared__ uint32_t base[kDetMaxBricks];
  const uint32_t tile = xcd_remap(blockIdx.x, gridDim.x);
  // tiles whose count pass listed their bricks are filled by k_plan_fill_canon_slots
  if (pairs[(size_t)blockIdx.x * (kWgPairBytes / 4)] != kPairsOverflow) return;
  const int words = g.nbricks * (kPlanBlock / 32);
  for (int k = threadIdx.x; k < words; k += blockDim.x) mask[k] = 0;
  __syncthreads();
  const uint32_t lane = threadIdx.x;
  const uint32_t bit = 1u << (lane & 31);
  const uint32_t *row = rowmat + (size_t)tile * (size_t)g.nbricks;
  trace_bricks<T, SUN>(
      a, g, tile,
      [&](uint32_t key, int bin, const Ray<T> &, uint32_t, bool) {
        atomicOr(&mask[brick_of_key(g, key, bin) * 8 + (lane >> 5)], bit);
      },
      [](uint32_t) {});
  if (g.direct)
    for (int b = threadIdx.x; b < g.nbricks; b += blockDim.x)
      base[b] = (uint32_t)(offsets[b] + row[b]);
  __syncthreads();
  unsigned long long pcur = 0;
  uint32_t lcur = 0;
  trace_bricks<T, SUN>(a, g, tile, [&](uint32_t key, int bin, const Ray<T> &r, uint32_t rid,
                                       bool first) {
    if (first) {
      RayEntry<T> e;
      e.pabs[0] = r.pabs[0];
      e.pabs[1] = r.pabs[1];
      e.pabs[2] = r.pabs[2];
      e.up = pack_up(r.up);
      rec.ray[rid] = e;
    }
    const int b = brick_of_key(g, key, bin);
    // the lanes before this one that enter brick b: whole 64-bit mask words
    // of the earlier wavefronts (a count uniform per wavefront), then this
    // wavefront's word below the lane
    const unsigned long long *m = reinterpret_cast<const unsigned long long *>(&mask[b * 8]);
    const uint32_t wv = lane >> 6;
    uint32_t rank = 0;
    for (uint32_t w = 0; w < wv; w++) rank += (uint32_t)__popcll(m[w]);
    rank += (uint32_t)__popcll(m[wv] & ((1ull << (lane & 63)) - 1ull));
    const unsigned long long p =
        g.direct ? (unsigned long long)(base[b] + rank) : offsets[b] + row[b] + rank;
    rec.rid[p] = rid;
    rec.d0[p] = r.D[0];
    rec.d1[p] = r.D[1];
    rec.d2[p] = r.D[2];
    pcur = p;
    lcur = brick_local(g, r.iv);
  }, [&](uint32_t steps) { rec.lq[pcur] = lcur | (steps << 16); });
}

// Fill of a canonical plan for the tiles whose count pass listed their bricks:
// one LDS slot per listed brick (brick -> list position, its lane mask and its
// first record position: about 18 KiB instead of a mask per brick of the
// map), a first trace that sets each lane's bit in the masks of the bricks it
// enters, then a trace that writes each record at its brick's position + the
// number of earlier lanes in the brick's mask.  Dynamic LDS: 4 bytes per brick.
constexpr uint32_t kNoSlot = 0xffffffffu;
constexpr unsigned long long kNoRecord = ~0ull;

template <typename T, bool SUN>
__global__ __launch_bounds__(kPlanBlock) void k_plan_fill_canon_slots(
    KArgs<T> a, BrickGeom g, const unsigned long long *offsets, const uint32_t *rowmat,
    const uint32_t *pairs, RecPtrs<T> rec) {
  extern __shared__ uint32_t slot[];
  __shared__ __align__(16) uint32_t smask[kWgPairs * 8];
  __shared__ unsigned long long sbase[kWgPairs];
  const uint32_t tile = xcd_remap(blockIdx.x, gridDim.x);
  const uint32_t *pl = pairs + (size_t)blockIdx.x * (kWgPairBytes / 4);
  const uint32_t nl = pl[0];
  if (nl == kPairsOverflow) return;  // k_plan_fill_canon fills this tile
  const uint32_t *row = rowmat + (size_t)tile * (size_t)g.nbricks;
  for (int k = threadIdx.x; k < g.nbricks; k += blockDim.x) slot[k] = kNoSlot;
  for (uint32_t e = threadIdx.x; e < 8 * nl; e += blockDim.x) smask[e] = 0;
  __syncthreads();
  for (uint32_t q = threadIdx.x; q < nl; q += blockDim.x) {
    const uint32_t b = pl[2 + 2 * q];
    slot[b] = q;
    sbase[q] = offsets[b] + row[b];
  }
  __syncthreads();
  const uint32_t lane = threadIdx.x;
  const uint32_t bit = 1u << (lane & 31);
  trace_bricks<T, SUN>(
      a, g, tile,
      [&](uint32_t key, int bin, const Ray<T> &, uint32_t, bool) {
        const uint32_t q = slot[brick_of_key(g, key, bin)];
        if (q < nl) atomicOr(&smask[q * 8 + (lane >> 5)], bit);
      },
      [](uint32_t) {});
  __syncthreads();
  unsigned long long pcur = kNoRecord;
  uint32_t lcur = 0;
  trace_bricks<T, SUN>(a, g, tile, [&](uint32_t key, int bin, const Ray<T> &r, uint32_t rid,
                                       bool first) {
    if (first) {
      RayEntry<T> e;
      e.pabs[0] = r.pabs[0];
      e.pabs[1] = r.pabs[1];
      e.pabs[2] = r.pabs[2];
      e.up = pack_up(r.up);
      rec.ray[rid] = e;
    }
    const uint32_t q = slot[brick_of_key(g, key, bin)];
    if (q >= nl) {
      // a brick the count pass did not list: never for the same traversal;
      // flagged, and nothing is written
      flag_status(a.status, TG_EGEOMETRY);
      pcur = kNoRecord;
      return;
    }
    // the lanes before this one that enter the brick: whole 64-bit mask words
    // of the earlier wavefronts (a count uniform per wavefront), then this
    // wavefront's word below the lane
    const unsigned long long *m = reinterpret_cast<const unsigned long long *>(&smask[q * 8]);
    const uint32_t wv = lane >> 6;
    uint32_t rank = 0;
    for (uint32_t w = 0; w < wv; w++) rank += (uint32_t)__popcll(m[w]);
    rank += (uint32_t)__popcll(m[wv] & ((1ull << (lane & 63)) - 1ull));
    const unsigned long long p = sbase[q] + rank;
    rec.rid[p] = rid;
    rec.d0[p] = r.D[0];
    rec.d1[p] = r.D[1];
    rec.d2[p] = r.D[2];
    pcur = p;
    lcur = brick_local(g, r.iv);
  }, [&](uint32_t steps) {
    if (pcur != kNoRecord) rec.lq[pcur] = lcur | (steps << 16);
  });
}

// pairs: this workgroup's slot of the count pass's pairs (or null): when it
// holds the workgroup's per-brick record counts, the ranges are reserved from
// them and the rays are traced once; otherwise they are counted again here.
// With g.direct the LDS cursor of a brick key holds the position of the
// workgroup's next record in the plan (its range's start + the records
// written), so a record is placed by one LDS atomic, without the brick index
// and its offset.
template <typename T, bool SUN>
__global__ __launch_bounds__(kPlanBlock) void k_plan_fill(KArgs<T> a, BrickGeom g,
                                                          const unsigned long long *offsets,
                                                          uint32_t *cursor, const uint32_t *mult,
                                                          RecPtrs<T> rec, const uint32_t *pairs) {
  extern __shared__ uint32_t slot[];  // hist_lds_bytes(g)
  const bool lds = g.keyed;
  const bool direct = lds && g.direct;
  const int nkeys = g.nkeys * g.nt;
  const uint32_t tile = xcd_remap(blockIdx.x, gridDim.x);
  if (pairs) pairs += (size_t)blockIdx.x * (kWgPairBytes / 4);
  if (lds) {
    for (int k = threadIdx.x; k < nkeys; k += blockDim.x) slot[k] = 0;
    __syncthreads();
    const uint32_t np = pairs ? pairs[0] : kPairsOverflow;
    if (np != kPairsOverflow) {
      // reserve one contiguous range per brick for this workgroup's records
      for (uint32_t q = threadIdx.x; q < np; q += blockDim.x) {
        const uint32_t b = pairs[2 + 2 * q];
        const uint32_t start = atomicAdd(&cursor[b], pairs[3 + 2 * q]);
        slot[key_of_brick(g, (int)b)] = direct ? (uint32_t)(offsets[b] + start) : start;
      }
    } else {
      trace_bricks<T, SUN>(
          a, g, tile,
          [&](uint32_t key, int, const Ray<T> &, uint32_t, bool) { atomicAdd(&slot[key], 1u); },
          [](uint32_t) {});
      __syncthreads();
      for (int k = threadIdx.x; k < nkeys; k += blockDim.x) {
        if (!slot[k]) continue;
        const int b = brick_of_key(g, (uint32_t)k, g.nt > 1 ? k / g.nkeys : 0);
        const uint32_t start = atomicAdd(&cursor[b], slot[k]);
        slot[k] = direct ? (uint32_t)(offsets[b] + start) : start;
      }
    }
    __syncthreads();
  }
  unsigned long long pcur = 0;
  uint32_t lcur = 0;
  trace_bricks<T, SUN>(a, g, tile, [&](uint32_t key, int bin, const Ray<T> &r, uint32_t rid,
                                       bool first) {
    if (first) {
      RayEntry<T> e;
      e.pabs[0] = r.pabs[0];
      e.pabs[1] = r.pabs[1];
      e.pabs[2] = r.pabs[2];
      e.up = pack_up(r.up);
      rec.ray[rid] = e;
    }
    unsigned long long p;
    if (direct) {
      const uint32_t q = atomicAdd(&slot[key], 1u);
      store_at(rec.rid, q, rid);
      store_at(rec.d0, q, r.D[0]);
      store_at(rec.d1, q, r.D[1]);
      store_at(rec.d2, q, r.D[2]);
      p = q;
    } else {
      const int b = brick_of_key(g, key, bin);
      const uint32_t s = lds ? atomicAdd(&slot[key], 1u) : atomicAdd(&cursor[b], 1u);
      unsigned long long q = s;
      if (g.shuffle) {  // a bijection of [0, count): mult[b] is coprime with the count
        const unsigned long long cnt = offsets[b + 1] - offsets[b];
        q = (q * mult[b]) % cnt;
      }
      p = offsets[b] + q;
      rec.rid[p] = rid;
      rec.d0[p] = r.D[0];
      rec.d1[p] = r.D[1];
      rec.d2[p] = r.D[2];
    }
    pcur = p;
    lcur = brick_local(g, r.iv);
  }, [&](uint32_t steps) {
    if (direct) store_at(rec.lq, (uint32_t)pcur, lcur | (steps << 16));
    else rec.lq[pcur] = lcur | (steps << 16);
  });
}

// The brick is accumulated in double for both element types: on gfx950 the
// LDS float64 atomic add runs about 2.6x faster than the float32 one, and a
// wider partial sum only reduces rounding.  Each contribution d * w is still
// formed in T exactly as the reference forms it.
using Acc = double;

// Chunk c -> its brick b and record range [r0, r1) (uniform per workgroup).
__device__ __forceinline__ int chunk_range(const BrickGeom &g, const unsigned long long *offsets,
                                           const unsigned long long *chunk_prefix,
                                           unsigned long long c, unsigned long long &r0,
                                           unsigned long long &r1) {
  int lo = 0, hi = g.nbricks - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (chunk_prefix[mid] <= c) lo = mid;
    else hi = mid - 1;
  }
  r0 = offsets[lo] + (c - chunk_prefix[lo]) * (unsigned long long)g.chunk;
  r1 = r0 + g.chunk;
  if (r1 > offsets[lo + 1]) r1 = offsets[lo + 1];
  return lo;
}

// Plan order 2: in every pass of the backprojection workgroup over a chunk
// (kBpjBlock consecutive records, one per lane), the records are reordered by
// decreasing step count, so that each wavefront's 64 walks are about equally
// long; the reordering stays inside the pass, whose records come from nearby
// rays.  Ties are ordered by the pass position with its low `spread` bits
// reversed (tg_bpj_plan.tie_spread), so the plan is deterministic.
// x with its low s bits reversed (an involution on [0, 1024) for s <= 10).
__device__ __forceinline__ uint32_t tie_pos(uint32_t x, int s) {
  if (s <= 0) return x;
  const uint32_t m = (1u << s) - 1u;
  return (x & ~m) | (__brev(x & m) >> (32 - s));
}

// Inclusive scan of v over the workgroup (kBpjBlock lanes); wsum holds
// kBpjBlock / 64 + 1 words of LDS.
__device__ __forceinline__ uint32_t block_incl_scan(uint32_t v, uint32_t *wsum) {
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t u = __shfl_up(v, o, 64);
    if (lane >= (uint32_t)o) v += u;
  }
  if (lane == 63u) wsum[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    for (uint32_t w = 0; w < kBpjBlock / 64; w++) {
      const uint32_t x = wsum[w];
      wsum[w] = run;
      run += x;
    }
    wsum[kBpjBlock / 64] = run;
  }
  __syncthreads();
  return v + wsum[wave];
}

// Sort of the records of each pass (kBpjBlock consecutive records of a
// chunk) by decreasing step count, ties by tie_pos of the pass position:
// the ascending order of the keys ((0xffff - steps) << 10) | tie_pos(position).
// A pass whose records all take at most `top` steps (every pass: a ray visits
// at most b0 + b1 + b2 - 2 voxels of a brick, its indices moving one way per
// axis; top = b0 + b1 + b2 when that is under 256, else 255) is sorted by one
// stable counting pass over the digit top - steps in tie_pos order, which
// scans top + 2 buckets per wavefront; any other pass by a bitonic network
// over the full keys.  The two give the same order.
constexpr int kSortWaves = kBpjBlock / 64;
constexpr int kSortBuckets = 257;  // digits 0..255, and 256 for lanes past the chunk

// (at most 64 VGPRs: two workgroups per CU)
template <typename T>
__global__ __launch_bounds__(kBpjBlock) __attribute__((amdgpu_waves_per_eu(8))) void k_plan_sort(
    BrickGeom g,
                                                         const unsigned long long *offsets,
                                                         const unsigned long long *chunk_prefix,
                                                         RecPtrs<T> rec) {
  __shared__ uint32_t key[kBpjBlock];                         // keys, or each pass's source lanes
  __shared__ uint32_t hist[kSortBuckets * kSortWaves];        // (digit, wavefront) counts
  __shared__ uint32_t wsum[kSortWaves + 1];
  __shared__ uint16_t dig[kBpjBlock];
  __shared__ uint32_t s_rid[kBpjBlock], s_lq[kBpjBlock];
  __shared__ T s_d0[kBpjBlock], s_d1[kBpjBlock], s_d2[kBpjBlock];
  unsigned long long r0, r1;
  chunk_range(g, offsets, chunk_prefix, blockIdx.x, r0, r1);
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  const unsigned long long below = (1ull << lane) - 1ull;
  // the most steps a record takes in a brick, and the (digit, wavefront)
  // counters in use: digits 0..top, then top + 1 for lanes past the chunk
  const uint32_t edges = (uint32_t)(g.b[0] + g.b[1] + g.b[2]);
  const uint32_t top = edges < 256u ? edges : 255u;
  const int nhist = (int)(top + 2) * kSortWaves;
  // each lane's record of the next pass, loaded while this pass is sorted (the
  // passes are disjoint ranges, each rewritten in place)
  uint32_t n_rid = 0, n_lq = 0;
  T n_d0 = 0, n_d1 = 0, n_d2 = 0;
  if (r0 + tid < r1) {
    n_rid = rec.rid[r0 + tid];
    n_lq = rec.lq[r0 + tid];
    n_d0 = rec.d0[r0 + tid];
    n_d1 = rec.d1[r0 + tid];
    n_d2 = rec.d2[r0 + tid];
  }
  for (unsigned long long ws = r0; ws < r1; ws += kBpjBlock) {
    const unsigned long long q = ws + tid;
    const bool valid = q < r1;
    uint32_t steps = 0;
    if (valid) {
      s_rid[tid] = n_rid;
      s_lq[tid] = n_lq;
      s_d0[tid] = n_d0;
      s_d1[tid] = n_d1;
      s_d2[tid] = n_d2;
      steps = n_lq >> 16;
    }
    if (q + kBpjBlock < r1) {
      n_rid = rec.rid[q + kBpjBlock];
      n_lq = rec.lq[q + kBpjBlock];
      n_d0 = rec.d0[q + kBpjBlock];
      n_d1 = rec.d1[q + kBpjBlock];
      n_d2 = rec.d2[q + kBpjBlock];
    }
    // longest first; ties by pass position with its low g.spread bits
    // reversed: spread, the walks a wavefront takes together come from rays
    // apart in the pass rather than from adjacent pixels, which in a coarse
    // cube step through the same voxels and serialise on the same LDS cells
    const uint32_t tp = tie_pos(tid, g.spread);
    if (__syncthreads_or(valid && steps > top)) {
      key[tid] = valid ? ((0xffffu - steps) << 10) | tp : 0xffffffffu;
      __syncthreads();
      for (uint32_t k = 2; k <= kBpjBlock; k <<= 1) {  // bitonic sort, ascending
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
          const uint32_t o = tid ^ j;
          if (o > tid) {
            const uint32_t x = key[tid], y = key[o];
            const bool up = (tid & k) == 0;
            if ((x > y) == up) {
              key[tid] = y;
              key[o] = x;
            }
          }
          __syncthreads();
        }
      }
      // tie_pos is an involution: the key's position part names the source lane
      key[tid] = tie_pos(key[tid] & (kBpjBlock - 1), g.spread);
    } else {
      // lane tid's digit at its place tp in tie order
      dig[tp] = (uint16_t)(valid ? top - steps : top + 1u);
      for (int e = tid; e < nhist; e += kBpjBlock) hist[e] = 0;
      __syncthreads();
      // element v = tid of the tie order: its rank among the equal digits of
      // its wavefront (lanes agree on all 9 digit bits), and the wavefront's
      // count of that digit
      const uint32_t d = dig[tid];
      unsigned long long peers = ~0ull;
#pragma unroll
      for (int bit = 0; bit < 9; bit++) {
        const unsigned long long b = __ballot((d >> bit) & 1u);
        peers &= ((d >> bit) & 1u) ? b : ~b;
      }
      const uint32_t rank = (uint32_t)__popcll(peers & below);
      if (rank == 0) hist[d * kSortWaves + wave] = (uint32_t)__popcll(peers);
      __syncthreads();
      // exclusive scan of the counts in (digit, wavefront) order
      const int per = (nhist + kBpjBlock - 1) / kBpjBlock;
      const int e0 = (int)tid * per;
      uint32_t local = 0;
      for (int e = e0; e < e0 + per && e < nhist; e++) local += hist[e];
      uint32_t run = block_incl_scan(local, wsum) - local;
      for (int e = e0; e < e0 + per && e < nhist; e++) {
        const uint32_t c = hist[e];
        hist[e] = run;
        run += c;
      }
      __syncthreads();
      // sorted position of element v; its record is lane tie_pos(v)
      key[hist[d * kSortWaves + wave] + rank] = tie_pos(tid, g.spread);
    }
    __syncthreads();
    if (valid) {
      const uint32_t src = key[tid];
      rec.rid[q] = s_rid[src];
      rec.lq[q] = s_lq[src];
      rec.d0[q] = s_d0[src];
      rec.d1[q] = s_d1[src];
      rec.d2[q] = s_d2[src];
    }
    __syncthreads();
  }
}

// base[i] for a wavefront-uniform global base and a lane's 32-bit index: the
// base is moved to scalar registers so that the load takes the scalar base +
// 32-bit offset form, with no per-lane 64-bit address arithmetic.
template <typename V>
__device__ __forceinline__ V lane_load(const V *base, uint32_t i) {
  typedef const __attribute__((address_space(1))) V *global_ptr;
  const uint64_t v = (uint64_t)base;
  // (the builtin returns int: each half is taken as unsigned before widening)
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
  const uint64_t u = ((uint64_t)hi << 32) | (uint64_t)lo;
  return ((global_ptr)u)[i];
}

// The work of one workgroup: chunk c of one brick's records, accumulated in
// the LDS array acc and flushed to the cube (or to the float64 accumulator).
template <typename T>
__device__ __forceinline__ void plan_bpj_chunk(const KArgs<T> &a, const BrickGeom &g,
                                               const unsigned long long *offsets,
                                               const unsigned long long *chunk_prefix,
                                               const RecPtrs<T> &rec, unsigned long long c,
                                               Acc *acc) {
  const int vol = g.b[0] * g.b[1] * g.b[2];
  unsigned long long r0, r1;
  const int b = chunk_range(g, offsets, chunk_prefix, c, r0, r1);
  const int nsp = g.nb[0] * g.nb[1] * g.nb[2];
  const int bin = b / nsp;
  const int bs = b % nsp;
  const int bz = bs % g.nb[2], by = (bs / g.nb[2]) % g.nb[1], bx = bs / (g.nb[2] * g.nb[1]);
  const int lo0 = bx << g.lg[0], lo1 = by << g.lg[1], lo2 = bz << g.lg[2];
  const int cells = (int)brick_lds_cells(g);
  for (int k = threadIdx.x; k < cells; k += blockDim.x) acc[k] = (Acc)0;
  __syncthreads();
  // The record loop is uniform per wavefront (every lane takes part in the
  // ballot below); lanes past the chunk's end walk no steps.
  const uint32_t lane = threadIdx.x & 63u;
  const unsigned long long wave0 = __builtin_amdgcn_readfirstlane(threadIdx.x - lane);
  // a view-major stack holds the pixel of ray id r at r: no decoding
  const bool view_major = a.ds[1] == 1 && a.ds[0] == (int64_t)a.n2 &&
                          a.ds[2] == (int64_t)a.n1 * (int64_t)a.n2;
  for (unsigned long long pw = r0 + wave0; pw < r1; pw += blockDim.x) {
    const unsigned long long left = r1 - pw;  // records from this group's first
    const bool valid = left >= 64u || lane < (uint32_t)left;
    // the record's own fields in one round trip (a scalar base for the
    // wavefront's 64 records, the lane's constant byte offset), then the two
    // gathers that depend on its ray id (pixel value, ray constants) in a
    // second one.  Lanes past the chunk's end load nothing and walk no steps.
    uint32_t lq = 0, rid;
    T D0, D1, D2, w;
    RayEntry<T> e;
    if (valid) {
      rid = lane_load(rec.rid + pw, lane);
      lq = lane_load(rec.lq + pw, lane);
      D0 = lane_load(rec.d0 + pw, lane);
      D1 = lane_load(rec.d1 + pw, lane);
      D2 = lane_load(rec.d2 + pw, lane);
      int64_t did = rid;
      if (!view_major) {
        int t, i, j;
        ray_pixel((uint32_t)a.n1, (uint32_t)a.n2, rid, t, i, j);
        did = a.data32 ? (int64_t)((uint32_t)i * (uint32_t)a.ds[0] + (uint32_t)j * (uint32_t)a.ds[1] +
                                   (uint32_t)t * (uint32_t)a.ds[2])
                       : i * a.ds[0] + j * a.ds[1] + (int64_t)t * a.ds[2];
      }
      w = a.data[did];
      e = rec.ray[rid];
      if (a.pixw) w = a.pixw[did] * w;
    }
    // a zero pixel value walks its steps adding zeros, which keeps the
    // wavefront together and leaves every sum unchanged
    // byte offsets in the double accumulators of one step along each axis
    const int u0 = up_axis(e.up, 0), u1 = up_axis(e.up, 1), u2 = up_axis(e.up, 2);
    // (24-bit products: |pitch| x 8 < 2^23 for bricks of at most 65536 cells)
    const int dm0 = __mul24(u0, brick_pitch0(g) * (int)sizeof(Acc));
    const int dm1 = __mul24(u1, brick_pitch1(g) * (int)sizeof(Acc));
    const int dm2 = u2 * (int)sizeof(Acc);
    unsigned char *cell = (unsigned char *)(acc + (lq & 0xffffu));
    const uint32_t n = valid ? lq >> 16 : 0u;
    auto step = [&]() {
      const T d = tmin(tmin(D0, D1), D2);
      const bool sx = (D0 == d), sy = (D1 == d), sz = (D2 == d);
      D0 = sx ? e.pabs[0] : D0 - d;
      D1 = sy ? e.pabs[1] : D1 - d;
      D2 = sz ? e.pabs[2] : D2 - d;
      atomicAdd((Acc *)cell, (Acc)(d * w));
      cell += (sx ? dm0 : 0) + (sy ? dm1 : 0) + (sz ? dm2 : 0);
    };
    // Steps that every lane walks run on a scalar counter.  In a plan sorted
    // by decreasing step count the last lane holds the wavefront's fewest
    // steps, so the search starts there.
    const uint32_t common = wave_floor(n, (uint32_t)__builtin_amdgcn_readlane((int)n, 63));
    for (uint32_t k = 0; k < common; k++) step();
    for (uint32_t k = common; k < n; k++) step();
  }
  __syncthreads();
  const int64_t boff = (int64_t)bin * a.cs[3];
  for (int k = threadIdx.x; k < vol; k += blockDim.x) {
    const int l2 = k & (g.b[2] - 1), l1 = (k >> g.lg[2]) & (g.b[1] - 1), l0 = k >> (g.lg[2] + g.lg[1]);
    const Acc sum = acc[l0 * brick_pitch0(g) + l1 * brick_pitch1(g) + l2];
    if (sum == (Acc)0) continue;
    const int64_t i0 = lo0 + l0, i1 = lo1 + l1, i2 = lo2 + l2;
    if (a.voxmask && a.voxmask[i0 * a.vs[0] + i1 * a.vs[1] + i2 * a.vs[2]]) continue;
    const int64_t off = boff + i0 * a.cs[0] + i1 * a.cs[1] + i2 * a.cs[2];
    // TG_ACC_F64 keeps the brick's partial sum in float64 until the finish pass
    if (a.acc) unsafeAtomicAdd(a.acc + off, (double)sum);
    else unsafeAtomicAdd(a.cube + off, (T)sum);
  }
}

// ---- deterministic backprojection through a canonical plan ---------------
// One workgroup per brick walks the brick's records in ray order, in batches
// of up to kBpjBlock records (one per lane) whose steps fit `mcap`
// contributions.  Per batch: every lane writes its contributions d * w (in T,
// as C:923-928 forms them) at consecutive slots, numbered in (ray, step)
// order by a scan of the step counts; the slots are bucketed by cell
// (counting sort in LDS); and one lane per cell adds its bucket, sorted back
// into slot order, to the cell's running sum in T.  Each voxel therefore adds
// its contributions one at a time in the reference's single-thread order
// (view, row, column, C:489-494), starting from the cube's value, and the
// result is bit-identical to that run.


// Ascending order of 8 values: Batcher's odd-even merge network (19 exchanges).
__device__ __forceinline__ void sort8(uint32_t (&v)[8]) {
  auto cx = [&](int i, int j) {
    const uint32_t lo = v[i] < v[j] ? v[i] : v[j], hi = v[i] < v[j] ? v[j] : v[i];
    v[i] = lo;
    v[j] = hi;
  };
  cx(0, 1); cx(2, 3); cx(4, 5); cx(6, 7);
  cx(0, 2); cx(1, 3); cx(4, 6); cx(5, 7);
  cx(1, 2); cx(5, 6);
  cx(0, 4); cx(1, 5); cx(2, 6); cx(3, 7);
  cx(2, 4); cx(3, 5);
  cx(1, 2); cx(3, 4); cx(5, 6);
}

// The bucket words of the deterministic kernel are padded by one word per 8
// cells: the bucket scans read each lane's 8 consecutive cells, and with the
// padding the 64 lanes' words of one read fall in different banks instead of
// 16 to a bank.
__host__ __device__ __forceinline__ int det_cnt_index(int c) { return c + (c >> 3); }
__host__ __device__ __forceinline__ int det_cnt_words(int cells) { return cells + (cells >> 3) + 1; }

template <typename T>
__host__ __device__ __forceinline__ size_t det_lds_bytes(int cells, int mcap) {
  return (size_t)cells * sizeof(T) + (size_t)mcap * sizeof(T) + (size_t)det_cnt_words(cells) * 4 +
         4 * (kBpjBlock / 64 + 4) + (size_t)mcap * 4;
}

template <typename T>
__global__ __launch_bounds__(kBpjBlock) void k_plan_bpj_det(KArgs<T> a, BrickGeom g,
                                                            const unsigned long long *offsets,
                                                            const uint32_t *border,
                                                            RecPtrs<T> rec, int mcap) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int cells = g.b[0] * g.b[1] * g.b[2];
  T *acc = reinterpret_cast<T *>(smem_raw);
  T *uval = acc + cells;
  uint32_t *cnt = reinterpret_cast<uint32_t *>(uval + mcap);
  uint32_t *wsum = cnt + det_cnt_words(cells);  // kBpjBlock / 64 + 1 words, then the batch's counts
  uint32_t *batch = wsum + (kBpjBlock / 64 + 1);
  uint16_t *ucell = reinterpret_cast<uint16_t *>(batch + 3);
  uint16_t *sidx = ucell + mcap;
  const int b = (int)border[blockIdx.x];
  if (offsets[b] == offsets[b + 1]) return;  // no records: the brick is left as it is
  const int nsp = g.nb[0] * g.nb[1] * g.nb[2];
  const int bin = b / nsp, bs = b % nsp;
  const int bz = bs % g.nb[2], by = (bs / g.nb[2]) % g.nb[1], bx = bs / (g.nb[2] * g.nb[1]);
  const int lo0 = bx << g.lg[0], lo1 = by << g.lg[1], lo2 = bz << g.lg[2];
  const int64_t boff = (int64_t)bin * a.cs[3];
  const uint32_t tid = threadIdx.x;
  // the brick's voxels as they are: the sums continue from them (in place)
  for (int c = tid; c < cells; c += kBpjBlock) {
    const int l2 = c & (g.b[2] - 1), l1 = (c >> g.lg[2]) & (g.b[1] - 1), l0 = c >> (g.lg[2] + g.lg[1]);
    const int64_t i0 = lo0 + l0, i1 = lo1 + l1, i2 = lo2 + l2;
    const bool inside = i0 < (int64_t)a.n[0] && i1 < (int64_t)a.n[1] && i2 < (int64_t)a.n[2];
    acc[c] = inside ? a.cube[boff + i0 * a.cs[0] + i1 * a.cs[1] + i2 * a.cs[2]] : (T)0;
    cnt[det_cnt_index(c)] = 0u;
  }
  const bool view_major = a.ds[1] == 1 && a.ds[0] == (int64_t)a.n2 &&
                          a.ds[2] == (int64_t)a.n1 * (int64_t)a.n2;
  const int dm0 = brick_pitch0(g), dm1 = brick_pitch1(g);
  unsigned long long r = offsets[b];
  const unsigned long long r1 = offsets[b + 1];
  // a lane's record: streamed (ray id, lq, entry distances), then gathered by
  // ray id (pixel value, ray entry).  The next batch's records are loaded
  // while this batch is summed.
  auto stream_rec = [&](unsigned long long p, uint32_t &rid, uint32_t &lq, T &d0, T &d1, T &d2) {
    rid = rec.rid[p];
    lq = rec.lq[p];
    d0 = rec.d0[p];
    d1 = rec.d1[p];
    d2 = rec.d2[p];
  };
  auto gather_rec = [&](uint32_t rid, T &w, RayEntry<T> &e) {
    int64_t did = rid;
    if (!view_major) {
      int t, i, j;
      ray_pixel((uint32_t)a.n1, (uint32_t)a.n2, rid, t, i, j);
      did = i * a.ds[0] + j * a.ds[1] + (int64_t)t * a.ds[2];
    }
    w = a.data[did];
    if (a.pixw) w = a.pixw[did] * w;
    e = rec.ray[rid];
  };
  uint32_t rid = 0, lq = 0;
  T D0 = 0, D1 = 0, D2 = 0, w = 0;
  RayEntry<T> e;
  e.pabs[0] = e.pabs[1] = e.pabs[2] = (T)0;
  e.up = 0;
  if (r + tid < r1) {
    stream_rec(r + tid, rid, lq, D0, D1, D2);
    gather_rec(rid, w, e);
  }
  __syncthreads();
  while (r < r1) {
    const bool valid = r + tid < r1;
    // a zero pixel adds zeros: its steps are left out
    const uint32_t n = valid && w != (T)0 ? lq >> 16 : 0u;
    if (tid == 0) {
      batch[0] = 0;
      batch[1] = 0;
    }
    const uint32_t incl = block_incl_scan(n, wsum);  // synchronises: batch[] is zero
    const bool take = valid && incl <= (uint32_t)mcap;
    if (take) {
      atomicAdd(&batch[0], 1u);
      atomicMax(&batch[1], incl);
    }
    __syncthreads();  // the buckets cnt[] start at zero (left so by the last batch)
    const uint32_t nrec = batch[0], nslot = batch[1];
    const unsigned long long pn = r + nrec + tid;
    const bool vn = pn < r1;
    uint32_t rid_n = 0, lq_n = 0;
    T E0 = 0, E1 = 0, E2 = 0;
    if (vn) stream_rec(pn, rid_n, lq_n, E0, E1, E2);
    if (take && n) {
      const int u0 = up_axis(e.up, 0), u1 = up_axis(e.up, 1), u2 = up_axis(e.up, 2);
      int cell = (int)(lq & 0xffffu);
      uint32_t m = incl - n;
      for (uint32_t k = 0; k < n; k++, m++) {
        const T d = tmin(tmin(D0, D1), D2);
        const bool sx = (D0 == d), sy = (D1 == d), sz = (D2 == d);
        D0 = sx ? e.pabs[0] : D0 - d;
        D1 = sy ? e.pabs[1] : D1 - d;
        D2 = sz ? e.pabs[2] : D2 - d;
        ucell[m] = (uint16_t)cell;
        uval[m] = d * w;
        atomicAdd(&cnt[det_cnt_index(cell)], 1u);
        cell += (sx ? u0 * dm0 : 0) + (sy ? u1 * dm1 : 0) + (sz ? u2 : 0);
      }
    }
    __syncthreads();
    // bucket starts: cnt[c] = (exclusive prefix << 16) | count; the cells with
    // contributions are counted in the low half of the same scan
    const int per = (cells + kBpjBlock - 1) / kBpjBlock;
    const int c0 = (int)tid * per;
    uint32_t local = 0;
    for (int c = c0; c < c0 + per && c < cells; c++) {
      const uint32_t k = cnt[det_cnt_index(c)];
      local += (k << 16) | (k != 0u ? 1u : 0u);
    }
    const uint32_t excl = block_incl_scan(local, wsum) - local;
    const uint32_t nact = wsum[kBpjBlock / 64] & 0xffffu;
    uint32_t run = excl >> 16;
    for (int c = c0; c < c0 + per && c < cells; c++) {
      const uint32_t k = cnt[det_cnt_index(c)];
      cnt[det_cnt_index(c)] = k ? (run << 16) | k : 0u;
      run += k;
    }
    __syncthreads();
    // four slots per lane at a time: their reads and atomics overlap
    for (uint32_t m0 = tid; m0 < nslot; m0 += 4 * kBpjBlock) {
      uint32_t cs[4], pos[4];
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const uint32_t m = m0 + j * kBpjBlock;
        cs[j] = m < nslot ? (uint32_t)ucell[m] : 0xffffffffu;
      }
#pragma unroll
      for (int j = 0; j < 4; j++)
        if (cs[j] != 0xffffffffu)
          pos[j] = atomicAdd(&cnt[det_cnt_index((int)cs[j])], 1u << 16) >> 16;
#pragma unroll
      for (int j = 0; j < 4; j++)
        if (cs[j] != 0xffffffffu) sidx[pos[j]] = (uint16_t)(m0 + j * kBpjBlock);
    }
    __syncthreads();
    T w_n = 0;
    RayEntry<T> e_n;
    e_n.pabs[0] = e_n.pabs[1] = e_n.pabs[2] = (T)0;
    e_n.up = 0;
    if (vn) gather_rec(rid_n, w_n, e_n);
    // the cells with contributions, listed over the slots' cells (nact <= nslot)
    uint32_t ap = excl & 0xffffu;
    for (int c = c0; c < c0 + per && c < cells; c++)
      if (cnt[det_cnt_index(c)] & 0xffffu) ucell[ap++] = (uint16_t)c;
    __syncthreads();
    // each listed cell adds its bucket in slot order, one contribution at a
    // time, and leaves its bucket word zero for the next batch.  The slot
    // numbers of a bucket of up to 8 are sorted in registers (a sorting
    // network), longer buckets in place.
    for (uint32_t q = tid; q < nact; q += kBpjBlock) {
      const int c = ucell[q];
      const uint32_t v = cnt[det_cnt_index(c)];
      cnt[det_cnt_index(c)] = 0u;
      const uint32_t k = v & 0xffffu;
      const uint32_t end = v >> 16, st = end - k;
      T sum = acc[c];
      if (k <= 8u) {
        uint32_t sl[8];
#pragma unroll
        for (int i = 0; i < 8; i++) sl[i] = (uint32_t)i < k ? (uint32_t)sidx[st + i] : 0xffffu;
        sort8(sl);
        T u[8];
#pragma unroll
        for (int i = 0; i < 8; i++) u[i] = (uint32_t)i < k ? uval[sl[i]] : (T)0;
#pragma unroll
        for (int i = 0; i < 8; i++)
          if ((uint32_t)i < k) sum = sum + u[i];
      } else {
        for (uint32_t i = st + 1; i < end; i++) {  // insertion sort
          const uint16_t x = sidx[i];
          uint32_t z = i;
          while (z > st && sidx[z - 1] > x) {
            sidx[z] = sidx[z - 1];
            z--;
          }
          sidx[z] = x;
        }
        for (uint32_t i = st; i < end; i++) sum = sum + uval[sidx[i]];
      }
      acc[c] = sum;
    }
    r += nrec;
    rid = rid_n;
    lq = lq_n;
    D0 = E0;
    D1 = E1;
    D2 = E2;
    w = w_n;
    e = e_n;
  }
  // the last batch's sums are complete (inside the loop the next batch's
  // scans order them before the slots are rewritten)
  __syncthreads();
  for (int c = tid; c < cells; c += kBpjBlock) {
    const int l2 = c & (g.b[2] - 1), l1 = (c >> g.lg[2]) & (g.b[1] - 1), l0 = c >> (g.lg[2] + g.lg[1]);
    const int64_t i0 = lo0 + l0, i1 = lo1 + l1, i2 = lo2 + l2;
    if (i0 >= (int64_t)a.n[0] || i1 >= (int64_t)a.n[1] || i2 >= (int64_t)a.n[2]) continue;
    if (a.voxmask && a.voxmask[i0 * a.vs[0] + i1 * a.vs[1] + i2 * a.vs[2]]) continue;
    a.cube[boff + i0 * a.cs[0] + i1 * a.cs[1] + i2 * a.cs[2]] = acc[c];
  }
}

// One workgroup per chunk of one brick's records.
template <typename T>
__global__ __launch_bounds__(kBpjBlock) void k_plan_bpj(KArgs<T> a, BrickGeom g,
                                                        const unsigned long long *offsets,
                                                        const unsigned long long *chunk_prefix,
                                                        RecPtrs<T> rec) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  plan_bpj_chunk<T>(a, g, offsets, chunk_prefix, rec, blockIdx.x, reinterpret_cast<Acc *>(smem_raw));
}

// Many plans in one launch: the chunks of all jobs, one after the other.
template <typename T>
struct PlanJob {
  KArgs<T> a;
  BrickGeom g;
  const unsigned long long *offsets;
  const unsigned long long *chunk_prefix;
  RecPtrs<T> rec;
  unsigned long long chunk_begin;  // first global chunk of this job
};

template <typename T>
__global__ __launch_bounds__(kBpjBlock) void k_plan_bpj_batch(const PlanJob<T> *__restrict__ jobs,
                                                              int njobs) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const unsigned long long c = blockIdx.x;
  int lo = 0, hi = njobs - 1;  // job of this chunk (uniform per workgroup)
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].chunk_begin <= c) lo = mid;
    else hi = mid - 1;
  }
  const PlanJob<T> &j = jobs[lo];
  plan_bpj_chunk<T>(j.a, j.g, j.offsets, j.chunk_prefix, j.rec, c - j.chunk_begin,
                    reinterpret_cast<Acc *>(smem_raw));
}

__global__ void k_zero_u32(uint32_t *p, int n) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) p[k] = 0;
}

// ceil(log2 x): the exponent of the smallest power of two >= x
int ilog2(int x) {
  int l = 0;
  while ((1 << l) < x) l++;
  return l;
}

int make_geom(const tg_call *c, const tg_bpj_plan *p, BrickGeom &g) {
  memset(&g, 0, sizeof(g));
  const bool time = (c->kind == TG_PJT || c->kind == TG_BPJT);
  g.nt = time ? (c->map.nt > 0 ? c->map.nt : 1) : 1;
  for (int k = 0; k < 3; k++) {
    int b = p->brick[k];
    if (b <= 0 || (b & (b - 1))) return set_error(TG_EINVAL, "brick edges must be powers of two");
    g.b[k] = b;
    g.lg[k] = ilog2(b);
    g.nb[k] = (c->map.n[k] + b - 1) / b;
    if (c->map.n[k] > 1024) return set_error(TG_EINVAL, "plan backprojection supports n <= 1024");
  }
  // a record packs the voxel's LDS cell and the steps in the brick in 16 bits
  // each, and the brick's float64 accumulators must fit one workgroup's LDS
  if (brick_lds_cells(g) * (long long)sizeof(Acc) > kPlanLdsBytes ||
      g.b[0] + g.b[1] + g.b[2] > 65535)
    return set_error(TG_EINVAL, "brick of %d x %d x %d voxels exceeds the %d bytes of LDS", g.b[0],
                     g.b[1], g.b[2], kPlanLdsBytes);
  const long long nb = (long long)g.nb[0] * g.nb[1] * g.nb[2] * g.nt;
  if (nb > (1 << 30)) return set_error(TG_EINVAL, "too many bricks");
  g.nbricks = (int)nb;
  g.ks[1] = ilog2(g.nb[2]);
  g.ks[0] = g.ks[1] + ilog2(g.nb[1]);
  const long long keys = ((long long)g.nb[0] << g.ks[0]) * g.nt;
  if (keys >= (1ll << 31)) return set_error(TG_EINVAL, "too many bricks");
  g.nkeys = g.nb[0] << g.ks[0];
  g.keyed = keys <= kHistMax ? 1 : 0;
  g.direct = 0;
  g.chunk = p->chunk_records > 0 ? p->chunk_records : 32768;
  g.shuffle = p->order == 1 ? 1 : 0;
  g.sort = p->order == 2 ? 1 : 0;
  g.canon = p->order == 3 ? 1 : 0;
  if (p->order < 0 || p->order > 3) return set_error(TG_EINVAL, "plan order %d outside [0, 3]", p->order);
  if (p->tie_spread < 0 || p->tie_spread > 10)
    return set_error(TG_EINVAL, "tie_spread %d outside [0, 10]", p->tie_spread);
  g.spread = p->tie_spread;
  if (g.canon) {
    // lane masks of every brick in LDS; trace tiles of 256 pixels of one row;
    // cells and slots numbered in 16 bits
    if (g.nbricks > kDetMaxBricks)
      return set_error(TG_EINVAL, "a canonical plan holds at most %d bricks x time bins, not %d",
                       kDetMaxBricks, g.nbricks);
    // the count pass writes each tile's row of counts from its LDS histogram
    if (!g.keyed)
      return set_error(TG_EINVAL, "a canonical plan holds at most %d brick keys x time bins",
                       kHistMax);
    if (p->trace_tile != 0 && p->trace_tile != kPlanTraceTile)
      return set_error(TG_EINVAL, "a canonical plan traces rows of %d pixels", kPlanTraceTile);
    if (g.b[0] * g.b[1] * g.b[2] > 8192)
      return set_error(TG_EINVAL, "a canonical plan's bricks hold at most 8192 voxels");
    // one record's steps (at most the three edges' sum) must fit the smallest
    // batch of the deterministic kernel, or a batch could take no record
    if (g.b[0] + g.b[1] + g.b[2] > kDetMinSlots)
      return set_error(TG_EINVAL, "a canonical plan's brick edges sum to at most %d", kDetMinSlots);
  }
  return TG_OK;
}

// Target number of chunks (one 1024-lane workgroup each, one resident per CU
// with 128 KiB of LDS accumulators) when the caller leaves the chunk size open.
constexpr unsigned long long kAutoChunks = 1024;

struct MetaLayout {
  size_t counts, offsets, cursor, chunk_prefix, mult, border, total;
};

// border: the bricks in decreasing order of their record counts (the launch
// order of the deterministic backprojection, longest first).
MetaLayout meta_layout(int nbricks) {
  MetaLayout m;
  m.counts = 0;
  m.offsets = align_up(m.counts + 4 * (size_t)nbricks, 256);
  m.cursor = align_up(m.offsets + 8 * ((size_t)nbricks + 1), 256);
  m.chunk_prefix = align_up(m.cursor + 4 * (size_t)nbricks, 256);
  m.mult = align_up(m.chunk_prefix + 8 * ((size_t)nbricks + 1), 256);
  m.border = align_up(m.mult + 4 * (size_t)nbricks, 256);
  m.total = align_up(m.border + 4 * (size_t)nbricks, 256);
  return m;
}

// Scratch of a canonical plan: the tile x brick count matrix and the segment
// sums of its column scan.
// Scratch of a canonical plan's build: the count matrix (tiles x bricks),
// its column segment sums and every tile's pairs (the bricks it enters).
struct CanonScratch {
  size_t mat, seg, pairs, total;
  int64_t rows, nseg;
};

CanonScratch canon_scratch(int64_t tiles, int nbricks) {
  CanonScratch c;
  c.rows = tiles;
  c.nseg = (tiles + kDetSegRows - 1) / kDetSegRows;
  c.mat = 0;
  c.seg = align_up(4 * (size_t)tiles * (size_t)nbricks, 256);
  c.pairs = align_up(c.seg + 4 * (size_t)c.nseg * (size_t)nbricks, 256);
  c.total = align_up(c.pairs + kWgPairBytes * (size_t)tiles, 256);
  return c;
}

template <typename T>
RecPtrs<T> rec_ptrs(const tg_bpj_plan *p) {
  RecPtrs<T> r;
  const size_t R = (size_t)p->records;
  char *base = (char *)p->rec;
  size_t o = 0;
  r.rid = (uint32_t *)(base + o);
  o = align_up(o + 4 * R, 256);
  r.lq = (uint32_t *)(base + o);
  o = align_up(o + 4 * R, 256);
  r.d0 = (T *)(base + o);
  o = align_up(o + sizeof(T) * R, 256);
  r.d1 = (T *)(base + o);
  o = align_up(o + sizeof(T) * R, 256);
  r.d2 = (T *)(base + o);
  o = align_up(o + sizeof(T) * R, 256);
  r.ray = (RayEntry<T> *)(base + o);
  return r;
}

template <typename T>
size_t rec_bytes(size_t R, size_t rays) {
  return 2 * align_up(4 * R, 256) + 3 * align_up(sizeof(T) * R, 256) +
         align_up(sizeof(RayEntry<T>) * rays, 256) + 256;
}

size_t call_rays(const tg_call *c) { return (size_t)c->nv * c->n1 * c->n2; }

// Bytes of the count pass's per-workgroup pairs for the plan's trace grid.
template <typename T>
size_t plan_scratch_bytes(const KArgs<T> &a) {
  return (size_t)a.nv * (size_t)a.tiles_i * (size_t)a.tiles_j * kWgPairBytes;
}

// The same from the call alone (tg_bpj_plan_init): the trace grid of plan_count.
size_t plan_scratch_bytes_for(const tg_call *c, const tg_bpj_plan *p) {
  const int tj = p->trace_tile > 0 ? p->trace_tile : kPlanTraceTile;
  if (tj <= 0 || tj > 256 || (tj & (tj - 1))) return 0;
  const int bi = 256 / tj;
  return (size_t)c->nv * (size_t)((c->n1 + bi - 1) / bi) * (size_t)((c->n2 + tj - 1) / tj) *
         kWgPairBytes;
}

template <typename T>
int plan_count(const tg_call *c, tg_bpj_plan *p, hipStream_t s) {
  KArgs<T> a;
  int rc = make_args<T>(c, 0, c->nv, a, p->trace_tile > 0 ? p->trace_tile : kPlanTraceTile);
  if (rc) return rc;
  BrickGeom g;
  if ((rc = make_geom(c, p, g))) return rc;
  const MetaLayout m = meta_layout(g.nbricks);
  if (!p->meta || p->meta_bytes < m.total)
    return set_error(TG_EINVAL, "plan meta buffer needs %zu bytes", m.total);
  if ((uint64_t)c->nv * c->n1 * c->n2 > 0xffffffffull)
    return set_error(TG_EINVAL, "plan backprojection supports < 2^32 rays");
  char *meta = (char *)p->meta;
  uint32_t *counts = (uint32_t *)(meta + m.counts);
  hipLaunchKernelGGL(k_zero_u32, dim3(256), dim3(256), 0, s, counts, g.nbricks);
  static_assert(kPlanBlock == 256, "view-range launches assume 256-lane workgroups");
  const bool sun = (c->flags & TG_OBSTACLE_SUN) != 0;
  const uint64_t per_view = (uint64_t)a.tiles_i * (uint64_t)a.tiles_j;
  const CanonScratch cs = canon_scratch((int64_t)c->nv * (int64_t)per_view, g.nbricks);
  if (g.canon && (!p->scratch || p->scratch_bytes < cs.total))
    return set_error(TG_EINVAL, "a canonical plan needs a scratch of %zu bytes", cs.total);
  uint32_t *pairs = g.canon ? (uint32_t *)((char *)p->scratch + cs.pairs)
                 : (p->scratch && p->scratch_bytes >= plan_scratch_bytes(a)) ? (uint32_t *)p->scratch
                                                                             : nullptr;
  uint32_t *rowmat = g.canon ? (uint32_t *)((char *)p->scratch + cs.mat) : nullptr;
  uint32_t *per_ray = p->ray_steps_out;
  rc = for_view_launches(a, [&](const KArgs<T> &av, uint32_t blocks) {
    uint32_t *pl = pairs ? pairs + (size_t)av.t0 * per_view * (kWgPairBytes / 4) : nullptr;
    uint32_t *rm = rowmat ? rowmat + (size_t)av.t0 * per_view * (size_t)g.nbricks : nullptr;
    if (sun)
      hipLaunchKernelGGL((k_plan_count<T, true>), dim3(blocks), dim3(kPlanBlock), hist_lds_bytes(g),
                         s, av, g, counts, pl, per_ray, rm);
    else
      hipLaunchKernelGGL((k_plan_count<T, false>), dim3(blocks), dim3(kPlanBlock), hist_lds_bytes(g),
                         s, av, g, counts, pl, per_ray, rm);
  });
  // bit 0: each trace workgroup's per-brick counts (a canonical plan: also
  // the count matrix) in scratch; bit 1: every ray's step count in
  // ray_steps_out
  p->pairs_ready = ((pairs || rowmat) ? 1 : 0) | (per_ray ? 2 : 0);
  if (rc) return rc;
  if ((rc = hip_check(hipGetLastError(), "plan count launch"))) return rc;
  std::vector<uint32_t> hc(g.nbricks);
  if ((rc = hip_check(hipMemcpyAsync(hc.data(), counts, 4 * (size_t)g.nbricks,
                                     hipMemcpyDeviceToHost, s), "plan counts copy")))
    return rc;
  if ((rc = hip_check(hipStreamSynchronize(s), "plan count sync"))) return rc;
  unsigned long long tot = 0, ch = 0;
  for (int b = 0; b < g.nbricks; b++) tot += hc[b];
  if (p->chunk_records <= 0) {
    // automatic: about kAutoChunks workgroups, each with 2K..64K records
    unsigned long long want = (tot + kAutoChunks - 1) / kAutoChunks;
    int chunk = 2048;
    while (chunk < 65536 && (unsigned long long)chunk < want) chunk <<= 1;
    g.chunk = chunk;
    p->chunk_records = chunk;
  }
  std::vector<unsigned long long> off(g.nbricks + 1), cp(g.nbricks + 1);
  tot = 0;
  for (int b = 0; b < g.nbricks; b++) {
    off[b] = tot;
    cp[b] = ch;
    tot += hc[b];
    ch += (hc[b] + g.chunk - 1) / g.chunk;
  }
  off[g.nbricks] = tot;
  cp[g.nbricks] = ch;
  // one 1024-lane workgroup per chunk, and a dispatch holds 2^32 - 1 work-items
  if (ch > kMaxChunksPerLaunch)
    return set_error(TG_EINVAL, "plan of %llu chunks exceeds one launch; raise chunk_records", ch);
  // per-brick multiplier coprime with the brick's record count (bijective shuffle)
  std::vector<uint32_t> mult(g.nbricks, 1u);
  if (g.shuffle) {
    for (int b = 0; b < g.nbricks; b++) {
      const unsigned long long n = hc[b];
      if (n < 2) continue;
      unsigned long long m = 2654435761ull % n;
      if (m == 0) m = 1;
      auto gcd = [](unsigned long long x, unsigned long long y) {
        while (y) { unsigned long long r = x % y; x = y; y = r; }
        return x;
      };
      while (gcd(m, n) != 1) m = (m + 1) % n ? (m + 1) % n : 1;
      mult[b] = (uint32_t)m;
    }
  }
  if ((rc = hip_check(hipMemcpyAsync(meta + m.offsets, off.data(), 8 * off.size(),
                                     hipMemcpyHostToDevice, s), "plan offsets copy")))
    return rc;
  if ((rc = hip_check(hipMemcpyAsync(meta + m.chunk_prefix, cp.data(), 8 * cp.size(),
                                     hipMemcpyHostToDevice, s), "plan chunk copy")))
    return rc;
  if ((rc = hip_check(hipMemcpyAsync(meta + m.mult, mult.data(), 4 * mult.size(),
                                     hipMemcpyHostToDevice, s), "plan multiplier copy")))
    return rc;
  // bricks with records first, most records first (stable: ties by index)
  std::vector<uint32_t> border(g.nbricks);
  for (int b = 0; b < g.nbricks; b++) border[b] = (uint32_t)b;
  std::stable_sort(border.begin(), border.end(),
                   [&](uint32_t x, uint32_t y) { return hc[x] > hc[y]; });
  if ((rc = hip_check(hipMemcpyAsync(meta + m.border, border.data(), 4 * border.size(),
                                     hipMemcpyHostToDevice, s), "plan brick order copy")))
    return rc;
  if (g.canon && cs.rows > 0) {
    // column scan of the count matrix: each tile's first record in each brick
    uint32_t *seg = (uint32_t *)((char *)p->scratch + cs.seg);
    const dim3 grid2((uint32_t)cs.nseg, (uint32_t)((g.nbricks + 255) / 256));
    hipLaunchKernelGGL(k_det_colsum, grid2, dim3(256), 0, s, rowmat, cs.rows, g.nbricks, seg);
    hipLaunchKernelGGL(k_det_segscan, dim3((uint32_t)((g.nbricks + 255) / 256)), dim3(256), 0, s,
                       seg, cs.nseg, g.nbricks);
    hipLaunchKernelGGL(k_det_colapply, grid2, dim3(256), 0, s, rowmat, cs.rows, g.nbricks, seg);
    if ((rc = hip_check(hipGetLastError(), "canonical plan scan launch"))) return rc;
  }
  if ((rc = hip_check(hipStreamSynchronize(s), "plan meta sync"))) return rc;
  p->records = (int64_t)tot;
  p->chunks = (int64_t)ch;
  p->rec_bytes = rec_bytes<T>(tot, call_rays(c));
  for (int k = 0; k < 3; k++) p->nb_axis[k] = g.nb[k];
  p->nbricks = g.nbricks;
  return TG_OK;
}

template <typename T>
int plan_build(const tg_call *c, const tg_bpj_plan *p, hipStream_t s) {
  KArgs<T> a;
  int rc = make_args<T>(c, 0, c->nv, a, p->trace_tile > 0 ? p->trace_tile : kPlanTraceTile);
  if (rc) return rc;
  // the count pass's per-ray step counts, when it wrote them: each ray then
  // runs for its count without the map and obstacle tests
  a.ray_steps = ((p->pairs_ready & 2) && p->ray_steps_out) ? p->ray_steps_out : nullptr;
  BrickGeom g;
  if ((rc = make_geom(c, p, g))) return rc;
  // record positions fit the fill's 32-bit LDS cursors, and byte offsets into
  // the record arrays 32 bits
  g.direct = (!g.shuffle && (unsigned long long)p->records * sizeof(T) < (1ull << 32)) ? 1 : 0;
  const MetaLayout m = meta_layout(g.nbricks);
  if (!p->meta || p->meta_bytes < m.total) return set_error(TG_EINVAL, "plan meta buffer too small");
  if (!p->rec || p->rec_bytes < rec_bytes<T>((size_t)p->records, call_rays(c)))
    return set_error(TG_EINVAL, "plan record buffer needs %zu bytes",
                     rec_bytes<T>((size_t)p->records, call_rays(c)));
  char *meta = (char *)p->meta;
  uint32_t *cursor = (uint32_t *)(meta + m.cursor);
  const unsigned long long *offsets = (const unsigned long long *)(meta + m.offsets);
  const uint32_t *mult = (const uint32_t *)(meta + m.mult);
  hipLaunchKernelGGL(k_zero_u32, dim3(256), dim3(256), 0, s, cursor, g.nbricks);
  RecPtrs<T> rp = rec_ptrs<T>(p);
  const bool sun = (c->flags & TG_OBSTACLE_SUN) != 0;
  // the count pass's per-workgroup pairs, when it wrote them into this scratch
  const uint64_t per_view = (uint64_t)a.tiles_i * (uint64_t)a.tiles_j;
  if (g.canon) {
    // ray order inside every brick, from the count pass's scanned count matrix
    const CanonScratch cs = canon_scratch((int64_t)c->nv * (int64_t)per_view, g.nbricks);
    if (!(p->pairs_ready & 1) || !p->scratch || p->scratch_bytes < cs.total)
      return set_error(TG_EINVAL, "a canonical plan's build needs the scratch its count filled");
    const uint32_t *rowmat = (const uint32_t *)((const char *)p->scratch + cs.mat);
    const uint32_t *cpairs = (const uint32_t *)((const char *)p->scratch + cs.pairs);
    rc = for_view_launches(a, [&](const KArgs<T> &av, uint32_t blocks) {
      const uint32_t *rm = rowmat + (size_t)av.t0 * per_view * (size_t)g.nbricks;
      const uint32_t *ls = cpairs + (size_t)av.t0 * per_view * (kWgPairBytes / 4);
      const size_t lds = (size_t)g.nbricks * 4;
      if (sun) {
        hipLaunchKernelGGL((k_plan_fill_canon_slots<T, true>), dim3(blocks), dim3(kPlanBlock), lds,
                           s, av, g, offsets, rm, ls, rp);
        hipLaunchKernelGGL((k_plan_fill_canon<T, true>), dim3(blocks), dim3(kPlanBlock), 0, s, av,
                           g, offsets, rm, ls, rp);
      } else {
        hipLaunchKernelGGL((k_plan_fill_canon_slots<T, false>), dim3(blocks), dim3(kPlanBlock), lds,
                           s, av, g, offsets, rm, ls, rp);
        hipLaunchKernelGGL((k_plan_fill_canon<T, false>), dim3(blocks), dim3(kPlanBlock), 0, s, av,
                           g, offsets, rm, ls, rp);
      }
    });
    if (rc) return rc;
    return hip_check(hipGetLastError(), "canonical plan fill launch");
  }
  const uint32_t *pairs = ((p->pairs_ready & 1) && p->scratch &&
                           p->scratch_bytes >= plan_scratch_bytes(a))
                              ? (const uint32_t *)p->scratch
                              : nullptr;
  rc = for_view_launches(a, [&](const KArgs<T> &av, uint32_t blocks) {
    const uint32_t *pl = pairs ? pairs + (size_t)av.t0 * per_view * (kWgPairBytes / 4) : nullptr;
    if (sun)
      hipLaunchKernelGGL((k_plan_fill<T, true>), dim3(blocks), dim3(kPlanBlock), hist_lds_bytes(g),
                         s, av, g, offsets, cursor, mult, rp, pl);
    else
      hipLaunchKernelGGL((k_plan_fill<T, false>), dim3(blocks), dim3(kPlanBlock), hist_lds_bytes(g),
                         s, av, g, offsets, cursor, mult, rp, pl);
  });
  if (rc) return rc;
  if ((rc = hip_check(hipGetLastError(), "plan fill launch"))) return rc;
  if (g.sort && p->chunks > 0) {
    hipLaunchKernelGGL((k_plan_sort<T>), dim3((uint32_t)p->chunks), dim3(kBpjBlock), 0, s, g,
                       offsets, (const unsigned long long *)(meta + m.chunk_prefix), rp);
    if ((rc = hip_check(hipGetLastError(), "plan sort launch"))) return rc;
  }
  return TG_OK;
}

// Lets a kernel use up to 160 KiB of dynamic LDS (once per kernel).
template <class K>
int allow_full_lds(K kernel, bool &done) {
  if (done) return TG_OK;
  hipError_t e = hipFuncSetAttribute((const void *)kernel,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, kPlanLdsBytes);
  if (e != hipSuccess) return hip_check(e, "plan bpj LDS attribute");
  done = true;
  return TG_OK;
}

template <typename T>
int launch_plan_bpj(const KArgs<T> &a, const BrickGeom &g, const tg_bpj_plan *p, hipStream_t s) {
  const MetaLayout m = meta_layout(g.nbricks);
  const char *meta = (const char *)p->meta;
  const size_t lds = sizeof(Acc) * (size_t)brick_lds_cells(g);
  static bool attr_set = false;
  int rc = allow_full_lds(k_plan_bpj<T>, attr_set);
  if (rc) return rc;
  hipLaunchKernelGGL((k_plan_bpj<T>), dim3((uint32_t)p->chunks), dim3(kBpjBlock), lds, s, a, g,
                     (const unsigned long long *)(meta + m.offsets),
                     (const unsigned long long *)(meta + m.chunk_prefix), rec_ptrs<T>(p));
  return hip_check(hipGetLastError(), "plan bpj launch");
}

// Slots of contributions per batch of the deterministic kernel: as many as
// the LDS left by the brick's running sums and bucket words holds (a multiple
// of 256, at most 16384).  Each batch scans every bucket word once, so the
// more slots a batch takes, the less that scan costs per contribution.
template <typename T>
int det_slots(int cells) {
  int mcap = 16384;
  while (mcap > kDetMinSlots && det_lds_bytes<T>(cells, mcap) > (size_t)kPlanLdsBytes) mcap -= 256;
  return mcap;
}

template <typename T>
int launch_plan_bpj_det(const KArgs<T> &a, const BrickGeom &g, const tg_bpj_plan *p,
                        hipStream_t s) {
  if (p->records == 0) return TG_OK;
  const int cells = g.b[0] * g.b[1] * g.b[2];
  const int mcap = det_slots<T>(cells);
  const size_t lds = det_lds_bytes<T>(cells, mcap);
  if (lds > (size_t)kPlanLdsBytes)
    return set_error(TG_EINVAL, "a brick of %d voxels leaves no LDS for the deterministic sums", cells);
  static bool attr_set = false;
  int rc = allow_full_lds(k_plan_bpj_det<T>, attr_set);
  if (rc) return rc;
  const MetaLayout m = meta_layout(g.nbricks);
  const char *meta = (const char *)p->meta;
  hipLaunchKernelGGL((k_plan_bpj_det<T>), dim3((uint32_t)g.nbricks), dim3(kBpjBlock), lds, s, a, g,
                     (const unsigned long long *)(meta + m.offsets),
                     (const uint32_t *)(meta + m.border), rec_ptrs<T>(p), mcap);
  return hip_check(hipGetLastError(), "deterministic plan bpj launch");
}

template <typename T>
int plan_bpj(const tg_call *c, const tg_bpj_plan *p, hipStream_t s) {
  if (c->kind != TG_BPJ && c->kind != TG_BPJT)
    return set_error(TG_EINVAL, "plan backprojection needs kind bpj or bpjt");
  KArgs<T> a;
  int rc = make_args<T>(c, 0, c->nv, a);
  if (rc) return rc;
  if (!c->cube || !c->data) return set_error(TG_EINVAL, "cube and data pointers are required");
  BrickGeom g;
  if ((rc = make_geom(c, p, g))) return rc;
  if (g.nbricks != p->nbricks) return set_error(TG_EINVAL, "plan was built for another geometry");
  if (!p->meta || p->meta_bytes < meta_layout(g.nbricks).total || !p->rec ||
      p->rec_bytes < rec_bytes<T>((size_t)p->records, call_rays(c)))
    return set_error(TG_EINVAL, "plan buffers missing or too small");
  if (c->flags & TG_DETERMINISTIC) {
    if (!g.canon)
      return set_error(TG_EINVAL, "deterministic plan backprojection needs a plan of order 3");
    return launch_plan_bpj_det<T>(a, g, p, s);
  }
  const size_t lds = sizeof(Acc) * (size_t)brick_lds_cells(g);
  if (lds > (size_t)kPlanLdsBytes) return set_error(TG_EINVAL, "brick of %zu bytes exceeds LDS", lds);
  if (p->chunks == 0) return TG_OK;
  // the Sun's effect is in the records' step counts: one kernel for both obstacles
  if ((rc = launch_plan_bpj<T>(a, g, p, s))) return rc;
  return finish_acc<T>(a, s);
}

size_t plan_job_bytes() {
  return align_up(sizeof(PlanJob<double>) > sizeof(PlanJob<float>) ? sizeof(PlanJob<double>)
                                                                    : sizeof(PlanJob<float>),
                  16);
}

// Checks one job of a plan batch and fills its table entry (chunk_begin aside).
template <typename T>
int plan_job(const tg_call *c, const tg_bpj_plan *p, PlanJob<T> &j, size_t &lds) {
  if (c->kind != TG_BPJ && c->kind != TG_BPJT)
    return set_error(TG_EINVAL, "plan backprojection needs kind bpj or bpjt");
  if (c->flags & TG_DETERMINISTIC)
    return set_error(TG_EINVAL, "deterministic backprojection is not batched");
  int rc = make_args<T>(c, 0, c->nv, j.a);
  if (rc) return rc;
  if (!c->cube || !c->data) return set_error(TG_EINVAL, "cube and data pointers are required");
  if ((rc = make_geom(c, p, j.g))) return rc;
  if (j.g.nbricks != p->nbricks) return set_error(TG_EINVAL, "plan was built for another geometry");
  if (!p->meta || p->meta_bytes < meta_layout(j.g.nbricks).total || !p->rec ||
      p->rec_bytes < rec_bytes<T>((size_t)p->records, call_rays(c)))
    return set_error(TG_EINVAL, "plan buffers missing or too small");
  lds = sizeof(Acc) * (size_t)brick_lds_cells(j.g);
  if (lds > (size_t)kPlanLdsBytes) return set_error(TG_EINVAL, "brick of %zu bytes exceeds LDS", lds);
  const MetaLayout m = meta_layout(j.g.nbricks);
  const char *meta = (const char *)p->meta;
  j.offsets = (const unsigned long long *)(meta + m.offsets);
  j.chunk_prefix = (const unsigned long long *)(meta + m.chunk_prefix);
  j.rec = rec_ptrs<T>(p);
  return TG_OK;
}

template <typename T>
int plan_bpj_batch(const tg_call *calls, const tg_bpj_plan *plans, int njobs, void *queue,
                   void *staging, hipStream_t s) {
  memset(staging, 0, kPlanQueueHeader);
  PlanJob<T> *table = (PlanJob<T> *)((char *)staging + kPlanQueueHeader);
  unsigned long long total = 0;
  size_t lds = 0;
  for (int k = 0; k < njobs; k++) {
    PlanJob<T> j;
    memset(&j, 0, sizeof(j));
    size_t l = 0;
    int rc = plan_job<T>(&calls[k], &plans[k], j, l);
    if (rc) {
      const std::string msg = tg_last_error();
      return set_error(rc, "job %d: %s", k, msg.c_str());
    }
    if (calls[k].dtype != calls[0].dtype) return set_error(TG_EINVAL, "batch jobs must share a dtype");
    j.chunk_begin = total;
    total += (unsigned long long)plans[k].chunks;
    lds = l > lds ? l : lds;
    memcpy(&table[k], &j, sizeof(j));
  }
  {
    std::vector<const KArgs<T> *> pa(njobs);
    for (int k = 0; k < njobs; k++) pa[k] = &table[k].a;
    int rc = check_acc_jobs<T>(pa.data(), njobs);
    if (rc) return rc;
  }
  if (!queue) return TG_OK;
  const size_t bytes = kPlanQueueHeader + (size_t)njobs * sizeof(PlanJob<T>);
  int rc;
  if ((rc = hip_check(hipMemcpyAsync(queue, staging, bytes, hipMemcpyHostToDevice, s),
                      "plan batch table copy")))
    return rc;
  if (total > 0) {
    if (total > kMaxChunksPerLaunch) return set_error(TG_EINVAL, "plan batch has too many chunks");
    static bool attr_set = false;
    if ((rc = allow_full_lds(k_plan_bpj_batch<T>, attr_set))) return rc;
    hipLaunchKernelGGL((k_plan_bpj_batch<T>), dim3((uint32_t)total), dim3(kBpjBlock), lds, s,
                       (const PlanJob<T> *)((const char *)queue + kPlanQueueHeader), njobs);
    if ((rc = hip_check(hipGetLastError(), "plan batch launch"))) return rc;
  }
  // the jobs' float64 accumulators, finished in one launch over the device table
  static_assert(offsetof(PlanJob<T>, a) == 0, "finish_acc_table reads KArgs at entry start");
  std::vector<const KArgs<T> *> hargs(njobs);
  for (int k = 0; k < njobs; k++) hargs[k] = &table[k].a;
  return finish_acc_table<T>((const char *)queue + kPlanQueueHeader, sizeof(PlanJob<T>),
                             hargs.data(), njobs, s);
}

}  // namespace
}  // namespace tg

using namespace tg;

extern "C" {

int tg_bpj_plan_init(const tg_call *c, tg_bpj_plan *p) {
  if (!c || !p) return set_error(TG_EINVAL, "null argument");
  if (p->order == 3 && p->brick[0] == 0 && p->brick[1] == 0 && p->brick[2] == 0) {
    // canonical plans: 8192-voxel bricks, long across the image rows' rays
    p->brick[0] = 32;
    p->brick[1] = 16;
    p->brick[2] = 16;
  }
  if (p->brick[0] == 0 && p->brick[1] == 0 && p->brick[2] == 0) {
    // default: 32 x 16 x 32 voxels = 128 KiB of double accumulators; for cubes
    // of edge <= 96, 16 x 16 x 32 (64 KiB, two workgroups per CU), whose
    // per-chunk zeroing and flush cost less against the few records of a
    // small cube (48^3, 16 x 96^2: 0.025 -> 0.022 ms; 96^3, 32 x 128^2:
    // 0.105 -> 0.090 ms; 128^3, 64 x 512^2: 2.03 against 2.11 ms)
    const int most = c ? (c->map.n[0] > c->map.n[1] ? c->map.n[0] : c->map.n[1]) : 0;
    const int edge = c ? (most > c->map.n[2] ? most : c->map.n[2]) : 0;
    p->brick[0] = edge <= 96 ? 16 : 32;
    p->brick[1] = 16;
    p->brick[2] = 32;
  }
  BrickGeom g;
  int rc = make_geom(c, p, g);
  if (rc) return rc;
  p->nbricks = g.nbricks;
  for (int k = 0; k < 3; k++) p->nb_axis[k] = g.nb[k];
  p->meta_bytes = meta_layout(g.nbricks).total;
  if (g.canon) {
    const size_t per_tile = plan_scratch_bytes_for(c, p) / kWgPairBytes;
    p->scratch_bytes = canon_scratch((int64_t)per_tile, g.nbricks).total;
  } else {
    p->scratch_bytes = g.keyed ? plan_scratch_bytes_for(c, p) : 0;
  }
  p->pairs_ready = 0;
  p->records = 0;
  p->chunks = 0;
  p->rec_bytes = 0;
  return TG_OK;
}

int tg_bpj_plan_count(const tg_call *c, tg_bpj_plan *p, void *stream) {
  if (!c || !p) return set_error(TG_EINVAL, "null argument");
  hipStream_t s = (hipStream_t)stream;
  if (c->dtype == TG_F32) return plan_count<float>(c, p, s);
  if (c->dtype == TG_F64) return plan_count<double>(c, p, s);
  return set_error(TG_EINVAL, "bad dtype %d", c->dtype);
}

int tg_bpj_plan_build(const tg_call *c, const tg_bpj_plan *p, void *stream) {
  if (!c || !p) return set_error(TG_EINVAL, "null argument");
  hipStream_t s = (hipStream_t)stream;
  if (c->dtype == TG_F32) return plan_build<float>(c, p, s);
  if (c->dtype == TG_F64) return plan_build<double>(c, p, s);
  return set_error(TG_EINVAL, "bad dtype %d", c->dtype);
}

int tg_bpj_plan_backproject(const tg_call *c, const tg_bpj_plan *p, void *stream) {
  if (!c || !p) return set_error(TG_EINVAL, "null argument");
  hipStream_t s = (hipStream_t)stream;
  if (c->dtype == TG_F32) return plan_bpj<float>(c, p, s);
  if (c->dtype == TG_F64) return plan_bpj<double>(c, p, s);
  return set_error(TG_EINVAL, "bad dtype %d", c->dtype);
}

size_t tg_bpj_plan_batch_bytes(int32_t njobs) {
  if (njobs < 0) return 0;
  return kPlanQueueHeader + (size_t)njobs * plan_job_bytes();
}

int tg_bpj_plan_batch(const tg_call *calls, const tg_bpj_plan *plans, int32_t njobs, void *queue,
                      void *host_staging, void *stream) {
  if (njobs <= 0 || !calls || !plans) return set_error(TG_EINVAL, "empty plan batch");
  if (!queue || !host_staging) return set_error(TG_EINVAL, "plan batch needs queue and staging buffers");
  hipStream_t s = (hipStream_t)stream;
  if (calls[0].dtype == TG_F32) return plan_bpj_batch<float>(calls, plans, njobs, queue, host_staging, s);
  if (calls[0].dtype == TG_F64) return plan_bpj_batch<double>(calls, plans, njobs, queue, host_staging, s);
  return set_error(TG_EINVAL, "bad dtype %d", calls[0].dtype);
}

size_t tg_det_plan_workspace_bytes(const tg_call *c) {
  if (!c) return 0;
  return tg::det_plan_workspace_bytes(c);
}

}  // extern "C"

namespace tg {

// The deterministic backprojection of a full call (all views) through a
// canonical plan built inside the caller's workspace: meta, count matrix and
// records carved from it.  Returns TG_OK when done, kDetPlanUnfit when the
// geometry or the workspace does not allow it (nothing has been changed in
// the cube then), or an error.
template <typename T>
int run_det_plan(const tg_call *c, hipStream_t s) {
  tg_call cb = *c;
  cb.flags &= ~TG_DETERMINISTIC;
  cb.acc = nullptr;
  tg_bpj_plan p;
  memset(&p, 0, sizeof(p));
  p.order = 3;
  if (tg_bpj_plan_init(&cb, &p) != TG_OK) return kDetPlanUnfit;
  char *ws = (char *)c->workspace;
  const size_t meta = align_up(p.meta_bytes, 256), scratch = align_up(p.scratch_bytes, 256);
  // every ray's step count, written by the count pass for the fill's traces
  const size_t steps = align_up(4 * call_rays(c), 256);
  if (!ws || c->workspace_bytes < meta + scratch + steps + 256) return kDetPlanUnfit;
  p.meta = ws;
  p.scratch = ws + meta;
  p.ray_steps_out = (uint32_t *)(ws + meta + scratch);
  int rc = plan_count<T>(&cb, &p, s);
  if (rc) return rc;
  const size_t rec = rec_bytes<T>((size_t)p.records, call_rays(c));
  if (c->workspace_bytes < meta + scratch + steps + rec) return kDetPlanUnfit;
  p.rec = ws + meta + scratch + steps;
  p.rec_bytes = rec;
  if ((rc = plan_build<T>(&cb, &p, s))) return rc;
  return plan_bpj<T>(c, &p, s);
}

template int run_det_plan<float>(const tg_call *, hipStream_t);
template int run_det_plan<double>(const tg_call *, hipStream_t);

// Workspace bytes of run_det_plan for *c: meta, count matrix, per-ray step
// counts and an upper bound of the records (every traced ray entering every
// brick on its way is bounded by n0 / b0 + n1 / b1 + n2 / b2 + 3 bricks per
// ray); 0 when the geometry does not fit a canonical plan.
size_t det_plan_workspace_bytes(const tg_call *c) {
  tg_call cb = *c;
  cb.flags &= ~TG_DETERMINISTIC;
  tg_bpj_plan p;
  memset(&p, 0, sizeof(p));
  p.order = 3;
  if (tg_bpj_plan_init(&cb, &p) != TG_OK) return 0;
  const size_t rays = call_rays(c);
  const size_t per_ray = (size_t)(c->map.n[0] / p.brick[0] + c->map.n[1] / p.brick[1] +
                                  c->map.n[2] / p.brick[2] + 3);
  const size_t rec = c->dtype == TG_F64 ? rec_bytes<double>(rays * per_ray, rays)
                                        : rec_bytes<float>(rays * per_ray, rays);
  return align_up(p.meta_bytes, 256) + align_up(p.scratch_bytes, 256) +
         align_up(4 * rays, 256) + rec;
}

}  // namespace tg
