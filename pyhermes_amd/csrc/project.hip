// Catalogue -> scaling-function coefficients (replaces wavelet_grid.py:17-70).
//
// Deterministic, atomics-free in floating point:
//   1. count:  each particle is assigned to every output tile its 3x3x3 stencil
//              touches (1..8 tiles; tiles are 4x8x64 cells, z fastest); integer
//              counts per tile (wave-aggregated atomics).
//   2. scan:   exclusive scan of the tile counts.
//   3. fill:   bucket 16-byte records {cell, frac, w} per tile.
//   4. accumulate: one workgroup per tile sums the 27 FP64 contributions of its
//              records ((record, tap) pairs spread over the workgroup) into an LDS
//              int64 fixed-point accumulator (scale chosen per tile from the record
//              count and the catalogue's max |w|, so no overflow), then
//              writes the tile once, coalesced, as FP32 and/or FP64.
// Integer addition is associative, so the result is bit-identical whatever the
// order in which records were bucketed or accumulated.
#include <cmath>

#include "../../include/hermes_amd.h"
#include "hermes_common.h"

namespace {

constexpr int MAX_TX = 4, MAX_TY = 8, MAX_TZ = 64;
constexpr int TILE_CELLS = MAX_TX * MAX_TY * MAX_TZ;
constexpr int ACC_THREADS = 256;
constexpr int SCAN_THREADS = 1024, SCAN_ITEMS = 4;  // counts per scan block: 4096
constexpr int ACC_RCHUNK = 512;                      // records staged per LDS chunk
constexpr int OFFCH = 128;                           // tile ranges staged per LDS block

struct ProjRec {
    uint32_t cell;  // wrapped cell coordinates, 10 bits each (x | y<<10 | z<<20)
    uint32_t frac;  // fraction indices, 10 bits each
    double w;
};

// Power-of-two grid and tile extents (log2), tiles z fastest.
struct TileGeom {
    int L, lgL;
    int lgtx, lgty, lgtz;
    int lgnty, lgntz;
};

// Cell and fraction index exactly as wavelet_grid.py:32-34 (int32 floor, truncation).
__device__ __forceinline__ bool cell_frac(const double* pos, int64_t p, double sf, int res, int c[3], int f[3]) {
    bool ok = true;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double s = pos[3 * p + d] * sf;
        if (!(s > -1073741824.0 && s < 1073741824.0)) {
            ok = false;
            s = 0.0;
        }
        c[d] = (int)floor(s);
        f[d] = (int)((s - (double)c[d]) * (double)res);
        // s in (-2^-54, 0) rounds s - floor(s) to 1.0; keep the phi index inside the table
        if (f[d] > res - 1) f[d] = res - 1;
    }
    return ok;
}

// Tiles along one axis covered by cells c-2..c (periodic): 1 or 2.
__device__ __forceinline__ int tiles_of(int c, int m, int lgT, int out[2]) {
    out[0] = (c & m) >> lgT;
    out[1] = ((c - 2) & m) >> lgT;
    return out[0] == out[1] ? 1 : 2;
}

__device__ __forceinline__ int tile_id(const TileGeom& g, int X, int Y, int Z) {
    return (X << (g.lgnty + g.lgntz)) | (Y << g.lgntz) | Z;
}

// Counter increment for every active lane: ctr[key] += 1, returning the value before this lane's
// increment (old value + rank among the lanes with the same key).  The lanes are first grouped by key
// with register-only ballots (catalogues stored in spatial order put a wave's particles in few
// tiles), at most WAVE_GROUPS groups (scattered keys would need one ballot round per lane); lanes left
// over add 1 each.  Every group leader and left-over lane then issues its atomic in the same
// instruction, so the wave waits for one atomic round trip.  Must be reached by all lanes of the wave.
constexpr int WAVE_GROUPS = 8;
__device__ __forceinline__ int wave_add(int* __restrict__ ctr, int key, bool active) {
    const int lane = threadIdx.x % 64;
    unsigned long long pending = __ballot(active);
    int leader_of = lane, rank = 0, group = 1;
    for (int it = 0; it < WAVE_GROUPS && pending; ++it) {
        const int leader = __ffsll((long long)pending) - 1;
        const int lkey = __shfl(key, leader, 64);
        const unsigned long long same = __ballot(active && key == lkey) & pending;
        if ((same >> lane) & 1ull) {
            leader_of = leader;
            rank = __popcll(same & ((1ull << lane) - 1ull));
        }
        if (lane == leader) group = __popcll(same);
        pending &= ~same;
    }
    int base = 0;
    if (active && lane == leader_of) base = atomicAdd(&ctr[key], group);
    base = __shfl(base, leader_of, 64);
    return base + rank;
}

// Replica k (0..7) of a particle: the tile of stencil corner (X[k>>2], Y[(k>>1)&1], Z[k&1]) when that
// corner is distinct along every axis, else -1.
__device__ __forceinline__ int replica_tile(const TileGeom& g, const int X[2], const int Y[2], const int Z[2], int nx,
                                            int ny, int nz, int k) {
    const int a = k >> 2, b = (k >> 1) & 1, d = k & 1;
    if (a >= nx || b >= ny || d >= nz) return -1;
    return tile_id(g, X[a], Y[b], Z[d]);
}

__global__ void k_proj_count(const double* __restrict__ pos, int64_t n, double sf, int res, TileGeom g,
                             int* __restrict__ tile_count, int32_t* __restrict__ cell_out,
                             int32_t* __restrict__ frac_out, int* __restrict__ err) {
    const int m = g.L - 1;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t nround = (n + stride - 1) / stride * stride;  // uniform trip count for the wave-wide ops
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < nround; p += stride) {
        const bool act = p < n;
        int c[3] = {0, 0, 0}, f[3] = {0, 0, 0};
        if (act) {
            if (!cell_frac(pos, p, sf, res, c, f)) atomicOr(err, 1);
            if (cell_out) { cell_out[3 * p] = c[0]; cell_out[3 * p + 1] = c[1]; cell_out[3 * p + 2] = c[2]; }
            if (frac_out) { frac_out[3 * p] = f[0]; frac_out[3 * p + 1] = f[1]; frac_out[3 * p + 2] = f[2]; }
        }
        int X[2], Y[2], Z[2];
        const int nx = tiles_of(c[0], m, g.lgtx, X), ny = tiles_of(c[1], m, g.lgty, Y),
                  nz = tiles_of(c[2], m, g.lgtz, Z);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int t = act ? replica_tile(g, X, Y, Z, nx, ny, nz, k) : -1;
            (void)wave_add(tile_count, t < 0 ? 0 : t, t >= 0);
        }
    }
}

// Exclusive scan of the tile counts in three launches: block sums, a scan of the block sums
// (one workgroup; at most 2^19 tiles -> 128 blocks), then each block's local scan plus its offset.
// The last launch also clears the fill cursors and writes the total to off[n].
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_reduce(const int* __restrict__ count, int n,
                                                              long long* __restrict__ bsum) {
    __shared__ long long sh[SCAN_THREADS / 64];
    const int base = blockIdx.x * SCAN_THREADS * SCAN_ITEMS;
    long long s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        const int k = base + threadIdx.x * SCAN_ITEMS + i;
        if (k < n) s += count[k];
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (threadIdx.x % 64 == 0) sh[threadIdx.x / 64] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long t = 0;
        for (int i = 0; i < SCAN_THREADS / 64; ++i) t += sh[i];
        bsum[blockIdx.x] = t;
    }
}

__global__ void k_scan_blocks(long long* __restrict__ bsum, int nb) {
    // exclusive, in place; nb is small
    if (threadIdx.x == 0) {
        long long run = 0;
        for (int i = 0; i < nb; ++i) {
            const long long v = bsum[i];
            bsum[i] = run;
            run += v;
        }
        bsum[nb] = run;
    }
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_apply(const int* __restrict__ count, int n,
                                                             const long long* __restrict__ bsum, int nb,
                                                             int* __restrict__ off, int* __restrict__ cursor) {
    __shared__ long long part[SCAN_THREADS];
    const int base = blockIdx.x * SCAN_THREADS * SCAN_ITEMS;
    const int tid = threadIdx.x;
    int v[SCAN_ITEMS];
    long long s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        const int k = base + tid * SCAN_ITEMS + i;
        v[i] = k < n ? count[k] : 0;
        s += v[i];
    }
    part[tid] = s;
    __syncthreads();
    for (int o = 1; o < SCAN_THREADS; o <<= 1) {
        const long long a = tid >= o ? part[tid - o] : 0;
        __syncthreads();
        part[tid] += a;
        __syncthreads();
    }
    long long run = bsum[blockIdx.x] + part[tid] - s;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        const int k = base + tid * SCAN_ITEMS + i;
        if (k < n) {
            off[k] = (int)run;
            cursor[k] = 0;
        }
        run += v[i];
    }
    if (blockIdx.x == 0 && tid == 0) off[n] = (int)bsum[nb];
}

// Also reduces max |w| (as the bit pattern of a non-negative double: integer max of such patterns
// orders them like the values, so the result is the same in any order) with one atomic per workgroup.
__global__ __launch_bounds__(256) void k_proj_fill(const double* __restrict__ pos, const double* __restrict__ w,
                                                   int64_t n, double sf, int res, TileGeom g,
                                                   const int* __restrict__ off, int* __restrict__ cursor,
                                                   ProjRec* __restrict__ rec, unsigned long long* __restrict__ wmax) {
    __shared__ double wred[4];
    double mw = 0.0;
    const int m = g.L - 1;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t nround = (n + stride - 1) / stride * stride;
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < nround; p += stride) {
        const bool act = p < n;
        int c[3] = {0, 0, 0}, f[3] = {0, 0, 0};
        ProjRec r;
        r.cell = 0;
        r.frac = 0;
        r.w = 0.0;
        if (act) {
            cell_frac(pos, p, sf, res, c, f);
            r.cell = (uint32_t)(c[0] & m) | ((uint32_t)(c[1] & m) << 10) | ((uint32_t)(c[2] & m) << 20);
            r.frac = (uint32_t)f[0] | ((uint32_t)f[1] << 10) | ((uint32_t)f[2] << 20);
            r.w = w[p];
            mw = fmax(mw, fabs(r.w));
        }
        int X[2], Y[2], Z[2];
        const int nx = tiles_of(c[0], m, g.lgtx, X), ny = tiles_of(c[1], m, g.lgty, Y),
                  nz = tiles_of(c[2], m, g.lgtz, Z);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int t = act ? replica_tile(g, X, Y, Z, nx, ny, nz, k) : -1;
            const int slot = wave_add(cursor, t < 0 ? 0 : t, t >= 0);
            if (t >= 0) rec[off[t] + slot] = r;
        }
    }
    for (int o = 32; o > 0; o >>= 1) mw = fmax(mw, __shfl_xor(mw, o, 64));
    if (threadIdx.x % 64 == 0) wred[threadIdx.x / 64] = mw;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < (int)blockDim.x / 64; ++i) mw = fmax(mw, wred[i]);
        atomicMax(wmax, (unsigned long long)__double_as_longlong(mw));
    }
}


// Persistent workgroups loop over tiles (t = block, block + grid, ...; consecutive workgroups take
// consecutive tiles, so concurrent stores stay close in memory).  Zeroing and output phases: thread t
// takes pairs of consecutive z cells (16-byte LDS accesses at unit lane stride: no bank conflicts),
// stored as 8-byte FP32 / 16-byte FP64 pairs.  Fixed-point scale 2^S per tile: every cell receives at
// most one contribution per record, each bounded by unit = max|w| * max|phi|^3 (global max |w|); S keeps
// every scaled contribution below 2^51 (so round_to_ll applies) and the tile's sum below 2^61.
// Fixed-point value q * 2^-S as FP32: the signed high word and the unsigned low word converted
// exactly, joined by one FP64 fma (q itself, rounded once: |q| < 2^61), scaled exactly by 2^-S and
// rounded to FP32; within one ulp of the correctly rounded value.
__device__ __forceinline__ float fixed_to_f32(long long q, int S) {
    const double hi = (double)(int)(q >> 32), lo = (double)(unsigned)(q & 0xffffffffll);
    return (float)ldexp(fma(hi, 4294967296.0, lo), -S);
}

// Round to the nearest integer (ties to even) for |x| < 2^51: x + 1.5 * 2^52 lies in [2^52, 2^53),
// where the spacing of doubles is 1, so the addition itself rounds, and the difference of the bit
// patterns is the integer.  Two instructions instead of the generic double -> int64 conversion.
__device__ __forceinline__ long long round_to_ll(double x) {
    constexpr double MAGIC = 6755399441055744.0;  // 1.5 * 2^52
    return __double_as_longlong(x + MAGIC) - __double_as_longlong(MAGIC);
}

// LDS slot of tile cell (row, lz), row = lx * ty + ly.  With FIXED tiles (64 z cells per row) the z
// index is XOR-ed with an even mask that depends on the row, 2 (row & 31): a row is 512 bytes, so without
// it the cells at one z of every row share an LDS bank, and the three x taps of a record (three rows,
// one z) as well as records at nearby z would conflict.  Pairs (lz, lz + 1) stay adjacent and aligned,
// so the 16-byte accesses of the clearing and output phases keep their layout within the row.
template <bool FIXED>
__device__ __forceinline__ int acc_slot(int row, int lz, int lgtz) {
    return FIXED ? (row << 6) | (lz ^ ((row & 31) << 1)) : (row << lgtz) + lz;
}

// OUTS: bit 0 = FP32 field, bit 1 = FP64 field.  FIXED: the full 4 x 8 x 64 tile (grids of 64^3 and
// larger) with compile-time extents.
//
// Memory pipelining: stores count in the same in-order counter as loads, so a wait for a load issued
// after a tile's output stores also waits for those stores to reach memory.  The only global loads of
// the tile loop (the next tile's first record chunk) are therefore issued unconditionally (clamped
// index) before the current tile's stores, and every tile, empty or not, issues the same number of
// stores; the wait at the top of the next tile then leaves this tile's stores in flight.
template <bool COUNTS, bool FIXED, int OUTS>
__global__ __launch_bounds__(ACC_THREADS) void k_proj_accumulate(
    const ProjRec* __restrict__ rec, const int* __restrict__ off, const double* __restrict__ phi, int res, TileGeom gg,
    const unsigned long long* __restrict__ wmax_bits, double phimax3, int ntiles, float* __restrict__ out32,
    double* __restrict__ out64, int32_t* __restrict__ counts_out) {
    TileGeom g = gg;
    if (FIXED) {
        g.lgtx = 2;
        g.lgty = 3;
        g.lgtz = 6;
    }
    __shared__ __align__(16) long long acc[TILE_CELLS];
    __shared__ __align__(16) int cnt[COUNTS ? TILE_CELLS : 4];
    // the scaling function, resident for the workgroup's life, tap-interleaved: sphi[3 f + s] = phi[f + 1024 s],
    // so the three x taps of one record (three lanes of a wave) read adjacent words instead of one bank
    __shared__ double sphi[3 * 1024];
    __shared__ __align__(16) ProjRec srec[ACC_RCHUNK];  // one chunk of the tile's records
    __shared__ int2 soff[OFFCH];
    const int tid = threadIdx.x;
    const int m = g.L - 1;
    const int tx = 1 << g.lgtx, ty = 1 << g.lgty, tz = 1 << g.lgtz;
    const int ncell = tx * ty * tz;
    const int64_t L = g.L;
    // the host requires res == 1024 (3 * 1024 samples): all loads of the table in flight at once
    static_assert((3 * 1024) % ACC_THREADS == 0, "phi table staging");
    double ph[3 * 1024 / ACC_THREADS];
#pragma unroll
    for (int j = 0; j < 3 * 1024 / ACC_THREADS; ++j) ph[j] = phi[tid + j * ACC_THREADS];
    const double unit = __longlong_as_double((long long)*wmax_bits) * phimax3;
#pragma unroll
    for (int j = 0; j < 3 * 1024 / ACC_THREADS; ++j) {
        const int i = tid + j * ACC_THREADS;  // phi[i] = phi[f + 1024 s]
        sphi[3 * (i & 1023) + (i >> 10)] = ph[j];
    }

    // Software pipeline over this workgroup's tiles t, t+G, t+2G, ...: while tile t is processed, the
    // first record chunk of tile t+G is in flight into registers.
    const int G = gridDim.x;
    constexpr int RPT = ACC_RCHUNK / ACC_THREADS;  // records per thread per chunk
    // This workgroup's tile ranges (begin, end) staged in LDS OFFCH tiles at a time: the loop then
    // issues no global loads for them (a refill, every OFFCH tiles, waits for the stores in flight).
    auto refill = [&](int k0) {
        for (int i = tid; i < OFFCH; i += ACC_THREADS) {
            const int tt = blockIdx.x + (k0 + i) * G;
            soff[i] = tt < ntiles ? make_int2(off[tt], off[tt + 1]) : make_int2(0, 0);
        }
    };
    // cell pairs per thread: a compile-time count, so the store count of the output phase is known to
    // the compiler and its waits can leave those stores in flight
    constexpr int NPAIR = TILE_CELLS / (2 * ACC_THREADS);
#pragma unroll
    for (int j = 0; j < NPAIR; ++j) {
        const int i2 = tid * 2 + j * 2 * ACC_THREADS;
        *(longlong2*)&acc[i2] = make_longlong2(0, 0);
        if (COUNTS) *(int2*)&cnt[i2] = make_int2(0, 0);
    }
    refill(0);
    __syncthreads();
    int beg = soff[0].x, end = soff[0].y;
    // the next chunk in registers, two records per thread, moved as raw 16-byte vectors (a struct copy
    // would be placed in private or LDS memory, whose store waits for the load)
    static_assert(RPT == 2 && sizeof(ProjRec) == 16, "prefetch registers assume two 16-byte records per thread");
    const uint4* __restrict__ rec4 = (const uint4*)rec;
    uint4* srec4 = (uint4*)srec;
    {  // the first tile's first chunk, directly
        const int i0 = beg + tid, i1 = beg + tid + ACC_THREADS;
        srec4[tid] = rec4[i0 < end ? i0 : 0];  // index 0 always exists; such entries are never read
        srec4[tid + ACC_THREADS] = rec4[i1 < end ? i1 : 0];
    }
    // offset of each of this thread's cell pairs from its tile's first cell (the same for every tile)
    int loff[NPAIR];
#pragma unroll
    for (int j = 0; j < NPAIR; ++j) {
        const int i2 = tid * 2 + j * 2 * ACC_THREADS;
        const int lz = i2 & (tz - 1), ly = (i2 >> g.lgtz) & (ty - 1), lx = i2 >> (g.lgtz + g.lgty);
        loff[j] = (lx * g.L + ly) * g.L + lz;
    }
    int k = 0;  // this workgroup's tile counter: t = blockIdx.x + k G
    for (int t = blockIdx.x; t < ntiles; t += G, ++k) {
        if ((k + 1) % OFFCH == 0) {  // the next tile's range starts a new staged block
            __syncthreads();
            refill(k + 1);
        }
        const int tzid = t & ((1 << g.lgntz) - 1), tyid = (t >> g.lgntz) & ((1 << g.lgnty) - 1);
        const int txid = t >> (g.lgntz + g.lgnty);
        const int ox = txid << g.lgtx, oy = tyid << g.lgty, oz = tzid << g.lgtz;
        const int64_t tbase = ((int64_t)ox * L + oy) * L + oz;  // the tile's first cell
        __syncthreads();  // this tile's first chunk is in srec; the previous output phase is done with acc
        const int2 nr2 = soff[(k + 1) % OFFCH];
        const int nbeg = nr2.x, nend = nr2.y;
        // prefetch the next tile's first chunk into registers; it is moved to srec after this tile's
        // output stores (written at the end of the body, so the wait for it sees only in-loop paths)
        uint4 pref0, pref1;
        {
            const int i0 = nbeg + tid, i1 = nbeg + tid + ACC_THREADS;
            pref0 = rec4[i0 < nend ? i0 : 0];
            pref1 = rec4[i1 < nend ? i1 : 0];
        }
        const double bound = (double)(end - beg) * unit;
        int e = 0, eu = 0;
        frexp(bound, &e);
        frexp(unit, &eu);
        int S = 61 - e < 51 - eu ? 61 - e : 51 - eu;
        if (S > 1000) S = 1000;  // keep 2^S finite for vanishing weights
        const double mult = ldexp(1.0, S), inv = ldexp(1.0, -S);
        // (the accumulators are zero here: cleared before the loop and by each tile's output phase)
        // (record, tap) pairs spread over the workgroup, product order ((w phi_x) phi_y) phi_z; records
        // beyond the first chunk (rare) are staged synchronously
        for (int rb = beg; rb < end; rb += ACC_RCHUNK) {
            const int nr = end - rb < ACC_RCHUNK ? end - rb : ACC_RCHUNK;
            if (rb != beg) {
                for (int r = tid; r < nr; r += ACC_THREADS) srec[r] = rec[rb + r];
                __syncthreads();
            }
            // items (record, x tap); y and z taps limited to the tile; 2^S folded into w (exact), so
            // each contribution is round((((w 2^S) phi_x) phi_y) phi_z)
            for (int it = tid; it < nr * 3; it += ACC_THREADS) {
                const int r = it / 3, sx = it - 3 * r;
                const ProjRec R = srec[r];
                const int cx = R.cell & 1023;
                const int lx = ((cx - sx) & m) - ox;
                if (lx < 0 || lx >= tx) continue;
                const int cy = (R.cell >> 10) & 1023, cz = (R.cell >> 20) & 1023;
                const int fx = R.frac & 1023, fy = (R.frac >> 10) & 1023, fz = (R.frac >> 20) & 1023;
                const double phx = (R.w * mult) * sphi[3 * fx + sx];
                const int rowx = lx << g.lgty;
#pragma unroll
                for (int sy = 0; sy < 3; ++sy) {
                    const int ly = ((cy - sy) & m) - oy;
                    if (ly < 0 || ly >= ty) continue;
                    const double phxy = phx * sphi[3 * fy + sy];
                    const int row = rowx + ly;
#pragma unroll
                    for (int sz = 0; sz < 3; ++sz) {
                        const int lz = ((cz - sz) & m) - oz;
                        if (lz < 0 || lz >= tz) continue;
                        const long long q = round_to_ll(phxy * sphi[3 * fz + sz]);
                        atomicAdd((unsigned long long*)&acc[acc_slot<FIXED>(row, lz, g.lgtz)], (unsigned long long)q);
                    }
                }
            }
            if (COUNTS) {
                for (int r = tid; r < nr; r += ACC_THREADS) {
                    const ProjRec R = srec[r];
                    const int hx = (int)(R.cell & 1023) - ox, hy = (int)((R.cell >> 10) & 1023) - oy;
                    const int hz = (int)((R.cell >> 20) & 1023) - oz;
                    if (hx >= 0 && hx < tx && hy >= 0 && hy < ty && hz >= 0 && hz < tz)
                        atomicAdd(&cnt[(((hx << g.lgty) + hy) << g.lgtz) + hz], 1);
                }
            }
            __syncthreads();  // chunk consumed; after the last chunk the accumulators are complete
        }
        // output: the same stores for every tile (an empty tile's accumulators are zero); each pair is
        // cleared for the next tile as it is read
#pragma unroll
        for (int j = 0; j < NPAIR; ++j) {
            const int i2 = tid * 2 + j * 2 * ACC_THREADS;
            if (!FIXED && i2 >= ncell) break;
            const int64_t gi = tbase + loff[j];
            const int ai = acc_slot<FIXED>(i2 >> g.lgtz, i2 & (tz - 1), g.lgtz);
            const longlong2 a = *(const longlong2*)&acc[ai];
            *(longlong2*)&acc[ai] = make_longlong2(0, 0);
            if (OUTS == 1) {
                *(float2*)(out32 + gi) = make_float2(fixed_to_f32(a.x, S), fixed_to_f32(a.y, S));
            } else if (OUTS != 0) {
                const double v0 = (double)a.x * inv, v1 = (double)a.y * inv;
                if (OUTS & 1) *(float2*)(out32 + gi) = make_float2((float)v0, (float)v1);
                *(double2*)(out64 + gi) = make_double2(v0, v1);
            }
            if (COUNTS) {
                *(int2*)(counts_out + gi) = *(const int2*)&cnt[i2];
                *(int2*)&cnt[i2] = make_int2(0, 0);
            }
        }
        // srec was last read by the tap loop, before the barrier that closed the last chunk
        srec4[tid] = pref0;
        srec4[tid + ACC_THREADS] = pref1;
        beg = nbeg;
        end = nend;
    }
}

template <bool COUNTS, bool FIXED>
void launch_accumulate(int outs, int grid, hipStream_t st, const ProjRec* rec, const int* off, const double* phi,
                       int res, TileGeom g, const unsigned long long* wmax, double phimax3, int ntiles, float* out32,
                       double* out64, int32_t* counts) {
    switch (outs) {
        case 0: k_proj_accumulate<COUNTS, FIXED, 0><<<grid, ACC_THREADS, 0, st>>>(rec, off, phi, res, g, wmax, phimax3, ntiles, out32, out64, counts); break;
        case 1: k_proj_accumulate<COUNTS, FIXED, 1><<<grid, ACC_THREADS, 0, st>>>(rec, off, phi, res, g, wmax, phimax3, ntiles, out32, out64, counts); break;
        case 2: k_proj_accumulate<COUNTS, FIXED, 2><<<grid, ACC_THREADS, 0, st>>>(rec, off, phi, res, g, wmax, phimax3, ntiles, out32, out64, counts); break;
        default: k_proj_accumulate<COUNTS, FIXED, 3><<<grid, ACC_THREADS, 0, st>>>(rec, off, phi, res, g, wmax, phimax3, ntiles, out32, out64, counts); break;
    }
}

int ilog2(int v) {
    int r = 0;
    while ((1 << r) < v) ++r;
    return r;
}

}  // namespace

extern "C" int hermes_project(hermes_ctx* ctx, const double* pos, const double* w, int64_t n, int J, double box,
                              const double* phi, int phi_len, int phi_res, double phi_abs_max, float* eps32,
                              double* eps64, int32_t* cell_out, int32_t* frac_out, int32_t* cell_counts) {
    HERMES_REQUIRE(ctx, "null context");
    HERMES_REQUIRE(J >= 3 && J <= 10, "hermes_project: J=%d outside the supported range 3..10", J);
    HERMES_REQUIRE(phi_res == 1024 && phi_len == 3 * phi_res,
                   "hermes_project: expects a support-3 scaling function sampled at 1024 (got %d samples, res %d)",
                   phi_len, phi_res);
    HERMES_REQUIRE(box > 0.0 && std::isfinite(box), "hermes_project: box_size must be finite and positive");
    HERMES_REQUIRE(n >= 0 && n < (int64_t)1 << 28, "hermes_project: particle count %lld unsupported", (long long)n);
    HERMES_REQUIRE(eps32 || eps64 || cell_counts, "hermes_project: no output requested");
    HERMES_CHECK_HIP(hipSetDevice(ctx->device));
    const int L = 1 << J;
    TileGeom g;
    g.L = L;
    g.lgL = J;
    g.lgtx = ilog2(L < MAX_TX ? L : MAX_TX);
    g.lgty = ilog2(L < MAX_TY ? L : MAX_TY);
    g.lgtz = ilog2(L < MAX_TZ ? L : MAX_TZ);
    g.lgnty = J - g.lgty;
    g.lgntz = J - g.lgtz;
    const int ntiles = 1 << (3 * J - g.lgtx - g.lgty - g.lgtz);
    const double sf = (double)L / box;
    hipStream_t st = ctx->stream;

    // tile workspace: (unused u64 slot) | count | off(ntiles+1) | cursor | err | scan block sums (i64)
    const int nscan = (ntiles + SCAN_THREADS * SCAN_ITEMS - 1) / (SCAN_THREADS * SCAN_ITEMS);
    size_t tile_bytes = (size_t)ntiles * 8 + (size_t)ntiles * 4 * 3 + 64 + (size_t)(nscan + 1) * 8 + 64;
    char* tw = (char*)hermes_workspace(ctx, WS_PROJ_TILE, tile_bytes);
    HERMES_REQUIRE(tw, "%s", hermes_last_error());
    int* count = (int*)(tw + (size_t)ntiles * 8);
    int* off = count + ntiles;
    int* cursor = off + ntiles + 1;
    int* err = cursor + ntiles;
    long long* bsum = (long long*)(tw + (((size_t)ntiles * 8 + (size_t)ntiles * 4 * 3 + 16 + 15) / 16) * 16);
    HERMES_CHECK_HIP(hermes_fill_u32(st, count, (size_t)ntiles, 0u));
    unsigned long long* wmax = (unsigned long long*)tw;  // first 8 bytes of the u64 slot
    HERMES_CHECK_HIP(hermes_fill_u32(st, wmax, 2, 0u));
    // Invalid positions are projected at the origin (in bounds) and flagged: under stream capture in the
    // context's sticky error word, eagerly in a word the call reads back before it returns (and fails).
    int* errp = ctx->capturing ? ctx->dev_err : err;
    HERMES_REQUIRE(errp, "hermes_project: no device error word");
    if (!ctx->capturing) HERMES_CHECK_HIP(hermes_fill_u32(st, err, 1, 0u));

    k_proj_count<<<hermes_blocks(n, 256, 4096), 256, 0, st>>>(pos, n, sf, phi_res, g, count, cell_out, frac_out,
                                                              errp);
    HERMES_LAUNCH_CHECK();
    k_scan_reduce<<<nscan, SCAN_THREADS, 0, st>>>(count, ntiles, bsum);
    HERMES_LAUNCH_CHECK();
    k_scan_blocks<<<1, 64, 0, st>>>(bsum, nscan);
    HERMES_LAUNCH_CHECK();
    k_scan_apply<<<nscan, SCAN_THREADS, 0, st>>>(count, ntiles, bsum, nscan, off, cursor);
    HERMES_LAUNCH_CHECK();
    // Records: a particle's 3-cell footprint meets at most 2 tiles per axis (tile sides >= 4), so 8n
    // records always suffice.  Up to RESERVE_BYTES that worst case is reserved and the kernels run back
    // to back, the positions checked once at the end; larger catalogues are sized exactly after the
    // count (a host synchronisation).  Capturing always uses the worst case (hermes_project_reserve).
    constexpr size_t RESERVE_BYTES = (size_t)2 << 30;
    const size_t worst = (size_t)(n > 0 ? 8 * n : 1) * sizeof(ProjRec);
    ProjRec* rec = nullptr;
    bool checked = false;
    if (ctx->capturing || worst <= RESERVE_BYTES) {
        rec = (ProjRec*)hermes_workspace(ctx, WS_PROJ_REC, worst);
        HERMES_REQUIRE(rec, "%s%s", hermes_last_error(), ctx->capturing ? " (hermes_project_reserve sizes it)" : "");
    } else {
        int host[2] = {0, 0};
        HERMES_CHECK_HIP(hipMemcpyAsync(&host[0], err, sizeof(int), hipMemcpyDeviceToHost, st));
        HERMES_CHECK_HIP(hipMemcpyAsync(&host[1], off + ntiles, sizeof(int), hipMemcpyDeviceToHost, st));
        HERMES_CHECK_HIP(hipStreamSynchronize(st));
        HERMES_REQUIRE(host[0] == 0, "hermes_project: non-finite or out-of-range particle positions");
        checked = true;
        const int total = host[1];
        rec = (ProjRec*)hermes_workspace(ctx, WS_PROJ_REC, (size_t)(total > 0 ? total : 1) * sizeof(ProjRec));
        HERMES_REQUIRE(rec, "%s", hermes_last_error());
    }
    if (n > 0) {
        k_proj_fill<<<hermes_blocks(n, 256, 4096), 256, 0, st>>>(pos, w, n, sf, phi_res, g, off, cursor, rec,
                                                                 wmax);
        HERMES_LAUNCH_CHECK();
    }
    const double phimax3 = phi_abs_max * phi_abs_max * phi_abs_max * (1.0 + 1e-12);
    int n_cu = 256;
    (void)hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, ctx->device);
    const int acc_grid = ntiles < 3 * n_cu ? ntiles : 3 * n_cu;
    const bool full = g.lgtx == 2 && g.lgty == 3 && g.lgtz == 6;
    const int outs = (eps32 ? 1 : 0) | (eps64 ? 2 : 0);
    if (cell_counts) {
        if (full)
            launch_accumulate<true, true>(outs, acc_grid, st, rec, off, phi, phi_res, g, wmax, phimax3, ntiles, eps32,
                                          eps64, cell_counts);
        else
            launch_accumulate<true, false>(outs, acc_grid, st, rec, off, phi, phi_res, g, wmax, phimax3, ntiles, eps32,
                                           eps64, cell_counts);
    } else {
        if (full)
            launch_accumulate<false, true>(outs, acc_grid, st, rec, off, phi, phi_res, g, wmax, phimax3, ntiles, eps32,
                                           eps64, nullptr);
        else
            launch_accumulate<false, false>(outs, acc_grid, st, rec, off, phi, phi_res, g, wmax, phimax3, ntiles, eps32,
                                            eps64, nullptr);
    }
    HERMES_LAUNCH_CHECK();
    if (!ctx->capturing && !checked) {  // invalid positions were projected at the origin (in bounds)
        int host_err = 0;
        if (hermes_read_word(ctx, err, &host_err)) return 1;
        HERMES_REQUIRE(host_err == 0, "hermes_project: non-finite or out-of-range particle positions");
    }
    return 0;
}

// Grows the projection workspaces to what a captured projection of up to n_max particles at level
// J needs (the worst-case record count, 8 per particle), so the capture itself allocates nothing.
extern "C" int hermes_project_reserve(hermes_ctx* ctx, int J, int64_t n_max) {
    HERMES_REQUIRE(ctx, "null context");
    HERMES_REQUIRE(!ctx->capturing, "hermes_project_reserve: the stream is being captured");
    HERMES_REQUIRE(J >= 3 && J <= 10, "hermes_project_reserve: J=%d outside 3..10", J);
    HERMES_REQUIRE(n_max >= 0 && n_max < (int64_t)1 << 28, "hermes_project_reserve: n_max out of range");
    HERMES_CHECK_HIP(hipSetDevice(ctx->device));
    const int L = 1 << J;
    const int lgtx = ilog2(L < MAX_TX ? L : MAX_TX), lgty = ilog2(L < MAX_TY ? L : MAX_TY),
              lgtz = ilog2(L < MAX_TZ ? L : MAX_TZ);
    const int ntiles = 1 << (3 * J - lgtx - lgty - lgtz);
    const int nscan = (ntiles + SCAN_THREADS * SCAN_ITEMS - 1) / (SCAN_THREADS * SCAN_ITEMS);
    const size_t tile_bytes = (size_t)ntiles * 8 + (size_t)ntiles * 4 * 3 + 64 + (size_t)(nscan + 1) * 8 + 64;
    HERMES_REQUIRE(hermes_workspace(ctx, WS_PROJ_TILE, tile_bytes), "%s", hermes_last_error());
    HERMES_REQUIRE(hermes_workspace(ctx, WS_PROJ_REC, (size_t)(n_max > 0 ? 8 * n_max : 1) * sizeof(ProjRec)), "%s",
                   hermes_last_error());
    return 0;
}
