// Catalogue -> scaling-function coefficients (replaces wavelet_grid.py:17-70).
//
// Deterministic, atomics-free in floating point:
//   1. count:  each particle is assigned to every output tile its 3x3x3 stencil
//              touches (1..8 tiles; tiles are 4x8x64 cells, z fastest); integer
//              counts and the max |w| bit pattern per tile.
//   2. scan:   exclusive scan of the tile counts.
//   3. fill:   bucket 16-byte records {cell, frac, w} per tile.
//   4. accumulate: one workgroup per tile sums the 27 FP64 contributions of its
//              records into an LDS int64 fixed-point accumulator (scale chosen per
//              tile from the integer count and max |w|, so no overflow), then
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

struct ProjRec {
    uint32_t cell;  // wrapped cell coordinates, 10 bits each (x | y<<10 | z<<20)
    uint32_t frac;  // fraction indices, 10 bits each
    double w;
};

struct TileGeom {
    int L, tx, ty, tz, ntx, nty, ntz;
};

__device__ __forceinline__ int wrapi(int c, int L) {
    int r = c % L;
    return r < 0 ? r + L : r;
}

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

__device__ __forceinline__ int tiles_of(int c, int L, int T, int out[2]) {
    int t0 = wrapi(c, L) / T;
    int t1 = wrapi(c - 2, L) / T;
    out[0] = t0;
    out[1] = t1;
    return t0 == t1 ? 1 : 2;
}

__global__ void k_proj_count(const double* __restrict__ pos, const double* __restrict__ w, int64_t n, double sf,
                             int res, TileGeom g, int* __restrict__ tile_count,
                             unsigned long long* __restrict__ tile_maxw, int32_t* __restrict__ cell_out,
                             int32_t* __restrict__ frac_out, int* __restrict__ err) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
        int c[3], f[3];
        if (!cell_frac(pos, p, sf, res, c, f)) atomicOr(err, 1);
        if (cell_out) { cell_out[3 * p] = c[0]; cell_out[3 * p + 1] = c[1]; cell_out[3 * p + 2] = c[2]; }
        if (frac_out) { frac_out[3 * p] = f[0]; frac_out[3 * p + 1] = f[1]; frac_out[3 * p + 2] = f[2]; }
        unsigned long long wb = (unsigned long long)__double_as_longlong(fabs(w[p]));
        int X[2], Y[2], Z[2];
        int nx = tiles_of(c[0], g.L, g.tx, X), ny = tiles_of(c[1], g.L, g.ty, Y), nz = tiles_of(c[2], g.L, g.tz, Z);
        for (int a = 0; a < nx; ++a)
            for (int b = 0; b < ny; ++b)
                for (int d = 0; d < nz; ++d) {
                    int t = (X[a] * g.nty + Y[b]) * g.ntz + Z[d];
                    atomicAdd(&tile_count[t], 1);
                    atomicMax(&tile_maxw[t], wb);
                }
    }
}

// Exclusive scan of tile counts with one workgroup; also clears the fill cursors.
__global__ void k_proj_scan(const int* __restrict__ count, int n, int* __restrict__ off, int* __restrict__ cursor) {
    __shared__ long long part[1024];
    int tid = threadIdx.x;
    int chunk = (n + blockDim.x - 1) / blockDim.x;
    int beg = tid * chunk, end = min(n, beg + chunk);
    long long s = 0;
    for (int i = beg; i < end; ++i) s += count[i];
    part[tid] = s;
    __syncthreads();
    for (int o = 1; o < (int)blockDim.x; o <<= 1) {
        long long v = (tid >= o) ? part[tid - o] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    long long run = part[tid] - s;
    for (int i = beg; i < end; ++i) {
        off[i] = (int)run;
        run += count[i];
        cursor[i] = 0;
    }
    if (tid == blockDim.x - 1) off[n] = (int)part[tid];
}

__global__ void k_proj_fill(const double* __restrict__ pos, const double* __restrict__ w, int64_t n, double sf,
                            int res, TileGeom g, const int* __restrict__ off, int* __restrict__ cursor,
                            ProjRec* __restrict__ rec) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
        int c[3], f[3];
        cell_frac(pos, p, sf, res, c, f);
        ProjRec r;
        r.cell = (uint32_t)wrapi(c[0], g.L) | ((uint32_t)wrapi(c[1], g.L) << 10) | ((uint32_t)wrapi(c[2], g.L) << 20);
        r.frac = (uint32_t)f[0] | ((uint32_t)f[1] << 10) | ((uint32_t)f[2] << 20);
        r.w = w[p];
        int X[2], Y[2], Z[2];
        int nx = tiles_of(c[0], g.L, g.tx, X), ny = tiles_of(c[1], g.L, g.ty, Y), nz = tiles_of(c[2], g.L, g.tz, Z);
        for (int a = 0; a < nx; ++a)
            for (int b = 0; b < ny; ++b)
                for (int d = 0; d < nz; ++d) {
                    int t = (X[a] * g.nty + Y[b]) * g.ntz + Z[d];
                    int slot = off[t] + atomicAdd(&cursor[t], 1);
                    rec[slot] = r;
                }
    }
}

__global__ __launch_bounds__(ACC_THREADS) void k_proj_accumulate(
    const ProjRec* __restrict__ rec, const int* __restrict__ off, const unsigned long long* __restrict__ tile_maxw,
    const double* __restrict__ phi, int res, TileGeom g, double phimax3, float* __restrict__ out32,
    double* __restrict__ out64, int32_t* __restrict__ counts_out) {
    __shared__ unsigned long long acc[TILE_CELLS];
    __shared__ int cnt[TILE_CELLS];
    const int t = blockIdx.x;
    const int tid = threadIdx.x;
    const int tzid = t % g.ntz, tyid = (t / g.ntz) % g.nty, txid = t / (g.ntz * g.nty);
    const int ox = txid * g.tx, oy = tyid * g.ty, oz = tzid * g.tz;
    const int ncell = g.tx * g.ty * g.tz;
    const int beg = off[t], end = off[t + 1];
    const int64_t L = g.L;

    if (beg == end) {
        for (int i = tid; i < ncell; i += blockDim.x) {
            int lz = i % g.tz, ly = (i / g.tz) % g.ty, lx = i / (g.tz * g.ty);
            int64_t gi = ((ox + lx) * L + (oy + ly)) * L + (oz + lz);
            if (out32) out32[gi] = 0.0f;
            if (out64) out64[gi] = 0.0;
            if (counts_out) counts_out[gi] = 0;
        }
        return;
    }
    // Fixed-point scale: every cell receives at most one contribution per record,
    // each bounded by max|w| * max|phi|^3, so |sum| < 2^61 after scaling.
    double bound = (double)(end - beg) * __longlong_as_double((long long)tile_maxw[t]) * phimax3;
    int e = 0;
    frexp(bound, &e);
    int S = 61 - e;
    if (S > 1000) S = 1000;  // keep 2^S finite for vanishing weights
    const double mult = ldexp(1.0, S), inv = ldexp(1.0, -S);

    for (int i = tid; i < ncell; i += blockDim.x) {
        acc[i] = 0ull;
        cnt[i] = 0;
    }
    __syncthreads();
    for (int r = beg + tid; r < end; r += blockDim.x) {
        ProjRec R = rec[r];
        int cx = R.cell & 1023, cy = (R.cell >> 10) & 1023, cz = (R.cell >> 20) & 1023;
        int fx = R.frac & 1023, fy = (R.frac >> 10) & 1023, fz = (R.frac >> 20) & 1023;
        for (int sx = 0; sx < 3; ++sx) {
            int lx = wrapi(cx - sx, g.L) - ox;
            if (lx < 0 || lx >= g.tx) continue;
            double phx = R.w * phi[fx + res * sx];
            for (int sy = 0; sy < 3; ++sy) {
                int ly = wrapi(cy - sy, g.L) - oy;
                if (ly < 0 || ly >= g.ty) continue;
                double phxy = phx * phi[fy + res * sy];
                for (int sz = 0; sz < 3; ++sz) {
                    int lz = wrapi(cz - sz, g.L) - oz;
                    if (lz < 0 || lz >= g.tz) continue;
                    double v = phxy * phi[fz + res * sz];
                    long long q = __double2ll_rn(v * mult);
                    atomicAdd(&acc[(lx * g.ty + ly) * g.tz + lz], (unsigned long long)q);
                }
            }
        }
        if (counts_out) {
            int hx = cx - ox, hy = cy - oy, hz = cz - oz;
            if (hx >= 0 && hx < g.tx && hy >= 0 && hy < g.ty && hz >= 0 && hz < g.tz)
                atomicAdd(&cnt[(hx * g.ty + hy) * g.tz + hz], 1);
        }
    }
    __syncthreads();
    for (int i = tid; i < ncell; i += blockDim.x) {
        int lz = i % g.tz, ly = (i / g.tz) % g.ty, lx = i / (g.tz * g.ty);
        int64_t gi = ((ox + lx) * L + (oy + ly)) * L + (oz + lz);
        double v = (double)(long long)acc[i] * inv;
        if (out32) out32[gi] = (float)v;
        if (out64) out64[gi] = v;
        if (counts_out) counts_out[gi] = cnt[i];
    }
}

}  // namespace

extern "C" int hermes_project(hermes_ctx* ctx, const double* pos, const double* w, int64_t n, int J, double box,
                              const double* phi, int phi_len, int phi_res, double phi_abs_max, float* eps32,
                              double* eps64, int32_t* cell_out, int32_t* frac_out, int32_t* cell_counts) {
    HERMES_REQUIRE(ctx, "null context");
    HERMES_REQUIRE(J >= 2 && J <= 10, "hermes_project: J=%d outside the supported range 2..10", J);
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
    g.tx = L < MAX_TX ? L : MAX_TX;
    g.ty = L < MAX_TY ? L : MAX_TY;
    g.tz = L < MAX_TZ ? L : MAX_TZ;
    g.ntx = L / g.tx;
    g.nty = L / g.ty;
    g.ntz = L / g.tz;
    const int ntiles = g.ntx * g.nty * g.ntz;
    const double sf = (double)L / box;
    hipStream_t st = ctx->stream;

    // tile workspace: count | maxw(u64) | off(ntiles+1) | cursor | err + total readback
    size_t tile_bytes = (size_t)ntiles * 4 * 3 + 64 + (size_t)ntiles * 8 + 64;
    char* tw = (char*)hermes_workspace(ctx, WS_PROJ_TILE, tile_bytes);
    HERMES_REQUIRE(tw, "%s", hermes_last_error());
    unsigned long long* maxw = (unsigned long long*)tw;
    int* count = (int*)(tw + (size_t)ntiles * 8);
    int* off = count + ntiles;
    int* cursor = off + ntiles + 1;
    int* err = cursor + ntiles;
    HERMES_CHECK_HIP(hipMemsetAsync(tw, 0, (size_t)ntiles * 8 + (size_t)ntiles * 4, st));
    HERMES_CHECK_HIP(hipMemsetAsync(err, 0, sizeof(int), st));

    k_proj_count<<<hermes_blocks(n, 256, 4096), 256, 0, st>>>(pos, w, n, sf, phi_res, g, count, maxw, cell_out,
                                                              frac_out, err);
    HERMES_LAUNCH_CHECK();
    k_proj_scan<<<1, 1024, 0, st>>>(count, ntiles, off, cursor);
    HERMES_LAUNCH_CHECK();
    int host[2] = {0, 0};
    HERMES_CHECK_HIP(hipMemcpyAsync(&host[0], err, sizeof(int), hipMemcpyDeviceToHost, st));
    HERMES_CHECK_HIP(hipMemcpyAsync(&host[1], off + ntiles, sizeof(int), hipMemcpyDeviceToHost, st));
    HERMES_CHECK_HIP(hipStreamSynchronize(st));
    HERMES_REQUIRE(host[0] == 0, "hermes_project: non-finite or out-of-range particle positions");
    const int total = host[1];
    ProjRec* rec = (ProjRec*)hermes_workspace(ctx, WS_PROJ_REC, (size_t)(total > 0 ? total : 1) * sizeof(ProjRec));
    HERMES_REQUIRE(rec, "%s", hermes_last_error());
    if (n > 0) {
        k_proj_fill<<<hermes_blocks(n, 256, 4096), 256, 0, st>>>(pos, w, n, sf, phi_res, g, off, cursor, rec);
        HERMES_LAUNCH_CHECK();
    }
    const double phimax3 = phi_abs_max * phi_abs_max * phi_abs_max * (1.0 + 1e-12);
    k_proj_accumulate<<<ntiles, ACC_THREADS, 0, st>>>(rec, off, maxw, phi, phi_res, g, phimax3, eps32, eps64,
                                                      cell_counts);
    HERMES_LAUNCH_CHECK();
    return 0;
}
