// Monte-Carlo 3PCF kernels.
//
// * hermes_interpolate: n(x) = sum_27 eps[(c - s) & mask] phi phi phi at scaled
//   positions (+ offset), FP64 (wavelet_grid.py:143-186).
// * hermes_mc_triplet_box: the box-random-centre estimator of
//   corr3pcf_kernels.py:100-140 for one (R1, R2, mu): sum over rotations and
//   centres of n1 n2 n3 / (n_rot N), rotation uniforms supplied by the caller,
//   triangle geometry of corr3pcf_kernels.py:21-63, deterministic FP64 sums.
// * hermes_mc_binned: binned triangle sampler over (r1, r2, cos theta) bins with
//   an optional importance proposal (tables built by the host from the learned
//   proposal), Philox4x32-10 counter-based random numbers, per-bin sums of
//   w f, (w f)^2 and counts accumulated with deterministic int64 fixed-point
//   atomics.  Optionally records per-sample features for proposal training.
// * hermes_mc_bucketize: the proposal's density buckets on a sub-cell grid and
//   the stable per-bucket list of sub-cells the sampler draws from.
#include "../../include/hermes_amd.h"
#include "hermes_common.h"

#include <cmath>

namespace {

constexpr double TWO_PI = 6.283185307179586476925286766559;

__device__ __forceinline__ double interp_at(const double* __restrict__ g, const double* __restrict__ phi, int L,
                                            int res, double sx, double sy, double sz) {
    const int64_t mask = L - 1;
    const int64_t xc = (int64_t)floor(sx), yc = (int64_t)floor(sy), zc = (int64_t)floor(sz);
    // s in (-2^-54, 0) rounds s - floor(s) to 1.0; keep the phi index inside the table
    const int64_t xf = min((int64_t)((sx - (double)xc) * res), (int64_t)res - 1);
    const int64_t yf = min((int64_t)((sy - (double)yc) * res), (int64_t)res - 1);
    const int64_t zf = min((int64_t)((sz - (double)zc) * res), (int64_t)res - 1);
    double acc = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int64_t xg = (xc - a) & mask;
        const double px = phi[xf + a * res];
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const int64_t yg = (yc - b) & mask;
            const double py = phi[yf + b * res];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const int64_t zg = (zc - c) & mask;
                acc += g[(xg * L + yg) * L + zg] * px * py * phi[zf + c * res];
            }
        }
    }
    return acc;
}

__global__ void k_interpolate(const double* __restrict__ g, const double* __restrict__ spos, int64_t n,
                              const double* __restrict__ phi, int L, int res, double dx, double dy, double dz,
                              double* __restrict__ out) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x)
        out[p] = interp_at(g, phi, L, res, spos[3 * p] + dx, spos[3 * p + 1] + dy, spos[3 * p + 2] + dz);
}

// corr3pcf_kernels.py:21-63
__device__ __forceinline__ void triangle_offsets(double R1, double R2, double mu, double ph, double ct1, double al,
                                                 double o[6]) {
    const double st1 = sqrt(1.0 - ct1 * ct1);
    double sp, cp;
    sincos(ph, &sp, &cp);
    const double dx1 = st1 * cp, dy1 = st1 * sp, dz1 = ct1;
    o[0] = R1 * dx1;
    o[1] = R1 * dy1;
    o[2] = R1 * dz1;
    double ox1, oy1, oz1;
    if (fabs(dx1) > 1e-10 || fabs(dy1) > 1e-10) {
        ox1 = -dy1; oy1 = dx1; oz1 = 0.0;
    } else {
        ox1 = 0.0; oy1 = -dz1; oz1 = dy1;
    }
    const double nrm = sqrt(ox1 * ox1 + oy1 * oy1 + oz1 * oz1);
    ox1 /= nrm; oy1 /= nrm; oz1 /= nrm;
    const double ox2 = dy1 * oz1 - dz1 * oy1, oy2 = dz1 * ox1 - dx1 * oz1, oz2 = dx1 * oy1 - dy1 * ox1;
    double sa, ca;
    sincos(al, &sa, &ca);
    double st2 = 1.0 - mu * mu;
    if (st2 < 0.0) st2 = 0.0;
    const double st = sqrt(st2);
    o[3] = R2 * (mu * dx1 + st * (ca * ox1 + sa * ox2));
    o[4] = R2 * (mu * dy1 + st * (ca * oy1 + sa * oy2));
    o[5] = R2 * (mu * dz1 + st * (ca * oz1 + sa * oz2));
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

constexpr int MC_THREADS = 256;

// grid over (rotation, centre); per-block partial sums
__global__ __launch_bounds__(MC_THREADS) void k_triplet_box(double R1, double R2, double mu,
                                                            const double* __restrict__ spos,
                                                            const double* __restrict__ n1, int64_t n, int n_rot,
                                                            const double* __restrict__ abc,
                                                            const double* __restrict__ g2,
                                                            const double* __restrict__ g3,
                                                            const double* __restrict__ phi, int L, int res,
                                                            double* __restrict__ partial) {
    __shared__ double sh[MC_THREADS / 64];
    const int64_t total = n * (int64_t)n_rot;
    double s = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(i / n);
        const int64_t p = i - (int64_t)r * n;
        double o[6];
        triangle_offsets(R1, R2, mu, TWO_PI * abc[3 * r], 2.0 * abc[3 * r + 1] - 1.0, TWO_PI * abc[3 * r + 2], o);
        const double x = spos[3 * p], y = spos[3 * p + 1], z = spos[3 * p + 2];
        const double v2 = interp_at(g2, phi, L, res, x + o[0], y + o[1], z + o[2]);
        const double v3 = interp_at(g3, phi, L, res, x + o[3], y + o[4], z + o[5]);
        s += n1[p] * v2 * v3;
    }
    s = wave_sum(s);
    if (threadIdx.x % 64 == 0) sh[threadIdx.x / 64] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < MC_THREADS / 64; ++w) t += sh[w];
        partial[blockIdx.x] = t;
    }
}

__global__ void k_sum_1(const double* __restrict__ partial, int nb, double* __restrict__ out) {
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < nb; ++i) t += partial[i];
        out[0] = t;
    }
}

// ---------------------------------------------------------------- Philox4x32-10
struct u4 {
    uint32_t x, y, z, w;
};

__device__ __forceinline__ u4 philox(uint64_t counter, uint64_t key) {
    uint32_t c0 = (uint32_t)counter, c1 = (uint32_t)(counter >> 32), c2 = 0x9E3779B9u, c3 = 0x85EBCA6Bu;
    uint32_t k0 = (uint32_t)key, k1 = (uint32_t)(key >> 32);
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t h0 = (uint32_t)(p0 >> 32), l0 = (uint32_t)p0, h1 = (uint32_t)(p1 >> 32), l1 = (uint32_t)p1;
        const uint32_t n0 = h1 ^ c1 ^ k0, n2 = h0 ^ c3 ^ k1;
        c0 = n0; c1 = l1; c2 = n2; c3 = l0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return {c0, c1, c2, c3};
}

__device__ __forceinline__ double u53(uint32_t a, uint32_t b) {
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

struct BinnedArgs {
    int L, res, nr1, nr2, nmu;
    double rmin1, rmax1, rmin2, rmax2;  // grid units, uniform (linear) bins
    double cap;                          // |w f| above cap goes to the overflow list (main pass)
    double scale_wf, scale_wf2;          // fixed-point scales for sum(w f) and sum((w f)^2)
    int64_t n;                           // samples (pilot, main) or overflow entries (fix-up)
    uint64_t seed;
    int importance;
    double alpha;                        // defensive mixture weight of the uniform proposal
    int nbucket;
    int sub;                             // proposal sub-cells per cell and axis (1, 2, 4, 8)
    int anchors;                         // 1: the drawn point is the apex; 3: any vertex, balance-heuristic weight
    int record;                          // write per-sample features for the first n_record samples
    int64_t n_record;
    int64_t over_cap;                    // capacity of the overflow list
};

struct ProposalTables {
    const unsigned char* cell_bucket;
    const int* bucket_cells;
    const int64_t* bucket_off;
    const float* bucket_cdf;
    const double* bucket_density;
    const float* bin_cdf;
    const float* bin_pdf;
};

__device__ __forceinline__ int upper_bound_f(const float* __restrict__ cdf, int n, float u) {
    int lo = 0, hi = n;  // first index with cdf[i] > u
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (cdf[mid] > u) hi = mid;
        else lo = mid + 1;
    }
    return lo < n ? lo : n - 1;
}

// Two uniforms in [0, 1) from stream `stream` of sample i.
__device__ __forceinline__ void uniforms2(uint64_t i, uint64_t seed, uint64_t stream, double& a, double& b) {
    const u4 r = philox(i, seed ^ (0x9E3779B97F4A7C15ULL * (stream + 1)));
    a = u53(r.x, r.y);
    b = u53(r.z, r.w);
}

// Sub-cell of a point (cell units) on the periodic grid of side Ls = L * sub.
__device__ __forceinline__ int subcell_of(double x, double y, double z, int Ls, double sub) {
    const int m = Ls - 1;
    return ((((int)floor(x * sub)) & m) * Ls + (((int)floor(y * sub)) & m)) * Ls + (((int)floor(z * sub)) & m);
}

struct Sample {
    int bin;
    int k[3];  // proposal bucket of each vertex's sub-cell (-1 without tables)
    double wf, r1, r2, mu, f;
};

// Sample i: apex c, bin (r1, r2, mu), orientation; vertices c, c + o1, c + o2;
// f = n1 n2 n3.  Plain: apex uniform in the box, bin uniform.  Importance
// (defensive mixture): with probability alpha as plain, otherwise bucket
// k ~ bucket_cdf, a uniform sub-cell of that bucket, a uniform point p in it,
// bin ~ bin_cdf[k] and, with anchors = 3, a uniformly chosen vertex a placed at
// p (c = p - o_a); with anchors = 1 the apex is p.  Each anchor proposes
// q_a = p_uniform * R(k(v_a), b) with R(k, b) = bucket_density[k] * bin_pdf[k][b],
// so the weight is w = 1 / (alpha + (1 - alpha) * mean_a R(k(v_a), b)) over the
// vertices v_a of the drawn triangle (one anchor: the apex only).  The same i
// always gives the same sample, whichever pass evaluates it.
__device__ __forceinline__ void draw_sample(const BinnedArgs& A, const double* __restrict__ g1,
                                            const double* __restrict__ g2, const double* __restrict__ g3,
                                            const double* __restrict__ phi, const ProposalTables& T, int64_t i,
                                            Sample& S) {
    const int L = A.L;
    const int B = A.nr1 * A.nr2 * A.nmu;
    const int Ls = L * A.sub;
    const double sub = (double)A.sub;
    const double dr1 = (A.rmax1 - A.rmin1) / A.nr1, dr2 = (A.rmax2 - A.rmin2) / A.nr2, dmu = 2.0 / A.nmu;
    double ux, uy, uz, ubin, upick, umix, ubin2, w1, w2, w3, oph, oct, oal, uanc;
    uniforms2((uint64_t)i, A.seed, 0, ux, uy);
    uniforms2((uint64_t)i, A.seed, 1, uz, ubin);
    uniforms2((uint64_t)i, A.seed, 2, upick, umix);
    uniforms2((uint64_t)i, A.seed, 3, ubin2, w1);
    uniforms2((uint64_t)i, A.seed, 4, w2, w3);
    uniforms2((uint64_t)i, A.seed, 5, oph, oct);
    uniforms2((uint64_t)i, A.seed, 6, oal, uanc);
    double px, py, pz;
    int bin, anchor = 0;
    if (!A.importance || umix < A.alpha) {
        px = ux * L;
        py = uy * L;
        pz = uz * L;
        bin = (int)(ubin * B);
        if (bin >= B) bin = B - 1;
    } else {
        const int k = upper_bound_f(T.bucket_cdf, A.nbucket, (float)ubin);
        const int64_t c0 = T.bucket_off[k], c1 = T.bucket_off[k + 1];
        int64_t pick = c0 + (int64_t)(upick * (double)(c1 - c0));
        if (pick >= c1) pick = c1 - 1;
        // well-formed tables never select an empty bucket; keep malformed ones in bounds
        const int64_t ncell = (int64_t)Ls * Ls * Ls;
        pick = pick < 0 ? 0 : (pick >= ncell ? ncell - 1 : pick);
        int64_t sc = T.bucket_cells[pick];
        sc = sc < 0 ? 0 : (sc >= ncell ? ncell - 1 : sc);
        const int zs = (int)(sc % Ls), ys = (int)((sc / Ls) % Ls), xs = (int)(sc / ((int64_t)Ls * Ls));
        px = (xs + ux) / sub;
        py = (ys + uy) / sub;
        pz = (zs + uz) / sub;
        bin = upper_bound_f(T.bin_cdf + (int64_t)k * B, B, (float)ubin2);
        if (A.anchors == 3) anchor = min((int)(uanc * 3.0), 2);
    }
    const int imu = bin % A.nmu, ir2 = (bin / A.nmu) % A.nr2, ir1 = bin / (A.nmu * A.nr2);
    const double r1 = A.rmin1 + (ir1 + w1) * dr1, r2 = A.rmin2 + (ir2 + w2) * dr2;
    const double mu = -1.0 + (imu + w3) * dmu;
    double o[6];
    triangle_offsets(r1, r2, mu, TWO_PI * oph, 2.0 * oct - 1.0, TWO_PI * oal, o);
    double cx = px, cy = py, cz = pz;
    if (anchor == 1) {
        cx = px - o[0];
        cy = py - o[1];
        cz = pz - o[2];
    } else if (anchor == 2) {
        cx = px - o[3];
        cy = py - o[4];
        cz = pz - o[5];
    }
    const double n1 = interp_at(g1, phi, L, A.res, cx, cy, cz);
    const double n2 = interp_at(g2, phi, L, A.res, cx + o[0], cy + o[1], cz + o[2]);
    const double n3 = interp_at(g3, phi, L, A.res, cx + o[3], cy + o[4], cz + o[5]);
    const double f = n1 * n2 * n3;
    double w = 1.0;
    S.k[0] = S.k[1] = S.k[2] = -1;
    if (T.cell_bucket) {
        S.k[0] = min((int)T.cell_bucket[subcell_of(cx, cy, cz, Ls, sub)], A.nbucket - 1);
        S.k[1] = min((int)T.cell_bucket[subcell_of(cx + o[0], cy + o[1], cz + o[2], Ls, sub)], A.nbucket - 1);
        S.k[2] = min((int)T.cell_bucket[subcell_of(cx + o[3], cy + o[4], cz + o[5], Ls, sub)], A.nbucket - 1);
        if (A.importance) {
            double rel = T.bucket_density[S.k[0]] * (double)T.bin_pdf[(int64_t)S.k[0] * B + bin];
            if (A.anchors == 3) {
                rel += T.bucket_density[S.k[1]] * (double)T.bin_pdf[(int64_t)S.k[1] * B + bin];
                rel += T.bucket_density[S.k[2]] * (double)T.bin_pdf[(int64_t)S.k[2] * B + bin];
                rel *= (1.0 / 3.0);
            }
            w = 1.0 / (A.alpha + (1.0 - A.alpha) * rel);
        }
    }
    S.bin = bin;
    S.wf = w * f;
    S.r1 = r1;
    S.r2 = r2;
    S.mu = mu;
    S.f = f;
}

enum McPass { MC_PILOT = 0, MC_MAIN = 1, MC_FIXUP = 2 };

// PASS = MC_PILOT: only max |w f| over samples [0, A.n).
// PASS = MC_MAIN: samples [0, A.n): counts; w f and (w f)^2 into acc at scale
//   (cap) when |w f| <= cap, else the sample index goes to the overflow list.
// PASS = MC_FIXUP: the A.n listed samples into acc_hi at scale (max |w f|).
// Every pass folds max |w f| (inf for a non-finite value) into stats[0] as the
// bit pattern of a non-negative double, whose integer order is the numeric order;
// stats[1] counts overflow samples.  Fixed-point sums make the result
// independent of the order of the atomics.
template <int PASS>
__global__ __launch_bounds__(MC_THREADS) void k_mc_binned(
    const double* __restrict__ g1, const double* __restrict__ g2, const double* __restrict__ g3,
    const double* __restrict__ phi, BinnedArgs A, ProposalTables T, const int64_t* __restrict__ fix_list,
    unsigned long long* __restrict__ acc, unsigned long long* __restrict__ acc_hi,
    unsigned long long* __restrict__ stats, int64_t* __restrict__ over_list, float* __restrict__ rec) {
    __shared__ double red[MC_THREADS / 64];
    const int B = A.nr1 * A.nr2 * A.nmu;
    double mx = 0.0;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < A.n; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = PASS == MC_FIXUP ? fix_list[t] : t;
        Sample S;
        draw_sample(A, g1, g2, g3, phi, T, i, S);
        const double a = fabs(S.wf);
        mx = (a == a) ? fmax(mx, a) : INFINITY;
        if (PASS == MC_FIXUP) {
            atomicAdd(&acc_hi[S.bin], (unsigned long long)__double2ll_rn(S.wf * A.scale_wf));
            atomicAdd(&acc_hi[B + S.bin], (unsigned long long)__double2ll_rn(S.wf * S.wf * A.scale_wf2));
        } else if (PASS == MC_MAIN) {
            atomicAdd(&acc[2 * B + S.bin], 1ull);
            if (a <= A.cap) {
                atomicAdd(&acc[S.bin], (unsigned long long)__double2ll_rn(S.wf * A.scale_wf));
                atomicAdd(&acc[B + S.bin], (unsigned long long)__double2ll_rn(S.wf * S.wf * A.scale_wf2));
            } else {
                const unsigned long long slot = atomicAdd(&stats[1], 1ull);
                if ((int64_t)slot < A.over_cap) over_list[slot] = i;
            }
            if (A.record && i < A.n_record) {
                float* rr = rec + 8 * i;
                rr[0] = (float)S.r1;
                rr[1] = (float)S.r2;
                rr[2] = (float)S.mu;
                rr[3] = (float)S.f;
                rr[4] = (float)S.bin;
                rr[5] = (float)S.k[0];
                rr[6] = (float)S.k[1];
                rr[7] = (float)S.k[2];
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
    if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = red[0];
        for (int w = 1; w < MC_THREADS / 64; ++w) m = fmax(m, red[w]);
        if (m > 0.0) atomicMax(&stats[0], (unsigned long long)__double_as_longlong(m));
    }
}

// ---------------------------------------------------------------- proposal buckets
// The proposal's sub-cell grid (side Ls = L * sub): each sub-cell's bucket is
// the number of edges <= the interpolated field at its centre; bucket_cells
// lists the sub-cells bucket by bucket, ascending within a bucket (a stable,
// deterministic counting sort).  A block owns BK_CHUNK consecutive sub-cells.
constexpr int BK_THREADS = 256;
constexpr int BK_PER_THREAD = 64;
constexpr int BK_CHUNK = BK_THREADS * BK_PER_THREAD;

__global__ __launch_bounds__(BK_THREADS) void k_bucket_assign(const double* __restrict__ g,
                                                              const double* __restrict__ phi, int L, int res,
                                                              int sub, const double* __restrict__ edges,
                                                              int nbucket, int64_t ncell,
                                                              unsigned char* __restrict__ cell_bucket,
                                                              int* __restrict__ hist) {
    __shared__ int h[256];
    for (int k = threadIdx.x; k < 256; k += BK_THREADS) h[k] = 0;
    __syncthreads();
    const int64_t Ls = (int64_t)L * sub;
    const double inv = 1.0 / sub;
    const int64_t base = (int64_t)blockIdx.x * BK_CHUNK;
    for (int j = 0; j < BK_PER_THREAD; ++j) {
        const int64_t i = base + (int64_t)j * BK_THREADS + threadIdx.x;
        if (i < ncell) {
            const int64_t zs = i % Ls, ys = (i / Ls) % Ls, xs = i / (Ls * Ls);
            const double v = interp_at(g, phi, L, res, (xs + 0.5) * inv, (ys + 0.5) * inv, (zs + 0.5) * inv);
            int lo = 0, hi = nbucket - 1;  // first edge > v
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (edges[mid] > v) hi = mid;
                else lo = mid + 1;
            }
            cell_bucket[i] = (unsigned char)lo;
            atomicAdd(&h[lo], 1);
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < nbucket; k += BK_THREADS) hist[(int64_t)k * gridDim.x + blockIdx.x] = h[k];
}

// One block per bucket: exclusive prefix over blocks of hist[k][*] in place; total -> tot[k].
__global__ __launch_bounds__(1024) void k_bucket_scan(int* __restrict__ hist, int nblocks,
                                                      int64_t* __restrict__ tot) {
    __shared__ int64_t wsum[16];
    int* row = hist + (int64_t)blockIdx.x * nblocks;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int64_t carry = 0;
    for (int base = 0; base < nblocks; base += 1024) {
        const int i = base + threadIdx.x;
        const int64_t v = i < nblocks ? row[i] : 0;
        int64_t x = v;
        for (int o = 1; o < 64; o <<= 1) {
            const int64_t y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane == 63) wsum[wave] = x;
        __syncthreads();
        int64_t pre = 0, all = 0;
        for (int w = 0; w < 16; ++w) {
            if (w < wave) pre += wsum[w];
            all += wsum[w];
        }
        if (i < nblocks) row[i] = (int)(carry + pre + x - v);
        carry += all;
        __syncthreads();
    }
    if (threadIdx.x == 0) tot[blockIdx.x] = carry;
}

// Thread t of block b owns sub-cells [b*CHUNK + 64 t, +64).  For each bucket in
// turn the block ranks its matches (per-thread popcount, block exclusive scan)
// and writes their indices in ascending order.  ncell is a multiple of 64.
__global__ __launch_bounds__(BK_THREADS) void k_bucket_scatter(const unsigned char* __restrict__ cell_bucket,
                                                               int64_t ncell, int nbucket,
                                                               const int* __restrict__ offs,
                                                               const int64_t* __restrict__ bucket_off,
                                                               int* __restrict__ bucket_cells) {
    __shared__ int wtot[BK_THREADS / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t base = (int64_t)blockIdx.x * BK_CHUNK + (int64_t)threadIdx.x * BK_PER_THREAD;
    const bool valid = base < ncell;
    uint32_t wv[16];
    if (valid) {
        const uint4* p = (const uint4*)(cell_bucket + base);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint4 t = p[q];
            wv[4 * q] = t.x;
            wv[4 * q + 1] = t.y;
            wv[4 * q + 2] = t.z;
            wv[4 * q + 3] = t.w;
        }
    } else {
#pragma unroll
        for (int q = 0; q < 16; ++q) wv[q] = 0;
    }
    for (int k = 0; k < nbucket; ++k) {
        const uint32_t kk = 0x01010101u * (uint32_t)k;
        uint32_t m[16];
        int cnt = 0;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const uint32_t x = wv[q] ^ kk;  // zero byte <=> match
            uint32_t t = (x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu;
            t = ~(t | x | 0x7F7F7F7Fu);    // bit 7 of each zero byte
            m[q] = valid ? t : 0u;
            cnt += __popc(m[q]);
        }
        int x = cnt;
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane == 63) wtot[wave] = x;
        __syncthreads();
        int pre = x - cnt;
        for (int w = 0; w < wave; ++w) pre += wtot[w];
        __syncthreads();
        if (cnt) {
            int64_t dst = bucket_off[k] + offs[(int64_t)k * gridDim.x + blockIdx.x] + pre;
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                uint32_t t = m[q];
                while (t) {
                    const int byte = (__ffs(t) - 1) >> 3;
                    bucket_cells[dst++] = (int)(base + 4 * q + byte);
                    t &= t - 1;
                }
            }
        }
    }
}

}  // namespace

extern "C" {

int hermes_interpolate(hermes_ctx* ctx, int J, const double* grid, const double* spos, int64_t n, const double* phi,
                       int phi_res, double dx, double dy, double dz, double* out) {
    HERMES_REQUIRE(ctx && grid && phi && (n == 0 || (spos && out)), "hermes_interpolate: null argument");
    HERMES_REQUIRE(J >= 2 && J <= 10, "J=%d unsupported", J);
    HERMES_REQUIRE(phi_res == 1024, "phi_res must be 1024");
    HERMES_CHECK_HIP(hipSetDevice(ctx->device));
    if (n == 0) return 0;
    k_interpolate<<<hermes_blocks(n, 256, 65536), 256, 0, ctx->stream>>>(grid, spos, n, phi, 1 << J, phi_res, dx, dy,
                                                                         dz, out);
    HERMES_LAUNCH_CHECK();
    return 0;
}

int hermes_mc_triplet_box(hermes_ctx* ctx, int J, double R1, double R2, double mu, const double* spos,
                          const double* n1, int64_t n, int n_rot, const double* abc_host, const double* g2,
                          const double* g3, const double* phi, int phi_res, double* out_host) {
    HERMES_REQUIRE(ctx && spos && n1 && abc_host && g2 && g3 && phi && out_host, "hermes_mc_triplet_box: null arg");
    HERMES_REQUIRE(J >= 2 && J <= 10, "J=%d unsupported", J);
    HERMES_REQUIRE(n > 0 && n_rot > 0, "need centres and rotations");
    HERMES_CHECK_HIP(hipSetDevice(ctx->device));
    const int64_t total = n * (int64_t)n_rot;
    const int nb = (int)hermes_blocks(total, MC_THREADS, 8192);
    double* ws = (double*)hermes_workspace(ctx, WS_MC, sizeof(double) * ((size_t)nb + 8 + 3 * (size_t)n_rot));
    HERMES_REQUIRE(ws, "%s", hermes_last_error());
    double* d_abc = ws + nb + 8;
    HERMES_CHECK_HIP(hipMemcpyAsync(d_abc, abc_host, sizeof(double) * 3 * n_rot, hipMemcpyHostToDevice, ctx->stream));
    k_triplet_box<<<nb, MC_THREADS, 0, ctx->stream>>>(R1, R2, mu, spos, n1, n, n_rot, d_abc, g2, g3, phi, 1 << J,
                                                      phi_res, ws);
    HERMES_LAUNCH_CHECK();
    k_sum_1<<<1, 64, 0, ctx->stream>>>(ws, nb, ws + nb);
    HERMES_LAUNCH_CHECK();
    double s = 0.0;
    HERMES_CHECK_HIP(hipMemcpyAsync(&s, ws + nb, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HERMES_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    *out_host = s / ((double)n_rot * (double)n);
    return 0;
}

int hermes_mc_binned(hermes_ctx* ctx, int J, const double* g1, const double* g2, const double* g3, const double* phi,
                     int phi_res, int nr1, int nr2, int nmu, double rmin1, double rmax1, double rmin2, double rmax2,
                     int64_t n_samples, uint64_t seed, double wf_cap, const hermes_mc_proposal* prop, float* record,
                     int64_t n_record, double* sum_wf_host, double* sum_wf2_host, int64_t* count_host) {
    HERMES_REQUIRE(ctx && g1 && g2 && g3 && phi && sum_wf_host && sum_wf2_host && count_host,
                   "hermes_mc_binned: null argument");
    HERMES_REQUIRE(J >= 2 && J <= 10, "J=%d unsupported", J);
    HERMES_REQUIRE(phi_res == 1024, "phi_res must be 1024");
    HERMES_REQUIRE(nr1 >= 1 && nr2 >= 1 && nmu >= 1 && (int64_t)nr1 * nr2 * nmu <= (1 << 20), "invalid binning");
    HERMES_REQUIRE(rmax1 > rmin1 && rmax2 > rmin2 && rmin1 >= 0.0 && rmin2 >= 0.0, "invalid radial ranges");
    HERMES_REQUIRE(n_samples >= 1 && wf_cap >= 0.0, "n_samples must be positive and wf_cap non-negative");
    HERMES_REQUIRE(record == nullptr || n_record >= 0, "invalid record count");
    if (prop) {
        HERMES_REQUIRE(prop->cell_bucket && prop->bucket_cells && prop->bucket_off && prop->bucket_cdf &&
                           prop->bucket_density && prop->bin_cdf && prop->bin_pdf,
                       "hermes_mc_binned: incomplete proposal tables");
        HERMES_REQUIRE(prop->nbucket >= 1 && prop->nbucket <= 256, "nbucket must be in 1..256");
        HERMES_REQUIRE(prop->alpha > 0.0 && prop->alpha <= 1.0, "alpha must be in (0, 1]");
        HERMES_REQUIRE(prop->sub == 1 || prop->sub == 2 || prop->sub == 4 || prop->sub == 8,
                       "sub must be 1, 2, 4 or 8");
        HERMES_REQUIRE((1 << J) * prop->sub <= 1024, "L * sub must not exceed 1024");
        HERMES_REQUIRE(prop->anchors == 1 || prop->anchors == 3, "anchors must be 1 or 3");
    }
    HERMES_CHECK_HIP(hipSetDevice(ctx->device));
    const int B = nr1 * nr2 * nmu;
    constexpr int64_t OVER_CAP = 1 << 20;
    constexpr int64_t PILOT = 1 << 20;
    BinnedArgs A;
    A.L = 1 << J;
    A.res = phi_res;
    A.nr1 = nr1;
    A.nr2 = nr2;
    A.nmu = nmu;
    A.rmin1 = rmin1;
    A.rmax1 = rmax1;
    A.rmin2 = rmin2;
    A.rmax2 = rmax2;
    A.seed = seed;
    A.importance = prop != nullptr;
    A.alpha = prop ? prop->alpha : 1.0;
    A.nbucket = prop ? prop->nbucket : 0;
    A.sub = prop ? prop->sub : 1;
    A.anchors = prop ? prop->anchors : 1;
    A.record = record != nullptr;
    A.n_record = record ? n_record : 0;
    A.over_cap = OVER_CAP;
    ProposalTables T{};
    if (prop)
        T = {prop->cell_bucket, prop->bucket_cells, prop->bucket_off, prop->bucket_cdf, prop->bucket_density,
             prop->bin_cdf, prop->bin_pdf};
    // workspace: acc (3B: sum wf, sum wf^2, count) | acc_hi (2B) | stats (2) | overflow list
    const size_t n_acc = 5 * (size_t)B + 2;
    char* ws = (char*)hermes_workspace(ctx, WS_MC, sizeof(unsigned long long) * n_acc + sizeof(int64_t) * OVER_CAP);
    HERMES_REQUIRE(ws, "%s", hermes_last_error());
    unsigned long long* acc = (unsigned long long*)ws;
    unsigned long long* acc_hi = acc + 3 * (size_t)B;
    unsigned long long* stats = acc_hi + 2 * (size_t)B;
    int64_t* over_list = (int64_t*)(ws + sizeof(unsigned long long) * n_acc);
    unsigned long long st[2];
    auto read_stats = [&]() -> int {
        HERMES_CHECK_HIP(hipMemcpyAsync(st, stats, sizeof(st), hipMemcpyDeviceToHost, ctx->stream));
        HERMES_CHECK_HIP(hipStreamSynchronize(ctx->stream));
        return 0;
    };
    auto max_of = [&]() { return __builtin_bit_cast(double, st[0]); };

    // The fixed-point scale comes from a cap on |w f|: given by the caller, or the
    // maximum over the first PILOT samples (the main pass draws the same ones).
    double cap = wf_cap;
    if (!(cap > 0.0)) {
        HERMES_CHECK_HIP(hipMemsetAsync(stats, 0, 2 * sizeof(unsigned long long), ctx->stream));
        A.n = n_samples < PILOT ? n_samples : PILOT;
        k_mc_binned<MC_PILOT><<<hermes_blocks(A.n, MC_THREADS, 16384), MC_THREADS, 0, ctx->stream>>>(
            g1, g2, g3, phi, A, T, nullptr, acc, acc_hi, stats, over_list, nullptr);
        HERMES_LAUNCH_CHECK();
        if (read_stats()) return 1;
        cap = max_of();
        HERMES_REQUIRE(std::isfinite(cap), "hermes_mc_binned: non-finite triangle weight");
        if (!(cap > 0.0)) cap = 1.0;
    }
    for (int attempt = 0;; ++attempt) {
        HERMES_CHECK_HIP(hipMemsetAsync(acc, 0, sizeof(unsigned long long) * n_acc, ctx->stream));
        A.n = n_samples;
        A.cap = cap;
        // |w f| <= cap for every accumulated sample: each bin sum stays below 2^62
        A.scale_wf = ldexp(1.0, 62) / ((double)n_samples * cap);
        A.scale_wf2 = ldexp(1.0, 62) / ((double)n_samples * cap * cap);
        k_mc_binned<MC_MAIN><<<hermes_blocks(n_samples, MC_THREADS, 16384), MC_THREADS, 0, ctx->stream>>>(
            g1, g2, g3, phi, A, T, nullptr, acc, acc_hi, stats, over_list, record);
        HERMES_LAUNCH_CHECK();
        if (read_stats()) return 1;
        if (st[1] == 0) break;
        HERMES_REQUIRE(std::isfinite(max_of()), "hermes_mc_binned: non-finite triangle weight");
        if ((int64_t)st[1] <= OVER_CAP) break;
        // too many samples above the cap for the list: redo with the exact maximum
        HERMES_REQUIRE(attempt == 0, "hermes_mc_binned: overflow list exceeded twice");
        cap = max_of();
    }
    const int64_t n_over = (int64_t)st[1];
    double scale_hi = 0.0, scale_hi2 = 0.0;
    if (n_over > 0) {
        const double mx = max_of();
        BinnedArgs F = A;
        F.n = n_over;
        F.record = 0;
        F.scale_wf = scale_hi = ldexp(1.0, 62) / ((double)n_over * mx);
        F.scale_wf2 = scale_hi2 = ldexp(1.0, 62) / ((double)n_over * mx * mx);
        k_mc_binned<MC_FIXUP><<<hermes_blocks(n_over, MC_THREADS, 16384), MC_THREADS, 0, ctx->stream>>>(
            g1, g2, g3, phi, F, T, over_list, acc, acc_hi, stats, over_list, nullptr);
        HERMES_LAUNCH_CHECK();
    }
    std::vector<long long> h(5 * (size_t)B);
    HERMES_CHECK_HIP(hipMemcpyAsync(h.data(), acc, sizeof(long long) * 5 * (size_t)B, hipMemcpyDeviceToHost,
                                    ctx->stream));
    HERMES_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    for (int b = 0; b < B; ++b) {
        sum_wf_host[b] = (double)h[b] / A.scale_wf;
        sum_wf2_host[b] = (double)h[B + b] / A.scale_wf2;
        if (n_over > 0) {
            sum_wf_host[b] += (double)h[3 * (size_t)B + b] / scale_hi;
            sum_wf2_host[b] += (double)h[4 * (size_t)B + b] / scale_hi2;
        }
        count_host[b] = h[2 * B + b];
    }
    return 0;
}

int hermes_mc_bucketize(hermes_ctx* ctx, int J, const double* grid, const double* phi, int phi_res, int sub,
                        const double* edges_host, int nbucket, uint8_t* cell_bucket, int32_t* bucket_cells,
                        int64_t* bucket_off, int64_t* counts_host) {
    HERMES_REQUIRE(ctx && grid && phi && cell_bucket && bucket_cells && bucket_off && counts_host,
                   "hermes_mc_bucketize: null argument");
    HERMES_REQUIRE(J >= 2 && J <= 10, "J=%d unsupported", J);
    HERMES_REQUIRE(phi_res == 1024, "phi_res must be 1024");
    HERMES_REQUIRE(sub == 1 || sub == 2 || sub == 4 || sub == 8, "sub must be 1, 2, 4 or 8");
    HERMES_REQUIRE((1 << J) * sub <= 1024, "L * sub must not exceed 1024");
    HERMES_REQUIRE(nbucket >= 1 && nbucket <= 256, "nbucket must be in 1..256");
    HERMES_REQUIRE(nbucket == 1 || edges_host, "hermes_mc_bucketize: edges missing");
    for (int k = 1; k + 1 < nbucket; ++k)
        HERMES_REQUIRE(edges_host[k] >= edges_host[k - 1], "edges must be ascending");
    HERMES_CHECK_HIP(hipSetDevice(ctx->device));
    const int64_t Ls = (int64_t)(1 << J) * sub;
    const int64_t ncell = Ls * Ls * Ls;  // a multiple of 64 since Ls >= 4
    const int nblocks = (int)((ncell + BK_CHUNK - 1) / BK_CHUNK);
    const size_t hist_bytes = sizeof(int) * (size_t)nbucket * nblocks;
    char* ws = (char*)hermes_workspace(ctx, WS_MC, hist_bytes + sizeof(int64_t) * 256 + sizeof(double) * 256);
    HERMES_REQUIRE(ws, "%s", hermes_last_error());
    int* hist = (int*)ws;
    int64_t* tot = (int64_t*)(ws + hist_bytes);
    double* edges = (double*)(ws + hist_bytes + sizeof(int64_t) * 256);
    if (nbucket > 1)
        HERMES_CHECK_HIP(hipMemcpyAsync(edges, edges_host, sizeof(double) * (nbucket - 1), hipMemcpyHostToDevice,
                                        ctx->stream));
    k_bucket_assign<<<nblocks, BK_THREADS, 0, ctx->stream>>>(grid, phi, 1 << J, phi_res, sub, edges, nbucket, ncell,
                                                             cell_bucket, hist);
    HERMES_LAUNCH_CHECK();
    k_bucket_scan<<<nbucket, 1024, 0, ctx->stream>>>(hist, nblocks, tot);
    HERMES_LAUNCH_CHECK();
    std::vector<int64_t> t(nbucket), off(nbucket + 1);
    HERMES_CHECK_HIP(hipMemcpyAsync(t.data(), tot, sizeof(int64_t) * nbucket, hipMemcpyDeviceToHost, ctx->stream));
    HERMES_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    off[0] = 0;
    for (int k = 0; k < nbucket; ++k) {
        off[k + 1] = off[k] + t[k];
        counts_host[k] = t[k];
    }
    HERMES_REQUIRE(off[nbucket] == ncell, "hermes_mc_bucketize: bucket counts do not add up");
    HERMES_CHECK_HIP(hipMemcpyAsync(bucket_off, off.data(), sizeof(int64_t) * (nbucket + 1), hipMemcpyHostToDevice,
                                    ctx->stream));
    k_bucket_scatter<<<nblocks, BK_THREADS, 0, ctx->stream>>>(cell_bucket, ncell, nbucket, hist, bucket_off,
                                                              bucket_cells);
    HERMES_LAUNCH_CHECK();
    HERMES_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

}  // extern "C"
