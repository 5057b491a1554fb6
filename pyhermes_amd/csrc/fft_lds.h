// In-LDS complex FFTs of power-of-two length for the 3-D grid transforms.
//
// Mixed-radix Stockham autosort: stage (R, Ns) reads butterfly j's inputs
// x[j + r L/R], multiplies input r by w^(r (j mod Ns) L/(Ns R)) (w = e^{DIR 2 pi i / L}),
// applies a length-R DFT and writes output r to (j div Ns) Ns R + (j mod Ns) + r Ns.
// Stages run R0 = 2^(log2 L mod 3) (or 8) first, then radix 8, so Ns = 1, R0, 8 R0, ...
// Every thread loads all of its butterflies into registers before a barrier and
// stores after it, so one LDS buffer holds the lines in place.
//
// Lines are stored either row-major with a pad of one element per 16
// (LINES_FAST = false: a line is contiguous, consecutive threads walk along it)
// or element-major (LINES_FAST = true: consecutive threads take consecutive
// lines, which for a tile of contiguous grid lines keeps LDS accesses unit-stride).
//
// DIR = +1 is the unnormalised inverse transform (e^{+2 pi i jk/L}), DIR = -1 the forward one.
#pragma once

#include <hip/hip_runtime.h>

namespace hermes_fft {

// Workgroup barrier ordering LDS accesses only.  __syncthreads() also waits until every global store of
// the wave has completed, which in a kernel that alternates LDS work with output stores (one field after
// another) serialises each output phase with the next compute phase; the LDS-scope fences let the stores
// drain while the waves go on.  Only for barriers that protect LDS data.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Loads and stores with the nontemporal hint, for data that a kernel reads or writes once and that no later
// access finds in the caches (passes over grids of hundreds of MB).
typedef float v2f_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float2 stream_load(const float2* p) {
    const v2f_t v = __builtin_nontemporal_load((const v2f_t*)p);
    return make_float2(v.x, v.y);
}
__device__ __forceinline__ void stream_store(float2* p, float2 a) {
    const v2f_t v = {a.x, a.y};
    __builtin_nontemporal_store(v, (v2f_t*)p);
}

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * (DIR i)
template <int DIR>
__device__ __forceinline__ float2 mul_diri(float2 a) {
    return DIR > 0 ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

template <int R, int DIR>
__device__ __forceinline__ void dft(float2* v);

template <>
__device__ __forceinline__ void dft<2, 1>(float2* v) {
    const float2 a = v[0];
    v[0] = cadd(a, v[1]);
    v[1] = csub(a, v[1]);
}
template <>
__device__ __forceinline__ void dft<2, -1>(float2* v) {
    dft<2, 1>(v);
}

template <int DIR>
__device__ __forceinline__ void dft4(float2& v0, float2& v1, float2& v2, float2& v3) {
    const float2 t0 = cadd(v0, v2), t1 = csub(v0, v2), t2 = cadd(v1, v3), t3 = mul_diri<DIR>(csub(v1, v3));
    v0 = cadd(t0, t2);
    v2 = csub(t0, t2);
    v1 = cadd(t1, t3);
    v3 = csub(t1, t3);
}
template <>
__device__ __forceinline__ void dft<4, 1>(float2* v) {
    dft4<1>(v[0], v[1], v[2], v[3]);
}
template <>
__device__ __forceinline__ void dft<4, -1>(float2* v) {
    dft4<-1>(v[0], v[1], v[2], v[3]);
}

// Radix 8 by decimation in time: X_k = E_k + w8^k O_k, X_{k+4} = E_k - w8^k O_k.
template <int DIR>
__device__ __forceinline__ void dft8(float2* v) {
    constexpr float c = 0.70710678118654752f;
    float2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    float2 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4<DIR>(e0, e1, e2, e3);
    dft4<DIR>(o0, o1, o2, o3);
    // w8^1 = c (1 + DIR i), w8^2 = DIR i, w8^3 = c (-1 + DIR i)
    o1 = make_float2(c * (o1.x - DIR * o1.y), c * (DIR * o1.x + o1.y));
    o2 = mul_diri<DIR>(o2);
    o3 = make_float2(c * (-o3.x - DIR * o3.y), c * (DIR * o3.x - o3.y));
    v[0] = cadd(e0, o0);
    v[4] = csub(e0, o0);
    v[1] = cadd(e1, o1);
    v[5] = csub(e1, o1);
    v[2] = cadd(e2, o2);
    v[6] = csub(e2, o2);
    v[3] = cadd(e3, o3);
    v[7] = csub(e3, o3);
}
template <>
__device__ __forceinline__ void dft<8, 1>(float2* v) {
    dft8<1>(v);
}
template <>
__device__ __forceinline__ void dft<8, -1>(float2* v) {
    dft8<-1>(v);
}

// NLINES lines of length 2^LOG2L in LDS, transformed by NT threads.
template <int LOG2L, int NLINES, int NT, int DIR, bool LINES_FAST>
struct LdsFft {
    static constexpr int L = 1 << LOG2L;
    static constexpr int ROW = L + L / 16;  // padded row pitch (row-major layout)
    static constexpr int ELEMS = LINES_FAST ? NLINES * L : NLINES * ROW;

    __device__ static __forceinline__ int addr(int line, int e) {
        if (LINES_FAST) return e * NLINES + line;
        return line * ROW + e + (e >> 4);
    }

    template <int R, int NS>
    __device__ static __forceinline__ void stage(float2* lds, const float2* tw, int tid) {
        constexpr int NB = NLINES * L / R;  // butterflies in this stage
        constexpr int BPT = (NB + NT - 1) / NT;
        float2 v[BPT][R];
#pragma unroll
        for (int b = 0; b < BPT; ++b) {
            const int u = tid + b * NT;
            if (NB % NT == 0 || u < NB) {
                const int line = LINES_FAST ? u % NLINES : u / (L / R);
                const int j = LINES_FAST ? u / NLINES : u % (L / R);
#pragma unroll
                for (int r = 0; r < R; ++r) v[b][r] = lds[addr(line, j + r * (L / R))];
                if (NS > 1) {
                    const int k = j % NS;
#pragma unroll
                    for (int r = 1; r < R; ++r) v[b][r] = cmul(v[b][r], tw[r * k * (L / (NS * R))]);
                }
                dft<R, DIR>(v[b]);
            }
        }
        lds_barrier();
#pragma unroll
        for (int b = 0; b < BPT; ++b) {
            const int u = tid + b * NT;
            if (NB % NT == 0 || u < NB) {
                const int line = LINES_FAST ? u % NLINES : u / (L / R);
                const int j = LINES_FAST ? u / NLINES : u % (L / R);
                const int d = (j / NS) * NS * R + (j % NS);
#pragma unroll
                for (int r = 0; r < R; ++r) lds[addr(line, d + r * NS)] = v[b][r];
            }
        }
        lds_barrier();
    }

    // Full transform; the caller has synchronised after filling lds and tw.
    __device__ static __forceinline__ void run(float2* lds, const float2* tw, int tid) {
        constexpr int R0 = (LOG2L % 3) ? (1 << (LOG2L % 3)) : 8;
        stage<R0, 1>(lds, tw, tid);
        if constexpr (LOG2L > 3) stage<8, R0>(lds, tw, tid);
        if constexpr (LOG2L > 6) stage<8, R0 * 8>(lds, tw, tid);
        if constexpr (LOG2L > 9) stage<8, R0 * 64>(lds, tw, tid);
    }
};

// Twiddles w^k = e^{DIR 2 pi i k / L}, k < L, from the FP64 table rounded once (tw64 holds cos, sin).
template <int LOG2L>
__device__ __forceinline__ void load_twiddles(float2* tw, const float2* __restrict__ table, int tid, int nt) {
    for (int k = tid; k < (1 << LOG2L); k += nt) tw[k] = table[k];
}

}  // namespace hermes_fft
