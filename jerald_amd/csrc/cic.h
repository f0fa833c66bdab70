// Device-side CIC building blocks shared by the CIC entry points and the
// fused LDL layer kernels (see cic.hip for the algorithm notes).
#pragma once

#include <cstring>

#include "ldl_common.h"

namespace ldl {

// ---------------------------------------------------------------- helpers
// Fixed-point exponent s (deposits are rint(w m 2^s)), from max |m| < 2^e:
//  * FP32 path: s = 30 - e, so every deposit fits in 31 bits and converts with
//    the native double -> int32 instruction; a cell (or the total) stays below
//    Np 2^30 < 2^62 for Np < 2^32;
//  * FP64 path: s = 61 - e - ceil(log2 Np), the largest scale for which the
//    total stays below 2^61.
template <typename T>
__device__ __forceinline__ int fixed_shift(unsigned long long maxbits, int64_t np) {
    double mx = __longlong_as_double((long long)maxbits);
    if (mx == 0.0 || np <= 0) return 0;
    int e = ilogb(mx) + 1;  // mx < 2^e (frexp exponent)
    if (sizeof(T) == 4) return 30 - e;
    int lg = np > 1 ? 64 - __clzll((unsigned long long)(np - 1)) : 0;
    return 61 - e - lg;
}

// The 8 fixed-point CIC deposits of one particle with fractions f and mass
// m at scale 2^s: q[4a+2b+c] = rint((w_x[a] w_y[b]) (w_z[c] (m 2^s))), and
// q[7] = rint(m 2^s) - (q[0] + ... + q[6]) so the deposits sum exactly to
// the rounded mass.  The CPU emulation (oracle/cic_oracle.c) performs the
// same operations in the same order.
template <typename T>
__device__ __forceinline__ void deposits(const T f[3], double m, int s, long long q[8]) {
    const double wx[2] = {(double)(T(1) - f[0]), (double)f[0]};
    const double wy[2] = {(double)(T(1) - f[1]), (double)f[1]};
    const double ms = ldexp(m, s);
    const double wzm[2] = {(double)(T(1) - f[2]) * ms, (double)f[2] * ms};
    const double wxy[4] = {wx[0] * wy[0], wx[0] * wy[1], wx[1] * wy[0], wx[1] * wy[1]};
    long long part = 0;
#pragma unroll
    for (int k = 0; k < 7; k++) {
        const double w = wxy[k >> 1] * wzm[k & 1];
        long long v;
        if (sizeof(T) == 4) v = (long long)__double2int_rn(w);  // |w| < 2^30
        else v = __double2ll_rn(w);
        q[k] = v;
        part += v;
    }
    const long long tot = sizeof(T) == 4 ? (long long)__double2int_rn(ms) : __double2ll_rn(ms);
    q[7] = tot - part;
}

static __global__ void k_fill_u64(unsigned long long* __restrict__ p, int64_t n, unsigned long long v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

template <typename T>
__global__ void k_fill(T* __restrict__ p, int64_t n, T v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// header[c] = bits of max |mass[:, c]| (as double); positive doubles order
// like their bit patterns, so an integer atomicMax is exact and order free.
// The (np, NC) array is read flat as coalesced 16-byte vectors of W = 16 /
// sizeof(T) elements, U vectors in flight per thread; flat element e belongs
// to component e % NC.  Maxima of |m| are exact in T.
template <typename T, int NC>
__global__ void __launch_bounds__(256) k_absmax(const T* __restrict__ mass, int64_t np, unsigned long long* __restrict__ hdr) {
    constexpr int W = 16 / sizeof(T), U = 4;
    struct alignas(16) Vec { T v[W]; };
    T mx[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) mx[c] = T(0);
    const int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, nth = (int64_t)gridDim.x * blockDim.x;
    const int64_t ne = np * NC;
    const int64_t nv = (reinterpret_cast<uintptr_t>(mass) & 15) == 0 ? ne / W : 0;
    const Vec* mv = reinterpret_cast<const Vec*>(mass);
    for (int64_t v0 = tid; v0 < nv; v0 += U * nth) {
        Vec q[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int64_t v = v0 + u * nth;
            q[u] = mv[v < nv ? v : v0];  // a repeated vector leaves the maxima unchanged
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int64_t v = v0 + u * nth < nv ? v0 + u * nth : v0;
            T rel[NC];  // rel[k]: max over the elements j of q[u] with j % NC == k
#pragma unroll
            for (int k = 0; k < NC; k++) rel[k] = T(0);
#pragma unroll
            for (int j = 0; j < W; j++) rel[j % NC] = fmax(rel[j % NC], fabs(q[u].v[j]));
            const int r0 = (int)((W * v) % NC);  // component of element 0
#pragma unroll
            for (int c = 0; c < NC; c++) {
                const int k = c - r0 < 0 ? c - r0 + NC : c - r0;
                T val = rel[0];
#pragma unroll
                for (int kk = 1; kk < NC; kk++) val = k == kk ? rel[kk] : val;
                mx[c] = fmax(mx[c], val);
            }
        }
    }
    for (int64_t e = nv * W + tid; e < ne; e += nth) {
        const int c = (int)(e % NC);
#pragma unroll
        for (int k = 0; k < NC; k++)
            if (k == c) mx[k] = fmax(mx[k], fabs(mass[e]));
    }
#pragma unroll
    for (int c = 0; c < NC; c++) {
        double v = (double)mx[c];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
        if ((threadIdx.x & 63) == 0 && v > 0.0) atomicMax(&hdr[c], (unsigned long long)__double_as_longlong(v));
    }
}

// Fixed-point deposit.  acc has NC consecutive meshes.
template <typename T, int NC>
__global__ void __launch_bounds__(256) k_paint_fixed(const T* __restrict__ pos, PosDec<T> pd, int64_t np,
                                                     const T* __restrict__ mass, T mscalar, Mesh m,
                                                     unsigned long long* __restrict__ acc,
                                                     const unsigned long long* __restrict__ hdr) {
    int s[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) s[c] = fixed_shift<T>(hdr[c], np);
    const int64_t cells = m.cells();
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < np; p += (int64_t)gridDim.x * blockDim.x) {
        int64_t idx[3];
        T f[3];
        pd.decode(pos, p, idx, f);
        const int64_t ii[2] = {idx[0], idx[0] + 1 == m.n0 ? 0 : idx[0] + 1};
        const int64_t jj[2] = {idx[1], idx[1] + 1 == m.n1 ? 0 : idx[1] + 1};
        const int64_t kk[2] = {idx[2], idx[2] + 1 == m.n2 ? 0 : idx[2] + 1};
#pragma unroll
        for (int c = 0; c < NC; c++) {
            const double mm = mass ? (double)mass[p * NC + c] : (double)mscalar;
            long long q[8];
            deposits<T>(f, mm, s[c], q);
            unsigned long long* a = acc + c * cells;
#pragma unroll
            for (int k = 0; k < 8; k++)
                if (q[k] != 0) atomicAdd(&a[(ii[k >> 2] * m.n1 + jj[(k >> 1) & 1]) * m.n2 + kk[k & 1]], (unsigned long long)q[k]);
        }
    }
}

// Floating-point atomic deposit (non-deterministic order), for comparison.
template <typename T, int NC>
__global__ void __launch_bounds__(256) k_paint_float(const T* __restrict__ pos, PosDec<T> pd, int64_t np,
                                                     const T* __restrict__ mass, T mscalar, T scale, Mesh m,
                                                     T* __restrict__ out) {
    const int64_t cells = m.cells();
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < np; p += (int64_t)gridDim.x * blockDim.x) {
        int64_t idx[3];
        T f[3];
        pd.decode(pos, p, idx, f);
        const T wx[2] = {T(1) - f[0], f[0]}, wy[2] = {T(1) - f[1], f[1]}, wz[2] = {T(1) - f[2], f[2]};
        const int64_t ii[2] = {idx[0], idx[0] + 1 == m.n0 ? 0 : idx[0] + 1};
        const int64_t jj[2] = {idx[1], idx[1] + 1 == m.n1 ? 0 : idx[1] + 1};
        const int64_t kk[2] = {idx[2], idx[2] + 1 == m.n2 ? 0 : idx[2] + 1};
#pragma unroll
        for (int c = 0; c < NC; c++) {
            const T mm = (mass ? mass[p * NC + c] : mscalar) * scale;
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int ax = q >> 2, by = (q >> 1) & 1, cz = q & 1;
                atomicAdd(&out[c * cells + (ii[ax] * m.n1 + jj[by]) * m.n2 + kk[cz]], wx[ax] * wy[by] * wz[cz] * mm);
            }
        }
    }
}

// out = acc * 2^-s * scale (rounded once), and acc reset to zero.
template <typename T, int NC>
__global__ void __launch_bounds__(256) k_fixed_to_float(unsigned long long* __restrict__ acc,
                                                        const unsigned long long* __restrict__ hdr, int64_t np,
                                                        int64_t cells, double scale, T* __restrict__ out) {
    double f[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) f[c] = ldexp(scale, -fixed_shift<T>(hdr[c], np));
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < cells; i += (int64_t)gridDim.x * blockDim.x) {
#pragma unroll
        for (int c = 0; c < NC; c++) {
            const long long v = (long long)acc[c * cells + i];
            out[c * cells + i] = (T)((double)v * f[c]);
            acc[c * cells + i] = 0ull;
        }
    }
}

// ---------------------------------------------------------------- gathers
// CIC stencil of one particle: the four (x, y) rows of the 8 corners and the
// lower z index.  Corner q = 4a + 2b + c sits at row[2a + b] + (k + c) mod N2.
template <typename T> struct Stencil {
    int64_t row[4];
    int64_t k;
    T f[3];
};

template <typename T>
__device__ __forceinline__ void locate(const PosDec<T>& pd, const Mesh& m, T x0, T x1, T x2, int64_t p, Stencil<T>& st) {
    int64_t idx[3];
    pd.decode_xyz(x0, x1, x2, p, idx, st.f);
    const int64_t i1 = idx[0] + 1 == m.n0 ? 0 : idx[0] + 1;
    const int64_t j1 = idx[1] + 1 == m.n1 ? 0 : idx[1] + 1;
    st.row[0] = (idx[0] * m.n1 + idx[1]) * m.n2;
    st.row[1] = (idx[0] * m.n1 + j1) * m.n2;
    st.row[2] = (i1 * m.n1 + idx[1]) * m.n2;
    st.row[3] = (i1 * m.n1 + j1) * m.n2;
    st.k = idx[2];
}

template <typename T>
__device__ __forceinline__ void locate(const PosDec<T>& pd, const Mesh& m, const T* __restrict__ pos, int64_t p, Stencil<T>& st) {
    const T* q = pos + 3 * p;
    locate(pd, m, q[0], q[1], q[2], p, st);
}

// Pair of adjacent values at an element-aligned (not pair-aligned) address.
template <typename T> struct alignas(sizeof(T)) Pair { T a, b; };

// The 8 corner values: one two-element load per row (z-adjacent corners are
// contiguous), plus a separate load of the wrapped corner when k = N2 - 1.
template <typename T>
__device__ __forceinline__ void gather8(const T* __restrict__ F, const Stencil<T>& st, int64_t n2, T v[8]) {
    const bool wrap = st.k + 1 == n2;
    const int64_t kk = wrap ? (n2 >= 2 ? n2 - 2 : 0) : st.k;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        if (n2 >= 2) {
            const Pair<T> pr = *reinterpret_cast<const Pair<T>*>(F + st.row[r] + kk);
            v[2 * r] = wrap ? pr.b : pr.a;
            v[2 * r + 1] = pr.b;
        } else {
            v[2 * r] = v[2 * r + 1] = F[st.row[r]];
        }
    }
    if (wrap && n2 >= 2) {
#pragma unroll
        for (int r = 0; r < 4; r++) v[2 * r + 1] = F[st.row[r]];
    }
}

// ---- power-of-two meshes (every FFT mesh: axes 2^k <= 1024, at most 2^30
// cells): 32-bit byte offsets from a uniform base pointer, wraps by masking.
struct MeshP2 {
    uint32_t msk[3];   // n - 1 per axis
    uint32_t lg1, lg2;  // log2 n1, log2 n2
};
inline MeshP2 mesh_p2(const Mesh& m) {
    auto lg = [](int64_t v) { uint32_t k = 0; while ((int64_t(1) << k) < v) k++; return k; };
    MeshP2 q;
    q.msk[0] = (uint32_t)(m.n0 - 1);
    q.msk[1] = (uint32_t)(m.n1 - 1);
    q.msk[2] = (uint32_t)(m.n2 - 1);
    q.lg1 = lg(m.n1);
    q.lg2 = lg(m.n2);
    return q;
}

template <typename T> struct StencilP2 {
    uint32_t row[4];  // byte offsets of the four (x, y) rows
    uint32_t k;       // lower z index
    T f[3];
};

template <typename T>
__device__ __forceinline__ void locate_p2(const PosDec<T>& pd, const MeshP2& mp, T x0, T x1, T x2, int64_t p, StencilP2<T>& st) {
    uint32_t idx[3];
    pd.decode_p2(x0, x1, x2, p, mp.msk, idx, st.f);
    const uint32_t i1 = (idx[0] + 1u) & mp.msk[0], j1 = (idx[1] + 1u) & mp.msk[1];
    constexpr uint32_t lb = sizeof(T) == 4 ? 2 : 3;
    st.row[0] = ((idx[0] << mp.lg1) | idx[1]) << (mp.lg2 + lb);
    st.row[1] = ((idx[0] << mp.lg1) | j1) << (mp.lg2 + lb);
    st.row[2] = ((i1 << mp.lg1) | idx[1]) << (mp.lg2 + lb);
    st.row[3] = ((i1 << mp.lg1) | j1) << (mp.lg2 + lb);
    st.k = idx[2];
}

// Corner values of NF fields on a power-of-two mesh: one two-element load per
// (field, row) -- z-adjacent corners are contiguous -- all issued before any
// is used; lanes with k = N2 - 1 load the pair ending at k and then the wrapped
// corner (N2 >= 2).
template <typename T, int NF>
__device__ __forceinline__ void gather_p2(const T* const (&F)[NF], const StencilP2<T>& st, const MeshP2& mp, T (&v)[NF][8]) {
    const bool wrap = st.k == mp.msk[2];
    const uint32_t kb = (wrap ? st.k - 1u : st.k) * (uint32_t)sizeof(T);
    Pair<T> pr[NF][4];
#pragma unroll
    for (int f = 0; f < NF; f++)
#pragma unroll
        for (int r = 0; r < 4; r++) pr[f][r] = *reinterpret_cast<const Pair<T>*>(reinterpret_cast<const char*>(F[f]) + (st.row[r] + kb));
    T w[NF][4];
#pragma unroll
    for (int f = 0; f < NF; f++)
#pragma unroll
        for (int r = 0; r < 4; r++) w[f][r] = T(0);
    if (wrap) {
#pragma unroll
        for (int f = 0; f < NF; f++)
#pragma unroll
            for (int r = 0; r < 4; r++) w[f][r] = *reinterpret_cast<const T*>(reinterpret_cast<const char*>(F[f]) + st.row[r]);
    }
#pragma unroll
    for (int f = 0; f < NF; f++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            v[f][2 * r] = wrap ? pr[f][r].b : pr[f][r].a;
            v[f][2 * r + 1] = wrap ? w[f][r] : pr[f][r].b;
        }
}

// Three field pointers passed by value: each is a uniform kernel argument, so
// the gathers address all three as base + 32-bit offset.
template <typename T> struct Fields3 {
    const T* f[3];
};
template <typename T>
inline Fields3<T> fields3(const T* F3, int64_t cells) {
    return Fields3<T>{{F3, F3 + cells, F3 + 2 * cells}};
}

// Mesh geometry policies of the gather kernels: GeomAny works on every mesh;
// GeomP2 on power-of-two meshes whose byte size fits 32 bits (p2_ok).
template <typename T> struct GeomAny {
    using St = Stencil<T>;
    Mesh m;
    __device__ __forceinline__ void locate(const PosDec<T>& pd, T x0, T x1, T x2, int64_t p, St& st) const {
        ldl::locate(pd, m, x0, x1, x2, p, st);
    }
    template <int NF>
    __device__ __forceinline__ void gather(const T* const (&F)[NF], const St& st, T (&v)[NF][8]) const {
#pragma unroll
        for (int f = 0; f < NF; f++) gather8(F[f], st, m.n2, v[f]);
    }
};
template <typename T> struct GeomP2 {
    using St = StencilP2<T>;
    MeshP2 mp;
    __device__ __forceinline__ void locate(const PosDec<T>& pd, T x0, T x1, T x2, int64_t p, St& st) const {
        locate_p2(pd, mp, x0, x1, x2, p, st);
    }
    template <int NF>
    __device__ __forceinline__ void gather(const T* const (&F)[NF], const St& st, T (&v)[NF][8]) const { gather_p2<T, NF>(F, st, mp, v); }
};
template <typename T>
inline bool p2_ok(const Mesh& m) {
    return is_pow2(m.n0) && is_pow2(m.n1) && is_pow2(m.n2) && m.n2 >= 2 &&
           m.cells() * (int64_t)sizeof(T) <= (int64_t(1) << 32);
}
// Calls f(geom) with the fastest policy that fits the mesh.
template <typename T, typename Fn>
inline void with_geom(const Mesh& m, Fn&& f) {
    if (p2_ok<T>(m)) f(GeomP2<T>{mesh_p2(m)});
    else f(GeomAny<T>{m});
}

// value (deriv < 0) or derivative along ``deriv`` (box units)
template <typename T>
__device__ __forceinline__ T interp(const T v[8], const T f[3], int deriv, const T cellinv[3]) {
    T w[3][2];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        if (d == deriv) { w[d][0] = -cellinv[d]; w[d][1] = cellinv[d]; }
        else { w[d][0] = T(1) - f[d]; w[d][1] = f[d]; }
    }
    T acc = T(0);
#pragma unroll
    for (int q = 0; q < 8; q++) acc += w[0][q >> 2] * w[1][(q >> 1) & 1] * w[2][q & 1] * v[q];
    return acc;
}

// value and full gradient from the same 8 corner values
template <typename T>
__device__ __forceinline__ void interp_vg(const T v[8], const T f[3], const T cellinv[3], T& val, T g[3]) {
    const T t[3] = {T(1) - f[0], T(1) - f[1], T(1) - f[2]};
    T a00 = t[2] * v[0] + f[2] * v[1], a01 = t[2] * v[2] + f[2] * v[3];
    T a10 = t[2] * v[4] + f[2] * v[5], a11 = t[2] * v[6] + f[2] * v[7];
    T dz00 = v[1] - v[0], dz01 = v[3] - v[2], dz10 = v[5] - v[4], dz11 = v[7] - v[6];
    T b0 = t[1] * a00 + f[1] * a01, b1 = t[1] * a10 + f[1] * a11;
    val = t[0] * b0 + f[0] * b1;
    g[0] = (b1 - b0) * cellinv[0];
    g[1] = (t[0] * (a01 - a00) + f[0] * (a11 - a10)) * cellinv[1];
    g[2] = (t[0] * (t[1] * dz00 + f[1] * dz01) + f[0] * (t[1] * dz10 + f[1] * dz11)) * cellinv[2];
}

template <typename T, typename G>
__global__ void __launch_bounds__(256) k_readout(const T* __restrict__ pos, PosDec<T> pd, int64_t np, Sweep sw, const T* __restrict__ F,
                                                 G geo, int deriv, T* __restrict__ out) {
    LDL_SWEEP(sw, np, p) {
        typename G::St st;
        geo.locate(pd, pos[3 * p], pos[3 * p + 1], pos[3 * p + 2], p, st);
        const T* fs[1] = {F};
        T v[1][8];
        geo.template gather<1>(fs, st, v);
        out[p] = interp(v[0], st.f, deriv, pd.cellinv);
    }
}

// 3 fields -> (Np,3)
template <typename T, typename G>
__global__ void __launch_bounds__(256) k_readout3(const T* __restrict__ pos, PosDec<T> pd, int64_t np, Sweep sw, Fields3<T> F3,
                                                  G geo, T* __restrict__ out) {
    LDL_SWEEP(sw, np, p) {
        typename G::St st;
        geo.locate(pd, pos[3 * p], pos[3 * p + 1], pos[3 * p + 2], p, st);
        T v[3][8];
        geo.template gather<3>(F3.f, st, v);
#pragma unroll
        for (int d = 0; d < 3; d++) out[3 * p + d] = interp(v[d], st.f, -1, pd.cellinv);
    }
}

// gradient of one field -> (Np,3), times pscale[p]
template <typename T, typename G>
__global__ void __launch_bounds__(256) k_readout_grad(const T* __restrict__ pos, PosDec<T> pd, int64_t np, Sweep sw, const T* __restrict__ F,
                                                      G geo, const T* __restrict__ pscale, T* __restrict__ out) {
    LDL_SWEEP(sw, np, p) {
        typename G::St st;
        geo.locate(pd, pos[3 * p], pos[3 * p + 1], pos[3 * p + 2], p, st);
        const T* fs[1] = {F};
        T v[1][8], val, g[3];
        geo.template gather<1>(fs, st, v);
        interp_vg(v[0], st.f, pd.cellinv, val, g);
        const T s = pscale ? pscale[p] : T(1);
#pragma unroll
        for (int d = 0; d < 3; d++) out[3 * p + d] = g[d] * s;
    }
}

// sum_d v[:,d] grad(F_d) -> (Np,3)
template <typename T, typename G>
__global__ void __launch_bounds__(256) k_readout_grad3(const T* __restrict__ pos, PosDec<T> pd, int64_t np, Sweep sw, Fields3<T> F3,
                                                       const T* __restrict__ v3, G geo, T* __restrict__ out) {
    LDL_SWEEP(sw, np, p) {
        typename G::St st;
        geo.locate(pd, pos[3 * p], pos[3 * p + 1], pos[3 * p + 2], p, st);
        T v[3][8];
        geo.template gather<3>(F3.f, st, v);
        T acc[3] = {T(0), T(0), T(0)};
#pragma unroll
        for (int d = 0; d < 3; d++) {
            T val, g[3];
            interp_vg(v[d], st.f, pd.cellinv, val, g);
            const T w = v3[3 * p + d];
#pragma unroll
            for (int e = 0; e < 3; e++) acc[e] += w * g[e];
        }
#pragma unroll
        for (int e = 0; e < 3; e++) out[3 * p + e] = acc[e];
    }
}

struct TileGeom {
    int ts;              // tile side in cells (a power of two)
    int tsh;             // log2(ts)
    int64_t nt[3];       // tiles per axis (last tile may be partial)
    int64_t ntiles;
    int64_t hstride;     // halo entries per tile, 3 ts^2 + 3 ts + 1
    int64_t lstride;     // low-face entries per tile, 3 ts^2 - 3 ts + 1
    __host__ __device__ int64_t extent(int d, int64_t t, const Mesh& m) const {
        const int64_t n = d == 0 ? m.n0 : d == 1 ? m.n1 : m.n2;
        const int64_t lo = t * ts;
        return n - lo < ts ? n - lo : ts;
    }
};

inline TileGeom make_tiles(const Mesh& m, int ts) {
    TileGeom g;
    g.ts = ts;
    g.tsh = 0;
    while ((1 << g.tsh) < ts) g.tsh++;
    g.nt[0] = (m.n0 + ts - 1) / ts;
    g.nt[1] = (m.n1 + ts - 1) / ts;
    g.nt[2] = (m.n2 + ts - 1) / ts;
    g.ntiles = g.nt[0] * g.nt[1] * g.nt[2];
    g.hstride = 3ll * ts * ts + 3ll * ts + 1;
    g.lstride = 3ll * ts * ts - 3ll * ts + 1;
    return g;
}

// Position of box cell (a, b, c) -- on one of the +1 faces -- in a tile's
// halo record; ex, ey, ez are the tile's extents.
__device__ __forceinline__ int64_t halo_index(int64_t a, int64_t b, int64_t c, int64_t ex, int64_t ey, int64_t ez) {
    if (a == ex) return b * (ez + 1) + c;
    if (b == ey) return (ey + 1) * (ez + 1) + a * (ez + 1) + c;
    return (ey + 1) * (ez + 1) + ex * (ez + 1) + a * ey + b;
}

// Counting-sort pass 1: tile key of each particle and its rank among the
// particles of the same tile.  Consecutive lanes hold consecutive particles,
// and particles stored in lattice order fall into a few runs of equal keys per
// wavefront: each run's first lane reserves the run's ranks with one global
// atomic and broadcasts the base to the run.  Rank order inside a tile depends
// on atomic order, which no paint result depends on (integer sums).
template <typename T, int PPT>
__global__ void __launch_bounds__(256) k_bin_keys(const T* __restrict__ pos, PosDec<T> pd, int64_t np, Mesh m, TileGeom tg,
                                                  unsigned* __restrict__ counts, unsigned* __restrict__ key,
                                                  unsigned* __restrict__ slot) {
    constexpr int CH = 256 * PPT;
    constexpr unsigned EMPTY = 0xFFFFFFFFu;
    const int lane = threadIdx.x & 63;
    const unsigned long long upto = lane == 63 ? ~0ull : (2ull << lane) - 1ull;  // lanes 0..lane
    const int64_t nchunks = (np + CH - 1) / CH;
    for (int64_t ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
        unsigned k[PPT], base[PPT];
        int hl[PPT];
#pragma unroll
        for (int q = 0; q < PPT; q++) {
            const int64_t p = ch * CH + q * 256 + threadIdx.x;
            k[q] = EMPTY;
            if (p < np) {
                int64_t idx[3];
                T f[3];
                pd.decode(pos, p, idx, f);
                k[q] = (((unsigned)idx[0] >> tg.tsh) * (unsigned)tg.nt[1] + ((unsigned)idx[1] >> tg.tsh)) * (unsigned)tg.nt[2] +
                       ((unsigned)idx[2] >> tg.tsh);
            }
        }
        // run heads, run lengths and one atomic per run (all issued before any result is used)
#pragma unroll
        for (int q = 0; q < PPT; q++) {
            const unsigned prev = __shfl_up(k[q], 1);
            const bool head = lane == 0 || prev != k[q];
            const unsigned long long heads = __ballot(head);
            hl[q] = 63 - __clzll(heads & upto);
            const unsigned long long after = heads & ~upto;
            const int next = after ? __ffsll((long long)after) - 1 : 64;
            base[q] = 0u;
            if (head && k[q] != EMPTY) base[q] = atomicAdd(&counts[k[q]], (unsigned)(next - lane));
        }
#pragma unroll
        for (int q = 0; q < PPT; q++) {
            const unsigned b = __shfl(base[q], hl[q]);
            const int64_t p = ch * CH + q * 256 + threadIdx.x;
            if (p < np) {
                key[p] = k[q];
                slot[p] = b + (unsigned)(lane - hl[q]);
            }
        }
    }
}

// Exclusive scan of the tile counts into offsets[0..ntiles]; counts reset.
static __global__ void __launch_bounds__(1024) k_scan_counts(unsigned* __restrict__ counts, unsigned* __restrict__ offsets, int64_t n) {
    __shared__ unsigned long long part[1024];
    const int t = threadIdx.x;
    const int64_t per = (n + 1023) / 1024;
    const int64_t b = t * per, e = b + per < n ? b + per : n;
    unsigned long long s = 0;
    for (int64_t i = b; i < e; i++) s += counts[i];
    part[t] = s;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {  // inclusive Hillis-Steele scan
        unsigned long long v = t >= o ? part[t - o] : 0ull;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    unsigned long long run = t ? part[t - 1] : 0ull;
    for (int64_t i = b; i < e; i++) {
        offsets[i] = (unsigned)run;
        run += counts[i];
        counts[i] = 0u;
    }
    if (t == 1023) offsets[n] = (unsigned)part[1023];
}

// Counting-sort pass 2: particle indices grouped by tile.
static __global__ void __launch_bounds__(256) k_bin_scatter(int64_t np, const unsigned* __restrict__ key, const unsigned* __restrict__ slot,
                                                     const unsigned* __restrict__ offsets, unsigned* __restrict__ sorted) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < np; p += (int64_t)gridDim.x * blockDim.x)
        sorted[offsets[key[p]] + slot[p]] = (unsigned)p;
}

// Position of core cell (a, b, c) with a == 0, b == 0 or c == 0 in a tile's
// low-face record (ex, ey, ez: the tile's extents); inverse: low_cell.
__device__ __forceinline__ int low_index(int a, int b, int c, int ex, int ey, int ez) {
    if (a == 0) return b * ez + c;
    if (b == 0) return ey * ez + (a - 1) * ez + c;
    return ey * ez + (ex - 1) * ez + (a - 1) * (ey - 1) + (b - 1);
}
__device__ __forceinline__ void low_cell(int i, int ex, int ey, int ez, int& a, int& b, int& c) {
    if (i < ey * ez) { a = 0; b = i / ez; c = i - b * ez; return; }
    i -= ey * ez;
    if (i < (ex - 1) * ez) { a = 1 + i / ez; b = 0; c = i - (a - 1) * ez; return; }
    i -= (ex - 1) * ez;
    a = 1 + i / (ey - 1);
    b = 1 + (i - (a - 1) * (ey - 1));
    c = 0;
}

// One workgroup per tile: LDS fixed-point deposit.  Interior cells (no
// neighbour contributes) are final and written as FP values; low-face cells
// go to the ``low`` record and the +1 faces to the ``halo`` record, both as
// integers, for k_tiles_fixup.
template <typename T, int NC, int TS, int TB = 256>
__global__ void __launch_bounds__(TB) k_paint_tiles(const T* __restrict__ pos, PosDec<T> pd, int64_t np, const T* __restrict__ mass,
                                                     T mscalar, Mesh m, TileGeom tg, const unsigned* __restrict__ offsets,
                                                     const unsigned* __restrict__ sorted, unsigned long long* __restrict__ low,
                                                     unsigned long long* __restrict__ halo, const unsigned long long* __restrict__ hdr,
                                                     double scale, T* __restrict__ out) {
    constexpr int B = TS + 1, B3 = B * B * B;
    __shared__ unsigned long long box[NC * B3];
    // XCD-contiguous tile order: the tiles an XCD runs together are z-neighbours,
    // whose particles share cache lines of the Lagrangian-ordered arrays
    const int64_t t = xcd_tile(blockIdx.x, gridDim.x);
    const int64_t tz = t % tg.nt[2], ty = (t / tg.nt[2]) % tg.nt[1], tx = t / (tg.nt[2] * tg.nt[1]);
    const int64_t lo[3] = {tx * TS, ty * TS, tz * TS};
    const int ex = (int)tg.extent(0, tx, m), ey = (int)tg.extent(1, ty, m), ez = (int)tg.extent(2, tz, m);
    for (int i = threadIdx.x; i < NC * B3; i += TB) box[i] = 0ull;
    int s[NC];
    double fs[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) {
        s[c] = fixed_shift<T>(hdr[c], np);
        fs[c] = ldexp(scale, -s[c]);
    }
    __syncthreads();
    const unsigned j0 = offsets[t], j1 = offsets[t + 1];
    // KB particles per thread per round; all their loads are issued before any
    // deposit so the dependent sorted[j] -> pos[p] chains overlap
    constexpr int KB = 4;
    for (unsigned jb = j0; jb < j1; jb += TB * KB) {
        unsigned pi[KB];
#pragma unroll
        for (int k = 0; k < KB; k++) {
            const unsigned j = jb + k * TB + threadIdx.x;
            pi[k] = sorted[j < j1 ? j : j1 - 1];
        }
        T xs[KB][3];
        T ms[KB][NC];
#pragma unroll
        for (int k = 0; k < KB; k++) {
            const T* q = pos + 3 * (int64_t)pi[k];
            xs[k][0] = q[0];
            xs[k][1] = q[1];
            xs[k][2] = q[2];
#pragma unroll
            for (int c = 0; c < NC; c++) ms[k][c] = mass ? mass[(int64_t)pi[k] * NC + c] : mscalar;
        }
#pragma unroll
        for (int k = 0; k < KB; k++) {
        if (jb + k * TB + threadIdx.x >= j1) continue;
        const int64_t p = pi[k];
        int64_t idx[3];
        T f[3];
        pd.decode_xyz(xs[k][0], xs[k][1], xs[k][2], p, idx, f);
        const int a = (int)(idx[0] - lo[0]), b = (int)(idx[1] - lo[1]), cz = (int)(idx[2] - lo[2]);
        const int base = (a * B + b) * B + cz;
#pragma unroll
        for (int c = 0; c < NC; c++) {
            long long q[8];
            deposits<T>(f, (double)ms[k][c], s[c], q);
#pragma unroll
            for (int qq = 0; qq < 8; qq++)
                if (q[qq] != 0) atomicAdd(&box[c * B3 + base + (qq >> 2) * B * B + ((qq >> 1) & 1) * B + (qq & 1)], (unsigned long long)q[qq]);
        }
        }
    }
    __syncthreads();
    const int64_t cells = m.cells();
    const int64_t lstride = tg.lstride, hstride = tg.hstride;
    const int e1z = ez + 1, e1y = ey + 1;
    const int nbox = (ex + 1) * e1y * e1z;
    for (int i = threadIdx.x; i < nbox; i += TB) {
        const int cc = i % e1z, bb = (i / e1z) % e1y, aa = i / (e1z * e1y);
        const int li = (aa * B + bb) * B + cc;
        if (aa < ex && bb < ey && cc < ez) {
            if (aa && bb && cc) {
                const int64_t g = ((lo[0] + aa) * m.n1 + (lo[1] + bb)) * m.n2 + (lo[2] + cc);
#pragma unroll
                for (int c = 0; c < NC; c++) out[c * cells + g] = (T)((double)(long long)box[c * B3 + li] * fs[c]);
            } else {
                const int64_t h = t * lstride + low_index(aa, bb, cc, ex, ey, ez);
#pragma unroll
                for (int c = 0; c < NC; c++) low[c * tg.ntiles * lstride + h] = box[c * B3 + li];
            }
        } else {
            const int64_t h = t * hstride + halo_index(aa, bb, cc, ex, ey, ez);
#pragma unroll
            for (int c = 0; c < NC; c++) halo[c * tg.ntiles * hstride + h] = box[c * B3 + li];
        }
    }
}

// Low-face cells: own value + the +1 faces of the lower neighbour tiles,
// then the FP conversion.  One thread per low-face cell.
template <typename T, int NC>
__global__ void __launch_bounds__(256) k_tiles_fixup(const unsigned long long* __restrict__ low,
                                                     const unsigned long long* __restrict__ halo,
                                                     const unsigned long long* __restrict__ hdr, int64_t np, Mesh m,
                                                     TileGeom tg, double scale, T* __restrict__ out) {
    double f[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) f[c] = ldexp(scale, -fixed_shift<T>(hdr[c], np));
    const int64_t cells = m.cells();
    const unsigned total = (unsigned)(tg.ntiles * tg.lstride);  // < 2^31 for meshes up to 2048^3
    const unsigned ls = (unsigned)tg.lstride, nt1 = (unsigned)tg.nt[1], nt2 = (unsigned)tg.nt[2];
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
        const unsigned t = i / ls;
        const int li = (int)(i - t * ls);
        const unsigned tq = t / nt2;
        const int64_t tz = t - tq * nt2, ty = tq % nt1, tx = tq / nt1;
        const int ex = (int)tg.extent(0, tx, m), ey = (int)tg.extent(1, ty, m), ez = (int)tg.extent(2, tz, m);
        if (li >= ex * ey * ez - (ex - 1) * (ey - 1) * (ez - 1)) continue;  // partial tile: shorter record
        int a, b, c;
        low_cell(li, ex, ey, ez, a, b, c);
        long long v[NC];
#pragma unroll
        for (int k = 0; k < NC; k++) v[k] = (long long)low[k * tg.ntiles * tg.lstride + i];
#pragma unroll
        for (int q = 1; q < 8; q++) {
            const int sx = q >> 2, sy = (q >> 1) & 1, sz = q & 1;
            if ((sx && a) || (sy && b) || (sz && c)) continue;
            const int64_t nx = sx ? (tx == 0 ? tg.nt[0] - 1 : tx - 1) : tx;
            const int64_t ny = sy ? (ty == 0 ? tg.nt[1] - 1 : ty - 1) : ty;
            const int64_t nz = sz ? (tz == 0 ? tg.nt[2] - 1 : tz - 1) : tz;
            const int nex = (int)tg.extent(0, nx, m), ney = (int)tg.extent(1, ny, m), nez = (int)tg.extent(2, nz, m);
            const int64_t nt = (nx * tg.nt[1] + ny) * tg.nt[2] + nz;
            const int64_t h = nt * tg.hstride + halo_index(sx ? nex : a, sy ? ney : b, sz ? nez : c, nex, ney, nez);
#pragma unroll
            for (int k = 0; k < NC; k++) v[k] += (long long)halo[k * tg.ntiles * tg.hstride + h];
        }
        const int64_t g = ((tx * tg.ts + a) * m.n1 + (ty * tg.ts + b)) * m.n2 + (tz * tg.ts + c);
#pragma unroll
        for (int k = 0; k < NC; k++) out[k * cells + g] = (T)((double)v[k] * f[k]);
    }
}

size_t paint_ws_bytes(const Mesh& m, int nc, int64_t np);

template <typename T>
int paint_impl(const T* pos, const ldl_pos_desc* pdesc, int64_t np, const T* mass, double mscalar, int nc,
               const int64_t nmesh[3], double out_scale, int32_t flags, T* out, void* ws, size_t ws_bytes,
               hipStream_t st, unsigned* bins_out = nullptr, const unsigned* bins_in = nullptr);
// Particle sort by 16^3-cell tile that a 1-component paint can hand to a later
// 3-component paint of the same positions: offsets (ntiles + 1) then sorted
// particle indices (np), uint32.
size_t paint_bins_bytes(const Mesh& m, int64_t np);
template <typename T>
int readout_impl(const T* pos, const ldl_pos_desc* pdesc, int64_t np, const T* field, const int64_t nmesh[3],
                 int kind, int vjpdim, const T* aux, T* out, hipStream_t st);

}  // namespace ldl
