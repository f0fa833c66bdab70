// Device-side CIC building blocks shared by the CIC entry points and the
// fused LDL layer kernels (see cic.hip for the algorithm notes).
#pragma once

#include <cstring>
#include <type_traits>

#include "ldl_common.h"

namespace ldl {

// ---------------------------------------------------------------- helpers
// Fixed-point exponent s (deposits are splits of rint(m 2^s)), from max |m| < 2^e:
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
// m at scale 2^s, split hierarchically from the rounded mass M = rint(m 2^s):
// the x split gives X1 = trunc(f_x M) and X0 = M - X1, each X_a splits along
// y the same way, each of those along z, and q[4a+2b+c] is the z part c of
// (a, b).  Each split takes trunc(f V) (clamped into [0, V] or [V, 0]) and
// V minus it, so every deposit has the sign of m (a paint of non-negative
// masses is non-negative in every cell) and the deposits sum exactly to M.
// The products are formed in T (their rounding moves a deposit by at most a
// few units of 2^-s times f V 2^-24 in FP32; sign and sum are exact
// regardless).  FP32 deposits are int32 (|m 2^s| < 2^30);
// FP64 deposits are int64.  The CPU emulation (oracle/cic_oracle.c) performs
// the same operations in the same order.
template <typename T> using DepositT = typename std::conditional<sizeof(T) == 4, int, long long>::type;

template <typename T, typename D>
__device__ __forceinline__ D split_part(T fr, D V) {  // trunc(fr V), clamped between 0 and V
    const D v = (D)(fr * (T)V);
    if constexpr (sizeof(D) == 4) {
        // the median of v, 0 and V (one v_med3_i32): v has the sign of V
        // (fr >= 0), so this is the clamp below
        return max(min(v, D(0)), min(max(v, D(0)), V));
    } else {
        return V >= 0 ? (v > V ? V : v) : (v < V ? V : v);
    }
}

// The particle's rounded fixed-point mass rint(m 2^s).  FP32: m 2^s is exact
// in FP32 (|m 2^s| < 2^30; a result below the normal range rounds to 0 either
// way), so the rounding equals that of the FP64 product.
template <typename T>
__device__ __forceinline__ DepositT<T> fixed_mass(T m, int s) {
    if constexpr (sizeof(T) == 4) return (int)rintf(ldexpf(m, s));
    else return __double2ll_rn(ldexp(m, s));
}

// The 8 corner deposits of fixed-point mass M.
template <typename T>
__device__ __forceinline__ void deposits_of(const T f[3], DepositT<T> M, DepositT<T> q[8]) {
    using D = DepositT<T>;
    const D x1 = split_part<T, D>(f[0], M);
    const D xa[2] = {M - x1, x1};
#pragma unroll
    for (int a = 0; a < 2; a++) {
        const D y1 = split_part<T, D>(f[1], xa[a]);
        const D yb[2] = {xa[a] - y1, y1};
#pragma unroll
        for (int b = 0; b < 2; b++) {
            const D z1 = split_part<T, D>(f[2], yb[b]);
            q[4 * a + 2 * b] = yb[b] - z1;
            q[4 * a + 2 * b + 1] = z1;
        }
    }
}

template <typename T>
__device__ __forceinline__ void deposits(const T f[3], T m, int s, DepositT<T> q[8]) {
    deposits_of<T>(f, fixed_mass<T>(m, s), q);
}

static __global__ void k_copy_u64(const unsigned long long* __restrict__ src, unsigned long long* __restrict__ dst, int64_t n) {
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
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

// e % NC for NC in {1, 3} with 32-bit arithmetic (2^32 = 1 mod 3).
template <int NC>
__device__ __forceinline__ int flat_mod(uint64_t e) {
    static_assert(NC == 1 || NC == 3, "1 or 3 components");
    if (NC == 1) return 0;
    const uint32_t s = (uint32_t)(e >> 32) % 3u + (uint32_t)e % 3u;
    return (int)(s % 3u);
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
            const int r0 = flat_mod<NC>((uint64_t)v * W);  // component of element 0
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
        const int c = flat_mod<NC>((uint64_t)e);
#pragma unroll
        for (int k = 0; k < NC; k++)
            if (k == c) mx[k] = fmax(mx[k], fabs(mass[e]));
    }
    // wavefront, then workgroup maxima: one atomic per workgroup and component
    __shared__ double wmax[NC][4];
#pragma unroll
    for (int c = 0; c < NC; c++) {
        double v = (double)mx[c];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
        if ((threadIdx.x & 63) == 0) wmax[c][threadIdx.x >> 6] = v;
    }
    __syncthreads();
    if (threadIdx.x < NC) {
        const int c = threadIdx.x;
        const double v = fmax(fmax(wmax[c][0], wmax[c][1]), fmax(wmax[c][2], wmax[c][3]));
        if (v > 0.0) atomicMax(&hdr[c], (unsigned long long)__double_as_longlong(v));
    }
}

// max |w| * scale over a mesh into *slot as double bits: for a 3-component
// paint of the positions whose unit-mass density (times 1/scale) is w, the
// largest sum of CIC weights any cell receives.
// 16-byte loads, four in flight per thread (a 16-byte-aligned mesh; else one
// element at a time).
template <typename T>
__global__ void __launch_bounds__(256) k_mesh_absmax(const T* __restrict__ w, int64_t n, double scale, unsigned long long* __restrict__ slot) {
    using VT = typename std::conditional<sizeof(T) == 4, float4, double2>::type;
    constexpr int V = 16 / (int)sizeof(T);
    auto vmax = [](VT a) -> T {
        if constexpr (sizeof(T) == 4) return fmaxf(fmaxf(fabsf(a.x), fabsf(a.y)), fmaxf(fabsf(a.z), fabsf(a.w)));
        else return fmax(fabs(a.x), fabs(a.y));
    };
    const int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, nth = (int64_t)gridDim.x * blockDim.x;
    const int64_t nv = (reinterpret_cast<uintptr_t>(w) % 16 == 0) ? n / V : 0;
    const VT* wv = reinterpret_cast<const VT*>(w);
    T mx = T(0);
    int64_t i = tid;
    for (; i + 3 * nth < nv; i += 4 * nth) {
        const VT a = wv[i], b = wv[i + nth], c = wv[i + 2 * nth], d = wv[i + 3 * nth];
        mx = fmax(mx, fmax(fmax(vmax(a), vmax(b)), fmax(vmax(c), vmax(d))));
    }
    for (; i < nv; i += nth) mx = fmax(mx, vmax(wv[i]));
    for (int64_t j = nv * V + tid; j < n; j += nth) mx = fmax(mx, fabs(w[j]));
    double v = (double)mx * scale;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    __shared__ double wm[4];
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double b = fmax(fmax(wm[0], wm[1]), fmax(wm[2], wm[3]));
        if (b > 0.0) atomicMax(slot, (unsigned long long)__double_as_longlong(b));
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
            const T mm = mass ? mass[p * NC + c] : mscalar;
            DepositT<T> q[8];
            deposits<T>(f, mm, s[c], q);
            unsigned long long* a = acc + c * cells;
#pragma unroll
            for (int k = 0; k < 8; k++)
                if (q[k] != 0) atomicAdd(&a[(ii[k >> 2] * m.n1 + jj[(k >> 1) & 1]) * m.n2 + kk[k & 1]], (unsigned long long)(long long)q[k]);
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

// LU: lattice positions with unit steps and power-of-two strides
// (PosDec::lattice_unit_p2), decoded without run-time branches.
template <typename T, bool LU = false>
__device__ __forceinline__ void locate_p2(const PosDec<T>& pd, const MeshP2& mp, T x0, T x1, T x2, int64_t p, StencilP2<T>& st) {
    uint32_t idx[3];
    if constexpr (LU) pd.decode_lu(x0, x1, x2, p, mp.msk, idx, st.f);
    else pd.decode_p2(x0, x1, x2, p, mp.msk, idx, st.f);
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

// One cell of an interleaved 3-component field (x, y, z packed, 3 T per
// cell): one 12-byte load in FP32 (elements are 4-byte aligned), 24 bytes in
// FP64.
template <typename T> struct Cell3 {
    T v[3];
};
template <typename T>
__device__ __forceinline__ Cell3<T> load_cell3(const T* p) {
    return Cell3<T>{{p[0], p[1], p[2]}};
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
    // the three components of an interleaved field (cell c at F3i[3c .. 3c+2])
    __device__ __forceinline__ void gather3i(const T* __restrict__ F3i, const St& st, T (&v)[3][8]) const {
        const int64_t k1 = st.k + 1 == m.n2 ? 0 : st.k + 1;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const Cell3<T> a = load_cell3(F3i + 3 * (st.row[r] + st.k)), b = load_cell3(F3i + 3 * (st.row[r] + k1));
#pragma unroll
            for (int c = 0; c < 3; c++) {
                v[c][2 * r] = a.v[c];
                v[c][2 * r + 1] = b.v[c];
            }
        }
    }
};
template <typename T, bool LU = false> struct GeomP2 {
    using St = StencilP2<T>;
    MeshP2 mp;
    __device__ __forceinline__ void locate(const PosDec<T>& pd, T x0, T x1, T x2, int64_t p, St& st) const {
        locate_p2<T, LU>(pd, mp, x0, x1, x2, p, st);
    }
    template <int NF>
    __device__ __forceinline__ void gather(const T* const (&F)[NF], const St& st, T (&v)[NF][8]) const { gather_p2<T, NF>(F, st, mp, v); }
    // the three components of an interleaved field; byte offsets are the
    // 1-slot row offsets times 3, so the field's bytes must fit 32 bits (p2_ok3i)
    __device__ __forceinline__ void gather3i(const T* __restrict__ F3i, const St& st, T (&v)[3][8]) const {
        constexpr uint32_t cb = 3 * sizeof(T);
        const uint32_t k0 = st.k * cb, k1 = ((st.k + 1u) & mp.msk[2]) * cb;
        const char* b = reinterpret_cast<const char*>(F3i);
        Cell3<T> a[4], c[4];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            a[r] = load_cell3(reinterpret_cast<const T*>(b + (st.row[r] * 3u + k0)));
            c[r] = load_cell3(reinterpret_cast<const T*>(b + (st.row[r] * 3u + k1)));
        }
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
            for (int q = 0; q < 3; q++) {
                v[q][2 * r] = a[r].v[q];
                v[q][2 * r + 1] = c[r].v[q];
            }
    }
};
template <typename T>
inline bool p2_ok(const Mesh& m) {
    return is_pow2(m.n0) && is_pow2(m.n1) && is_pow2(m.n2) && m.n2 >= 2 &&
           m.cells() * (int64_t)sizeof(T) <= (int64_t(1) << 32);
}
// Calls f(geom) with the fastest policy that fits the mesh and the positions.
template <typename T, typename Fn>
inline void with_geom(const Mesh& m, const PosDec<T>& pd, Fn&& f) {
    if (p2_ok<T>(m) && pd.lattice_unit_p2()) f(GeomP2<T, true>{mesh_p2(m)});
    else if (p2_ok<T>(m)) f(GeomP2<T>{mesh_p2(m)});
    else f(GeomAny<T>{m});
}
// Same for kernels that also gather an interleaved 3-component field.
template <typename T>
inline bool p2_ok3i(const Mesh& m) {
    return p2_ok<T>(m) && m.cells() * 3 * (int64_t)sizeof(T) <= (int64_t(1) << 32);
}
template <typename T, typename Fn>
inline void with_geom3i(const Mesh& m, const PosDec<T>& pd, Fn&& f) {
    if (p2_ok3i<T>(m) && pd.lattice_unit_p2()) f(GeomP2<T, true>{mesh_p2(m)});
    else if (p2_ok3i<T>(m)) f(GeomP2<T>{mesh_p2(m)});
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
    int sh[3];      // log2 of the tile's side along each axis (cells)
    int64_t nt[3];  // tiles per axis (the last tile of an axis may be partial)
    int64_t ntiles;
    uint32_t n32[3];  // mesh sides (each < 2^31 on the tiled path)
    uint32_t fm[3];   // power-of-two mesh: min(tile side, n) - 1
    int lnt[3];       // power-of-two mesh: log2 nt
    int64_t nkeys() const { return ntiles * 8; }
    __host__ __device__ int64_t extent(int d, int64_t t, const Mesh& m) const {
        const int64_t n = d == 0 ? m.n0 : d == 1 ? m.n1 : m.n2;
        const int64_t ts = (int64_t)1 << sh[d];
        const int64_t lo = t * ts;
        return n - lo < ts ? n - lo : ts;
    }
};

// Tiles of 2^sh[0] x 2^sh[1] x 2^sh[2] cells.
inline TileGeom make_tiles(const Mesh& m, const int sh[3]) {
    TileGeom g;
    const int64_t n[3] = {m.n0, m.n1, m.n2};
    g.ntiles = 1;
    for (int d = 0; d < 3; d++) {
        g.sh[d] = sh[d];
        g.nt[d] = (n[d] + ((int64_t)1 << sh[d]) - 1) >> sh[d];
        g.ntiles *= g.nt[d];
        const int64_t ts = (int64_t)1 << sh[d];
        g.n32[d] = (uint32_t)n[d];
        g.fm[d] = (uint32_t)((ts < n[d] ? ts : n[d]) - 1);
        int l = 0;
        while (((int64_t)1 << l) < g.nt[d]) l++;
        g.lnt[d] = l;
    }
    return g;
}

// Position decode of the paint's sort and deposit kernels, chosen on the host
// (both kernels use the same one): any mesh; a power-of-two mesh; a
// power-of-two mesh with lattice positions of unit steps and power-of-two
// strides (PosDec::lattice_unit_p2), decoded without run-time branches.
constexpr int kDecAny = 0, kDecP2 = 1, kDecLU = 2;

// Lower CIC corner (32-bit cell indices) and fractions of particle p: on a
// power-of-two mesh with mask wraps, else with the general decode.
template <typename T, int DM>
__device__ __forceinline__ void cell_of(const PosDec<T>& pd, const MeshP2& mp, T x0, T x1, T x2, int64_t p, uint32_t id[3], T f[3]) {
    if constexpr (DM == kDecLU) {
        pd.decode_lu(x0, x1, x2, p, mp.msk, id, f);
    } else if constexpr (DM == kDecP2) {
        pd.decode_p2(x0, x1, x2, p, mp.msk, id, f);
    } else {
        int64_t idx[3];
        pd.decode_xyz(x0, x1, x2, p, idx, f);
#pragma unroll
        for (int d = 0; d < 3; d++) id[d] = (uint32_t)idx[d];
    }
}

// Sort key of a particle whose CIC lower corner is cell id: 8 * tile + face
// class, where face-class bit d is set when the corner lies on the tile's
// upper face along axis d -- exactly the particles whose stencil also reaches
// the next tile along d.  32-bit arithmetic: sides below 2^31, keys below
// 2^32.  On a power-of-two mesh the tile counts are powers of two and the
// upper face is where id + 1 is a multiple of min(tile side, n).
template <int DM>
__device__ __forceinline__ unsigned tile_key(const uint32_t id[3], const TileGeom& tg) {
    unsigned t[3], fc = 0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        t[d] = id[d] >> tg.sh[d];
        if constexpr (DM == kDecAny) {
            const uint32_t hi = (t[d] + 1u) << tg.sh[d];  // one past the tile's last cell
            if (id[d] + 1u == (hi < tg.n32[d] ? hi : tg.n32[d])) fc |= 1u << d;
        } else {
            if (((id[d] + 1u) & tg.fm[d]) == 0u) fc |= 1u << d;
        }
    }
    if constexpr (DM == kDecAny) return (((t[0] * (unsigned)tg.nt[1] + t[1]) * (unsigned)tg.nt[2] + t[2]) << 3) | fc;
    else return (((((t[0] << tg.lnt[1]) | t[1]) << tg.lnt[2]) | t[2]) << 3) | fc;
}

// Counting-sort pass 1: key of each particle, and the particle counts per
// key.  Consecutive lanes hold consecutive particles, and particles stored in
// lattice order fall into a few runs of equal keys per wavefront: each run's
// first lane adds the run's length to an LDS window of kSortWin counters
// from the chunk's smallest key (or, outside it, to the global count), and
// the window's nonzero counters are added to the global counts once per
// chunk.  All PPT positions of a thread are loaded before any is decoded
// (np >= 1).
constexpr int kSortWin = 1024;  // key window of the sort's per-workgroup counters

template <typename T, int PPT, int DM>
__global__ void __launch_bounds__(256) k_bin_keys(const T* __restrict__ pos, PosDec<T> pd, int64_t np, MeshP2 mp, TileGeom tg,
                                                  unsigned* __restrict__ counts, unsigned* __restrict__ key) {
    constexpr int CH = 256 * PPT, W = kSortWin;
    constexpr unsigned EMPTY = 0xFFFFFFFFu;
    __shared__ unsigned win[W];
    __shared__ unsigned wmin[4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const unsigned long long upto = lane == 63 ? ~0ull : (2ull << lane) - 1ull;  // lanes 0..lane
    const int64_t nchunks = (np + CH - 1) / CH;
    for (int64_t ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
        unsigned k[PPT];
        T xs[PPT][3];
#pragma unroll
        for (int q = 0; q < PPT; q++) {
            int64_t p = ch * CH + q * 256 + threadIdx.x;
            p = p < np ? p : np - 1;
#pragma unroll
            for (int d = 0; d < 3; d++) xs[q][d] = ld_stream(pos + 3 * p + d);
        }
#pragma unroll
        for (int q = 0; q < PPT; q++) {
            const int64_t p = ch * CH + q * 256 + threadIdx.x;
            k[q] = EMPTY;
            if (p < np) {
                uint32_t id[3];
                T f[3];
                cell_of<T, DM>(pd, mp, xs[q][0], xs[q][1], xs[q][2], p, id, f);
                k[q] = tile_key<DM>(id, tg);
                st_stream(key + p, k[q]);
            }
        }
        unsigned m = EMPTY;
#pragma unroll
        for (int q = 0; q < PPT; q++) m = k[q] < m ? k[q] : m;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const unsigned y = __shfl_xor(m, o);
            m = y < m ? y : m;
        }
        for (int i = threadIdx.x; i < W; i += 256) win[i] = 0u;
        if (lane == 0) wmin[wv] = m;
        __syncthreads();
        // the chunk's first particle exists (ch < nchunks), so kb is a key
        const unsigned kb0 = wmin[0] < wmin[1] ? wmin[0] : wmin[1], kb1 = wmin[2] < wmin[3] ? wmin[2] : wmin[3];
        const unsigned kb = kb0 < kb1 ? kb0 : kb1;
#pragma unroll
        for (int q = 0; q < PPT; q++) {
            const unsigned prev = __shfl_up(k[q], 1);
            const bool head = lane == 0 || prev != k[q];
            const unsigned long long heads = __ballot(head);
            const unsigned long long after = heads & ~upto;
            const int next = after ? __ffsll((long long)after) - 1 : 64;
            if (head && k[q] != EMPTY) {
                const unsigned len = (unsigned)(next - lane);
                if (k[q] - kb < (unsigned)W) atomicAdd(&win[k[q] - kb], len);
                else atomicAdd(&counts[k[q]], len);
            }
        }
        __syncthreads();
        for (int i = threadIdx.x; i < W; i += 256) {
            const unsigned c = win[i];
            if (c) atomicAdd(&counts[kb + (unsigned)i], c);
        }
        __syncthreads();  // the next chunk clears the window
    }
}

// Exclusive scan of the key counts into offsets[0..n] in two passes over
// segments of kScanSeg counts: segment sums, then each segment adds the sums
// of the segments before it and scans itself.
constexpr int kScanSeg = 8192;  // 1024 threads x 8 counts

__device__ __forceinline__ unsigned block_exclusive_scan_1024(unsigned v, unsigned* sh, unsigned& total) {
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    unsigned x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh[w] = x;
    __syncthreads();
    if (t < 64) {
        unsigned s = t < 16 ? sh[t] : 0u;
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) {
            const unsigned y = __shfl_up(s, o);
            if (lane >= o) s += y;
        }
        if (t < 16) sh[16 + t] = s;
    }
    __syncthreads();
    total = sh[16 + 15];
    const unsigned r = x - v + (w ? sh[16 + w - 1] : 0u);
    __syncthreads();
    return r;
}

static __global__ void __launch_bounds__(1024) k_scan_sums(const unsigned* __restrict__ counts, int64_t n, unsigned* __restrict__ segsum) {
    __shared__ unsigned sh[32];
    const int64_t b0 = blockIdx.x * (int64_t)kScanSeg;
    unsigned v = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const int64_t j = b0 + i * 1024 + threadIdx.x;
        if (j < n) v += counts[j];
    }
    unsigned tot;
    block_exclusive_scan_1024(v, sh, tot);
    if (threadIdx.x == 0) segsum[blockIdx.x] = tot;
}

// Also writes the offsets to cursor (which may alias counts: a workgroup
// reads its segment of counts whole before it writes).
static __global__ void __launch_bounds__(1024) k_scan_apply(const unsigned* counts, int64_t n, const unsigned* __restrict__ segsum,
                                                          unsigned* __restrict__ offsets, unsigned* cursor) {
    __shared__ unsigned sh[32];
    __shared__ unsigned seg[kScanSeg];
    __shared__ unsigned before;
    const int64_t b0 = blockIdx.x * (int64_t)kScanSeg;
    if (threadIdx.x < 64) {  // sum of the earlier segments' totals
        unsigned s = 0;
        for (unsigned i = threadIdx.x; i < blockIdx.x; i += 64) s += segsum[i];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
        if (threadIdx.x == 0) before = s;
    }
#pragma unroll
    for (int i = 0; i < 8; i++) {  // coalesced load of the segment
        const int64_t j = b0 + i * 1024 + threadIdx.x;
        seg[i * 1024 + threadIdx.x] = j < n ? counts[j] : 0u;
    }
    __syncthreads();
    unsigned c[8], v = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {  // thread t scans counts 8t..8t+7 of the segment
        c[i] = seg[8 * threadIdx.x + i];
        v += c[i];
    }
    unsigned tot;
    unsigned run = block_exclusive_scan_1024(v, sh, tot) + before;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        seg[8 * threadIdx.x + i] = run;
        run += c[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const int64_t j = b0 + i * 1024 + threadIdx.x;
        if (j < n) {
            offsets[j] = seg[i * 1024 + threadIdx.x];
            cursor[j] = seg[i * 1024 + threadIdx.x];
        }
    }
    if (b0 + kScanSeg >= n && threadIdx.x == 0) offsets[n] = before + tot;
}

// Counting-sort pass 2: particle indices grouped by key.  cursor[k] starts
// at the key's offset.  A workgroup reserves the ranks of a chunk in two
// levels: each run of equal keys in a wavefront takes its local ranks with one
// LDS atomic on a window of kSortWin counters from the chunk's smallest key
// (the chunk's rows share their tiles, so its runs fall on a few dozen keys),
// then each key the window holds takes a block of global ranks with one
// device atomic.  A run whose key lies outside the window reserves with a
// device atomic of its own.  Particles are placed in lane order within a run;
// rank order inside a key depends on atomic order, which no paint result
// depends on (integer sums).  Whole wavefronts step together (shuffles over
// all lanes); every workgroup runs the same number of barriers per chunk.
template <int PPT>
__global__ void __launch_bounds__(256) k_bin_scatter(int64_t np, const unsigned* __restrict__ key, unsigned* __restrict__ cursor,
                                                     unsigned* __restrict__ sorted) {
    constexpr int CH = 256 * PPT, W = kSortWin;
    constexpr unsigned EMPTY = 0xFFFFFFFFu;
    __shared__ unsigned win[W];  // local counts, then the global base of each key
    __shared__ unsigned wmin[4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const unsigned long long upto = lane == 63 ? ~0ull : (2ull << lane) - 1ull;  // lanes 0..lane
    // XCD-contiguous chunks: each XCD writes the ranks of its own band of
    // particles, so the partial cache lines of a key's range fill in one L2
    const int64_t nchunks = (np + CH - 1) / CH, per = (nchunks + gridDim.x - 1) / gridDim.x;
    const int64_t c0 = xcd_tile(blockIdx.x, gridDim.x) * per, c1 = c0 + per < nchunks ? c0 + per : nchunks;
    for (int64_t ch = c0; ch < c1; ch++) {
        unsigned k[PPT], loc[PPT];
        unsigned m = EMPTY;
#pragma unroll
        for (int q = 0; q < PPT; q++) {
            const int64_t p = ch * CH + q * 256 + threadIdx.x;
            k[q] = p < np ? key[p] : EMPTY;
            m = k[q] < m ? k[q] : m;
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const unsigned y = __shfl_xor(m, o);
            m = y < m ? y : m;
        }
        for (int i = threadIdx.x; i < W; i += 256) win[i] = 0u;
        if (lane == 0) wmin[wv] = m;
        __syncthreads();
        // the chunk's first particle exists (ch < nchunks), so kb is a key
        const unsigned kb0 = wmin[0] < wmin[1] ? wmin[0] : wmin[1], kb1 = wmin[2] < wmin[3] ? wmin[2] : wmin[3];
        const unsigned kb = kb0 < kb1 ? kb0 : kb1;
#pragma unroll
        for (int q = 0; q < PPT; q++) {
            const unsigned prev = __shfl_up(k[q], 1);
            const bool head = lane == 0 || prev != k[q];
            const unsigned long long heads = __ballot(head);
            const int hl = 63 - __clzll(heads & upto);
            const unsigned long long after = heads & ~upto;
            const int next = after ? __ffsll((long long)after) - 1 : 64;
            unsigned b = 0u;
            if (head && k[q] != EMPTY) {
                const unsigned len = (unsigned)(next - lane);
                b = k[q] - kb < (unsigned)W ? atomicAdd(&win[k[q] - kb], len) : atomicAdd(&cursor[k[q]], len);
            }
            loc[q] = __shfl(b, hl) + (unsigned)(lane - hl);
        }
        __syncthreads();
        for (int i = threadIdx.x; i < W; i += 256) {
            const unsigned c = win[i];
            if (c) win[i] = atomicAdd(&cursor[kb + (unsigned)i], c);
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < PPT; q++) {
            const int64_t p = ch * CH + q * 256 + threadIdx.x;
            if (p < np) sorted[(k[q] - kb < (unsigned)W ? win[k[q] - kb] : 0u) + loc[q]] = (unsigned)p;
        }
        __syncthreads();  // the next chunk clears the window
    }
}

// One workgroup per tile: the tile's cells are summed completely in LDS and
// written once.  The workgroup reads its own particles (all 8 face classes)
// and, from the 7 lower neighbour tiles, the particles whose stencil reaches
// into it (the face classes holding those neighbours' upper faces); every
// particle deposits only its corners that lie in the tile.  A particle read by
// several tiles computes the same 8 integer deposits in each.
//
// ACC32 (FP32 only): 32-bit accumulators, half the LDS.  hdr[3] then holds W,
// the largest CIC weight sum of any cell, and every shift drops by g with
// 2^g > 2 W + 1, so a cell's |sum| stays below 2^29 plus 4 per contribution:
// the int32 sums equal the int64 ones of the same deposits.
//
// vmax (FP32, one component, may be NULL): max |cell| of the painted mesh, as
// FP32 bits, by one atomicMax per workgroup into a zeroed slot.
template <typename T, int NC, int TX, int TY, int TZ, int TB, int DM, bool ACC32 = false>
__global__ void __launch_bounds__(TB) k_paint_tiles(const T* __restrict__ pos, PosDec<T> pd, int64_t np, const T* __restrict__ mass,
                                                     T mscalar, Mesh m, MeshP2 mp, TileGeom tg, const unsigned* __restrict__ offsets,
                                                     const unsigned* __restrict__ sorted, const unsigned long long* __restrict__ hdr,
                                                     double scale, T* __restrict__ out, unsigned* __restrict__ vmax = nullptr) {
    static_assert(!ACC32 || sizeof(T) == 4, "32-bit accumulation takes the 32-bit FP32 deposits");
    using Acc = typename std::conditional<ACC32, unsigned int, unsigned long long>::type;
    using SAcc = typename std::conditional<ACC32, int, long long>::type;
    constexpr int B3 = TX * TY * TZ;
    constexpr int NR = 20;  // 1 own range + 12 + 6 + 1 neighbour ranges
    __shared__ Acc box[NC * B3];
    __shared__ unsigned rstart[NR], rpre[NR + 1], rn0;
    __shared__ unsigned char rcode[NR];
    // XCD-contiguous tile order: the tiles an XCD runs together are z-neighbours,
    // whose particles share cache lines of the Lagrangian-ordered arrays
    const int64_t t = xcd_tile(blockIdx.x, gridDim.x);
    const int64_t tz = t % tg.nt[2], ty = (t / tg.nt[2]) % tg.nt[1], tx = t / (tg.nt[2] * tg.nt[1]);
    const int64_t lo[3] = {tx * TX, ty * TY, tz * TZ};
    const int ex = (int)tg.extent(0, tx, m), ey = (int)tg.extent(1, ty, m), ez = (int)tg.extent(2, tz, m);
    for (int i = threadIdx.x; i < NC * B3; i += TB) box[i] = Acc(0);
    if (threadIdx.x < 64) {
        // range r (lane r < NR): r = 0 is the own tile (all face classes); the
        // others are (dir, fc) pairs -- bit d of dir: the tile one lower along
        // axis d (periodic) -- with (fc & dir) == dir, packed 6 bits per range
        // as dir | fc << 3
        constexpr unsigned long long kLo = 0x6fac9a4b9a59240ull, kHi = 0xffedbdb7cd2c93bull;
        const int r = threadIdx.x;
        unsigned len = 0, start = 0, dir = 0;
        if (r < NR) {
            const unsigned v = (unsigned)((r < 10 ? kLo >> (6 * r) : kHi >> (6 * (r - 10))) & 63ull);
            dir = v & 7u;
            const unsigned fc = v >> 3;
            if (r == 0) {
                start = offsets[8 * t];
                len = offsets[8 * t + 8] - start;
                rn0 = offsets[8 * t + 1] - start;  // own face class 0 comes first
            } else {
                const int64_t nx = (dir & 1) ? (tx == 0 ? tg.nt[0] - 1 : tx - 1) : tx;
                const int64_t ny = (dir & 2) ? (ty == 0 ? tg.nt[1] - 1 : ty - 1) : ty;
                const int64_t nz = (dir & 4) ? (tz == 0 ? tg.nt[2] - 1 : tz - 1) : tz;
                const int64_t k = 8 * ((nx * tg.nt[1] + ny) * tg.nt[2] + nz) + fc;
                start = offsets[k];
                len = offsets[k + 1] - start;
            }
        }
        unsigned inc = len;  // inclusive scan of the range lengths over the wave
#pragma unroll
        for (int o = 1; o < 32; o <<= 1) {
            const unsigned y = __shfl_up(inc, o);
            if (r >= o) inc += y;
        }
        if (r < NR) {
            rstart[r] = start;
            rcode[r] = (unsigned char)dir;
            rpre[r] = inc - len;
            if (r == NR - 1) rpre[NR] = inc;
        }
    }
    int g = 0;
    if constexpr (ACC32) {
        const double W = __longlong_as_double((long long)hdr[3]);
        g = ilogb(2.0 * W + 1.0) + 1;  // 2^g > 2 W + 1
    }
    int s[NC];
    double fs[NC];
    DepositT<T> Ms[NC];  // the fixed-point mass of a scalar mass
#pragma unroll
    for (int c = 0; c < NC; c++) {
        s[c] = fixed_shift<T>(hdr[c], np) - g;
        fs[c] = ldexp(scale, -s[c]);
        Ms[c] = fixed_mass<T>(mscalar, s[c]);
    }
    __syncthreads();
    const unsigned total = rpre[NR];
    // The own tile's face class 0 (its first rn0 particles, most of them) has
    // all 8 corners of every particle in the tile.
    const unsigned r0 = __builtin_amdgcn_readfirstlane(rstart[0]), nfast = __builtin_amdgcn_readfirstlane(rn0);
    // KB particles per thread per round; all their loads are issued before any
    // deposit so the dependent sorted[j] -> pos[p] chains overlap (measured at
    // 512^3: the one-component paint fastest at 6, the 32-bit three-component
    // paint at 4, 59 VGPRs; FP64 three-component 4, whose 1024-thread groups
    // would spill at 6).  A round takes TB KB particles, wave w the 64
    // at w 64 + k TB for k < KB; a wave skips its sub-batches that start past
    // the end (the waves of a tile stay within one sub-batch of each other),
    // and a wave whose sub-batches all lie in face class 0 deposits without
    // range lookups or corner tests.
    constexpr int KB = NC == 1 ? 6 : (ACC32 || sizeof(T) == 8) ? 4 : 6;
    const unsigned lane = threadIdx.x & 63u;
    auto round = [&](const unsigned wb, const unsigned nk, auto fast_tag) {
        constexpr bool FAST = decltype(fast_tag)::value;
        unsigned pi[KB];
        int code[KB];
#pragma unroll
        for (int k = 0; k < KB; k++) {
            if constexpr (FAST) {
                pi[k] = sorted[r0 + wb + k * TB + lane];
                code[k] = 0;
            } else {
                if (k >= (int)nk) break;
                unsigned j = wb + k * TB + lane;
                j = j < total ? j : total - 1;
                int r = 0;
                if (j >= rpre[1]) {  // binary search over the neighbour ranges 1..NR-1
                    int a = 1, b = NR - 1;
                    while (a < b) {
                        const int mid = (a + b + 1) >> 1;
                        if (rpre[mid] <= j) a = mid;
                        else b = mid - 1;
                    }
                    r = a;
                }
                pi[k] = sorted[rstart[r] + (j - rpre[r])];
                code[k] = rcode[r];
            }
        }
        T xs[KB][3];
        T ms[KB][NC];
#pragma unroll
        for (int k = 0; k < KB; k++) {
            if (!FAST && k >= (int)nk) break;
            const T* q = pos + 3 * (int64_t)pi[k];
            xs[k][0] = q[0];
            xs[k][1] = q[1];
            xs[k][2] = q[2];
        }
        if (mass) {  // uniform: a particle's NC masses in one load
#pragma unroll
            for (int k = 0; k < KB; k++) {
                if (!FAST && k >= (int)nk) break;
#pragma unroll
                for (int c = 0; c < NC; c++) ms[k][c] = mass[(int64_t)pi[k] * NC + c];
            }
        } else {
#pragma unroll
            for (int k = 0; k < KB; k++)
#pragma unroll
                for (int c = 0; c < NC; c++) ms[k][c] = mscalar;
        }
#pragma unroll
        for (int k = 0; k < KB; k++) {
            if (!FAST && k >= (int)nk) break;
            const int64_t p = pi[k];
            uint32_t id[3];
            T f[3];
            cell_of<T, DM>(pd, mp, xs[k][0], xs[k][1], xs[k][2], p, id, f);
            if constexpr (FAST) {
                const int base = ((int)(id[0] - (uint32_t)lo[0]) * TY + (int)(id[1] - (uint32_t)lo[1])) * TZ +
                                 (int)(id[2] - (uint32_t)lo[2]);
#pragma unroll
                for (int c = 0; c < NC; c++) {
                    DepositT<T> q[8];
                    deposits_of<T>(f, mass ? fixed_mass<T>(ms[k][c], s[c]) : Ms[c], q);
#pragma unroll
                    for (int qq = 0; qq < 8; qq++) {
                        const int dx = qq >> 2, dy = (qq >> 1) & 1, dz = qq & 1;
                        atomicAdd(&box[c * B3 + base + dx * TY * TZ + dy * TZ + dz], (Acc)(SAcc)q[qq]);
                    }
                }
            } else {
                // every lane stays active to the wave-wide test below; lanes
                // past the end (reading a valid duplicate) deposit nothing
                const bool valid = wb + k * TB + lane < total;
                // local lower corner: -1 along the axes the particle was read
                // from a lower neighbour, else its offset in this tile
                const int a = (code[k] & 1) ? -1 : (int)(id[0] - (uint32_t)lo[0]);
                const int b = (code[k] & 2) ? -1 : (int)(id[1] - (uint32_t)lo[1]);
                const int cz = (code[k] & 4) ? -1 : (int)(id[2] - (uint32_t)lo[2]);
                const bool in_x[2] = {valid && a >= 0, valid && a + 1 < ex};
                const bool in_y[2] = {b >= 0, b + 1 < ey};
                const bool in_z[2] = {cz >= 0, cz + 1 < ez};
                const int base = (a * TY + b) * TZ + cz;
                // every lane of the wave has all 8 corners in the tile:
                // unconditional deposits; else each corner is tested.  One of
                // the two per wave.
                const bool all_in = in_x[0] && in_x[1] && in_y[0] && in_y[1] && in_z[0] && in_z[1];
                const bool wave_in = __all(all_in);
#pragma unroll
                for (int c = 0; c < NC; c++) {
                    DepositT<T> q[8];
                    deposits_of<T>(f, mass ? fixed_mass<T>(ms[k][c], s[c]) : Ms[c], q);
#pragma unroll
                    for (int qq = 0; qq < 8; qq++) {
                        const int dx = qq >> 2, dy = (qq >> 1) & 1, dz = qq & 1;
                        if (wave_in || (in_x[dx] && in_y[dy] && in_z[dz]))
                            atomicAdd(&box[c * B3 + base + dx * TY * TZ + dy * TZ + dz], (Acc)(SAcc)q[qq]);
                    }
                }
            }
        }
    };
    for (unsigned jb = 0; jb < total; jb += TB * KB) {
        const unsigned wb = __builtin_amdgcn_readfirstlane(jb + (threadIdx.x - lane));  // the wave's first particle
        if (wb >= total) break;                                                       // and of every later round
        const unsigned nk = (total - wb + TB - 1) / TB;  // sub-batches with a particle (KB or more: all)
        if (wb + (KB - 1) * TB + 64 <= nfast) round(wb, (unsigned)KB, std::true_type{});
        else round(wb, nk < (unsigned)KB ? nk : (unsigned)KB, std::false_type{});
    }
    __syncthreads();
    const int64_t cells = m.cells();
    T mx = T(0);
    for (int li = threadIdx.x; li < B3; li += TB) {
        const int cc = li % TZ, bb = (li / TZ) % TY, aa = li / (TZ * TY);
        if (aa >= ex || bb >= ey || cc >= ez) continue;
        const int64_t gi = ((lo[0] + aa) * m.n1 + (lo[1] + bb)) * m.n2 + (lo[2] + cc);
#pragma unroll
        for (int c = 0; c < NC; c++) {
            const T v = (T)((double)(SAcc)box[c * B3 + li] * fs[c]);
            out[c * cells + gi] = v;
            if constexpr (NC == 1 && sizeof(T) == 4) mx = fmaxf(mx, fabsf(v));
        }
    }
    if constexpr (NC == 1 && sizeof(T) == 4) {
        if (vmax) {  // uniform: the workgroup's largest cell, one atomic
            __shared__ float wmx[TB / 64];
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
            if ((threadIdx.x & 63) == 0) wmx[threadIdx.x >> 6] = mx;
            __syncthreads();
            if (threadIdx.x == 0) {
                float b = wmx[0];
#pragma unroll
                for (int i = 1; i < TB / 64; i++) b = fmaxf(b, wmx[i]);
                atomicMax(vmax, __float_as_uint(b));  // non-negative floats order like their bits
            }
        }
    }
}

// slot (FP32 bits of max |w| from the 1-component paint that wrote w) -> the
// double bits k_mesh_absmax would give for w and scale: (double)max|w| * scale.
static __global__ void k_vmax_scaled(const unsigned* __restrict__ slot, double scale, unsigned long long* __restrict__ out) {
    const double v = (double)__uint_as_float(*slot) * scale;
    if (v > 0.0) *out = (unsigned long long)__double_as_longlong(v);
}

size_t paint_ws_bytes(const Mesh& m, int nc, int64_t np, int32_t flags);

template <typename T>
int paint_impl(const T* pos, const ldl_pos_desc* pdesc, int64_t np, const T* mass, double mscalar, int nc,
               const int64_t nmesh[3], double out_scale, int32_t flags, T* out, void* ws, size_t ws_bytes,
               hipStream_t st, unsigned* bins_out = nullptr, const unsigned* bins_in = nullptr, const T* wmesh = nullptr,
               double wscale = 0.0, const unsigned long long* absmax_in = nullptr);
// Particle sort by tile and face class that a 1-component paint can
// hand to a later 3-component paint of the same positions: offsets
// (8 ntiles + 1), sorted particle indices (np), then (FP32) the FP32 bits of
// the painted mesh's largest |cell|, uint32.
size_t paint_bins_bytes(const Mesh& m, int64_t np);
template <typename T>
int readout_impl(const T* pos, const ldl_pos_desc* pdesc, int64_t np, const T* field, const int64_t nmesh[3],
                 int kind, int vjpdim, const T* aux, T* out, hipStream_t st);

}  // namespace ldl
