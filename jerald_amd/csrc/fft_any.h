// Batched FFTs of any length whose prime factors are 2, 3 and 5, for mesh
// axes that are not powers of two.  Same conventions and pass structure as
// fft_core.h (Stockham autosort in LDS, unnormalised, forward exp(-2 pi i/N));
// the stage list (radices 8, 4, 2, 5, 3) is a runtime plan, so butterflies
// read and write LDS each stage (ping-pong buffers) instead of staying in
// registers.  Host helpers pick the power-of-two kernels when they apply.
#pragma once

#include "fft_core.h"

namespace ldl {

struct AnyPlan {
    int n, nst;
    int radix[20];
};

// Radix list for n (8s first, then 4, 2, 5, 3); false if n has another prime factor.
inline bool any_plan(int64_t n, AnyPlan* pl) {
    if (n < 2 || n > 1024) return false;
    int64_t r = n;
    pl->n = (int)n;
    pl->nst = 0;
    auto take = [&](int f) {
        while (r % f == 0 && pl->nst < 20) {
            pl->radix[pl->nst++] = f;
            r /= f;
        }
    };
    take(8);
    take(4);
    take(2);
    take(5);
    take(3);
    return r == 1;
}

// 3- and 5-point DFTs with w = exp(DIR 2 pi i / R).
template <int DIR, typename C> __device__ __forceinline__ void dft3(C* v) {
    using S = decltype(v[0].x);
    const S s3 = S(0.86602540378443864676372317075294);  // sin(2 pi / 3)
    const C a = cadd(v[1], v[2]), b = csub(v[1], v[2]);
    const C t = C{v[0].x - S(0.5) * a.x, v[0].y - S(0.5) * a.y};
    const C u = rot90<DIR>(C{s3 * b.x, s3 * b.y});
    v[0] = cadd(v[0], a);
    v[1] = cadd(t, u);
    v[2] = csub(t, u);
}

template <int DIR, typename C> __device__ __forceinline__ void dft5(C* v) {
    using S = decltype(v[0].x);
    const S c1 = S(0.30901699437494742410229341718282);   // cos(2 pi / 5)
    const S c2 = S(-0.80901699437494742410229341718282);  // cos(4 pi / 5)
    const S s1 = S(0.95105651629515357211643933337938);   // sin(2 pi / 5)
    const S s2 = S(0.58778525229247312916870595463907);   // sin(4 pi / 5)
    const C a1 = cadd(v[1], v[4]), a2 = cadd(v[2], v[3]);
    const C b1 = csub(v[1], v[4]), b2 = csub(v[2], v[3]);
    const C t1 = C{v[0].x + c1 * a1.x + c2 * a2.x, v[0].y + c1 * a1.y + c2 * a2.y};
    const C t2 = C{v[0].x + c2 * a1.x + c1 * a2.x, v[0].y + c2 * a1.y + c1 * a2.y};
    const C u1 = rot90<DIR>(C{s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y});
    const C u2 = rot90<DIR>(C{s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y});
    v[0] = cadd(v[0], cadd(a1, a2));
    v[1] = cadd(t1, u1);
    v[4] = csub(t1, u1);
    v[2] = cadd(t2, u2);
    v[3] = csub(t2, u2);
}

template <int DIR, typename C> __device__ __forceinline__ void dft_any(C* v, int R) {
    switch (R) {
        case 2: dft2<DIR>(v); break;
        case 3: dft3<DIR>(v); break;
        case 4: dft4<DIR>(v[0], v[1], v[2], v[3]); break;
        case 5: dft5<DIR>(v); break;
        default: dft8<DIR>(v); break;
    }
}

// Transform the P pencils of a (element n of pencil p at a[n * (P + 1) + p]);
// b is scratch of the same size.  tw[m * tws] = exp(-2 pi i m / n).  Returns
// the buffer holding the result.  All threads of the workgroup take part.
template <typename T, int DIR>
__device__ c2<T>* lds_fft_any(c2<T>* a, c2<T>* b, const c2<T>* tw, int tws, int P, const AnyPlan& pl) {
    using C = c2<T>;
    const int n = pl.n, LS = P + 1;
    C* src = a;
    C* dst = b;
    int Ns = 1;
    for (int st = 0; st < pl.nst; st++) {
        const int R = pl.radix[st], NB = n / R, tstep = n / (Ns * R);
        for (int w = threadIdx.x; w < NB * P; w += blockDim.x) {
            const int p = w % P, j = w / P, k = j % Ns;
            C v[8];
#pragma unroll
            for (int r = 0; r < 8; r++)
                if (r < R) v[r] = src[(j + r * NB) * LS + p];
            if (Ns > 1) {
#pragma unroll
                for (int r = 1; r < 8; r++)
                    if (r < R) {
                        C t = tw[r * k * tstep * tws];
                        if (DIR > 0) t.y = -t.y;
                        v[r] = cmul(v[r], t);
                    }
            }
            dft_any<DIR>(v, R);
            const int base = (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < 8; r++)
                if (r < R) dst[(base + r * Ns) * LS + p] = v[r];
        }
        __syncthreads();
        C* t = src;
        src = dst;
        dst = t;
        Ns *= R;
    }
    return src;
}

constexpr int kAnyThreads = 256;
constexpr size_t kAnyLds = 64 * 1024;

// Pencils per workgroup: the largest power of two <= 64 whose two buffers of
// ``len`` elements per pencil plus a table of ``tlen`` entries fit 64 KB (one
// pencil may need more: at most 2 * 1025 * 2 + 2048 FP64 complex = 97 KB).
template <typename T>
inline int any_pencils(int len, int tlen) {
    int P = 64;
    while (P > 1 && (2 * (size_t)len * (P + 1) + tlen) * sizeof(c2<T>) > kAnyLds) P >>= 1;
    return P;
}
template <typename T>
inline size_t any_lds_bytes(int len, int tlen, int P) {
    return (2 * (size_t)len * (P + 1) + tlen) * sizeof(c2<T>);
}
// Dynamic LDS above 64 KB must be enabled per kernel.
template <typename K>
inline void allow_lds(K* kernel, size_t bytes) {
    if (bytes > kAnyLds) (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

template <typename T, int DIR, class Op>
__global__ void __launch_bounds__(kAnyThreads) k_colpass_any(ColGeom g, const c2<T>* __restrict__ tw, AnyPlan pl, int P, Op op) {
    using C = c2<T>;
    extern __shared__ unsigned char any_smem[];
    const int n = pl.n, LS = P + 1;
    C* a = reinterpret_cast<C*>(any_smem);
    C* b = a + n * LS;
    C* twl = b + n * LS;
    const int64_t tile = xcd_tile(blockIdx.x, gridDim.x);
    const int64_t plane = tile / g.tiles_per_plane;
    const int64_t col0 = (tile - plane * g.tiles_per_plane) * P;
    const int64_t base = plane * g.plane_stride + col0;
    const int pv = (int)((g.ncols - col0) < P ? (g.ncols - col0) : P);
    for (int i = threadIdx.x; i < n; i += blockDim.x) twl[i] = tw[i];
    for (int i = threadIdx.x; i < n * P; i += blockDim.x) {
        const int nn = i / P, p = i - nn * P;
        const int pc = p < pv ? p : pv - 1;
        const C v = op.load(base + nn * g.es + pc, nn, plane, col0 + pc);
        a[nn * LS + p] = p < pv ? v : C{T(0), T(0)};
    }
    __syncthreads();
    const C* r = lds_fft_any<T, DIR>(a, b, twl, 1, P, pl);
    for (int i = threadIdx.x; i < n * P; i += blockDim.x) {
        const int nn = i / P, p = i - nn * P;
        if (p < pv) op.store(base + nn * g.es + p, nn, plane, col0 + p, r[nn * LS + p]);
    }
}

// Real rows of N2 = 2H -> half spectrum (pl is the plan of H; tw2 the N2 table).
template <typename T, class Pro, class Epi>
__global__ void __launch_bounds__(kAnyThreads) k_zpass_r2c_any(int64_t nrows, const c2<T>* __restrict__ tw2, AnyPlan pl, int P, Pro pro,
                                                              Epi epi) {
    using C = c2<T>;
    extern __shared__ unsigned char any_smem[];
    const int H = pl.n, N2 = 2 * H, LS = P + 1;
    C* a = reinterpret_cast<C*>(any_smem);
    C* b = a + (H + 1) * LS;
    C* twl = b + (H + 1) * LS;
    const int64_t row0 = xcd_tile(blockIdx.x, gridDim.x) * (int64_t)P;
    for (int i = threadIdx.x; i < N2; i += blockDim.x) twl[i] = tw2[i];
    for (int i = threadIdx.x; i < H * P; i += blockDim.x) {
        const int p = i / H, m = i - p * H;
        const C v = pro.load2(row0 + p < nrows ? row0 + p : nrows - 1, m);
        a[m * LS + p] = row0 + p < nrows ? v : C{T(0), T(0)};
    }
    __syncthreads();
    const C* z = lds_fft_any<T, -1>(a, b, twl, 2, P, pl);
    for (int i = threadIdx.x; i < (H + 1) * P; i += blockDim.x) {
        const int p = i / (H + 1), k = i - p * (H + 1);
        if (row0 + p >= nrows) continue;
        const C zk = z[(k == H ? 0 : k) * LS + p];
        const C zc = cconj(z[(k == 0 ? 0 : H - k) * LS + p]);
        // E = (Z[k] + conj Z[H-k]) / 2, O = (Z[k] - conj Z[H-k]) / (2i), X = E + W^k O
        const C e = C{T(0.5) * (zk.x + zc.x), T(0.5) * (zk.y + zc.y)};
        const C d = csub(zk, zc);
        const C o = C{T(0.5) * d.y, T(-0.5) * d.x};
        epi.store(row0 + p, k, cadd(e, cmul(o, twl[k])));
    }
}

// Half spectrum rows -> real rows of N2 = 2H, with the Epi::NR reductions of
// k_zpass_c2r (workgroup totals to Epi::reduce_out(blockIdx.x, ...)).
template <typename T, class Pro, class Epi>
__global__ void __launch_bounds__(kAnyThreads) k_zpass_c2r_any(int64_t nrows, const c2<T>* __restrict__ tw2, AnyPlan pl, int P, Pro pro,
                                                              Epi epi) {
    using C = c2<T>;
    extern __shared__ unsigned char any_smem[];
    const int H = pl.n, N2 = 2 * H, LS = P + 1;
    C* a = reinterpret_cast<C*>(any_smem);
    C* b = a + (H + 1) * LS;
    C* twl = b + (H + 1) * LS;
    const int64_t row0 = xcd_tile(blockIdx.x, gridDim.x) * (int64_t)P;
    for (int i = threadIdx.x; i < N2; i += blockDim.x) twl[i] = tw2[i];
    for (int i = threadIdx.x; i < (H + 1) * P; i += blockDim.x) {
        const int p = i / (H + 1), k = i - p * (H + 1);
        C w = pro.load(row0 + p < nrows ? row0 + p : nrows - 1, k);
        if (row0 + p >= nrows) w = C{T(0), T(0)};
        if (k == 0 || k == H) w.y = T(0);  // FFTW ignores these imaginary parts
        a[k * LS + p] = w;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < H * P; i += blockDim.x) {
        const int p = i / H, k = i - p * H;
        const C xk = a[k * LS + p];
        const C xc = a[(H - k) * LS + p];
        // Z'[k] = (X[k] + conj X[H-k]) + i (X[k] - conj X[H-k]) e^{+2 pi i k / N2}
        C w = twl[k];
        w.y = -w.y;
        const C e = cadd(xk, cconj(xc)), o = cmul(csub(xk, cconj(xc)), w);
        b[k * LS + p] = cadd(e, C{-o.y, o.x});
    }
    __syncthreads();
    const C* x = lds_fft_any<T, +1>(b, a, twl, 2, P, pl);
    constexpr int NR = Epi::NR;
    double acc[NR > 0 ? NR : 1] = {};
    for (int i = threadIdx.x; i < H * P; i += blockDim.x) {
        const int p = i / H, m = i - p * H;
        if (row0 + p < nrows) epi.store2(row0 + p, m, x[m * LS + p], acc);
    }
    if constexpr (NR > 0) {
        block_sum<NR, kAnyThreads>(acc);
        if (threadIdx.x == 0) epi.reduce_out(blockIdx.x, acc);
    }
}

// ------------------------------------------------------ host pass helpers
// Each runs the power-of-two kernel when the length allows it, else the
// any-length kernel.  ``nblocks`` (c2r) receives the workgroup count, which
// indexes the Epi reduction partials.

inline bool pow2_len(int64_t n, int64_t maxn) { return is_pow2(n) && n >= 4 && n <= maxn; }

template <typename T, int DIR, class Op>
inline int colpass(int64_t n, int64_t planes, int64_t ncols, int64_t plane_stride, int64_t es, const c2<T>* tw, const Op& op,
                   hipStream_t st, const char* tname = "colpass") {
    if (pow2_len(n, 1024)) {
        dispatch_n(n, [&](auto N) {
            const ColGeom g = col_geom<T, decltype(N)::value>(planes, ncols, plane_stride, es);
            launch_colpass<T, decltype(N)::value, DIR>(g, tw, op, st, tname);
        });
        return LDL_OK;
    }
    AnyPlan pl;
    if (!any_plan(n, &pl)) return fail(LDL_EUNSUPPORTED, "FFT length must have no prime factor other than 2, 3 and 5");
    const c2<T>* twn;
    int rc = twiddles_n<T>((int)n, &twn);
    if (rc) return rc;
    const int P = any_pencils<T>((int)n, (int)n);
    ColGeom g{planes, ncols, plane_stride, es, (ncols + P - 1) / P, 0};
    g.ntiles = g.tiles_per_plane * planes;
    const size_t lb = any_lds_bytes<T>((int)n, (int)n, P);
    allow_lds(&k_colpass_any<T, DIR, Op>, lb);
    KTimer _t(tname, st);
    hipLaunchKernelGGL((k_colpass_any<T, DIR, Op>), dim3((unsigned)g.ntiles), dim3(kAnyThreads), lb, st, g, twn, pl, P, op);
    return LDL_OK;
}

template <typename T, class Pro, class Epi>
inline int zpass_r2c(int64_t nrows, int64_t n2, const c2<T>* tw, const Pro& pro, const Epi& epi, hipStream_t st,
                     const char* tname = "zpass_r2c") {
    if (pow2_len(n2 / 2, 512) && n2 % 2 == 0) {
        dispatch_h(n2 / 2, [&](auto H) { launch_zr2c<T, decltype(H)::value>(nrows, tw, pro, epi, st, tname); });
        return LDL_OK;
    }
    AnyPlan pl;
    if (n2 % 2 || !any_plan(n2 / 2, &pl)) return fail(LDL_EUNSUPPORTED, "FFT last axis must be even with no prime factor other than 2, 3 and 5");
    const c2<T>* tw2;
    int rc = twiddles_n<T>((int)n2, &tw2);
    if (rc) return rc;
    const int H = (int)(n2 / 2), P = any_pencils<T>(H + 1, (int)n2);
    const size_t lb = any_lds_bytes<T>(H + 1, (int)n2, P);
    allow_lds(&k_zpass_r2c_any<T, Pro, Epi>, lb);
    KTimer _t(tname, st);
    hipLaunchKernelGGL((k_zpass_r2c_any<T, Pro, Epi>), dim3((unsigned)((nrows + P - 1) / P)), dim3(kAnyThreads), lb, st, nrows, tw2, pl, P,
                       pro, epi);
    return LDL_OK;
}

// Workgroup count of zpass_c2r over nrows rows of n2 (its reduction partials).
template <typename T>
inline int64_t zpass_c2r_blocks(int64_t nrows, int64_t n2) {
    int64_t nb = 0;
    if (pow2_len(n2 / 2, 512) && n2 % 2 == 0) {
        dispatch_h(n2 / 2, [&](auto H) { nb = (nrows + FftCfg<T, decltype(H)::value>::P - 1) / FftCfg<T, decltype(H)::value>::P; });
    } else {
        const int P = any_pencils<T>((int)(n2 / 2) + 1, (int)n2);
        nb = (nrows + P - 1) / P;
    }
    return nb;
}

template <typename T, class Pro, class Epi>
inline int zpass_c2r(int64_t nrows, int64_t n2, const c2<T>* tw, const Pro& pro, const Epi& epi, hipStream_t st, int64_t* nblocks,
                     const char* tname = "zpass_c2r") {
    if (pow2_len(n2 / 2, 512) && n2 % 2 == 0) {
        dispatch_h(n2 / 2, [&](auto H) {
            constexpr int HH = decltype(H)::value;
            constexpr int RPB = zc2r_rows_per_block<T, HH, Epi>();
            if (nblocks) *nblocks = (nrows + RPB - 1) / RPB;
            launch_zc2r<T, HH>(nrows, tw, pro, epi, st, tname);
        });
        return LDL_OK;
    }
    AnyPlan pl;
    if (n2 % 2 || !any_plan(n2 / 2, &pl)) return fail(LDL_EUNSUPPORTED, "FFT last axis must be even with no prime factor other than 2, 3 and 5");
    const c2<T>* tw2;
    int rc = twiddles_n<T>((int)n2, &tw2);
    if (rc) return rc;
    const int H = (int)(n2 / 2), P = any_pencils<T>(H + 1, (int)n2);
    const int64_t nb = (nrows + P - 1) / P;
    if (nblocks) *nblocks = nb;
    const size_t lb = any_lds_bytes<T>(H + 1, (int)n2, P);
    allow_lds(&k_zpass_c2r_any<T, Pro, Epi>, lb);
    KTimer _t(tname, st);
    hipLaunchKernelGGL((k_zpass_c2r_any<T, Pro, Epi>), dim3((unsigned)nb), dim3(kAnyThreads), lb, st, nrows, tw2, pl, P, pro, epi);
    return LDL_OK;
}

}  // namespace ldl
