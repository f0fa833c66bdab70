// Batched 1D complex FFTs in LDS (Stockham autosort, radix 8/4/2) and the three
// pass kernels of a 3D r2c / c2r transform on gfx950.
//
// Layout of a tile in LDS: element n of pencil p at lds[n * (P + 1) + p] (the
// +1 pad keeps 64-bit accesses conflict free).  A workgroup owns P pencils;
// TPP threads share one pencil, each holding N / TPP values per stage in
// registers.  Forward uses exp(-2 pi i / N), inverse exp(+2 pi i / N); both are
// unnormalised.  Twiddles come from tables built in FP64 (runtime.hip).
//
// The 3D transforms follow FFTW's r2c / c2r conventions (lib/parfft.h:45-46):
// row-major (N0, N1, N2) real <-> (N0, N1, N2/2+1) complex:
//   z pass  : real rows of N2 -> N2/2 complex FFT of (x[2m] + i x[2m+1]) ->
//             split into the N2/2+1 half-spectrum bins;
//   y pass  : complex FFT of length N1 along columns of stride N2/2+1;
//   x pass  : complex FFT of length N0 along columns of stride N1 (N2/2+1).
// Load/store functors let the LDL kernels fuse their pointwise work into
// the passes (ldl_sm.hip).
#pragma once

#include <type_traits>

#include "ldl_common.h"

namespace ldl {

// ------------------------------------------------------------- config
constexpr int radix_of(int rem) { return rem == 2 ? 2 : (rem == 4 || rem == 16) ? 4 : 8; }

template <typename T, int N> struct FftCfg {
    // threads per pencil: each thread holds N / TPP values per stage (16 at
    // N = 512), which keeps the fused kernels below 128 VGPRs
    static constexpr int TPP = N <= 8 ? 1 : N <= 32 ? 4 : N == 64 ? 8 : N == 128 ? 16 : N <= 512 ? 32 : 64;
    // FP64 at N = 1024 keeps 4 pencils so data + twiddles fit the 160 KB of LDS
    static constexpr int P = sizeof(T) == 4 ? (N <= 8 ? 64 : N <= 64 ? 32 : 16)
                                            : (N <= 8 ? 32 : N <= 64 ? 16 : N <= 512 ? 8 : 4);
    static constexpr int THREADS = P * TPP;
    static constexpr int LS = P + 1;
    static constexpr int LDS_ELEMS = N * LS;
};

// ------------------------------------------------------------ butterflies
// a * (DIR * i)
template <int DIR, typename C> __device__ __forceinline__ C rot90(C a) {
    return DIR < 0 ? C{a.y, -a.x} : C{-a.y, a.x};
}

template <int DIR, typename C> __device__ __forceinline__ void dft2(C* v) {
    C a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <int DIR, typename C> __device__ __forceinline__ void dft4(C& x0, C& x1, C& x2, C& x3) {
    C t0 = cadd(x0, x2), t1 = csub(x0, x2), t2 = cadd(x1, x3), t3 = rot90<DIR>(csub(x1, x3));
    x0 = cadd(t0, t2);
    x2 = csub(t0, t2);
    x1 = cadd(t1, t3);
    x3 = csub(t1, t3);
}

template <int DIR, typename C> __device__ __forceinline__ void dft8(C* v) {
    using S = decltype(v[0].x);
    const S s = S(0.70710678118654752440084436210485);
    C e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    C o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4<DIR>(e0, e1, e2, e3);
    dft4<DIR>(o0, o1, o2, o3);
    // o_m *= w8^m, w8 = exp(DIR 2 pi i / 8)
    o1 = C{s * (o1.x - S(DIR) * o1.y), s * (S(DIR) * o1.x + o1.y)};
    o2 = rot90<DIR>(o2);
    o3 = C{s * (-o3.x - S(DIR) * o3.y), s * (S(DIR) * o3.x - o3.y)};
    v[0] = cadd(e0, o0); v[4] = csub(e0, o0);
    v[1] = cadd(e1, o1); v[5] = csub(e1, o1);
    v[2] = cadd(e2, o2); v[6] = csub(e2, o2);
    v[3] = cadd(e3, o3); v[7] = csub(e3, o3);
}

template <int R, int DIR, typename C> __device__ __forceinline__ void dft(C* v) {
    if constexpr (R == 2) dft2<DIR>(v);
    else if constexpr (R == 4) dft4<DIR>(v[0], v[1], v[2], v[3]);
    else dft8<DIR>(v);
}

// ------------------------------------------------------------ LDS FFT
// In-place transform of the P pencils held in lds.  twN[m * TWS] =
// exp(-2 pi i m / N): the workgroup's LDS copy of the twiddle table.
// Stages Ns .. NE (exclusive) of the transform: all of them by default.
template <typename T, int N, int P, int TPP, int DIR, int TWS = 1, int Ns = 1, int NE = N>
__device__ __forceinline__ void lds_fft(c2<T>* lds, const c2<T>* twN) {
    if constexpr (Ns < NE) {
        using C = c2<T>;
        constexpr int R = radix_of(N / Ns);
        constexpr int NB = N / R;
        static_assert(NB % TPP == 0, "threads per pencil must divide the butterfly count");
        constexpr int BPT = NB / TPP;
        constexpr int LS = P + 1;
        const int p = threadIdx.x % P, q = threadIdx.x / P;
        C v[BPT][R];
#pragma unroll
        for (int b = 0; b < BPT; b++) {
            const int j = q + b * TPP;
#pragma unroll
            for (int r = 0; r < R; r++) v[b][r] = lds[(j + r * NB) * LS + p];
        }
        __syncthreads();
#pragma unroll
        for (int b = 0; b < BPT; b++) {
            const int j = q + b * TPP;
            const int k = j & (Ns - 1);
            if constexpr (Ns > 1) {
#pragma unroll
                for (int r = 1; r < R; r++) {
                    C w = twN[r * k * (N / (Ns * R)) * TWS];
                    if constexpr (DIR > 0) w.y = -w.y;
                    v[b][r] = cmul(v[b][r], w);
                }
            }
            dft<R, DIR>(v[b]);
            const int base = (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) lds[(base + r * Ns) * LS + p] = v[b][r];
        }
        __syncthreads();
        lds_fft<T, N, P, TPP, DIR, TWS, Ns * R, NE>(lds, twN);
    }
}

// Ns of the last stage of an N-point transform.
constexpr int fft_last_ns(int n) {
    int ns = 1;
    while (ns * radix_of(n / ns) < n) ns *= radix_of(n / ns);
    return ns;
}

// The same transform with its first and last stages in registers.  Thread
// (q, p) = (threadIdx / P, threadIdx % P) holds x[s] = element q + TPP s of
// pencil p: the elements reg_load gives it and reg_store takes from it.  The
// first-stage butterflies j = q + TPP b read elements j + (N / R) r, which are
// x[b + (N / R / TPP) r], and the last-stage ones write elements j + (N / R) r
// in natural order, which are x[b + (N / R / TPP) r] again; so a load, the
// transform and a store make two LDS round trips fewer than with lds_fft, and
// an elementwise step between two transforms none.
//
// fft_from_regs: the first stage on x, written to lds, then the stages below
// NE in lds (NE = N: the whole transform, in lds in natural order).  Other
// threads may still read lds on entry (the first write waits at a barrier).
template <typename T, int N, int P, int TPP, int DIR, int TWS = 1, int NE = N>
__device__ __forceinline__ void fft_from_regs(c2<T> (&x)[N / TPP], c2<T>* lds, const c2<T>* twN) {
    using C = c2<T>;
    constexpr int R1 = radix_of(N), LS = P + 1, B1 = N / R1 / TPP;
    static_assert(B1 * R1 * TPP == N, "threads per pencil must divide the butterfly count");
    const int p = threadIdx.x % P, q = threadIdx.x / P;
    __syncthreads();  // earlier readers of lds are done
#pragma unroll
    for (int b = 0; b < B1; b++) {
        C v[R1];
#pragma unroll
        for (int r = 0; r < R1; r++) v[r] = x[b + B1 * r];
        dft<R1, DIR>(v);
        const int j = q + b * TPP;
#pragma unroll
        for (int r = 0; r < R1; r++) lds[(j * R1 + r) * LS + p] = v[r];
    }
    __syncthreads();
    lds_fft<T, N, P, TPP, DIR, TWS, R1, NE>(lds, twN);
}

// fft_to_regs: the last stage from lds into x, every earlier stage done (and
// followed by a barrier).  Other threads may still read lds on return.
template <typename T, int N, int P, int TPP, int DIR, int TWS = 1>
__device__ __forceinline__ void fft_to_regs(c2<T> (&x)[N / TPP], const c2<T>* lds, const c2<T>* twN) {
    using C = c2<T>;
    constexpr int NL = fft_last_ns(N), RL = N / NL, BL = NL / TPP, LS = P + 1;
    static_assert(BL * RL * TPP == N, "threads per pencil must divide the butterfly count");
    const int p = threadIdx.x % P, q = threadIdx.x / P;
#pragma unroll
    for (int b = 0; b < BL; b++) {
        const int j = q + b * TPP;  // < NL: its twiddle index is j itself
        C v[RL];
#pragma unroll
        for (int r = 0; r < RL; r++) v[r] = lds[(j + r * NL) * LS + p];
        if constexpr (NL > 1) {
#pragma unroll
            for (int r = 1; r < RL; r++) {
                C w = twN[r * j * TWS];
                if constexpr (DIR > 0) w.y = -w.y;
                v[r] = cmul(v[r], w);
            }
        }
        dft<RL, DIR>(v);
#pragma unroll
        for (int r = 0; r < RL; r++) x[b + BL * r] = v[r];
    }
}

// The whole transform of x, in place.
template <typename T, int N, int P, int TPP, int DIR, int TWS = 1>
__device__ __forceinline__ void reg_fft(c2<T> (&x)[N / TPP], c2<T>* lds, const c2<T>* twN) {
    if constexpr (radix_of(N) == N) {  // one stage: the thread holds the whole pencil
        static_assert(TPP == 1, "a one-stage transform is one thread per pencil");
        dft<N, DIR>(x);
    } else {
        fft_from_regs<T, N, P, TPP, DIR, TWS, fft_last_ns(N)>(x, lds, twN);
        fft_to_regs<T, N, P, TPP, DIR, TWS>(x, lds, twN);
    }
}

// Copy the length-L twiddle table (tw + L) into LDS.
template <typename T, int L, int TH>
__device__ __forceinline__ void stage_twiddles(c2<T>* twl, const c2<T>* __restrict__ tw) {
    for (int i = threadIdx.x; i < L; i += TH) twl[i] = tw[L + i];
}

// Tile of N x P complex values: every thread first issues all of its global
// loads (N / TPP of them), then writes LDS, so each wave keeps many
// requests in flight.  f(n, p) loads element n of pencil p.  Pencils past
// pv (a partial tile) load a valid pencil and are zeroed afterwards, so the
// loads stay branch free (a branch per load would serialise them).
template <typename T, int N, int P, int TH, class F>
__device__ __forceinline__ void tile_load(c2<T>* lds, int pv, F&& f) {
    constexpr int NL = N * P / TH;
    c2<T> v[NL];
#pragma unroll
    for (int s = 0; s < NL; s++) {
        const int i = threadIdx.x + s * TH;
        const int n = i / P, p = i - n * P;
        v[s] = f(n, p < pv ? p : pv - 1);
    }
#pragma unroll
    for (int s = 0; s < NL; s++) {
        const int i = threadIdx.x + s * TH;
        const int n = i / P, p = i - n * P;
        lds[n * (P + 1) + p] = p < pv ? v[s] : c2<T>{T(0), T(0)};
    }
}

// tile_load into registers: x[s] = element q + TPP s of pencil p (reg_fft).
template <typename T, int N, int P, int TH, class F>
__device__ __forceinline__ void reg_load(c2<T> (&x)[N * P / TH], int pv, F&& f) {
    constexpr int NL = N * P / TH;
#pragma unroll
    for (int s = 0; s < NL; s++) {
        const int i = threadIdx.x + s * TH;
        const int n = i / P, p = i - n * P;
        x[s] = f(n, p < pv ? p : pv - 1);
    }
    if ((int)threadIdx.x % P >= pv) {
#pragma unroll
        for (int s = 0; s < NL; s++) x[s] = c2<T>{T(0), T(0)};
    }
}

// tile_store from registers.
template <typename T, int N, int P, int TH, class F>
__device__ __forceinline__ void reg_store(const c2<T> (&x)[N * P / TH], int pv, F&& f) {
    constexpr int NL = N * P / TH;
#pragma unroll
    for (int s = 0; s < NL; s++) {
        const int i = threadIdx.x + s * TH;
        const int n = i / P, p = i - n * P;
        if (p < pv) f(n, p, x[s]);
    }
}

// f(n, p, value) stores element n of pencil p (p < pv).
template <typename T, int N, int P, int TH, class F>
__device__ __forceinline__ void tile_store(const c2<T>* lds, int pv, F&& f) {
    constexpr int NL = N * P / TH;
#pragma unroll
    for (int s = 0; s < NL; s++) {
        const int i = threadIdx.x + s * TH;
        const int n = i / P, p = i - n * P;
        if (p < pv) f(n, p, lds[n * (P + 1) + p]);
    }
}

// Sum NR per-thread FP64 values over the workgroup; thread 0 gets the totals.
// block_sum into caller-provided shared memory: NR * ceil(TH / 64) doubles
// that no thread reads any more (say the caller's tile after a barrier).
template <int NR, int TH>
__device__ __forceinline__ void block_sum_in(double (&v)[NR], double* red) {
    constexpr int NW = (TH + 63) / 64;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int r = 0; r < NR; r++) {
        double x = wave_sum(v[r]);
        if (lane == 0) red[r * NW + w] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int r = 0; r < NR; r++) {
            double x = 0.0;
            for (int i = 0; i < NW; i++) x += red[r * NW + i];
            v[r] = x;
        }
    }
}

template <int NR, int TH>
__device__ __forceinline__ void block_sum(double (&v)[NR]) {
    __shared__ double red[NR * ((TH + 63) / 64)];
    block_sum_in<NR, TH>(v, red);
}

// ------------------------------------------------------- column passes
// A column pass transforms, for each of ``planes`` planes, the ``ncols``
// contiguous columns of length N with element stride ``es``.
struct ColGeom {
    int64_t planes, ncols, plane_stride, es, tiles_per_plane, ntiles;
};

template <typename T, int N, int P = FftCfg<T, N>::P>
inline ColGeom col_geom(int64_t planes, int64_t ncols, int64_t plane_stride, int64_t es) {
    ColGeom g{planes, ncols, plane_stride, es, (ncols + P - 1) / P, 0};
    g.ntiles = g.tiles_per_plane * planes;
    return g;
}

// Plain load/store functor: out = in (times scale on store).
template <typename T> struct PlainColOp {
    const c2<T>* in;
    c2<T>* out;
    T scale;
    __device__ __forceinline__ c2<T> load(int64_t idx, int n, int64_t plane, int64_t col) const { return in[idx]; }
    __device__ __forceinline__ void store(int64_t idx, int n, int64_t plane, int64_t col, c2<T> v) const {
        out[idx] = c2<T>{v.x * scale, v.y * scale};
    }
};

template <typename T, int N, int DIR, class Op>
__global__ void __launch_bounds__((FftCfg<T, N>::THREADS)) k_colpass(ColGeom g, const c2<T>* __restrict__ tw, Op op) {
    using Cfg = FftCfg<T, N>;
    constexpr int P = Cfg::P, TH = Cfg::THREADS;
    __shared__ c2<T> lds[Cfg::LDS_ELEMS];
    __shared__ c2<T> twl[N];
    // walks each XCD's range backward: between z passes (which walk forward,
    // like the rows they write) each column pass and the z pass after it start
    // on the lines written last, which the 256 MB memory-side cache still
    // holds (y passes 0.2205 -> 0.212 ms, the z c2r after them 0.241 -> 0.229)
    const int64_t tile = xcd_tile_rev(blockIdx.x, gridDim.x);
    const int64_t plane = tile / g.tiles_per_plane;
    const int64_t col0 = (tile - plane * g.tiles_per_plane) * P;
    const int64_t base = plane * g.plane_stride + col0;
    const int pv = (int)((g.ncols - col0) < P ? (g.ncols - col0) : P);
    stage_twiddles<T, N, TH>(twl, tw);
    c2<T> x[N * P / TH];
    reg_load<T, N, P, TH>(x, pv, [&](int n, int p) { return op.load(base + n * g.es + p, n, plane, col0 + p); });
    reg_fft<T, N, P, Cfg::TPP, DIR>(x, lds, twl);
    reg_store<T, N, P, TH>(x, pv, [&](int n, int p, c2<T> v) { op.store(base + n * g.es + p, n, plane, col0 + p, v); });
}

// ------------------------------------------------------------ z passes
// Real rows -> half spectrum.  Pro::load2(row, m) returns (x[2m], x[2m+1]);
// Epi::store(row, k, X) stores half-spectrum bin k in [0, H].
template <typename T, int H, class Pro, class Epi>
__global__ void __launch_bounds__((FftCfg<T, H>::THREADS)) k_zpass_r2c(int64_t nrows, const c2<T>* __restrict__ tw, Pro pro, Epi epi) {
    using Cfg = FftCfg<T, H>;
    using C = c2<T>;
    constexpr int P = Cfg::P, LS = Cfg::LS, TH = Cfg::THREADS;
    constexpr int N2 = 2 * H;
    __shared__ C lds[Cfg::LDS_ELEMS];
    __shared__ C twl[N2];
    const int64_t row0 = xcd_tile(blockIdx.x, gridDim.x) * P;
    stage_twiddles<T, N2, TH>(twl, tw);
    {
        constexpr int NL = H * P / TH;
        C v[NL];
#pragma unroll
        for (int s = 0; s < NL; s++) {
            const int i = threadIdx.x + s * TH;
            const int p = i / H, m = i - p * H;
            v[s] = pro.load2(row0 + p < nrows ? row0 + p : nrows - 1, m);  // branch free, see tile_load
        }
#pragma unroll
        for (int s = 0; s < NL; s++) {
            const int i = threadIdx.x + s * TH;
            const int p = i / H, m = i - p * H;
            lds[m * LS + p] = row0 + p < nrows ? v[s] : C{T(0), T(0)};
        }
    }
    __syncthreads();
    lds_fft<T, H, P, Cfg::TPP, -1, 2>(lds, twl);
    const C* twN2 = twl;
    for (int i = threadIdx.x; i < (H + 1) * P; i += TH) {
        const int p = i / (H + 1), k = i - p * (H + 1);
        if (row0 + p >= nrows) continue;
        const C zk = lds[(k == H ? 0 : k) * LS + p];
        const C zc = cconj(lds[(k == 0 ? 0 : H - k) * LS + p]);
        // E = (Z[k] + conj Z[H-k]) / 2, O = (Z[k] - conj Z[H-k]) / (2i), X = E + W^k O
        const C e = C{T(0.5) * (zk.x + zc.x), T(0.5) * (zk.y + zc.y)};
        const C d = csub(zk, zc);
        const C o = C{T(0.5) * d.y, T(-0.5) * d.x};
        epi.store(row0 + p, k, cadd(e, cmul(o, twN2[k])));
    }
}

// Epilogues that store rows in threads of three (Epi::TRIPLE != 0: rows 3r,
// 3r+1, 3r+2 are the three components of output row r) get whole threes per
// workgroup and store through Epi::store3(r, m, row 3r, row 3r+1, row 3r+2).
template <class E, class = void> struct epi_triple : std::false_type {};
template <class E> struct epi_triple<E, std::void_t<decltype(E::TRIPLE)>> : std::integral_constant<bool, E::TRIPLE != 0> {};
template <typename T, int H, class Epi> constexpr int zc2r_rows_per_block() {
    return epi_triple<Epi>::value ? FftCfg<T, H>::P / 3 * 3 : FftCfg<T, H>::P;
}

// Half spectrum rows -> real rows.  Pro::load(row, k) returns bin k in
// [0, H]; Epi::store2(row, m, (x[2m], x[2m+1]), acc) may add to Epi::NR
// per-thread FP64 accumulators, whose workgroup totals go to
// Epi::reduce_out(blockIdx.x, totals).
template <typename T, int H, class Pro, class Epi>
__global__ void __launch_bounds__((FftCfg<T, H>::THREADS)) k_zpass_c2r(int64_t nrows, const c2<T>* __restrict__ tw, Pro pro, Epi epi) {
    using Cfg = FftCfg<T, H>;
    using C = c2<T>;
    constexpr int P = Cfg::P, LS = Cfg::LS, TH = Cfg::THREADS;
    constexpr int N2 = 2 * H;
    constexpr int RPB = zc2r_rows_per_block<T, H, Epi>();
    static_assert(RPB > 0, "a workgroup must hold at least one row (three for triple epilogues)");
    __shared__ C lds[(H + 1) * LS];
    __shared__ C twl[N2];
    const int64_t row0 = xcd_tile(blockIdx.x, gridDim.x) * RPB;
    stage_twiddles<T, N2, TH>(twl, tw);
    {
        constexpr int NL = ((H + 1) * P + TH - 1) / TH;
        C v[NL];
#pragma unroll
        for (int s = 0; s < NL; s++) {
            const int i = threadIdx.x + s * TH;
            const int ii = i < (H + 1) * P ? i : (H + 1) * P - 1;
            const int p = ii / (H + 1), k = ii - p * (H + 1);
            v[s] = pro.load(row0 + p < nrows ? row0 + p : nrows - 1, k);  // branch free, see tile_load
        }
#pragma unroll
        for (int s = 0; s < NL; s++) {
            const int i = threadIdx.x + s * TH;
            const int p = i / (H + 1), k = i - p * (H + 1);
            C w = row0 + p < nrows ? v[s] : C{T(0), T(0)};
            if (k == 0 || k == H) w.y = T(0);  // FFTW ignores these imaginary parts
            if (i < (H + 1) * P) lds[k * LS + p] = w;
        }
    }
    __syncthreads();
    const C* twN2 = twl;
    constexpr int HP = H / 2 + 1;
    C za[(HP * P + TH - 1) / TH], zb[(HP * P + TH - 1) / TH];
    int ia[(HP * P + TH - 1) / TH], ib[(HP * P + TH - 1) / TH];
#pragma unroll
    for (int s = 0; s < (HP * P + TH - 1) / TH; s++) {
        const int i = threadIdx.x + s * TH;
        ia[s] = -1;
        ib[s] = -1;
        if (i < HP * P) {
            const int p = i / HP, k = i - p * HP;
            const C xk = lds[k * LS + p];
            const C xc = lds[(H - k) * LS + p];
            // Z'[k] = (X[k] + conj X[H-k]) + i (X[k] - conj X[H-k]) e^{+2 pi i k / N2}
            C w = twN2[k];
            w.y = -w.y;
            C e = cadd(xk, cconj(xc)), o = cmul(csub(xk, cconj(xc)), w);
            za[s] = cadd(e, C{-o.y, o.x});
            ia[s] = k * LS + p;
            if (k != 0 && 2 * k != H) {
                C w2 = twN2[H - k];
                w2.y = -w2.y;
                C e2 = cadd(xc, cconj(xk)), o2 = cmul(csub(xc, cconj(xk)), w2);
                zb[s] = cadd(e2, C{-o2.y, o2.x});
                ib[s] = (H - k) * LS + p;
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < (HP * P + TH - 1) / TH; s++) {
        if (ia[s] >= 0) lds[ia[s]] = za[s];
        if (ib[s] >= 0) lds[ib[s]] = zb[s];
    }
    __syncthreads();
    lds_fft<T, H, P, Cfg::TPP, +1, 2>(lds, twl);
    constexpr int NR = Epi::NR;
    double acc[NR > 0 ? NR : 1] = {};
    if constexpr (epi_triple<Epi>::value) {
        for (int i = threadIdx.x; i < H * (RPB / 3); i += TH) {
            const int q = i / H, m = i - q * H;
            if (row0 + 3 * q < nrows)
                epi.store3((row0 + 3 * q) / 3, m, lds[m * LS + 3 * q], lds[m * LS + 3 * q + 1], lds[m * LS + 3 * q + 2]);
        }
    } else {
        for (int i = threadIdx.x; i < H * RPB; i += TH) {
            const int p = i / H, m = i - p * H;
            if (row0 + p < nrows) epi.store2(row0 + p, m, lds[m * LS + p], acc);
        }
    }
    if constexpr (NR > 0) {
        block_sum<NR, TH>(acc);
        if (threadIdx.x == 0) epi.reduce_out(blockIdx.x, acc);
    }
}

// Plain z-pass functors.
template <typename T> struct RealRowLoad {  // rows of N2 reals
    const T* in;
    int64_t n2;
    __device__ __forceinline__ c2<T> load2(int64_t row, int m) const {
        return *reinterpret_cast<const c2<T>*>(in + row * n2 + 2 * m);
    }
};
template <typename T> struct HalfRowStore {
    c2<T>* out;
    int64_t nh;
    T scale;
    __device__ __forceinline__ void store(int64_t row, int k, c2<T> v) const {
        out[row * nh + k] = c2<T>{v.x * scale, v.y * scale};
    }
};
template <typename T> struct HalfRowLoad {
    const c2<T>* in;
    int64_t nh;
    __device__ __forceinline__ c2<T> load(int64_t row, int k) const { return in[row * nh + k]; }
};
template <typename T> struct RealRowStore {
    static constexpr int NR = 0;
    T* out;
    int64_t n2;
    T scale;
    __device__ __forceinline__ void reduce_out(int64_t, const double*) const {}
    __device__ __forceinline__ void store2(int64_t row, int m, c2<T> v, double*) const {
        *reinterpret_cast<c2<T>*>(out + row * n2 + 2 * m) = c2<T>{v.x * scale, v.y * scale};
    }
};

// ------------------------------------------------------------ dispatch
template <int N> using IC = std::integral_constant<int, N>;

// Calls f(IC<N>{}) for a runtime power of two N in [4, MAXN]; false otherwise.
template <int MAXN = 1024, class F> inline bool dispatch_n(int64_t n, F&& f) {
    switch (n) {
        case 4: f(IC<4>{}); return true;
        case 8: f(IC<8>{}); return true;
        case 16: f(IC<16>{}); return true;
        case 32: f(IC<32>{}); return true;
        case 64: f(IC<64>{}); return true;
        case 128: f(IC<128>{}); return true;
        case 256: f(IC<256>{}); return true;
        case 512: f(IC<512>{}); return true;
        case 1024:
            if constexpr (MAXN >= 1024) { f(IC<1024>{}); return true; }
            return false;
        default: return false;
    }
}

// z passes run half-length FFTs: H = N2 / 2 <= 512.
template <class F> inline bool dispatch_h(int64_t h, F&& f) { return dispatch_n<512>(h, f); }

template <typename T, int N, int DIR, class Op>
inline void launch_colpass(const ColGeom& g, const c2<T>* tw, const Op& op, hipStream_t st, const char* tname = "colpass") {
    KTimer _t(tname, st);
    hipLaunchKernelGGL((k_colpass<T, N, DIR, Op>), dim3((unsigned)g.ntiles), dim3(FftCfg<T, N>::THREADS), 0, st, g, tw, op);
}

template <typename T, int H, class Pro, class Epi>
inline void launch_zr2c(int64_t nrows, const c2<T>* tw, const Pro& pro, const Epi& epi, hipStream_t st,
                        const char* tname = "zpass_r2c") {
    constexpr int P = FftCfg<T, H>::P;
    KTimer _t(tname, st);
    hipLaunchKernelGGL((k_zpass_r2c<T, H, Pro, Epi>), dim3((unsigned)((nrows + P - 1) / P)), dim3(FftCfg<T, H>::THREADS), 0, st,
                       nrows, tw, pro, epi);
}

template <typename T, int H, class Pro, class Epi>
inline void launch_zc2r(int64_t nrows, const c2<T>* tw, const Pro& pro, const Epi& epi, hipStream_t st,
                        const char* tname = "zpass_c2r") {
    constexpr int RPB = zc2r_rows_per_block<T, H, Epi>();
    KTimer _t(tname, st);
    hipLaunchKernelGGL((k_zpass_c2r<T, H, Pro, Epi>), dim3((unsigned)((nrows + RPB - 1) / RPB)), dim3(FftCfg<T, H>::THREADS), 0, st,
                       nrows, tw, pro, epi);
}

}  // namespace ldl
