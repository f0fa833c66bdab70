// Fused LDL displacement layer (sm kind), response and loss on gfx950, and
// their hand-written adjoints.
//
// Forward of one layer (JERALDModel.potential_gradient_sm, model.py:185-212):
//   rho = fact * paint(X, 1)                       fixed-point paint (cic.h)
//   U   = r2c((rho + 1e-8)^gamma) / N^3            z pass (power fused) + y pass
//   P   = 1e-2 alpha G(k) U,  G = -exp(-k^2/kl^2 - kh^2/k^2) k^nu
//   F_d = c2r(-i D_d(k) P)                          x pass: forward, filter, three
//                                                   inverse transforms in one kernel
//   X  += readout(X, F_d)                           fused gather + update
// Backward (exact adjoint; r2c^T = c2r(./c), c2r^T = c r2c with c = 1 on the
// k_z = 0, N/2 planes and 2 elsewhere):
//   gF_d = paint(X, g_d)                            3-component fixed-point paint
//   Q    = sum_d i D_d r2c_unnorm(gF_d)             z, y passes; x pass fused with
//   dL/dalpha = 1e-2 sum_H c G Re(conj(Q) U)        the FP64 parameter reductions
//   dL/dG     = 1e-2 alpha c Re(conj(Q) U)  -> kh, kl, nu
//   g_u  = c2r(1e-2 alpha G Q) / N^3                x inverse, y inverse, z c2r
//   g_rho = g_u gamma (rho+1e-8)^(gamma-1),  dL/dgamma = sum g_u u ln(rho+1e-8)
//   g_X  = g + sum_d g_d grad F_d + fact grad g_rho  fused gather
// Parameter reductions are per-workgroup FP64 partials summed in a fixed
// order, so gradients are bit-reproducible run to run.
//
// The dm kind (model.py:136-183) swaps in its kernel and an identity response.
// The HI kind (model.py:214-247) scales the potential by 1e-3 and adds a
// second source, the star map:
//   P = 1e-3 (alpha1 G1 U + alpha2 G2 S2),  S2 = r2c(starmap^gamma2) / N^3,
//   G2 = +exp(-k^2/kl2^2 - kh2^2/k^2) k^nu2,
// whose parameter partials the backward x pass adds against S2 and against
// T2 = r2c(starmap^gamma2 ln starmap) / N^3 (for gamma2).  Its response
// (model.py:313-334) is F = ReLU(w ReLU(rho + zs)^mu - dep + b) with the
// depletion map dep = ReLU(beta1 c2r(r2c(starmap) W) + eta) + beta2
// starmap^gamma3, W = exp(-k/kli - khi/k) k^nui + xi.
#include "cic.h"
#include "fft.h"

namespace ldl {

static constexpr int kBlk = 256;
static_assert(kBlk == kChunk, "gathers take one particle per thread");

// ---------------------------------------------------------------- tables
// Per-axis wavenumbers k = 2 pi fftfreq(N, cell) (pmesh.py:174-186; the z
// axis keeps N/2+1 entries, so its last one is -pi/cell) and the 6th-order
// finite-difference gradient symbol (45 sin kD - 9 sin 2kD + sin 3kD)/(30 D)
// (model.py:200-201), evaluated in FP64.
template <typename T> struct AxisTabs {
    T* k[3];
    T* D[3];
    int64_t n[3], count[3];
    double cell[3];
};
template <typename T>
__global__ void k_axis_tables(AxisTabs<T> a) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int d = 0;  // the three axes' tables, one after the other
    while (d < 2 && i >= a.count[d]) i -= a.count[d++];
    if (i >= a.count[d]) return;
    const int64_t n = a.n[d];
    const double cell = a.cell[d];
    T* __restrict__ k = a.k[d];
    T* __restrict__ D = a.D[d];
    const int64_t m = 2 * i < n ? i : i - n;  // fftfreq order, odd n included: 0..(n-1)/2, -(n-1)/2..-1
    const double th = 2.0 * (double)m / (double)n;  // k cell / pi
    k[i] = (T)(2.0 * M_PI * (double)m / ((double)n * cell));
    D[i] = (T)((45.0 * sinpi(th) - 9.0 * sinpi(2.0 * th) + sinpi(3.0 * th)) / (30.0 * cell));
}

template <typename T> struct KTabs {
    const T* k[3];
    const T* D[3];
    // |k|^2 with |k| = 1e-8 at k = 0 (model.py:102)
    __device__ __forceinline__ T knorm2(int64_t ix, int64_t iy, int64_t iz) const {
        const T a = k[0][ix], b = k[1][iy], c = k[2][iz];
        const T k2 = a * a + b * b + c * c;
        return k2 == T(0) ? T(1e-8) * T(1e-8) : k2;
    }
};

// Green's kernel of a layer, from k^2 (|k| = 1e-8 at k = 0):
//   sm (model.py:195): G = -exp(-k^2/kl^2 - kh^2/k^2) |k|^nu
//   dm (model.py:156): G = -exp(-|k|/kl  - kh/|k|)    |k|^nu
// with the terms its parameter gradients reuse: th, tl with
// dG/dkh = G gkh th and dG/dkl = G gkl tl (sm: th = 1/k^2, tl = k^2; dm:
// th = 1/|k|, tl = |k|; gkh, gkl from the host), and ln|k|.  FP32: hardware
// reciprocal, square root, log2 and exp2; FP64: library routines.
template <typename T> struct SmFilter {
    T kh, inv_kl, nu;  // sm: kh^2, 1/kl^2; dm: kh, 1/kl
    int dm;
    __device__ __forceinline__ void eval(T k2, T& G, T& th, T& tl, T& lnk) const {
        if constexpr (sizeof(T) == 4) {
            const float l2 = __builtin_amdgcn_logf(k2);  // log2 k^2
            lnk = 0.5f * kLn2f * l2;
            const float pw = nu == 0.0f ? 1.0f : __builtin_amdgcn_exp2f(0.5f * nu * l2);
            if (dm) {
                tl = __builtin_amdgcn_sqrtf(k2);
                th = __builtin_amdgcn_rcpf(tl);
            } else {
                tl = k2;
                th = __builtin_amdgcn_rcpf(k2);
            }
            G = -__expf(-tl * inv_kl - kh * th) * pw;
        } else {
            lnk = 0.5 * log(k2);
            const double pw = nu == 0.0 ? 1.0 : pow(k2, 0.5 * nu);
            tl = dm ? sqrt(k2) : k2;
            th = 1.0 / tl;
            G = -exp(-tl * inv_kl - kh * th) * pw;
        }
    }
    __device__ __forceinline__ T G(T k2) const {
        T g, th, tl, lnk;
        eval(k2, g, th, tl, lnk);
        return g;
    }
};

// Filter parameters and gradient coefficients of one layer from prm =
// [alpha, gamma, kh, kl, nu].
template <typename T>
inline SmFilter<T> layer_filter(const double* prm, int dm) {
    return dm ? SmFilter<T>{(T)prm[2], (T)(1.0 / prm[3]), (T)prm[4], 1}
              : SmFilter<T>{(T)(prm[2] * prm[2]), (T)(1.0 / (prm[3] * prm[3])), (T)prm[4], 0};
}
inline void filter_grad_coeffs(const double* prm, int dm, double* gkh, double* gkl) {
    if (dm) {
        *gkh = -1.0;
        *gkl = 1.0 / (prm[3] * prm[3]);
    } else {
        *gkh = -2.0 * prm[2];
        *gkl = 2.0 / (prm[3] * prm[3] * prm[3]);
    }
}

// Row pitch of the half spectra the layer, loss and depletion entry points
// keep (U, S2 and all workspace spectra): N2/2+1 complex rounded up to 64
// bytes.  The column passes move 16-column strips of every row; with rows
// 64-byte aligned those are whole 64-byte sectors (an odd N2/2+1 pitch makes
// the strided x pass about twice slower).
template <typename T> inline int64_t spec_pitch(const Mesh& m) {
    constexpr int64_t q = 64 / (2 * (int64_t)sizeof(T));
    return (m.nh() + q - 1) / q * q;
}
template <typename T> inline int64_t spec_cells(const Mesh& m) { return m.n0 * m.n1 * spec_pitch<T>(m); }

// Workspace carve-up of ldl_sm_* calls.
struct SmWs {
    size_t paint, spec3, mesh3, tabs, part, total;
    int64_t part_doubles;
};
static SmWs sm_ws(const Mesh& m, int tsize, int64_t np, int32_t pflags) {
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    SmWs w;
    w.paint = al(paint_ws_bytes(m, 3, np, pflags));
    const int64_t hq = (64 / (2 * tsize)), hcp = m.n0 * m.n1 * ((m.nh() + hq - 1) / hq * hq);  // spec_cells
    w.spec3 = al((size_t)3 * hcp * 2 * tsize);
    // three meshes; also holds one pitched spectrum (the HI layer's T2)
    w.mesh3 = al((size_t)std::max<int64_t>(3 * m.cells(), 2 * hcp) * tsize);
    w.tabs = al((size_t)2 * (m.n0 + m.n1 + m.nh()) * tsize);
    w.part_doubles = 12 * (m.n0 * m.n1 + m.n1 * (m.nh() + 1) + 4096);
    w.part = al((size_t)w.part_doubles * 8);
    w.total = w.paint + w.spec3 + w.mesh3 + w.tabs + w.part;
    return w;
}

template <typename T> struct SmPtrs {
    void* paint;
    c2<T>* spec3;
    T* mesh3;
    T* tabs;
    double* part;
};
template <typename T>
static SmPtrs<T> sm_ptrs(void* ws, const SmWs& w) {
    char* b = reinterpret_cast<char*>(ws);
    SmPtrs<T> p;
    p.paint = b;
    p.spec3 = reinterpret_cast<c2<T>*>(b + w.paint);
    p.mesh3 = reinterpret_cast<T*>(b + w.paint + w.spec3);
    p.tabs = reinterpret_cast<T*>(b + w.paint + w.spec3 + w.mesh3);
    p.part = reinterpret_cast<double*>(b + w.paint + w.spec3 + w.mesh3 + w.tabs);
    return p;
}

template <typename T>
static KTabs<T> build_tables(T* tabs, const Mesh& m, const double box[3], hipStream_t st) {
    const int64_t cnt[3] = {m.n0, m.n1, m.nh()};
    const int64_t nn[3] = {m.n0, m.n1, m.n2};
    KTabs<T> kt;
    AxisTabs<T> a;
    T* p = tabs;
    for (int d = 0; d < 3; d++) {
        a.k[d] = p;
        a.D[d] = p + cnt[d];
        a.n[d] = nn[d];
        a.count[d] = cnt[d];
        a.cell[d] = box[d] / (double)nn[d];
        kt.k[d] = a.k[d];
        kt.D[d] = a.D[d];
        p += 2 * cnt[d];
    }
    hipLaunchKernelGGL(k_axis_tables<T>, dim3(grid_for(cnt[0] + cnt[1] + cnt[2], kBlk)), dim3(kBlk), 0, st, a);
    return kt;
}

// Sum column j of a (nblocks x stride) FP64 array and add it to out[map[j]]
// (or store when ``assign``); fixed order, one workgroup.
struct RedMap {
    int n;
    int map[12];
};
// One pass over the rows: thread t sums column j of rows t, t + 256, ... in
// that order, then the waves and the four wave sums in a fixed order.  Four
// rows' loads are issued before they are added (in the same order).
__global__ void __launch_bounds__(256) k_reduce_cols(const double* __restrict__ part, int64_t nblocks, int stride, RedMap rm,
                                                     double* __restrict__ out, int assign) {
    constexpr int MX = 12, U = 4;
    __shared__ double red[MX][4];
    double s[MX];
#pragma unroll
    for (int j = 0; j < MX; j++) s[j] = 0.0;
    int64_t b = threadIdx.x;
    for (; b + (U - 1) * 256 < nblocks; b += U * 256) {
        double v[U][MX];
#pragma unroll
        for (int u = 0; u < U; u++)
#pragma unroll
            for (int j = 0; j < MX; j++) v[u][j] = j < rm.n ? part[(b + u * 256) * stride + j] : 0.0;
#pragma unroll
        for (int u = 0; u < U; u++)
#pragma unroll
            for (int j = 0; j < MX; j++)
                if (j < rm.n) s[j] += v[u][j];
    }
    for (; b < nblocks; b += 256) {
        const double* row = part + b * stride;
#pragma unroll
        for (int j = 0; j < MX; j++)
            if (j < rm.n) s[j] += row[j];
    }
#pragma unroll
    for (int j = 0; j < MX; j++) {
        if (j < rm.n) {
            const double v = wave_sum(s[j]);
            if ((threadIdx.x & 63) == 0) red[j][threadIdx.x >> 6] = v;
        }
    }
    __syncthreads();
    if ((int)threadIdx.x < rm.n) {
        const int j = threadIdx.x;
        const double t = (red[j][0] + red[j][1]) + (red[j][2] + red[j][3]);
        if (assign) out[rm.map[j]] = t;
        else out[rm.map[j]] += t;
    }
}

// --------------------------------------------------------- z-pass functors
template <typename T> struct PowRowLoad {  // (rho + 1e-8)^gamma, model.py:193
    const T* rho;
    int64_t n2;
    T gamma;
    __device__ __forceinline__ c2<T> load2(int64_t row, int m) const {
        const c2<T> r = *reinterpret_cast<const c2<T>*>(rho + row * n2 + 2 * m);
        return c2<T>{ppow(r.x + T(1e-8), gamma), ppow(r.y + T(1e-8), gamma)};
    }
};

// starmap^g with 0^0 = 1, and starmap^g ln(starmap) with the value 0 at a zero
// base: the derivative of x^g in g as JAX forms it (log of a zero base taken
// as log 1), which the reference's gamma2 and gamma3 gradients inherit.
template <typename T> __device__ __forceinline__ T spow(T x, T g) { return (x == T(0) && g == T(0)) ? T(1) : ppow(x, g); }
template <typename T> __device__ __forceinline__ T spow_log(T x, T g) { return x == T(0) ? T(0) : ppow(x, g) * plog(x); }

template <typename T> struct StarRowLoad {  // starmap^gamma2 (model.py:228), or times ln starmap
    const T* star;
    int64_t n2;
    T gamma;
    int with_log;
    __device__ __forceinline__ c2<T> load2(int64_t row, int m) const {
        const c2<T> r = *reinterpret_cast<const c2<T>*>(star + row * n2 + 2 * m);
        return with_log ? c2<T>{spow_log(r.x, gamma), spow_log(r.y, gamma)} : c2<T>{spow(r.x, gamma), spow(r.y, gamma)};
    }
};

// Three half spectra (rows of pitch nh) as rows taken component-fastest: row
// R is row R / 3 of spectrum R % 3, so the rows of one cell's three
// components are transformed by the same workgroup.  With Dz set, the third
// spectrum is multiplied by -i D_z(k_z) on load.
template <typename T> struct Tri3RowLoad {
    const c2<T>* in0;
    const c2<T>* in1;
    const c2<T>* in2;
    int64_t nh;
    __device__ __forceinline__ c2<T> load(int64_t R, int k) const {
        const int64_t r = R / 3;
        const int d = (int)(R - 3 * r);
        const c2<T>* in = d == 0 ? in0 : d == 1 ? in1 : in2;  // (no dynamically indexed array: it would live in scratch)
        return in[r * nh + k];
    }
};


// Rows taken component-fastest (Tri3RowLoad) into the interleaved field F3i
// (x, y, z packed per cell): the power-of-two z pass hands over whole threes
// of rows (store3, two cells per call); the any-length pass single rows
// (store2).
template <typename T> struct Tri3Store3 {
    static constexpr int NR = 0;
    static constexpr int TRIPLE = 1;
    T* out;
    int64_t n2;
    __device__ __forceinline__ void reduce_out(int64_t, const double*) const {}
    __device__ __forceinline__ void store2(int64_t R, int m, c2<T> v, double*) const {
        const int64_t r = R / 3;
        T* o = out + 3 * (r * n2 + 2 * m) + (R - 3 * r);
        o[0] = v.x;
        o[3] = v.y;
    }
    __device__ __forceinline__ void store3(int64_t r, int m, c2<T> a, c2<T> b, c2<T> c) const {
        T* o = out + 3 * (r * n2 + 2 * m);  // two adjacent cells: 6 consecutive values
        o[0] = a.x;
        o[1] = b.x;
        o[2] = c.x;
        o[3] = a.y;
        o[4] = b.y;
        o[5] = c.y;
    }
};

// g_u -> g_rho = g_u gamma (rho+1e-8)^(gamma-1); dL/dgamma partial.
template <typename T> struct GammaBwdStore {
    static constexpr int NR = 1;
    const T* rho;
    T* grho;
    int64_t n2;
    T scale, gamma;
    double* part;
    __device__ __forceinline__ void store2(int64_t row, int m, c2<T> v, double* acc) const {
        const int64_t i = row * n2 + 2 * m;
        const c2<T> r = *reinterpret_cast<const c2<T>*>(rho + i);
        const T g0 = v.x * scale, g1 = v.y * scale;
        T p0, q0, l0, p1, q1, l1;
        pow_pair_log(r.x + T(1e-8), gamma, p0, q0, l0);
        pow_pair_log(r.y + T(1e-8), gamma, p1, q1, l1);
        acc[0] += (double)(g0 * p0) * (double)l0 + (double)(g1 * p1) * (double)l1;
        *reinterpret_cast<c2<T>*>(grho + i) = c2<T>{g0 * gamma * q0, g1 * gamma * q1};
    }
    __device__ __forceinline__ void reduce_out(int64_t b, const double* acc) const { part[b] = acc[0]; }
};

// ------------------------------------------------------- fused x passes
// Second potential source of the HI kind: P += a G_sm(kh2, kl2, nu2) S with
// a = -1e-3 alpha2 (the sign turns the sm kernel into G2).  The backward adds
// the partials of alpha2, kh2, kl2, nu2 (against S) and gamma2 (against Tl).
template <typename T> struct Src2 {
    const c2<T>* S;   // r2c(starmap^gamma2) / N^3
    const c2<T>* Tl;  // r2c(starmap^gamma2 ln starmap) / N^3 (backward)
    SmFilter<T> fl;
    T a;
    double ad, cs, gkh, gkl;  // ad = a in FP64, cs = -1e-3
};

// Half-spectrum weights of the backward x pass at z bin kz (n2 the z
// length, zdbl = min(n1 / 2, n2 / 2 + 1)):
//   c  = 1 on the k_z = 0, n2/2 planes, 2 elsewhere (the exact adjoint);
//   ca = weight of the alpha partials, cg = weight of the kernel-parameter
//        partials (kh, kl, nu).
// grad_rule "exact": ca = cg = c.  grad_rule "reference" reproduces the
// reference's rules: its FFT VJPs (jaxpops_bind.py:699-701, 758-759) leave
// every half-spectrum bin at weight 1, and c2f's backward (model.py:23-29)
// doubles the kernel's cotangent on z bins [1, int(n1/2)) -- the second axis
// length, clipped to the half spectrum -- so ca = 1 and cg = 2 there, 1
// elsewhere (cg = c on meshes with n1 == n2).
struct XbwdW {
    double c, ca, cg;
};
__device__ __forceinline__ XbwdW xbwd_weights(int64_t kz, int64_t n2, int64_t zdbl, int alpha_rule) {
    const double c = (kz == 0 || 2 * kz == n2) ? 1.0 : 2.0;
    if (!alpha_rule) return XbwdW{c, c, c};
    return XbwdW{c, 1.0, (kz >= 1 && kz < zdbl) ? 2.0 : 1.0};
}

// Per-cell parameter partials of the backward x pass.  acc[0..3]: alpha,
// kh, kl, nu of the first source; with HI acc[4..8]: alpha2, kh2, kl2, nu2,
// gamma2 (gamma2 reaches the potential through r2c and the three c2r, whose
// rule pair composes to the exact adjoint: weight c in both rules).  Returns
// G of the first source.
template <typename T, bool HI>
__device__ __forceinline__ T xbwd_cell(const SmFilter<T>& fl, const Src2<T>& s2, T k2, c2<T> q, c2<T> u, int64_t idx, XbwdW w,
                                       double cpot, double apot, double gkh, double gkl, double* acc) {
    T G, th, tl, lnk;
    fl.eval(k2, G, th, tl, lnk);
    const double re = (double)q.x * (double)u.x + (double)q.y * (double)u.y;
    const double Gd = (double)G;
    acc[0] += w.ca * cpot * Gd * re;
    const double gG = apot * w.cg * re * Gd;  // dL/dG * G
    acc[1] += gG * gkh * (double)th;
    acc[2] += gG * gkl * (double)tl;
    acc[3] += gG * (double)lnk;
    if constexpr (HI) {  // both kernels are of sm form: th, tl, lnk are shared
        const double G2 = (double)s2.fl.G(k2);
        const c2<T> sv = s2.S[idx], tv = s2.Tl[idx];
        const double re2 = (double)q.x * (double)sv.x + (double)q.y * (double)sv.y;
        const double re3 = (double)q.x * (double)tv.x + (double)q.y * (double)tv.y;
        acc[4] += w.ca * s2.cs * G2 * re2;
        const double gG2 = s2.ad * w.cg * re2 * G2;
        acc[5] += gG2 * s2.gkh * (double)th;
        acc[6] += gG2 * s2.gkl * (double)tl;
        acc[7] += gG2 * (double)lnk;
        acc[8] += s2.ad * w.c * G2 * re3;
    }
    return G;
}

// Forward: X FFT, save U = . / N^3, P = 1e-2 alpha G U, inverse X FFT of P
// into out.  The derivative symbols are applied in real space afterwards
// (k_grad6).  FP32 512-thread workgroups without the HI source keep the
// whole tile in registers from the load to the store (RE: 128 VGPRs, the
// two workgroups per CU the LDS allows); the others run the filter through
// LDS, which would otherwise spill.  Element indices of a tile fit 32 bits
// (every axis is at most 1024), so its accesses are 32-bit offsets from a
// uniform base.
template <typename T, int N, bool HI, int P = FftCfg<T, N>::P>
__global__ void __launch_bounds__((P * FftCfg<T, N>::TPP), (P * FftCfg<T, N>::TPP == 512 && sizeof(T) == 4 && !HI ? 4 : 1))
k_sm_x_fwd(ColGeom g, const c2<T>* __restrict__ tw, c2<T>* __restrict__ U, c2<T>* __restrict__ out, KTabs<T> kt,
           SmFilter<T> fl, T a1e2, T invM, Src2<T> s2) {
    using Cfg = FftCfg<T, N>;
    using C = c2<T>;
    constexpr int LS = P + 1, TH = P * Cfg::TPP, EPT = N * P / TH;
    constexpr bool RE = TH == 512 && sizeof(T) == 4 && !HI;
    __shared__ C lds[N * LS];
    __shared__ C twl[N];
    const int64_t tile = xcd_tile(blockIdx.x, gridDim.x);
    const int64_t plane = tile / g.tiles_per_plane;
    const int64_t col0 = (tile - plane * g.tiles_per_plane) * P;
    const int64_t base = plane * g.plane_stride + col0;
    const int pv = (int)((g.ncols - col0) < P ? (g.ncols - col0) : P);
    C* Ub = U + base;
    C* Ob = out + base;
    const uint32_t es = (uint32_t)g.es;
    stage_twiddles<T, N, TH>(twl, tw);
    C x[EPT];
    reg_load<T, N, P, TH>(x, pv, [&](int n, int p) { return Ub[(uint32_t)n * es + (uint32_t)p]; });
    if constexpr (RE) reg_fft<T, N, P, Cfg::TPP, -1>(x, lds, twl);
    else fft_from_regs<T, N, P, Cfg::TPP, -1>(x, lds, twl);
    // each thread rewrites only its own elements: no barrier between
#pragma unroll
    for (int s = 0; s < EPT; s++) {
        const int i = threadIdx.x + s * TH;
        const int n = i / P, p = i - n * P;
        C u = RE ? x[s] : lds[n * LS + p];
        u = C{u.x * invM, u.y * invM};
        C pk = C{T(0), T(0)};
        if (p < pv) {
            Ub[(uint32_t)n * es + (uint32_t)p] = u;
            const T k2 = kt.knorm2(n, plane, col0 + p);
            const T G = fl.G(k2);
            pk = C{u.x * a1e2 * G, u.y * a1e2 * G};
            if constexpr (HI) {
                const C sv = s2.S[base + n * g.es + p];
                const T G2 = s2.a * s2.fl.G(k2);
                pk = C{pk.x + G2 * sv.x, pk.y + G2 * sv.y};
            }
        }
        if constexpr (RE) x[s] = pk;
        else lds[n * LS + p] = pk;
    }
    if constexpr (RE) {
        reg_fft<T, N, P, Cfg::TPP, +1>(x, lds, twl);
    } else {
        __syncthreads();
        lds_fft<T, N, P, Cfg::TPP, +1, 1, 1, fft_last_ns(N)>(lds, twl);
        fft_to_regs<T, N, P, Cfg::TPP, +1>(x, lds, twl);
    }
    reg_store<T, N, P, TH>(x, pv, [&](int n, int p, C v) { Ob[(uint32_t)n * es + (uint32_t)p] = v; });
}

// Backward: X FFT of Q's y,z-transformed spectrum (Q = sum_d i D_d times
// the transforms of the painted gradient meshes, formed as the transform of
// their real-space divergence, k_div6); the FP64 parameter partials (alpha,
// kh, kl, nu; HI also the second source's) against the saved U, then one
// inverse X FFT of cpot alpha G Q into out (may alias Sq: each workgroup
// reads its tile whole before it writes).
// Pencils per workgroup of the backward x pass: FP32 N = 512 takes 8 (its
// tile, with the reduction in the tile's LDS, is 40 KB: four workgroups per
// CU; 0.513 -> 0.496 ms at 512^3, where the forward pass is faster at 16).
template <typename T, int N> constexpr int xbwd_pencils() { return sizeof(T) == 4 && N == 512 ? 8 : FftCfg<T, N>::P; }

template <typename T, int N, bool HI, int P = xbwd_pencils<T, N>()>
__global__ void __launch_bounds__((P * FftCfg<T, N>::TPP))
k_sm_x_bwd(ColGeom g, const c2<T>* Sq, const c2<T>* __restrict__ tw, const c2<T>* __restrict__ U, c2<T>* out,
           int64_t n2, int64_t zdbl, KTabs<T> kt, SmFilter<T> fl, T a1e2, double gkh, double gkl, double cpot, double alpha,
           int alpha_rule, Src2<T> s2, double* __restrict__ part) {
    using Cfg = FftCfg<T, N>;
    using C = c2<T>;
    constexpr int LS = P + 1, TH = P * Cfg::TPP, EPT = N * P / TH;
    __shared__ C lds[N * LS];
    __shared__ C twl[N];
    const int64_t tile = xcd_tile(blockIdx.x, gridDim.x);
    const int64_t plane = tile / g.tiles_per_plane;
    const int64_t col0 = (tile - plane * g.tiles_per_plane) * P;
    const int64_t base = plane * g.plane_stride + col0;
    const int pv = (int)((g.ncols - col0) < P ? (g.ncols - col0) : P);
    const C* Sb = Sq + base;
    const C* Ub = U + base;
    C* Ob = out + base;
    const uint32_t es = (uint32_t)g.es;  // tile element indices fit 32 bits, as in k_sm_x_fwd
    stage_twiddles<T, N, TH>(twl, tw);
    // the first forward and last inverse stages in registers; the epilogue
    // through LDS (in registers its FP64 partials would spill)
    C x[EPT];
    reg_load<T, N, P, TH>(x, pv, [&](int n, int p) { return Sb[(uint32_t)n * es + (uint32_t)p]; });
    fft_from_regs<T, N, P, Cfg::TPP, -1>(x, lds, twl);
    constexpr int NP = HI ? 9 : 4;
    double acc[NP] = {};
    const double apot = cpot * alpha;
    // each thread rewrites only its own elements: no barrier between
#pragma unroll
    for (int s = 0; s < EPT; s++) {
        const int i = threadIdx.x + s * TH;
        const int n = i / P, p = i - n * P;
        const C q = lds[n * LS + p];
        C y = C{T(0), T(0)};
        if (p < pv) {
            const int64_t kz = col0 + p;
            const uint32_t e = (uint32_t)n * es + (uint32_t)p;
            const T G = xbwd_cell<T, HI>(fl, s2, kt.knorm2(n, plane, kz), q, Ub[e], base + e, xbwd_weights(kz, n2, zdbl, alpha_rule),
                                         cpot, apot, gkh, gkl, acc);
            y = C{a1e2 * G * q.x, a1e2 * G * q.y};
        }
        lds[n * LS + p] = y;
    }
    __syncthreads();
    lds_fft<T, N, P, Cfg::TPP, +1, 1, 1, fft_last_ns(N)>(lds, twl);
    fft_to_regs<T, N, P, Cfg::TPP, +1>(x, lds, twl);
    reg_store<T, N, P, TH>(x, pv, [&](int n, int p, C v) { Ob[(uint32_t)n * es + (uint32_t)p] = v; });
    __syncthreads();  // the tile's last reads are done: its LDS holds the reduction
    static_assert(NP * (TH / 64) * sizeof(double) <= sizeof(lds), "reduction fits the tile");
    block_sum_in<NP, TH>(acc, reinterpret_cast<double*>(lds));
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < NP; j++) part[blockIdx.x * NP + j] = acc[j];
    }
}

// ------------------------------------------- real-space derivative stencils
// The derivative symbol of k_axis_tables, D_d = (45 sin t - 9 sin 2t +
// sin 3t) / (30 h_d) with t = 2 pi n / N_d, is that of the 6th-order central
// difference
//   diff_d(a)[x] = 45 (a[x+1] - a[x-1]) - 9 (a[x+2] - a[x-2]) + (a[x+3] - a[x-3])
// along axis d (circular): a circular shift by s multiplies transform bin n
// by exp(+-2 pi i n s / N), so the transform of diff_d(a) / (60 h_d) is
// i D_d times the transform of a, and the inverse transform of -i D_d P is
// -diff_d(phi) / (60 h_d) with phi the inverse transform of P.  The layer
// therefore transforms one mesh each way: k_grad6 turns the potential phi
// into the forces F_d = -diff_d(phi) / (60 h_d), and k_div6 turns the
// painted gradient meshes g_d into sum_d diff_d(g_d) / (60 h_d), whose
// transform is the sum_d i D_d of theirs.
//
// A workgroup of 4 st_tz<T>() threads covers kStTY (y) x st_tz<T>() (z) cells (each thread
// 4 cells along y at one z) and marches through kStXC x planes: the x
// difference comes from a 7-plane register queue per cell, the y and z
// differences from the current plane staged in LDS with 3-cell halos.  All
// indices are wrapped, so a tile larger than the mesh reads valid periodic
// values and only its stores are bounded.
constexpr int kStTY = 16, kStXC = 64;
// z width of a tile: its 3-cell z halo costs two extra cache lines per row, so
// wider rows fetch less (FP32, 512^3: grad6 0.500 -> 0.466 ms, div6 0.471 ->
// 0.432 ms from 64 to 256); FP64 keeps 128 (at 1024 threads its divergence
// would spill)
template <typename T> constexpr int st_tz() { return sizeof(T) == 4 ? 256 : 128; }

struct StGeom {
    int n0, n1, n2, tiles_y, tiles_z, chunks;
};

__device__ __forceinline__ int st_wrap(int v, int n) { return ((v % n) + n) % n; }

template <typename T>
__device__ __forceinline__ T st_diff(T m3, T m2, T m1, T p1, T p2, T p3) {
    return T(45) * (p1 - m1) - T(9) * (p2 - m2) + (p3 - m3);
}

template <int TZ> struct StTile {
    int x0, x1, y0, z0, tz, yg;
    int zw, yw[4];
    bool zok, yok[4];
    // halo cells this thread stages: y-halo rows hy[0..1] (row index of the
    // (kStTY + 6)-row plane, -1 none) at its own z; one z-halo cell (zrow, zcol)
    int hy[2], hyw[2], zrow, zcol, zyw, zzw;
    __device__ __forceinline__ StTile(const StGeom& g) {
        const int t = (int)xcd_tile(blockIdx.x, gridDim.x);
        const int tzt = t % g.tiles_z, tyt = (t / g.tiles_z) % g.tiles_y, xc = t / (g.tiles_z * g.tiles_y);
        x0 = xc * kStXC;
        x1 = x0 + kStXC < g.n0 ? x0 + kStXC : g.n0;
        y0 = tyt * kStTY;
        z0 = tzt * TZ;
        tz = threadIdx.x % TZ;
        yg = threadIdx.x / TZ;
        zw = st_wrap(z0 + tz, g.n2);
        zok = z0 + tz < g.n2;
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const int y = y0 + 4 * yg + c;
            yw[c] = st_wrap(y, g.n1);
            yok[c] = y < g.n1;
        }
        // y halo: rows 0..2 and kStTY+3..kStTY+5; thread rows yg and yg + 4
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int r = yg + 4 * h;
            hy[h] = r < 6 ? (r < 3 ? r : kStTY + r) : -1;
            hyw[h] = st_wrap(y0 + (hy[h] < 0 ? 0 : hy[h]) - 3, g.n1);
        }
        // z halo: kStTY rows x 6 columns (0..2, TZ+3..TZ+5), threads 0..95
        const int tt = threadIdx.x;
        zrow = tt < 6 * kStTY ? tt / 6 : -1;
        const int k = tt - (tt / 6) * 6;
        zcol = k < 3 ? k : TZ + k;
        zyw = st_wrap(y0 + (zrow < 0 ? 0 : zrow), g.n1);
        zzw = st_wrap(z0 + zcol - 3, g.n2);
    }
};

// Forces F (N0, N1, N2, 3), interleaved per cell, from the potential phi.
template <typename T>
__global__ void __launch_bounds__(4 * st_tz<T>()) k_grad6(const T* __restrict__ phi, T* __restrict__ F, StGeom g, T cx, T cy, T cz) {
    constexpr int TZ = st_tz<T>();
    __shared__ T pl[kStTY + 6][TZ + 6];
    const StTile<TZ> tl(g);
    if (tl.x0 >= g.n0) return;  // uniform over the workgroup
    const int64_t ps = (int64_t)g.n1 * g.n2;
    auto at = [&](int x, int y, int z) { return phi[(int64_t)st_wrap(x, g.n0) * ps + (int64_t)y * g.n2 + z]; };
    T q[4][7];
#pragma unroll
    for (int c = 0; c < 4; c++)
#pragma unroll
        for (int k = 0; k < 7; k++) q[c][k] = at(tl.x0 - 3 + k, tl.yw[c], tl.zw);
    T hy0 = at(tl.x0, tl.hyw[0], tl.zw), hy1 = tl.hy[1] >= 0 ? at(tl.x0, tl.hyw[1], tl.zw) : T(0);
    T hz = tl.zrow >= 0 ? at(tl.x0, tl.zyw, tl.zzw) : T(0);
    for (int x = tl.x0; x < tl.x1; x++) {
        __syncthreads();  // the previous plane's reads are done
#pragma unroll
        for (int c = 0; c < 4; c++) pl[4 * tl.yg + c + 3][tl.tz + 3] = q[c][3];
        pl[tl.hy[0]][tl.tz + 3] = hy0;
        if (tl.hy[1] >= 0) pl[tl.hy[1]][tl.tz + 3] = hy1;
        if (tl.zrow >= 0) pl[tl.zrow + 3][tl.zcol] = hz;
        __syncthreads();
        // next plane's halo and the queue head, in flight during this plane
        T head[4];
        const bool more = x + 1 < tl.x1;
        if (more) {
#pragma unroll
            for (int c = 0; c < 4; c++) head[c] = at(x + 4, tl.yw[c], tl.zw);
            hy0 = at(x + 1, tl.hyw[0], tl.zw);
            if (tl.hy[1] >= 0) hy1 = at(x + 1, tl.hyw[1], tl.zw);
            if (tl.zrow >= 0) hz = at(x + 1, tl.zyw, tl.zzw);
        }
        if (tl.zok) {
#pragma unroll
            for (int c = 0; c < 4; c++) {
                if (!tl.yok[c]) continue;
                const int ly = 4 * tl.yg + c + 3, lz = tl.tz + 3;
                const T dx = st_diff(q[c][0], q[c][1], q[c][2], q[c][4], q[c][5], q[c][6]);
                const T dy = st_diff(pl[ly - 3][lz], pl[ly - 2][lz], pl[ly - 1][lz], pl[ly + 1][lz], pl[ly + 2][lz], pl[ly + 3][lz]);
                const T dz = st_diff(pl[ly][lz - 3], pl[ly][lz - 2], pl[ly][lz - 1], pl[ly][lz + 1], pl[ly][lz + 2], pl[ly][lz + 3]);
                T* o = F + 3 * ((int64_t)x * ps + (int64_t)(tl.y0 + 4 * tl.yg + c) * g.n2 + (tl.z0 + tl.tz));
                o[0] = -cx * dx;
                o[1] = -cy * dy;
                o[2] = -cz * dz;
            }
        }
        if (more) {
#pragma unroll
            for (int c = 0; c < 4; c++) {
#pragma unroll
                for (int k = 0; k < 6; k++) q[c][k] = q[c][k + 1];
                q[c][6] = head[c];
            }
        }
    }
}

// Divergence out = sum_d c_d diff_d(g_d) of the meshes g (3 consecutive
// (N0, N1, N2) meshes) into a real mesh.
template <typename T>
__global__ void __launch_bounds__(4 * st_tz<T>()) k_div6(const T* __restrict__ gm, T* __restrict__ out, StGeom g, T cx, T cy, T cz) {
    constexpr int TZ = st_tz<T>();
    __shared__ T ly_[kStTY + 6][TZ];
    __shared__ T lz_[kStTY][TZ + 6];
    const StTile<TZ> tl(g);
    if (tl.x0 >= g.n0) return;  // uniform over the workgroup
    const int64_t ps = (int64_t)g.n1 * g.n2, cells = ps * g.n0;
    const T* gx = gm;
    const T* gy = gm + cells;
    const T* gz = gm + 2 * cells;
    auto at = [&](const T* a, int x, int y, int z) { return a[(int64_t)st_wrap(x, g.n0) * ps + (int64_t)y * g.n2 + z]; };
    T q[4][7], vy[4], vz[4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
#pragma unroll
        for (int k = 0; k < 7; k++) q[c][k] = at(gx, tl.x0 - 3 + k, tl.yw[c], tl.zw);
        vy[c] = at(gy, tl.x0, tl.yw[c], tl.zw);
        vz[c] = at(gz, tl.x0, tl.yw[c], tl.zw);
    }
    T hy0 = at(gy, tl.x0, tl.hyw[0], tl.zw), hy1 = tl.hy[1] >= 0 ? at(gy, tl.x0, tl.hyw[1], tl.zw) : T(0);
    T hz = tl.zrow >= 0 ? at(gz, tl.x0, tl.zyw, tl.zzw) : T(0);
    for (int x = tl.x0; x < tl.x1; x++) {
        __syncthreads();  // the previous plane's reads are done
#pragma unroll
        for (int c = 0; c < 4; c++) {
            ly_[4 * tl.yg + c + 3][tl.tz] = vy[c];
            lz_[4 * tl.yg + c][tl.tz + 3] = vz[c];
        }
        ly_[tl.hy[0]][tl.tz] = hy0;
        if (tl.hy[1] >= 0) ly_[tl.hy[1]][tl.tz] = hy1;
        if (tl.zrow >= 0) lz_[tl.zrow][tl.zcol] = hz;
        __syncthreads();
        T head[4];
        const bool more = x + 1 < tl.x1;
        if (more) {
#pragma unroll
            for (int c = 0; c < 4; c++) {
                head[c] = at(gx, x + 4, tl.yw[c], tl.zw);
                vy[c] = at(gy, x + 1, tl.yw[c], tl.zw);
                vz[c] = at(gz, x + 1, tl.yw[c], tl.zw);
            }
            hy0 = at(gy, x + 1, tl.hyw[0], tl.zw);
            if (tl.hy[1] >= 0) hy1 = at(gy, x + 1, tl.hyw[1], tl.zw);
            if (tl.zrow >= 0) hz = at(gz, x + 1, tl.zyw, tl.zzw);
        }
        if (tl.zok) {
#pragma unroll
            for (int c = 0; c < 4; c++) {
                if (!tl.yok[c]) continue;
                const int r = 4 * tl.yg + c, z = tl.tz;
                const T dx = st_diff(q[c][0], q[c][1], q[c][2], q[c][4], q[c][5], q[c][6]);
                const T dy = st_diff(ly_[r][z], ly_[r + 1][z], ly_[r + 2][z], ly_[r + 4][z], ly_[r + 5][z], ly_[r + 6][z]);
                const T dz = st_diff(lz_[r][z], lz_[r][z + 1], lz_[r][z + 2], lz_[r][z + 4], lz_[r][z + 5], lz_[r][z + 6]);
                out[(int64_t)x * ps + (int64_t)(tl.y0 + r) * g.n2 + (tl.z0 + z)] = cx * dx + cy * dy + cz * dz;
            }
        }
        if (more) {
#pragma unroll
            for (int c = 0; c < 4; c++) {
#pragma unroll
                for (int k = 0; k < 6; k++) q[c][k] = q[c][k + 1];
                q[c][6] = head[c];
            }
        }
    }
}

template <typename T> static StGeom st_geom(const Mesh& m) {
    StGeom g;
    g.n0 = (int)m.n0;
    g.n1 = (int)m.n1;
    g.n2 = (int)m.n2;
    g.tiles_y = (int)((m.n1 + kStTY - 1) / kStTY);
    g.tiles_z = (int)((m.n2 + st_tz<T>() - 1) / st_tz<T>());
    g.chunks = (int)((m.n0 + kStXC - 1) / kStXC);
    return g;
}

// One axis of a separable Gaussian smoothing, W = exp(-|k|^2 sigma^2 / 2) =
// prod_d exp(-k_d^2 sigma^2 / 2): forward FFT along the column axis, times
// exp(-hs2 k_d^2) (hs2 = sigma^2 / 2), inverse, unnormalised (the z pass's
// store applies 1 / (N0 N1 N2) for all three).  It runs on a real mesh
// viewed as complex pairs (z, z+1): the weight is real and even in k_d, so
// its kernel is real and the two lines of a pair are filtered independently.
// With ``msum`` the first pass subtracts the field's mean msum[0] * inv (see
// sm_loss: the mean is carried in FP64 around the FP32 transforms).
template <typename T, int N>
__global__ void __launch_bounds__((FftCfg<T, N>::THREADS))
k_gauss_line(ColGeom g, const c2<T>* __restrict__ tw, c2<T>* __restrict__ io, const T* __restrict__ kd, T hs2,
             const double* __restrict__ msum, double inv) {
    using Cfg = FftCfg<T, N>;
    using C = c2<T>;
    constexpr int P = Cfg::P, TH = Cfg::THREADS;
    __shared__ C lds[Cfg::LDS_ELEMS];
    __shared__ C twl[N];
    const int64_t tile = xcd_tile(blockIdx.x, gridDim.x);
    const int64_t plane = tile / g.tiles_per_plane;
    const int64_t col0 = (tile - plane * g.tiles_per_plane) * P;
    const int64_t base = plane * g.plane_stride + col0;
    const int pv = (int)((g.ncols - col0) < P ? (g.ncols - col0) : P);
    stage_twiddles<T, N, TH>(twl, tw);
    const T mu = msum ? (T)(msum[0] * inv) : T(0);
    C x[N * P / TH];
    reg_load<T, N, P, TH>(x, pv, [&](int n, int p) {
        const C v = io[base + n * g.es + p];
        return C{v.x - mu, v.y - mu};
    });
    reg_fft<T, N, P, Cfg::TPP, -1>(x, lds, twl);
#pragma unroll
    for (int s = 0; s < N * P / TH; s++) {
        const int n = (threadIdx.x + s * TH) / P;
        const T k = kd[n];
        const T w = pexp(-hs2 * k * k);  // unnormalised both ways: the r2c store applies 1/M
        x[s] = C{x[s].x * w, x[s].y * w};
    }
    reg_fft<T, N, P, Cfg::TPP, +1>(x, lds, twl);
    reg_store<T, N, P, TH>(x, pv, [&](int n, int p, C v) { io[base + n * g.es + p] = v; });
}

// z axis of the separable Gaussian: half-spectrum bins times
// exp(-hs2 k_z^2) scale as the r2c pass stores them.
template <typename T> struct GaussHalfRowStore {
    c2<T>* out;
    int64_t nh;
    const T* kz;
    T hs2, scale;
    __device__ __forceinline__ void store(int64_t row, int k, c2<T> v) const {
        const T w = pexp(-hs2 * kz[k] * kz[k]) * scale;
        out[row * nh + k] = c2<T>{v.x * w, v.y * w};
    }
};

// X pass with a real isotropic weight: forward, times W(k) scale, inverse.
// kind 1: W = knorm^-n + 1 (knorm(0) = 1e-8, model.py:376-381); kind 2:
// W = exp(-k^2 sigma^2 / 2).  The k = 0 bin is set to 0: the loss carries the
// field's mean separately in FP64 and adds W(0) times it back (sm_loss), so
// the rounding left in this bin by the FP32 transforms of the mean-free
// field is not amplified by W(0) (1e8^n + 1 for kind 1).
template <typename T, int N>
__global__ void __launch_bounds__((FftCfg<T, N>::THREADS))
k_x_realfilter(ColGeom g, const c2<T>* __restrict__ tw, c2<T>* __restrict__ io, KTabs<T> kt, int kind, T par, T scale) {
    using Cfg = FftCfg<T, N>;
    using C = c2<T>;
    constexpr int P = Cfg::P, LS = Cfg::LS, TH = Cfg::THREADS;
    __shared__ C lds[Cfg::LDS_ELEMS];
    __shared__ C twl[N];
    const int64_t tile = xcd_tile(blockIdx.x, gridDim.x);
    const int64_t plane = tile / g.tiles_per_plane;
    const int64_t col0 = (tile - plane * g.tiles_per_plane) * P;
    const int64_t base = plane * g.plane_stride + col0;
    const int pv = (int)((g.ncols - col0) < P ? (g.ncols - col0) : P);
    stage_twiddles<T, N, TH>(twl, tw);
    C x[N * P / TH];
    reg_load<T, N, P, TH>(x, pv, [&](int n, int p) { return io[base + n * g.es + p]; });
    fft_from_regs<T, N, P, Cfg::TPP, -1>(x, lds, twl);
    for (int i = threadIdx.x; i < N * P; i += TH) {
        const int n = i / P, p = i - n * P;
        if (p >= pv) continue;
        T w;
        if (kind == 1) {
            w = ppow(kt.knorm2(n, plane, col0 + p), T(-0.5) * par) + T(1);
        } else {
            const T a = kt.k[0][n], b = kt.k[1][plane], c = kt.k[2][col0 + p];
            w = pexp(T(-0.5) * (a * a + b * b + c * c) * par * par);
        }
        w = (n == 0 && plane == 0 && col0 + p == 0) ? T(0) : w * scale;
        const C v = lds[n * LS + p];
        lds[n * LS + p] = C{v.x * w, v.y * w};
    }
    __syncthreads();
    lds_fft<T, N, P, Cfg::TPP, +1, 1, 1, fft_last_ns(N)>(lds, twl);
    fft_to_regs<T, N, P, Cfg::TPP, +1>(x, lds, twl);
    reg_store<T, N, P, TH>(x, pv, [&](int n, int p, C v) { io[base + n * g.es + p] = v; });
}

// ------------------------------------ spectral kernels of the any-length path
// Meshes whose axes are not all powers of two run the x passes unfused: plain
// column passes around these elementwise kernels over the half spectrum
// (index i -> (ix, iy, iz) with iz fastest, rows of pitch hp; the pad columns
// iz >= nh are skipped), which do what the fused x kernels do between their
// transforms.

// out3[d] = -i D_d (cpot alpha G U [+ a G2 S2])
template <typename T, bool HI>
__global__ void __launch_bounds__(256) k_sm_filter_fwd(const c2<T>* __restrict__ U, c2<T>* __restrict__ out3, int64_t hcells, int64_t n1,
                                                       int64_t nh, int64_t hp, KTabs<T> kt, SmFilter<T> fl, T a1e2, Src2<T> s2) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < hcells; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t iz = i % hp, iy = (i / hp) % n1, ix = i / (hp * n1);
        if (iz >= nh) continue;
        const T k2 = kt.knorm2(ix, iy, iz);
        const T G = fl.G(k2);
        const c2<T> u = U[i];
        c2<T> p = c2<T>{u.x * a1e2 * G, u.y * a1e2 * G};
        if constexpr (HI) {
            const c2<T> sv = s2.S[i];
            const T G2 = s2.a * s2.fl.G(k2);
            p = c2<T>{p.x + G2 * sv.x, p.y + G2 * sv.y};
        }
        const T D[3] = {kt.D[0][ix], kt.D[1][iy], kt.D[2][iz]};
#pragma unroll
        for (int d = 0; d < 3; d++) out3[d * hcells + i] = c2<T>{D[d] * p.y, -D[d] * p.x};
    }
}

// Q = sum_d i D_d S_d; FP64 partials (alpha, kh, kl, nu; HI also alpha2,
// kh2, kl2, nu2, gamma2) per workgroup into part[NP * block + j]; out =
// cpot alpha G Q (out may alias S3's first spectrum).
template <typename T, bool HI>
__global__ void __launch_bounds__(256) k_sm_filter_bwd(const c2<T>* S3, const c2<T>* __restrict__ U, c2<T>* out, int64_t hcells, int64_t n1,
                                                       int64_t nh, int64_t hp, int64_t n2, int64_t zdbl, KTabs<T> kt, SmFilter<T> fl, T a1e2, double gkh,
                                                       double gkl, double cpot, double alpha, int alpha_rule, Src2<T> s2,
                                                       double* __restrict__ part) {
    constexpr int NP = HI ? 9 : 4;
    double acc[NP] = {};
    const double apot = cpot * alpha;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < hcells; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t iz = i % hp, iy = (i / hp) % n1, ix = i / (hp * n1);
        if (iz >= nh) continue;
        const T D[3] = {kt.D[0][ix], kt.D[1][iy], kt.D[2][iz]};
        c2<T> q = c2<T>{T(0), T(0)};
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const c2<T> v = S3[d * hcells + i];
            q = c2<T>{q.x - D[d] * v.y, q.y + D[d] * v.x};  // += i D v
        }
        const T G = xbwd_cell<T, HI>(fl, s2, kt.knorm2(ix, iy, iz), q, U[i], i, xbwd_weights(iz, n2, zdbl, alpha_rule), cpot, apot, gkh,
                                     gkl, acc);
        out[i] = c2<T>{a1e2 * G * q.x, a1e2 * G * q.y};
    }
    block_sum<NP, 256>(acc);
    if (threadIdx.x == 0)
        for (int j = 0; j < NP; j++) part[blockIdx.x * NP + j] = acc[j];
}

// S *= scale * W(k) (the weights of k_x_realfilter, k = 0 bin set to 0)
template <typename T>
__global__ void __launch_bounds__(256) k_realfilter_mult(c2<T>* __restrict__ S, int64_t hcells, int64_t n1, int64_t nh, int64_t hp, KTabs<T> kt,
                                                         int kind, T par, T scale) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < hcells; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t iz = i % hp, iy = (i / hp) % n1, ix = i / (hp * n1);
        if (iz >= nh) continue;
        T w;
        if (kind == 1) {
            w = ppow(kt.knorm2(ix, iy, iz), T(-0.5) * par) + T(1);
        } else {
            const T a = kt.k[0][ix], b = kt.k[1][iy], c = kt.k[2][iz];
            w = pexp(T(-0.5) * (a * a + b * b + c * c) * par * par);
        }
        w = i == 0 ? T(0) : w * scale;
        const c2<T> v = S[i];
        S[i] = c2<T>{v.x * w, v.y * w};
    }
}

// --------------------------------------------------------------- gathers
// Pair sweep: each thread gathers for two particles, both particles' position
// loads and then both particles' field loads issued before any is consumed.
// Loads of a missing second particle read particle np - 1; only its stores
// are skipped.  The forces F3i are interleaved (x, y, z) per cell.

// pos_out = pos + readout3(F) * step (step = 1/cell in mesh units, 1 in box units)
template <typename T, typename G>
__global__ void __launch_bounds__(256) k_readout3_update(const T* __restrict__ pos, PosDec<T> pd, int64_t np, Sweep sw, const T* __restrict__ F3i,
                                                         G geo, T s0, T s1, T s2, T* __restrict__ out) {
    LDL_SWEEP2(sw, np, p0, p1) {
        const int64_t pp[2] = {p0 < np ? p0 : np - 1, p1 < np ? p1 : np - 1};
        T x[2][3];
#pragma unroll
        for (int k = 0; k < 2; k++)
#pragma unroll
            for (int d = 0; d < 3; d++) x[k][d] = ld_stream(pos + 3 * pp[k] + d);
        typename G::St st[2];
        T v[2][3][8];
#pragma unroll
        for (int k = 0; k < 2; k++) geo.locate(pd, x[k][0], x[k][1], x[k][2], pp[k], st[k]);
#pragma unroll
        for (int k = 0; k < 2; k++) geo.gather3i(F3i, st[k], v[k]);
#pragma unroll
        for (int k = 0; k < 2; k++) {
            const T o0 = x[k][0] + interp(v[k][0], st[k].f, -1, pd.cellinv) * s0;
            const T o1 = x[k][1] + interp(v[k][1], st[k].f, -1, pd.cellinv) * s1;
            const T o2 = x[k][2] + interp(v[k][2], st[k].f, -1, pd.cellinv) * s2;
            if ((k ? p1 : p0) < np) {
                st_stream(out + 3 * pp[k], o0);
                st_stream(out + 3 * pp[k] + 1, o1);
                st_stream(out + 3 * pp[k] + 2, o2);
            }
        }
    }
}

// Per-component max |v| of the values a workgroup wrote, into gmax (double
// bits, one atomic per component and workgroup): the scale the next
// 3-component paint of these values needs, equal to what its own max scan
// (k_absmax) would find.
template <typename T>
__device__ __forceinline__ void block_absmax3(const T (&mx)[3], unsigned long long* gmax) {
    __shared__ double wm[3][4];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        double v = (double)mx[c];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
        if ((threadIdx.x & 63) == 0) wm[c][threadIdx.x >> 6] = v;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        // the maxima of the many workgroups are alike: an atomic only when this
        // one exceeds the current value keeps the three addresses uncontended
        const int c = threadIdx.x;
        const double v = fmax(fmax(wm[c][0], wm[c][1]), fmax(wm[c][2], wm[c][3]));
        const unsigned long long b = (unsigned long long)__double_as_longlong(v);
        if (v > 0.0 && b > *reinterpret_cast<volatile unsigned long long*>(&gmax[c])) atomicMax(&gmax[c], b);
    }
}

// g_in = g + sum_d g_d grad F_d + fact grad g_rho  (box units); with gmax,
// also the per-component max |g_in|.
template <typename T, typename G>
__global__ void __launch_bounds__(256) k_sm_bwd_gather(const T* __restrict__ pos, PosDec<T> pd, int64_t np, Sweep sw, const T* __restrict__ F3i,
                                                       const T* __restrict__ grho, G geo, T fact, const T* __restrict__ gout,
                                                       T* __restrict__ gin, unsigned long long* __restrict__ gmax) {
    T mx[3] = {T(0), T(0), T(0)};
    LDL_SWEEP2(sw, np, p0, p1) {
        const int64_t pp[2] = {p0 < np ? p0 : np - 1, p1 < np ? p1 : np - 1};
        T x[2][3], g[2][3];
#pragma unroll
        for (int k = 0; k < 2; k++)
#pragma unroll
            for (int d = 0; d < 3; d++) {
                x[k][d] = ld_stream(pos + 3 * pp[k] + d);
                g[k][d] = ld_stream(gout + 3 * pp[k] + d);
            }
        typename G::St st[2];
        T v[2][3][8], vr[2][1][8];
        const T* fr[1] = {grho};
#pragma unroll
        for (int k = 0; k < 2; k++) geo.locate(pd, x[k][0], x[k][1], x[k][2], pp[k], st[k]);
#pragma unroll
        for (int k = 0; k < 2; k++) {
            geo.gather3i(F3i, st[k], v[k]);
            geo.template gather<1>(fr, st[k], vr[k]);
        }
#pragma unroll
        for (int k = 0; k < 2; k++) {
            T acc[3] = {g[k][0], g[k][1], g[k][2]};
            T val, gr[3];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                interp_vg(v[k][d], st[k].f, pd.cellinv, val, gr);
#pragma unroll
                for (int e = 0; e < 3; e++) acc[e] += g[k][d] * gr[e];
            }
            interp_vg(vr[k][0], st[k].f, pd.cellinv, val, gr);
            if ((k ? p1 : p0) < np) {
#pragma unroll
                for (int e = 0; e < 3; e++) {
                    const T o = acc[e] + fact * gr[e];
                    st_stream(gin + 3 * pp[k] + e, o);
                    mx[e] = fmax(mx[e], fabs(o));
                }
            }
        }
    }
    if (gmax) block_absmax3<T>(mx, gmax);
}

// g = fact * grad(interp g_rho); with gmax, also the per-component max |g|.
template <typename T, typename G>
__global__ void __launch_bounds__(256) k_paint_vjp_pos(const T* __restrict__ pos, PosDec<T> pd, int64_t np, Sweep sw, const T* __restrict__ grho,
                                                       G geo, T fact, T* __restrict__ g, unsigned long long* __restrict__ gmax) {
    T mx[3] = {T(0), T(0), T(0)};
    LDL_SWEEP2(sw, np, p0, p1) {
        const int64_t pp[2] = {p0 < np ? p0 : np - 1, p1 < np ? p1 : np - 1};
        T x[2][3];
#pragma unroll
        for (int k = 0; k < 2; k++)
#pragma unroll
            for (int d = 0; d < 3; d++) x[k][d] = ld_stream(pos + 3 * pp[k] + d);
        typename G::St st[2];
        T v[2][1][8];
        const T* fs[1] = {grho};
#pragma unroll
        for (int k = 0; k < 2; k++) geo.locate(pd, x[k][0], x[k][1], x[k][2], pp[k], st[k]);
#pragma unroll
        for (int k = 0; k < 2; k++) geo.template gather<1>(fs, st[k], v[k]);
#pragma unroll
        for (int k = 0; k < 2; k++) {
            T val, gr[3];
            interp_vg(v[k][0], st[k].f, pd.cellinv, val, gr);
            if ((k ? p1 : p0) < np) {
#pragma unroll
                for (int e = 0; e < 3; e++) {
                    const T o = fact * gr[e];
                    st_stream(g + 3 * pp[k] + e, o);
                    mx[e] = fmax(mx[e], fabs(o));
                }
            }
        }
    }
    if (gmax) block_absmax3<T>(mx, gmax);
}

// ------------------------------------------------------------------ loss
// Response of the final density.  sm (model.py:338-344): F = ReLU(w (rho +
// 1e-8)^mu + b); dm (``ident``, model.py:346-347): F = rho; HI
// (model.py:313-334): F = ReLU(w ReLU(rho + zs)^mu - dep + b) with the
// depletion map ``dep`` precomputed; ``gdep`` (optional) receives dL/d dep.
template <typename T> struct RespPar {
    T mu, w, b, index;
    int l1;
    const T* field2;
    int ident;
    int hi;
    T zs;
    const T* dep;
    T* gdep;
};

__device__ __forceinline__ float relu12(float x) { return x > 1e-12f ? x : 1e-12f; }
__device__ __forceinline__ double relu12(double x) { return x > 1e-12 ? x : 1e-12; }

// Forward response value at cell i: y, F, Fi (model.py:344, :358-362).
// PLAIN as for response_adj below.
template <typename T, bool PLAIN = false>
__device__ __forceinline__ void response(const RespPar<T>& rp, int64_t i, T rho, T f2, T& base, T& y, T& F, T& Fi) {
    if (!PLAIN && rp.ident) {
        base = y = F = rho;
    } else {
        const bool hi = !PLAIN && rp.hi;
        base = ppow(hi ? relu12(rho + rp.zs) : rho + T(1e-8), rp.mu);
        y = rp.w * base + rp.b;
        if (hi) y -= rp.dep[i];
        F = relu12(y);
    }
    if constexpr (PLAIN) {
        Fi = F;
    } else {
        Fi = rp.index == T(1) ? F : ppow(F, rp.index);
        Fi = Fi * f2;
    }
}

// Same, also returning x^(mu-1) and ln x (x = rho+1e-8, HI ReLU(rho+zs)) for
// the adjoint.  PLAIN: the caller guarantees the sm response with index 1
// and no second field (plain_response), so those branches compile away; the
// arithmetic, and so every bit, is the general path's.
template <typename T, bool PLAIN = false>
__device__ __forceinline__ void response_adj(const RespPar<T>& rp, int64_t i, T rho, T f2, T& base, T& base1, T& lr, T& y, T& F,
                                             T& Fi) {
    if (!PLAIN && rp.ident) {
        base = y = F = rho;
        base1 = T(1);
        lr = T(0);
    } else {
        const bool hi = !PLAIN && rp.hi;
        pow_pair_log(hi ? relu12(rho + rp.zs) : rho + T(1e-8), rp.mu, base, base1, lr);
        y = rp.w * base + rp.b;
        if (hi) y -= rp.dep[i];
        F = relu12(y);
    }
    if constexpr (PLAIN) {
        Fi = F;  // F^1 times a unit field
    } else {
        Fi = rp.index == T(1) ? F : ppow(F, rp.index);
        Fi = Fi * f2;
    }
}

template <typename T>
__host__ __device__ __forceinline__ bool plain_response(const RespPar<T>& rp) {
    return !rp.ident && !rp.hi && rp.index == T(1) && rp.field2 == nullptr;
}

// d loss / d y, given d loss / d Fi (times field2 and the index power)
template <typename T, bool PLAIN = false>
__device__ __forceinline__ T response_bwd(const RespPar<T>& rp, T gFi, T f2, T y, T F) {
    if constexpr (PLAIN) return y > T(1e-12) ? gFi : T(0);
    T g = gFi * f2;
    if (rp.index != T(1)) g = g * rp.index * ppow(F, rp.index - T(1));
    return (rp.ident || y > T(1e-12)) ? g : T(0);
}

// d loss / d rho from d loss / d y (the mu, w, b partials are zero for
// ident); HI also writes dL/d dep = -gy.
template <typename T, bool PLAIN = false>
__device__ __forceinline__ T response_grho(const RespPar<T>& rp, int64_t i, T gy, T base1, T rho) {
    if constexpr (PLAIN) return gy * rp.w * rp.mu * base1;
    if (rp.ident) return gy;
    if (rp.hi) {
        if (rp.gdep) rp.gdep[i] = -gy;
        return rho + rp.zs > T(1e-12) ? gy * rp.w * rp.mu * base1 : T(0);
    }
    return gy * rp.w * rp.mu * base1;
}


// mask sums -> part (2 per block)
template <typename T>
__global__ void __launch_bounds__(256) k_mask_sums(const T* __restrict__ mask, const T* __restrict__ maskv, int64_t n, double* __restrict__ part) {
    double a[2] = {0.0, 0.0};
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (mask) a[0] += (double)mask[i];
        if (maskv) a[1] += (double)maskv[i];
    }
    block_sum<2, 256>(a);
    if (threadIdx.x == 0) {
        part[blockIdx.x * 2] = a[0];
        part[blockIdx.x * 2 + 1] = a[1];
    }
}

// Loss sums per block: {loss, g_mu, g_w, g_b, loss_v, g_zs, sum of g_ds}
static constexpr int kLossCols = 7;

// Unsmoothed loss and its gradient in one pass (npix from device memory).
template <typename T>
__global__ void __launch_bounds__(256) k_loss_direct(const T* __restrict__ rho, const T* __restrict__ target, const T* __restrict__ mask,
                                                     const T* __restrict__ maskv, RespPar<T> rp, int64_t n, const double* __restrict__ npix,
                                                     T* __restrict__ grho, double* __restrict__ part) {
    const double inv = 1.0 / npix[0];
    double a[kLossCols] = {};
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const T r = rho[i];
        const T f2 = rp.field2 ? rp.field2[i] : T(1);
        T base, base1, lr, y, F, Fi;
        response_adj(rp, i, r, f2, base, base1, lr, y, F, Fi);
        const T diff = Fi - target[i];
        const T ad = fabs(diff);
        const T mk = mask ? mask[i] : T(1);
        const double l = rp.l1 ? (double)ad : (double)ad * (double)ad;
        a[0] += (double)mk * l;
        if (maskv) a[4] += (double)maskv[i] * l;
        if (grho) {
            const T sgn = diff > T(0) ? T(1) : diff < T(0) ? T(-1) : T(0);
            const T gd = (T)((double)mk * inv) * (rp.l1 ? sgn : T(2) * diff);
            const T gy = response_bwd(rp, gd, f2, y, F);
            if (!rp.ident) {
                a[1] += (double)(gy * rp.w * base) * (double)lr;
                a[2] += (double)(gy * base);
                a[3] += (double)gy;
            }
            const T gr = response_grho(rp, i, gy, base1, r);
            a[5] += (double)gr;
            grho[i] = gr;
        }
    }
    block_sum<kLossCols, 256>(a);
    if (threadIdx.x == 0)
        for (int j = 0; j < kLossCols; j++) part[blockIdx.x * kLossCols + j] = a[j];
}

// diff = Fi - target (smoothed-loss path) and per-block FP64 sums of diff
template <typename T, bool PLAIN = false>
__global__ void __launch_bounds__(256) k_loss_diff(const T* __restrict__ rho, const T* __restrict__ target, RespPar<T> rp, int64_t n,
                                                   T* __restrict__ diff, double* __restrict__ dpart) {
    double a[1] = {0.0};
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const T f2 = PLAIN ? T(1) : rp.field2 ? rp.field2[i] : T(1);
        T base, y, F, Fi;
        response<T, PLAIN>(rp, i, rho[i], f2, base, y, F, Fi);
        const T d = Fi - target[i];
        diff[i] = d;
        a[0] += (double)d;
    }
    block_sum<1, 256>(a);
    if (threadIdx.x == 0) dpart[blockIdx.x] = a[0];
}

// rows of N2 reals minus their mesh's mean sum[0] * inv (FP64 sum)
template <typename T> struct ShiftRowLoad {
    const T* in;
    int64_t n2;
    const double* sum;
    double inv;
    __device__ __forceinline__ c2<T> load2(int64_t row, int m) const {
        const T mu = (T)(sum[0] * inv);
        const c2<T> v = *reinterpret_cast<const c2<T>*>(in + row * n2 + 2 * m);
        return c2<T>{v.x - mu, v.y - mu};
    }
};

// The model map alone (JERALDModel.evaluate, model.py:313-344): out = F.
template <typename T>
__global__ void __launch_bounds__(256) k_response(const T* __restrict__ rho, RespPar<T> rp, int64_t n, T* __restrict__ out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        T base, y, F, Fi;
        response(rp, i, rho[i], T(1), base, y, F, Fi);
        out[i] = F;
    }
}

// z c2r epilogue of the smoothed residual: loss partials and g_ds (and the
// FP64 sum of g_ds, the mean its smoothing carries).  dc[0] * dcw is added
// to every value: W(0) times the mean of the field the smoothing took out.
template <typename T> struct LossStore {
    static constexpr int NR = 3;
    const T* mask;
    const T* maskv;
    T* gds;
    const double* npix;
    int64_t n2;
    T scale;
    int l1;
    double* part;
    const double* dc;
    double dcw;
    __device__ __forceinline__ void store2(int64_t row, int m, c2<T> v, double* acc) const {
        const double inv = 1.0 / npix[0];
        const int64_t i0 = row * n2 + 2 * m;
        const T sh = (T)(dc[0] * dcw);
        const T ds[2] = {v.x * scale + sh, v.y * scale + sh};
#pragma unroll
        for (int e = 0; e < 2; e++) {
            const int64_t i = i0 + e;
            const T ad = fabs(ds[e]);
            const double l = l1 ? (double)ad : (double)ad * (double)ad;
            const T mk = mask ? mask[i] : T(1);
            acc[0] += (double)mk * l;
            if (maskv) acc[1] += (double)maskv[i] * l;
            if (gds) {
                const T sgn = ds[e] > T(0) ? T(1) : ds[e] < T(0) ? T(-1) : T(0);
                const T gv = (T)((double)mk * inv) * (l1 ? sgn : T(2) * ds[e]);
                gds[i] = gv;
                acc[2] += (double)gv;
            }
        }
    }
    __device__ __forceinline__ void reduce_out(int64_t b, const double* acc) const {
        part[b * kLossCols + 0] = acc[0];
        part[b * kLossCols + 4] = acc[1];
        part[b * kLossCols + 6] = acc[2];
    }
};

// z c2r epilogue of the smoothed gradient: response backward -> g_rho.
template <typename T> struct RespBwdStore {
    static constexpr int NR = 4;
    const T* rho;
    T* grho;
    RespPar<T> rp;
    int64_t n2;
    T scale;
    double* part;
    const double* dc;  // as LossStore's
    double dcw;
    __device__ __forceinline__ void store2(int64_t row, int m, c2<T> v, double* acc) const {
        const int64_t i0 = row * n2 + 2 * m;
        const T sh = (T)(dc[0] * dcw);
        const T gv[2] = {v.x * scale + sh, v.y * scale + sh};
#pragma unroll
        for (int e = 0; e < 2; e++) {
            const int64_t i = i0 + e;
            const T r = rho[i];
            const T f2 = rp.field2 ? rp.field2[i] : T(1);
            T base, base1, lr, y, F, Fi;
            response_adj(rp, i, r, f2, base, base1, lr, y, F, Fi);
            const T gy = response_bwd(rp, gv[e], f2, y, F);
            if (!rp.ident) {
                acc[0] += (double)(gy * rp.w * base) * (double)lr;
                acc[1] += (double)(gy * base);
                acc[2] += (double)gy;
            }
            const T gr = response_grho(rp, i, gy, base1, r);
            acc[3] += (double)gr;
            grho[i] = gr;
        }
    }
    __device__ __forceinline__ void reduce_out(int64_t b, const double* acc) const {
        part[b * kLossCols + 1] = acc[0];
        part[b * kLossCols + 2] = acc[1];
        part[b * kLossCols + 3] = acc[2];
        part[b * kLossCols + 5] = acc[3];
    }
};

// z c2r epilogue of q = S^2 r for the unmasked L2 loss with a real, even
// smoothing weight W (S = F^-1 W F is symmetric): sum_x (S r)^2 = sum_x r q and
// the residual's gradient is 2 q / npix, so one smoothing with W^2 gives the
// loss partial, and the response backward follows in the same pass.  dc[0] *
// dcw adds W(0)^2 times the residual's mean back (as LossStore's).  The
// residual r = Fi - target is formed again here (the smoothing's line passes
// overwrite their input), with k_loss_diff's arithmetic.
template <typename T, bool PLAIN = false> struct LossSq2Store {
    static constexpr int NR = 5;
    const T* target;
    const T* rho;
    T* grho;
    RespPar<T> rp;
    const double* npix;
    int64_t n2;
    T scale;
    double* part;
    const double* dc;
    double dcw;
    __device__ __forceinline__ void store2(int64_t row, int m, c2<T> v, double* acc) const {
        const int64_t i0 = row * n2 + 2 * m;
        const T sh = (T)(dc[0] * dcw);
        const T q[2] = {v.x * scale + sh, v.y * scale + sh};
        const T ginv = (T)(2.0 / npix[0]);
#pragma unroll
        for (int e = 0; e < 2; e++) {
            const int64_t i = i0 + e;
            const T r = rho[i];
            const T f2 = PLAIN ? T(1) : rp.field2 ? rp.field2[i] : T(1);
            T base, base1, lr, y, F, Fi;
            response_adj<T, PLAIN>(rp, i, r, f2, base, base1, lr, y, F, Fi);
            acc[0] += (double)(Fi - target[i]) * (double)q[e];
            const T gy = response_bwd<T, PLAIN>(rp, q[e] * ginv, f2, y, F);
            if (PLAIN || !rp.ident) {
                acc[1] += (double)(gy * rp.w * base) * (double)lr;
                acc[2] += (double)(gy * base);
                acc[3] += (double)gy;
            }
            const T gr = response_grho<T, PLAIN>(rp, i, gy, base1, r);
            acc[4] += (double)gr;
            grho[i] = gr;
        }
    }
    __device__ __forceinline__ void reduce_out(int64_t b, const double* acc) const {
        part[b * kLossCols + 0] = acc[0];
        part[b * kLossCols + 1] = acc[1];
        part[b * kLossCols + 2] = acc[2];
        part[b * kLossCols + 3] = acc[3];
        part[b * kLossCols + 5] = acc[4];
    }
};

// out = {loss / npix, g_mu, g_w, g_b, loss_v / npix_v[, g_zs]}
__global__ void k_loss_finish(const double* __restrict__ sums, const double* __restrict__ npix, double* __restrict__ out, int nout) {
    if (threadIdx.x == 0) {
        out[0] = sums[0] / npix[0];
        out[1] = sums[1];
        out[2] = sums[2];
        out[3] = sums[3];
        out[4] = npix[1] > 0.0 ? sums[4] / npix[1] : 0.0;
        if (nout > 5) out[5] = sums[5];
    }
}

// ------------------------------------------------------- HI depletion
// dep = ReLU(beta1 S + eta) + beta2 starmap^gamma3 with S = c2r(Sk W), Sk =
// r2c(starmap) / N^3 and W = exp(-k/kli - khi/k) k^nui + xi (model.py:326-332).
// W is xi minus the dm-form kernel at (khi, kli, nui).
template <typename T> struct DepPar {
    T beta1, beta2, gamma3, eta;
};

// out (rows of pitch hp) = Sk W, Sk in the unpadded PMesh.r2c layout
template <typename T>
__global__ void __launch_bounds__(256) k_hi_wmult(const c2<T>* __restrict__ Sk, c2<T>* __restrict__ out, int64_t hcells, int64_t n1,
                                                  int64_t nh, int64_t hp, KTabs<T> kt, SmFilter<T> fl, T xi) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < hcells; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t iz = i % nh, iy = (i / nh) % n1, ix = i / (nh * n1);
        const T w = xi - fl.G(kt.knorm2(ix, iy, iz));
        const c2<T> v = Sk[i];
        out[(ix * n1 + iy) * hp + iz] = c2<T>{v.x * w, v.y * w};
    }
}

// z c2r epilogue: S (kept for the backward) and dep.
template <typename T> struct DepStore {
    static constexpr int NR = 0;
    const T* star;
    T* S;
    T* dep;
    int64_t n2;
    DepPar<T> dp;
    __device__ __forceinline__ void reduce_out(int64_t, const double*) const {}
    __device__ __forceinline__ void store2(int64_t row, int m, c2<T> v, double*) const {
        const int64_t i = row * n2 + 2 * m;
        const c2<T> x = *reinterpret_cast<const c2<T>*>(star + i);
        *reinterpret_cast<c2<T>*>(S + i) = v;
        const T d0 = relu12(dp.beta1 * v.x + dp.eta) + dp.beta2 * spow(x.x, dp.gamma3);
        const T d1 = relu12(dp.beta1 * v.y + dp.eta) + dp.beta2 * spow(x.y, dp.gamma3);
        *reinterpret_cast<c2<T>*>(dep + i) = c2<T>{d0, d1};
    }
};

// Depletion backward, elementwise: g_S = beta1 gdep [beta1 S + eta > 1e-12]
// and the partials {beta1, beta2, gamma3, eta} per block.
template <typename T>
__global__ void __launch_bounds__(256) k_hi_dep_bwd(const T* __restrict__ star, const T* __restrict__ S, const T* __restrict__ gdep, int64_t n,
                                                    DepPar<T> dp, T* __restrict__ gS, double* __restrict__ part) {
    double a[4] = {};
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const T g = gdep[i], sv = S[i], x = star[i];
        a[1] += (double)g * (double)spow(x, dp.gamma3);
        a[2] += (double)g * (double)dp.beta2 * (double)spow_log(x, dp.gamma3);
        const T gA = dp.beta1 * sv + dp.eta > T(1e-12) ? g : T(0);
        a[3] += (double)gA;
        a[0] += (double)gA * (double)sv;
        gS[i] = dp.beta1 * gA;
    }
    block_sum<4, 256>(a);
    if (threadIdx.x == 0)
        for (int j = 0; j < 4; j++) part[blockIdx.x * 4 + j] = a[j];
}

// dL/d(khi, kli, nui, xi) = sum_H c Re(conj(ghat) Sk) dW/dtheta, ghat the
// unnormalised r2c of g_S: dW/dkhi = G/k, dW/dkli = -G k/kli^2,
// dW/dnui = -G ln k, dW/dxi = 1 (G the dm-form kernel).  The filter passes
// through c2f (model.py:329-330), so the reference rule weights these bins
// like the layer kernels' parameters (xbwd_weights' cg).
template <typename T>
__global__ void __launch_bounds__(256) k_hi_wgrad(const c2<T>* __restrict__ ghat, const c2<T>* __restrict__ Sk, int64_t hcells, int64_t n1,
                                                  int64_t nh, int64_t hp, int64_t n2, int64_t zdbl, int rule, KTabs<T> kt, SmFilter<T> fl,
                                                  double inv_kli2, double* __restrict__ part) {
    double a[4] = {};
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < hcells; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t iz = i % nh, iy = (i / nh) % n1, ix = i / (nh * n1);
        T G, th, tl, lnk;
        fl.eval(kt.knorm2(ix, iy, iz), G, th, tl, lnk);
        const c2<T> g = ghat[(ix * n1 + iy) * hp + iz], sk = Sk[i];
        const double re = xbwd_weights(iz, n2, zdbl, rule).cg * ((double)g.x * (double)sk.x + (double)g.y * (double)sk.y);
        const double Gd = (double)G;
        a[0] += re * Gd * (double)th;
        a[1] -= re * Gd * (double)tl * inv_kli2;
        a[2] -= re * Gd * (double)lnk;
        a[3] += re;
    }
    block_sum<4, 256>(a);
    if (threadIdx.x == 0)
        for (int j = 0; j < 4; j++) part[blockIdx.x * 4 + j] = a[j];
}

// ------------------------------------------------------------ host side
// out (rows of pitch hp) = scale r2c(rows of ``load``): z pass, y pass, x
// pass (plain kernels).
template <typename T, class Load>
static int r2c_full(const Mesh& m, const c2<T>* tw, Load load, c2<T>* out, int64_t hp, T scale, hipStream_t st) {
    const int64_t nh = m.nh();
    int rc;
    if ((rc = zpass_r2c<T>(m.n0 * m.n1, m.n2, tw, load, HalfRowStore<T>{out, hp, T(1)}, st))) return rc;
    if ((rc = colpass<T, -1>(m.n1, m.n0, nh, m.n1 * hp, hp, tw, PlainColOp<T>{out, out, T(1)}, st))) return rc;
    return colpass<T, -1>(m.n0, m.n1, nh, hp, m.n1 * hp, tw, PlainColOp<T>{out, out, scale}, st);
}

// Second source of an HI layer from prm[5..9] = [alpha2, gamma2, kh2, kl2, nu2].
template <typename T>
static Src2<T> hi_src2(const double* prm, const c2<T>* S, const c2<T>* Tl) {
    Src2<T> s2{};
    s2.S = S;
    s2.Tl = Tl;
    s2.fl = layer_filter<T>(prm + 5, 0);
    s2.ad = -1e-3 * prm[5];
    s2.a = (T)s2.ad;
    s2.cs = -1e-3;
    filter_grad_coeffs(prm + 5, 0, &s2.gkh, &s2.gkl);
    return s2;
}
template <typename T>
static int check_prm(const double* prm, int n) {
    if (!prm) return fail(LDL_EINVAL, "parameter array is NULL");
    for (int i = 0; i < n; i++)
        if (!std::isfinite(prm[i])) return fail(LDL_EINVAL, "parameters must be finite");
    return LDL_OK;
}

// Layer kinds: 0 sm, 1 dm, 2 HI (prm has 10 entries; star and S2buf used).
enum { kKindSm = 0, kKindDm = 1, kKindHi = 2 };

template <typename T>
static int sm_layer_fwd(int kind, const T* pos, const ldl_pos_desc* pdesc, int64_t np, const int64_t nmesh[3], const double* prm,
                        double fact, int32_t pflags, const T* star, T* pos_out, T* rho, T* Ubuf, T* S2buf, T* F, void* bins, void* ws,
                        size_t wsb, hipStream_t st) {
    Mesh m;
    int rc = check_fft_mesh(nmesh, &m);
    if (rc) return rc;
    const bool hi = kind == kKindHi;
    if ((rc = check_prm<T>(prm, hi ? 10 : 5))) return rc;
    if (!pos || !rho || !Ubuf || !F) return fail(LDL_EINVAL, "NULL array argument");
    if (hi && (!star || !S2buf)) return fail(LDL_EINVAL, "NULL array argument");
    PosDec<T> pd;
    if ((rc = make_posdec<T>(pdesc, nmesh, np, &pd))) return rc;
    const Sweep sw = make_sweep(pd, np);
    const SmWs w = sm_ws(m, sizeof(T), np, pflags);
    if (!ws || wsb < w.total) return fail(LDL_EINVAL, "sm workspace too small (see ldl_sm_workspace_bytes)");
    const SmPtrs<T> P = sm_ptrs<T>(ws, w);
    const c2<T>* tw;
    if ((rc = twiddles<T>(&tw))) return rc;
    // 1. density (with LDL_PAINT_BINS_VALID from the sort an earlier forward of
    // these positions wrote into bins)
    const bool tiled = !(pflags & (LDL_PAINT_FLOAT_ATOMIC | LDL_PAINT_DIRECT));
    const bool bins_in = tiled && bins && (pflags & LDL_PAINT_BINS_VALID);
    if ((rc = paint_impl<T>(pos, pdesc, np, nullptr, 1.0, 1, nmesh, fact, pflags & ~LDL_PAINT_BINS_VALID, rho, P.paint, w.paint, st,
                            tiled && !bins_in ? reinterpret_cast<unsigned*>(bins) : nullptr,
                            bins_in ? reinterpret_cast<const unsigned*>(bins) : nullptr)))
        return rc;
    const KTabs<T> kt = build_tables<T>(P.tabs, m, pdesc->box, st);
    // half spectra: nh columns in rows of pitch hp, hc apart
    const int64_t nrows = m.n0 * m.n1, nh = m.nh(), hp = spec_pitch<T>(m), hc = spec_cells<T>(m);
    c2<T>* U = reinterpret_cast<c2<T>*>(Ubuf);
    // 2. z pass with the gamma power, y pass
    if ((rc = zpass_r2c<T>(nrows, m.n2, tw, PowRowLoad<T>{rho, m.n2, (T)prm[1]}, HalfRowStore<T>{U, hp, T(1)}, st))) return rc;
    if ((rc = colpass<T, -1>(m.n1, m.n0, nh, m.n1 * hp, hp, tw, PlainColOp<T>{U, U, T(1)}, st))) return rc;
    const T invM = (T)(1.0 / (double)m.cells());
    // HI: second source S2 = r2c(starmap^gamma2) / N^3 (model.py:228)
    Src2<T> s2{};
    if (hi) {
        c2<T>* S2 = reinterpret_cast<c2<T>*>(S2buf);
        if ((rc = r2c_full<T>(m, tw, StarRowLoad<T>{star, m.n2, (T)prm[6], 0}, S2, hp, invM, st))) return rc;
        s2 = hi_src2<T>(prm, S2, nullptr);
    }
    // 3. x pass + filter + inverse x pass (and the x difference)
    const SmFilter<T> fl = layer_filter<T>(prm, kind == kKindDm);
    const T a1e2 = (T)((hi ? 1e-3 : 1e-2) * prm[0]);
    if (fft_mesh_pow2(m)) {
        dispatch_n(m.n0, [&](auto N) {
            constexpr int NN = decltype(N)::value;
            ColGeom g = col_geom<T, NN>(m.n1, nh, hp, m.n1 * hp);
            constexpr int XTH = FftCfg<T, NN>::THREADS;
            KTimer _t("sm_x_fwd", st);
            if (hi)
                hipLaunchKernelGGL((k_sm_x_fwd<T, NN, true>), dim3((unsigned)g.ntiles), dim3(XTH), 0, st, g, tw, U, P.spec3, kt,
                                   fl, a1e2, invM, s2);
            else
                hipLaunchKernelGGL((k_sm_x_fwd<T, NN, false>), dim3((unsigned)g.ntiles), dim3(XTH), 0, st, g, tw, U, P.spec3,
                                   kt, fl, a1e2, invM, s2);
        });
    } else {
        if ((rc = colpass<T, -1>(m.n0, m.n1, nh, hp, m.n1 * hp, tw, PlainColOp<T>{U, U, invM}, st))) return rc;
        const unsigned gf = grid_for(hc, 256, 1 << 16);
        if (hi)
            hipLaunchKernelGGL((k_sm_filter_fwd<T, true>), dim3(gf), dim3(256), 0, st, U, P.spec3, hc, m.n1, nh, hp, kt, fl, a1e2, s2);
        else
            hipLaunchKernelGGL((k_sm_filter_fwd<T, false>), dim3(gf), dim3(256), 0, st, U, P.spec3, hc, m.n1, nh, hp, kt, fl, a1e2, s2);
        for (int d = 0; d < 3; d++)
            if ((rc = colpass<T, +1>(m.n0, m.n1, nh, hp, m.n1 * hp, tw, PlainColOp<T>{P.spec3 + d * hc, P.spec3 + d * hc, T(1)}, st))) return rc;
    }
    // 4. inverse y passes (3 force spectra as 3*N0 planes) and z c2r into F
    if (fft_mesh_pow2(m)) {  // phi = inverse transform of P (x pass output in slot 0), F = -grad6(phi)
        T* phi = P.mesh3;
        if ((rc = colpass<T, +1>(m.n1, m.n0, nh, m.n1 * hp, hp, tw, PlainColOp<T>{P.spec3, P.spec3, T(1)}, st))) return rc;
        if ((rc = zpass_c2r<T>(nrows, m.n2, tw, HalfRowLoad<T>{P.spec3, hp}, RealRowStore<T>{phi, m.n2, T(1)}, st, nullptr)))
            return rc;
        const StGeom sg = st_geom<T>(m);
        KTimer _t("grad6", st);
        hipLaunchKernelGGL(k_grad6<T>, dim3((unsigned)(sg.chunks * sg.tiles_y * sg.tiles_z)), dim3(4 * st_tz<T>()), 0, st, phi, F, sg,
                           (T)((double)m.n0 / (60.0 * pdesc->box[0])), (T)((double)m.n1 / (60.0 * pdesc->box[1])),
                           (T)((double)m.n2 / (60.0 * pdesc->box[2])));
    } else {
        if ((rc = colpass<T, +1>(m.n1, 3 * m.n0, nh, m.n1 * hp, hp, tw, PlainColOp<T>{P.spec3, P.spec3, T(1)}, st, "colpass3"))) return rc;
        if ((rc = zpass_c2r<T>(3 * nrows, m.n2, tw, Tri3RowLoad<T>{P.spec3, P.spec3 + hc, P.spec3 + 2 * hc, hp},
                               Tri3Store3<T>{F, m.n2}, st, nullptr, "zpass_c2r3")))
            return rc;
    }
    // 5. readout + position update (skipped without pos_out: the backward's
    // recomputation of a layer needs only the saved tensors)
    T s[3];
    for (int d = 0; d < 3; d++) s[d] = pdesc->kind == LDL_POS_LATTICE ? pd.cellinv[d] : T(1);
    if (np > 0 && pos_out)
        {
            KTimer _t("readout3_update", st);
            with_geom3i<T>(m, pd, [&](auto geo) {
                hipLaunchKernelGGL((k_readout3_update<T, decltype(geo)>), dim3(sweep2_grid(sw)), dim3(kBlk), 0, st, pos, pd, np, sw, F, geo,
                                   s[0], s[1], s[2], pos_out);
            });
        }
    LDL_LAUNCH_CHECK();
    return LDL_OK;
}

template <typename T>
static int sm_layer_bwd(int kind, const T* pos, const ldl_pos_desc* pdesc, int64_t np, const int64_t nmesh[3], const double* prm,
                        double fact, int32_t alpha_rule, int32_t pflags, const T* star, const T* rho, const T* Ubuf, const T* S2buf,
                        const T* F, const void* bins, const T* gout, T* gin, double* grad, double* gmaxd, void* ws, size_t wsb,
                        hipStream_t st) {
    auto* gmax = reinterpret_cast<unsigned long long*>(gmaxd);
    Mesh m;
    int rc = check_fft_mesh(nmesh, &m);
    if (rc) return rc;
    const bool hi = kind == kKindHi;
    if ((rc = check_prm<T>(prm, hi ? 10 : 5))) return rc;
    if (!pos || !rho || !Ubuf || !F || !gout || !gin || !grad) return fail(LDL_EINVAL, "NULL array argument");
    if (hi && (!star || !S2buf)) return fail(LDL_EINVAL, "NULL array argument");
    PosDec<T> pd;
    if ((rc = make_posdec<T>(pdesc, nmesh, np, &pd))) return rc;
    const Sweep sw = make_sweep(pd, np);
    const SmWs w = sm_ws(m, sizeof(T), np, pflags);
    if (!ws || wsb < w.total) return fail(LDL_EINVAL, "sm workspace too small (see ldl_sm_workspace_bytes)");
    const SmPtrs<T> P = sm_ptrs<T>(ws, w);
    const c2<T>* tw;
    if ((rc = twiddles<T>(&tw))) return rc;
    const int64_t nrows = m.n0 * m.n1, nh = m.nh(), hp = spec_pitch<T>(m), hc = spec_cells<T>(m), cells = m.cells();
    // 1. gF_d = paint(X, g_d)  (3 meshes in mesh3)
    const bool reuse = bins && !(pflags & (LDL_PAINT_FLOAT_ATOMIC | LDL_PAINT_DIRECT));
    // (the forward density rho = fact * sum of CIC weights bounds the cells' sums)
    if ((rc = paint_impl<T>(pos, pdesc, np, gout, 0.0, 3, nmesh, 1.0, pflags, P.mesh3, P.paint, w.paint, st, nullptr,
                            reuse ? reinterpret_cast<const unsigned*>(bins) : nullptr, rho, 1.0 / fact, gmax)))
        return rc;
    const KTabs<T> kt = build_tables<T>(P.tabs, m, pdesc->box, st);
    // 2. unnormalised r2c: power-of-two meshes transform the real-space
    // divergence of the three meshes (one spectrum, slot 0); others the three
    // meshes (z and y passes; the filter kernel applies the i D_d)
    const bool p2 = fft_mesh_pow2(m);
    if (p2) {
        T* dv = reinterpret_cast<T*>(P.spec3 + 2 * hc);  // a real mesh in the free third spectrum
        const StGeom sg = st_geom<T>(m);
        {
            KTimer _t("div6", st);
            hipLaunchKernelGGL(k_div6<T>, dim3((unsigned)(sg.chunks * sg.tiles_y * sg.tiles_z)), dim3(4 * st_tz<T>()), 0, st, P.mesh3, dv, sg,
                               (T)((double)m.n0 / (60.0 * pdesc->box[0])), (T)((double)m.n1 / (60.0 * pdesc->box[1])),
                               (T)((double)m.n2 / (60.0 * pdesc->box[2])));
        }
        if ((rc = zpass_r2c<T>(nrows, m.n2, tw, RealRowLoad<T>{dv, m.n2}, HalfRowStore<T>{P.spec3, hp, T(1)}, st))) return rc;
        if ((rc = colpass<T, -1>(m.n1, m.n0, nh, m.n1 * hp, hp, tw, PlainColOp<T>{P.spec3, P.spec3, T(1)}, st))) return rc;
    } else {
        if ((rc = zpass_r2c<T>(3 * nrows, m.n2, tw, RealRowLoad<T>{P.mesh3, m.n2}, HalfRowStore<T>{P.spec3, hp, T(1)}, st,
                               "zpass_r2c3")))
            return rc;
        if ((rc = colpass<T, -1>(m.n1, 3 * m.n0, nh, m.n1 * hp, hp, tw, PlainColOp<T>{P.spec3, P.spec3, T(1)}, st, "colpass3"))) return rc;
    }
    // HI: T2 = r2c(starmap^gamma2 ln starmap) / N^3 for the gamma2 partial, in
    // the mesh scratch (free after the z pass above; sized to hold it)
    Src2<T> s2{};
    if (hi) {
        c2<T>* T2 = reinterpret_cast<c2<T>*>(P.mesh3);
        if ((rc = r2c_full<T>(m, tw, StarRowLoad<T>{star, m.n2, (T)prm[6], 1}, T2, hp, (T)(1.0 / (double)cells), st))) return rc;
        s2 = hi_src2<T>(prm, reinterpret_cast<const c2<T>*>(S2buf), T2);
    }
    // 3. x pass: Q, parameter partials, inverse x of cpot alpha G Q
    const SmFilter<T> fl = layer_filter<T>(prm, kind == kKindDm);
    double gkh, gkl;
    filter_grad_coeffs(prm, kind == kKindDm, &gkh, &gkl);
    const double cpot = hi ? 1e-3 : 1e-2;
    const T a1e2 = (T)(cpot * prm[0]);
    const c2<T>* U = reinterpret_cast<const c2<T>*>(Ubuf);
    const int64_t zdbl = std::min<int64_t>(m.n1 / 2, nh);  // end of c2f's doubled z bins (reference rule)
    c2<T>* S = P.spec3;  // result reuses the first spectrum (read before written)
    int64_t ntx = 0;
    if (fft_mesh_pow2(m)) {
        dispatch_n(m.n0, [&](auto N) {
            constexpr int NN = decltype(N)::value;
            ColGeom g = col_geom<T, NN, xbwd_pencils<T, NN>()>(m.n1, nh, hp, m.n1 * hp);
            constexpr int XTH = xbwd_pencils<T, NN>() * FftCfg<T, NN>::TPP;
            ntx = g.ntiles;
            KTimer _t("sm_x_bwd", st);
            if (hi)
                hipLaunchKernelGGL((k_sm_x_bwd<T, NN, true>), dim3((unsigned)g.ntiles), dim3(XTH), 0, st, g, P.spec3, tw, U, S,
                                   m.n2, zdbl, kt, fl, a1e2, gkh, gkl, cpot, prm[0], (int)alpha_rule, s2, P.part);
            else
                hipLaunchKernelGGL((k_sm_x_bwd<T, NN, false>), dim3((unsigned)g.ntiles), dim3(XTH), 0, st, g, P.spec3, tw, U,
                                   S, m.n2, zdbl, kt, fl, a1e2, gkh, gkl, cpot, prm[0], (int)alpha_rule, s2, P.part);
        });
    } else {
        for (int d = 0; d < 3; d++)
            if ((rc = colpass<T, -1>(m.n0, m.n1, nh, hp, m.n1 * hp, tw, PlainColOp<T>{P.spec3 + d * hc, P.spec3 + d * hc, T(1)}, st))) return rc;
        const unsigned gf = grid_for(hc, 256, 4096);
        ntx = gf;
        if (hi)
            hipLaunchKernelGGL((k_sm_filter_bwd<T, true>), dim3(gf), dim3(256), 0, st, P.spec3, U, S, hc, m.n1, nh, hp, m.n2, zdbl, kt, fl, a1e2,
                               gkh, gkl, cpot, prm[0], (int)alpha_rule, s2, P.part);
        else
            hipLaunchKernelGGL((k_sm_filter_bwd<T, false>), dim3(gf), dim3(256), 0, st, P.spec3, U, S, hc, m.n1, nh, hp, m.n2, zdbl, kt, fl, a1e2,
                               gkh, gkl, cpot, prm[0], (int)alpha_rule, s2, P.part);
        if ((rc = colpass<T, +1>(m.n0, m.n1, nh, hp, m.n1 * hp, tw, PlainColOp<T>{S, S, T(1)}, st))) return rc;
    }
    // partial columns -> parameter slots: alpha, kh, kl, nu (HI: then alpha2,
    // kh2, kl2, nu2, gamma2)
    const RedMap rx = hi ? RedMap{9, {0, 2, 3, 4, 5, 7, 8, 9, 6}} : RedMap{4, {0, 2, 3, 4}};
    hipLaunchKernelGGL(k_reduce_cols, dim3(1), dim3(256), 0, st, P.part, ntx, rx.n, rx, grad, 0);
    // 4. inverse y, z c2r with the gamma-power adjoint -> g_rho (mesh3[0])
    if ((rc = colpass<T, +1>(m.n1, m.n0, nh, m.n1 * hp, hp, tw, PlainColOp<T>{S, S, T(1)}, st))) return rc;
    T* grho = P.mesh3;
    int64_t nzb = 0;
    if ((rc = zpass_c2r<T>(nrows, m.n2, tw, HalfRowLoad<T>{S, hp},
                           GammaBwdStore<T>{rho, grho, m.n2, (T)(1.0 / (double)cells), (T)prm[1], P.part}, st, &nzb,
                           "zpass_c2r_gamma")))
        return rc;
    RedMap rz{1, {1}};
    hipLaunchKernelGGL(k_reduce_cols, dim3(1), dim3(256), 0, st, P.part, nzb, 1, rz, grad, 0);
    // 5. position adjoint
    if (np > 0)
        {
            if (gmax) hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, st, gmax, (int64_t)3, 0ull);  // the paint has read it
            KTimer _t("sm_bwd_gather", st);
            with_geom3i<T>(m, pd, [&](auto geo) {
                hipLaunchKernelGGL((k_sm_bwd_gather<T, decltype(geo)>), dim3(sweep2_grid(sw)), dim3(kBlk), 0, st, pos, pd, np, sw, F, grho, geo,
                                   (T)fact, gout, gin, gmax);
            });
        }
    LDL_LAUNCH_CHECK();
    return LDL_OK;
}

// kind: 0 sm (prm = [mu, w, b]), 1 dm (prm unused), 2 HI (prm = [mu, w, b,
// zs], dep required, gdep optional); out has 5 entries, 6 for HI (+ g_zs).
template <typename T>
static int sm_response(int kind, const T* rho, const T* dep, const int64_t nmesh[3], const double* prm, T* out, hipStream_t st) {
    Mesh m;
    int rc = check_mesh(nmesh, &m);
    if (rc) return rc;
    if ((rc = check_prm<T>(prm, kind == kKindHi ? 4 : 3))) return rc;
    if (!rho || !out || (kind == kKindHi && !dep)) return fail(LDL_EINVAL, "NULL array argument");
    RespPar<T> rp{};
    rp.index = T(1);
    rp.mu = (T)prm[0];
    rp.w = (T)prm[1];
    rp.b = (T)prm[2];
    rp.hi = kind == kKindHi;
    rp.zs = rp.hi ? (T)prm[3] : T(0);
    rp.dep = dep;
    hipLaunchKernelGGL(k_response<T>, dim3(grid_for(m.cells(), 256, 4096)), dim3(256), 0, st, rho, rp, m.cells(), out);
    LDL_LAUNCH_CHECK();
    return LDL_OK;
}

template <typename T>
static int sm_loss(int kind, const T* rho, const T* target, const T* mask, const T* maskv, const T* field2, const T* dep, const int64_t nmesh[3],
                   const double box[3], const double* prm, double index, int32_t p_norm, int32_t smooth_kind, double smooth_par,
                   double* out, T* grho, T* gdep, void* ws, size_t wsb, hipStream_t st) {
    Mesh m;
    int rc = smooth_kind ? check_fft_mesh(nmesh, &m) : check_mesh(nmesh, &m);
    if (rc) return rc;
    if (kind != kKindDm && (rc = check_prm<T>(prm, kind == kKindHi ? 4 : 3))) return rc;
    if (!rho || !target || !out) return fail(LDL_EINVAL, "NULL array argument");
    if (kind == kKindHi && !dep) return fail(LDL_EINVAL, "NULL array argument");
    if (p_norm != 1 && p_norm != 2) return fail(LDL_EINVAL, "p_norm must be 1 or 2");
    if (smooth_kind < 0 || smooth_kind > 2) return fail(LDL_EINVAL, "smooth_kind must be 0, 1 or 2");
    if (smooth_kind && (!box || !(box[0] > 0 && box[1] > 0 && box[2] > 0))) return fail(LDL_EINVAL, "box must be positive");
    const SmWs w = sm_ws(m, sizeof(T), 0, LDL_PAINT_FLOAT_ATOMIC);  // no paint in the loss: no paint region
    if (!ws || wsb < w.total) return fail(LDL_EINVAL, "sm workspace too small (see ldl_sm_workspace_bytes)");
    const SmPtrs<T> P = sm_ptrs<T>(ws, w);
    const int64_t cells = m.cells();
    // part layout: [0, 16) npix scratch and sums; block partials from 64
    double* npix = P.part;          // {npix, npix_v}
    double* sums = P.part + 8;      // kLossCols sums
    double* bpart = P.part + 64;
    const unsigned gb = 2048;
    hipLaunchKernelGGL(k_mask_sums<T>, dim3(gb), dim3(256), 0, st, mask, maskv, cells, bpart);
    RedMap rn{2, {0, 1}};
    hipLaunchKernelGGL(k_reduce_cols, dim3(1), dim3(256), 0, st, bpart, (int64_t)gb, 2, rn, npix, 1);
    if (!mask) {  // npix = number of cells
        const double nc = (double)cells;
        hipLaunchKernelGGL(k_fill<double>, dim3(1), dim3(1), 0, st, npix, (int64_t)1, nc);
    }
    RespPar<T> rp{};
    rp.index = (T)index;
    rp.l1 = p_norm == 1;
    rp.field2 = field2;
    rp.ident = kind == kKindDm;
    if (kind != kKindDm) {
        rp.mu = (T)prm[0];
        rp.w = (T)prm[1];
        rp.b = (T)prm[2];
    }
    if (kind == kKindHi) {
        rp.hi = 1;
        rp.zs = (T)prm[3];
        rp.dep = dep;
        rp.gdep = grho ? gdep : nullptr;
    }
    RedMap r6{kLossCols, {0, 1, 2, 3, 4, 5, 6}};
    if (smooth_kind == 0) {
        { KTimer _t("loss", st); hipLaunchKernelGGL(k_loss_direct<T>, dim3(gb), dim3(256), 0, st, rho, target, mask, maskv, rp, cells, npix, grho, bpart); }
        hipLaunchKernelGGL(k_reduce_cols, dim3(1), dim3(256), 0, st, bpart, (int64_t)gb, kLossCols, r6, sums, 1);
    } else {
        const c2<T>* tw;
        if ((rc = twiddles<T>(&tw))) return rc;
        const KTabs<T> kt = build_tables<T>(P.tabs, m, box, st);
        const int64_t nrows = m.n0 * m.n1, nh = m.nh(), hp = spec_pitch<T>(m);
        T* diff = P.mesh3;
        T* gds = P.mesh3 + cells;
        c2<T>* S = P.spec3;
        const T invM = (T)(1.0 / (double)cells);
        // The FP32 transforms lose the mean of a field whose rms is much larger
        // than it (a residual near convergence): its DC bin comes out with an
        // error relative to sum |x|, not to sum x, and dL/db is that bin.  So
        // the smoothing works on the mean-free field and the field's FP64 sum
        // (dsum for the residual, gsum for g_ds) is added back as W(0) mean
        // by the c2r epilogues.
        double* dsum = P.part + 16;
        double* gsum = P.part + 17;
        const double invMd = 1.0 / (double)cells;
        const double W0 = smooth_kind == 1 ? std::pow(1e-8, -smooth_par) + 1.0 : 1.0;  // W at knorm(0) = 1e-8
        auto smooth = [&](T* src, const double* msum, double par) -> int {  // S = W(k) r2c(src - mean) / M, inverse x and y passes
            int rc2;
            if (smooth_kind == 2 && fft_mesh_pow2(m)) {
                // separable Gaussian: y and x line filters in place on the real
                // mesh (as complex pairs along z), then the z factor and 1/M as
                // the r2c pass stores the half spectrum
                c2<T>* sc = reinterpret_cast<c2<T>*>(src);
                const int64_t n2c = m.n2 / 2;
                const T hs2 = (T)(0.5 * par * par);
                dispatch_n(m.n1, [&](auto N) {
                    constexpr int NN = decltype(N)::value;
                    const ColGeom g = col_geom<T, NN>(m.n0, n2c, m.n1 * n2c, n2c);
                    KTimer _t("gauss_line", st);
                    hipLaunchKernelGGL((k_gauss_line<T, NN>), dim3((unsigned)g.ntiles), dim3(FftCfg<T, NN>::THREADS), 0, st, g, tw, sc, kt.k[1], hs2,
                                       msum, invMd);
                });
                dispatch_n(m.n0, [&](auto N) {
                    constexpr int NN = decltype(N)::value;
                    const ColGeom g = col_geom<T, NN>(m.n1, n2c, n2c, m.n1 * n2c);
                    KTimer _t("gauss_line", st);
                    hipLaunchKernelGGL((k_gauss_line<T, NN>), dim3((unsigned)g.ntiles), dim3(FftCfg<T, NN>::THREADS), 0, st, g, tw, sc, kt.k[0], hs2,
                                       (const double*)nullptr, 0.0);
                });
                return zpass_r2c<T>(nrows, m.n2, tw, RealRowLoad<T>{src, m.n2}, GaussHalfRowStore<T>{S, hp, kt.k[2], hs2, invM}, st);
            }
            if ((rc2 = zpass_r2c<T>(nrows, m.n2, tw, ShiftRowLoad<T>{src, m.n2, msum, invMd}, HalfRowStore<T>{S, hp, T(1)}, st))) return rc2;
            if ((rc2 = colpass<T, -1>(m.n1, m.n0, nh, m.n1 * hp, hp, tw, PlainColOp<T>{S, S, T(1)}, st))) return rc2;
            if (fft_mesh_pow2(m)) {
                dispatch_n(m.n0, [&](auto N) {
                    constexpr int NN = decltype(N)::value;
                    ColGeom g = col_geom<T, NN>(m.n1, nh, hp, m.n1 * hp);
                    KTimer _t("x_realfilter", st);
                    hipLaunchKernelGGL((k_x_realfilter<T, NN>), dim3((unsigned)g.ntiles), dim3(FftCfg<T, NN>::THREADS), 0, st, g, tw, S, kt,
                                       (int)smooth_kind, (T)par, invM);
                });
            } else {
                const int64_t hc = spec_cells<T>(m);
                if ((rc2 = colpass<T, -1>(m.n0, m.n1, nh, hp, m.n1 * hp, tw, PlainColOp<T>{S, S, T(1)}, st))) return rc2;
                hipLaunchKernelGGL(k_realfilter_mult<T>, dim3(grid_for(hc, 256, 1 << 16)), dim3(256), 0, st, S, hc, m.n1, nh, hp, kt, (int)smooth_kind,
                                   (T)par, invM);
                if ((rc2 = colpass<T, +1>(m.n0, m.n1, nh, hp, m.n1 * hp, tw, PlainColOp<T>{S, S, T(1)}, st))) return rc2;
            }
            return colpass<T, +1>(m.n1, m.n0, nh, m.n1 * hp, hp, tw, PlainColOp<T>{S, S, T(1)}, st);
        };
        const int64_t nzb = zpass_c2r_blocks<T>(nrows, m.n2);
        // every block writes the loss columns; the gradient columns stay zero without g_rho
        hipLaunchKernelGGL(k_fill<double>, dim3(grid_for(nzb * kLossCols, 256)), dim3(256), 0, st, bpart, nzb * kLossCols, 0.0);
        double* dpart = bpart + nzb * kLossCols;
        if (plain_response(rp))
            hipLaunchKernelGGL((k_loss_diff<T, true>), dim3(gb), dim3(256), 0, st, rho, target, rp, cells, diff, dpart);
        else
            hipLaunchKernelGGL((k_loss_diff<T, false>), dim3(gb), dim3(256), 0, st, rho, target, rp, cells, diff, dpart);
        const RedMap r1{1, {0}};
        hipLaunchKernelGGL(k_reduce_cols, dim3(1), dim3(256), 0, st, dpart, (int64_t)gb, 1, r1, dsum, 1);
        if (smooth_kind == 2 && p_norm == 2 && !mask && !maskv && grho) {
            // unmasked L2 with the Gaussian: one smoothing with W^2 = the
            // Gaussian of width sigma sqrt(2) (W(0)^2 = 1) gives the loss and
            // g_rho (LossSq2Store)
            if ((rc = smooth(diff, dsum, smooth_par * std::sqrt(2.0)))) return rc;
            if (plain_response(rp))
                rc = zpass_c2r<T>(nrows, m.n2, tw, HalfRowLoad<T>{S, hp},
                                  LossSq2Store<T, true>{target, rho, grho, rp, npix, m.n2, T(1), bpart, dsum, invMd}, st, nullptr,
                                  "zpass_c2r_loss2");
            else
                rc = zpass_c2r<T>(nrows, m.n2, tw, HalfRowLoad<T>{S, hp},
                                  LossSq2Store<T>{target, rho, grho, rp, npix, m.n2, T(1), bpart, dsum, invMd}, st, nullptr,
                                  "zpass_c2r_loss2");
            if (rc) return rc;
        } else {
            if ((rc = smooth(diff, dsum, smooth_par))) return rc;
            if ((rc = zpass_c2r<T>(nrows, m.n2, tw, HalfRowLoad<T>{S, hp},
                                   LossStore<T>{mask, maskv, grho ? gds : nullptr, npix, m.n2, T(1), p_norm == 1, bpart, dsum, W0 * invMd},
                                   st, nullptr, "zpass_c2r_loss")))
                return rc;
            if (grho) {
                hipLaunchKernelGGL(k_reduce_cols, dim3(1), dim3(256), 0, st, bpart + 6, nzb, kLossCols, r1, gsum, 1);
                if ((rc = smooth(gds, gsum, smooth_par))) return rc;
                if ((rc = zpass_c2r<T>(nrows, m.n2, tw, HalfRowLoad<T>{S, hp}, RespBwdStore<T>{rho, grho, rp, m.n2, T(1), bpart, gsum, W0 * invMd},
                                       st, nullptr, "zpass_c2r_resp")))
                    return rc;
            }
        }
        hipLaunchKernelGGL(k_reduce_cols, dim3(1), dim3(256), 0, st, bpart, nzb, kLossCols, r6, sums, 1);
    }
    hipLaunchKernelGGL(k_loss_finish, dim3(1), dim3(64), 0, st, sums, npix, out, kind == kKindHi ? 6 : 5);
    LDL_LAUNCH_CHECK();
    return LDL_OK;
}

// prm = [beta1, khi, kli, nui, beta2, gamma3, xi, eta] (model.py:322-331)
template <typename T>
static int hi_dep_setup(const T* star, const c2<T>* Sk, const int64_t nmesh[3], const double box[3], const double* prm, void* ws, size_t wsb,
                        hipStream_t st, Mesh* m, SmPtrs<T>* P, const c2<T>** tw, KTabs<T>* kt, const c2<T>** sk) {
    int rc = check_fft_mesh(nmesh, m);
    if (rc) return rc;
    if ((rc = check_prm<T>(prm, 8))) return rc;
    if (!star) return fail(LDL_EINVAL, "NULL array argument");
    if (!box || !(box[0] > 0 && box[1] > 0 && box[2] > 0)) return fail(LDL_EINVAL, "box must be positive");
    if (!(prm[2] != 0.0)) return fail(LDL_EINVAL, "kli must be nonzero");
    const SmWs w = sm_ws(*m, sizeof(T), 0, LDL_PAINT_FLOAT_ATOMIC);  // no paint region
    if (!ws || wsb < w.total) return fail(LDL_EINVAL, "sm workspace too small (see ldl_sm_workspace_bytes)");
    *P = sm_ptrs<T>(ws, w);
    if ((rc = twiddles<T>(tw))) return rc;
    *kt = build_tables<T>(P->tabs, *m, box, st);
    *sk = Sk;
    if (!Sk) {  // Sk = r2c(starmap) / N^3 (unpadded rows, the caller's layout) into the second spectrum slot
        c2<T>* o = P->spec3 + spec_cells<T>(*m);
        if ((rc = r2c_full<T>(*m, *tw, RealRowLoad<T>{star, m->n2}, o, m->nh(), (T)(1.0 / (double)m->cells()), st))) return rc;
        *sk = o;
    }
    return LDL_OK;
}

template <typename T>
static int hi_depletion_fwd(const T* star, const c2<T>* Sk, const int64_t nmesh[3], const double box[3], const double* prm, T* S, T* dep,
                            void* ws, size_t wsb, hipStream_t st) {
    Mesh m;
    SmPtrs<T> P;
    const c2<T>* tw;
    KTabs<T> kt;
    const c2<T>* sk;
    int rc = hi_dep_setup<T>(star, Sk, nmesh, box, prm, ws, wsb, st, &m, &P, &tw, &kt, &sk);
    if (rc) return rc;
    if (!S || !dep) return fail(LDL_EINVAL, "NULL array argument");
    const int64_t nh = m.nh(), hp = spec_pitch<T>(m), hc = m.hcells();
    const SmFilter<T> fl{(T)prm[1], (T)(1.0 / prm[2]), (T)prm[3], 1};
    c2<T>* W = P.spec3;
    hipLaunchKernelGGL(k_hi_wmult<T>, dim3(grid_for(hc, 256, 1 << 16)), dim3(256), 0, st, sk, W, hc, m.n1, nh, hp, kt, fl, (T)prm[6]);
    if ((rc = colpass<T, +1>(m.n0, m.n1, nh, hp, m.n1 * hp, tw, PlainColOp<T>{W, W, T(1)}, st))) return rc;
    if ((rc = colpass<T, +1>(m.n1, m.n0, nh, m.n1 * hp, hp, tw, PlainColOp<T>{W, W, T(1)}, st))) return rc;
    const DepPar<T> dp{(T)prm[0], (T)prm[4], (T)prm[5], (T)prm[7]};
    if ((rc = zpass_c2r<T>(m.n0 * m.n1, m.n2, tw, HalfRowLoad<T>{W, hp}, DepStore<T>{star, S, dep, m.n2, dp}, st, nullptr))) return rc;
    LDL_LAUNCH_CHECK();
    return LDL_OK;
}

template <typename T>
static int hi_depletion_bwd(const T* star, const c2<T>* Sk, const T* S, const T* gdep, const int64_t nmesh[3], const double box[3],
                            const double* prm, int32_t rule, double* grad8, void* ws, size_t wsb, hipStream_t st) {
    Mesh m;
    SmPtrs<T> P;
    const c2<T>* tw;
    KTabs<T> kt;
    const c2<T>* sk;
    int rc = hi_dep_setup<T>(star, Sk, nmesh, box, prm, ws, wsb, st, &m, &P, &tw, &kt, &sk);
    if (rc) return rc;
    if (!S || !gdep || !grad8) return fail(LDL_EINVAL, "NULL array argument");
    const int64_t nh = m.nh(), hp = spec_pitch<T>(m), hc = m.hcells(), cells = m.cells();
    const DepPar<T> dp{(T)prm[0], (T)prm[4], (T)prm[5], (T)prm[7]};
    T* gS = P.mesh3;
    const unsigned gb = 2048;
    hipLaunchKernelGGL(k_hi_dep_bwd<T>, dim3(gb), dim3(256), 0, st, star, S, gdep, cells, dp, gS, P.part);
    const RedMap re{4, {0, 4, 5, 7}};
    hipLaunchKernelGGL(k_reduce_cols, dim3(1), dim3(256), 0, st, P.part, (int64_t)gb, 4, re, grad8, 1);
    c2<T>* gh = P.spec3;
    if ((rc = r2c_full<T>(m, tw, RealRowLoad<T>{gS, m.n2}, gh, hp, T(1), st))) return rc;
    const SmFilter<T> fl{(T)prm[1], (T)(1.0 / prm[2]), (T)prm[3], 1};
    const unsigned gf = grid_for(hc, 256, 4096);
    const int64_t zdbl = std::min<int64_t>(m.n1 / 2, nh);
    hipLaunchKernelGGL(k_hi_wgrad<T>, dim3(gf), dim3(256), 0, st, gh, sk, hc, m.n1, nh, hp, m.n2, zdbl, (int)(rule != 0), kt, fl,
                       1.0 / (prm[2] * prm[2]), P.part);
    const RedMap rw{4, {1, 2, 3, 6}};
    hipLaunchKernelGGL(k_reduce_cols, dim3(1), dim3(256), 0, st, P.part, (int64_t)gf, 4, rw, grad8, 1);
    LDL_LAUNCH_CHECK();
    return LDL_OK;
}

template <typename T>
static int paint_vjp_pos(const T* pos, const ldl_pos_desc* pdesc, int64_t np, const T* grho, const int64_t nmesh[3], double fact, T* g,
                         double* gmaxd, hipStream_t st) {
    auto* gmax = reinterpret_cast<unsigned long long*>(gmaxd);
    if (gmax) hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, st, gmax, (int64_t)3, 0ull);
    Mesh m;
    int rc = check_mesh(nmesh, &m);
    if (rc) return rc;
    if (np == 0) return LDL_OK;
    if (!pos || !grho || !g) return fail(LDL_EINVAL, "NULL array argument");
    PosDec<T> pd;
    if ((rc = make_posdec<T>(pdesc, nmesh, np, &pd))) return rc;
    const Sweep sw = make_sweep(pd, np);
    KTimer _t("paint_vjp_pos", st);
    with_geom<T>(m, pd, [&](auto geo) {
        hipLaunchKernelGGL((k_paint_vjp_pos<T, decltype(geo)>), dim3(sweep2_grid(sw)), dim3(kBlk), 0, st, pos, pd, np, sw, grho, geo, (T)fact, g,
                           gmax);
    });
    LDL_LAUNCH_CHECK();
    return LDL_OK;
}

}  // namespace ldl

using namespace ldl;

extern "C" {

size_t ldl_sm_bins_bytes(const int64_t nmesh[3], int64_t np) {
    Mesh m;
    if (check_mesh(nmesh, &m) || np < 0) return 0;
    return paint_bins_bytes(m, np);
}

int64_t ldl_sm_spectrum_pitch(const int64_t nmesh[3], int32_t is_f64) {
    Mesh m;
    if (check_fft_mesh(nmesh, &m)) return 0;
    return is_f64 ? spec_pitch<double>(m) : spec_pitch<float>(m);
}

size_t ldl_sm_workspace_bytes(const int64_t nmesh[3], int32_t is_f64, int64_t np, int32_t paint_flags) {
    Mesh m;
    if (check_mesh(nmesh, &m) || np < 0) return 0;
    return sm_ws(m, is_f64 ? 8 : 4, np, paint_flags).total;
}

#define SM_ENTRY(SUF, T)                                                                                                    \
    int ldl_sm_layer_fwd_##SUF(const T* pos, const ldl_pos_desc* pd, int64_t np, const int64_t nmesh[3], const double prm[5], \
                               double fact, int32_t pflags, T* pos_out, T* rho, T* U, T* F, void* bins, void* ws, size_t wsb, \
                               void* stream) {                                                                              \
        return sm_layer_fwd<T>(kKindSm, pos, pd, np, nmesh, prm, fact, pflags, nullptr, pos_out, rho, U, nullptr, F, bins, ws, \
                               wsb, as_stream(stream));                                                                     \
    }                                                                                                                       \
    int ldl_dm_layer_fwd_##SUF(const T* pos, const ldl_pos_desc* pd, int64_t np, const int64_t nmesh[3], const double prm[5], \
                               double fact, int32_t pflags, T* pos_out, T* rho, T* U, T* F, void* bins, void* ws, size_t wsb, \
                               void* stream) {                                                                              \
        return sm_layer_fwd<T>(kKindDm, pos, pd, np, nmesh, prm, fact, pflags, nullptr, pos_out, rho, U, nullptr, F, bins, ws, \
                               wsb, as_stream(stream));                                                                     \
    }                                                                                                                       \
    int ldl_hi_layer_fwd_##SUF(const T* pos, const ldl_pos_desc* pd, int64_t np, const int64_t nmesh[3], const double prm[10], \
                               double fact, int32_t pflags, const T* starmap, T* pos_out, T* rho, T* U, T* S2, T* F,        \
                               void* bins, void* ws, size_t wsb, void* stream) {                                            \
        return sm_layer_fwd<T>(kKindHi, pos, pd, np, nmesh, prm, fact, pflags, starmap, pos_out, rho, U, S2, F, bins, ws, wsb, \
                               as_stream(stream));                                                                          \
    }                                                                                                                       \
    int ldl_sm_layer_bwd_##SUF(const T* pos, const ldl_pos_desc* pd, int64_t np, const int64_t nmesh[3], const double prm[5], \
                               double fact, int32_t alpha_rule, int32_t pflags, const T* rho, const T* U, const T* F,      \
                               const void* bins, const T* g_out, T* g_in, double* grad5, double* gmax, void* ws,           \
                               size_t wsb, void* stream) {                                                                  \
        return sm_layer_bwd<T>(kKindSm, pos, pd, np, nmesh, prm, fact, alpha_rule, pflags, nullptr, rho, U, nullptr, F, bins, \
                               g_out, g_in, grad5, gmax, ws, wsb, as_stream(stream));                                       \
    }                                                                                                                       \
    int ldl_dm_layer_bwd_##SUF(const T* pos, const ldl_pos_desc* pd, int64_t np, const int64_t nmesh[3], const double prm[5], \
                               double fact, int32_t alpha_rule, int32_t pflags, const T* rho, const T* U, const T* F,      \
                               const void* bins, const T* g_out, T* g_in, double* grad5, double* gmax, void* ws,           \
                               size_t wsb, void* stream) {                                                                  \
        return sm_layer_bwd<T>(kKindDm, pos, pd, np, nmesh, prm, fact, alpha_rule, pflags, nullptr, rho, U, nullptr, F, bins, \
                               g_out, g_in, grad5, gmax, ws, wsb, as_stream(stream));                                       \
    }                                                                                                                       \
    int ldl_hi_layer_bwd_##SUF(const T* pos, const ldl_pos_desc* pd, int64_t np, const int64_t nmesh[3], const double prm[10], \
                               double fact, int32_t alpha_rule, int32_t pflags, const T* starmap, const T* rho, const T* U, \
                               const T* S2, const T* F, const void* bins, const T* g_out, T* g_in, double* grad10,         \
                               double* gmax, void* ws, size_t wsb, void* stream) {                                          \
        return sm_layer_bwd<T>(kKindHi, pos, pd, np, nmesh, prm, fact, alpha_rule, pflags, starmap, rho, U, S2, F, bins,   \
                               g_out, g_in, grad10, gmax, ws, wsb, as_stream(stream));                                      \
    }                                                                                                                       \
    int ldl_sm_loss_##SUF(const T* rho, const T* target, const T* mask, const T* mask_valid, const T* field2,               \
                          const int64_t nmesh[3], const double box[3], const double prm3[3], double index, int32_t p_norm,  \
                          int32_t smooth_kind, double smooth_par, double* out5, T* g_rho, void* ws, size_t wsb,             \
                          void* stream) {                                                                                   \
        if (!prm3) return fail(LDL_EINVAL, "prm3 is NULL");                                                                \
        return sm_loss<T>(kKindSm, rho, target, mask, mask_valid, field2, nullptr, nmesh, box, prm3, index, p_norm,        \
                          smooth_kind, smooth_par, out5, g_rho, nullptr, ws, wsb, as_stream(stream));                       \
    }                                                                                                                       \
    int ldl_sm_response_##SUF(const T* rho, const int64_t nmesh[3], const double prm3[3], T* out, void* stream) {         \
        return sm_response<T>(kKindSm, rho, nullptr, nmesh, prm3, out, as_stream(stream));                                 \
    }                                                                                                                       \
    int ldl_hi_response_##SUF(const T* rho, const T* dep, const int64_t nmesh[3], const double prm4[4], T* out,           \
                              void* stream) {                                                                               \
        return sm_response<T>(kKindHi, rho, dep, nmesh, prm4, out, as_stream(stream));                                     \
    }                                                                                                                       \
    int ldl_dm_loss_##SUF(const T* rho, const T* target, const T* mask, const T* mask_valid, const T* field2,               \
                          const int64_t nmesh[3], const double box[3], double index, int32_t p_norm, int32_t smooth_kind,   \
                          double smooth_par, double* out5, T* g_rho, void* ws, size_t wsb, void* stream) {                 \
        return sm_loss<T>(kKindDm, rho, target, mask, mask_valid, field2, nullptr, nmesh, box, nullptr, index, p_norm,     \
                          smooth_kind, smooth_par, out5, g_rho, nullptr, ws, wsb, as_stream(stream));                       \
    }                                                                                                                       \
    int ldl_hi_loss_##SUF(const T* rho, const T* target, const T* mask, const T* mask_valid, const T* field2,               \
                          const T* dep, const int64_t nmesh[3], const double box[3], const double prm4[4], double index,    \
                          int32_t p_norm, int32_t smooth_kind, double smooth_par, double* out6, T* g_rho, T* g_dep,         \
                          void* ws, size_t wsb, void* stream) {                                                             \
        if (!prm4) return fail(LDL_EINVAL, "prm4 is NULL");                                                                \
        return sm_loss<T>(kKindHi, rho, target, mask, mask_valid, field2, dep, nmesh, box, prm4, index, p_norm,            \
                          smooth_kind, smooth_par, out6, g_rho, g_dep, ws, wsb, as_stream(stream));                         \
    }                                                                                                                       \
    int ldl_hi_depletion_fwd_##SUF(const T* starmap, const T* starmapk, const int64_t nmesh[3], const double box[3],       \
                                   const double prm8[8], T* sfield, T* dep, void* ws, size_t wsb, void* stream) {           \
        return hi_depletion_fwd<T>(starmap, reinterpret_cast<const c2<T>*>(starmapk), nmesh, box, prm8, sfield, dep, ws,   \
                                   wsb, as_stream(stream));                                                                 \
    }                                                                                                                       \
    int ldl_hi_depletion_bwd_##SUF(const T* starmap, const T* starmapk, const T* sfield, const T* g_dep,                   \
                                   const int64_t nmesh[3], const double box[3], const double prm8[8], int32_t grad_rule,    \
                                   double* grad8, void* ws, size_t wsb, void* stream) {                                     \
        return hi_depletion_bwd<T>(starmap, reinterpret_cast<const c2<T>*>(starmapk), sfield, g_dep, nmesh, box, prm8,     \
                                   grad_rule, grad8, ws, wsb, as_stream(stream));                                           \
    }                                                                                                                       \
    int ldl_paint_vjp_pos_##SUF(const T* pos, const ldl_pos_desc* pd, int64_t np, const T* g_rho, const int64_t nmesh[3],   \
                                double fact, T* g, double* gmax, void* stream) {                                            \
        return paint_vjp_pos<T>(pos, pd, np, g_rho, nmesh, fact, g, gmax, as_stream(stream));                               \
    }

SM_ENTRY(f32, float)
SM_ENTRY(f64, double)

}  // extern "C"
