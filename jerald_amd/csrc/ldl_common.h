// Shared definitions for the libldl HIP sources (gfx950 / CDNA4).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <string>

#include "../../include/ldl.h"

namespace ldl {

// ------------------------------------------------------------ error channel
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define LDL_HIP_TRY(expr)                                                              \
    do {                                                                               \
        hipError_t _e = (expr);                                                        \
        if (_e != hipSuccess)                                                          \
            return ::ldl::fail(LDL_EHIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

#define LDL_LAUNCH_CHECK()                                                             \
    do {                                                                               \
        hipError_t _e = hipGetLastError();                                             \
        if (_e != hipSuccess)                                                          \
            return ::ldl::fail(LDL_EHIP, std::string("kernel launch: ") + hipGetErrorString(_e)); \
    } while (0)

// ------------------------------------------------------------- vector types
template <typename T> struct V2;
template <> struct V2<float> { using t = float2; };
template <> struct V2<double> { using t = double2; };
template <typename T> using c2 = typename V2<T>::t;

template <typename C> __host__ __device__ __forceinline__ C cmul(C a, C b) {
    return C{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename C> __host__ __device__ __forceinline__ C cmulc(C a, C b) {  // a * conj(b)
    return C{a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
template <typename C> __host__ __device__ __forceinline__ C cadd(C a, C b) { return C{a.x + b.x, a.y + b.y}; }
template <typename C> __host__ __device__ __forceinline__ C csub(C a, C b) { return C{a.x - b.x, a.y - b.y}; }
template <typename C> __host__ __device__ __forceinline__ C cconj(C a) { return C{a.x, -a.y}; }

// ----------------------------------------------------------- mesh geometry
struct Mesh {
    int64_t n0, n1, n2;
    __host__ __device__ int64_t cells() const { return n0 * n1 * n2; }
    __host__ __device__ int64_t nh() const { return n2 / 2 + 1; }
    __host__ __device__ int64_t hcells() const { return n0 * n1 * (n2 / 2 + 1); }
};

__host__ __device__ __forceinline__ int64_t wrapi(int64_t i, int64_t n) {
    if ((uint64_t)i < (uint64_t)n) return i;  // common case: already in range
    if (i == n) return 0;
    if (i == -1) return n - 1;
    int64_t r = i % n;
    return r < 0 ? r + n : r;
}

// ------------------------------------------------------ position decoding
// Device-side copy of ldl_pos_desc, specialised to the working precision.
template <typename T> struct PosDec {
    int kind;
    T box[3];
    T cell[3];     // box / N in T (reference: BoxSize / Nmesh in the working dtype)
    T cellinv[3];  // N / box in T (reference: Nmesh[d] / BoxSize[d], lib/parops_fast.h:748-750)
    uint32_t l12, l2;  // lattice strides (Np < 2^32 for the lattice form)
    int sh12, sh2;     // log2 of the strides when they are powers of two, else -1
    double lstep[3];
    int64_t n[3];

    int unit;          // lattice form with lattice_step == 1 on every axis

    // CIC lower corner (wrapped) and fractional offsets of particle p.
    __device__ __forceinline__ void decode(const T* __restrict__ pos, int64_t p, int64_t idx[3], T f[3]) const {
        const T* q = pos + 3 * p;
        decode_xyz(q[0], q[1], q[2], p, idx, f);
    }

    // Mesh coordinates x (before the floor) of particle p from its loaded
    // coordinates, and the integer cell offset ``base`` they are relative to
    // (the lattice cell; 0 for physical positions).
    template <typename I>
    __device__ __forceinline__ void coords(T x0, T x1, T x2, int64_t p, T x[3], I base[3]) const {
        x[0] = x0;
        x[1] = x1;
        x[2] = x2;
        base[0] = base[1] = base[2] = 0;
        if (kind == LDL_POS_PHYSICAL) {
#pragma unroll
            for (int d = 0; d < 3; d++) {
                // numpy remainder semantics, then division (jaxpops_bind.py:163-164);
                // fmod is the identity on [0, L)
                T r = x[d];
                if (!(r >= T(0) && r < box[d])) {
                    r = fmod(r, box[d]);
                    if (r != T(0) && ((r < T(0)) != (box[d] < T(0)))) r += box[d];
                }
                x[d] = r / cell[d];
            }
        } else {
            uint32_t pp = (uint32_t)p;
            uint32_t li[3];
            if (sh12 >= 0 && sh2 >= 0) {
                li[0] = pp >> sh12;
                const uint32_t rem = pp & (l12 - 1u);
                li[1] = rem >> sh2;
                li[2] = rem & (l2 - 1u);
            } else {
                li[0] = pp / l12;
                const uint32_t rem = pp - li[0] * l12;
                li[1] = rem / l2;
                li[2] = rem - li[1] * l2;
            }
            if (unit) {
#pragma unroll
                for (int d = 0; d < 3; d++) base[d] = (I)li[d];
            } else {
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    double b = (double)li[d] * lstep[d];
                    double bf = floor(b);
                    base[d] = (I)bf;
                    x[d] = x[d] + (T)(b - bf);
                }
            }
        }
    }

    // Same, from the already loaded coordinates (x0, x1, x2) of particle p.
    __device__ __forceinline__ void decode_xyz(T x0, T x1, T x2, int64_t p, int64_t idx[3], T f[3]) const {
        T x[3];
        int64_t base[3];
        coords(x0, x1, x2, p, x, base);
#pragma unroll
        for (int d = 0; d < 3; d++) {
            T fl = floor(x[d]);
            f[d] = x[d] - fl;
            idx[d] = wrapi(base[d] + (int64_t)fl, n[d]);
        }
    }

    // Power-of-two mesh: 32-bit cell indices wrapped with the masks n - 1
    // (mesh coordinates within +-2^31 cells).
    __device__ __forceinline__ void decode_p2(T x0, T x1, T x2, int64_t p, const uint32_t msk[3], uint32_t idx[3], T f[3]) const {
        T x[3];
        int32_t base[3];
        coords(x0, x1, x2, p, x, base);
#pragma unroll
        for (int d = 0; d < 3; d++) {
            T fl = floor(x[d]);
            f[d] = x[d] - fl;
            idx[d] = (uint32_t)(base[d] + (int32_t)fl) & msk[d];
        }
    }

    // decode_p2 for lattice positions with unit steps and power-of-two strides
    // (lattice_unit_p2() on the host): no run-time branches on the form.
    __device__ __forceinline__ void decode_lu(T x0, T x1, T x2, int64_t p, const uint32_t msk[3], uint32_t idx[3], T f[3]) const {
        const uint32_t pp = (uint32_t)p;
        const uint32_t rem = pp & (l12 - 1u);
        const int32_t base[3] = {(int32_t)(pp >> sh12), (int32_t)(rem >> sh2), (int32_t)(rem & (l2 - 1u))};
        const T x[3] = {x0, x1, x2};
#pragma unroll
        for (int d = 0; d < 3; d++) {
            T fl = floor(x[d]);
            f[d] = x[d] - fl;
            idx[d] = (uint32_t)(base[d] + (int32_t)fl) & msk[d];
        }
    }
    bool lattice_unit_p2() const { return kind == LDL_POS_LATTICE && unit && sh12 >= 0 && sh2 >= 0; }
};

template <typename T>
inline int make_posdec(const ldl_pos_desc* pd, const int64_t nmesh[3], int64_t np, PosDec<T>* out) {
    if (!pd) return fail(LDL_EINVAL, "pos_desc is NULL");
    if (pd->kind != LDL_POS_PHYSICAL && pd->kind != LDL_POS_LATTICE)
        return fail(LDL_EINVAL, "pos_desc.kind must be LDL_POS_PHYSICAL or LDL_POS_LATTICE");
    out->kind = pd->kind;
    for (int d = 0; d < 3; d++) {
        if (!(pd->box[d] > 0.0) || !std::isfinite(pd->box[d])) return fail(LDL_EINVAL, "box sizes must be positive and finite");
        out->box[d] = (T)pd->box[d];
        out->cell[d] = out->box[d] / (T)nmesh[d];
        out->cellinv[d] = (T)nmesh[d] / out->box[d];
        out->n[d] = nmesh[d];
        out->lstep[d] = pd->lattice_step[d];
    }
    out->l12 = out->l2 = 1;
    out->sh12 = out->sh2 = 0;
    out->unit = 0;
    if (pd->kind == LDL_POS_LATTICE) {
        int64_t l0 = pd->lattice[0], l1 = pd->lattice[1], l2 = pd->lattice[2];
        if (l0 <= 0 || l1 <= 0 || l2 <= 0) return fail(LDL_EINVAL, "lattice sizes must be positive");
        if (l0 * l1 * l2 != np) return fail(LDL_EINVAL, "lattice[0]*lattice[1]*lattice[2] must equal np");
        if (np > (int64_t)0xFFFFFFFFll) return fail(LDL_EUNSUPPORTED, "lattice form supports np < 2^32");
        out->l12 = (uint32_t)(l1 * l2);
        out->l2 = (uint32_t)l2;
        auto lg = [](uint32_t v) { int k = 0; while ((1u << k) < v) k++; return (1u << k) == v ? k : -1; };
        out->sh12 = lg(out->l12);
        out->sh2 = lg(out->l2);
        for (int d = 0; d < 3; d++)
            if (!(pd->lattice_step[d] > 0.0)) return fail(LDL_EINVAL, "lattice_step must be positive");
        out->unit = pd->lattice_step[0] == 1.0 && pd->lattice_step[1] == 1.0 && pd->lattice_step[2] == 1.0;
    }
    return LDL_OK;
}

inline bool is_pow2(int64_t n) { return n > 0 && (n & (n - 1)) == 0; }

inline int check_mesh(const int64_t nmesh[3], Mesh* m) {
    if (!nmesh) return fail(LDL_EINVAL, "nmesh is NULL");
    for (int d = 0; d < 3; d++)
        if (nmesh[d] < 1) return fail(LDL_EINVAL, "mesh sizes must be >= 1");
    m->n0 = nmesh[0];
    m->n1 = nmesh[1];
    m->n2 = nmesh[2];
    return LDL_OK;
}

// FFT-capable mesh: each axis in [4, 1024] with no prime factor other than 2,
// 3 and 5; the last axis even.  (Powers of two take the register-resident
// kernels, other lengths the mixed-radix ones.)
inline bool fft_len_ok(int64_t n) {
    if (n < 4 || n > 1024) return false;
    for (int f : {2, 3, 5})
        while (n % f == 0) n /= f;
    return n == 1;
}
inline int check_fft_mesh(const int64_t nmesh[3], Mesh* m) {
    int rc = check_mesh(nmesh, m);
    if (rc) return rc;
    for (int d = 0; d < 3; d++)
        if (!fft_len_ok(nmesh[d]))
            return fail(LDL_EUNSUPPORTED, "FFT mesh axes must be in [4, 1024] with no prime factor other than 2, 3 and 5");
    if (nmesh[2] % 2) return fail(LDL_EUNSUPPORTED, "FFT mesh last axis must be even");
    return LDL_OK;
}
// Every axis a power of two (with a last axis >= 8): the fused LDL kernels apply.
inline bool fft_mesh_pow2(const Mesh& m) { return is_pow2(m.n0) && is_pow2(m.n1) && is_pow2(m.n2) && m.n2 >= 8; }

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline unsigned grid_for(int64_t n, int block, int64_t cap = 1 << 20) {
    int64_t g = (n + block - 1) / block;
    if (g > cap) g = cap;
    if (g < 1) g = 1;
    return (unsigned)g;
}

// Twiddle table for the current device: tw[L + m] = exp(-2 pi i m / L) for
// every power of two L in [1, 2048], built in FP64 on the host.
template <typename T> int twiddles(const c2<T>** out);
template <typename T> int twiddles_n(int n, const c2<T>** out);

// Bijective blockIdx -> tile map that gives each of the 8 XCDs (which the
// hardware feeds round-robin) a contiguous range of tiles, so tiles that
// share cache lines share an L2.
__device__ __forceinline__ int64_t xcd_tile(int64_t bid, int64_t nb) {
    const int64_t x = bid & 7, local = bid >> 3;
    const int64_t q = nb >> 3, r = nb & 7;
    return x * q + (x < r ? x : r) + local;
}

// The same ranges, each walked from its end: a pass that reads what the
// previous pass wrote in xcd_tile order starts on the lines written last.
__device__ __forceinline__ int64_t xcd_tile_rev(int64_t bid, int64_t nb) {
    const int64_t x = bid & 7, local = bid >> 3;
    const int64_t q = nb >> 3, r = nb & 7;
    const int64_t len = q + (x < r ? 1 : 0);
    return x * q + (x < r ? x : r) + (len - 1 - local);
}

// Order in which the workgroups of a particle gather sweep the particles.
// A workgroup takes chunks of kChunk consecutive particles (one per thread);
// its chunk id comes from xcd_tile, so each XCD owns a contiguous run of chunk
// ids.  Linear order: chunk c starts at particle c * kChunk.  Blocked order,
// for a lattice of L0 x L1 x L2 particles stored in lattice order with L2 a
// multiple of kChunk: chunk ids run over slabs of S lattice planes (S = L0 / 8
// when 8 divides L0, so each XCD owns one slab), then blocks of J rows, then
// the S planes, then the J rows, then the L2 / kChunk chunks of a row.  The
// mesh rows a lattice plane gathers from are reused by the next plane J rows'
// worth of particles later, while they are still in the XCD's L2; in linear
// order that reuse distance is a whole lattice plane, whose mesh footprint
// (several fields of 2-3 mesh planes) exceeds the L2.
constexpr int kChunk = 256;
struct Sweep {
    int blocked;
    uint32_t H, J, S, NYB, L1, L2;
    int64_t nchunks;
    __device__ __forceinline__ int64_t first(int64_t c) const {
        if (!blocked) return c * kChunk;
        uint32_t q = (uint32_t)c;
        const uint32_t h = q % H; q /= H;
        const uint32_t jj = q % J; q /= J;
        const uint32_t i = q % S; q /= S;
        const uint32_t yb = q % NYB, slab = q / NYB;
        return (((int64_t)(slab * S + i) * L1 + yb * J + jj) * L2) + (int64_t)h * kChunk;
    }
};

template <typename T>
inline Sweep make_sweep(const PosDec<T>& pd, int64_t np) {
    Sweep w{};
    w.nchunks = (np + kChunk - 1) / kChunk;
    if (pd.kind != LDL_POS_LATTICE || np == 0) return w;
    const int64_t L2 = pd.l2, L1 = pd.l12 / pd.l2, L0 = np / pd.l12;
    if (L2 % kChunk) return w;
    w.blocked = 1;
    w.L1 = (uint32_t)L1;
    w.L2 = (uint32_t)L2;
    w.H = (uint32_t)(L2 / kChunk);
    w.J = (uint32_t)L1;
    for (int64_t j = 32; j >= 2; j >>= 1)
        if (L1 % j == 0) { w.J = (uint32_t)j; break; }
    w.S = (uint32_t)(L0 % 8 == 0 ? L0 / 8 : L0);
    w.NYB = (uint32_t)(L1 / w.J);
    return w;
}

// Grid for a Sweep-ordered kernel (kChunk threads per workgroup).
inline unsigned sweep_grid(const Sweep& w) { return grid_for(w.nchunks, 1, 1 << 22); }

// Pair sweep: a workgroup takes two consecutive chunks of the Sweep order and
// each thread one particle of each, so a thread has both particles' loads in
// flight at once.  The body sees p0 and p1; p1 >= np when the second chunk
// does not exist (odd chunk count).
inline unsigned sweep2_grid(const Sweep& w) { return grid_for((w.nchunks + 1) / 2, 1, 1 << 22); }
#define LDL_SWEEP2(sw, np, p0, p1)                                                                           \
    for (int64_t _base = 0, _npair = ((sw).nchunks + 1) / 2; _base < _npair; _base += gridDim.x)            \
        for (int64_t _q = _base + xcd_tile(blockIdx.x, gridDim.x), p0 = (sw).first(2 * _q) + threadIdx.x,   \
                     p1 = 2 * _q + 1 < (sw).nchunks ? (sw).first(2 * _q + 1) + threadIdx.x : (np), _once = 1; \
             _once && _q < _npair; _once = 0)

// Loop over this workgroup's particles in Sweep order; the body sees ``p``.
#define LDL_SWEEP(sw, np, p)                                                                  \
    for (int64_t _base = 0; _base < (sw).nchunks; _base += gridDim.x)                         \
        for (int64_t _c = _base + xcd_tile(blockIdx.x, gridDim.x), p = (sw).first(_c) + threadIdx.x, \
                     _once = 1;                                                                \
             _once && _c < (sw).nchunks && p < (np); _once = 0)

// ------------------------------------------------- kernel timing (bench)
bool timing_enabled();
void timer_begin(const char* name, hipStream_t st, void** token);
void timer_end(const char* name, hipStream_t st, void* token);

// Scope guard that brackets the launches inside it with HIP events on ``st``
// when ldl_timer_enable(1) is active.
struct KTimer {
    const char* name;
    hipStream_t st;
    void* tok;
    KTimer(const char* n, hipStream_t s) : name(n), st(s), tok(nullptr) { timer_begin(n, s, &tok); }
    ~KTimer() { timer_end(name, st, tok); }
};

// ----------------------------------------------------- elementwise math
// Power, log and exp of positive arguments: the FP32 path uses the hardware
// log2 / exp2 / reciprocal instructions (a few ulp), the FP64 path the
// accurate library routines.
constexpr float kLn2f = 0.693147180559945309f;
__device__ __forceinline__ float ppow(float x, float y) { return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x)); }
__device__ __forceinline__ double ppow(double x, double y) { return pow(x, y); }
__device__ __forceinline__ float plog(float x) { return __builtin_amdgcn_logf(x) * kLn2f; }
__device__ __forceinline__ double plog(double x) { return log(x); }
__device__ __forceinline__ float pexp(float x) { return __expf(x); }
__device__ __forceinline__ double pexp(double x) { return exp(x); }
// x^a, x^(a-1) and ln x together (x > 0)
__device__ __forceinline__ void pow_pair_log(float x, float a, float& p, float& p1, float& ln) {
    const float l2 = __builtin_amdgcn_logf(x);
    p = __builtin_amdgcn_exp2f(a * l2);
    p1 = p * __builtin_amdgcn_rcpf(x);
    ln = l2 * kLn2f;
}
__device__ __forceinline__ void pow_pair_log(double x, double a, double& p, double& p1, double& ln) {
    ln = log(x);
    p = pow(x, a);
    p1 = pow(x, a - 1.0);
}

// Per-particle streams of the gathers (positions, cotangents): read or
// written once per launch, marked non-temporal so they do not evict the mesh
// lines the gathers reuse from L2.
template <typename T> __device__ __forceinline__ T ld_stream(const T* p) { return __builtin_nontemporal_load(p); }
template <typename T> __device__ __forceinline__ void st_stream(T* p, T v) { __builtin_nontemporal_store(v, p); }

// --------------------------------------------------------- reductions
template <typename T> __device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

}  // namespace ldl
