// CIC paint (particle -> mesh) and readout (mesh -> particle) on gfx950.
//
// Algorithm (lib/parops_fast.h:353-481, :615-821, single rank): x = mesh
// coordinate, i = floor(x) wrapped mod N, f = x - i; corner (i+a, j+b, k+c)
// gets weight w_x[a] w_y[b] w_z[c] with w[0] = 1 - f, w[1] = f.
//
// Paint accumulates in fixed point: each deposit is an integer multiple of
// 2^-s and deposits are summed with integer adds, so the density is
// bit-identical for any particle order or scheduling.  s is chosen per call
// from max|m| and Np so that no cell (nor the total) can overflow.  The
// rounded mass rint(m 2^s) is split over the corners x, then y, then z, each
// split truncating toward zero, so the painted total equals the sum of the
// rounded masses exactly and every deposit has the sign of its mass.
#include "cic.h"

#include <mutex>
#include <vector>

namespace ldl {

// ---------------------------------------------------------- host drivers
static constexpr int kBlock = 256;
static constexpr int64_t kHdrBytes = 256;

// Tiles of the tiled paint: 8 x 16 x 32 cells (z fastest) = 32 KB of 64-bit
// LDS accumulators per component (one component: several workgroups per CU;
// three: 96 KB, one workgroup of 1024 threads per CU).  Long z runs keep the
// particles a tile reads from the lattice-ordered arrays in runs of 32
// consecutive rows' worth of bytes, so fewer cache lines straddle two tiles.
// The 32-bit three-component paint holds 48 KB: three workgroups of 512
// threads per CU (6 waves per SIMD), so one tile's LDS clear and write-out
// overlap the others' deposits.
static constexpr int kTileX = 8, kTileY = 16, kTileZ = 32;
static constexpr int kTileSh[3] = {3, 4, 5};
static constexpr int kAcc32Block = 512;
// One component (32 KB of 64-bit accumulators): 512 threads with 4 particles
// each per round, 62 VGPRs, four workgroups per CU (0.84 ms per 512^3 launch
// against 0.91 for 256 threads with 6).
static constexpr int kPaint1Block = 512;

static inline size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

// Workspace: header | direct-mode mesh (nc, LDL_PAINT_DIRECT only) | counts |
// offsets | segment sums | key | sorted
struct PaintWs {
    size_t acc, counts, offsets, segsum, key, sorted, total;
};
static bool paint_direct(int32_t flags) { return (flags & LDL_PAINT_DIRECT) && !(flags & LDL_PAINT_FLOAT_ATOMIC); }
static PaintWs paint_ws(const Mesh& m, int nc, int64_t np, int32_t flags) {
    const TileGeom tg = make_tiles(m, kTileSh);
    const int64_t nk = tg.nkeys();
    PaintWs w;
    w.acc = al256((size_t)kHdrBytes);
    w.counts = w.acc + (paint_direct(flags) ? al256((size_t)nc * m.cells() * 8) : 0);
    w.offsets = w.counts + al256((size_t)(nk + 1) * 4);
    w.segsum = w.offsets + al256((size_t)(nk + 1) * 4);
    w.key = w.segsum + al256((size_t)(nk / kScanSeg + 1) * 4);
    w.sorted = w.key + al256((size_t)np * 4);
    w.total = w.sorted + al256((size_t)np * 4);
    return w;
}

size_t paint_ws_bytes(const Mesh& m, int nc, int64_t np, int32_t flags) {
    if (flags & LDL_PAINT_FLOAT_ATOMIC) return 0;  // float atomics deposit straight into the output
    return paint_ws(m, nc, np < 0 ? 0 : np, flags).total;
}

size_t paint_bins_bytes(const Mesh& m, int64_t np) {
    const TileGeom tg = make_tiles(m, kTileSh);
    return (size_t)(tg.nkeys() + 1 + (np < 0 ? 0 : np) + 1) * 4;
}

template <typename T>
int paint_impl(const T* pos, const ldl_pos_desc* pdesc, int64_t np, const T* mass, double mscalar, int nc,
               const int64_t nmesh[3], double out_scale, int32_t flags, T* out, void* ws, size_t ws_bytes,
               hipStream_t st, unsigned* bins_out, const unsigned* bins_in, const T* wmesh, double wscale,
               const unsigned long long* absmax_in) {
    Mesh m;
    int rc = check_mesh(nmesh, &m);
    if (rc) return rc;
    if (np < 0) return fail(LDL_EINVAL, "np must be >= 0");
    if (!out) return fail(LDL_EINVAL, "out is NULL");
    if (np > 0 && !pos) return fail(LDL_EINVAL, "pos is NULL");
    if (nc == 3 && !mass && np > 0) return fail(LDL_EINVAL, "ppaint3D needs a (Np,3) mass array");
    PosDec<T> pd;
    if ((rc = make_posdec<T>(pdesc, nmesh, np, &pd))) return rc;
    const int64_t cells = m.cells();
    const int64_t tot = (int64_t)nc * cells;
    const unsigned gp = grid_for(np, kBlock, 1 << 16);
    const unsigned gm = grid_for(tot, kBlock, 1 << 16);
    const int mode = flags & 1;
    if (mode == LDL_PAINT_FLOAT_ATOMIC) {
        hipLaunchKernelGGL(k_fill<T>, dim3(gm), dim3(kBlock), 0, st, out, tot, T(0));
        if (np > 0) {
            if (nc == 1)
                { KTimer _t("paint_deposit", st); hipLaunchKernelGGL((k_paint_float<T, 1>), dim3(gp), dim3(kBlock), 0, st, pos, pd, np, mass, (T)mscalar, (T)out_scale, m, out); }
            else
                { KTimer _t("paint_deposit", st); hipLaunchKernelGGL((k_paint_float<T, 3>), dim3(gp), dim3(kBlock), 0, st, pos, pd, np, mass, (T)mscalar, (T)out_scale, m, out); }
        }
        LDL_LAUNCH_CHECK();
        return LDL_OK;
    }
    if (np > (int64_t)0xFFFFFFFFll) return fail(LDL_EUNSUPPORTED, "paint supports np < 2^32");
    const PaintWs w = paint_ws(m, nc, np, flags);
    if (!ws || ws_bytes < w.total) return fail(LDL_EINVAL, "paint workspace too small (see ldl_paint_workspace_bytes)");
    char* wb = reinterpret_cast<char*>(ws);
    auto* hdr = reinterpret_cast<unsigned long long*>(wb);
    auto* acc = reinterpret_cast<unsigned long long*>(wb + w.acc);
    // header: max |m| bits; for a scalar mass it is known on the host
    unsigned long long hbits = 0;
    if (!mass) {
        double a = std::fabs((double)(T)mscalar);
        std::memcpy(&hbits, &a, 8);
    }
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, st, hdr, (int64_t)(kHdrBytes / 8), hbits);
    if (mass && absmax_in) {  // max |m| per component handed over by the kernel that wrote the masses
        hipLaunchKernelGGL(k_copy_u64, dim3(1), dim3(64), 0, st, absmax_in, hdr, (int64_t)nc);
    } else if (mass && np > 0) {
        KTimer _t("paint_absmax", st);
        if (nc == 1) hipLaunchKernelGGL((k_absmax<T, 1>), dim3(grid_for(np, kBlock, 4096)), dim3(kBlock), 0, st, mass, np, hdr);
        else hipLaunchKernelGGL((k_absmax<T, 3>), dim3(grid_for(np, kBlock, 4096)), dim3(kBlock), 0, st, mass, np, hdr);
    }
    if (flags & LDL_PAINT_DIRECT) {
        // reference variant: one device-scope 64-bit atomic per corner
        if (!(flags & LDL_PAINT_WS_ZEROED))
            hipLaunchKernelGGL(k_fill_u64, dim3(gm), dim3(kBlock), 0, st, acc, tot, 0ull);
        if (np > 0) {
            if (nc == 1)
                { KTimer _t("paint_deposit", st); hipLaunchKernelGGL((k_paint_fixed<T, 1>), dim3(gp), dim3(kBlock), 0, st, pos, pd, np, mass, (T)mscalar, m, acc, hdr); }
            else
                { KTimer _t("paint_deposit", st); hipLaunchKernelGGL((k_paint_fixed<T, 3>), dim3(gp), dim3(kBlock), 0, st, pos, pd, np, mass, (T)mscalar, m, acc, hdr); }
        }
        if (nc == 1)
            { KTimer _t("paint_convert", st); hipLaunchKernelGGL((k_fixed_to_float<T, 1>), dim3(grid_for(cells, kBlock, 1 << 16)), dim3(kBlock), 0, st, acc, hdr, np, cells, out_scale, out); }
        else
            { KTimer _t("paint_convert", st); hipLaunchKernelGGL((k_fixed_to_float<T, 3>), dim3(grid_for(cells, kBlock, 1 << 16)), dim3(kBlock), 0, st, acc, hdr, np, cells, out_scale, out); }
        LDL_LAUNCH_CHECK();
        return LDL_OK;
    }
    // tiled path: counting sort by (tile, face class), then one workgroup per
    // tile sums the tile's cells completely in LDS.  A 3-component paint that
    // is handed the sort of an earlier 1-component paint of the same positions
    // (bins_in) skips the sort.
    const TileGeom tg = make_tiles(m, kTileSh);
    const int64_t nk = tg.nkeys();
    if (m.n0 >= (1ll << 31) || m.n1 >= (1ll << 31) || m.n2 >= (1ll << 31) || nk >= (1ll << 32))
        return fail(LDL_EUNSUPPORTED, "the tiled paint takes mesh axes below 2^31 cells and fewer than 2^29 tiles");
    // power-of-two meshes decode with 32-bit mask wraps, unit-step lattice
    // positions with power-of-two strides without run-time branches (the sort
    // and deposit kernels must use the same decode)
    const bool p2 = is_pow2(m.n0) && is_pow2(m.n1) && is_pow2(m.n2);
    const int dm = !p2 ? kDecAny : pd.lattice_unit_p2() ? kDecLU : kDecP2;
    auto with_dm = [dm](auto&& f) {
        if (dm == kDecLU) f(std::integral_constant<int, kDecLU>{});
        else if (dm == kDecP2) f(std::integral_constant<int, kDecP2>{});
        else f(std::integral_constant<int, kDecAny>{});
    };
    const MeshP2 mp = mesh_p2(m);
    auto* counts = reinterpret_cast<unsigned*>(wb + w.counts);
    auto* segsum = reinterpret_cast<unsigned*>(wb + w.segsum);
    auto* key = reinterpret_cast<unsigned*>(wb + w.key);
    unsigned* offsets = reinterpret_cast<unsigned*>(wb + w.offsets);
    unsigned* sorted = reinterpret_cast<unsigned*>(wb + w.sorted);
    // the FP32 1-component paint that hands its sort over also records its
    // largest cell there; a 3-component paint handed that sort reads it in place
    // of scanning the density
    unsigned* vmax_out = nullptr;
    if (bins_out) {
        offsets = bins_out;
        sorted = bins_out + nk + 1;
        if (nc == 1 && sizeof(T) == 4) {
            vmax_out = bins_out + nk + 1 + np;
            hipLaunchKernelGGL(k_fill<unsigned>, dim3(1), dim3(64), 0, st, vmax_out, (int64_t)1, 0u);
        }
    }
    if (bins_in) {
        offsets = const_cast<unsigned*>(bins_in);
        sorted = const_cast<unsigned*>(bins_in) + nk + 1;
    } else {
        hipLaunchKernelGGL(k_fill<unsigned>, dim3(grid_for(nk, kBlock, 1 << 16)), dim3(kBlock), 0, st, counts, nk, 0u);
        if (np > 0) {
            constexpr int PPT = 8;
            KTimer _t("paint_bin", st);
            const dim3 gb(grid_for((np + 256 * PPT - 1) / (256 * PPT), 1, 1 << 15));
            with_dm([&](auto D) {
                hipLaunchKernelGGL((k_bin_keys<T, PPT, decltype(D)::value>), gb, dim3(256), 0, st, pos, pd, np, mp, tg, counts, key);
            });
        }
        {
            KTimer _t("paint_scan", st);
            const unsigned nseg = (unsigned)((nk + kScanSeg - 1) / kScanSeg);
            hipLaunchKernelGGL(k_scan_sums, dim3(nseg), dim3(1024), 0, st, counts, nk, segsum);
            hipLaunchKernelGGL(k_scan_apply, dim3(nseg), dim3(1024), 0, st, counts, nk, segsum, offsets, counts);  // counts become the cursors
        }
        if (np > 0) {
            KTimer _t("paint_scatter", st);
            constexpr int PPT = 8;
            const dim3 gs(grid_for((np + 256 * PPT - 1) / (256 * PPT), 1, 1 << 15));
            hipLaunchKernelGGL(k_bin_scatter<PPT>, gs, dim3(256), 0, st, np, key, counts, sorted);
        }
    }
    const dim3 gt((unsigned)tg.ntiles);
    if (nc == 1) {
        KTimer _t("paint_deposit1", st);
        with_dm([&](auto D) {
            hipLaunchKernelGGL((k_paint_tiles<T, 1, kTileX, kTileY, kTileZ, kPaint1Block, decltype(D)::value>), gt, dim3(kPaint1Block), 0, st, pos, pd,
                               np, mass, (T)mscalar, m, mp, tg, offsets, sorted, hdr, out_scale, out, vmax_out);
        });
    } else if (sizeof(T) == 4 && wmesh) {
        // FP32 with the positions' unit-mass density known: 32-bit accumulators
        if (bins_in) {  // recorded by the 1-component paint that wrote the sort and wmesh
            hipLaunchKernelGGL(k_vmax_scaled, dim3(1), dim3(1), 0, st, bins_in + nk + 1 + np, wscale, hdr + 3);
        } else {
            KTimer _t("paint_absmax", st);
            hipLaunchKernelGGL(k_mesh_absmax<T>, dim3(grid_for(cells, kBlock, 2048)), dim3(kBlock), 0, st, wmesh, cells, wscale, hdr + 3);
        }
        KTimer _t("paint_deposit3", st);
        if constexpr (sizeof(T) == 4) {
            with_dm([&](auto D) {
                hipLaunchKernelGGL((k_paint_tiles<T, 3, kTileX, kTileY, kTileZ, kAcc32Block, decltype(D)::value, true>), gt, dim3(kAcc32Block), 0, st, pos,
                                   pd, np, mass, (T)mscalar, m, mp, tg, offsets, sorted, hdr, out_scale, out);
            });
        }
    } else {
        KTimer _t("paint_deposit3", st);
        with_dm([&](auto D) {
            hipLaunchKernelGGL((k_paint_tiles<T, 3, kTileX, kTileY, kTileZ, 1024, decltype(D)::value>), gt, dim3(1024), 0, st, pos, pd, np, mass,
                               (T)mscalar, m, mp, tg, offsets, sorted, hdr, out_scale, out);
        });
    }
    LDL_LAUNCH_CHECK();
    return LDL_OK;
}

template <typename T>
int readout_impl(const T* pos, const ldl_pos_desc* pdesc, int64_t np, const T* field, const int64_t nmesh[3],
                 int kind, int vjpdim, const T* aux, T* out, hipStream_t st) {
    Mesh m;
    int rc = check_mesh(nmesh, &m);
    if (rc) return rc;
    if (np < 0) return fail(LDL_EINVAL, "np must be >= 0");
    if (np == 0) return LDL_OK;
    if (!pos || !field || !out) return fail(LDL_EINVAL, "NULL array argument");
    PosDec<T> pd;
    if ((rc = make_posdec<T>(pdesc, nmesh, np, &pd))) return rc;
    const Sweep sw = make_sweep(pd, np);
    const unsigned g = sweep_grid(sw);
    static_assert(kBlock == kChunk, "gathers take one particle per thread");
    if (kind == 0 && (vjpdim < -1 || vjpdim > 2)) return fail(LDL_EINVAL, "vjpdim must be between -1 and 2");
    if (kind == 3 && !aux) return fail(LDL_EINVAL, "v3 is NULL");
    KTimer _t("readout", st);
    switch (kind) {
        case 0:
            with_geom<T>(m, pd, [&](auto geo) {
                hipLaunchKernelGGL((k_readout<T, decltype(geo)>), dim3(g), dim3(kBlock), 0, st, pos, pd, np, sw, field, geo, vjpdim, out);
            });
            break;
        case 1:
            with_geom<T>(m, pd, [&](auto geo) {
                hipLaunchKernelGGL((k_readout3<T, decltype(geo)>), dim3(g), dim3(kBlock), 0, st, pos, pd, np, sw, fields3(field, m.cells()), geo, out);
            });
            break;
        case 2:
            with_geom<T>(m, pd, [&](auto geo) {
                hipLaunchKernelGGL((k_readout_grad<T, decltype(geo)>), dim3(g), dim3(kBlock), 0, st, pos, pd, np, sw, field, geo, aux, out);
            });
            break;
        case 3:
            with_geom<T>(m, pd, [&](auto geo) {
                hipLaunchKernelGGL((k_readout_grad3<T, decltype(geo)>), dim3(g), dim3(kBlock), 0, st, pos, pd, np, sw, fields3(field, m.cells()), aux, geo, out);
            });
            break;
    }
    LDL_LAUNCH_CHECK();
    return LDL_OK;
}

template int paint_impl<float>(const float*, const ldl_pos_desc*, int64_t, const float*, double, int, const int64_t*, double, int32_t, float*, void*,
                               size_t, hipStream_t, unsigned*, const unsigned*, const float*, double, const unsigned long long*);
template int paint_impl<double>(const double*, const ldl_pos_desc*, int64_t, const double*, double, int, const int64_t*, double, int32_t, double*,
                                void*, size_t, hipStream_t, unsigned*, const unsigned*, const double*, double, const unsigned long long*);
template int readout_impl<float>(const float*, const ldl_pos_desc*, int64_t, const float*, const int64_t*, int, int, const float*, float*, hipStream_t);
template int readout_impl<double>(const double*, const ldl_pos_desc*, int64_t, const double*, const int64_t*, int, int, const double*, double*, hipStream_t);

// ------------------------------------------------ allocation-free paint
// ldl_ppaint_noalloc_* paint with this fixed device array as their workspace,
// for callers that cannot allocate (graph capture, serving loops).  Host code
// reaches it only through the address hipGetSymbolAddress returns for the
// current device (the array's name is a host-side stand-in there); every
// call on a device shares that device's copy.
static constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

static std::mutex g_scratch_mutex;
static std::vector<void*> g_scratch_ptr;  // per device id

static int noalloc_scratch(void** out) {
    int dev = 0;
    LDL_HIP_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(LDL_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return LDL_OK;
}

template <typename T>
static int paint_noalloc(const T* pos, const ldl_pos_desc* pdesc, int64_t np, const T* mass, double mscalar, const int64_t nmesh[3],
                         double out_scale, int32_t flags, T* out, hipStream_t st) {
    Mesh m;
    int rc = check_mesh(nmesh, &m);
    if (rc) return rc;
    if (np < 0) return fail(LDL_EINVAL, "np must be >= 0");
    // the array holds whatever the previous call left: never skip clearing
    flags &= ~LDL_PAINT_WS_ZEROED;
    const size_t need = paint_ws_bytes(m, 1, np, flags);
    if (need > kNoAllocScratchBytes)
        return fail(LDL_EINVAL, "ldl_ppaint_noalloc: this paint needs " + std::to_string(need) + " bytes of scratch, the fixed array holds " +
                                    std::to_string(kNoAllocScratchBytes) + " (use ldl_ppaint with a workspace)");
    void* ws = nullptr;
    if ((rc = noalloc_scratch(&ws))) return rc;
    return paint_impl<T>(pos, pdesc, np, mass, mscalar, 1, nmesh, out_scale, flags, out, ws, kNoAllocScratchBytes, st);
}

}  // namespace ldl

using namespace ldl;

extern "C" {

int ldl_ppaint_noalloc_f32(const float* pos, const ldl_pos_desc* pd, int64_t np, const float* mass, double mass_scalar,
                           const int64_t nmesh[3], double out_scale, int32_t flags, float* out, void* stream) {
    return paint_noalloc<float>(pos, pd, np, mass, mass_scalar, nmesh, out_scale, flags, out, as_stream(stream));
}
int ldl_ppaint_noalloc_f64(const double* pos, const ldl_pos_desc* pd, int64_t np, const double* mass, double mass_scalar,
                           const int64_t nmesh[3], double out_scale, int32_t flags, double* out, void* stream) {
    return paint_noalloc<double>(pos, pd, np, mass, mass_scalar, nmesh, out_scale, flags, out, as_stream(stream));
}

size_t ldl_paint_workspace_bytes(const int64_t nmesh[3], int32_t ncomp, int64_t np, int32_t flags) {
    Mesh m;
    if (check_mesh(nmesh, &m) || (ncomp != 1 && ncomp != 3) || np < 0) return 0;
    return paint_ws_bytes(m, ncomp, np, flags);
}

#define PAINT_ENTRY(SUF, T)                                                                                      \
    int ldl_ppaint_##SUF(const T* pos, const ldl_pos_desc* pd, int64_t np, const T* mass, double mass_scalar,    \
                         const int64_t nmesh[3], double out_scale, int32_t flags, T* out, void* ws, size_t wsb,  \
                         void* stream) {                                                                        \
        return paint_impl<T>(pos, pd, np, mass, mass_scalar, 1, nmesh, out_scale, flags, out, ws, wsb, as_stream(stream)); \
    }                                                                                                            \
    int ldl_ppaint3D_##SUF(const T* pos, const ldl_pos_desc* pd, int64_t np, const T* mass3, const int64_t nmesh[3], \
                           double out_scale, int32_t flags, T* out, void* ws, size_t wsb, void* stream) {        \
        return paint_impl<T>(pos, pd, np, mass3, 0.0, 3, nmesh, out_scale, flags, out, ws, wsb, as_stream(stream)); \
    }                                                                                                            \
    int ldl_preadout_##SUF(const T* pos, const ldl_pos_desc* pd, int64_t np, const T* field, const int64_t nmesh[3], \
                           int32_t vjpdim, T* out, void* stream) {                                               \
        return readout_impl<T>(pos, pd, np, field, nmesh, 0, vjpdim, nullptr, out, as_stream(stream));          \
    }                                                                                                            \
    int ldl_preadout3D_##SUF(const T* pos, const ldl_pos_desc* pd, int64_t np, const T* field,                  \
                             const int64_t nmesh[3], int32_t vjp, const T* pscale, T* out, void* stream) {       \
        if (!vjp && pscale) return fail(LDL_EINVAL, "pscale is only defined for vjp != 0");                     \
        return readout_impl<T>(pos, pd, np, field, nmesh, vjp ? 2 : 1, -1, pscale, out, as_stream(stream));     \
    }                                                                                                            \
    int ldl_readout_grad3_##SUF(const T* pos, const ldl_pos_desc* pd, int64_t np, const T* fields3, const T* v3, \
                                const int64_t nmesh[3], T* out, void* stream) {                                  \
        return readout_impl<T>(pos, pd, np, fields3, nmesh, 3, -1, v3, out, as_stream(stream));                 \
    }

PAINT_ENTRY(f32, float)
PAINT_ENTRY(f64, double)

}  // extern "C"
