// CIC paint (particle -> mesh) and readout (mesh -> particle) on gfx950.
//
// Algorithm (lib/parops_fast.h:353-481, :615-821, single rank): x = mesh
// coordinate, i = floor(x) wrapped mod N, f = x - i; corner (i+a, j+b, k+c)
// gets weight w_x[a] w_y[b] w_z[c] with w[0] = 1 - f, w[1] = f.
//
// Paint accumulates in 64-bit fixed point: each deposit is rounded once to
// an integer multiple of 2^-s and summed with integer atomics, so the density
// is bit-identical for any particle order or scheduling.  s is chosen per call
// from max|m| and Np so that no cell (nor the total) can exceed 2^61 * 2^-s.
// The eighth corner receives rint(m 2^s) minus the other seven deposits, so
// the painted total equals the sum of the rounded masses exactly.
#include "cic.h"

namespace ldl {

// ---------------------------------------------------------- host drivers
static constexpr int kBlock = 256;
static constexpr int64_t kHdrBytes = 256;

size_t paint_ws_bytes(const Mesh& m, int nc) { return (size_t)kHdrBytes + (size_t)nc * (size_t)m.cells() * 8; }

template <typename T>
int paint_impl(const T* pos, const ldl_pos_desc* pdesc, int64_t np, const T* mass, double mscalar, int nc,
               const int64_t nmesh[3], double out_scale, int32_t flags, T* out, void* ws, size_t ws_bytes,
               hipStream_t st) {
    Mesh m;
    int rc = check_mesh(nmesh, &m);
    if (rc) return rc;
    if (np < 0) return fail(LDL_EINVAL, "np must be >= 0");
    if (!out) return fail(LDL_EINVAL, "out is NULL");
    if (np > 0 && !pos) return fail(LDL_EINVAL, "pos is NULL");
    if (nc == 3 && !mass) return fail(LDL_EINVAL, "ppaint3D needs a (Np,3) mass array");
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
    if (!ws || ws_bytes < paint_ws_bytes(m, nc)) return fail(LDL_EINVAL, "paint workspace too small (see ldl_paint_workspace_bytes)");
    auto* hdr = reinterpret_cast<unsigned long long*>(ws);
    auto* acc = reinterpret_cast<unsigned long long*>(reinterpret_cast<char*>(ws) + kHdrBytes);
    // header: max |m| bits; for a scalar mass it is known on the host
    unsigned long long hbits = 0;
    if (!mass) {
        double a = std::fabs((double)(T)mscalar);
        std::memcpy(&hbits, &a, 8);
    }
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, st, hdr, (int64_t)(kHdrBytes / 8), hbits);
    if (!(flags & 2))  // LDL_PAINT_WS_ZEROED (internal): accumulator already zero
        hipLaunchKernelGGL(k_fill_u64, dim3(gm), dim3(kBlock), 0, st, acc, tot, 0ull);
    if (mass && np > 0) {
        if (nc == 1) hipLaunchKernelGGL((k_absmax<T, 1>), dim3(grid_for(np, kBlock, 4096)), dim3(kBlock), 0, st, mass, np, hdr);
        else hipLaunchKernelGGL((k_absmax<T, 3>), dim3(grid_for(np, kBlock, 4096)), dim3(kBlock), 0, st, mass, np, hdr);
    }
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
    const unsigned g = grid_for(np, kBlock, 1 << 16);
    switch (kind) {
        case 0:
            if (vjpdim < -1 || vjpdim > 2) return fail(LDL_EINVAL, "vjpdim must be between -1 and 2");
            { KTimer _t("readout", st); hipLaunchKernelGGL(k_readout<T>, dim3(g), dim3(kBlock), 0, st, pos, pd, np, field, m, vjpdim, out); }
            break;
        case 1: { KTimer _t("readout", st); hipLaunchKernelGGL(k_readout3<T>, dim3(g), dim3(kBlock), 0, st, pos, pd, np, field, m, out); } break;
        case 2: { KTimer _t("readout", st); hipLaunchKernelGGL(k_readout_grad<T>, dim3(g), dim3(kBlock), 0, st, pos, pd, np, field, m, aux, out); } break;
        case 3:
            if (!aux) return fail(LDL_EINVAL, "v3 is NULL");
            { KTimer _t("readout", st); hipLaunchKernelGGL(k_readout_grad3<T>, dim3(g), dim3(kBlock), 0, st, pos, pd, np, field, aux, m, out); }
            break;
    }
    LDL_LAUNCH_CHECK();
    return LDL_OK;
}

template int paint_impl<float>(const float*, const ldl_pos_desc*, int64_t, const float*, double, int, const int64_t*, double, int32_t, float*, void*, size_t, hipStream_t);
template int paint_impl<double>(const double*, const ldl_pos_desc*, int64_t, const double*, double, int, const int64_t*, double, int32_t, double*, void*, size_t, hipStream_t);
template int readout_impl<float>(const float*, const ldl_pos_desc*, int64_t, const float*, const int64_t*, int, int, const float*, float*, hipStream_t);
template int readout_impl<double>(const double*, const ldl_pos_desc*, int64_t, const double*, const int64_t*, int, int, const double*, double*, hipStream_t);

}  // namespace ldl

using namespace ldl;

extern "C" {

size_t ldl_paint_workspace_bytes(const int64_t nmesh[3], int32_t ncomp) {
    Mesh m;
    if (check_mesh(nmesh, &m) || (ncomp != 1 && ncomp != 3)) return 0;
    return paint_ws_bytes(m, ncomp);
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
