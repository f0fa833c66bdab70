/* CPU oracle: CIC paint / readout / readout-gradient in plain C.
 *
 * TEST INFRASTRUCTURE ONLY.  Used by tests/ (as the checker), by
 * __graft_entry__.smoke() and by the cpu_baseline leg of bench.py.  The product
 * library never links it.
 *
 * Restates lpaint / lreadout / lpaint3D / lreadout3D
 * (lib/parops_fast.h:353-481, :615-676, :727-821) on a single rank: the MPI
 * decompose/exchange/gather steps (lib/parops_fast.h:51-351) are the identity
 * at comm size 1.  Floor indices are wrapped mod N (the reference drops a
 * particle whose mesh coordinate rounds to exactly N).
 *
 * Two flavours:
 *   * double precision deposits/gathers (oracle_paint_f64, ...), OpenMP over
 *     particles (atomic adds for paint): the CPU baseline;
 *   * bit-exact emulations of the GPU's 64-bit fixed-point paint
 *     (oracle_paint_fixed_{f32,f64}): same floating-point operation order,
 *     integer sums, so the GPU result must match them bit for bit.
 *
 * Pinned by: tests/test_oracle_cic.py (brute force vs the NumPy oracle, mass
 * conservation, adjointness, reference golden vectors from oracle/_ref).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static inline int64_t wrap64(int64_t i, int64_t n) {
    int64_t r = i % n;
    return r < 0 ? r + n : r;
}

/* numpy float remainder semantics (np.mod), then division by the cell size,
 * in the working precision (jaxpops_bind.py:163-164). */
static inline double mesh_coord_f64(double p, double box, double cell) {
    double r = fmod(p, box);
    if (r != 0.0 && ((r < 0.0) != (box < 0.0))) r += box;
    return r / cell;
}
static inline float mesh_coord_f32(float p, float box, float cell) {
    float r = fmodf(p, box);
    if (r != 0.0f && ((r < 0.0f) != (box < 0.0f))) r += box;
    return r / cell;
}

/* ---------------- double precision (baseline) ---------------- */
void oracle_paint_f64(const double* x, int64_t np_, const double* mass, double mscalar,
                      int64_t n0, int64_t n1, int64_t n2, double* out, int nthreads) {
    int64_t nm = n0 * n1 * n2;
    memset(out, 0, sizeof(double) * nm);
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
#endif
    for (int64_t p = 0; p < np_; p++) {
        double m = mass ? mass[p] : mscalar;
        double fx = floor(x[3 * p]), fy = floor(x[3 * p + 1]), fz = floor(x[3 * p + 2]);
        double dx = x[3 * p] - fx, dy = x[3 * p + 1] - fy, dz = x[3 * p + 2] - fz;
        double wx[2] = {1.0 - dx, dx}, wy[2] = {1.0 - dy, dy}, wz[2] = {1.0 - dz, dz};
        int64_t i = wrap64((int64_t)fx, n0), j = wrap64((int64_t)fy, n1), k = wrap64((int64_t)fz, n2);
        int64_t ii[2] = {i, (i + 1) % n0}, jj[2] = {j, (j + 1) % n1}, kk[2] = {k, (k + 1) % n2};
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++)
                for (int c = 0; c < 2; c++) {
                    double v = wx[a] * wy[b] * wz[c] * m;
#ifdef _OPENMP
#pragma omp atomic
#endif
                    out[(ii[a] * n1 + jj[b]) * n2 + kk[c]] += v;
                }
    }
}

/* readout (deriv_dim < 0), derivative readout (deriv_dim in 0..2, weights
 * -cellinv/+cellinv along that axis). out has stride ostride. */
void oracle_readout_f64(const double* x, int64_t np_, const double* field,
                        int64_t n0, int64_t n1, int64_t n2, int deriv_dim,
                        const double* cellinv, double* out, int64_t ostride, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
#endif
    for (int64_t p = 0; p < np_; p++) {
        double fx = floor(x[3 * p]), fy = floor(x[3 * p + 1]), fz = floor(x[3 * p + 2]);
        double d[3] = {x[3 * p] - fx, x[3 * p + 1] - fy, x[3 * p + 2] - fz};
        double w[3][2];
        for (int a = 0; a < 3; a++) {
            if (a == deriv_dim) { w[a][0] = -cellinv[a]; w[a][1] = cellinv[a]; }
            else { w[a][0] = 1.0 - d[a]; w[a][1] = d[a]; }
        }
        int64_t i = wrap64((int64_t)fx, n0), j = wrap64((int64_t)fy, n1), k = wrap64((int64_t)fz, n2);
        int64_t ii[2] = {i, (i + 1) % n0}, jj[2] = {j, (j + 1) % n1}, kk[2] = {k, (k + 1) % n2};
        double acc = 0.0;
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++)
                for (int c = 0; c < 2; c++)
                    acc += w[0][a] * w[1][b] * w[2][c] * field[(ii[a] * n1 + jj[b]) * n2 + kk[c]];
        out[p * ostride] = acc;
    }
}

void oracle_mesh_units_f64(const double* pos, int64_t np_, const double* box, const int64_t* nmesh, double* x) {
    double cell[3];
    for (int d = 0; d < 3; d++) cell[d] = box[d] / (double)nmesh[d];
    for (int64_t p = 0; p < np_; p++)
        for (int d = 0; d < 3; d++) x[3 * p + d] = mesh_coord_f64(pos[3 * p + d], box[d], cell[d]);
}

/* ---------------- fixed-point paint emulation ----------------
 * Per particle: CIC fractions in the working precision T, then the rounded
 * mass M = rint(m 2^s) (m 2^s in double) split hierarchically: X1 =
 * trunc(f_x M) clamped between 0 and M, X0 = M - X1; each X_a split along y
 * the same way; each of those along z; corner k = 4a+2b+c receives the z
 * part c of (a, b).  Products in the working precision T.  Every deposit has the sign of m and
 * the deposits sum exactly to M.  The accumulator is summed in int64 (exact,
 * order free).  s comes from the caller (FP32: 30 - e, FP64: 61 - e -
 * ceil(log2 Np), with max|m| < 2^e).  kind 0: pos physical (wrapped and
 * divided in T);  kind 1: pos is a displacement in mesh units from the
 * lattice point idx*step.                                                  */
static int64_t split_part_f32(float fr, int64_t V) {
    int64_t v = (int64_t)(fr * (float)V);
    if (V >= 0) return v > V ? V : v;
    return v < V ? V : v;
}
static int64_t split_part_f64(double fr, int64_t V) {
    int64_t v = (int64_t)(fr * (double)V);
    if (V >= 0) return v > V ? V : v;
    return v < V ? V : v;
}

#define DEPOSIT_BODY(SPLIT)                                                        \
    int64_t M = llrint(ldexp((double)m, s));                                       \
    int64_t ii[2] = {i, (i + 1) % n0}, jj[2] = {j, (j + 1) % n1}, kk[2] = {k, (k + 1) % n2}; \
    int64_t x1 = SPLIT(fx_, M), xa[2] = {M - x1, x1};                              \
    for (int a = 0; a < 2; a++) {                                                  \
        int64_t y1 = SPLIT(fy_, xa[a]), yb[2] = {xa[a] - y1, y1};                  \
        for (int b = 0; b < 2; b++) {                                              \
            int64_t z1 = SPLIT(fz_, yb[b]);                                        \
            acc[(ii[a] * n1 + jj[b]) * n2 + kk[0]] += yb[b] - z1;                  \
            acc[(ii[a] * n1 + jj[b]) * n2 + kk[1]] += z1;                          \
        }                                                                          \
    }

void oracle_paint_fixed_f32(const float* pos, int64_t np_, int kind, const double* boxd,
                            const int64_t* lat, const double* lstep, const float* mass, double mscalar,
                            int64_t n0, int64_t n1, int64_t n2, int s, int64_t* acc) {
    float box[3], cell[3];
    int64_t nm[3] = {n0, n1, n2};
    for (int d = 0; d < 3; d++) { box[d] = (float)boxd[d]; cell[d] = box[d] / (float)nm[d]; }
    memset(acc, 0, sizeof(int64_t) * n0 * n1 * n2);
    for (int64_t p = 0; p < np_; p++) {
        float m = mass ? mass[p] : (float)mscalar;
        float xs[3];
        int64_t base[3] = {0, 0, 0};
        if (kind == 0) {
            for (int d = 0; d < 3; d++) xs[d] = mesh_coord_f32(pos[3 * p + d], box[d], cell[d]);
        } else {
            int64_t li[3] = {p / (lat[1] * lat[2]), (p / lat[2]) % lat[1], p % lat[2]};
            for (int d = 0; d < 3; d++) {
                double b = (double)li[d] * lstep[d];
                double bf = floor(b);
                base[d] = (int64_t)bf;
                xs[d] = pos[3 * p + d] + (float)(b - bf);
            }
        }
        float flx = floorf(xs[0]), fly = floorf(xs[1]), flz = floorf(xs[2]);
        float fx_ = xs[0] - flx, fy_ = xs[1] - fly, fz_ = xs[2] - flz;
        int64_t i = wrap64(base[0] + (int64_t)flx, n0), j = wrap64(base[1] + (int64_t)fly, n1),
                k = wrap64(base[2] + (int64_t)flz, n2);
        DEPOSIT_BODY(split_part_f32)
    }
}

void oracle_paint_fixed_f64(const double* pos, int64_t np_, int kind, const double* box,
                            const int64_t* lat, const double* lstep, const double* mass, double mscalar,
                            int64_t n0, int64_t n1, int64_t n2, int s, int64_t* acc) {
    double cell[3];
    int64_t nm[3] = {n0, n1, n2};
    for (int d = 0; d < 3; d++) cell[d] = box[d] / (double)nm[d];
    memset(acc, 0, sizeof(int64_t) * n0 * n1 * n2);
    for (int64_t p = 0; p < np_; p++) {
        double m = mass ? mass[p] : mscalar;
        double xs[3];
        int64_t base[3] = {0, 0, 0};
        if (kind == 0) {
            for (int d = 0; d < 3; d++) xs[d] = mesh_coord_f64(pos[3 * p + d], box[d], cell[d]);
        } else {
            int64_t li[3] = {p / (lat[1] * lat[2]), (p / lat[2]) % lat[1], p % lat[2]};
            for (int d = 0; d < 3; d++) {
                double b = (double)li[d] * lstep[d];
                double bf = floor(b);
                base[d] = (int64_t)bf;
                xs[d] = pos[3 * p + d] + (b - bf);
            }
        }
        double flx = floor(xs[0]), fly = floor(xs[1]), flz = floor(xs[2]);
        double fx_ = xs[0] - flx, fy_ = xs[1] - fly, fz_ = xs[2] - flz;
        int64_t i = wrap64(base[0] + (int64_t)flx, n0), j = wrap64(base[1] + (int64_t)fly, n1),
                k = wrap64(base[2] + (int64_t)flz, n2);
        DEPOSIT_BODY(split_part_f64)
    }
}
