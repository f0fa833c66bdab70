// Thin C driver around the reference's own CIC kernels, compiled from the
// read-only sources under the reference checkout (lib/parops_fast.h) by
// oracle/Makefile into oracle/_ref/.  TEST INFRASTRUCTURE ONLY: it produces
// golden vectors and cross-checks the CPU oracle; nothing in the product links it.
//
// Runs the reference's parallel drivers at MPI comm size 1 (singleton init):
//   ppaint     lib/parops_fast.h:483-531      ppaint3D   :678-725
//   preadout   lib/parops_fast.h:533-567      preadout3D :823-857
// Positions are passed in mesh units, as the reference's custom calls receive
// them (jaxpops_bind.py:163-166).  Each readout first calls ppaint with the same
// positions, because the reference readouts use the layout stored by the last
// paint (lib/cpu.cc:93).
#include <mpi.h>
#include <cstdint>
#include <cstring>
#include "parops_fast.h"

using namespace parops_jax;

namespace {
template <typename T> struct State { Layout layout; T* expos = nullptr; bool init = false; };
State<float> sf;
State<double> sd;

void ensure_mpi() {
    int flag = 0;
    MPI_Initialized(&flag);
    if (!flag) MPI_Init(nullptr, nullptr);
    // exchange1D/gather1D commit and free the predefined MPI_DOUBLE/MPI_FLOAT
    // (lib/parops_fast.h:231/:253 and :279/:301).  OpenMPI tolerates
    // it; MPICH reports an error, which the reference ignores once it is
    // returned instead of aborting.
    MPI_Comm_set_errhandler(MPI_COMM_WORLD, MPI_ERRORS_RETURN);
    MPI_Comm_set_errhandler(MPI_COMM_SELF, MPI_ERRORS_RETURN);
}

template <typename T> State<T>& st();
template <> State<float>& st<float>() { return sf; }
template <> State<double>& st<double>() { return sd; }

template <typename T, typename intT>
void do_paint(T* pos, int64_t np_, T* mass, const int64_t* n, T* out) {
    ensure_mpi();
    State<T>& s = st<T>();
    if (!s.init) { initLayout(&s.layout, 1); s.init = true; }
    intT nm[3] = {(intT)n[0], (intT)n[1], (intT)n[2]};
    intT ex[2] = {0, nm[0]}, ey[2] = {0, nm[1]}, ez[2] = {0, nm[2]};
    intT *pn = nm, *px = ex, *py = ey, *pz = ez;
    int64_t N = np_;
    bool use = false;
    MPI_Comm comm = MPI_COMM_WORLD;
    ppaint<T, intT>(pos, N, mass, pn, px, py, pz, &out, &s.layout, &s.expos, use, comm);
}

template <typename T, typename intT>
void do_readout(T* pos, int64_t np_, T* field, const int64_t* n, T* box, int vjpdim, T* out, T* scratch_mesh) {
    T* ones = new T[np_];
    for (int64_t i = 0; i < np_; i++) ones[i] = 1;
    do_paint<T, intT>(pos, np_, ones, n, scratch_mesh);
    delete[] ones;
    State<T>& s = st<T>();
    intT nm[3] = {(intT)n[0], (intT)n[1], (intT)n[2]};
    intT ex[2] = {0, nm[0]}, ey[2] = {0, nm[1]}, ez[2] = {0, nm[2]};
    intT *pn = nm, *px = ex, *py = ey, *pz = ez;
    int64_t N = np_;
    int32_t vd = vjpdim;
    MPI_Comm comm = MPI_COMM_WORLD;
    preadout<T, intT>(s.layout, s.expos, N, field, pn, box, px, py, pz, &out, comm, vd);
}

template <typename T, typename intT>
void do_readout3(T* pos, int64_t np_, T* field, const int64_t* n, T* box, int vjp, T* out, T* scratch_mesh) {
    T* ones = new T[np_];
    for (int64_t i = 0; i < np_; i++) ones[i] = 1;
    do_paint<T, intT>(pos, np_, ones, n, scratch_mesh);
    delete[] ones;
    State<T>& s = st<T>();
    intT nm[3] = {(intT)n[0], (intT)n[1], (intT)n[2]};
    intT ex[2] = {0, nm[0]}, ey[2] = {0, nm[1]}, ez[2] = {0, nm[2]};
    intT *pn = nm, *px = ex, *py = ey, *pz = ez;
    int64_t N = np_;
    bool v = vjp != 0;
    MPI_Comm comm = MPI_COMM_WORLD;
    preadout3D<T, intT>(s.layout, s.expos, N, field, pn, box, px, py, pz, &out, comm, v);
}

template <typename T, typename intT>
void do_paint3(T* pos, int64_t np_, T* mass3, const int64_t* n, T* out) {
    ensure_mpi();
    State<T>& s = st<T>();
    if (!s.init) { initLayout(&s.layout, 1); s.init = true; }
    intT nm[3] = {(intT)n[0], (intT)n[1], (intT)n[2]};
    intT ex[2] = {0, nm[0]}, ey[2] = {0, nm[1]}, ez[2] = {0, nm[2]};
    intT *pn = nm, *px = ex, *py = ey, *pz = ez;
    int64_t N = np_;
    MPI_Comm comm = MPI_COMM_WORLD;
    ppaint3D<T, intT>(pos, N, mass3, pn, px, py, pz, &out, &s.layout, &s.expos, comm);
}
}  // namespace

extern "C" {
void ref_paint_f64(double* pos, int64_t np_, double* mass, const int64_t* n, double* out) { do_paint<double, int64_t>(pos, np_, mass, n, out); }
void ref_paint_f32(float* pos, int64_t np_, float* mass, const int64_t* n, float* out) { do_paint<float, int32_t>(pos, np_, mass, n, out); }
void ref_readout_f64(double* pos, int64_t np_, double* field, const int64_t* n, double* box, int vjpdim, double* out, double* scratch) { do_readout<double, int64_t>(pos, np_, field, n, box, vjpdim, out, scratch); }
void ref_readout_f32(float* pos, int64_t np_, float* field, const int64_t* n, float* box, int vjpdim, float* out, float* scratch) { do_readout<float, int32_t>(pos, np_, field, n, box, vjpdim, out, scratch); }
void ref_readout3_f64(double* pos, int64_t np_, double* field, const int64_t* n, double* box, int vjp, double* out, double* scratch) { do_readout3<double, int64_t>(pos, np_, field, n, box, vjp, out, scratch); }
void ref_readout3_f32(float* pos, int64_t np_, float* field, const int64_t* n, float* box, int vjp, float* out, float* scratch) { do_readout3<float, int32_t>(pos, np_, field, n, box, vjp, out, scratch); }
void ref_paint3_f64(double* pos, int64_t np_, double* mass3, const int64_t* n, double* out) { do_paint3<double, int64_t>(pos, np_, mass3, n, out); }
void ref_paint3_f32(float* pos, int64_t np_, float* mass3, const int64_t* n, float* out) { do_paint3<float, int32_t>(pos, np_, mass3, n, out); }
}
