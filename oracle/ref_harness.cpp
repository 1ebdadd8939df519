// Harness over the reference's own hot-path sources that build without MPI (TEST
// INFRASTRUCTURE ONLY): include/Differ.hpp (limiters, MUSCL_1, MUSCL_2), src/TimeAdv.cpp
// (TimeAdv::EE) and src/Function.cpp (Function::sum). It only marshals plain arrays into the
// reference's Array<> containers and calls the reference functions unchanged; the sources are
// compiled where they lie under the reference tree by oracle/Makefile.ref.
#include <cstring>

#include "Differ.hpp"
#include "Function.hpp"
#include "TimeAdv.hpp"

namespace {
void fill1(Array<double, 1> &a, const double *src, int n) {
    for (int i = 0; i < n; i++) a(i) = src[i];
}
}  // namespace

extern "C" {

// MUSCL_1 / MUSCL_2 (include/Differ.hpp:110-136, 149-238) on a ghost-inclusive 3-D field in
// reference layout; FLR(i,j,k,0/1) is zero-initialised before the call.
void ref_muscl(int order, int dir, int ni, int nj, int nk, int bc, const double *fi_in, double *flr_out, int limit) {
    const int NX = ni + 2 * bc, NY = nj + 2 * bc, NZ = nk + 2 * bc;
    Array<double, 1> xn(NX), yn(NY), zn(NZ);
    xn.Fill(0.0), yn.Fill(0.0), zn.Fill(0.0);
    Array<double, 3> fi(NX, NY, NZ);
    Array<double, 4> FLR(NX, NY, NZ, 2);
    std::memcpy(fi.Getbuf(), fi_in, sizeof(double) * NX * NY * NZ);
    FLR.Fill(0.0);
    if (order == 1)
        MUSCL_1(dir, xn, yn, zn, FLR, fi, bc, limit);
    else
        MUSCL_2(dir, xn, yn, zn, FLR, fi, bc, limit);
    std::memcpy(flr_out, FLR.Getbuf(), sizeof(double) * NX * NY * NZ * 2);
}

// TimeAdv::EE (src/TimeAdv.cpp:29-43); RHS is zero-initialised before the call.
void ref_ee(int ns, double dt, int ni, int nj, int nk, int bc, const double *x, const double *y, const double *z,
            const double *F, const double *G, const double *Q, double *rhs_out) {
    const int NX = ni + 2 * bc, NY = nj + 2 * bc, NZ = nk + 2 * bc, NC = ns + 4;
    Array<double, 1> xn(NX), yn(NY), zn(NZ);
    fill1(xn, x, NX), fill1(yn, y, NY), fill1(zn, z, NZ);
    Array<double, 4> f(NX, NY, NZ, NC), g(NX, NY, NZ, NC), q(NX, NY, NZ, NC), r(NX, NY, NZ, NC);
    const size_t n = sizeof(double) * NX * NY * NZ * NC;
    std::memcpy(f.Getbuf(), F, n);
    std::memcpy(g.Getbuf(), G, n);
    std::memcpy(q.Getbuf(), Q, n);
    r.Fill(0.0);
    TimeAdv t;
    t.EE(ns, dt, xn, yn, zn, bc, f, g, q, r);
    std::memcpy(rhs_out, r.Getbuf(), n);
}

// Function::sum (src/Function.cpp:17-42)
double ref_sum(int flag, const double *a, const double *b, int n) {
    Array<double, 1> A(n), B(n);
    fill1(A, a, n), fill1(B, b, n);
    Function f;
    return f.sum(flag, A, B);
}

}  // extern "C"
