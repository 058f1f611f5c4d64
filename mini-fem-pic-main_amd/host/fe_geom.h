// Per-element linear shape-function gradients, as FESolver::computeNX evaluates them
// (src/FESolver.cpp:207-246, with inverse() at :249-287).  Shared by the field-from-potential
// helper and the host FE solver so both use the same arithmetic.  Built with
// -ffp-contract=off: every product and sum separately rounded, in the reference's order.
#pragma once
#include <cmath>

namespace mphost {

// src/maths.cpp:48-53
inline double det3(const double M[3][3]) {
    return M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) -
           M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
           M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
}

// x[a][d]: coordinates of the element's nodes in con[] order.  NX[a][d] = dN_a/dx_d;
// *detJ = det3(dx/dxi).  An inverse that fails the reference's conditioning test
// (|adj|_max / |det| > 1e12) is zero, as there.
inline void element_nx(const double x[4][3], double NX[4][3], double* detJ) {
    static const double na_xi[4][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {-1, -1, -1}};
    double x_xi[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            x_xi[i][j] = 0;
            for (int a = 0; a < 4; a++) x_xi[i][j] += na_xi[a][j] * x[a][i];
        }
    if (detJ) *detJ = det3(x_xi);
    const double a = x_xi[0][0], b = x_xi[0][1], c = x_xi[0][2];
    const double d = x_xi[1][0], e = x_xi[1][1], f = x_xi[1][2];
    const double g = x_xi[2][0], h = x_xi[2][1], i = x_xi[2][2];
    double V[3][3];
    V[0][0] = (e * i - f * h);
    V[1][0] = -(d * i - f * g);
    V[2][0] = (d * h - e * g);
    V[0][1] = -(b * i - c * h);
    V[1][1] = (a * i - c * g);
    V[2][1] = -(a * h - b * g);
    V[0][2] = (b * f - c * e);
    V[1][2] = -(a * f - c * d);
    V[2][2] = (a * e - b * d);
    const double det = a * V[0][0] + b * V[1][0] + c * V[2][0];
    double vmax = 0;
    for (int r = 0; r < 3; r++)
        for (int s = 0; s < 3; s++) vmax = std::fabs(V[r][s]) > vmax ? std::fabs(V[r][s]) : vmax;
    double idet = 0;
    if (!(std::fabs(vmax) / std::fabs(det) > 1e12)) idet = 1 / det;
    for (int r = 0; r < 3; r++)
        for (int s = 0; s < 3; s++) V[r][s] *= idet;
    for (int q = 0; q < 4; q++)
        for (int dd = 0; dd < 3; dd++) {
            NX[q][dd] = 0;
            for (int k = 0; k < 3; k++) NX[q][dd] += na_xi[q][k] * V[k][dd];
        }
}

// FESolver::updateEf for one element (src/FESolver.cpp:622-636): E = -sum_a NX[a] phi_a.
inline void element_ef(const double NX[4][3], const double phi[4], double out[3]) {
    for (int dd = 0; dd < 3; dd++) out[dd] = 0;
    for (int q = 0; q < 4; q++)
        for (int dd = 0; dd < 3; dd++) out[dd] -= NX[q][dd] * phi[q];
}

}  // namespace mphost
