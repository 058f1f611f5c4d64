/*
 * fe_dense.c - CPU oracle of FESolver (test infrastructure only): the reference's dense
 * arithmetic restated literally, to pin the compressed-row host solver of the library.
 *
 *   src/FESolver.cpp:47-127   constructor (ID, LM, quadrature), :207-287 computeNX / inverse
 *   src/FESolver.cpp:155-185  startAssembly / addKe / addFe, :434-485 preAssembly
 *   src/FESolver.cpp:290-364  solve / solveNonLinear, :367-431 buildJmatrix
 *   src/FESolver.cpp:488-526  buildF1Vector, :561-601 solveLinear, :604-636 computePhi / updateEf
 *   src/maths.cpp:48-70       det3, matVecMultiply, vecVecSubtract
 * K and J are dense neq x neq; every loop runs over all columns, as in the reference.
 * Build: -O3 -ffp-contract=off, never -ffast-math.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define OR_EPS0 8.8541878e-12
#define OR_QE 1.602e-19

typedef struct or_fe {
    int nn, ne, neq, reset_f1;
    int *ID, *LM, *con;
    double *NX, *detJ, *K, *J, *F0, *F1, *d, *g, *uh;
    double n0, phi0, kTe, l[2], W[2];
    int n_int;
} or_fe;

static double or_det3(double M[3][3]) {
    return M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) - M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
           M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
}

static void or_inverse(double M[3][3], double V[3][3]) {
    double a = M[0][0], b = M[0][1], c = M[0][2], d = M[1][0], e = M[1][1], f = M[1][2], g = M[2][0], h = M[2][1],
           i = M[2][2];
    V[0][0] = (e * i - f * h);
    V[1][0] = -(d * i - f * g);
    V[2][0] = (d * h - e * g);
    V[0][1] = -(b * i - c * h);
    V[1][1] = (a * i - c * g);
    V[2][1] = -(a * h - b * g);
    V[0][2] = (b * f - c * e);
    V[1][2] = -(a * f - c * d);
    V[2][2] = (a * e - b * d);
    double det = a * V[0][0] + b * V[1][0] + c * V[2][0];
    double Vmax = 0;
    for (int m = 0; m < 3; m++)
        for (int n = 0; n < 3; n++) Vmax = fabs(V[m][n]) > Vmax ? fabs(V[m][n]) : Vmax;
    double idet = 0;
    if (fabs(Vmax) / fabs(det) > 1e12) {
    } else
        idet = 1 / det;
    for (int r = 0; r < 3; r++)
        for (int s = 0; s < 3; s++) V[r][s] *= idet;
}

static double or_eval_na(int a, double xi, double eta, double zeta) {
    switch (a) {
        case 0: return xi;
        case 1: return eta;
        case 2: return zeta;
        case 3: return 1 - xi - eta - zeta;
        default: return 0;
    }
}

void or_fe_free(or_fe* s) {
    if (!s) return;
    free(s->ID); free(s->LM); free(s->con); free(s->NX); free(s->detJ); free(s->K); free(s->J);
    free(s->F0); free(s->F1); free(s->d); free(s->g); free(s->uh);
    free(s);
}

/* node_type: 0 NORMAL, 1 OPEN, 2 INLET, 3 FIXED (src/meshes.h:22) */
or_fe* or_fe_create(const double* pos, const int32_t* type, int32_t nn, const int32_t* con, int32_t ne, double n0,
                    double phi0, double kTe, double wall_potential, int reset_f1) {
    or_fe* s = (or_fe*)calloc(1, sizeof(or_fe));
    s->nn = nn;
    s->ne = ne;
    s->reset_f1 = reset_f1;
    s->n0 = n0;
    s->phi0 = phi0;
    s->kTe = kTe;
    s->ID = (int*)malloc(sizeof(int) * nn);
    int P = 0;
    for (int n = 0; n < nn; n++) s->ID[n] = (type[n] == 0 || type[n] == 1) ? P++ : -1;
    s->neq = P;
    const size_t q = (size_t)P;
    s->K = (double*)calloc(q * q + 1, sizeof(double));
    s->J = (double*)calloc(q * q + 1, sizeof(double));
    s->F0 = (double*)calloc(q + 1, sizeof(double));
    s->F1 = (double*)calloc(q + 1, sizeof(double));
    s->d = (double*)calloc(q + 1, sizeof(double));
    s->g = (double*)calloc(nn, sizeof(double));
    s->uh = (double*)calloc(nn, sizeof(double));
    s->con = (int*)malloc(sizeof(int) * 4 * ne);
    s->LM = (int*)malloc(sizeof(int) * 4 * ne);
    s->NX = (double*)malloc(sizeof(double) * 12 * ne);
    s->detJ = (double*)malloc(sizeof(double) * ne);
    for (int e = 0; e < ne; e++)
        for (int a = 0; a < 4; a++) {
            s->con[4 * e + a] = con[4 * e + a];
            s->LM[4 * e + a] = s->ID[con[4 * e + a]];
        }
    s->l[0] = -sqrt(1.0 / 3.0);
    s->l[1] = sqrt(1.0 / 3.0);
    s->W[0] = 1;
    s->W[1] = 1;
    s->n_int = 2;
    /* computeNX */
    double na_xi[4][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {-1, -1, -1}};
    for (int e = 0; e < ne; e++) {
        double x[4][3];
        for (int a = 0; a < 4; a++)
            for (int d = 0; d < 3; d++) x[a][d] = pos[3 * con[4 * e + a] + d];
        double x_xi[3][3];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                x_xi[i][j] = 0;
                for (int a = 0; a < 4; a++) x_xi[i][j] += na_xi[a][j] * x[a][i];
            }
        s->detJ[e] = or_det3(x_xi);
        double xi_x[3][3];
        or_inverse(x_xi, xi_x);
        for (int a = 0; a < 4; a++)
            for (int d = 0; d < 3; d++) {
                double* nx = &s->NX[12 * e + 3 * a + d];
                *nx = 0;
                for (int k = 0; k < 3; k++) *nx += na_xi[a][k] * xi_x[k][d];
            }
    }
    /* g (src/mini-pic.cpp:97-102) */
    for (int n = 0; n < nn; n++) s->g[n] = type[n] == 2 ? 0 : (type[n] == 3 ? -wall_potential : 0);
    /* preAssembly */
    for (int e = 0; e < ne; e++) {
        double ke[4][4];
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++) {
                ke[a][b] = 0;
                for (int k = 0; k < s->n_int; k++)
                    for (int j = 0; j < s->n_int; j++)
                        for (int i = 0; i < s->n_int; i++) {
                            double dot = 0;
                            for (int d = 0; d < 3; d++) dot += s->NX[12 * e + 3 * a + d] * s->NX[12 * e + 3 * b + d];
                            ke[a][b] += dot * s->detJ[e] * s->W[i] * s->W[j] * s->W[k];
                        }
            }
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++) {
                int Pp = s->LM[4 * e + a], Q = s->LM[4 * e + b];
                if (Pp < 0 || Q < 0) continue;
                s->K[(size_t)Pp * q + Q] += ke[a][b];
            }
        double fe[4];
        for (int a = 0; a < 4; a++) {
            double fh = 0, fg = 0;
            for (int b = 0; b < 4; b++) {
                double gb = s->g[s->con[4 * e + b]];
                fg -= ke[a][b] * gb;
            }
            fe[a] = fh + fg;
        }
        for (int a = 0; a < 4; a++) {
            int Pp = s->LM[4 * e + a];
            if (Pp < 0) continue;
            s->F0[Pp] += fe[a];
        }
    }
    return s;
}

static void or_build_f1(or_fe* s, const double* ion_den) {
    if (s->reset_f1) memset(s->F1, 0, sizeof(double) * (size_t)s->neq);
    double* f = (double*)calloc((size_t)s->neq + 1, sizeof(double));
    for (int n = 0; n < s->nn; n++) {
        int A = s->ID[n];
        if (A < 0) continue;
        f[A] = (OR_QE / OR_EPS0) * (ion_den[n] + s->n0 * exp((s->d[A] - s->phi0) / s->kTe));
    }
    for (int e = 0; e < s->ne; e++) {
        double fe[4] = {0, 0, 0, 0};
        for (int a = 0; a < 4; a++) {
            double ff = 0;
            int A = s->LM[4 * e + a];
            if (A >= 0) {
                for (int k = 0; k < s->n_int; k++)
                    for (int j = 0; j < s->n_int; j++)
                        for (int i = 0; i < s->n_int; i++) {
                            double xi = 0.5 * (s->l[i] + 1), eta = 0.5 * (s->l[j] + 1), zeta = 0.5 * (s->l[k] + 1);
                            double Na = or_eval_na(a, xi, eta, zeta);
                            ff += f[A] * Na * s->detJ[e] * s->W[i] * s->W[j] * s->W[k];
                        }
                ff *= (1.0 / 8.0);
                fe[a] = ff;
            }
        }
        for (int a = 0; a < 4; a++) {
            int Pp = s->LM[4 * e + a];
            if (Pp < 0) continue;
            s->F1[Pp] += fe[a];
        }
    }
    free(f);
}

static void or_build_j(or_fe* s) {
    const size_t q = (size_t)s->neq;
    double* fp_term = (double*)malloc(sizeof(double) * (q + 1));
    double* FP = (double*)calloc(q + 1, sizeof(double));
    for (int n = 0; n < s->neq; n++) fp_term[n] = -OR_QE / OR_EPS0 * s->n0 * exp((s->d[n] - s->phi0) / s->kTe) * (1 / s->kTe);
    for (size_t i = 0; i < q; i++)
        for (size_t j = 0; j < q; j++) s->J[i * q + j] = s->K[i * q + j];
    for (int e = 0; e < s->ne; e++) {
        double fe[4];
        for (int a = 0; a < 4; a++) {
            double ff = 0;
            int A = s->LM[4 * e + a];
            if (A >= 0) {
                for (int k = 0; k < s->n_int; k++)
                    for (int j = 0; j < s->n_int; j++)
                        for (int i = 0; i < s->n_int; i++) {
                            double xi = 0.5 * (s->l[i] + 1), eta = 0.5 * (s->l[j] + 1), zeta = 0.5 * (s->l[k] + 1);
                            double Na = or_eval_na(a, xi, eta, zeta);
                            ff += fp_term[A] * Na * s->detJ[e] * s->W[i] * s->W[j] * s->W[k];
                        }
                ff *= (1.0 / 8.0);
            }
            fe[a] = ff;
        }
        for (int a = 0; a < 4; a++) {
            int Pp = s->LM[4 * e + a];
            if (Pp < 0) continue;
            FP[Pp] += fe[a];
        }
    }
    for (size_t u = 0; u < q; u++) s->J[u * q + u] -= FP[u];
    free(fp_term);
    free(FP);
}

static void or_gauss_seidel(or_fe* s, const double* A, double* x, const double* b) {
    const size_t q = (size_t)s->neq;
    for (int it = 0; it < 10000; it++) {
        for (size_t u = 0; u < q; u++) {
            if (fabs(A[u * q + u]) < 1e-12) continue;
            double sum = 0;
            for (size_t v = 0; v < q; v++) {
                if (u == v) continue;
                sum += A[u * q + v] * x[v];
            }
            x[u] = (b[u] - sum) / A[u * q + u];
        }
        if (it % 25 == 0) {
            double L = 0;
            for (size_t u = 0; u < q; u++) {
                double sum = 0;
                for (size_t v = 0; v < q; v++) sum += A[u * q + v] * x[v];
                double r = b[u] - sum;
                L += r * r;
            }
            if (sqrt(L) / s->neq < 1e-4) break;
        }
    }
}

/* computePhi with method 0 (NonLinear) or 1 (GaussSeidel).  Returns 0, or -1 when the Newton
 * iteration has not converged in 10 steps (the reference exits). */
int or_fe_compute_phi(or_fe* s, const double* ion_den, int method) {
    const size_t q = (size_t)s->neq;
    double* y = (double*)calloc(q + 1, sizeof(double));
    double* G = (double*)calloc(q + 1, sizeof(double));
    int rc = 0;
    int iters = method == 0 ? 10 : 1, converged = method != 0;
    for (int it = 0; it < iters; it++) {
        or_build_f1(s, ion_den);
        for (size_t i = 0; i < q; i++) {
            G[i] = 0;
            for (size_t j = 0; j < q; j++) G[i] += s->K[i * q + j] * s->d[j];
        }
        for (size_t i = 0; i < q; i++) G[i] = G[i] - s->F0[i];
        for (size_t i = 0; i < q; i++) G[i] = G[i] - s->F1[i];
        or_build_j(s);
        or_gauss_seidel(s, s->J, y, G);
        for (size_t n = 0; n < q; n++) s->d[n] -= y[n];
        if (method != 0) break;
        double sum = 0;
        for (size_t u = 0; u < q; u++) sum += y[u] * y[u];
        if (sqrt(sum) / s->neq < 1e-2) {
            converged = 1;
            break;
        }
    }
    if (!converged) rc = -1;
    for (int n = 0; n < s->nn; n++) {
        s->uh[n] = s->g[n];
        int A = s->ID[n];
        if (A >= 0) s->uh[n] += s->d[A];
    }
    free(y);
    free(G);
    return rc;
}

/* uh[nn] and, when ef is not NULL, updateEf's ef[ne][3]. */
void or_fe_get(const or_fe* s, double* uh, double* ef) {
    if (uh) memcpy(uh, s->uh, sizeof(double) * (size_t)s->nn);
    if (!ef) return;
    for (int e = 0; e < s->ne; e++) {
        for (int d = 0; d < 3; d++) ef[3 * e + d] = 0;
        for (int a = 0; a < 4; a++) {
            int A = s->con[4 * e + a];
            for (int d = 0; d < 3; d++) ef[3 * e + d] -= s->NX[12 * e + 3 * a + d] * s->uh[A];
        }
    }
}

int32_t or_fe_neq(const or_fe* s) { return s->neq; }
