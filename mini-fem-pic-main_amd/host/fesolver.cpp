// Host FE Poisson-Boltzmann solver of mini-pic's coupled loop (declared in mp_fem.h).
//
// Follows src/FESolver.cpp operation for operation (built with -ffp-contract=off).  The
// reference keeps K and J as dense neq x neq arrays and sums over every column; here they are
// compressed rows over the element-sharing pattern, columns ascending.  Every left-out entry
// is an exact zero, whose product with a finite value is a signed zero, and adding a signed
// zero to a running sum that starts at +0.0 leaves it unchanged; so each dot product below
// equals the dense one bit for bit while touching ~15 entries per row instead of neq.
#include "mp_fem.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <string>
#include <vector>

#include "fe_geom.h"

namespace {

thread_local std::string g_fe_err;

int fail(int code, const std::string& m) {
    g_fe_err = m;
    return code;
}

// src/FESolver.h:18-19
constexpr double kEPS0 = 8.8541878e-12;
constexpr double kQE = 1.602e-19;

// The solver's methods as the reference's trace rows name them (src/FESolver.cpp TRACE_ME):
// the direct method stands in for solveLinearLapack.
enum FeFn {
    kFnFESolver, kFnComputeNX, kFnPreAssembly, kFnComputePhi, kFnSolve, kFnSolveNonLinear, kFnBuildF1Vector, kFnBuildJmatrix,
    kFnSolveLinear, kFnSolveLinearDirect, kFnUpdateEf, kFnCount
};
const char* const kFnName[kFnCount] = {"FESolver",       "computeNX",     "preAssembly",  "computePhi",   "solve",
                                       "solveNonLinear", "buildF1Vector", "buildJmatrix", "solveLinear",
                                       "solveLinearDirect", "updateEf"};

static_assert(kFnCount == MP_FE_PROFILE_ENTRIES, "profile rows");

struct FnProfile {
    int64_t calls = 0;
    double total = 0.0, child = 0.0;  // inclusive seconds, seconds of traced callees
};

}  // namespace

struct mp_fe {
    int32_t n_nodes = 0, n_el = 0, neq = 0;
    std::vector<int32_t> ID;    // [n_nodes]: unknown index or -1 (Dirichlet "g" node)
    std::vector<int32_t> con;   // [n_el][4]
    std::vector<int32_t> LM;    // [n_el][4] = ID[con]
    std::vector<double> NX;     // [n_el][4][3]
    std::vector<double> detJ;   // [n_el]
    // K and J over the unknowns in compressed rows (columns ascending)
    std::vector<int64_t> rp;
    std::vector<int32_t> ci;
    std::vector<int64_t> dg;    // [neq]: position of the diagonal entry
    std::vector<int64_t> slot;  // [n_el][4][4]: position of (LM[a], LM[b]), -1 if a g node
    std::vector<double> K, J;
    std::vector<double> F0, F1, d, g, uh;
    double n0 = 0, phi0 = 0, kTe = 1;
    int32_t flags = 0;
    double l[2] = {0, 0}, W[2] = {1, 1};
    int n_int = 2;
    // wall time per method and the stack of open ones (the reference's TRACE_ME bookkeeping)
    mutable FnProfile prof[kFnCount];
    mutable std::vector<int> open;
};

namespace {

// Times one method call: inclusive time to the method, and to its caller's callee time.
struct Traced {
    const mp_fe& fe;
    int fn;
    std::chrono::steady_clock::time_point t0;
    Traced(const mp_fe& f, int k) : fe(f), fn(k), t0(std::chrono::steady_clock::now()) { fe.open.push_back(k); }
    ~Traced() {
        const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        fe.prof[fn].calls++;
        fe.prof[fn].total += dt;
        fe.open.pop_back();
        if (!fe.open.empty()) fe.prof[fe.open.back()].child += dt;
    }
};

}  // namespace

namespace {

// FESolver::evalNa (src/FESolver.cpp:188-196)
double eval_na(int a, double xi, double eta, double zeta) {
    switch (a) {
        case 0: return xi;
        case 1: return eta;
        case 2: return zeta;
        case 3: return 1 - xi - eta - zeta;
        default: return 0;
    }
}

// The element integral the reference evaluates for buildF1Vector and buildJmatrix:
// (sum over the 2x2x2 points of v*Na*detJ*Wi*Wj*Wk) * (1/8).
double quad_na(const mp_fe& fe, int e, int a, double v) {
    double ff = 0;
    for (int k = 0; k < fe.n_int; k++)
        for (int j = 0; j < fe.n_int; j++)
            for (int i = 0; i < fe.n_int; i++) {
                const double xi = 0.5 * (fe.l[i] + 1);
                const double eta = 0.5 * (fe.l[j] + 1);
                const double zeta = 0.5 * (fe.l[k] + 1);
                const double Na = eval_na(a, xi, eta, zeta);
                ff += v * Na * fe.detJ[e] * fe.W[i] * fe.W[j] * fe.W[k];
            }
    ff *= (1.0 / 8.0);
    return ff;
}

// preAssembly (src/FESolver.cpp:434-485) after startAssembly (:155-163): K, F0.
void pre_assembly(mp_fe& fe) {
    Traced tr_(fe, kFnPreAssembly);
    std::fill(fe.K.begin(), fe.K.end(), 0.0);
    std::fill(fe.F0.begin(), fe.F0.end(), 0.0);
    std::fill(fe.F1.begin(), fe.F1.end(), 0.0);
    for (int e = 0; e < fe.n_el; e++) {
        const double* nx = &fe.NX[(size_t)e * 12];
        double ke[4][4];
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++) {
                ke[a][b] = 0;
                for (int k = 0; k < fe.n_int; k++)
                    for (int j = 0; j < fe.n_int; j++)
                        for (int i = 0; i < fe.n_int; i++) {
                            double dot = 0;
                            for (int q = 0; q < 3; q++) dot += nx[3 * a + q] * nx[3 * b + q];
                            ke[a][b] += dot * fe.detJ[e] * fe.W[i] * fe.W[j] * fe.W[k];
                        }
            }
        // addKe (src/FESolver.cpp:166-175)
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++) {
                const int64_t s = fe.slot[(size_t)e * 16 + 4 * a + b];
                if (s >= 0) fe.K[s] += ke[a][b];
            }
        // F0 = fh + fg, fg = -sum_b ke[a][b] g_b; addFe (:178-185)
        for (int a = 0; a < 4; a++) {
            const double fh = 0;
            double fg = 0;
            for (int b = 0; b < 4; b++) {
                const double gb = fe.g[fe.con[(size_t)e * 4 + b]];
                fg -= ke[a][b] * gb;
            }
            const int P = fe.LM[(size_t)e * 4 + a];
            if (P >= 0) fe.F0[P] += fh + fg;
        }
    }
}

// buildF1Vector (src/FESolver.cpp:488-526): adds this call's "ff" term into F1 (cleared
// first under MP_FE_RESET_F1).
void build_f1(mp_fe& fe, const double* ion_den) {
    Traced tr_(fe, kFnBuildF1Vector);
    if (fe.flags & MP_FE_RESET_F1) std::fill(fe.F1.begin(), fe.F1.end(), 0.0);
    std::vector<double> f((size_t)fe.neq);
    for (int n = 0; n < fe.n_nodes; n++) {
        const int A = fe.ID[n];
        if (A < 0) continue;
        f[A] = (kQE / kEPS0) * (ion_den[n] + fe.n0 * std::exp((fe.d[A] - fe.phi0) / fe.kTe));
    }
    for (int e = 0; e < fe.n_el; e++)
        for (int a = 0; a < 4; a++) {
            const int A = fe.LM[(size_t)e * 4 + a];
            if (A >= 0) fe.F1[A] += quad_na(fe, e, a, f[A]);
        }
}

// y = K x (matVecMultiply, src/maths.cpp:57-64) over the compressed rows.
void mat_vec(const mp_fe& fe, const std::vector<double>& A, const std::vector<double>& x, std::vector<double>& y) {
    for (int i = 0; i < fe.neq; i++) {
        double s = 0;
        for (int64_t p = fe.rp[i]; p < fe.rp[i + 1]; p++) s += A[p] * x[fe.ci[p]];
        y[i] = s;
    }
}

// buildJmatrix (src/FESolver.cpp:367-431): J = K - diag(FP).
void build_j(mp_fe& fe) {
    Traced tr_(fe, kFnBuildJmatrix);
    std::vector<double> fp_term((size_t)fe.neq), FP((size_t)fe.neq, 0.0);
    for (int n = 0; n < fe.neq; n++)
        fp_term[n] = -kQE / kEPS0 * fe.n0 * std::exp((fe.d[n] - fe.phi0) / fe.kTe) * (1 / fe.kTe);
    fe.J = fe.K;
    for (int e = 0; e < fe.n_el; e++) {
        double fev[4];
        for (int a = 0; a < 4; a++) {
            const int A = fe.LM[(size_t)e * 4 + a];
            fev[a] = A >= 0 ? quad_na(fe, e, a, fp_term[A]) : 0.0;
        }
        for (int a = 0; a < 4; a++) {
            const int P = fe.LM[(size_t)e * 4 + a];
            if (P >= 0) FP[P] += fev[a];
        }
    }
    for (int u = 0; u < fe.neq; u++) fe.J[fe.dg[u]] -= FP[u];
}

// solveLinear (src/FESolver.cpp:561-601): Gauss-Seidel on A x = b, x holds the initial guess;
// residual checked every 25 sweeps, tol 1e-4 on sqrt(sum r^2) / neq, at most 10000 sweeps.
void gauss_seidel(const mp_fe& fe, const std::vector<double>& A, std::vector<double>& x,
                  const std::vector<double>& b, mp_fe_stats& st) {
    Traced tr_(fe, kFnSolveLinear);
    const double tol = 1e-4;
    for (int u = 0; u < fe.neq; u++)
        if (std::fabs(A[fe.dg[u]]) < 1e-12) st.zero_diagonal++;
    bool converged = false;
    int it;
    for (it = 0; it < 10000; it++) {
        for (int u = 0; u < fe.neq; u++) {
            const double auu = A[fe.dg[u]];
            if (std::fabs(auu) < 1e-12) continue;
            double sum = 0;
            for (int64_t p = fe.rp[u]; p < fe.rp[u + 1]; p++) {
                const int v = fe.ci[p];
                if (v == u) continue;
                sum += A[p] * x[v];
            }
            x[u] = (b[u] - sum) / auu;
        }
        if (it % 25 == 0) {
            double L = 0;
            for (int u = 0; u < fe.neq; u++) {
                double sum = 0;
                for (int64_t p = fe.rp[u]; p < fe.rp[u + 1]; p++) sum += A[p] * x[fe.ci[p]];
                const double r = b[u] - sum;
                L += r * r;
            }
            const double L2 = std::sqrt(L) / fe.neq;
            if (L2 < tol) {
                converged = true;
                break;
            }
        }
    }
    st.gs_sweeps += converged ? it + 1 : it;
    if (!converged) st.gs_unconverged++;
}

// In place of solveLinearLapack's dsysv_: dense Cholesky of the symmetric definite J.  With
// the reference's node order in every tet detJ is negative, so K and J are negative definite:
// the factorisation runs on -J (and -b) then.
bool cholesky_solve(const mp_fe& fe, const std::vector<double>& A, std::vector<double>& x,
                    const std::vector<double>& b_in) {
    Traced tr_(fe, kFnSolveLinearDirect);
    const int n = fe.neq;
    if (n == 0) return true;
    const double sg = A[fe.dg[0]] < 0 ? -1.0 : 1.0;
    std::vector<double> L((size_t)n * n, 0.0), b((size_t)n);
    for (int i = 0; i < n; i++) {
        b[i] = sg * b_in[i];
        for (int64_t p = fe.rp[i]; p < fe.rp[i + 1]; p++) L[(size_t)i * n + fe.ci[p]] = sg * A[p];
    }
    for (int j = 0; j < n; j++) {
        double s = L[(size_t)j * n + j];
        for (int k = 0; k < j; k++) s -= L[(size_t)j * n + k] * L[(size_t)j * n + k];
        if (!(s > 0)) return false;
        const double ljj = std::sqrt(s);
        L[(size_t)j * n + j] = ljj;
        for (int i = j + 1; i < n; i++) {
            double t = L[(size_t)i * n + j];
            for (int k = 0; k < j; k++) t -= L[(size_t)i * n + k] * L[(size_t)j * n + k];
            L[(size_t)i * n + j] = t / ljj;
        }
    }
    std::vector<double> z((size_t)n);
    for (int i = 0; i < n; i++) {
        double t = b[i];
        for (int k = 0; k < i; k++) t -= L[(size_t)i * n + k] * z[k];
        z[i] = t / L[(size_t)i * n + i];
    }
    for (int i = n - 1; i >= 0; i--) {
        double t = z[i];
        for (int k = i + 1; k < n; k++) t -= L[(size_t)k * n + i] * x[k];
        x[i] = t / L[(size_t)i * n + i];
    }
    return true;
}

// G = K d - F0 - F1
void residual(const mp_fe& fe, std::vector<double>& G) {
    mat_vec(fe, fe.K, fe.d, G);
    for (int i = 0; i < fe.neq; i++) G[i] = G[i] - fe.F0[i];
    for (int i = 0; i < fe.neq; i++) G[i] = G[i] - fe.F1[i];
}

// solve (src/FESolver.cpp:290-321): y starts at zero and carries over between Newton
// iterations as the Gauss-Seidel initial guess; d -= y after each linear solve.
int solve(mp_fe& fe, const double* ion_den, int32_t method, mp_fe_stats& st, std::string& err) {
    Traced tr_(fe, kFnSolve);
    std::vector<double> y((size_t)fe.neq, 0.0), G((size_t)fe.neq);
    if (method == MP_FE_NONLINEAR) {
        Traced tr_nr(fe, kFnSolveNonLinear);
        double L2 = 0;
        for (int it = 0; it < 10; it++) {
            build_f1(fe, ion_den);
            residual(fe, G);
            build_j(fe);
            gauss_seidel(fe, fe.J, y, G, st);
            for (int n = 0; n < fe.neq; n++) fe.d[n] -= y[n];
            double sum = 0;
            for (int u = 0; u < fe.neq; u++) sum += y[u] * y[u];
            L2 = std::sqrt(sum) / fe.neq;
            st.nr_iterations = it + 1;
            st.nr_l2 = L2;
            if (L2 < 1e-2) return MP_OK;
        }
        err = "NR failed to converge, L2 = " + std::to_string(L2);
        return MP_ERR_STATE;
    }
    build_f1(fe, ion_den);
    residual(fe, G);
    build_j(fe);
    if (method == MP_FE_GAUSS_SEIDEL) {
        gauss_seidel(fe, fe.J, y, G, st);
    } else if (!cholesky_solve(fe, fe.J, y, G)) {
        err = "mp_fe_compute_phi: J is not definite";
        return MP_ERR_STATE;
    }
    for (int n = 0; n < fe.neq; n++) fe.d[n] -= y[n];
    st.nr_iterations = 1;
    return MP_OK;
}

// The constructor, startAssembly and preAssembly of src/FESolver.cpp:48-128, 155-164, 434-486
// over the mesh's tables; MP_ERR_ARG for a node index outside the mesh.
int fe_setup(mp_fe& fe, const mp_mesh* m, double wall_potential) {
    fe.n_nodes = mp_mesh_num_nodes(m);
    fe.n_el = mp_mesh_num_tets(m);
    const mp_tetra* tets = mp_mesh_tets(m);
    const double* pos = mp_mesh_node_pos(m);
    const int32_t* type = mp_mesh_node_type(m);
    // unknowns: NORMAL and OPEN nodes, in node order (src/FESolver.cpp:51-55, 101-107)
    fe.ID.assign((size_t)fe.n_nodes, -1);
    for (int n = 0; n < fe.n_nodes; n++)
        if (type[n] == MP_NODE_NORMAL || type[n] == MP_NODE_OPEN) fe.ID[n] = fe.neq++;
    fe.con.resize((size_t)fe.n_el * 4);
    fe.LM.resize((size_t)fe.n_el * 4);
    fe.NX.resize((size_t)fe.n_el * 12);
    fe.detJ.resize((size_t)fe.n_el);
    for (int e = 0; e < fe.n_el; e++)
        for (int a = 0; a < 4; a++) {
            const int32_t nd = tets[e].con[a];
            if (nd < 0 || nd >= fe.n_nodes) return MP_ERR_ARG;
            fe.con[(size_t)e * 4 + a] = nd;
            fe.LM[(size_t)e * 4 + a] = fe.ID[nd];
        }
    {
        Traced tr_nx(fe, kFnComputeNX);
        for (int e = 0; e < fe.n_el; e++) {
            double x[4][3], NX[4][3];
            for (int a = 0; a < 4; a++) {
                const int32_t nd = fe.con[(size_t)e * 4 + a];
                for (int q = 0; q < 3; q++) x[a][q] = pos[3 * (size_t)nd + q];
            }
            mphost::element_nx(x, NX, &fe.detJ[e]);
            for (int a = 0; a < 4; a++)
                for (int q = 0; q < 3; q++) fe.NX[(size_t)e * 12 + 3 * a + q] = NX[a][q];
        }
    }
    // quadrature (src/FESolver.cpp:120-123)
    fe.l[0] = -std::sqrt(1.0 / 3.0);
    fe.l[1] = std::sqrt(1.0 / 3.0);
    fe.W[0] = 1;
    fe.W[1] = 1;
    fe.n_int = 2;
    // row pattern: unknown pairs that share an element
    std::vector<std::vector<int32_t>> cols((size_t)fe.neq);
    for (int e = 0; e < fe.n_el; e++)
        for (int a = 0; a < 4; a++) {
            const int P = fe.LM[(size_t)e * 4 + a];
            if (P < 0) continue;
            for (int b = 0; b < 4; b++) {
                const int Q = fe.LM[(size_t)e * 4 + b];
                if (Q >= 0) cols[P].push_back(Q);
            }
        }
    fe.rp.assign((size_t)fe.neq + 1, 0);
    for (int P = 0; P < fe.neq; P++) {
        auto& c = cols[P];
        c.push_back(P);  // the diagonal exists even for a node in no element
        std::sort(c.begin(), c.end());
        c.erase(std::unique(c.begin(), c.end()), c.end());
        fe.rp[P + 1] = fe.rp[P] + (int64_t)c.size();
    }
    fe.ci.resize((size_t)fe.rp[fe.neq]);
    fe.dg.resize((size_t)fe.neq);
    for (int P = 0; P < fe.neq; P++) {
        std::copy(cols[P].begin(), cols[P].end(), fe.ci.begin() + fe.rp[P]);
        fe.dg[P] = fe.rp[P] + (std::lower_bound(cols[P].begin(), cols[P].end(), P) - cols[P].begin());
    }
    fe.slot.assign((size_t)fe.n_el * 16, -1);
    for (int e = 0; e < fe.n_el; e++)
        for (int a = 0; a < 4; a++) {
            const int P = fe.LM[(size_t)e * 4 + a];
            if (P < 0) continue;
            for (int b = 0; b < 4; b++) {
                const int Q = fe.LM[(size_t)e * 4 + b];
                if (Q < 0) continue;
                const auto b0 = fe.ci.begin() + fe.rp[P], b1 = fe.ci.begin() + fe.rp[P + 1];
                fe.slot[(size_t)e * 16 + 4 * a + b] = std::lower_bound(b0, b1, Q) - fe.ci.begin();
            }
        }
    const size_t nnz = fe.ci.size();
    fe.K.assign(nnz, 0.0);
    fe.J.assign(nnz, 0.0);
    fe.F0.assign((size_t)fe.neq, 0.0);
    fe.F1.assign((size_t)fe.neq, 0.0);
    fe.d.assign((size_t)fe.neq, 0.0);
    fe.uh.assign((size_t)fe.n_nodes, 0.0);
    // g: 0 on INLET, -wall_potential on FIXED, 0 elsewhere (src/mini-pic.cpp:97-102)
    fe.g.assign((size_t)fe.n_nodes, 0.0);
    for (int n = 0; n < fe.n_nodes; n++) {
        if (type[n] == MP_NODE_INLET) fe.g[n] = 0;
        else if (type[n] == MP_NODE_FIXED) fe.g[n] = -wall_potential;
        else fe.g[n] = 0;
    }
    pre_assembly(fe);
    return MP_OK;
}

}  // namespace

extern "C" {

const char* mp_fe_last_error(void) { return g_fe_err.c_str(); }

int mp_fe_create(const mp_mesh* m, double n0, double phi0, double kTe, double wall_potential, int32_t flags,
                 mp_fe** out) {
    if (!m || !out) return fail(MP_ERR_ARG, "mp_fe_create: null argument");
    *out = nullptr;
    if (!(kTe > 0)) return fail(MP_ERR_ARG, "mp_fe_create: kTe must be positive");
    if (flags & ~MP_FE_RESET_F1) return fail(MP_ERR_ARG, "mp_fe_create: unknown flags");
    mp_fe* fe = new mp_fe;
    fe->flags = flags;
    fe->n0 = n0;
    fe->phi0 = phi0;
    fe->kTe = kTe;
    int rc;
    {
        Traced tr_(*fe, kFnFESolver);
        rc = fe_setup(*fe, m, wall_potential);
    }
    if (rc) {
        delete fe;
        return fail(MP_ERR_ARG, "mp_fe_create: node index out of range");
    }
    *out = fe;
    return MP_OK;
}

void mp_fe_free(mp_fe* fe) { delete fe; }

int32_t mp_fe_num_unknowns(const mp_fe* fe) { return fe ? fe->neq : 0; }

int mp_fe_compute_phi(mp_fe* fe, const double* ion_den, int32_t method, mp_fe_stats* st_out) {
    if (!fe || !ion_den) return fail(MP_ERR_ARG, "mp_fe_compute_phi: null argument");
    if (method != MP_FE_NONLINEAR && method != MP_FE_GAUSS_SEIDEL && method != MP_FE_DIRECT)
        return fail(MP_ERR_ARG, "mp_fe_compute_phi: unknown method");
    Traced tr_phi(*fe, kFnComputePhi);
    mp_fe_stats st{0, 0, 0, 0, 0.0};
    std::string err;
    const int rc = solve(*fe, ion_den, method, st, err);
    if (rc) {
        if (st_out) *st_out = st;
        return fail(rc, err);
    }
    // computePhi (src/FESolver.cpp:604-619): uh = g + d on the unknowns
    for (int n = 0; n < fe->n_nodes; n++) {
        fe->uh[n] = fe->g[n];
        const int A = fe->ID[n];
        if (A >= 0) fe->uh[n] += fe->d[A];
    }
    if (st_out) *st_out = st;
    return MP_OK;
}

int mp_fe_update_ef(const mp_fe* fe, double* ef) {
    if (!fe || !ef) return fail(MP_ERR_ARG, "mp_fe_update_ef: null argument");
    Traced tr_(*fe, kFnUpdateEf);
    for (int e = 0; e < fe->n_el; e++) {
        double NX[4][3], u[4];
        for (int a = 0; a < 4; a++) {
            for (int q = 0; q < 3; q++) NX[a][q] = fe->NX[(size_t)e * 12 + 3 * a + q];
            u[a] = fe->uh[fe->con[(size_t)e * 4 + a]];
        }
        mphost::element_ef(NX, u, ef + 3 * (size_t)e);
    }
    return MP_OK;
}

int mp_fe_get_phi(const mp_fe* fe, double* uh) {
    if (!fe || !uh) return fail(MP_ERR_ARG, "mp_fe_get_phi: null argument");
    std::copy(fe->uh.begin(), fe->uh.end(), uh);
    return MP_OK;
}

int mp_fe_get_stiffness(const mp_fe* fe, int64_t* nnz, int64_t* row_ptr, int32_t* col, double* val) {
    if (!fe || !nnz) return fail(MP_ERR_ARG, "mp_fe_get_stiffness: null argument");
    *nnz = (int64_t)fe->ci.size();
    if (row_ptr) std::copy(fe->rp.begin(), fe->rp.end(), row_ptr);
    if (col) std::copy(fe->ci.begin(), fe->ci.end(), col);
    if (val) std::copy(fe->K.begin(), fe->K.end(), val);
    return MP_OK;
}

int mp_fe_get_profile(const mp_fe* fe, int32_t k, const char** name, int64_t* calls, double* total_s,
                      double* child_s) {
    if (!fe) return fail(MP_ERR_ARG, "mp_fe_get_profile: null solver");
    if (k < 0 || k >= kFnCount) return fail(MP_ERR_ARG, "mp_fe_get_profile: no such entry");
    if (name) *name = kFnName[k];
    if (calls) *calls = fe->prof[k].calls;
    if (total_s) *total_s = fe->prof[k].total;
    if (child_s) *child_s = fe->prof[k].child;
    return MP_OK;
}

}  // extern "C"
