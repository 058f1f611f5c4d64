// Host-side mesh tables, frozen fields and synthetic particles (declared in mp_mesh.h).
//
// The table arithmetic follows mini-pic's LoadVolumeMesh operation for operation so the
// tables are bit-identical to the reference's for the same node coordinates:
//   src/maths.cpp:22-53   det4 / det3 (cofactor expansion, left-to-right sums)
//   src/meshes.cpp:86-108 volume = (1/6) det4, flip con[1]<->con[3] when negative
//   src/meshes.cpp:110-176 alpha..delta with the per-vertex (v2,v3,v4) orders
//   src/meshes.cpp:178-227 face-neighbour table (here: hash of the sorted face triple)
//   src/meshes.cpp:229-251 node volume += lc(centroid)[v] * tet volume
//   src/meshes.cpp:253-262 nodes on faces without a neighbour become OPEN
// Compiled with -ffp-contract=off: no fused multiply-adds anywhere.
#include "mp_mesh.h"

#include "fe_geom.h"
#include "synth_rng.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

struct mp_mesh {
    std::vector<double>   pos;      // [n][3]
    std::vector<double>   nvol;     // [n]
    std::vector<int32_t>  ntype;    // [n]
    std::vector<mp_tetra> tets;     // [T]
    std::vector<mp_face>  inlet;    // inlet faces
    bool is_box = false;
    int32_t nx = 0, ny = 0, nz = 0;
    double lx = 0, ly = 0, lz = 0;
};

namespace {

thread_local std::string g_mesh_err;

// Static-partition parallel loop over [0, n) on the host's hardware threads.
template <typename F>
void parallel_for(int64_t n, F&& body) {
    unsigned hw = std::thread::hardware_concurrency();
    int64_t nt = hw ? (int64_t)hw : 1;
    if (nt > 64) nt = 64;
    if (n < 4096) nt = 1;
    if (nt <= 1) {
        for (int64_t i = 0; i < n; i++) body(i);
        return;
    }
    std::vector<std::thread> th;
    for (int64_t t = 0; t < nt; t++) {
        const int64_t lo = n * t / nt, hi = n * (t + 1) / nt;
        th.emplace_back([lo, hi, &body] {
            for (int64_t i = lo; i < hi; i++) body(i);
        });
    }
    for (auto& x : th) x.join();
}

int fail(const std::string& m) {
    g_mesh_err = m;
    return MP_ERR_ARG;
}

// src/maths.cpp:48-53
double det3(const double M[3][3]) {
    return M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) -
           M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
           M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
}

// src/maths.cpp:22-45
double det4(const double M[4][4]) {
    double M0[3][3], M1[3][3], M2[3][3], M3[3][3];
    for (int i = 0; i < 3; i++) {
        M0[i][0] = M[i + 1][1]; M0[i][1] = M[i + 1][2]; M0[i][2] = M[i + 1][3];
        M1[i][0] = M[i + 1][0]; M1[i][1] = M[i + 1][2]; M1[i][2] = M[i + 1][3];
        M2[i][0] = M[i + 1][0]; M2[i][1] = M[i + 1][1]; M2[i][2] = M[i + 1][3];
        M3[i][0] = M[i + 1][0]; M3[i][1] = M[i + 1][1]; M3[i][2] = M[i + 1][2];
    }
    return M[0][0] * det3(M0) - M[0][1] * det3(M1) + M[0][2] * det3(M2) - M[0][3] * det3(M3);
}

// Barycentric weight of vertex i, src/meshes.h:82-83.
inline double bary(const mp_tetra& t, int i, const double p[3]) {
    return (1.0 / 6.0) * (t.alpha[i] - p[0] * t.beta[i] + p[1] * t.gamma[i] - p[2] * t.delta[i]) /
           t.volume;
}

struct FaceKey {
    int32_t a, b, c;
    bool operator==(const FaceKey& o) const { return a == o.a && b == o.b && c == o.c; }
};
struct FaceHash {
    size_t operator()(const FaceKey& k) const {
        uint64_t h = (uint64_t)(uint32_t)k.a * 0x9E3779B97F4A7C15ull;
        h ^= (uint64_t)(uint32_t)k.b + 0x7F4A7C159E3779B9ull + (h << 6) + (h >> 2);
        h ^= (uint64_t)(uint32_t)k.c * 0xC2B2AE3D27D4EB4Full + (h << 6) + (h >> 2);
        return (size_t)h;
    }
};

// All per-tet tables from con[] and node positions (LoadVolumeMesh after parsing).
void compute_tables(mp_mesh& m) {
    const int32_t T = (int32_t)m.tets.size();
    const int32_t N = (int32_t)(m.pos.size() / 3);
    auto P = [&](int32_t n) { return &m.pos[3 * (size_t)n]; };

    // volumes, src/meshes.cpp:86-108
    for (int32_t l = 0; l < T; l++) {
        mp_tetra& tet = m.tets[l];
        double M[4][4];
        for (int i = 0; i < 4; i++) M[i][0] = 1;
        for (int v = 0; v < 4; v++)
            for (int d = 0; d < 3; d++) M[v][d + 1] = P(tet.con[v])[d];
        tet.volume = (1.0 / 6.0) * det4(M);
        if (tet.volume < 0) {
            int32_t t = tet.con[1];
            tet.con[1] = tet.con[3];
            tet.con[3] = t;
            tet.volume = -tet.volume;
        }
    }

    // 3x3 determinants, src/meshes.cpp:110-176
    static const int vmap[4][3] = {{1, 2, 3}, {3, 2, 0}, {3, 0, 1}, {1, 0, 2}};
    for (int32_t l = 0; l < T; l++) {
        mp_tetra& tet = m.tets[l];
        for (int v = 0; v < 4; v++) {
            const double* p2 = P(tet.con[vmap[v][0]]);
            const double* p3 = P(tet.con[vmap[v][1]]);
            const double* p4 = P(tet.con[vmap[v][2]]);
            double M[3][3];
            M[0][0] = p2[0]; M[0][1] = p2[1]; M[0][2] = p2[2];
            M[1][0] = p3[0]; M[1][1] = p3[1]; M[1][2] = p3[2];
            M[2][0] = p4[0]; M[2][1] = p4[1]; M[2][2] = p4[2];
            tet.alpha[v] = det3(M);
            M[0][0] = 1; M[0][1] = p2[1]; M[0][2] = p2[2];
            M[1][0] = 1; M[1][1] = p3[1]; M[1][2] = p3[2];
            M[2][0] = 1; M[2][1] = p4[1]; M[2][2] = p4[2];
            tet.beta[v] = det3(M);
            M[0][0] = 1; M[0][1] = p2[0]; M[0][2] = p2[2];
            M[1][0] = 1; M[1][1] = p3[0]; M[1][2] = p3[2];
            M[2][0] = 1; M[2][1] = p4[0]; M[2][2] = p4[2];
            tet.gamma[v] = det3(M);
            M[0][0] = 1; M[0][1] = p2[0]; M[0][2] = p2[1];
            M[1][0] = 1; M[1][1] = p3[0]; M[1][2] = p3[1];
            M[2][0] = 1; M[2][1] = p4[0]; M[2][2] = p4[1];
            tet.delta[v] = det3(M);
        }
    }

    // face neighbours (same result as the node-to-tet search of src/meshes.cpp:178-227
    // on a conforming mesh: the tet across the face opposite node v, else -1)
    std::unordered_map<FaceKey, int64_t, FaceHash> open;
    open.reserve((size_t)T * 2 + 16);
    for (int32_t l = 0; l < T; l++)
        for (int v = 0; v < 4; v++) m.tets[l].cell_con[v] = -1;
    for (int32_t l = 0; l < T; l++) {
        mp_tetra& tet = m.tets[l];
        for (int v = 0; v < 4; v++) {
            int32_t f[3], k = 0;
            for (int i = 0; i < 4; i++)
                if (i != v) f[k++] = tet.con[i];
            std::sort(f, f + 3);
            FaceKey key{f[0], f[1], f[2]};
            auto it = open.find(key);
            if (it == open.end()) {
                open.emplace(key, (int64_t)l * 4 + v);
            } else {
                int32_t o = (int32_t)(it->second >> 2);
                int ov = (int)(it->second & 3);
                tet.cell_con[v] = o;
                m.tets[o].cell_con[ov] = l;
                open.erase(it);
            }
        }
    }

    // node volumes, src/meshes.cpp:229-251
    m.nvol.assign(N, 0.0);
    for (int32_t i = 0; i < T; i++) {
        const mp_tetra& tet = m.tets[i];
        double c[3];
        for (int d = 0; d < 3; d++) {
            c[d] = 0;
            for (int v = 0; v < 4; v++) c[d] += 0.25 * P(tet.con[v])[d];
        }
        double lc[4];
        for (int v = 0; v < 4; v++) lc[v] = bary(tet, v, c);
        for (int v = 0; v < 4; v++) m.nvol[tet.con[v]] += lc[v] * tet.volume;
    }

    // OPEN marking, src/meshes.cpp:253-262
    m.ntype.assign(N, MP_NODE_NORMAL);
    for (int32_t e = 0; e < T; e++) {
        const mp_tetra& tet = m.tets[e];
        for (int v = 0; v < 4; v++)
            if (tet.cell_con[v] < 0)
                for (int i = 0; i < 4; i++)
                    if (i != v) m.ntype[tet.con[i]] = MP_NODE_OPEN;
    }
}

// Axis order of tet p inside a hex: path corner(0,0,0) -> +e[o0] -> +e[o1] -> +e[o2].
const int kPerm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};

using mprng::draw;
using mprng::mix64;
using mprng::normal2;

}  // namespace

extern "C" {

const char* mp_mesh_last_error(void) { return g_mesh_err.c_str(); }

int mp_mesh_box(int32_t nx, int32_t ny, int32_t nz, double lx, double ly, double lz,
                mp_mesh** out) {
    if (!out) return fail("mp_mesh_box: out is null");
    *out = nullptr;
    if (nx < 1 || ny < 1 || nz < 1) return fail("mp_mesh_box: hex counts must be >= 1");
    if (!(lx > 0) || !(ly > 0) || !(lz > 0)) return fail("mp_mesh_box: sides must be > 0");
    const int64_t nn = (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
    const int64_t nt = (int64_t)6 * nx * ny * nz;
    if (nn > INT32_MAX || nt > INT32_MAX) return fail("mp_mesh_box: mesh too large");
    mp_mesh* m = new (std::nothrow) mp_mesh;
    if (!m) return fail("mp_mesh_box: out of memory");
    m->is_box = true;
    m->nx = nx; m->ny = ny; m->nz = nz;
    m->lx = lx; m->ly = ly; m->lz = lz;
    m->pos.resize(3 * (size_t)nn);
    for (int32_t k = 0; k <= nz; k++)
        for (int32_t j = 0; j <= ny; j++)
            for (int32_t i = 0; i <= nx; i++) {
                size_t id = (size_t)i + (size_t)(nx + 1) * ((size_t)j + (size_t)(ny + 1) * k);
                m->pos[3 * id + 0] = lx * ((double)i / nx);
                m->pos[3 * id + 1] = ly * ((double)j / ny);
                m->pos[3 * id + 2] = lz * ((double)k / nz);
            }
    m->tets.resize((size_t)nt);
    auto node = [&](int32_t i, int32_t j, int32_t k) {
        return (int32_t)(i + (nx + 1) * (j + (ny + 1) * k));
    };
    for (int32_t k = 0; k < nz; k++)
        for (int32_t j = 0; j < ny; j++)
            for (int32_t i = 0; i < nx; i++) {
                size_t h = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k);
                for (int p = 0; p < 6; p++) {
                    mp_tetra& t = m->tets[6 * h + p];
                    std::memset(&t, 0, sizeof(t));
                    int c[3] = {i, j, k};
                    t.con[0] = node(c[0], c[1], c[2]);
                    for (int s = 0; s < 3; s++) {
                        c[kPerm[p][s]] += 1;
                        t.con[s + 1] = node(c[0], c[1], c[2]);
                    }
                }
            }
    compute_tables(*m);
    *out = m;
    return MP_OK;
}

int mp_mesh_load_dat(const char* path, mp_mesh** out) {
    if (!out || !path) return fail("mp_mesh_load_dat: null argument");
    *out = nullptr;
    std::ifstream in(path);
    if (!in.is_open()) return fail(std::string("Failed to open ") + path);
    int n_nodes = 0, n_elements = 0;
    in >> n_nodes >> n_elements;
    if (!in || n_nodes <= 0) return fail("mp_mesh_load_dat: bad header");
    mp_mesh* m = new (std::nothrow) mp_mesh;
    if (!m) return fail("mp_mesh_load_dat: out of memory");
    m->pos.reserve(3 * (size_t)n_nodes);
    for (int n = 0; n < n_nodes; n++) {
        int index;
        double x, y, z;
        in >> index >> x >> y >> z;
        m->pos.push_back(x / 1000.);
        m->pos.push_back(y / 1000.);
        m->pos.push_back(z / 1000.);
    }
    for (int e = 0; e < n_elements; e++) {
        int index, type;
        in >> index >> type;
        if (!in) break;
        if (type == 304) {
            int n1, n2, n3, n4;
            in >> n1 >> n2 >> n3 >> n4;
            mp_tetra t;
            std::memset(&t, 0, sizeof(t));
            t.con[0] = n1 - 1; t.con[1] = n2 - 1; t.con[2] = n3 - 1; t.con[3] = n4 - 1;
            for (int v = 0; v < 4; v++)
                if (t.con[v] < 0 || t.con[v] >= n_nodes) {
                    delete m;
                    return fail("mp_mesh_load_dat: node id out of range");
                }
            m->tets.push_back(t);
        } else {
            std::string s;
            std::getline(in, s);
        }
    }
    if (m->tets.empty()) {
        delete m;
        return fail("mp_mesh_load_dat: no tetrahedra");
    }
    compute_tables(*m);
    *out = m;
    return MP_OK;
}

int mp_mesh_load_surface(mp_mesh* m, const char* path, int32_t node_type, int32_t invert_normals) {
    if (!m || !path) return fail("mp_mesh_load_surface: null argument");
    std::ifstream in(path);
    if (!in.is_open()) return fail(std::string("Failed to open ") + path);
    int n_nodes = 0, n_elements = 0;
    in >> n_nodes >> n_elements;
    const int nn = (int)m->nvol.size();
    for (int n = 0; n < n_nodes; n++) {
        int index;
        double x, y, z;
        in >> index >> x >> y >> z;
        if (index < 1 || index > nn) continue;
        m->ntype[index - 1] = node_type;
    }
    if (node_type != MP_NODE_INLET) return MP_OK;
    // owning tet of each inlet triangle: hash of node triples instead of the O(faces*T)
    // scan of src/meshes.cpp:307-322 (same answer: the unique tet holding all 3 nodes)
    std::unordered_map<FaceKey, int32_t, FaceHash> owner;
    std::unordered_map<FaceKey, int32_t, FaceHash> dup;
    for (size_t l = 0; l < m->tets.size(); l++) {
        const mp_tetra& t = m->tets[l];
        for (int v = 0; v < 4; v++) {
            int32_t f[3], k = 0;
            for (int i = 0; i < 4; i++)
                if (i != v) f[k++] = t.con[i];
            std::sort(f, f + 3);
            FaceKey key{f[0], f[1], f[2]};
            if (!owner.emplace(key, (int32_t)l).second) dup[key] = (int32_t)l;
        }
    }
    for (int e = 0; e < n_elements; e++) {
        int index, type;
        in >> index >> type;
        if (!in) break;
        if (type != 203) {
            std::string s;
            std::getline(in, s);
            continue;
        }
        int n1, n2, n3;
        in >> n1 >> n2 >> n3;
        mp_face f;
        std::memset(&f, 0, sizeof(f));
        f.con[0] = n1 - 1; f.con[1] = n2 - 1; f.con[2] = n3 - 1;
        int32_t s3[3] = {n1 - 1, n2 - 1, n3 - 1};
        std::sort(s3, s3 + 3);
        FaceKey key{s3[0], s3[1], s3[2]};
        auto it = owner.find(key);
        if (it == owner.end()) return fail("No volume element attached to inlet surface");
        if (dup.count(key)) return fail("Inlet surface attached to more than one volume element");
        f.cell_con = it->second;
        const double* a = &m->pos[3 * (size_t)(n1 - 1)];
        const double* b = &m->pos[3 * (size_t)(n2 - 1)];
        const double* c = &m->pos[3 * (size_t)(n3 - 1)];
        for (int i = 0; i < 3; i++) {
            f.u[i] = b[i] - a[i];
            f.v[i] = c[i] - a[i];
        }
        double normal[3];
        normal[0] = f.u[2] * f.v[1] - f.u[1] * f.v[2];
        normal[1] = f.u[0] * f.v[2] - f.u[2] * f.v[0];
        normal[2] = f.u[1] * f.v[0] - f.u[0] * f.v[1];
        if (invert_normals) {
            normal[0] = -normal[0];
            normal[1] = -normal[1];
            normal[2] = -normal[2];
        }
        double len = std::sqrt(normal[0] * normal[0] + normal[1] * normal[1] + normal[2] * normal[2]);
        for (int i = 0; i < 3; i++) f.normal[i] = normal[i] / len;
        f.area = len / 2;
        m->inlet.push_back(f);
    }
    return MP_OK;
}

void mp_mesh_free(mp_mesh* m) { delete m; }

int32_t mp_mesh_num_nodes(const mp_mesh* m) { return m ? (int32_t)m->nvol.size() : 0; }
int32_t mp_mesh_num_tets(const mp_mesh* m) { return m ? (int32_t)m->tets.size() : 0; }
const mp_tetra* mp_mesh_tets(const mp_mesh* m) { return m ? m->tets.data() : nullptr; }
const double* mp_mesh_node_pos(const mp_mesh* m) { return m ? m->pos.data() : nullptr; }
const double* mp_mesh_node_volume(const mp_mesh* m) { return m ? m->nvol.data() : nullptr; }
const int32_t* mp_mesh_node_type(const mp_mesh* m) { return m ? m->ntype.data() : nullptr; }
int32_t mp_mesh_num_inlet_faces(const mp_mesh* m) { return m ? (int32_t)m->inlet.size() : 0; }
const mp_face* mp_mesh_inlet_faces(const mp_mesh* m) { return m ? m->inlet.data() : nullptr; }

int mp_mesh_locate_box(const mp_mesh* m, const double* pos, int64_t n, int32_t* cell) {
    if (!m || !m->is_box) return fail("mp_mesh_locate_box: not a box mesh");
    if (n < 0 || (n > 0 && (!pos || !cell))) return fail("mp_mesh_locate_box: bad arrays");
    const int32_t nxyz[3] = {m->nx, m->ny, m->nz};
    const double lxyz[3] = {m->lx, m->ly, m->lz};
    parallel_for(n, [&](int64_t p) {
        int32_t idx[3];
        double fr[3];
        bool inside = true;
        for (int d = 0; d < 3; d++) {
            double x = pos[3 * p + d];
            if (!(x >= 0.0 && x <= lxyz[d])) inside = false;
            double s = x / lxyz[d] * nxyz[d];
            int32_t i = (int32_t)std::floor(s);
            if (i < 0) i = 0;
            if (i > nxyz[d] - 1) i = nxyz[d] - 1;
            idx[d] = i;
            fr[d] = s - i;
        }
        if (!inside) {
            cell[p] = -1;
            return;
        }
        // axis order = fractional coordinates in descending order (ties: lower axis first)
        int o[3] = {0, 1, 2};
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2 - a; b++)
                if (fr[o[b]] < fr[o[b + 1]]) std::swap(o[b], o[b + 1]);
        int pidx = 0;
        for (int q = 0; q < 6; q++)
            if (kPerm[q][0] == o[0] && kPerm[q][1] == o[1]) pidx = q;
        size_t h = (size_t)idx[0] + (size_t)m->nx * ((size_t)idx[1] + (size_t)m->ny * idx[2]);
        cell[p] = (int32_t)(6 * h + pidx);
    });
    return MP_OK;
}

int mp_field_from_potential(const mp_mesh* m, const double* phi, double* ef) {
    if (!m || !phi || !ef) return fail("mp_field_from_potential: null argument");
    const int64_t T = (int64_t)m->tets.size();
    // computeNX (src/FESolver.cpp:207-246), inverse (:249-287), updateEf (:622-636)
    parallel_for(T, [&](int64_t e) {
        const mp_tetra& tet = m->tets[e];
        double x[4][3], NX[4][3], u[4];
        for (int a = 0; a < 4; a++) {
            for (int d = 0; d < 3; d++) x[a][d] = m->pos[3 * (size_t)tet.con[a] + d];
            u[a] = phi[tet.con[a]];
        }
        mphost::element_nx(x, NX, nullptr);
        mphost::element_ef(NX, u, ef + 3 * e);
    });
    return MP_OK;
}

int mp_field_sinusoid(const mp_mesh* m, double amp, double* ef) {
    if (!m || !ef) return fail("mp_field_sinusoid: null argument");
    const size_t N = m->nvol.size();
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (size_t n = 0; n < N; n++)
        for (int d = 0; d < 3; d++) {
            lo[d] = std::min(lo[d], m->pos[3 * n + d]);
            hi[d] = std::max(hi[d], m->pos[3 * n + d]);
        }
    std::vector<double> phi(N);
    for (size_t n = 0; n < N; n++) {
        double s = amp;
        for (int d = 0; d < 3; d++) s *= std::sin(M_PI * (m->pos[3 * n + d] - lo[d]) / (hi[d] - lo[d]));
        phi[n] = s;
    }
    return mp_field_from_potential(m, phi.data(), ef);
}

int mp_synth_uniform(uint64_t seed, int64_t n, const double lo[3], const double hi[3],
                     double sigma, const double drift[3], double* pos, double* vel) {
    if (n < 0 || (n > 0 && (!pos || !vel || !lo || !hi || !drift)))
        return fail("mp_synth_uniform: bad arguments");
    parallel_for(n, [&](int64_t p) {
        for (int d = 0; d < 3; d++) pos[3 * p + d] = lo[d] + (hi[d] - lo[d]) * draw(seed, p, d);
        double z[4];
        normal2(seed, p, 3, z[0], z[1]);
        normal2(seed, p, 5, z[2], z[3]);
        for (int d = 0; d < 3; d++) vel[3 * p + d] = drift[d] + sigma * z[d];
    });
    return MP_OK;
}

int mp_synth_mixture(uint64_t seed, int64_t n, const double lo[3], const double hi[3],
                     int32_t n_comp, const double* w, const double* centres, double sigma_pos,
                     double sigma_vel, const double drift[3], double* pos, double* vel) {
    if (n < 0 || n_comp < 1 || (n > 0 && (!pos || !vel || !lo || !hi || !drift || !w || !centres)))
        return fail("mp_synth_mixture: bad arguments");
    double wsum = 0;
    for (int k = 0; k < n_comp; k++) wsum += w[k];
    parallel_for(n, [&](int64_t p) {
        double z[4];
        normal2(seed, p, 0, z[0], z[1]);
        normal2(seed, p, 2, z[2], z[3]);
        for (int d = 0; d < 3; d++) vel[3 * p + d] = drift[d] + sigma_vel * z[d];
        // component choice, then resample position until inside (draw streams 4..)
        double u = draw(seed, p, 4) * wsum;
        int k = 0;
        double acc = w[0];
        while (k + 1 < n_comp && u >= acc) acc += w[++k];
        const double* c = centres + 3 * k;
        uint64_t s2 = mix64(seed ^ 0xA5A5A5A5DEADBEEFull);
        for (int attempt = 0;; attempt++) {
            double q[4], r[2];
            normal2(s2, p * 64 + attempt, 0, q[0], q[1]);
            normal2(s2, p * 64 + attempt, 2, q[2], r[0]);
            bool ok = true;
            double x[3];
            for (int d = 0; d < 3; d++) {
                x[d] = c[d] + sigma_pos * q[d];
                if (!(x[d] >= lo[d] && x[d] < hi[d])) ok = false;
            }
            if (ok || attempt == 63) {
                for (int d = 0; d < 3; d++)
                    pos[3 * p + d] = ok ? x[d] : std::min(std::max(x[d], lo[d]), std::nextafter(hi[d], lo[d]));
                break;
            }
        }
    });
    return MP_OK;
}

}  // extern "C"
