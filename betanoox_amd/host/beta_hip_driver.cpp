// beta_hip_driver.cpp -- linear static analysis through the C++ adapters,
// in the order of BasicModel::DoAnalysis (BasicModel.cpp:1169-1220):
// numbering, restraints, pattern (allocateForAnalysis), loads, assembleK,
// Equations::solve.
//
//   beta_hip_driver quad N       plane stress, N x N quad4, unit x-traction
//   beta_hip_driver hex N        unit cube, N^3 hex8 (E=1, nu=0.3), z=0 clamped,
//                                unit pressure on z=1
// options: --tol T --rule ref|rel --fp32ir --dump FILE (solution, binary doubles)
//
//   beta_hip_driver mesh FILE    a BETA mesh file (BasicMesh::ReadMesh format;
//                                --binary for ReadMeshBinary), E/nu from
//                                --E/--nu (default 1, 0.3); 3D: nodes at the
//                                lowest z clamped, unit pressure on the faces at
//                                the highest z; 2D: lowest x clamped, unit
//                                x-traction on the edges at the highest x
//     --settings FILE            SolverSettings commands up to "end"
//                                (setStorageMethod HIPPCG selects the FP32-inner
//                                device solver)
//     --parallel-assembly TOKEN  the ParallelAssembly token (HIP, OFA, ESA)
//     --solution FILE            Solution.txt of the result (equation.cpp:1378-1383)
//     --parse-only               read the inputs and print their sizes (no GPU)
//
// Prints key=value lines (n_dof, iters, rel_resid, tip, ms).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include <algorithm>
#include <fstream>

#include "beta_hip_adapters.hpp"
#include "beta_hip_frontend.hpp"

using namespace beta_hip;

static MeshData quad_mesh(int n) {
    MeshData m;
    m.dim = 2;
    m.nnpe = 4;
    m.ndof_per_node = 2;
    for (int j = 0; j <= n; j++)
        for (int i = 0; i <= n; i++) {
            m.xyz.push_back((double)i / n);
            m.xyz.push_back((double)j / n);
        }
    auto id = [n](int i, int j) { return i + (n + 1) * j; };
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
            int c[4] = {id(i, j), id(i + 1, j), id(i + 1, j + 1), id(i, j + 1)};
            m.conn.insert(m.conn.end(), c, c + 4);
        }
    return m;
}

static MeshData hex_mesh(int n) {
    MeshData m;
    m.dim = 3;
    m.nnpe = 8;
    m.ndof_per_node = 3;
    for (int k = 0; k <= n; k++)
        for (int j = 0; j <= n; j++)
            for (int i = 0; i <= n; i++) {
                m.xyz.push_back((double)i / n);
                m.xyz.push_back((double)j / n);
                m.xyz.push_back((double)k / n);
            }
    auto id = [n](int i, int j, int k) { return i + (n + 1) * (j + (n + 1) * k); };
    for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) {
                // reference hex8 node order (BasicElement.cpp:518-525)
                int c[8] = {id(i, j, k),         id(i, j + 1, k),     id(i, j + 1, k + 1), id(i, j, k + 1),
                            id(i + 1, j, k),     id(i + 1, j + 1, k), id(i + 1, j + 1, k + 1),
                            id(i + 1, j, k + 1)};
                m.conn.insert(m.conn.end(), c, c + 8);
            }
    return m;
}

// Consistent nodal loads of a uniform traction on the boundary faces (3D,
// area/4 per face node of each planar quadrilateral face) or edges (2D,
// length/2 per edge node) that lie on the plane coord[axis] = value.
static std::vector<std::pair<int, double>> boundary_weights(const MeshData &m, int axis, double value) {
    std::vector<double> w(m.n_nodes(), 0.0);
    const double tol = 1e-12 * std::max(1.0, std::fabs(value));
    auto on = [&](int n) { return std::fabs(m.xyz[(size_t)n * m.dim + axis] - value) <= tol; };
    for (int e = 0; e < m.n_el(); e++) {
        std::vector<int> f;
        for (int a = 0; a < m.nnpe; a++)
            if (on(m.conn[(size_t)e * m.nnpe + a])) f.push_back(m.conn[(size_t)e * m.nnpe + a]);
        if (m.dim == 2 && f.size() == 2) {
            double dx = m.xyz[2 * f[0]] - m.xyz[2 * f[1]], dy = m.xyz[2 * f[0] + 1] - m.xyz[2 * f[1] + 1];
            double len = std::sqrt(dx * dx + dy * dy);
            for (int n : f) w[n] += len / 2;
        } else if (m.dim == 3 && f.size() == 4) {
            double c[3] = {0, 0, 0};
            for (int n : f)
                for (int d = 0; d < 3; d++) c[d] += m.xyz[3 * n + d] / 4;
            const int u = (axis + 1) % 3, v = (axis + 2) % 3;
            std::sort(f.begin(), f.end(), [&](int a, int b) {
                return std::atan2(m.xyz[3 * a + v] - c[v], m.xyz[3 * a + u] - c[u]) <
                       std::atan2(m.xyz[3 * b + v] - c[v], m.xyz[3 * b + u] - c[u]);
            });
            double d1[3], d2[3];
            for (int d = 0; d < 3; d++) {
                d1[d] = m.xyz[3 * f[2] + d] - m.xyz[3 * f[0] + d];
                d2[d] = m.xyz[3 * f[3] + d] - m.xyz[3 * f[1] + d];
            }
            double cx = d1[1] * d2[2] - d1[2] * d2[1], cy = d1[2] * d2[0] - d1[0] * d2[2], cz = d1[0] * d2[1] - d1[1] * d2[0];
            double area = 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
            for (int n : f) w[n] += area / 4;
        }
    }
    std::vector<std::pair<int, double>> out;
    for (int n = 0; n < m.n_nodes(); n++)
        if (w[n] != 0.0) out.push_back({n, w[n]});
    return out;
}

static int run_mesh_file(int argc, char **argv) {
    std::string path = argv[2], settings_path, pa_token = "HIP", solution_path;
    bool binary = false, parse_only = false;
    double E = 1.0, nu = 0.3;
    for (int a = 3; a < argc; a++) {
        if (!std::strcmp(argv[a], "--binary")) binary = true;
        else if (!std::strcmp(argv[a], "--parse-only")) parse_only = true;
        else if (!std::strcmp(argv[a], "--settings") && a + 1 < argc) settings_path = argv[++a];
        else if (!std::strcmp(argv[a], "--parallel-assembly") && a + 1 < argc) pa_token = argv[++a];
        else if (!std::strcmp(argv[a], "--solution") && a + 1 < argc) solution_path = argv[++a];
        else if (!std::strcmp(argv[a], "--E") && a + 1 < argc) E = std::atof(argv[++a]);
        else if (!std::strcmp(argv[a], "--nu") && a + 1 < argc) nu = std::atof(argv[++a]);
        else {
            std::fprintf(stderr, "unknown option %s\n", argv[a]);
            return 2;
        }
    }
    try {
        MeshFile mf;
        if (binary) {
            mf = readMeshBinary(path);
        } else {
            std::ifstream in(path);
            if (!in) throw BhipError(BHIP_ERR_ARG, "cannot open " + path);
            mf = readMesh(in);
        }
        SolverSettings ss;
        if (!settings_path.empty()) {
            std::ifstream in(settings_path);
            if (!in) throw BhipError(BHIP_ERR_ARG, "cannot open " + settings_path);
            ss = readSolverSettings(in);
        }
        const bool hip_assembly = parseParallelAssembly(pa_token);
        MeshData mesh = mf.toMeshData();
        std::printf("n_nodes=%d\nn_el=%d\nnum_dims=%d\nnode_offset=%d\nstorage=%s\nprecision=%d\nmax_iterations=%d\n"
                    "tolerance=%.6e\nparallel_assembly_hip=%d\nzero_diag_replacement=%.6e\n",
                    mf.n_nodes(), mf.n_el(), mf.num_dims, mf.node_offset, ss.storage.c_str(), ss.precision,
                    ss.max_iterations, ss.tolerance, hip_assembly ? 1 : 0, ss.zero_diag_replacement);
        if (parse_only) return 0;
        Device dev(0);
        HipEquations eq(dev);
        eq.setup(mesh);
        const int dim = mesh.dim, nd = mesh.ndof_per_node;
        // clamp the lowest plane, load the highest (z in 3D, x in 2D)
        const int axis = dim == 3 ? 2 : 0;
        double lo = 1e300, hi = -1e300;
        for (int n = 0; n < mesh.n_nodes(); n++) {
            lo = std::min(lo, mesh.xyz[(size_t)n * dim + axis]);
            hi = std::max(hi, mesh.xyz[(size_t)n * dim + axis]);
        }
        const double tol_lo = 1e-12 * std::max(1.0, std::fabs(lo));
        for (int n = 0; n < mesh.n_nodes(); n++)
            if (std::fabs(mesh.xyz[(size_t)n * dim + axis] - lo) <= tol_lo)
                for (int c = 0; c < nd; c++) eq.setSolution(0.0, nd * n + c);
        eq.allocate();
        int tip = 0;
        double best = 1e300, ctr[3] = {0, 0, 0};
        for (int n = 0; n < mesh.n_nodes(); n++)
            for (int d = 0; d < dim; d++) ctr[d] += mesh.xyz[(size_t)n * dim + d] / mesh.n_nodes();
        for (auto &pw : boundary_weights(mesh, axis, hi)) {
            if (dim == 3) eq.addLoadToEqn(-pw.second, 3 * pw.first + 2);
            else eq.addLoadToEqn(pw.second, 2 * pw.first);
            double dd = 0;
            for (int d = 0; d < dim; d++)
                if (d != axis) dd += std::pow(mesh.xyz[(size_t)pw.first * dim + d] - ctr[d], 2);
            if (dd < best) {
                best = dd;
                tip = nd * pw.first + (dim == 3 ? 2 : 0);
            }
        }
        MaterialData mat;
        if (dim == 3) {
            mat.kind = BHIP_MAT_ISO_ELASTIC;
            mat.stride = 2;
            mat.params = {E, nu};
        } else {
            mat.kind = BHIP_MAT_PLANE_STRESS;
            mat.stride = 4;
            mat.params = {E, E, nu, E / (2 * (1 + nu))};
        }
        HipAssembler asmb(dev, eq);
        asmb.allocateAndInitialize(mesh, mat, 2, 1.0);
        asmb.assembleK(nullptr);
        applySolverSettings(ss, eq);
        const std::vector<double> &u = eq.solve(nullptr);
        const bhip_stats &s = eq.lastStats();
        std::printf("n_dof=%d\niters=%d\nouter=%d\nrel_resid=%.6e\ntip=%.17e\nms=%.3f\n", eq.numberOfEquations(),
                    s.iters, s.outer_iters, s.rel_resid, u[tip], s.ms_total);
        if (!solution_path.empty() &&
            bhip_write_vector(solution_path.c_str(), BHIP_VEC_SOLUTION, (int)u.size(), u.data()) != BHIP_OK)
            throw BhipError(BHIP_ERR_ARG, "cannot write " + solution_path);
    } catch (const BhipError &e) {
        std::fprintf(stderr, "error %d: %s\n", e.status(), e.what());
        return 1;
    }
    return 0;
}

int main(int argc, char **argv) {
    if (argc >= 3 && !std::strcmp(argv[1], "mesh")) return run_mesh_file(argc, argv);
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s quad|hex N [--tol T] [--rule ref|rel] [--fp32ir|--fp64blocked] [--dump FILE]\n"
                             "       %s mesh FILE [--binary] [--settings FILE] [--parallel-assembly T] [--E e] [--nu v]\n"
                             "            [--solution FILE] [--parse-only]\n", argv[0], argv[0]);
        return 2;
    }
    std::string kind = argv[1];
    int n = std::atoi(argv[2]);
    double tol = 1e-10;
    int rule = BHIP_STOP_REL_RESID;
    int prec = BHIP_PREC_FP64;
    const char *dump = nullptr;
    for (int a = 3; a < argc; a++) {
        if (!std::strcmp(argv[a], "--tol") && a + 1 < argc) tol = std::atof(argv[++a]);
        else if (!std::strcmp(argv[a], "--rule") && a + 1 < argc) rule = std::strcmp(argv[++a], "ref") ? BHIP_STOP_REL_RESID : BHIP_STOP_REF_R_OVER_X;
        else if (!std::strcmp(argv[a], "--fp32ir")) prec = BHIP_PREC_FP32_IR;
        else if (!std::strcmp(argv[a], "--fp64blocked")) prec = BHIP_PREC_FP64_BLOCKED;
        else if (!std::strcmp(argv[a], "--dump") && a + 1 < argc) dump = argv[++a];
    }
    if (n <= 0 || (kind != "quad" && kind != "hex")) return 2;
    try {
        Device dev(0);
        HipEquations eq(dev);
        MeshData mesh = kind == "quad" ? quad_mesh(n) : hex_mesh(n);
        eq.setup(mesh);
        MaterialData mat;
        const double h = 1.0 / n;
        int tip = 0;
        if (kind == "quad") {
            mat.kind = BHIP_MAT_PLANE_STRESS;
            mat.stride = 4;
            mat.params = {1.0, 1.0, 0.3, 1.0 / 2.6};
            for (int j = 0; j <= n; j++) {  // x = 0 clamped
                eq.setSolution(0.0, 2 * ((n + 1) * j));
                eq.setSolution(0.0, 2 * ((n + 1) * j) + 1);
            }
            eq.allocate();
            for (int j = 0; j <= n; j++) {  // unit traction on x = 1
                double w = (j == 0 || j == n) ? h / 2 : h;
                eq.addLoadToEqn(w, 2 * (n + (n + 1) * j));
            }
            tip = 2 * (n + (n + 1) * (n / 2));
        } else {
            mat.kind = BHIP_MAT_ISO_ELASTIC;
            mat.stride = 2;
            mat.params = {1.0, 0.3};
            int nn1 = (n + 1) * (n + 1);
            for (int a = 0; a < nn1; a++)
                for (int c = 0; c < 3; c++) eq.setSolution(0.0, 3 * a + c);
            eq.allocate();
            for (int j = 0; j <= n; j++)
                for (int i = 0; i <= n; i++) {
                    bool ex = (i == 0 || i == n), ey = (j == 0 || j == n);
                    double w = ex && ey ? h * h / 4 : (ex || ey ? h * h / 2 : h * h);
                    eq.addLoadToEqn(-w, 3 * (i + (n + 1) * (j + (n + 1) * n)) + 2);
                }
            tip = 3 * (n / 2 + (n + 1) * (n / 2 + (n + 1) * n)) + 2;
        }
        HipAssembler asmb(dev, eq);
        asmb.allocateAndInitialize(mesh, mat, 2, 1.0);
        asmb.assembleK(nullptr);
        eq.setTolerance(tol);
        eq.setStopRule(rule);
        eq.setPrecision(prec);
        const std::vector<double> &u = eq.solve(nullptr);
        const bhip_stats &s = eq.lastStats();
        std::printf("n_dof=%d\niters=%d\nouter=%d\nrel_resid=%.6e\ntip=%.17e\nms=%.3f\n", eq.numberOfEquations(),
                    s.iters, s.outer_iters, s.rel_resid, u[tip], s.ms_total);
        // LargeMatrix view: K(tip,tip) and a product, as the reference exposes them
        HipLargeMatrix K(dev);
        std::printf("k_tip_tip=%.17e\n", K.getaij(tip, tip));
        if (dump) {
            FILE *f = std::fopen(dump, "wb");
            if (!f) return 3;
            std::fwrite(u.data(), sizeof(double), u.size(), f);
            std::fclose(f);
        }
    } catch (const BhipError &e) {
        std::fprintf(stderr, "error %d: %s\n", e.status(), e.what());
        return 1;
    }
    return 0;
}
