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
// Prints key=value lines (n_dof, iters, rel_resid, tip, ms).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "beta_hip_adapters.hpp"

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

int main(int argc, char **argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s quad|hex N [--tol T] [--rule ref|rel] [--fp32ir] [--dump FILE]\n", argv[0]);
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
