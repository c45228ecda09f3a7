// beta_hip_adapters.hpp -- C++ plug-in adapters over the C ABI of
// libbeta_hip.so, shaped like the reference's plug-in surfaces so that a
// model selects the GPU path as a drop-in:
//
//   HipLargeMatrix   ~ LargeMatrix / SymmetricSparseMatrix
//                      (beta/math/equation/largeMatrix.cpp, sparse.cpp)
//   HipEquations     ~ Equations (beta/math/equation/equation.cpp): restraints,
//                      loads, RHS build, solve
//   HipAssembler     ~ BasicAssembler / SerialAssembler
//                      (beta/assemblers/SerialAssembler.cpp)
//
// Errors: every C-ABI status is turned into a BhipError exception; with
// exitOnError(true) the adapters instead print the message and exit(1), which
// is what the reference does (FatalError, sparse.cpp:225-229).
#pragma once

#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "beta_hip.h"

namespace beta_hip {

class BhipError : public std::runtime_error {
   public:
    BhipError(int status, const std::string &msg) : std::runtime_error(msg), status_(status) {}
    int status() const { return status_; }

   private:
    int status_;
};

// Owns one device context.
class Device {
   public:
    explicit Device(int device = 0);
    ~Device();
    Device(const Device &) = delete;
    Device &operator=(const Device &) = delete;
    bhip_ctx *ctx() const { return ctx_; }
    void check(int status, const char *what) const;
    static void exitOnError(bool on) { exit_on_error_ = on; }

   private:
    bhip_ctx *ctx_ = nullptr;
    static bool exit_on_error_;
};

// Mesh description as BasicMesh holds it after ReadMesh + setNumDofPerNode.
struct MeshData {
    int dim = 3;
    int nnpe = 8;
    std::vector<double> xyz;   // n_nodes * dim
    std::vector<int> conn;     // n_el * nnpe, reference node order
    int ndof_per_node = 3;
    int n_nodes() const { return (int)(xyz.size() / dim); }
    int n_el() const { return (int)(conn.size() / nnpe); }
};

// Equations-shaped facade (equation.cpp): restraint/solution/load vectors on
// the host, matrix and solve on the device.
class HipEquations {
   public:
    explicit HipEquations(Device &dev) : dev_(dev) {}
    void setup(const MeshData &mesh);  // numbering (node.cpp:206-214)
    int numberOfEquations() const { return n_; }
    void setRestraint(int eqn, bool val);                 // equation.cpp:306-318
    void setSolution(double val, int eqn);                // equation.cpp:320-341
    void addLoadToEqn(double load, int eqn);              // equation.cpp:851-868
    void zeroLoadVector();
    void allocate();                                      // pattern on the device
    const std::vector<double> &solve(const double *initGuess = nullptr);  // equation.cpp:1288-1405
    std::vector<double> &solution() { return solution_; }
    void setMaxIterations(int n) { maxit_ = n; }          // setSparseSolverMaxIterations
    void setTolerance(double eps) { eps_ = eps; }         // setSparseSolverTolerance
    void setStopRule(int r) { stop_rule_ = r; }
    void setPrecision(int p) { precision_ = p; }
    // ReplaceZeroDiagonal (equation.cpp:1819-1821 -> setReplaceZeroDiagonal_Option)
    void setReplaceZeroDiagonal_Option(double v) {
        dev_.check(bhip_set_zero_diag_replacement(dev_.ctx(), v), "setReplaceZeroDiagonal_Option");
    }
    const bhip_stats &lastStats() const { return stats_; }
    Device &device() { return dev_; }

   private:
    Device &dev_;
    int n_ = 0;
    std::vector<unsigned char> restraint_;
    std::vector<double> solution_, load_, lastSolution_;
    int maxit_ = 10000;
    double eps_ = 1e-4;  // equation.cpp:1566-1567
    int stop_rule_ = BHIP_STOP_REF_R_OVER_X;
    int precision_ = BHIP_PREC_FP64;
    bhip_stats stats_{};
    bool restraints_uploaded_ = false;
};

// LargeMatrix-shaped view of the device matrix (the "existing CG solver
// interface"): solve(b, x0) with the solution returned in b, vectorProduct,
// lower-triangle element access for diagnostics.
class HipLargeMatrix {
   public:
    explicit HipLargeMatrix(Device &dev) : dev_(dev) {}
    void setMAX_NUMBER_ITERATIONS(int n) { maxit_ = n; }
    void setSOLUTION_EPSILON(double e) { eps_ = e; }
    void setStopRule(int r) { stop_rule_ = r; }
    void setPrecision(int p) { precision_ = p; }
    void setReplaceZeroDiagonal_Option(double v) {          // largeMatrix.cpp:43-79
        dev_.check(bhip_set_zero_diag_replacement(dev_.ctx(), v), "setReplaceZeroDiagonal_Option");
    }
    void solve(double *b, double *initialGuess);           // sparse.cpp:114-242
    void vectorProduct(double *result, double *vector);    // sparse.cpp:452-478
    double getaij(int i, int j);                            // sparse.cpp:80-99, i >= j
    int numberOfEquations() const { return bhip_num_dof(dev_.ctx()); }
    const bhip_stats &lastStats() const { return stats_; }

   private:
    void cacheLower();
    Device &dev_;
    int maxit_ = 10000;
    double eps_ = 1e-4;
    int stop_rule_ = BHIP_STOP_REF_R_OVER_X;
    int precision_ = BHIP_PREC_FP64;
    bhip_stats stats_{};
    std::vector<int> rowptr_, colind_;
    std::vector<double> val_;
};

// Material record per element group (MaterialFactory + SetElementProperty).
struct MaterialData {
    int kind = BHIP_MAT_ISO_ELASTIC;
    int stride = 2;
    std::vector<double> params;     // n_groups * stride
    std::vector<int> elem_group;    // empty: one group, or one per element
};

// BasicAssembler-shaped: allocateAndInitialize / assembleK / assembleKandF.
class HipAssembler {
   public:
    HipAssembler(Device &dev, HipEquations &eq) : dev_(dev), eq_(eq) {}
    void allocateAndInitialize(const MeshData &mesh, const MaterialData &mat, int order = 2,
                               double factor = 1.0);
    void assembleK(const double *theSolution = nullptr);   // SerialAssembler.cpp:35-59
    void assembleHeatTransient(double alpha, double dt, const double *u_s);  // DiffusionModel.cpp:322-347
    std::vector<double> pseudoLoad();

   private:
    Device &dev_;
    HipEquations &eq_;
};

}  // namespace beta_hip
