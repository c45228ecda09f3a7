// beta_hip_adapters.cpp -- see beta_hip_adapters.hpp.
#include "beta_hip_adapters.hpp"

#include <algorithm>
#include <cstring>

namespace beta_hip {

bool Device::exit_on_error_ = false;

Device::Device(int device) {
    int st = bhip_create(device, &ctx_);
    if (st != BHIP_OK) {
        if (exit_on_error_) {
            std::fprintf(stderr, "bhip_create failed: %d\n", st);
            std::exit(1);
        }
        throw BhipError(st, "bhip_create failed (no usable GPU?)");
    }
}

Device::~Device() { bhip_destroy(ctx_); }

void Device::check(int status, const char *what) const {
    if (status == BHIP_OK) return;
    std::string msg = std::string(what) + ": " + bhip_last_error(ctx_);
    if (exit_on_error_) {
        std::fprintf(stderr, "%s\n", msg.c_str());
        std::exit(1);
    }
    throw BhipError(status, msg);
}

// ------------------------------------------------------------ HipEquations
void HipEquations::setup(const MeshData &mesh) {
    dev_.check(bhip_set_mesh(dev_.ctx(), mesh.dim, mesh.n_nodes(), mesh.xyz.data(), mesh.n_el(), mesh.nnpe,
                             mesh.conn.data(), nullptr, mesh.ndof_per_node),
               "set mesh");
    n_ = bhip_num_dof(dev_.ctx());
    restraint_.assign(n_, 0);
    solution_.assign(n_, 0.0);
    load_.assign(n_, 0.0);
    lastSolution_.assign(n_, 0.0);
    restraints_uploaded_ = false;
}

void HipEquations::setRestraint(int eqn, bool val) {
    restraint_.at(eqn) = val ? 1 : 0;
    restraints_uploaded_ = false;
}

void HipEquations::setSolution(double val, int eqn) {
    restraint_.at(eqn) = 1;
    solution_.at(eqn) = val;
    restraints_uploaded_ = false;
}

void HipEquations::addLoadToEqn(double load, int eqn) {
    if (restraint_.at(eqn)) dev_.check(BHIP_ERR_ARG, "Cannot add load to a restrained degree of freedom.");
    load_[eqn] += load;
}

void HipEquations::zeroLoadVector() { std::fill(load_.begin(), load_.end(), 0.0); }

void HipEquations::allocate() {
    dev_.check(bhip_set_restraints(dev_.ctx(), restraint_.data(), solution_.data()), "set restraints");
    dev_.check(bhip_build_pattern(dev_.ctx()), "build pattern");
    restraints_uploaded_ = true;
}

const std::vector<double> &HipEquations::solve(const double *initGuess) {
    if (!restraints_uploaded_)
        dev_.check(bhip_set_prescribed(dev_.ctx(), solution_.data()), "set prescribed values");
    dev_.check(bhip_set_load(dev_.ctx(), load_.data()), "set load");
    dev_.check(bhip_build_rhs(dev_.ctx()), "build rhs");
    // initial guess: restrained DOFs take their prescribed value (equation.cpp:1370-1380)
    std::vector<double> guess(n_);
    const double *src = initGuess ? initGuess : lastSolution_.data();
    for (int i = 0; i < n_; i++) guess[i] = restraint_[i] && initGuess ? solution_[i] : src[i];
    dev_.check(bhip_set_x0(dev_.ctx(), guess.data()), "set x0");
    bhip_solve_opts o{};
    o.tol = eps_;
    o.maxit = maxit_;
    o.stop_rule = stop_rule_;
    o.precision = precision_;
    o.inner_tol = 1e-3;
    o.max_outer = 50;
    o.check_every = 32;
    o.use_x0 = 1;
    dev_.check(bhip_solve(dev_.ctx(), &o, &stats_), "solve");
    dev_.check(bhip_get_solution(dev_.ctx(), solution_.data()), "get solution");
    lastSolution_ = solution_;
    return solution_;
}

// ---------------------------------------------------------- HipLargeMatrix
void HipLargeMatrix::solve(double *b, double *initialGuess) {
    bhip_solve_opts o{};
    o.tol = eps_;
    o.maxit = maxit_;
    o.stop_rule = stop_rule_;
    o.precision = precision_;
    o.inner_tol = 1e-3;
    o.max_outer = 50;
    o.check_every = 32;
    dev_.check(bhip_pcg_solve(dev_.ctx(), b, initialGuess, &o, &stats_), "Solution not converged");
}

void HipLargeMatrix::vectorProduct(double *result, double *vector) {
    dev_.check(bhip_spmv(dev_.ctx(), vector, result), "vectorProduct");
}

void HipLargeMatrix::cacheLower() {
    long long nl = 0, nf = 0;
    dev_.check(bhip_pattern_nnz(dev_.ctx(), &nl, &nf), "pattern size");
    int n = numberOfEquations();
    rowptr_.resize(n + 1);
    colind_.resize(nl);
    val_.resize(nl);
    dev_.check(bhip_get_pattern_lower(dev_.ctx(), rowptr_.data(), colind_.data()), "pattern");
    dev_.check(bhip_get_values_lower(dev_.ctx(), val_.data()), "values");
}

double HipLargeMatrix::getaij(int i, int j) {
    if (j > i) std::swap(i, j);
    if (rowptr_.empty()) cacheLower();
    auto b = colind_.begin() + rowptr_[i], e = colind_.begin() + rowptr_[i + 1];
    auto it = std::lower_bound(b, e, j);
    if (it == e || *it != j) dev_.check(BHIP_ERR_ARG, "Bad operand...");
    return val_[it - colind_.begin()];
}

// ------------------------------------------------------------ HipAssembler
void HipAssembler::allocateAndInitialize(const MeshData &mesh, const MaterialData &mat, int order,
                                         double factor) {
    (void)mesh;
    int ng = (int)(mat.params.size() / mat.stride);
    dev_.check(bhip_set_materials(dev_.ctx(), mat.kind, ng, mat.params.data(),
                                  mat.elem_group.empty() ? nullptr : mat.elem_group.data()),
               "set materials");
    dev_.check(bhip_set_integration(dev_.ctx(), order, factor), "set integration");
}

void HipAssembler::assembleK(const double *theSolution) {
    (void)theSolution;  // small-strain linear stiffness does not depend on u
    dev_.check(bhip_assemble(dev_.ctx(), BHIP_ASM_STIFFNESS, nullptr, nullptr), "assembleK");
}

void HipAssembler::assembleHeatTransient(double alpha, double dt, const double *u_s) {
    double coef[2] = {alpha, dt};
    dev_.check(bhip_assemble(dev_.ctx(), BHIP_ASM_HEAT_TRANSIENT, coef, u_s), "assembleK_F_s_alpha_family");
}

std::vector<double> HipAssembler::pseudoLoad() {
    std::vector<double> f(eq_.numberOfEquations());
    dev_.check(bhip_get_pseudo_load(dev_.ctx(), f.data()), "pseudo load");
    return f;
}

}  // namespace beta_hip
