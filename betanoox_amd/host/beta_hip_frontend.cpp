// beta_hip_frontend.cpp -- mesh files and solver-settings commands (see the
// header for the formats and the reference lines they follow).
#include "beta_hip_frontend.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <sstream>

namespace beta_hip {

namespace {

[[noreturn]] void bad(const std::string &what) { throw BhipError(BHIP_ERR_ARG, "mesh/settings input: " + what); }

std::string upper(std::string s) {
    for (auto &c : s) c = (char)std::toupper((unsigned char)c);
    return s;
}

// Next non-empty line without leading blanks (peekAtNextLine's view).
std::string peek_line(std::istream &in) {
    std::streampos pos = in.tellg();
    std::string line;
    while (std::getline(in, line)) {
        size_t a = line.find_first_not_of(" \t\r");
        if (a != std::string::npos) {
            size_t b = line.find_last_not_of(" \t\r");
            in.clear();
            in.seekg(pos);
            return line.substr(a, b - a + 1);
        }
    }
    in.clear();
    in.seekg(pos);
    return "";
}

void finish(MeshFile &m) {
    if (m.n_nodes() == 0) bad("no nodes");
    m.node_offset = m.node_number[0] == 1 ? 1 : 0;
    for (int i = 0; i < m.n_nodes(); i++)
        if (m.node_number[i] != i + m.node_offset)
            bad("node numbers must run contiguously from 0 or 1 (node " + std::to_string(i) + ")");
    for (auto &c : m.conn) {
        c -= m.node_offset;
        if (c < 0 || c >= m.n_nodes()) bad("connectivity refers to a node that does not exist");
    }
}

}  // namespace

MeshData MeshFile::toMeshData() const {
    MeshData d;
    d.dim = num_dims;
    if (num_dims == 3) {
        d.nnpe = 8;
        d.ndof_per_node = 3;
    } else if (num_dims == 2) {
        d.nnpe = 4;
        d.ndof_per_node = 2;
    } else {
        bad("only 2D (quad4) and 3D (hex8) meshes run on the device");
    }
    for (int e = 0; e < n_el(); e++)
        if (nnpe[e] != d.nnpe) bad("element " + std::to_string(elem_number[e]) + " has " + std::to_string(nnpe[e]) +
                                   " nodes; the device path has hex8 (3D) and quad4 (2D)");
    d.xyz.reserve((size_t)n_nodes() * num_dims);
    for (int i = 0; i < n_nodes(); i++)
        for (int c = 0; c < num_dims; c++) d.xyz.push_back(xyz[3 * (size_t)i + c]);
    d.conn = conn;
    return d;
}

MeshFile readMesh(std::istream &in) {
    MeshFile m;
    int nn = 0, ne = 0, nd = 0;
    if (!(in >> nn >> ne >> nd)) bad("cannot read 'numNodes numElements numDims'");
    if (nn <= 0 || ne < 0 || std::abs(nd) < 1 || std::abs(nd) > 3) bad("numDims is invalid");
    m.num_dims = std::abs(nd);
    // coordinate read option: from numDims, or a processCommands block
    std::string option = m.num_dims == 1 ? "x" : m.num_dims == 2 ? "xy" : "xyz";
    std::string rest_of_line;
    std::getline(in, rest_of_line);
    if (peek_line(in) == "processCommands") {
        std::string tok;
        in >> tok;
        for (;;) {
            if (!(in >> tok)) bad("processCommands without endCommands");
            if (tok == "endCommands") break;
            if (tok == "setCoordinateReadOption") {
                if (!(in >> tok)) bad("setCoordinateReadOption without a value");
                if (tok == "xyz" || tok == "xy" || tok == "yz" || tok == "x") option = tok;
                else bad("unknown coordinate read option " + tok);
            }
        }
    }
    m.node_number.resize(nn);
    m.xyz.assign(3 * (size_t)nn, 0.0);
    for (int i = 0; i < nn; i++) {
        double *p = &m.xyz[3 * (size_t)i];
        bool ok = (bool)(in >> m.node_number[i]);
        if (option == "x") ok = ok && (in >> p[0]);
        else if (option == "xy") ok = ok && (in >> p[0] >> p[1]);
        else if (option == "yz") ok = ok && (in >> p[1] >> p[2]);
        else ok = ok && (in >> p[0] >> p[1] >> p[2]);
        if (!ok) bad("bad node line " + std::to_string(i));
    }
    m.elem_number.resize(ne);
    m.nnpe.resize(ne);
    for (int e = 0; e < ne; e++) {
        int k = 0;
        if (!(in >> m.elem_number[e] >> k) || k <= 0 || k > 27) bad("bad element line " + std::to_string(e));
        m.nnpe[e] = k;
        for (int a = 0; a < k; a++) {
            int v;
            if (!(in >> v)) bad("bad element line " + std::to_string(e));
            m.conn.push_back(v);
        }
    }
    finish(m);
    return m;
}

MeshFile readMeshBinary(const std::string &path) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) bad("cannot open binary mesh " + path);
    MeshFile m;
    int nn = 0, ne = 0, nd = 0;
    bool ok = std::fread(&nn, sizeof(int), 1, f) == 1 && std::fread(&ne, sizeof(int), 1, f) == 1 &&
              std::fread(&nd, sizeof(int), 1, f) == 1;
    if (!ok || nn <= 0 || ne < 0 || std::abs(nd) < 1 || std::abs(nd) > 3) {
        std::fclose(f);
        bad("bad binary mesh header");
    }
    m.num_dims = std::abs(nd);
    m.node_number.resize(nn);
    m.xyz.resize(3 * (size_t)nn);
    for (int i = 0; i < nn && ok; i++)
        ok = std::fread(&m.node_number[i], sizeof(int), 1, f) == 1 && std::fread(&m.xyz[3 * (size_t)i], sizeof(double), 3, f) == 3;
    m.elem_number.resize(ne);
    m.nnpe.resize(ne);
    for (int e = 0; e < ne && ok; e++) {
        ok = std::fread(&m.elem_number[e], sizeof(int), 1, f) == 1 && std::fread(&m.nnpe[e], sizeof(int), 1, f) == 1 &&
             m.nnpe[e] > 0 && m.nnpe[e] <= 27;
        if (!ok) break;
        size_t at = m.conn.size();
        m.conn.resize(at + m.nnpe[e]);
        ok = std::fread(&m.conn[at], sizeof(int), m.nnpe[e], f) == (size_t)m.nnpe[e];
    }
    std::fclose(f);
    if (!ok) bad("binary mesh is truncated: " + path);
    finish(m);
    return m;
}

void writeMesh(std::ostream &out, const MeshFile &m) {
    out << m.n_nodes() << " " << m.n_el() << " " << m.num_dims << "\n";
    out << std::setprecision(17);
    for (int i = 0; i < m.n_nodes(); i++) {
        out << m.node_number[i];
        for (int c = 0; c < m.num_dims; c++) out << " " << m.xyz[3 * (size_t)i + c];
        out << "\n";
    }
    size_t at = 0;
    for (int e = 0; e < m.n_el(); e++) {
        out << m.elem_number[e] << " " << m.nnpe[e];
        for (int a = 0; a < m.nnpe[e]; a++) out << " " << m.conn[at + a] + m.node_offset;
        at += m.nnpe[e];
        out << "\n";
    }
}

void writeMeshBinary(const std::string &path, const MeshFile &m) {
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) bad("cannot write " + path);
    int h[3] = {m.n_nodes(), m.n_el(), m.num_dims};
    std::fwrite(h, sizeof(int), 3, f);
    for (int i = 0; i < m.n_nodes(); i++) {
        std::fwrite(&m.node_number[i], sizeof(int), 1, f);
        std::fwrite(&m.xyz[3 * (size_t)i], sizeof(double), 3, f);
    }
    size_t at = 0;
    for (int e = 0; e < m.n_el(); e++) {
        std::fwrite(&m.elem_number[e], sizeof(int), 1, f);
        std::fwrite(&m.nnpe[e], sizeof(int), 1, f);
        for (int a = 0; a < m.nnpe[e]; a++) {
            int v = m.conn[at + a] + m.node_offset;
            std::fwrite(&v, sizeof(int), 1, f);
        }
        at += m.nnpe[e];
    }
    std::fclose(f);
}

SolverSettings readSolverSettings(std::istream &in) {
    SolverSettings s;
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::string cmd, arg;
        if (!(ls >> cmd)) continue;
        ls >> arg;
        if (cmd == "end") return s;
        if (cmd == "setStorageMethod") {
            const std::string t = upper(arg);
            s.storage = t;
            if (t == "HIPPCG") {
                s.precision = BHIP_PREC_FP32_IR;
                s.stop_rule = BHIP_STOP_REL_RESID;
            } else if (t == "HIPPCG_FP64") {
                s.precision = BHIP_PREC_FP64;
                s.stop_rule = BHIP_STOP_REL_RESID;
            } else if (t == "HIPPCG_FP64_BLOCKED") {
                s.precision = BHIP_PREC_FP64_BLOCKED;
                s.stop_rule = BHIP_STOP_REL_RESID;
            } else if (t == "SPARSE" || t == "PARCG" || t == "MKLCG" || t == "PROFILE" || t == "OLAF" ||
                       t == "PARDISO" || t == "MKLPARDISO" || t == "WSMPSYMM" || t == "PETSC") {
                s.precision = BHIP_PREC_FP64;  // equation.cpp:36-140 tokens; run as the FP64 Jacobi-PCG
                s.stop_rule = BHIP_STOP_REF_R_OVER_X;
            } else {
                bad("Solver Method: " + arg + " selected does not exist for this platform.");
            }
        } else if (cmd == "setSparseSolverMaxIterations") {
            s.max_iterations = std::atoi(arg.c_str());
        } else if (cmd == "setSparseSolverTolerance") {
            s.tolerance = std::atof(arg.c_str());
        } else if (cmd == "setSolverVerboseFlag") {
            s.verbose = std::atoi(arg.c_str());
        } else if (cmd == "ReplaceZeroDiagonal") {
            s.zero_diag_replacement = std::atof(arg.c_str());
        } else if (cmd == "setReorderingScheme" || cmd == "setUsePreviousFillInOrdering" || cmd == "UseMultiCore") {
            // direct-solver and threading options: nothing to do for the device PCG
        } else {
            bad("Did not recognize command: " + cmd);
        }
    }
    bad("SolverSettings without 'end'");
}

void applySolverSettings(const SolverSettings &s, HipEquations &eq) {
    eq.setMaxIterations(s.max_iterations);
    eq.setTolerance(s.tolerance);
    eq.setStopRule(s.stop_rule);
    eq.setPrecision(s.precision);
    eq.setReplaceZeroDiagonal_Option(s.zero_diag_replacement);
}

bool parseParallelAssembly(const std::string &token) {
    const std::string t = upper(token);
    if (t == "HIP") return true;
    if (t == "OFA" || t == "ESA") return false;  // BasicModel.cpp:391-413; assembled on the device as well
    throw BhipError(BHIP_ERR_ARG, "Requested parallel assembly algorithm not recognized");
}

}  // namespace beta_hip
