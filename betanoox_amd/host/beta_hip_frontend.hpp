// beta_hip_frontend.hpp -- input files of a BETA model that select and feed
// the device path: the mesh file and the solver-settings commands.
//
//  - Mesh file, text (BasicMesh::ReadMesh, BasicMesh.cpp:79-188):
//      numNodes numElements numDims
//      [processCommands  setCoordinateReadOption xyz|xy|yz|x ...  endCommands]
//      numNodes lines:    nodeNumber coordinates (per the read option)
//      numElements lines: elementNumber nnpe node_1 ... node_nnpe
//    Node numbers start at 0 or 1 (the first node's number decides, as
//    nodeOffset does); connectivity refers to node numbers.
//  - Mesh file, binary (BasicMesh::ReadMeshBinary, BasicMesh.cpp:190-268,
//    Node::readBinary node.cpp:113-122, BasicElement::readBinary
//    BasicElement.cpp:228-248): int numNodes, numElements, numDims; per node
//    int number and three doubles; per element int number, int nnpe, nnpe ints.
//  - Solver settings (Equations::processCommand, equation.cpp:1769-1903),
//    one command per line until "end": setStorageMethod,
//    setSparseSolverMaxIterations, setSparseSolverTolerance,
//    setSolverVerboseFlag, ReplaceZeroDiagonal, setReorderingScheme,
//    setUsePreviousFillInOrdering, UseMultiCore.  setStorageMethod HIPPCG
//    selects the device FP32-inner solver with FP64 refinement, HIPPCG_FP64
//    the plain FP64 device PCG; the reference's own tokens (SPARSE, PARCG,
//    MKLCG, ...) are accepted and also run on the device in FP64 with the
//    reference's stopping rule.
//  - "ParallelAssembly HIP" (the token BasicModel::processCommand reads,
//    BasicModel.cpp:391-413) selects the device assembler.
#pragma once

#include <istream>
#include <ostream>
#include <string>
#include <vector>

#include "beta_hip_adapters.hpp"

namespace beta_hip {

struct MeshFile {
    int num_dims = 3;
    std::vector<int> node_number;  // as in the file
    std::vector<double> xyz;       // 3 per node (missing components are 0)
    std::vector<int> elem_number;
    std::vector<int> nnpe;         // per element
    std::vector<int> conn;         // node indices (0-based positions in the node list)
    int node_offset = 0;           // 1 when the first node is numbered 1
    int n_nodes() const { return (int)node_number.size(); }
    int n_el() const { return (int)elem_number.size(); }
    // Element/node data for the device path (hex8 in 3D, quad4 in 2D).
    MeshData toMeshData() const;
};

// Throw BhipError(BHIP_ERR_ARG, ...) on a malformed file.
MeshFile readMesh(std::istream &in);
MeshFile readMeshBinary(const std::string &path);
void writeMesh(std::ostream &out, const MeshFile &m);           // text, as ReadMesh reads it
void writeMeshBinary(const std::string &path, const MeshFile &m);

struct SolverSettings {
    std::string storage = "SPARSE";  // setStorageMethod token
    int precision = BHIP_PREC_FP64;
    int stop_rule = BHIP_STOP_REF_R_OVER_X;
    int max_iterations = 10000;      // equation.cpp:1566-1567 defaults
    double tolerance = 1e-4;
    int verbose = 0;
    double zero_diag_replacement = 0.0;  // ReplaceZeroDiagonal (equation.cpp:1819-1821)
};

// Reads commands until "end" (Equations::readCommands).  Unknown commands
// are an error, as in the reference.
SolverSettings readSolverSettings(std::istream &in);
// Applies the settings to an Equations facade.
void applySolverSettings(const SolverSettings &s, HipEquations &eq);
// "ParallelAssembly <token>": true for HIP; the reference's OFA/ESA tokens
// are recognised (and assembled on the device too); anything else throws.
bool parseParallelAssembly(const std::string &token);

}  // namespace beta_hip
