"""MI355X-native BETA finite-element hot path (element integration, CSR
assembly, Jacobi-PCG).  The compute path is libbeta_hip.so (HIP, gfx950),
reached through the C ABI in include/beta_hip.h; see betanoox_amd.capi."""
from . import capi, meshgen  # noqa: F401

__all__ = ["capi", "meshgen", "setup_problem"]


def setup_problem(ctx, prob, elem_group=None):
    """Mesh, restraints, pattern, materials and integration rule of a
    meshgen.Problem on a capi.Context (BasicModel::allocateForAnalysis)."""
    ctx.set_mesh(prob.dim, prob.xyz, prob.conn, ndof_uniform=prob.ndof)
    ctx.set_restraints(prob.restrained, prob.prescribed)
    ctx.build_pattern()
    if elem_group is None and "group" in prob.extra:
        elem_group = prob.extra["group"]
    ctx.set_materials(prob.mat_kind, prob.mat, elem_group)
    ctx.set_integration(prob.order, prob.factor)
    return ctx
