"""wavetoy_hip -- MI355X (gfx950) HIP implementation of the 3D SBP-SAT scalar
wave RHS and RK4 step of eschnett/WaveToySecondOrder.jl, behind the C ABI of
libwt2.so (include/wt2.h) and a host mirror of the reference's call names."""
from ._lib import EXPORTS, LIB_PATH, WT2Error, lib  # noqa: F401
from .core import (  # noqa: F401
    SOMMERFELD_BDRY_TAG, AnalyticBackground3D, Batch, Element, Geometry, Mesh, Params3d, SBPOps, Workspace,
    adm_bump_background, ctx_footprint, mesh_desc, discrete_energy, eigenmode_cartesian, gaussian_pulse, host_coords,
    make_coef3d, make_element, make_geometry, make_operators, make_uniform_hex,
    make_warped_uniform_hex, make_workspace, recommended_dt_rk4, rhs_wave3d,
    sample_background3d, tau_effective, to_device, wave3d_rhs_full,
    BC3d, make_bc3d, wave3d_curved_rhs, wave3d_energy, eigenmode_radial, outgoing_pulse,
    FACE_SUBLUMINAL, FACE_OUTFLOW, FACE_INFLOW, FACE_SONIC, classify_face1d, classify_face3d,
    validate_bc1d, HexMesh, connect_by_coords, make_cubed_ball_mesh,
    make_cubed_cube_mesh, make_inflated_cube_mesh, make_radial_shell_mesh, EXCISION_BDRY_TAG,
)
from .evolve import evolve3d  # noqa: F401
