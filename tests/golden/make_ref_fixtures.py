#!/usr/bin/env python3
"""Write the golden fixtures of the CD3D/TCD3D path from the reference build.

Runs ``oracle/_ref/cd3d_ref`` (the reference compiled from its own sources by
``oracle/build_ref.py``) on a set of small cases and stores what the reference
computed -- cell vertex lists, cell-to-DOF maps, CSR pattern and entries, rhs,
a V-cycle probe, the FGMRES solution and its iteration count, and TCD3D
Crank-Nicolson states -- as compressed ``.npz`` files next to this script.

Only data is stored; the reference itself never travels.  This script runs
where the reference sources exist (it needs ``/root/reference/data/Wuerfel.*``).

usage: python tests/golden/make_ref_fixtures.py [case ...]
"""
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

# The reference reads data/Wuerfel.GEO into the box [0,1] x [0,0.346] x [0,0.346]
# (ReadGeo.C:1753-1755 scales parametrised y and z by 0.346).  Cases marked
# "unit" use oracle/data/unitcube.{GEO,PRM} instead: the same cube with
# explicit vertex coordinates (KNPR 0), which the reader does not scale.
BOX = (1.0, 0.346, 0.346)
UNIT = (1.0, 1.0, 1.0)

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
EXE = os.path.join(REPO, "oracle", "_ref", "cd3d_ref")
REF = os.environ.get("PARMOON_REF", "/root/reference")

# Solver settings shared by every case (Jacobi 0.8,
# 2 pre + 2 post, V-cycle, FGMRES(20); coarse smoother SSOR because the
# reference's exact coarse solve needs an external LAPACK).
SOLVER = dict(
    SOLVER_TYPE=1, SC_MG_TYPE_SCALAR=0, SC_SOLVER_SCALAR=16, SC_PRECONDITIONER_SCALAR=5,
    SC_GMRES_RESTART=20, SC_LIN_MAXIT_SCALAR=200, SC_LIN_RED_FACTOR_SCALAR=1e-10,
    SC_LIN_RES_NORM_MIN_SCALAR=1e-30, SC_MG_CYCLE_SCALAR=1, SC_SMOOTHER_SCALAR=1,
    SC_PRE_SMOOTH_SCALAR=2, SC_POST_SMOOTH_SCALAR=2, SC_SMOOTH_DAMP_FACTOR_SCALAR=0.8,
    SC_SMOOTH_DAMP_FACTOR_FINE_SCALAR=0.8, SC_COARSE_SMOOTHER_SCALAR=3,
    SC_COARSE_MAXIT_SCALAR=10, SC_COARSE_RED_FACTOR_SCALAR=0.1, SC_GMG_DAMP_FACTOR_SCALAR=1.0,
    SC_GMG_DAMP_FACTOR_FINE_SCALAR=1.0, SC_STEP_LENGTH_CONTROL_FINE_SCALAR=0,
    SC_STEP_LENGTH_CONTROL_ALL_SCALAR=0, SC_AMG_PREC_IT=1, SC_MINIT=0, SC_VERBOSE=1,
    SC_FLEXIBLE_KRYLOV_SPACE_SOLVER=1,
)

# name -> (mode, .dat overrides)
CASES = {
    # Poisson with manufactured solution: configuration 1 at 2 and 3 levels
    "q1_l2_poisson": ("cd", dict(ANSATZ_ORDER=1, LEVELS=2, UNIFORM_STEPS=2, DISCTYPE=1,
                                 P0=1.0, P1=0, P2=0, P3=0, P4=0, P5=1)),
    "q1_c1": ("cd", dict(ANSATZ_ORDER=1, LEVELS=3, UNIFORM_STEPS=2, DISCTYPE=1,
                         P0=1.0, P1=0, P2=0, P3=0, P4=0, P5=1)),
    # SUPG convection-diffusion, configuration 2 problem at reduced size
    "q2_l2_supg": ("cd", dict(ANSATZ_ORDER=2, LEVELS=2, UNIFORM_STEPS=1, DISCTYPE=2,
                              CELL_MEASURE=4, SDFEM_TYPE=2, DELTA0=1.0,
                              P0=1e-3, P1=1.0, P2=0.5, P3=0.25, P4=0, P5=0, P6=1.0)),
    "q1_l3_supg": ("cd", dict(ANSATZ_ORDER=1, LEVELS=3, UNIFORM_STEPS=2, DISCTYPE=2,
                              CELL_MEASURE=4, SDFEM_TYPE=2, DELTA0=1.0,
                              P0=1e-6, P1=-0.3, P2=0.7, P3=0.4, P4=0.5, P5=0, P6=1.0)),
    "q2_l2_galerkin_conv": ("cd", dict(ANSATZ_ORDER=2, LEVELS=2, UNIFORM_STEPS=1, DISCTYPE=1,
                                       P0=0.1, P1=1.0, P2=0.5, P3=0.25, P4=1.0, P5=1)),
    # TCD3D, Crank-Nicolson with the rotating convection field (configuration 3 problem)
    # (rotation centre and pulse centre placed inside the reference's box)
    # on the unit cube: configuration 1 exactly, and the configuration 2
    # problem (Q2, SUPG, eps 1e-3, b = (1, 0.5, 0.25), f = 1) at 8^3 cells
    "unit_c1": ("cd", dict(geo="unit", ANSATZ_ORDER=1, LEVELS=3, UNIFORM_STEPS=2, DISCTYPE=1,
                           P0=1.0, P1=0, P2=0, P3=0, P4=0, P5=1)),
    "unit_c2_n8": ("cd", dict(geo="unit", ANSATZ_ORDER=2, LEVELS=3, UNIFORM_STEPS=1, DISCTYPE=2,
                              CELL_MEASURE=4, SDFEM_TYPE=2, DELTA0=1.0,
                              P0=1e-3, P1=1.0, P2=0.5, P3=0.25, P4=0, P5=0, P6=1.0)),
    "q1_tcd": ("tcd", dict(ANSATZ_ORDER=1, LEVELS=2, UNIFORM_STEPS=2, DISCTYPE=1,
                           P0=1e-4, P4=0, P5=0, P6=0.0, P7=1, P8=3,
                           P9=0.3, P10=0.173, P11=0.173, P12=0.5, P13=0.173,
                           TIME_DISC=2, STARTTIME=0.0, ENDTIME=1.0, TIMESTEPLENGTH=1e-3)),
    # unit cube split into m^3 macro bricks (the meshes of the batched and
    # serving workloads: coarsest level m^3 cells), SUPG with non-axis
    # convection over several decades of eps
    "m2_q1_supg": ("cd", dict(geo="m2", ANSATZ_ORDER=1, LEVELS=3, UNIFORM_STEPS=1, DISCTYPE=2,
                              CELL_MEASURE=4, SDFEM_TYPE=2, DELTA0=1.0,
                              P0=1e-6, P1=0.48, P2=-0.36, P3=0.64, P4=0, P5=0, P6=1.0)),
    "m3_q1_supg": ("cd", dict(geo="m3", ANSATZ_ORDER=1, LEVELS=3, UNIFORM_STEPS=0, DISCTYPE=2,
                              CELL_MEASURE=4, SDFEM_TYPE=2, DELTA0=1.0,
                              P0=1e-2, P1=-0.5, P2=0.2, P3=0.6, P4=0, P5=0, P6=1.0)),
    "m5_q1_supg": ("cd", dict(geo="m5", ANSATZ_ORDER=1, LEVELS=2, UNIFORM_STEPS=0, DISCTYPE=2,
                              CELL_MEASURE=4, SDFEM_TYPE=2, DELTA0=1.0,
                              P0=3e-4, P1=0.1, P2=0.7, P3=-0.2, P4=0, P5=0, P6=1.0)),
    "m2_q2_supg": ("cd", dict(geo="m2", ANSATZ_ORDER=2, LEVELS=3, UNIFORM_STEPS=0, DISCTYPE=2,
                              CELL_MEASURE=4, SDFEM_TYPE=2, DELTA0=1.0,
                              P0=1e-3, P1=0.6, P2=0.5, P3=-0.4, P4=0, P5=0, P6=1.0)),
    "m3_q2_supg": ("cd", dict(geo="m3", ANSATZ_ORDER=2, LEVELS=2, UNIFORM_STEPS=0, DISCTYPE=2,
                              CELL_MEASURE=4, SDFEM_TYPE=2, DELTA0=1.0,
                              P0=1e-5, P1=-0.2, P2=-0.3, P3=0.5, P4=0, P5=0, P6=1.0)),
    # W- and F-cycles (SC_MG_CYCLE_SCALAR 2 and 0, MultiGrid3D.C:690-703)
    "m2_q1_wcycle": ("cd", dict(geo="m2", ANSATZ_ORDER=1, LEVELS=4, UNIFORM_STEPS=0, DISCTYPE=2,
                                CELL_MEASURE=4, SDFEM_TYPE=2, DELTA0=1.0, SC_MG_CYCLE_SCALAR=2,
                                P0=1e-2, P1=0.3, P2=0.4, P3=-0.5, P4=0, P5=0, P6=1.0)),
    "m2_q1_fcycle": ("cd", dict(geo="m2", ANSATZ_ORDER=1, LEVELS=4, UNIFORM_STEPS=0, DISCTYPE=2,
                                CELL_MEASURE=4, SDFEM_TYPE=2, DELTA0=1.0, SC_MG_CYCLE_SCALAR=0,
                                P0=1e-2, P1=0.3, P2=0.4, P3=-0.5, P4=0, P5=0, P6=1.0)),
    # SOR and SSOR smoothers in the reference's row order (SC_SMOOTHER_SCALAR
    # 2 and 3, MGLevel3D.C:254-441), as the reference's ReadIn/cd3d.dat (SSOR
    # 3 + 3, omega 1) and ReadIn/tcd3d.dat (SOR 2 + 2) set them; SOR also as
    # the coarse smoother
    "m2_q1_ssor": ("cd", dict(geo="m2", ANSATZ_ORDER=1, LEVELS=3, UNIFORM_STEPS=0, DISCTYPE=2,
                              CELL_MEASURE=4, SDFEM_TYPE=2, DELTA0=1.0, SC_SMOOTHER_SCALAR=3,
                              SC_PRE_SMOOTH_SCALAR=3, SC_POST_SMOOTH_SCALAR=3,
                              SC_SMOOTH_DAMP_FACTOR_SCALAR=1.0, SC_SMOOTH_DAMP_FACTOR_FINE_SCALAR=1.0,
                              P0=1e-2, P1=0.3, P2=0.4, P3=-0.5, P4=0, P5=0, P6=1.0)),
    "unit_q2_sor": ("cd", dict(geo="unit", ANSATZ_ORDER=2, LEVELS=3, UNIFORM_STEPS=1, DISCTYPE=2,
                               CELL_MEASURE=4, SDFEM_TYPE=2, DELTA0=1.0, SC_SMOOTHER_SCALAR=2,
                               SC_COARSE_SMOOTHER_SCALAR=2,
                               SC_SMOOTH_DAMP_FACTOR_SCALAR=1.0, SC_SMOOTH_DAMP_FACTOR_FINE_SCALAR=1.0,
                               P0=1e-3, P1=1.0, P2=0.5, P3=0.25, P4=0, P5=0, P6=1.0)),
    "m2_q2_ssor_w12": ("cd", dict(geo="m2", ANSATZ_ORDER=2, LEVELS=2, UNIFORM_STEPS=0, DISCTYPE=2,
                                  CELL_MEASURE=4, SDFEM_TYPE=2, DELTA0=1.0, SC_SMOOTHER_SCALAR=3,
                                  SC_SMOOTH_DAMP_FACTOR_SCALAR=1.2, SC_SMOOTH_DAMP_FACTOR_FINE_SCALAR=1.2,
                                  P0=1e-2, P1=-0.4, P2=0.6, P3=0.3, P4=0.5, P5=0, P6=1.0)),
    # the smoother of the GPU benchmark (damped Jacobi omega 0.5, 3 + 3 sweeps)
    "unit_c2_n8_bench": ("cd", dict(geo="unit", ANSATZ_ORDER=2, LEVELS=3, UNIFORM_STEPS=1, DISCTYPE=2,
                                    CELL_MEASURE=4, SDFEM_TYPE=2, DELTA0=1.0,
                                    SC_SMOOTH_DAMP_FACTOR_SCALAR=0.5, SC_SMOOTH_DAMP_FACTOR_FINE_SCALAR=0.5,
                                    SC_PRE_SMOOTH_SCALAR=3, SC_POST_SMOOTH_SCALAR=3,
                                    P0=1e-3, P1=1.0, P2=0.5, P3=0.25, P4=0, P5=0, P6=1.0)),
    # SUPG with the rotating field (convection varying from cell to cell, so
    # the element matrices and SUPG delta differ per cell)
    "unit_q1_supg_rot": ("cd", dict(geo="unit", ANSATZ_ORDER=1, LEVELS=3, UNIFORM_STEPS=2, DISCTYPE=2,
                                    CELL_MEASURE=4, SDFEM_TYPE=2, DELTA0=1.0,
                                    SC_SMOOTH_DAMP_FACTOR_SCALAR=0.5, SC_SMOOTH_DAMP_FACTOR_FINE_SCALAR=0.5,
                                    SC_PRE_SMOOTH_SCALAR=3, SC_POST_SMOOTH_SCALAR=3,
                                    P0=1e-3, P4=0, P5=0, P6=1.0, P7=1, P12=0.45, P13=0.55)),
    "m2_q2_supg_rot": ("cd", dict(geo="m2", ANSATZ_ORDER=2, LEVELS=3, UNIFORM_STEPS=0, DISCTYPE=2,
                                  CELL_MEASURE=4, SDFEM_TYPE=2, DELTA0=1.0,
                                  SC_SMOOTH_DAMP_FACTOR_SCALAR=0.5, SC_SMOOTH_DAMP_FACTOR_FINE_SCALAR=0.5,
                                  SC_PRE_SMOOTH_SCALAR=3, SC_POST_SMOOTH_SCALAR=3,
                                  P0=1e-4, P4=0.2, P5=0, P6=1.0, P7=1, P12=0.45, P13=0.55)),
    # coefficients varying in space, evaluated by the reference's CoeffFct3D
    # at every quadrature point (eps = P0 (1 + xyz), b = (P1 sin(pi y),
    # P2 cos(pi x), P3 z), c = P4 x, f = P6 (1 + x)): the per-point
    # coefficient path
    "unit_q1_var_gal": ("cd", dict(geo="unit", ANSATZ_ORDER=1, LEVELS=3, UNIFORM_STEPS=2, DISCTYPE=1,
                                   P0=0.05, P1=0.8, P2=-0.6, P3=0.5, P4=0.3, P5=0, P6=1.0, P7=2)),
    "unit_q1_var_supg": ("cd", dict(geo="unit", ANSATZ_ORDER=1, LEVELS=3, UNIFORM_STEPS=2, DISCTYPE=2,
                                    CELL_MEASURE=4, SDFEM_TYPE=2, DELTA0=1.0,
                                    SC_SMOOTH_DAMP_FACTOR_SCALAR=0.5, SC_SMOOTH_DAMP_FACTOR_FINE_SCALAR=0.5,
                                    SC_PRE_SMOOTH_SCALAR=3, SC_POST_SMOOTH_SCALAR=3,
                                    P0=1e-3, P1=0.8, P2=-0.6, P3=0.5, P4=0.3, P5=0, P6=1.0, P7=2)),
    "m2_q2_var_supg": ("cd", dict(geo="m2", ANSATZ_ORDER=2, LEVELS=3, UNIFORM_STEPS=0, DISCTYPE=2,
                                  CELL_MEASURE=4, SDFEM_TYPE=2, DELTA0=1.0,
                                  SC_SMOOTH_DAMP_FACTOR_SCALAR=0.5, SC_SMOOTH_DAMP_FACTOR_FINE_SCALAR=0.5,
                                  SC_PRE_SMOOTH_SCALAR=3, SC_POST_SMOOTH_SCALAR=3,
                                  P0=1e-3, P1=0.8, P2=-0.6, P3=0.5, P4=0.3, P5=0, P6=1.0, P7=2)),
    # Crank-Nicolson with the rotating field on macro-brick meshes, Q1 and Q2
    "m2_q1_tcd": ("tcd", dict(geo="m2", ANSATZ_ORDER=1, LEVELS=2, UNIFORM_STEPS=1, DISCTYPE=1,
                              P0=1e-4, P4=0, P5=0, P6=0.0, P7=1, P8=3,
                              P9=0.3, P10=0.5, P11=0.5, P12=0.5, P13=0.5,
                              TIME_DISC=2, STARTTIME=0.0, ENDTIME=1.0, TIMESTEPLENGTH=1e-3)),
    # Crank-Nicolson with SOR 2 + 2 (omega 1) and the SSOR coarse smoother, as
    # the reference's ReadIn/tcd3d.dat sets them
    "m2_q1_tcd_sor": ("tcd", dict(geo="m2", ANSATZ_ORDER=1, LEVELS=2, UNIFORM_STEPS=1, DISCTYPE=1,
                                  SC_SMOOTHER_SCALAR=2, SC_SMOOTH_DAMP_FACTOR_SCALAR=1.0,
                                  SC_SMOOTH_DAMP_FACTOR_FINE_SCALAR=1.0,
                                  P0=1e-4, P4=0, P5=0, P6=0.0, P7=1, P8=3,
                                  P9=0.3, P10=0.5, P11=0.5, P12=0.5, P13=0.5,
                                  TIME_DISC=2, STARTTIME=0.0, ENDTIME=1.0, TIMESTEPLENGTH=1e-3)),
    "m2_q2_tcd": ("tcd", dict(geo="m2", ANSATZ_ORDER=2, LEVELS=2, UNIFORM_STEPS=0, DISCTYPE=1,
                              P0=1e-4, P4=0, P5=0, P6=0.0, P7=1, P8=3,
                              P9=0.3, P10=0.5, P11=0.5, P12=0.5, P13=0.5,
                              TIME_DISC=2, STARTTIME=0.0, ENDTIME=1.0, TIMESTEPLENGTH=1e-3)),
    # SUPG-stabilised Crank-Nicolson (MatrixMRhsAssemble_SUPG at t0,
    # MatricesAKRhsAssemble_SUPG per step, the SUPG matrix K beside M): the
    # rotating field with f = 1 (per-cell delta, SUPG rhs), Q1 and Q2, and a
    # constant field (the same element matrices in every cell); the
    # benchmark's smoother (omega 0.5, 3 + 3): with omega 0.8, 2 + 2 the
    # nonsymmetric M + K + tau theta1 A stalls near 1e-7 within 200 iterations
    "m2_q1_tcd_supg": ("tcdsupg", dict(SC_SMOOTH_DAMP_FACTOR_SCALAR=0.5, SC_SMOOTH_DAMP_FACTOR_FINE_SCALAR=0.5,
                                       SC_PRE_SMOOTH_SCALAR=3, SC_POST_SMOOTH_SCALAR=3, geo="m2", ANSATZ_ORDER=1, LEVELS=2, UNIFORM_STEPS=1, DISCTYPE=2,
                                       CELL_MEASURE=4, SDFEM_TYPE=2, DELTA0=1.0,
                                       P0=1e-4, P4=0, P5=0, P6=1.0, P7=1, P8=3,
                                       P9=0.3, P10=0.5, P11=0.5, P12=0.5, P13=0.5,
                                       TIME_DISC=2, STARTTIME=0.0, ENDTIME=1.0, TIMESTEPLENGTH=1e-3)),
    "m2_q2_tcd_supg": ("tcdsupg", dict(SC_SMOOTH_DAMP_FACTOR_SCALAR=0.5, SC_SMOOTH_DAMP_FACTOR_FINE_SCALAR=0.5,
                                       SC_PRE_SMOOTH_SCALAR=3, SC_POST_SMOOTH_SCALAR=3, geo="m2", ANSATZ_ORDER=2, LEVELS=2, UNIFORM_STEPS=0, DISCTYPE=2,
                                       CELL_MEASURE=4, SDFEM_TYPE=2, DELTA0=1.0,
                                       P0=1e-4, P4=0, P5=0, P6=1.0, P7=1, P8=3,
                                       P9=0.3, P10=0.5, P11=0.5, P12=0.5, P13=0.5,
                                       TIME_DISC=2, STARTTIME=0.0, ENDTIME=1.0, TIMESTEPLENGTH=1e-3)),
    "unit_q1_tcd_supg_const": ("tcdsupg", dict(SC_SMOOTH_DAMP_FACTOR_SCALAR=0.5, SC_SMOOTH_DAMP_FACTOR_FINE_SCALAR=0.5,
                                               SC_PRE_SMOOTH_SCALAR=3, SC_POST_SMOOTH_SCALAR=3, geo="unit", ANSATZ_ORDER=1, LEVELS=3, UNIFORM_STEPS=1, DISCTYPE=2,
                                               CELL_MEASURE=4, SDFEM_TYPE=2, DELTA0=1.0,
                                               P0=1e-3, P1=0.6, P2=-0.3, P3=0.4, P4=0.1, P5=0, P6=1.0, P7=0,
                                               P8=3, P9=0.3, P10=0.5, P11=0.5,
                                               TIME_DISC=2, STARTTIME=0.0, ENDTIME=1.0, TIMESTEPLENGTH=1e-2)),
    # general hexahedra: the interior macro vertices of the m^3 box moved
    # (make_box_geo.py "w"), every cell mapped trilinearly (THexaTrilinear)
    "m3w_q1_poisson": ("cd", dict(geo="m3w", ANSATZ_ORDER=1, LEVELS=3, UNIFORM_STEPS=0, DISCTYPE=1,
                                  P0=1.0, P1=0, P2=0, P3=0, P4=0, P5=1)),
    "m3w_q2_supg": ("cd", dict(geo="m3w", ANSATZ_ORDER=2, LEVELS=2, UNIFORM_STEPS=0, DISCTYPE=2,
                               CELL_MEASURE=4, SDFEM_TYPE=2, DELTA0=1.0,
                               SC_SMOOTH_DAMP_FACTOR_SCALAR=0.5, SC_SMOOTH_DAMP_FACTOR_FINE_SCALAR=0.5,
                               SC_PRE_SMOOTH_SCALAR=3, SC_POST_SMOOTH_SCALAR=3,
                               P0=1e-3, P1=1.0, P2=0.5, P3=0.25, P4=0, P5=0, P6=1.0)),
    "m2w_q1_var_supg": ("cd", dict(geo="m2w", ANSATZ_ORDER=1, LEVELS=3, UNIFORM_STEPS=1, DISCTYPE=2,
                                   CELL_MEASURE=4, SDFEM_TYPE=2, DELTA0=1.0,
                                   SC_SMOOTH_DAMP_FACTOR_SCALAR=0.5, SC_SMOOTH_DAMP_FACTOR_FINE_SCALAR=0.5,
                                   SC_PRE_SMOOTH_SCALAR=3, SC_POST_SMOOTH_SCALAR=3,
                                   P0=1e-3, P1=0.8, P2=-0.6, P3=0.5, P4=0.3, P5=0, P6=1.0, P7=2)),
    "m2w_q2_supg_rot": ("cd", dict(geo="m2w", ANSATZ_ORDER=2, LEVELS=3, UNIFORM_STEPS=0, DISCTYPE=2,
                                   CELL_MEASURE=4, SDFEM_TYPE=2, DELTA0=1.0,
                                   SC_SMOOTH_DAMP_FACTOR_SCALAR=0.5, SC_SMOOTH_DAMP_FACTOR_FINE_SCALAR=0.5,
                                   SC_PRE_SMOOTH_SCALAR=3, SC_POST_SMOOTH_SCALAR=3,
                                   P0=1e-4, P4=0.2, P5=0, P6=1.0, P7=1, P12=0.45, P13=0.55)),
    "m2w_q2_tcd_supg": ("tcdsupg", dict(SC_SMOOTH_DAMP_FACTOR_SCALAR=0.5, SC_SMOOTH_DAMP_FACTOR_FINE_SCALAR=0.5,
                                        SC_PRE_SMOOTH_SCALAR=3, SC_POST_SMOOTH_SCALAR=3, geo="m2w", ANSATZ_ORDER=2, LEVELS=2, UNIFORM_STEPS=0, DISCTYPE=2,
                                        CELL_MEASURE=4, SDFEM_TYPE=2, DELTA0=1.0,
                                        P0=1e-4, P4=0, P5=0, P6=1.0, P7=1, P8=3,
                                        P9=0.3, P10=0.5, P11=0.5, P12=0.5, P13=0.5,
                                        TIME_DISC=2, STARTTIME=0.0, ENDTIME=1.0, TIMESTEPLENGTH=1e-3)),
}

# Records only (iteration count and residual history, no level data): the
# configuration-2 problem at 16^3 cells with the configuration-1 smoother
# (omega 0.8, 2 + 2) and with the benchmark's (omega 0.5, 3 + 3)
_C2_16 = dict(geo="unit", ANSATZ_ORDER=2, LEVELS=3, UNIFORM_STEPS=2, DISCTYPE=2,
              CELL_MEASURE=4, SDFEM_TYPE=2, DELTA0=1.0,
              P0=1e-3, P1=1.0, P2=0.5, P3=0.25, P4=0, P5=0, P6=1.0)
RECORDS = {
    "record_c2_n16_omega08": ("cd", dict(_C2_16)),
    "record_c2_n16_omega05": ("cd", dict(_C2_16, SC_SMOOTH_DAMP_FACTOR_SCALAR=0.5,
                                         SC_SMOOTH_DAMP_FACTOR_FINE_SCALAR=0.5,
                                         SC_PRE_SMOOTH_SCALAR=3, SC_POST_SMOOTH_SCALAR=3)),
}


def parse_vtk(path):
    """Points, hexahedra, point data and the cell region ids of a legacy
    ASCII VTK file as TOutput3D::WriteVtk writes it."""
    lines = open(path).read().split("\n")
    out, i = {}, 0
    while i < len(lines):
        w = lines[i].split()
        if w and w[0] == "POINTS":
            n = int(w[1])
            out["vtk_points"] = np.array([[float(v) for v in lines[i + 1 + k].split()] for k in range(n)])
            i += n
        elif w and w[0] == "CELLS":
            n = int(w[1])
            out["vtk_cells"] = np.array([[int(v) for v in lines[i + 1 + k].split()[1:]] for k in range(n)], np.int32)
            i += n
        elif w and w[0] == "CELL_TYPES":
            out["vtk_cell_types"] = np.array([int(v) for v in lines[i + 1].split()], np.int32)
        elif w and w[0] == "SCALARS" and w[1] != "RegionID":
            n = len(out["vtk_points"])
            out["vtk_values"] = np.array([float(lines[i + 2 + k]) for k in range(n)])
            out["vtk_name"] = np.array(w[1])
            i += n
        elif w and w[0] == "SCALARS" and w[1] == "RegionID":
            n = len(out["vtk_cells"])
            out["vtk_region"] = np.array([int(lines[i + 2 + k]) for k in range(n)], np.int32)
            i += n
        i += 1
    out["vtk_text_head"] = np.array("\n".join(lines[:6]))
    return out


def write_dat(path, overrides):
    overrides = dict(overrides)
    geo_kind = overrides.pop("geo", None)
    data = os.path.join(REPO, "oracle", "data")
    if geo_kind == "unit":
        geo, prm = os.path.join(data, "unitcube.GEO"), os.path.join(data, "unitcube.PRM")
    elif geo_kind is not None and geo_kind.startswith("m"):
        # unit cube split into m^3 macro bricks (oracle/data/make_box_geo.py)
        sys.path.insert(0, data)
        from make_box_geo import geo_for_kind
        geo = os.path.join(os.path.dirname(path), f"box_{geo_kind}.GEO")
        with open(geo, "w") as f:
            f.write(geo_for_kind(geo_kind)[1])
        prm = os.path.join(data, "unitcube.PRM")
    else:
        geo, prm = f"{REF}/data/Wuerfel.GEO", f"{REF}/data/Wuerfel.PRM"
    kv = dict(GEOFILE=geo, BNDFILE=prm,
              MESH_TYPE=0, OUTFILE="ref.out", LOGFILE="ref.log", MEASURE_ERRORS=0,
              WRITE_VTK=0, WRITE_PS=0, RE_NR=1.0, USE_ISOPARAMETRIC=0)
    kv.update(SOLVER)
    kv.update(overrides)
    with open(path, "w") as f:
        for k, v in kv.items():
            f.write(f"{k}: {v}\n")


def run_case(name):
    record = name in RECORDS
    mode, ov = RECORDS[name] if record else CASES[name]
    with tempfile.TemporaryDirectory() as td:
        dat = os.path.join(td, "case.dat")
        write_dat(dat, ov)
        p = subprocess.run([EXE, dat, td, mode], cwd=td, capture_output=True, text=True)
        if p.returncode:
            print(p.stdout[-3000:], p.stderr[-3000:])
            raise SystemExit(f"{name}: reference run failed ({p.returncode})")
        out = {"params": np.array(repr(ov)), "box": np.array(UNIT if ov.get("geo") else BOX)}
        gk = ov.get("geo") or ""
        if gk.startswith("m") and not gk[-1].isdigit():  # deformed: the macro vertices the reference read
            sys.path.insert(0, os.path.join(REPO, "oracle", "data"))
            from make_box_geo import geo_for_kind, macro_vertices
            out["macro_xyz"] = macro_vertices(geo_for_kind(gk)[1])
        levels = [tuple(map(int, re.findall(r"-?\d+", l)[:8]))
                  for l in p.stdout.splitlines() if l.startswith("LEVEL ")]
        for lev, nc, nbf, ndof, act, hb, ndir, nnz in levels:
            pre = f"L{lev}_"
            rd = lambda f, dt: np.fromfile(os.path.join(td, pre + f), dtype=dt)
            out[pre + "meta"] = np.array([nc, nbf, ndof, act, hb, ndir, nnz], np.int64)
            out[pre + "cells"] = rd("cells.f64", np.float64).reshape(nc, 8, 3)
            out[pre + "gn"] = rd("gn.i32", np.int32).reshape(nc, nbf)
            out[pre + "rowptr"] = rd("rowptr.i32", np.int32)
            out[pre + "kcol"] = rd("kcol.i32", np.int32)
            out[pre + "entries"] = rd("entries.f64", np.float64)
            out[pre + "rhs"] = rd("rhs.f64", np.float64)
        if mode == "cd":
            for f in ("vcycle_in", "vcycle_out", "sol"):
                out[f] = np.fromfile(os.path.join(td, f + ".f64"), dtype=np.float64)
            m = re.search(r"SOLVE its (\d+) r0 (\S+) rfinal (\S+)", p.stdout)
            out["its"] = np.array(int(m.group(1)))
            out["r0"] = np.array(float(m.group(2)))
            out["rfinal"] = np.array(float(m.group(3)))
            hist = [float(x) for x in re.findall(r"fgmres iteration \d+ (\S+)", p.stdout)]
            out["res_hist"] = np.array(hist)
            t = re.search(r"TIME space (\S+) struct (\S+) assemble (\S+) solve (\S+)", p.stdout)
            out["ref_times"] = np.array([float(x) for x in t.groups()])
            # diagnostics of the solution: L2/H1 errors against the sine of
            # Examples/CD_3D/Laplace.h, its nodal interpolant and that one's
            # errors, the VTK file (parsed) and the binary matrix's header
            for f in ("errors_sol", "errors_interp", "interp_sine"):
                out[f] = np.fromfile(os.path.join(td, f + ".f64"), dtype=np.float64)
            out.update(parse_vtk(os.path.join(td, "sol.vtk")))
            raw = np.fromfile(os.path.join(td, "matrix.bin"), dtype=np.uint8)
            hdr = raw[:12].view(np.int32)
            nrow, nnz = int(hdr[0]), int(hdr[2])
            lf = len(levels) - 1
            rp = raw[12:12 + 4 * (nrow + 1)].view(np.int32)
            kc = raw[12 + 4 * (nrow + 1):12 + 4 * (nrow + 1) + 4 * nnz].view(np.int32)
            ent = raw[12 + 4 * (nrow + 1) + 4 * nnz:].view(np.float64)
            assert np.array_equal(rp, out[f"L{lf}_rowptr"]) and np.array_equal(kc, out[f"L{lf}_kcol"])
            assert np.array_equal(ent, out[f"L{lf}_entries"])
            out["matrix_header"] = hdr.copy()
            out["matrix_bytes"] = np.array(len(raw))
        else:
            steps = int(ov["P8"])
            for k in range(steps + 1):
                out[f"u{k}"] = np.fromfile(os.path.join(td, f"tcd_u{k}.f64"), dtype=np.float64)
            its = [int(x) for x in re.findall(r"TCD step \d+ t \S+ its (\d+)", p.stdout)]
            if its:
                out["its_steps"] = np.array(its)
        if record:
            keep = ("params", "box", "its", "r0", "rfinal", "res_hist", "ref_times")
            out = {k: v for k, v in out.items() if k in keep}
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
        print(f"{name}: {[l[:8] for l in levels]}", "its" in out and int(out["its"]) or "")


def main():
    if not os.path.exists(EXE):
        raise SystemExit("build oracle/_ref first: python oracle/build_ref.py")
    for name in (sys.argv[1:] or list(CASES) + list(RECORDS)):
        run_case(name)


if __name__ == "__main__":
    main()
