"""Algorithmic work model of one RK stage (flops and compulsory HBM bytes).

Flops count + - * / and one per log/exp/pow/sqrt, per element and stage.  The
per-piece constants are exact operation counts of the reference's formulas,
obtained by instantiating the CPU oracle's templated flux / contraction /
node-conversion routines with a counting scalar type (tests check that these
constants still equal those counts):

    c_F    one two-point entropy-conservative flux from node records
           (okl/Euler3D.okl:122-162, okl/Euler2D.okl:139-161)
    c_node conservative -> (rho, u, beta) conversion of one node
    c_vv   averaged-metric contraction of one ordered volume-volume pair
           (okl/Euler3D.okl:2317-2340)
    c_vf   averaged-normal contraction of one volume-face pair, both coupling
           blocks (okl/Euler3D.okl:2383-2403 and :2488-2507), with the
           geofac average, the normal and F.n formed once per pair and
           weighted into both rows: 3D 18 + 15 + 25 + 4 x 5 = 78, 2D
           8 + 6 + 12 + 4 x 4 = 42

Volume kernel work counts each distinct flux once (the EC flux is symmetric):

    F_vol = P_F c_F + Nq^2 c_vv + Nq Nf c_vf + (Nq + Nf) c_node,
    P_F   = Nq (Nq + 1)/2 + Nq Nf.

Compulsory FP64 traffic of the three-kernel dataflow (bytes per element):

    B_vol  = 8 (14 (Nq + Nf) + 5 (Nq + Nf))      state + geofacs in, rhs out
    B_surf = 8 (19 Nf + 10 Nq) + 4 Nf             traces, fgeo, rhsQf, rhsQ RMW, mapPq
    B_upd  = 8 (31 Nq + 5 Nf)                     rhs, res, Q, J in; res, Q, rhsQ, Qf out
"""

from __future__ import annotations

# exact counts (add+mul+div+log), see module docstring
PIECES = {
    3: {"c_F": 75, "c_node": 15, "c_vv": 63, "c_vf": 78},
    2: {"c_F": 63, "c_node": 11, "c_vv": 48, "c_vf": 42},
}

# exact counts of the surface and update kernels per element (reference loop
# structure: per face point two node conversions, one flux, the LF penalty and
# the Nq x Nf lift; four Nq x Nq and one Nf x Nq mat-vec per field plus the
# variable maps), keyed by (dim, Nq, Nf)
SURF_UPDATE = {
    (2, 3, 6): (1098, 852),
    (2, 6, 9): (1869, 2544),
    (2, 12, 12): (3084, 8280),
    (2, 16, 15): (4339, 14092),
    (2, 25, 18): (6526, 31960),
    (3, 4, 12): (2648, 2000),
    (3, 11, 24): (6991, 10590),
    (3, 23, 48): (19747, 41958),
    (3, 44, 64): (39836, 134632),
}

FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 (vector and matrix) peak
HBM_PEAK_GBS = 8000.0     # MI355X HBM3E


def vol_flops(dim, Nq, Nf):
    p = PIECES[dim]
    PF = Nq * (Nq + 1) // 2 + Nq * Nf
    return PF * p["c_F"] + Nq * Nq * p["c_vv"] + Nq * Nf * p["c_vf"] + (Nq + Nf) * p["c_node"]


def vol_flops_reference_structure(dim, Nq, Nf):
    """Work of the reference kernel's loop structure: Nq^2 + 2 Nq Nf flux calls, each
    converting its partner node (okl/Euler3D.okl:2263-2273)."""
    p = PIECES[dim]
    nflux = Nq * Nq + 2 * Nq * Nf
    return nflux * (p["c_F"] + p["c_node"]) + Nq * Nq * p["c_vv"] + Nq * Nf * p["c_vf"]


def update_flops(dim, Nq, Nf):
    """Update work without the last projection VqPq v~ (VqPq is a projection and
    v~ already lies in its range, so u(VqPq v~) = u(v~)): the reference's count
    less one Nq x Nq mat-vec per field."""
    return SURF_UPDATE[(dim, Nq, Nf)][1] - 2 * Nq * Nq * (dim + 2)


def stage_flops(dim, Nq, Nf):
    """Volume (symmetric count) + surface + update flops of one element-stage."""
    return vol_flops(dim, Nq, Nf) + SURF_UPDATE[(dim, Nq, Nf)][0] + update_flops(dim, Nq, Nf)


def bytes_stage_fused(dim, Nq, Nf):
    """Compulsory traffic of the fused stage: state, traces (own + neighbour),
    geometry, map and J in; Q, res, projected state and new traces out."""
    nfld, ng = dim + 2, dim * dim
    rd = nfld * Nq + 2 * nfld * Nf + ng * (Nq + Nf) + (dim + 1) * Nf + Nq + 2 * nfld * Nq
    wr = 3 * nfld * Nq + nfld * Nf
    return 8 * (rd + wr) + 4 * Nf


def bytes_vol(dim, Nq, Nf):
    nfld = dim + 2
    return 8 * ((nfld + dim * dim) + nfld) * (Nq + Nf)


def bytes_surf(dim, Nq, Nf):
    nfld = dim + 2
    return 8 * ((3 * nfld + dim + 1) * Nf + 2 * nfld * Nq) + 4 * Nf


def bytes_update(dim, Nq, Nf):
    nfld = dim + 2
    return 8 * ((6 * nfld + 1) * Nq + nfld * Nf)


def dofs(dim, N, K):
    Np = (N + 1) * (N + 2) // 2 if dim == 2 else (N + 1) * (N + 2) * (N + 3) // 6
    return K * Np * (dim + 2)
