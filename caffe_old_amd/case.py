"""Case description and mesh-node generation for the CAFFE SIMPLE/PISO hot path.

A :class:`Case` carries what the reference reads from ``case/case.info``,
``case/initialConditions.info`` and ``mesh/structuredMesh.dat``:

* ``RunControl.timeStep`` / ``maxNumberOfIterations`` (RunControl.cpp:38-51),
* ``Solver.{timeAccurate, relaxationFactorMomentum, relaxationFactorPCorr, rho, mu,
  numberOfInnerIterations, numberOfPressureCorrections}`` (Solver.cpp:8,
  Simple.cpp:61-67, 104-108, Piso.cpp:9),
* ``Boundaries.{east..bottom}.{type, refVector, refValue}``
  (FlowBoundaryConditions.cpp:25-42),
* ``ImmersedBoundaries.Sphere.{radius, center}`` (IbPiso.cpp:37-38),
* uniform initial ``u`` and ``p`` (InitialConditions.cpp:12-74, uniform only); for the
  Diffusion solver ``p_init`` is the initial ``phi`` and the boundaries are fixed /
  zeroGradient / empty with refValue,
* the node array of the structured mesh, (nK+1, nJ+1, nI+1, 3) in C order so that
  i runs fastest in memory, as ``Array3D`` stores it (Array3D.tpp:40-58).

Node generators restate StructuredMesh::initializeCartesianMesh
(StructuredMesh.cpp:131-148) and the transfinite fill of HexaMeshGen::generateMesh
(HexaMeshGen.cpp:32-79) with the same IEEE operation order, plus the 6-significant-
digit round trip that HexaMeshGen::writeMeshFile (HexaMeshGen.cpp:81-110, default
ostream precision) and StructuredMesh::initialize (StructuredMesh.cpp:27-67) put
the nodes through.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

FACES = ("east", "west", "north", "south", "top", "bottom")
# FlowBoundaryConditions::Type (FlowBoundaryConditions.h:10), same strings as case.info;
# "fixed" is the PrimitiveBoundaryCondition type the Diffusion solver reads
# (PrimitiveBoundaryCondition.cpp:33-58), where zeroGradient and empty also apply
BC_TYPES = {"inlet": 0, "outlet": 1, "wall": 2, "empty": 3, "zeroGradient": 4, "fixed": 5}
FLOW_BC = ("inlet", "outlet", "wall", "empty", "zeroGradient")
SCALAR_BC = ("fixed", "zeroGradient", "empty")
ALGOS = {"simple": 0, "piso": 1, "ibpiso": 2, "diffusion": 3, "coupled": 4}


@dataclass
class Boundary:
    type: str = "wall"
    refVector: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    refValue: float = 0.0

    def __post_init__(self):
        if self.type not in BC_TYPES:
            # FlowBoundaryConditions.cpp:41-42 raises on an unknown type
            raise ValueError(f'invalid boundary type "{self.type}"')


@dataclass
class Case:
    nodes: np.ndarray
    algo: str = "simple"
    time_accurate: bool = False
    rho: float = 1.0
    mu: float = 0.01
    alpha_u: float = 0.7
    alpha_p: float = 0.3
    n_inner: int = 1
    n_pcorr: int = 1
    boundaries: List[Boundary] = field(default_factory=lambda: [Boundary() for _ in FACES])
    u_init: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    p_init: float = 0.0
    sphere: Optional[Tuple[float, Tuple[float, float, float]]] = None
    time_step: float = 1.0
    max_iterations: int = 2
    # corrected mode: the non-orthogonal term mu_f (grad u)_f . c_f of Simple.cpp:153-158 with
    # the faces of grad u filled (interior: volume-weighted interpolation, domain faces: the
    # boundary cell's gradient); False keeps the reference, where those faces stay zero
    non_orth: bool = False

    def __post_init__(self):
        self.nodes = np.ascontiguousarray(self.nodes, dtype=np.float64)
        if self.nodes.ndim != 4 or self.nodes.shape[3] != 3:
            raise ValueError("nodes must be (nK+1, nJ+1, nI+1, 3)")
        if self.algo not in ALGOS:
            raise ValueError(f"unknown algorithm {self.algo}")
        if len(self.boundaries) != 6:
            raise ValueError("six boundaries (east, west, north, south, top, bottom) are required")
        if self.algo == "ibpiso" and self.sphere is None:
            raise ValueError("ibpiso needs an immersed sphere")
        allowed = SCALAR_BC if self.algo == "diffusion" else FLOW_BC
        for b in self.boundaries:
            if b.type not in allowed:
                # FlowBoundaryConditions.cpp:41-42 / PrimitiveBoundaryCondition.cpp:49-50
                raise ValueError(f'invalid boundary type "{b.type}" for the {self.algo} solver')

    @property
    def shape(self) -> Tuple[int, int, int]:
        """(nI, nJ, nK) cell counts."""
        nk, nj, ni, _ = self.nodes.shape
        return ni - 1, nj - 1, nk - 1

    @property
    def n_cells(self) -> int:
        a, b, c = self.shape
        return a * b * c

    @property
    def n_solves(self) -> int:
        """RunControl::continueRun increments then tests >=, so M iterations run M-1 solves
        (RunControl.cpp:53-63)."""
        return max(self.max_iterations - 1, 0)


# ------------------------------------------------------------------ node generators
def cartesian_nodes(lx: float, ly: float, lz: float, ni: int, nj: int, nk: int) -> np.ndarray:
    """StructuredMesh::initializeCartesianMesh (StructuredMesh.cpp:131-148): node (i,j,k) =
    (i*hx, j*hy, k*hz) with h = L/(nNodes-1); ni, nj, nk are cell counts."""
    hx, hy, hz = lx / ni, ly / nj, lz / nk
    k, j, i = np.meshgrid(np.arange(nk + 1), np.arange(nj + 1), np.arange(ni + 1), indexing="ij")
    return np.stack([i * hx, j * hy, k * hz], axis=-1).astype(np.float64)


def transfinite_nodes(vertices: Sequence[Sequence[float]], ni: int, nj: int, nk: int) -> np.ndarray:
    """HexaMeshGen::generateMesh (HexaMeshGen.cpp:32-79) for ni x nj x nk cells, vertices v1..v8
    in the order of structuredMesh.info.  Each expression is evaluated as the reference writes
    it, (a - b)*s + b, so the result is bit-identical."""
    v = np.asarray(vertices, dtype=np.float64)
    ui, uj, uk = ni, nj, nk
    nd = np.zeros((nk + 1, nj + 1, ni + 1, 3))
    for k in range(uk + 1):
        s = float(k) / float(uk)
        nd[k, 0, 0] = (v[4] - v[0]) * s + v[0]
        nd[k, uj, 0] = (v[7] - v[3]) * s + v[3]
        nd[k, 0, ui] = (v[5] - v[1]) * s + v[1]
        nd[k, uj, ui] = (v[6] - v[2]) * s + v[2]
        t1 = nd[k, uj, 0] - nd[k, 0, 0]
        t2 = nd[k, uj, ui] - nd[k, 0, ui]
        sj = (np.arange(uj + 1, dtype=np.float64) / float(uj))[:, None]
        nd[k, :, 0] = t1[None, :] * sj + nd[k, 0, 0][None, :]
        nd[k, :, ui] = t2[None, :] * sj + nd[k, 0, ui][None, :]
    si = (np.arange(ui + 1, dtype=np.float64) / float(ui))[None, None, :, None]
    t1 = nd[:, :, ui, :] - nd[:, :, 0, :]
    nd[:, :, :, :] = t1[:, :, None, :] * si + nd[:, :, 0, :][:, :, None, :]
    return nd


def tecplot_roundtrip(nodes: np.ndarray) -> np.ndarray:
    """The node values after hexaMesh writes them with the default ostream precision of six
    significant digits (HexaMeshGen.cpp:97-110) and the solver reads them back
    (StructuredMesh.cpp:47-59)."""
    flat = np.asarray(nodes, dtype=np.float64).ravel()
    out = np.array([float("%g" % x) for x in flat], dtype=np.float64)
    return out.reshape(nodes.shape)


def box_vertices(lx: float, ly: float, lz: float):
    return [(0, 0, 0), (lx, 0, 0), (lx, ly, 0), (0, ly, 0), (0, 0, lz), (lx, 0, lz), (lx, ly, lz), (0, ly, lz)]


# Philox-4x32-10 (Salmon et al., SC'11): counter-based, so host and device draw the same
# number for (seed, node, component) without shared state.
_PH_M0, _PH_M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
_PH_W0, _PH_W1 = np.uint32(0x9E3779B9), np.uint32(0xBB67AE85)


def philox4x32(ctr: np.ndarray, key: Tuple[int, int]) -> np.ndarray:
    """ctr: (..., 4) uint32 counters -> (..., 4) uint32 random words."""
    c = [ctr[..., q].astype(np.uint32) for q in range(4)]
    k0 = np.uint32(key[0])
    k1 = np.uint32(key[1])
    mask = np.uint64(0xFFFFFFFF)
    for _ in range(10):
        p0 = c[0].astype(np.uint64) * _PH_M0
        p1 = c[2].astype(np.uint64) * _PH_M1
        hi0 = (p0 >> np.uint64(32)).astype(np.uint32)
        lo0 = (p0 & mask).astype(np.uint32)
        hi1 = (p1 >> np.uint64(32)).astype(np.uint32)
        lo1 = (p1 & mask).astype(np.uint32)
        c = [hi1 ^ c[1] ^ k0, lo1, hi0 ^ c[3] ^ k1, lo0]
        k0 = np.uint32((int(k0) + int(_PH_W0)) & 0xFFFFFFFF)
        k1 = np.uint32((int(k1) + int(_PH_W1)) & 0xFFFFFFFF)
    return np.stack(c, axis=-1)


def perturb_interior(nodes: np.ndarray, fraction: float, h: float, seed: int = 0) -> np.ndarray:
    """Move every interior node by fraction*h*U(-1,1) per component; boundary nodes stay, so the
    domain faces stay planar (HexaMeshGen::checkMesh, HexaMeshGen.cpp:118-124).  The uniform
    draw for (node n, component c) is word c of Philox-4x32-10 at counter (n, 0, 0, 0) with key
    (seed, 0), mapped to (w + 0.5) / 2^32 in (0, 1)."""
    nd = np.array(nodes, dtype=np.float64, copy=True)
    nk1, nj1, ni1, _ = nd.shape
    n = np.arange(nk1 * nj1 * ni1, dtype=np.uint64)
    ctr = np.zeros((n.size, 4), dtype=np.uint32)
    ctr[:, 0] = (n & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    ctr[:, 1] = (n >> np.uint64(32)).astype(np.uint32)
    w = philox4x32(ctr, (seed, 0))[:, :3].astype(np.float64)
    uni = (w + 0.5) / 4294967296.0
    delta = (fraction * h) * (2.0 * uni - 1.0)
    delta = delta.reshape(nk1, nj1, ni1, 3)
    interior = np.zeros((nk1, nj1, ni1), dtype=bool)
    interior[1:-1, 1:-1, 1:-1] = True
    nd[interior] += delta[interior]
    return nd


# ------------------------------------------------------------------ the benchmark cases
def _walls(lid=(1.0, 0.0, 0.0)):
    b = [Boundary("wall") for _ in FACES]
    b[2] = Boundary("wall", tuple(lid), 0.0)  # north wall (y = 1) moves
    return b


def lid_cavity(n: int, mu: float, lid: float = 1.0, algo: str = "simple", *, rho: float = 1.0,
               alpha_u: float = 0.7, alpha_p: float = 0.3, time_accurate: bool = False,
               dt: float = 1.0, n_pcorr: int = 1, n_inner: int = 1, max_iterations: int = 201,
               nodes: Optional[np.ndarray] = None) -> Case:
    """3D lid-driven cavity on the unit cube with n^3 uniform hex cells, lid u = (lid,0,0) on
    y = 1 (north), no-slip elsewhere, zero initial u and p."""
    if nodes is None:
        nodes = cartesian_nodes(1.0, 1.0, 1.0, n, n, n)
    return Case(nodes=nodes, algo=algo, time_accurate=time_accurate, rho=rho, mu=mu, alpha_u=alpha_u,
                alpha_p=alpha_p, n_inner=n_inner, n_pcorr=n_pcorr, boundaries=_walls((lid, 0.0, 0.0)),
                time_step=dt, max_iterations=max_iterations)


def cavity_simple_32() -> Case:
    """Plain method, small: 32^3, rho 1, mu 0.01 (Re 100), SIMPLE alpha_u 0.7 alpha_p 0.3,
    steady, 200 outer iterations (maxNumberOfIterations 201)."""
    return lid_cavity(32, 0.01, 1.0, "simple", alpha_u=0.7, alpha_p=0.3, max_iterations=201)


def cavity_piso_128(n: int = 128, steps: int = 100) -> Case:
    """128^3 cavity, Re 1000 (mu 0.001, lid 1), PISO with 2 correctors, dt = CFL*h/U with
    CFL 0.5, one inner iteration, alpha_u = alpha_p = 1 (recorded choice)."""
    h = 1.0 / n
    return lid_cavity(n, 1e-3, 1.0, "piso", alpha_u=1.0, alpha_p=1.0, time_accurate=True,
                      dt=0.5 * h / 1.0, n_pcorr=2, n_inner=1, max_iterations=steps + 1)


def channel_ib_sphere(ni: int = 256, nj: int = 64, nk: int = 64, steps: int = 500) -> Case:
    """Box 4x1x1, uniform inlet u = 1 (west), zero-gradient outlet (east, p = 0), no-slip walls,
    sphere diameter 0.2 at (1.0, 0.5, 0.5), Re_D 200 (mu = 1e-3), PISO 2 correctors,
    dt = 0.5 h / U, alpha_u = 0.7, alpha_p = 0.3 (recorded choice: the configuration leaves
    them open; with 1 / 1 IbPiso, the oracle's restatement and this library alike, diverges
    within about ten steps on the 64x16x16, 128x32x32 and 256x64x64 channels, with 0.7 / 0.3
    it stays bounded over the 40 steps measured)."""
    nodes = cartesian_nodes(4.0, 1.0, 1.0, ni, nj, nk)
    b = [Boundary("outlet", (0, 0, 0), 0.0), Boundary("inlet", (1.0, 0, 0), 0.0)] + [Boundary("wall") for _ in range(4)]
    h = 4.0 / ni
    return Case(nodes=nodes, algo="ibpiso", time_accurate=True, rho=1.0, mu=1e-3, alpha_u=0.7, alpha_p=0.3,
                n_inner=1, n_pcorr=2, boundaries=b, u_init=(1.0, 0.0, 0.0), p_init=0.0,
                sphere=(0.1, (1.0, 0.5, 0.5)), time_step=0.5 * h, max_iterations=steps + 1)


def skewed_cavity(n: int = 256, steps: int = 50, fraction: float = 0.15, seed: int = 0) -> Case:
    """n^3 cavity whose interior nodes are moved by +/-fraction of the cell size (seeded Philox),
    Re 1000, PISO 2 correctors, dt = 0.5 h."""
    h = 1.0 / n
    nodes = perturb_interior(cartesian_nodes(1.0, 1.0, 1.0, n, n, n), fraction, h, seed)
    return lid_cavity(n, 1e-3, 1.0, "piso", alpha_u=1.0, alpha_p=1.0, time_accurate=True, dt=0.5 * h,
                      n_pcorr=2, n_inner=1, max_iterations=steps + 1, nodes=nodes)
