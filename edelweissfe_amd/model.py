"""Model tree for the GPU path: nodes, node sets, element blocks with their material and quadrature-point
state, and the boxGen structured-mesh generator.

Mirrors the parts of the reference's model that the element-assembly path reads:
  * FEModel (edelweissfe/models/femodel.py:45-297): nodes in insertion order, nodeSets, elements,
    ``advanceToTime`` -> every element accepts its last state (femodel.py:245-258,
    marmotelement/element.pyx:278-281);
  * boxGen (edelweissfe/generators/boxgen.py:91-448): node order ix -> iy -> iz, C3D8 / C3D20
    connectivity in Abaqus order, the face / edge / vertex node sets and element sets;
  * sections with per-element material parameters (edelweissfe/sections/base/sectionbase.py:86-128).

Elements are held as blocks (one connectivity array per element type + material) rather than one Python
object per element, so a 2-million-element mesh is built in seconds; ``ElementBlock`` carries the
per-element API the solver needs (nNodes, nDof, fields, acceptLastState).
"""

from __future__ import annotations

import numpy as np

from ._lib import ELEMENT_KINDS, MATERIAL_KINDS

ELEMENT_LIBRARY = {
    # nNodes, nInt (quadrature points), formulation (edelweissfe/elements/library.py:212-227, 260-275, 453-516)
    "C3D8": dict(nNodes=8, nInt=8, tl=False),
    "C3D20": dict(nNodes=20, nInt=27, tl=False),
    "C3D8TL": dict(nNodes=8, nInt=8, tl=True),
    "C3D20TL": dict(nNodes=20, nInt=27, tl=True),
    "C3D20R": dict(nNodes=20, nInt=8, tl=False),  # library.py:276-290, 2x2x2 points
    "C3D20RTL": dict(nNodes=20, nInt=8, tl=True),  # library.py:517-531
}
# other names of the same elements (library.py:539-550)
ELEMENT_ALIASES = {"C3D8N": "C3D8", "C3D20N": "C3D20", "C3D8NTL": "C3D8TL", "C3D20NTL": "C3D20TL"}


def elementType(name: str) -> str:
    """The element library's name of an element type (aliases resolved, upper case)."""
    et = name.upper()
    return ELEMENT_ALIASES.get(et, et)

MATERIAL_PARAMETERS = {
    # name -> names of the material properties (the reference material classes' constructors)
    "linearelastic": ("E", "nu"),  # linearelastic.py
    "vonmises": ("E", "nu", "fy0", "HLin", "dfy", "delta"),  # vonmises.py:105-124
    "neohookeanwa": ("mu", "K"),  # neohookepencegouformulationa.py:72-78
    # St. Venant-Kirchhoff (the Marmot material StVenantKirchhoffIsotropic of the C3D8TL regression deck)
    "stvenantkirchhoff": ("E", "nu"),
}
# materials of the total-Lagrangian elements (the small-strain elements take the others)
FINITE_STRAIN_MATERIALS = ("neohookeanwa", "stvenantkirchhoff")


class Material:
    """A material definition: name + property vector (the reference's ``*material, name=..`` block)."""

    def __init__(self, name: str, materialProperties):
        key = name.lower()
        if key not in MATERIAL_PARAMETERS:
            raise AttributeError(f"material {name} is not available on the GPU path "
                                 f"(available: {', '.join(MATERIAL_PARAMETERS)})")
        props = np.asarray(materialProperties, dtype=np.float64).ravel()
        need = len(MATERIAL_PARAMETERS[key])
        if key == "vonmises" and len(props) == 4:  # linear hardening given without the nonlinear terms
            props = np.concatenate([props, [0.0, 0.0]])
        if len(props) < need:
            raise ValueError(f"material {name} needs {need} properties {MATERIAL_PARAMETERS[key]}, got {len(props)}")
        self.name = key
        self.materialProperties = props
        self.kind = MATERIAL_KINDS[key]


class ElementBlock:
    """All elements of one type with materials of one kind (sections applied to element sets).

    conn           int32 [nEl, nNodes]   model node indices, element-local (Abaqus) node order
    labels         int64 [nEl]           element labels
    elParams       float64 [nEl] or None per-element value of the material's first property
                   (section ``materialParameterFromField``, sectionbase.py:86-128)
    materials      list of Material      the block's materials (one per section), all of one kind;
                   ``material`` is the first
    materialIndex  int32 [nEl] or None   the material of every element (None: all take ``material``;
                   every section assigns its material to its elements, sectionbase.py:84-112)
    """

    def __init__(self, elType: str, conn, labels=None, material: Material | None = None, elParams=None):
        et = elementType(elType)
        if et not in ELEMENT_LIBRARY:
            raise AttributeError(f"element type {elType} is not available on the GPU path")
        self.elType = et
        self.kind = ELEMENT_KINDS[et]
        self.conn = np.ascontiguousarray(conn, dtype=np.int32)
        lib = ELEMENT_LIBRARY[et]
        if self.conn.ndim != 2 or self.conn.shape[1] != lib["nNodes"]:
            raise ValueError(f"{et} connectivity must have shape [nEl, {lib['nNodes']}]")
        self.nNodes = lib["nNodes"]
        self.nDof = 3 * self.nNodes
        self.nInt = lib["nInt"]
        self.labels = np.arange(1, len(self.conn) + 1) if labels is None else np.asarray(labels)
        self.material = material
        self.materials = [] if material is None else [material]
        self.materialIndex = None
        self.elParams = None if elParams is None else np.ascontiguousarray(elParams, dtype=np.float64)
        self.fields = [["displacement"]] * self.nNodes
        # quadrature-point state (device tensors, plan order), owned here like the reference's elements own
        # theirs; created by the solver's assembly on first use
        self.stateVars = None
        self.stateVarsTemp = None
        self.planOrder = None

    def __len__(self):
        return len(self.conn)

    @property
    def hasMaterial(self):
        return self.material is not None

    def setMaterial(self, material: Material):
        self.setMaterials([material])

    def setMaterials(self, materials, materialIndex=None):
        """Materials of one kind and, with several, the material index of every element."""
        materials = list(materials)
        if not materials:
            raise ValueError("a block needs at least one material")
        for material in materials:
            if self.elType.endswith("TL") != (material.name in FINITE_STRAIN_MATERIALS):
                raise AttributeError(f"{self.elType} cannot take material {material.name}: small-strain elements "
                                     "take linearelastic / vonmises, total-Lagrangian elements take neohookeanwa / "
                                     "stvenantkirchhoff")
        if len({m.name for m in materials}) != 1:
            raise NotImplementedError("one element block takes materials of one kind (one kernel per evaluation)")
        if len(materials) > 1:
            idx = np.ascontiguousarray(materialIndex, dtype=np.int32)
            if idx.shape != (len(self),) or idx.min() < 0 or idx.max() >= len(materials):
                raise ValueError(f"materialIndex must hold one index in [0, {len(materials)}) per element")
            self.materialIndex = idx
        else:
            self.materialIndex = None
        self.materials = materials
        self.material = materials[0]

    def materialParameterTable(self):
        """float64 [nMaterials, 8]: the parameter sets of the block's materials (zero padded)."""
        t = np.zeros((len(self.materials), 8))
        for i, m in enumerate(self.materials):
            t[i, : min(8, len(m.materialProperties))] = m.materialProperties[:8]
        return t

    def elementMaterialProperties(self):
        """float64 [nEl, 8]: every element's parameters, the per-element first property included."""
        t = self.materialParameterTable()
        out = t[self.materialIndex] if self.materialIndex is not None else np.repeat(t[:1], len(self), axis=0)
        if self.elParams is not None:
            out[:, 0] = self.elParams
        return out

    def acceptLastState(self):
        """state_n <- state_trial (marmotelement/element.pyx:278-281)."""
        if self.stateVarsTemp is not None:
            self.stateVars.copy_(self.stateVarsTemp)


class FEModel:
    """Nodes (insertion order), node sets, element blocks, element sets, time."""

    def __init__(self, domainSize: int = 3):
        if domainSize != 3:
            raise NotImplementedError("the GPU path covers 3D solid elements only")
        self.domainSize = 3
        self.nodeCoordinates = np.zeros((0, 3))
        self.nodeLabels = np.zeros(0, dtype=np.int64)
        self.nodeSets: dict[str, np.ndarray] = {}
        self.elementBlocks: list[ElementBlock] = []
        self.elementSets: dict[str, tuple[int, np.ndarray]] = {}  # name -> (block index, element indices)
        # name -> {face number (1..6, Abaqus numbering): element set name} (abqmodelconstructor.py:213-225)
        self.surfaces: dict[str, dict[int, str]] = {}
        self.time = 0.0
        self.constraints = {}
        self.scalarVariables = {}
        self.analyticalFields = {}  # name -> f(x, y, z)
        self.fieldOutputs = {}  # name -> (result, element set, quadrature points)

    def addNodes(self, coords, labels=None):
        coords = np.asarray(coords, dtype=np.float64).reshape(-1, 3)
        start = len(self.nodeCoordinates)
        if labels is None:
            first = int(self.nodeLabels.max()) + 1 if len(self.nodeLabels) else 1
            labels = np.arange(first, first + len(coords))
        self.nodeCoordinates = np.concatenate([self.nodeCoordinates, coords])
        self.nodeLabels = np.concatenate([self.nodeLabels, np.asarray(labels, dtype=np.int64)])
        return np.arange(start, start + len(coords))

    def addElementBlock(self, block: ElementBlock, setName: str | None = None):
        self.elementBlocks.append(block)
        if setName:
            self.elementSets[setName] = (len(self.elementBlocks) - 1, np.arange(len(block)))
        return block

    @property
    def nNodes(self):
        return len(self.nodeCoordinates)

    def advanceToTime(self, time: float):
        """femodel.py:245-258."""
        self.time = time
        for b in self.elementBlocks:
            b.acceptLastState()


# --------------------------------------------------------------------------------------------------
# quadrature points
# --------------------------------------------------------------------------------------------------
_HEX_NODES = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1], [-1, -1, 1], [1, -1, 1], [1, 1, 1],
                       [-1, 1, 1], [0, -1, -1], [1, 0, -1], [0, 1, -1], [-1, 0, -1], [0, -1, 1], [1, 0, 1],
                       [0, 1, 1], [-1, 0, 1], [-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]], dtype=float)


def quadraturePoints(nInt: int) -> np.ndarray:
    """Parametric coordinates [nInt, 3] in the reference's order (library.py:212-227, 260-275): 2x2x2: xi = -g
    for the first four, (eta, zeta) = (-,-), (-,+), (+,+), (+,-); 3x3x3: q = 9 i + 3 j + k over (xi, eta, zeta)."""
    if nInt == 8:
        g = 1.0 / np.sqrt(3.0)
        eta, zeta = np.array([-1, -1, 1, 1]), np.array([-1, 1, 1, -1])
        return g * np.stack([np.repeat([-1.0, 1.0], 4), np.tile(eta, 2), np.tile(zeta, 2)], axis=1)
    g = np.sqrt(0.6)
    pts = np.array([-g, 0.0, g])
    i, j, k = np.meshgrid(np.arange(3), np.arange(3), np.arange(3), indexing="ij")
    return np.stack([pts[i.ravel()], pts[j.ravel()], pts[k.ravel()]], axis=1)


def shapeFunctions(nNodes: int, xi: np.ndarray) -> np.ndarray:
    """N [nPoints, nNodes] of the trilinear (8) or serendipity (20) hexahedron, Abaqus node order."""
    c = _HEX_NODES[:nNodes]
    x = np.atleast_2d(xi)[:, None, :]  # [p, 1, 3]
    f = 1 + c[None] * x  # [p, a, 3]
    if nNodes == 8:
        return 0.125 * f.prod(axis=2)
    N = np.empty((x.shape[0], 20))
    N[:, :8] = 0.125 * f[:, :8].prod(axis=2) * ((c[None, :8] * x).sum(axis=2) - 2)
    mid = np.where(c[None, 8:] == 0, 1 - x * x, f[:, 8:])
    N[:, 8:] = 0.25 * mid.prod(axis=2)
    return N


def shapeFunctionDerivatives(nNodes: int, xi: np.ndarray) -> np.ndarray:
    """dN/dxi [nPoints, nNodes, 3] of the trilinear (8) or serendipity (20) hexahedron, Abaqus node order."""
    c = _HEX_NODES[:nNodes]
    x = np.atleast_2d(xi)[:, None, :]  # [p, 1, 3]
    f = 1 + c[None] * x  # [p, a, 3]
    dN = np.empty((x.shape[0], nNodes, 3))
    for d in range(3):
        o = [k for k in range(3) if k != d]
        if nNodes == 8:
            dN[:, :, d] = 0.125 * c[None, :, d] * f[:, :, o[0]] * f[:, :, o[1]]
            continue
        s = (c[None, :8] * x).sum(axis=2) - 2
        dN[:, :8, d] = 0.125 * c[None, :8, d] * f[:, :8, o[0]] * f[:, :8, o[1]] * (s + f[:, :8, d])
        cm = c[None, 8:]
        fm = np.where(cm == 0, 1 - x * x, f[:, 8:])
        dfm = np.where(cm == 0, -2 * x, cm)
        dN[:, 8:, d] = 0.25 * dfm[:, :, d] * fm[:, :, o[0]] * fm[:, :, o[1]]
    return dN


def quadraturePointStrains(block, coords, U) -> np.ndarray:
    """[nEl, nInt, 6] small strains at the quadrature points (engineering Voigt [11,22,33,12,13,23]) of the
    nodal displacements U ([nNodes, 3] in model node order)."""
    dN = shapeFunctionDerivatives(block.nNodes, quadraturePoints(block.nInt))  # [q, a, i]
    Xe = np.asarray(coords)[block.conn]  # [e, a, j]
    Ue = np.asarray(U)[block.conn]
    J = np.einsum("qai,eaj->eqij", dN, Xe)
    g = np.einsum("eqik,qak->eqai", np.linalg.inv(J), dN)  # g_a = J^-1 dN_a
    H = np.einsum("eak,eqai->eqki", Ue, g)  # du_k / dx_i
    return np.stack([H[..., 0, 0], H[..., 1, 1], H[..., 2, 2], H[..., 0, 1] + H[..., 1, 0],
                     H[..., 0, 2] + H[..., 2, 0], H[..., 1, 2] + H[..., 2, 1]], axis=-1)


def quadraturePointCoordinates(block, coords) -> np.ndarray:
    """[nEl, nInt, 3] coordinates of the quadrature points of every element of the block (reference
    configuration; getCoordinatesAtQuadraturePoints of the reference's elements)."""
    N = shapeFunctions(block.nNodes, quadraturePoints(block.nInt))  # [q, a]
    return np.einsum("qa,eac->eqc", N, np.asarray(coords)[block.conn])


# --------------------------------------------------------------------------------------------------
# boxGen
# --------------------------------------------------------------------------------------------------
def boxGenArrays(nX, nY, nZ, lX=1.0, lY=1.0, lZ=1.0, x0=0.0, y0=0.0, z0=0.0, elType="C3D8", name="gen"):
    """Vectorised boxGen (boxgen.py:125-371). Returns (coords [nNodes,3], conn [nEl,npe] int32, nodeSets,
    gridIndex [nNodes,3] int (ix, iy, iz) on the generator's node grid)."""
    npe = ELEMENT_LIBRARY[elementType(elType)]["nNodes"]
    if npe == 8:
        nx, ny, nz = nX + 1, nY + 1, nZ + 1
    else:
        nx, ny, nz = 2 * nX + 1, 2 * nY + 1, 2 * nZ + 1
    ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    if npe == 8:
        present = np.ones((nx, ny, nz), dtype=bool)
    else:  # C3D20 meshes skip nodes with two or more odd grid indices (boxgen.py:139)
        present = ((ix % 2) + (iy % 2) + (iz % 2)) < 2
    g2m = np.full((nx, ny, nz), -1, dtype=np.int64)
    g2m[present] = np.arange(int(present.sum()))
    xs, ys, zs = np.linspace(x0, x0 + lX, nx), np.linspace(y0, y0 + lY, ny), np.linspace(z0, z0 + lZ, nz)
    grid = np.stack([ix[present], iy[present], iz[present]], axis=1)
    coords = np.stack([xs[grid[:, 0]], ys[grid[:, 1]], zs[grid[:, 2]]], axis=1)

    ex, ey, ez = (a.ravel() for a in np.meshgrid(np.arange(nX), np.arange(nY), np.arange(nZ), indexing="ij"))
    if npe == 8:
        offs = [(0, 0, 0), (0, 0, 1), (1, 0, 1), (1, 0, 0), (0, 1, 0), (0, 1, 1), (1, 1, 1), (1, 1, 0)]
        base = (ex, ey, ez)
    else:
        offs = [(0, 0, 0), (0, 0, 2), (2, 0, 2), (2, 0, 0), (0, 2, 0), (0, 2, 2), (2, 2, 2), (2, 2, 0),
                (0, 0, 1), (1, 0, 2), (2, 0, 1), (1, 0, 0), (0, 2, 1), (1, 2, 2), (2, 2, 1), (1, 2, 0),
                (0, 1, 0), (0, 1, 2), (2, 1, 2), (2, 1, 0)]
        base = (2 * ex, 2 * ey, 2 * ez)
    conn = np.stack([g2m[base[0] + a, base[1] + b, base[2] + c] for a, b, c in offs], axis=1).astype(np.int32)

    def filt(sl):
        sub = g2m[sl]
        return sub[sub >= 0].ravel()

    sets = {}
    for nm, sl in [("top", np.s_[:, -1, :]), ("bottom", np.s_[:, 0, :]), ("right", np.s_[-1, :, :]),
                   ("left", np.s_[0, :, :]), ("front", np.s_[:, :, -1]), ("back", np.s_[:, :, 0]),
                   ("bottomRight", np.s_[-1, 0, :]), ("bottomLeft", np.s_[0, 0, :]),
                   ("bottomFront", np.s_[:, 0, -1]), ("bottomBack", np.s_[:, 0, 0]),
                   ("topRight", np.s_[-1, -1, :]), ("topLeft", np.s_[0, -1, :]), ("topFront", np.s_[:, -1, -1]),
                   ("topBack", np.s_[:, -1, 0]), ("rightBack", np.s_[-1, :, 0]), ("rightFront", np.s_[-1, :, -1]),
                   ("leftBack", np.s_[0, :, 0]), ("leftFront", np.s_[0, :, -1]),
                   ("centerX", np.s_[nx // 2, :, :]), ("centerY", np.s_[:, ny // 2, :]),
                   ("centerZ", np.s_[:, :, nz // 2])]:
        sets[f"{name}_{nm}"] = filt(sl)
    for nm, (i, j, k) in [("bottomRightFront", (-1, 0, -1)), ("bottomRightBack", (-1, 0, 0)),
                          ("bottomLeftFront", (0, 0, -1)), ("bottomLeftBack", (0, 0, 0)),
                          ("topRightFront", (-1, -1, -1)), ("topRightBack", (-1, -1, 0)),
                          ("topLeftFront", (0, -1, -1)), ("topLeftBack", (0, -1, 0))]:
        sets[f"{name}_{nm}"] = np.array([g2m[i, j, k]])
    return coords, conn, sets, grid


def pipeGenArrays(nT, nY, nC, radius, lT=0.5, lY=1.0, phi=360.0, x0=0.0, y0=0.0, z0=0.0, outer=True,
                  reduceMidRadius=False, elType="C3D8", name="gen"):
    """pipeGen (pipegen.py:91-357): a pipe segment about the y axis, nT elements through the thickness lT, nY
    along the height lY, nC around the angle phi (degrees; 360 closes the ring). radius(h) is the outer
    (outer=True) or inner radius at height h = y - y0. Node grid (iy, it, ic): x = x0 + r sin(c),
    z = z0 + r cos(c), with r on nNodesT points from R - lT to R (outer) or R to R + lT; 20-node meshes skip
    nodes with two or more odd grid indices and, with reduceMidRadius, place the circumferentially mid-side nodes
    at r cos(dc). Elements in the order it, iy, ic. Returns (coords, conn, nodeSets, elementSets, surfaces)
    with node and element indices; surfaces map a set name to {face: element set name}."""
    npe = ELEMENT_LIBRARY[elementType(elType)]["nNodes"]
    m = 1 if npe == 8 else 2
    nNT, nNY, nNC = m * nT + 1, m * nY + 1, m * nC + 1
    closed = phi == 360
    nCp = nNC - (1 if closed else 0)  # node columns around the ring (the closing column is the first one)
    cLayers = np.linspace(0.0, np.pi * phi / 180.0, nNC)
    yLayers = np.linspace(y0, y0 + lY, nNY)
    iy, it, ic = np.meshgrid(np.arange(nNY), np.arange(nNT), np.arange(nCp), indexing="ij")
    present = np.ones(iy.shape, dtype=bool) if npe == 8 else ((iy % 2) + (it % 2) + (ic % 2)) < 2
    R = np.array([radius(y - y0) for y in yLayers], dtype=np.float64)
    # radial positions [iy, it]: the generator's linspace per layer, between R - lT and R (outer) or R and R + lT
    t = np.stack([np.linspace(Ry - lT, Ry, nNT) if outer else np.linspace(Ry, Ry + lT, nNT) for Ry in R])
    r = np.broadcast_to(t[:, :, None], iy.shape).copy()
    if npe == 20 and reduceMidRadius:
        r[:, :, 1::2] *= np.cos((np.pi * phi / 180.0) / nCp)
    c = cLayers[ic]
    X = np.stack([x0 + r * np.sin(c), yLayers[iy], z0 + r * np.cos(c)], axis=-1)
    g2m = np.full(iy.shape, -1, dtype=np.int64)
    g2m[present] = np.arange(int(present.sum()))
    coords = X[present]

    if npe == 8:
        offs = [(0, 0, 0), (0, 0, 1), (1, 0, 1), (1, 0, 0), (0, 1, 0), (0, 1, 1), (1, 1, 1), (1, 1, 0)]
    else:
        offs = [(0, 0, 0), (0, 0, 2), (2, 0, 2), (2, 0, 0), (0, 2, 0), (0, 2, 2), (2, 2, 2), (2, 2, 0),
                (0, 0, 1), (1, 0, 2), (2, 0, 1), (1, 0, 0), (0, 2, 1), (1, 2, 2), (2, 2, 1), (1, 2, 0),
                (0, 1, 0), (0, 1, 2), (2, 1, 2), (2, 1, 0)]
    et, ey, ec = (a.ravel() for a in np.meshgrid(np.arange(nT), np.arange(nY), np.arange(nC), indexing="ij"))
    conn = np.stack([g2m[m * ey + dy, m * et + dt, (m * ec + dc) % nCp] for dy, dt, dc in offs], axis=1)
    conn = conn.astype(np.int32)

    def filt(sl):
        sub = g2m[sl]
        return sub[sub >= 0].ravel()

    sets = {}
    for nm, sl in [("outer", np.s_[:, -1, :]), ("inner", np.s_[:, 0, :]), ("top", np.s_[-1, :, :]),
                   ("bottom", np.s_[0, :, :]), ("end", np.s_[:, :, -1]), ("begin", np.s_[:, :, 0]),
                   ("beginTop", np.s_[-1, :, 0]), ("endTop", np.s_[-1, :, -1]), ("beginBottom", np.s_[0, :, 0]),
                   ("endBottom", np.s_[0, :, -1]), ("innerEnd", np.s_[:, 0, -1]), ("innerBegin", np.s_[:, 0, 0]),
                   ("innerTop", np.s_[-1, 0, :]), ("innerBottom", np.s_[0, 0, :]), ("outerEnd", np.s_[:, -1, -1]),
                   ("outerBegin", np.s_[:, -1, 0]), ("outerTop", np.s_[-1, -1, :]),
                   ("outerBottom", np.s_[0, -1, :]), ("centerY", np.s_[int(nNY / 2), :, :]),
                   ("centerT", np.s_[:, int(nNT / 2), :]), ("centerC", np.s_[:, :, int(nCp / 2)])]:
        sets[f"{name}_{nm}"] = filt(sl)
    for nm, (i, j, k) in [("topEndInner", (-1, 0, -1)), ("topBeginInner", (-1, 0, 0)),
                          ("bottomEndInner", (0, 0, -1)), ("bottomBeginInner", (0, 0, 0)),
                          ("topEndOuter", (-1, -1, -1)), ("topBeginOuter", (-1, -1, 0)),
                          ("bottomEndOuter", (0, -1, -1)), ("bottomBeginOuter", (0, -1, 0))]:
        sets[f"{name}_{nm}"] = np.array([g2m[i, j, k]])

    eg = np.arange(nT * nY * nC).reshape(nT, nY, nC)
    esets = {f"{name}_all": eg.ravel(), f"{name}_bottom": eg[:, 0, :].ravel(), f"{name}_top": eg[:, -1, :].ravel(),
             f"{name}_outer": eg[-1, :, :].ravel(), f"{name}_inner": eg[0, :, :].ravel(),
             f"{name}_end": eg[:, :, -1].ravel(), f"{name}_begin": eg[:, :, 0].ravel(),
             f"{name}_centerSliceY": eg[int(nT / 2), :, :].ravel(), f"{name}_centerSliceT": eg[:, int(nY / 2), :].ravel(),
             f"{name}_centerSliceC": eg[:, :, int(nC / 2)].ravel()}
    nBand = min(nT, nY)
    if nBand > 3:
        band = np.concatenate([eg[int(nT / 2 + i - nBand / 2), int(nY / 2 + i - nBand / 2), 0:nC] for i in range(nBand)])
        esets[f"{name}_shearBandInnerToOuter"] = band
        esets[f"{name}_shearBandCenterInnerToOuter"] = band[(int(nBand / 2) - 1) * nC:(int(nBand / 2) + 2) * nC]
    surfaces = {f"{name}_{k}": {face: f"{name}_{k}"} for k, face in
                (("bottom", 1), ("top", 2), ("outer", 5), ("inner", 3), ("end", 4), ("begin", 6))}
    return coords, conn, sets, esets, surfaces


# boxGen's element-based surfaces: set name suffix -> face number (boxgen.py:438-446); with boxGen's element node
# order (parametric axes xi, eta, zeta along z, x, y) face 1 (zeta = -1) is y = y0, face 4 (xi = +1) is z = z0 + lZ
BOXGEN_SURFACES = {"bottom": 1, "top": 2, "left": 3, "front": 4, "right": 5, "back": 6}


def boxGenElementSets(nX, nY, nZ, name="gen"):
    """boxGen's element sets (boxgen.py:376-416) as element indices: element (ix, iy, iz) has index
    (ix nY + iy) nZ + iz."""
    g = np.arange(nX * nY * nZ).reshape(nX, nY, nZ)
    sets = {f"{name}_all": g.ravel(), f"{name}_bottom": g[:, 0, :].ravel(), f"{name}_top": g[:, -1, :].ravel(),
            f"{name}_right": g[-1, :, :].ravel(), f"{name}_left": g[0, :, :].ravel(),
            f"{name}_front": g[:, :, -1].ravel(), f"{name}_back": g[:, :, 0].ravel(),
            f"{name}_centralFrontToBack": g[int(nX / 2), int(nY / 2), 0:nZ].ravel(),
            f"{name}_centerSliceX": g[int(nX / 2), :, :].ravel(), f"{name}_centerSliceY": g[:, int(nY / 2), :].ravel(),
            f"{name}_centerSliceZ": g[:, :, int(nZ / 2)].ravel()}
    nBand = min(nX, nY)
    if nBand > 3:
        band = np.concatenate([g[int(nX / 2 + i - nBand / 2), int(nY / 2 + i - nBand / 2), 0:nZ].ravel()
                               for i in range(nBand)])
        sets[f"{name}_shearBandFrontToBack"] = band
        sets[f"{name}_shearBandCenterFrontToBack"] = band[(int(nBand / 2) - 1) * nZ: (int(nBand / 2) + 2) * nZ]
    return sets


def jitterNodes(coords, grid, gridShape, h, amplitude, seed):
    """Moves every node by U(-amplitude*h, +amplitude*h) per component, except along the normal of a
    boundary face it lies on, so loaded and fixed faces stay planar. Seeded numpy.random.default_rng."""
    rng = np.random.default_rng(seed)
    d = rng.uniform(-1.0, 1.0, size=coords.shape) * (amplitude * np.asarray(h))[None, :]
    for c in range(3):
        onFace = (grid[:, c] == 0) | (grid[:, c] == gridShape[c] - 1)
        d[onFace, c] = 0.0
    return coords + d


def generateModelData(generatorDefinition: dict, model: FEModel, journal=None) -> FEModel:
    """The boxGen generator plugin (boxgen.py:91) for this package: options nX, nY, nZ, lX, lY, lZ, x0, y0,
    z0, elType; adds nodes, one element block (material assigned later by a section) and the node /
    element sets ``<name>_<face|edge|vertex>``, ``<name>_all``."""
    o = {k: v for k, v in generatorDefinition.get("data", {}).items()}
    name = generatorDefinition.get("name", "boxGen")
    coords, conn, sets, _ = boxGenArrays(int(o.get("nX", 1)), int(o.get("nY", 1)), int(o.get("nZ", 1)),
                                         float(o.get("lX", 1.0)), float(o.get("lY", 1.0)), float(o.get("lZ", 1.0)),
                                         float(o.get("x0", 0.0)), float(o.get("y0", 0.0)), float(o.get("z0", 0.0)),
                                         o["elType"], name)
    first = model.nNodes
    model.addNodes(coords)
    block = ElementBlock(o["elType"], conn + first)
    model.addElementBlock(block, f"{name}_all")
    blockIndex = len(model.elementBlocks) - 1
    for k, v in boxGenElementSets(int(o.get("nX", 1)), int(o.get("nY", 1)), int(o.get("nZ", 1)), name).items():
        model.elementSets[k] = (blockIndex, v)
    for suffix, face in BOXGEN_SURFACES.items():
        model.surfaces[f"{name}_{suffix}"] = {face: f"{name}_{suffix}"}
    for k, v in sets.items():
        model.nodeSets[k] = v + first
    model.nodeSets.setdefault("all", np.arange(model.nNodes))
    model.nodeSets["all"] = np.arange(model.nNodes)
    return model
