"""The UFEM Navier-Stokes workloads: meshes, seeded initial fields, boundary-condition lists and
physics of the benchmark configurations, plus the implicit time step driven through the C ABI.

Boundary conditions follow plugins/UFEM/src/UFEM/BoundaryConditions.cpp (constant vector/scalar BCs,
applied per region in creation order over the region's nodes in ascending order, in increment form
new - old: solver/actions/Proto/DirichletBC.hpp:34-49).  The time step is the action sequence of
ns_implementation/NavierStokes.cpp:103-121: ZeroLSS, LinearizeU, Assembly, BoundaryConditions,
SolveLSS, UpdateActions.
"""
from dataclasses import dataclass, field

import numpy as np

from . import cabi, meshgen

DEFAULT_SEED = 20261202


@dataclass
class Case:
    name: str
    etype: int
    coords: np.ndarray
    conn: np.ndarray
    dims: tuple
    nu: float
    dt: float
    rho: float = 1.0
    theta: float = 1.0
    u0: np.ndarray = None
    p0: np.ndarray = None
    bc_regions: list = field(default_factory=list)   # [(name, nodes, {eq: value or callable(coords)})]
    solver: str = "bicgstab"
    tol: float = 1e-10
    periodic: tuple = None   # (links, active) per mesh node: cf3hip_lss_create_periodic
    t: float = 0.0           # time at the start of the next step (time-dependent conditions)

    @property
    def dim(self):
        return self.coords.shape[1]

    @property
    def n_nodes(self):
        return self.coords.shape[0]

    @property
    def n_elems(self):
        return self.conn.shape[0]


def _face_nodes(dims, axis, side):
    shape = [d + 1 for d in dims]
    idx = np.unravel_index(np.arange(int(np.prod(shape))), shape[::-1])[::-1]
    target = 0 if side == 0 else dims[axis]
    return np.nonzero(idx[axis] == target)[0].astype(np.uint32)


def lid_cavity_2d(n=32, seed=DEFAULT_SEED, nu=0.01, dt=0.01):
    """Unit square, n x n cells split in 2 P1 triangles; lid u=(1,0) on top applied last, no-slip on
    the other edges, p = 0 at node 0 (corner (0,0)); velocity U[-1,1] per node, p = 0."""
    coords, conn = meshgen.rect_tris(n, n)
    N = coords.shape[0]
    u0 = meshgen.uniform(seed, N * 2, -1.0, 1.0).reshape(N, 2)
    dims = (n, n)
    zero = {0: 0.0, 1: 0.0}
    regions = [("left", _face_nodes(dims, 0, 0), zero), ("right", _face_nodes(dims, 0, 1), zero),
               ("bottom", _face_nodes(dims, 1, 0), zero), ("top", _face_nodes(dims, 1, 1), {0: 1.0, 1: 0.0}),
               ("pressure", np.array([0], np.uint32), {2: 0.0})]
    return Case("lid_cavity_2d_%d" % n, cabi.TRIAG2D, coords, conn, dims, nu, dt, u0=u0, p0=np.zeros(N),
                bc_regions=regions, solver="gmres")


SKEW_SITE_PROB = 0.0125   # ~1% of the tetrahedra flattened to aspect ratio ~20


def inflow_cube_tets(n, seed=DEFAULT_SEED, nu=1e-3, dt=5e-3, jitter=0.1, amp=0.0, skew=False, outlet=False):
    """Unit cube, n^3 hexahedra in 6 Kuhn tetrahedra, interior jitter +-10% h (tangential on the
    boundary); no-slip on all faces except x = 1 with u = (-16 y(1-y) z(1-z)(1 + A sin(pi y) sin(pi z)), 0, 0)
    applied last; p = 0 at the node (0, n/2, n/2); velocity U[-1,1] per node, p = 0.  With skew, about
    1% of the elements are then flattened to aspect ratio ~20 (meshgen.skew_kuhn, draws after the
    velocity draws).

    With outlet, the x = 0 face is an outlet instead of a wall: p = 0 on all its nodes and the velocity
    left free there (the outlet of plugins/UFEM/test/atest-flatplate2d-laminar.py), no-slip on the four
    side faces.  Without it the box is closed apart from the inflow face, so the imposed inflow has no
    way out: the continuity equation is only satisfied through the PSPG pressure term and the pressure
    grows like the net flux over tau_PS (|p| ~ 1e5 at 36^3), which leaves a system whose solution is very
    sensitive to perturbations of its matrix."""
    coords, conn = meshgen.box_tetras(n, n, n)
    dims = (n, n, n)
    coords, used = meshgen.jitter(coords, dims, jitter, seed)
    N = coords.shape[0]
    u0 = meshgen.uniform(seed, N * 3, -1.0, 1.0, offset=used).reshape(N, 3)
    if skew:
        coords, _, _ = meshgen.skew_kuhn(coords, dims, SKEW_SITE_PROB, seed, used + 3 * N)
    zero = {0: 0.0, 1: 0.0, 2: 0.0}

    def inflow(c):
        y, z = c[:, 1], c[:, 2]
        return -16.0 * y * (1 - y) * z * (1 - z) * (1 + amp * np.sin(np.pi * y) * np.sin(np.pi * z))
    sides = [("ymin", _face_nodes(dims, 1, 0), zero), ("ymax", _face_nodes(dims, 1, 1), zero),
             ("zmin", _face_nodes(dims, 2, 0), zero), ("zmax", _face_nodes(dims, 2, 1), zero)]
    inflow_bc = ("inflow", _face_nodes(dims, 0, 1), {0: inflow, 1: 0.0, 2: 0.0})
    if outlet:
        regions = sides + [("outlet", _face_nodes(dims, 0, 0), {3: 0.0}), inflow_bc]
    else:
        regions = [("xmin", _face_nodes(dims, 0, 0), zero)] + sides + [
            inflow_bc, ("pressure", np.array([(n // 2 * (n + 1) + n // 2) * (n + 1)], np.uint32), {3: 0.0})]
    name = "inflow_cube_tets_%d%s%s" % (n, "_outlet" if outlet else "", "_skew" if skew else "")
    return Case(name, cabi.TETRA3D, coords, conn, dims, nu, dt, u0=u0, p0=np.zeros(N), bc_regions=regions)


def lid_cavity_cube_hexas(n, seed=DEFAULT_SEED, nu=2e-3, dt=5e-2, jitter=0.05, random_velocity=False):
    """Unit cube, n^3 trilinear Q1 hexahedra (create_box node order), node jitter +-5% h (tangential on
    the boundary); lid u = (1, 0, 0) on z = 1 applied last, no-slip on the other faces, p = 0 at node 0.
    Velocity starts at rest (or U[-1,1] per node with random_velocity, for parity checks), p = 0.  The
    steady solve is a sequence of Picard iterations, each an implicit step of size dt with u_adv = u."""
    coords, conn = meshgen.box_hexas(n, n, n)
    dims = (n, n, n)
    coords, used = meshgen.jitter(coords, dims, jitter, seed)
    N = coords.shape[0]
    if random_velocity:
        u0 = meshgen.uniform(seed, N * 3, -1.0, 1.0, offset=used).reshape(N, 3)
    else:
        u0 = np.zeros((N, 3))
    zero = {0: 0.0, 1: 0.0, 2: 0.0}
    regions = [("xmin", _face_nodes(dims, 0, 0), zero), ("xmax", _face_nodes(dims, 0, 1), zero),
               ("ymin", _face_nodes(dims, 1, 0), zero), ("ymax", _face_nodes(dims, 1, 1), zero),
               ("zmin", _face_nodes(dims, 2, 0), zero), ("lid", _face_nodes(dims, 2, 1), {0: 1.0, 1: 0.0, 2: 0.0}),
               ("pressure", np.array([0], np.uint32), {3: 0.0})]
    return Case("lid_cavity_cube_hexas_%d" % n, cabi.HEXA3D, coords, conn, dims, nu, dt, u0=u0, p0=np.zeros(N),
                bc_regions=regions)


def lid_cavity_cube_prisms(n, seed=DEFAULT_SEED, nu=2e-3, dt=5e-2, jitter=0.05, random_velocity=False):
    """The cavity of lid_cavity_cube_hexas on prisms: each of the n^3 cells split into two prisms along its
    xy diagonal (meshgen.box_prisms), the same node jitter, conditions and initial fields."""
    coords, conn = meshgen.box_prisms(n, n, n)
    dims = (n, n, n)
    coords, used = meshgen.jitter(coords, dims, jitter, seed)
    N = coords.shape[0]
    if random_velocity:
        u0 = meshgen.uniform(seed, N * 3, -1.0, 1.0, offset=used).reshape(N, 3)
    else:
        u0 = np.zeros((N, 3))
    zero = {0: 0.0, 1: 0.0, 2: 0.0}
    regions = [("xmin", _face_nodes(dims, 0, 0), zero), ("xmax", _face_nodes(dims, 0, 1), zero),
               ("ymin", _face_nodes(dims, 1, 0), zero), ("ymax", _face_nodes(dims, 1, 1), zero),
               ("zmin", _face_nodes(dims, 2, 0), zero), ("lid", _face_nodes(dims, 2, 1), {0: 1.0, 1: 0.0, 2: 0.0}),
               ("pressure", np.array([0], np.uint32), {3: 0.0})]
    return Case("lid_cavity_cube_prisms_%d" % n, cabi.PRISM3D, coords, conn, dims, nu, dt, u0=u0, p0=np.zeros(N),
                bc_regions=regions)


def bc_lists(case, u, p):
    """Ordered (nodes, eqs, values) of the Dirichlet increments new - old for the current fields."""
    nodes, eqs, vals = [], [], []
    for _, rn, spec in case.bc_regions:
        c = case.coords[rn]
        for eq, v in sorted(spec.items()):
            new = v(c) if callable(v) else np.full(len(rn), float(v))
            old = p[rn] if eq == case.dim else u[rn, eq]
            nodes.append(rn)
            eqs.append(np.full(len(rn), eq, np.uint32))
            vals.append(new - old)
    # region by region, node by node: within a region the equations of one node are consecutive
    out_n, out_e, out_v = [], [], []
    for rn, spec_eqs, spec_vals in _group(case, nodes, eqs, vals):
        out_n.append(rn); out_e.append(spec_eqs); out_v.append(spec_vals)
    return (np.concatenate(out_n).astype(np.uint32), np.concatenate(out_e).astype(np.uint32),
            np.concatenate(out_v))


def _group(case, nodes, eqs, vals):
    i = 0
    for _, rn, spec in case.bc_regions:
        ne = len(spec)
        n_stack = np.stack(nodes[i:i + ne], axis=1).ravel()
        e_stack = np.stack(eqs[i:i + ne], axis=1).ravel()
        v_stack = np.stack(vals[i:i + ne], axis=1).ravel()
        i += ne
        yield n_stack, e_stack, v_stack


def periodic_links(coords, pairs):
    """LinkPeriodicNodes::execute (cf3/mesh/actions/LinkPeriodicNodes.cpp:136-180) for each
    (source nodes, destination nodes, translation) in order: source node s is linked to the destination
    node d with |x_s + t - x_d|^2 < 1e-8 (the last such d; a later link of the same node replaces an
    earlier one).  Returns (links, active) per mesh node."""
    n = coords.shape[0]
    links = np.zeros(n, dtype=np.uint32)
    active = np.zeros(n, dtype=np.uint8)
    for src, dst, t in pairs:
        dst = np.asarray(dst)
        for s in src:
            d2 = ((coords[dst] - (coords[s] + np.asarray(t))) ** 2).sum(axis=1)
            m = np.nonzero(d2 < 1e-8)[0]
            if len(m):
                links[s] = dst[m[-1]]
                active[s] = 1
    return links, active


def periodic_p2m(links, active):
    """Block row of every mesh node under periodic links, as create_map_data on one rank
    (cf3/math/LSS/Trilinos/TrilinosDetail.cpp:147-227): the unlinked nodes in node order, a linked node
    the row of the end of its chain of links."""
    n = len(links)
    p2m = np.zeros(n, dtype=np.uint32)
    free = np.flatnonzero(active == 0)
    p2m[free] = np.arange(len(free), dtype=np.uint32)
    for i in np.flatnonzero(active):
        f, hops = int(links[i]), 0
        while active[f]:
            f, hops = int(links[f]), hops + 1
            if hops > n:
                raise ValueError("periodic links form a cycle")
        p2m[i] = p2m[f]
    return p2m


TG_UA, TG_VA, TG_D, TG_NU = 0.3, 0.2, 0.5, 0.001


def taylor_green_exact(coords, t, ua=TG_UA, va=TG_VA, d=TG_D, nu=TG_NU):
    """The translating Taylor-Green vortex of plugins/UFEM/test/atest-ufem-navier-stokes-taylor-green.py
    (check_result, :289-291): velocity (N x 2) and pressure at time t."""
    x, y = coords[:, 0], coords[:, 1]
    e = np.exp(-2.0 * nu * np.pi ** 2 / d ** 2 * t)
    u = ua - np.cos(np.pi / d * (x - ua * t)) * np.sin(np.pi / d * (y - va * t)) * e
    v = va + np.sin(np.pi / d * (x - ua * t)) * np.cos(np.pi / d * (y - va * t)) * e
    p = -0.25 * (np.cos(2 * np.pi / d * (x - ua * t)) + np.cos(2 * np.pi / d * (y - va * t))) * \
        np.exp(-4.0 * nu * np.pi ** 2 / d ** 2 * t)
    return np.stack([u, v], axis=1), p


def taylor_green_case(segments=32, dt=0.004, theta=0.5, element="quad"):
    """The Taylor-Green acceptance case of plugins/UFEM/test/atest-ufem-navier-stokes-taylor-green.py
    (plugins/UFEM/test/CMakeLists.txt:94-97 runs it with --dt=0.004 --elem=quad --segs=32 --theta=0.5
    --tsteps=5): the square [-0.5, 0.5]^2 in segments^2 quadrilaterals (the script's two BlockArrays
    blocks), or each cell split in two triangles
    (--elem=triag), periodic right -> left and top -> bottom (:85-97), density 1 and dynamic viscosity
    0.001 (:163-164), the generic expression (use_specializations false, :183), the exact velocity and
    pressure at t = 0 (setup_ic, :124-141) and the exact pressure at t + dt imposed at the centre node
    (add_pressure_bc, :166-169).  The reference bounds the errors after the fifth step by 1.4e-3 (u, v)
    and 3.3e-3 (p)."""
    if element == "quad":
        # BlockArrays of the script (:39-76): two blocks of segments x segments/2 cells, four patches
        from . import blockmesh
        blocks = blockmesh.BlockArrays()
        blocks.create_points(2, 6)[:] = [[-0.5, -0.5], [0.5, -0.5], [-0.5, 0.0], [0.5, 0.0], [-0.5, 0.5], [0.5, 0.5]]
        bl = blocks.create_blocks(2)
        bl[0], bl[1] = [0, 1, 3, 2], [2, 3, 5, 4]
        sd = blocks.create_block_subdivisions()
        sd[0], sd[1] = [segments, segments // 2], [segments, segments // 2]
        blocks.create_block_gradings()
        lp = blocks.create_patch_nb_faces("left", 2)
        lp[0], lp[1] = [2, 0], [4, 2]
        blocks.create_patch_nb_faces("bottom", 1)[0] = [0, 1]
        blocks.create_patch_nb_faces("top", 1)[0] = [5, 4]
        rp = blocks.create_patch_nb_faces("right", 2)
        rp[0], rp[1] = [1, 3], [3, 5]
        m = blocks.create_mesh()
        coords, conn = m.coords, m.blocks["Quad2D"]
        etype = cabi.QUAD2D
    elif element == "triag":
        coords, conn = meshgen.rect_tris(segments, segments)   # the script's MeshTriangulator split is not restated
        coords = coords - 0.5
        etype = cabi.TRIAG2D
    else:
        raise ValueError("element is 'quad' or 'triag'")
    x, y = coords[:, 0], coords[:, 1]
    tol = 1e-12
    right, left = np.flatnonzero(np.abs(x - 0.5) < tol), np.flatnonzero(np.abs(x + 0.5) < tol)
    top, bottom = np.flatnonzero(np.abs(y - 0.5) < tol), np.flatnonzero(np.abs(y + 0.5) < tol)
    links = periodic_links(coords, [(right, left, (-1.0, 0.0)), (top, bottom, (0.0, -1.0))])
    center = np.flatnonzero((np.abs(x) < tol) & (np.abs(y) < tol)).astype(np.uint32)
    if len(center) != 1:
        raise ValueError("the centre (0, 0) must be a mesh node: use an even number of segments")
    u0, p0 = taylor_green_exact(coords, 0.0)
    case = Case("taylor_green_%s_%d" % (element, segments), etype, coords, conn, (segments, segments), TG_NU, dt,
                theta=theta, u0=np.ascontiguousarray(u0), p0=p0, solver="gmres", tol=1e-12, periodic=links)
    case.bc_regions = [("center", center, {2: lambda c, case=case: taylor_green_exact(c, case.t + case.dt)[1]})]
    return case


def si_laminar_channel_case(refinement=1):
    """The channel of plugins/UFEM/test/atest-ufem-navier-stokes-semi-implicit-laminar-channel-2d.py: two
    BlockArrays blocks [0, 10] x [0, 1] and [0, 10] x [1, 2] of (20 r) x (16 r) quadrilaterals (:35-58),
    linked periodically right -> left (:77-82), density and dynamic viscosity 1 (:84-85), time step 0.5
    (:87), velocity 0 and body force (2, 0) initially (:107-117), no-slip on the bottom and top and the
    pressure 10 at the point region (5, 1) (:119-123).  The flow it settles to is the Poiseuille profile
    u = y (2 - y), v = 0, which the script checks at every node (:150-156).  The velocity conditions are
    given once per LSS row: a periodic partner of a wall node carries the same zero value."""
    from . import blockmesh
    blocks = blockmesh.BlockArrays()
    blocks.create_points(2, 6)[:] = [[0.0, 0.0], [10.0, 0.0], [0.0, 1.0], [10.0, 1.0], [0.0, 2.0], [10.0, 2.0]]
    bl = blocks.create_blocks(2)
    bl[0], bl[1] = [0, 1, 3, 2], [2, 3, 5, 4]
    sd = blocks.create_block_subdivisions()
    sd[0], sd[1] = [20 * refinement, 16 * refinement], [20 * refinement, 16 * refinement]
    blocks.create_block_gradings()
    lp = blocks.create_patch_nb_faces("left", 2)
    lp[0], lp[1] = [2, 0], [4, 2]
    blocks.create_patch_nb_faces("bottom", 1)[0] = [0, 1]
    blocks.create_patch_nb_faces("top", 1)[0] = [5, 4]
    rp = blocks.create_patch_nb_faces("right", 2)
    rp[0], rp[1] = [1, 3], [3, 5]
    m = blocks.create_mesh()
    coords, conn = m.coords, m.blocks["Quad2D"]
    x, y = coords[:, 0], coords[:, 1]
    tol = 1e-12
    right, left = np.flatnonzero(np.abs(x - 10.0) < tol), np.flatnonzero(np.abs(x) < tol)
    links = periodic_links(coords, [(right, left, (-10.0, 0.0))])
    p2m = periodic_p2m(*links)
    walls = np.flatnonzero((np.abs(y) < tol) | (np.abs(y - 2.0) < tol))
    _, first = np.unique(p2m[walls], return_index=True)
    walls = np.sort(walls[first]).astype(np.uint32)
    center = np.flatnonzero((np.abs(x - 5.0) < tol) & (np.abs(y - 1.0) < tol)).astype(np.uint32)
    N = coords.shape[0]
    case = Case("si_laminar_channel_%d" % refinement, cabi.QUAD2D, coords, conn, None, 1.0, 0.5, theta=0.5,
                u0=np.zeros((N, 2)), p0=np.zeros(N), periodic=links)
    case.bc_regions = [("walls", walls, {0: 0.0, 1: 0.0}), ("center", center, {2: 10.0})]
    case.body_force = np.tile([2.0, 0.0], (N, 1))
    return case


def mesh_case(path, etype_name=None, nu=1e-2, dt=1e-2, seed=DEFAULT_SEED, bcs=(), pin=None):
    """A case on a mesh file (meshio.read_mesh: .neu or .msh) with one element type: velocity U[-1,1] per
    node from the seed, p = 0.  bcs: [(boundary set, {eq: value or callable(coords)})], applied in that
    order over each set's nodes in ascending order; pin: (boundary set, eq) pins that variable to 0 at
    the set's first node."""
    from . import meshio
    m = meshio.read_mesh(path)
    if etype_name is None:
        if len(m.blocks) != 1:
            raise ValueError("mesh holds several element types; name one of %s" % sorted(m.blocks))
        etype_name = next(iter(m.blocks))
    etype = {"Triag2D": cabi.TRIAG2D, "Tetra3D": cabi.TETRA3D, "Hexa3D": cabi.HEXA3D, "Quad2D": cabi.QUAD2D}[etype_name]
    conn = m.blocks[etype_name]
    used = np.unique(conn.ravel())
    if used.size != m.n_nodes:
        raise ValueError("the %s cells do not use every node of the mesh" % etype_name)
    N, dim = m.n_nodes, m.dim
    u0 = meshgen.uniform(seed, N * dim, -1.0, 1.0).reshape(N, dim)
    regions = [(name, m.boundary_nodes(name), dict(spec)) for name, spec in bcs]
    if pin is not None:
        regions.append(("pin", m.boundary_nodes(pin[0])[:1], {pin[1]: 0.0}))
    return Case("%s_%s" % (m.name, etype_name), etype, m.coords, conn, None, nu, dt, u0=u0, p0=np.zeros(N),
                bc_regions=regions)


class GpuCase:
    """A case resident on the device: mesh, fields, LSS system; one implicit step per call."""

    def __init__(self, ctx, case, precision=cabi.FP64, use_specializations=True, supg_type=cabi.SUPG_METRIC,
                 fallback=True):
        self.ctx = ctx
        self.case = case
        self.mesh = cabi.Mesh(ctx, case.etype, case.coords, case.conn)
        self.fields = cabi.Fields(self.mesh)
        self.system = cabi.System(ctx, self.mesh, periodic=case.periodic)
        self.p2m = self.system.p2m()   # block row of each mesh node (conditions are given by mesh node)
        self.precision = precision
        if precision == cabi.FP32:
            self.system.enable_fp32()
        self.fields.set(cabi.FIELD_U, case.u0)
        self.fields.set(cabi.FIELD_P, case.p0)
        self.fields.initialize(case.nu)
        self.prm = cabi.ns_params(rho=case.rho, dt=case.dt, theta=case.theta,
                                  use_specializations=use_specializations, supg_type=supg_type)
        self.u = case.u0.copy()
        self.p = case.p0.copy()
        self.fallback = fallback
        self.fallback_params = None   # solver parameters of the FP64 redo (default: those of the step)
        self.fallbacks = 0

    def assemble(self, precision=cabi.FP64):
        prm = cabi.NSParams.from_buffer_copy(self.prm)
        prm.precision = precision
        self.system.assemble(self.fields, prm)

    def step(self, solver_params=None, raise_on_failure=True):
        self.fields.linearize()                                   # LinearizeU
        self.assemble(self.precision)                             # ZeroLSS + Assembly (FP64 or FP32 values)
        return self._solve_update(solver_params, raise_on_failure)

    def post_assembly_step(self, solver_params=None, raise_on_failure=True):
        """BoundaryConditions, SolveLSS and UpdateActions of a step whose LinearizeU and Assembly ran
        elsewhere (the batched assembly of the serving loop)."""
        return self._solve_update(solver_params, raise_on_failure)

    def picard_step(self, solver_params=None, raise_on_failure=True):
        """One Picard iteration of the steady solve: u_adv := u, then the implicit step."""
        self.fields.copy(cabi.FIELD_UADV, cabi.FIELD_U)
        self.assemble(self.precision)
        return self._solve_update(solver_params, raise_on_failure)

    def _solve_update(self, solver_params, raise_on_failure):
        """BoundaryConditions, SolveLSS, UpdateActions.  In the precision form a solve that does not
        reach the tolerance (the solver stops early when its FP64 residual stagnates) is redone with
        FP64 values assembled from the same state, from a zero initial guess; `fallbacks` counts them."""
        s = self.system
        n, e, v = bc_lists(self.case, self.u, self.p)             # BoundaryConditions
        n = self.p2m[n]
        s.dirichlet(n, e, v)
        if solver_params is None:
            if self.case.solver == "gmres" and self.precision == cabi.FP64:
                solver_params = cabi.solver_params(cabi.GMRES_ILU0, tol=self.case.tol)
            else:
                solver_params = cabi.solver_params(tol=self.case.tol, precision=self.precision)
        prm = cabi.NSParams.from_buffer_copy(self.prm)
        prm.precision = solver_params.precision
        fallback = self.fallback and solver_params.precision == cabi.FP32
        st = s.ns_solve(self.fields, prm, solver_params, raise_on_failure and not fallback)   # SolveLSS
        if fallback and not st.converged:
            self.fallbacks += 1
            self.assemble(cabi.FP64)
            s.reset_vector(cabi.SOLUTION)
            s.dirichlet(n, e, v)
            sp64 = cabi.SolverParams.from_buffer_copy(self.fallback_params or solver_params)
            sp64.precision = cabi.FP64
            prm.precision = cabi.FP64
            it32 = st.iterations
            st = s.ns_solve(self.fields, prm, sp64, raise_on_failure)
            st.iterations += it32
        s.update(self.fields)                                     # UpdateActions
        self.u = self.fields.get(cabi.FIELD_U)
        self.p = self.fields.get(cabi.FIELD_P)
        self.case.t += self.case.dt
        return st

    def close(self):
        self.system.close()
        self.fields.close()
        self.mesh.close()
