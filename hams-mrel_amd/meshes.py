"""Panel meshes and problem set-up for the boundary-integral hot path (host side).

Builds the inputs the C-ABI takes: hull panels of all bodies combined (the
reference's *_MULTI_COMB arrays, CalGreenFuncMulti.f90:53-106), optional interior
waterplane panels for irregular-frequency removal, and the per-panel geometry of
src/NormalProcess.f90 (centroid :134, area :171/:373, longest edge :401, unit
normal :74, rotational normals about the body's reference centre :46).

Conventions (those of the reference's .pnl meshes): vertices are 1-based node
numbers, panel normals point into the body (out of the fluid), z = 0 is the mean
free surface and the fluid is z < 0.

Meshes are generated deterministically; the only random draws are the ones a
workload names with its seed.
"""
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Body:
    xyz: np.ndarray            # (nnode, 3) global coordinates
    ncn: np.ndarray            # (npanel,)  3 or 4
    ncon: np.ndarray           # (npanel, 4) 1-based node numbers (0 pads triangles)
    xr: tuple = (0.0, 0.0, 0.0)   # rotation centre of the rigid-body modes
    ixyz: np.ndarray = None    # waterplane nodes (IRR), or None
    incn: np.ndarray = None
    incon: np.ndarray = None
    # body frame: the nodes in the body's own frame (local, ilocal) and its
    # horizontal placement, xyz = local + (origin, 0) (the reference's local
    # meshes placed by LCS_MULTI, ReadPanelMeshMulti.f90:52-60); None: the body
    # frame is the global frame
    local: np.ndarray = None
    ilocal: np.ndarray = None
    origin: tuple = (0.0, 0.0)

    @property
    def n(self):
        return len(self.ncn)

    @property
    def ni(self):
        return 0 if self.incn is None else len(self.incn)


# ---------------------------------------------------------------------------
# geometry, NormalProcess.f90
# ---------------------------------------------------------------------------
def _tri_area(p1, p2, p3):
    """CalDeltaArea_Improved (NormalProcess.f90:373-396), Kahan's stable Heron."""
    a = np.sqrt(np.sum((p1 - p2) ** 2, axis=-1))
    b = np.sqrt(np.sum((p3 - p1) ** 2, axis=-1))
    c = np.sqrt(np.sum((p3 - p2) ** 2, axis=-1))
    l1, l2, l3, l4 = a + (b + c), c - (a - b), c + (a - b), a + (b - c)
    return np.sqrt(l1 * l2 * l3 * l4) / 4.0


def panel_geometry(xyz, ncn, ncon, rg=(0.0, 0.0, 0.0)):
    """Centroid, area, longest edge and the 6 generalised normals of each panel."""
    xyz = np.asarray(xyz, dtype=np.float64)
    ncn = np.asarray(ncn)
    idx = np.asarray(ncon) - 1
    quad = ncn == 4
    idx = np.where(idx < 0, idx[:, :1], idx)          # pad triangles with vertex 1
    p = xyz[idx]                                      # (ne, 4, 3)
    cnt = np.where(quad, 4.0, 3.0)[:, None]
    xyz_p = np.where(quad[:, None], p.sum(1), p[:, :3].sum(1)) / cnt
    ds = _tri_area(p[:, 0], p[:, 1], p[:, 2])
    ds = ds + np.where(quad, _tri_area(p[:, 0], p[:, 3], p[:, 2]), 0.0)
    e4 = np.sqrt(np.sum((np.roll(p, -1, axis=1) - p) ** 2, axis=-1))
    e3 = np.stack([np.sqrt(np.sum((p[:, 1] - p[:, 0]) ** 2, -1)),
                   np.sqrt(np.sum((p[:, 2] - p[:, 1]) ** 2, -1)),
                   np.sqrt(np.sum((p[:, 0] - p[:, 2]) ** 2, -1))], axis=1)
    pnsz = np.where(quad, e4.max(1), e3.max(1))
    un_q = np.cross(p[:, 3] - p[:, 1], p[:, 2] - p[:, 0])
    un_t = np.cross(p[:, 0] - p[:, 1], p[:, 2] - p[:, 1])
    un = np.where(quad[:, None], un_q, un_t)
    nrm = un / np.sqrt(np.sum(un * un, axis=1))[:, None]
    rg = np.asarray(rg, dtype=np.float64)
    d = xyz_p - rg
    rot = np.stack([d[:, 1] * nrm[:, 2] - d[:, 2] * nrm[:, 1],
                    d[:, 2] * nrm[:, 0] - d[:, 0] * nrm[:, 2],
                    d[:, 0] * nrm[:, 1] - d[:, 1] * nrm[:, 0]], axis=1)
    return xyz_p, ds, pnsz, np.concatenate([nrm, rot], axis=1)


# ---------------------------------------------------------------------------
# synthetic bodies
# ---------------------------------------------------------------------------
def _orient_inward(xyz, ncon, ncn, inside_of):
    """Reverse the vertex order of panels whose normal points into the fluid."""
    xp, _, _, dx = panel_geometry(xyz, ncn, ncon)
    out = np.sum((xp - inside_of(xp)) * dx[:, :3], axis=1) > 0
    ncon = ncon.copy()
    q = out & (ncn == 4)
    ncon[q] = ncon[q][:, ::-1]
    t = out & (ncn == 3)
    ncon[t, :3] = ncon[t, :3][:, ::-1]
    return ncon


def _disk(nodes, cx, cy, z, radius, n_theta, n_r):
    """Polar disk of n_theta x n_r panels; the innermost ring is degenerate quads
    (two vertices on the centre node), as allowed by SingularIntgr.f90:416."""
    base = len(nodes)
    nodes.append((cx, cy, z))
    th = 2.0 * np.pi * np.arange(n_theta) / n_theta
    for k in range(1, n_r + 1):
        r = radius * k / n_r
        for t in th:
            nodes.append((cx + r * np.cos(t), cy + r * np.sin(t), z))
    ring = lambda k, t: base + 1 + (k - 1) * n_theta + (t % n_theta) + 1   # 1-based
    pan = []
    for t in range(n_theta):
        c = base + 1
        pan.append((c, ring(1, t), ring(1, t + 1), c))
        for k in range(1, n_r):
            pan.append((ring(k, t), ring(k + 1, t), ring(k + 1, t + 1), ring(k, t + 1)))
    return pan


def cylinder(radius, draft, n_theta, n_z, n_r, center=(0.0, 0.0), waterplane_nr=0):
    """Truncated vertical circular cylinder: n_theta x n_z side quads (uniform in
    azimuth x depth) plus an n_theta x n_r polar bottom disk; optional interior
    waterplane disk of n_theta x waterplane_nr panels at z = 0."""
    cx, cy = center
    nodes = []
    th = 2.0 * np.pi * np.arange(n_theta) / n_theta
    zs = -draft * np.arange(n_z + 1) / n_z
    for z in zs:
        for t in th:
            nodes.append((radius * np.cos(t), radius * np.sin(t), z))
    side = lambda k, t: k * n_theta + (t % n_theta) + 1
    pan = []
    for k in range(n_z):
        for t in range(n_theta):
            pan.append((side(k, t), side(k, t + 1), side(k + 1, t + 1), side(k + 1, t)))
    pan += _disk(nodes, 0.0, 0.0, -draft, radius, n_theta, n_r)
    local = np.array(nodes)
    ncon = np.array(pan, dtype=np.int32)
    ncn = np.full(len(pan), 4, dtype=np.int32)

    def inside(xp):
        c = np.empty_like(xp)
        c[:, 0], c[:, 1], c[:, 2] = 0.0, 0.0, -0.5 * draft
        return c
    ncon = _orient_inward(local, ncon, ncn, inside)
    body = Body(xyz=_placed(local, center), ncn=ncn, ncon=ncon, xr=(cx, cy, 0.0), local=local,
                origin=(float(cx), float(cy)))
    if waterplane_nr:
        wn = []
        wp = _disk(wn, 0.0, 0.0, 0.0, radius, n_theta, waterplane_nr)
        body.ilocal = np.array(wn)
        body.ixyz = _placed(body.ilocal, center)
        body.incon = np.array(wp, dtype=np.int32)
        body.incn = np.full(len(wp), 4, dtype=np.int32)
    return body


def _placed(local, center):
    """global nodes of a body frame placed at center"""
    g = np.array(local, dtype=np.float64)
    g[:, 0] = center[0] + g[:, 0]
    g[:, 1] = center[1] + g[:, 1]
    return g


def hemisphere(radius, n_theta, n_phi, center=(0.0, 0.0)):
    """Floating hemisphere (the sphere's lower half, waterline at z = 0): flat
    quads between n_phi latitude rings from the waterline down, closed by
    degenerate quads whose first and last vertex are the bottom pole (the
    bottom disk's convention)."""
    cx, cy = center
    nodes = []
    th = 2.0 * np.pi * np.arange(n_theta) / n_theta
    for k in range(n_phi):
        a = 0.5 * np.pi * k / n_phi
        for t in th:
            nodes.append((radius * np.cos(a) * np.cos(t), radius * np.cos(a) * np.sin(t),
                          -radius * np.sin(a)))
    nodes.append((0.0, 0.0, -radius))
    pole = len(nodes)                                   # 1-based index of the pole
    ring = lambda k, t: k * n_theta + (t % n_theta) + 1
    pan = []
    for k in range(n_phi - 1):
        for t in range(n_theta):
            pan.append((ring(k, t), ring(k, t + 1), ring(k + 1, t + 1), ring(k + 1, t)))
    for t in range(n_theta):
        pan.append((pole, ring(n_phi - 1, t), ring(n_phi - 1, t + 1), pole))
    local = np.array(nodes)
    ncon = np.array(pan, dtype=np.int32)
    ncn = np.full(len(pan), 4, dtype=np.int32)

    def inside(xp):
        c = np.zeros_like(xp)
        return c
    ncon = _orient_inward(local, ncon, ncn, inside)
    return Body(xyz=_placed(local, center), ncn=ncn, ncon=ncon, xr=(cx, cy, 0.0), local=local,
                origin=(float(cx), float(cy)))


def box(length, beam, draft, nx, ny, nz, center=(0.0, 0.0)):
    """Rectangular floating box: sides and bottom meshed with quads."""
    cx, cy = center
    nodes = {}
    pts = []

    def node(p):
        key = tuple(np.round(p, 12))
        if key not in nodes:
            nodes[key] = len(pts) + 1
            pts.append(p)
        return nodes[key]

    xs = cx + length * (np.arange(nx + 1) / nx - 0.5)
    ys = cy + beam * (np.arange(ny + 1) / ny - 0.5)
    zs = -draft * np.arange(nz + 1) / nz
    pan = []
    for i in range(nx):                       # bottom
        for j in range(ny):
            pan.append([node((xs[i], ys[j], -draft)), node((xs[i + 1], ys[j], -draft)),
                        node((xs[i + 1], ys[j + 1], -draft)), node((xs[i], ys[j + 1], -draft))])
    for k in range(nz):                       # sides y = +-beam/2
        for i in range(nx):
            for y in (ys[0], ys[-1]):
                pan.append([node((xs[i], y, zs[k])), node((xs[i + 1], y, zs[k])),
                            node((xs[i + 1], y, zs[k + 1])), node((xs[i], y, zs[k + 1]))])
        for j in range(ny):                   # sides x = +-length/2
            for x in (xs[0], xs[-1]):
                pan.append([node((x, ys[j], zs[k])), node((x, ys[j + 1], zs[k])),
                            node((x, ys[j + 1], zs[k + 1])), node((x, ys[j], zs[k + 1]))])
    xyz = np.array(pts)
    ncon = np.array(pan, dtype=np.int32)
    ncn = np.full(len(pan), 4, dtype=np.int32)

    def inside(xp):
        c = np.empty_like(xp)
        c[:, 0], c[:, 1], c[:, 2] = cx, cy, -0.5 * draft
        return c
    ncon = _orient_inward(xyz, ncon, ncn, inside)
    return Body(xyz=xyz, ncn=ncn, ncon=ncon, xr=(cx, cy, 0.0))


def read_pnl(path, lcs=(0.0, 0.0, 0.0, 0.0), xr=(0.0, 0.0, 0.0)):
    """HAMS .pnl hull or waterplane mesh (ReadPanelMeshMulti.f90:46-111), placed
    by the local coordinate system (x, y, z, rotation about z)."""
    lines = open(path).read().split("\n")
    i = 0
    while "Number of Panels" not in lines[i]:
        i += 1
    ne, nn = map(int, lines[i + 1].split()[:2])
    while "Node Coordinates" not in lines[i]:
        i += 1
    i += 1
    loc = np.array([list(map(float, lines[i + k].split()[1:4])) for k in range(nn)])
    i += nn
    while "Node Relations" not in lines[i]:
        i += 1
    i += 1
    ncn = np.zeros(ne, dtype=np.int32)
    ncon = np.zeros((ne, 4), dtype=np.int32)
    for k in range(ne):
        t = list(map(int, lines[i + k].split()))
        ncn[k] = t[1]
        ncon[k, :t[1]] = t[2:2 + t[1]]
    c, s = np.cos(lcs[3]), np.sin(lcs[3])
    xyz = np.empty_like(loc)
    xyz[:, 0] = lcs[0] + c * loc[:, 0] - s * loc[:, 1]
    xyz[:, 1] = lcs[1] + s * loc[:, 0] + c * loc[:, 1]
    xyz[:, 2] = loc[:, 2]
    # the body frame: the local nodes turned by the placement's rotation
    local = np.empty_like(loc)
    local[:, 0] = c * loc[:, 0] - s * loc[:, 1]
    local[:, 1] = s * loc[:, 0] + c * loc[:, 1]
    local[:, 2] = loc[:, 2]
    return Body(xyz=xyz, ncn=ncn, ncon=ncon, xr=tuple(xr), local=local, origin=(float(lcs[0]), float(lcs[1])))


# ---------------------------------------------------------------------------
# a multi-body problem
# ---------------------------------------------------------------------------
@dataclass
class Problem:
    bodies: list
    h: float                    # depth, < 0 = infinite
    omegas: np.ndarray          # wave frequencies [rad/s]
    beta: np.ndarray            # headings [rad]
    rho: float = 1025.0
    g: float = 9.81
    g_disp: float = 9.81        # gravity inside the dispersion solver
    xw: tuple = (0.0, 0.0)
    amp: float = 1.0
    name: str = ""
    # derived (filled by __post_init__)
    n: int = field(init=False)
    ni: int = field(init=False)
    nbody: int = field(init=False)

    def __post_init__(self):
        self.omegas = np.asarray(self.omegas, dtype=np.float64)
        self.beta = np.asarray(self.beta, dtype=np.float64)
        self.nbody = len(self.bodies)
        starts, xyz, ncn, ncon, xp, ds, pn, dx = [0], [], [], [], [], [], [], []
        ixp = []
        off = 0
        for b in self.bodies:
            g = panel_geometry(b.xyz, b.ncn, b.ncon, b.xr)
            xyz.append(b.xyz)
            ncn.append(b.ncn)
            nc = b.ncon.copy()
            nc[nc > 0] += off
            ncon.append(nc)
            off += len(b.xyz)
            xp.append(g[0]); ds.append(g[1]); pn.append(g[2]); dx.append(g[3])
            starts.append(starts[-1] + b.n)
            if b.ni:
                ixp.append(panel_geometry(b.ixyz, b.incn, b.incon, b.xr)[0])
        self.body_start = np.array(starts, dtype=np.int32)
        self.xyz = np.concatenate(xyz)
        self.ncn = np.concatenate(ncn).astype(np.int32)
        self.ncon = np.concatenate(ncon).astype(np.int32)
        self.xyz_p = np.concatenate(xp)
        self.ds = np.concatenate(ds)
        self.pnsz = np.concatenate(pn)
        self.dxyz_p = np.concatenate(dx)
        self.ixyz_p = np.concatenate(ixp) if ixp else np.zeros((0, 3))
        self.n = len(self.ncn)
        self.ni = len(self.ixyz_p)
        # body frames: the same inputs from each body's local nodes (the panel
        # geometry the device uses when frames are on) and the placements
        self.frames = any(b.local is not None for b in self.bodies)
        if self.frames:
            lx, lxp, lds, lpn, ldx, lix, org, wps = [], [], [], [], [], [], [], [0]
            for b in self.bodies:
                o = np.array([b.origin[0], b.origin[1], 0.0]) if b.local is not None else np.zeros(3)
                loc = b.local if b.local is not None else b.xyz
                if b.local is not None:
                    scale = max(1.0, float(np.max(np.abs(b.xyz))))
                    if b.local.shape != b.xyz.shape or np.max(np.abs(b.local + o - b.xyz)) > 1e-9 * scale:
                        raise ValueError("a body's local nodes placed at its origin are not its global nodes "
                                         "(set local=None, or move local with xyz)")
                g = panel_geometry(loc, b.ncn, b.ncon, np.asarray(b.xr, dtype=np.float64) - o)
                lx.append(loc)
                lxp.append(g[0]); lds.append(g[1]); lpn.append(g[2]); ldx.append(g[3])
                if b.ni:
                    iloc = b.ilocal if b.local is not None and b.ilocal is not None else b.ixyz - o
                    lix.append(panel_geometry(iloc, b.incn, b.incon)[0])
                org.append(o[:2])
                wps.append(wps[-1] + b.ni)
            self.local = dict(xyz=np.concatenate(lx), xyz_p=np.concatenate(lxp), ds=np.concatenate(lds),
                              pnsz=np.concatenate(lpn), dxyz_p=np.concatenate(ldx),
                              ixyz_p=np.concatenate(lix) if lix else np.zeros((0, 3)),
                              origin=np.array(org, dtype=np.float64), wp_start=np.array(wps, dtype=np.int32))

    @property
    def m(self):
        return 6 * self.nbody

    @property
    def nbeta(self):
        return len(self.beta)


def reference_parity_constants(p):
    """The reference's hard-wired constants (WavDynMods.f90:146, WavDynSubs.f90:58)."""
    p.rho, p.g, p.g_disp = 1000.0, 9.80665, 9.807
    return p


# ---------------------------------------------------------------------------
# the named workloads
# ---------------------------------------------------------------------------
def c1_single_cylinder(n_freq=16):
    """One truncated cylinder r=5, T=5, infinite depth; 64x12 side + 64x4 bottom
    = 1024 hull panels, 64x4 = 256 waterplane panels; 16 omegas in 0.1-2.0, beta=0."""
    b = cylinder(5.0, 5.0, 64, 12, 4, waterplane_nr=4)
    return Problem([b], h=-1.0, omegas=np.linspace(0.1, 2.0, n_freq), beta=np.array([0.0]),
                   name="C1 cylinder r5 T5 deep, 1024+256 panels")


def c2_three_cylinders(n_freq=40):
    """Three cylinders r=5, T=5 on an equilateral triangle (20 m side), h=50;
    1536 panels/body (64x16 side + 64x8 bottom) = 4608; 18 modes; 40 omegas in
    0.1-2.5; 8 headings 0:45:315."""
    s = 20.0
    cs = [(0.0, 0.0), (s, 0.0), (0.5 * s, s * np.sqrt(3) / 2)]
    bodies = [cylinder(5.0, 5.0, 64, 16, 8, center=c) for c in cs]
    return Problem(bodies, h=50.0, omegas=np.linspace(0.1, 2.5, n_freq),
                   beta=np.deg2rad(np.arange(0, 360, 45.0)),
                   name="C2 3 cylinders r5 T5 h50, 4608 panels")


def c4_eight_cylinders(n_freq=48):
    """4x2 grid of cylinders r=4, T=6, 16 m spacing, h=40; 2048 panels/body
    (64x24 side + 64x8 bottom) = 16384; 48 omegas; 4 headings."""
    cs = [(16.0 * i, 16.0 * j) for j in range(2) for i in range(4)]
    bodies = [cylinder(4.0, 6.0, 64, 24, 8, center=c) for c in cs]
    return Problem(bodies, h=40.0, omegas=np.linspace(0.1, 2.5, n_freq),
                   beta=np.deg2rad(np.arange(0, 360, 90.0)),
                   name="C4 8 cylinders r4 T6 h40, 16384 panels")


def c5_sixteen_cylinders(n_freq=64):
    """4x4 grid of cylinders r=3, T=4, 12 m spacing, h=60; 2048 panels/body =
    32768; 96 modes; 64 omegas in 0.2-2.5; 2 headings."""
    cs = [(12.0 * i, 12.0 * j) for j in range(4) for i in range(4)]
    bodies = [cylinder(3.0, 4.0, 64, 24, 8, center=c) for c in cs]
    return Problem(bodies, h=60.0, omegas=np.linspace(0.2, 2.5, n_freq),
                   beta=np.deg2rad(np.array([0.0, 90.0])),
                   name="C5 16 cylinders r3 T4 h60, 32768 panels")


def c3_random_problems(seed=20261203, count=64, n_freq=32):
    """64 independent random boxes/cylinders (dims U(2,20), draft U(1,10), panel
    count log-uniform in 512-6144, depth infinite or U(30,200))."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        is_box = rng.random() < 0.5
        d1, d2 = rng.uniform(2, 20, 2)
        draft = rng.uniform(1, 10)
        npan = int(round(np.exp(rng.uniform(np.log(512), np.log(6144)))))
        deep = rng.random() < 0.5
        h = -1.0 if deep else rng.uniform(30, 200)
        h = max(h, 1.5 * draft) if h > 0 else h
        if is_box:
            # bottom nx*ny + sides 2*nz*(nx+ny) ~ npan with cells ~ square
            a = np.sqrt(npan / (d1 * d2 + 2 * draft * (d1 + d2)))
            nx, ny, nz = max(2, round(a * d1)), max(2, round(a * d2)), max(1, round(a * draft))
            b = box(d1, d2, draft, nx, ny, nz)
        else:
            r = 0.5 * d1
            nt = max(8, int(round(np.sqrt(npan * 2 * np.pi * r / (draft + r)) / 4) * 4))
            nz = max(1, round(npan * draft / (draft + r) / nt))
            nr = max(1, round((npan - nt * nz) / nt))
            b = cylinder(r, draft, nt, nz, nr)
        out.append(Problem([b], h=h, omegas=np.linspace(0.1, 3.0, n_freq),
                           beta=np.deg2rad(np.arange(0, 360, 90.0)),
                           name=f"C3[{k}] {'box' if is_box else 'cyl'} N={b.n}"))
    return out


def reference_config1_from(inp):
    """test/data/config1 of the reference (input directory `inp`): 3 bodies x 234
    panels at y = 0, 54, 108 m, h = 70 m, omega = 0.05, 0.06, 0.07 rad/s,
    headings 0, 10, 20 deg, with the reference's constants."""
    import os
    ys = (0.0, 54.0, 108.0)
    bodies = [read_pnl(os.path.join(inp, f"HullMesh_{k + 1}.pnl"), lcs=(0.0, y, 0.0, 0.0),
                       xr=(0.0, y, 0.0)) for k, y in enumerate(ys)]
    p = Problem(bodies, h=70.0, omegas=np.array([0.05, 0.06, 0.07]),
                beta=np.deg2rad(np.array([0.0, 10.0, 20.0])), name="reference config1")
    return reference_parity_constants(p)


def reference_config1(root):
    import os
    return reference_config1_from(os.path.join(root, "test", "data", "config1", "input"))
