"""Seeded synthetic scenes, receiver sets and trace settings for the benchmark configurations.

Scenes are triangle soups in the layout a binary STL holds (9 vertex floats and 3 normal floats per
facet, plus a material id), so they enter the library exactly as a file read by
Preprocessor::StlToGeometry would.  Randomness comes from SplitMix64 with an explicit uniform
mapping, so every platform generates the same scene for a seed.
"""
import numpy as np

MASK64 = (1 << 64) - 1
SEED_BASE = 20261203


class SplitMix64:
    def __init__(self, seed):
        self.s = seed & MASK64

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & MASK64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, a=0.0, b=1.0):
        return a + (b - a) * ((self.next() >> 11) * (1.0 / (1 << 53)))

    def randint(self, lo, hi):
        """Integer in [lo, hi]."""
        return lo + int(self.uniform() * (hi - lo + 1)) if hi > lo else lo


CONCRETE = [0.0, -1.0, -1.0, 5.31, 0.0, 0.0326]   # [frequency, transLoss, reflLoss, er, er_im, sigma]


class Soup:
    def __init__(self):
        self.v, self.n, self.m = [], [], []

    def tri(self, a, b, c, n, mat=0):
        self.v.append(list(a) + list(b) + list(c))
        self.n.append(list(n))
        self.m.append(mat)

    def quad(self, a, b, c, d, n, mat=0):
        """Quad a-b-c-d as triangles (a, b, c) and (a, c, d)."""
        self.tri(a, b, c, n, mat)
        self.tri(a, c, d, n, mat)

    def box(self, x0, y0, z0, x1, y1, z1, inward=False, mat=0):
        s = -1.0 if inward else 1.0
        self.quad((x0, y0, z0), (x0, y1, z0), (x1, y1, z0), (x1, y0, z0), (0, 0, -s), mat)   # bottom
        self.quad((x0, y0, z1), (x1, y0, z1), (x1, y1, z1), (x0, y1, z1), (0, 0, s), mat)    # top
        self.quad((x0, y0, z0), (x1, y0, z0), (x1, y0, z1), (x0, y0, z1), (0, -s, 0), mat)   # y = y0
        self.quad((x0, y1, z0), (x0, y1, z1), (x1, y1, z1), (x1, y1, z0), (0, s, 0), mat)    # y = y1
        self.quad((x0, y0, z0), (x0, y0, z1), (x0, y1, z1), (x0, y1, z0), (-s, 0, 0), mat)   # x = x0
        self.quad((x1, y0, z0), (x1, y1, z0), (x1, y1, z1), (x1, y0, z1), (s, 0, 0), mat)    # x = x1

    def grid(self, o, u, v, nu, nv, n, mat=0):
        """Rectangle o + [0,1] u + [0,1] v split into nu x nv cells of two triangles each; the
        triangles wind (o, o + u, o + u + v) so that n should be u x v normalised."""
        o, u, v = np.asarray(o, float), np.asarray(u, float), np.asarray(v, float)
        for i in range(nu):
            for j in range(nv):
                a = o + u * (i / nu) + v * (j / nv)
                b = o + u * ((i + 1) / nu) + v * (j / nv)
                c = o + u * ((i + 1) / nu) + v * ((j + 1) / nv)
                d = o + u * (i / nu) + v * ((j + 1) / nv)
                self.quad(a, b, c, d, n, mat)

    def tessellated_box(self, x0, y0, z0, x1, y1, z1, cell, mat=0, bottom=False, top=True):
        """Outward box whose faces are split into cells of about `cell` metres."""
        nx, ny, nz = (max(1, int(round((x1 - x0) / cell))), max(1, int(round((y1 - y0) / cell))),
                      max(1, int(round((z1 - z0) / cell))))
        dx, dy, dz = (x1 - x0, 0, 0), (0, y1 - y0, 0), (0, 0, z1 - z0)
        if bottom:
            self.grid((x0, y0, z0), dy, dx, ny, nx, (0, 0, -1), mat)
        if top:
            self.grid((x0, y0, z1), dx, dy, nx, ny, (0, 0, 1), mat)
        self.grid((x0, y0, z0), dx, dz, nx, nz, (0, -1, 0), mat)     # y = y0
        self.grid((x0, y1, z0), dz, dx, nz, nx, (0, 1, 0), mat)      # y = y1
        self.grid((x0, y0, z0), dz, dy, nz, ny, (-1, 0, 0), mat)     # x = x0
        self.grid((x1, y0, z0), dy, dz, ny, nz, (1, 0, 0), mat)      # x = x1

    def arrays(self):
        return (np.asarray(self.v, np.float32).reshape(-1, 9), np.asarray(self.n, np.float32).reshape(-1, 3),
                np.asarray(self.m, np.int32))


def config1():
    """Closed box room 10x8x3 m (12 triangles, inward normals), concrete, 2.4 GHz, half-wave
    dipole Tx at (2, 2, 1.5), icosahedron level 5, 4 reflections, 10x10 receivers at z = 1.2."""
    s = Soup()
    s.box(0.0, 0.0, 0.0, 10.0, 8.0, 3.0, inward=True)
    rx = np.array([[0.5 + i, 0.4 + 0.8 * j, 1.2] for i in range(10) for j in range(10)])
    mats = np.array([[2.4e9] + CONCRETE[1:]])
    trace = dict(tx=(2.0, 2.0, 1.5), tess=5, maxR=4, maxT=0, frequency=2.4e9, pt=1.0, antennaType=2)
    return dict(name="box_room_12tri_L5_R4", soup=s.arrays(), rx=rx, materials=mats, trace=trace)


def config3(tess=10, floor_cell=0.25, wall_cell=0.5, n_walls=40, n_furniture=300, rx_grid=(200, 100),
            maxR=8, maxT=8, maxTotal=8):
    """Dense indoor triangle soup (seed 20261203 + 3): floor, ceiling and perimeter of a 50x50x4 m
    room, n_walls seeded interior walls (0.15 m thick, 3-15 m long, full height) and n_furniture
    boxes (0.4-2 m), all faces tessellated (about 200k triangles at the defaults); 16 materials with
    er ~ U[2, 7] and sigma ~ logU[1e-3, 1e-1] S/m at 2.4 GHz, one per object; Tx at 2 m in free
    space; icosahedron level 10; reflections and transmissions up to 8 each, at most 8 in total;
    200x100 receivers at z = 1 (pitch 0.25 x 0.5 m)."""
    rng = SplitMix64(SEED_BASE + 3)
    L, H = 50.0, 4.0
    mats = np.array([[2.4e9, -1.0, -1.0, rng.uniform(2.0, 7.0), 0.0, float(np.exp(rng.uniform(np.log(1e-3), np.log(1e-1))))]
                     for _ in range(16)])
    s = Soup()
    nf = int(round(L / floor_cell))
    s.grid((0, 0, 0), (L, 0, 0), (0, L, 0), nf, nf, (0, 0, 1), rng.randint(0, 15))          # floor, facing up
    s.grid((0, 0, H), (0, L, 0), (L, 0, 0), nf, nf, (0, 0, -1), rng.randint(0, 15))         # ceiling, facing down
    nw, nh = int(round(L / wall_cell)), int(round(H / wall_cell))
    pm = rng.randint(0, 15)
    s.grid((0, 0, 0), (0, 0, H), (L, 0, 0), nh, nw, (0, 1, 0), pm)                           # y = 0, inward
    s.grid((0, L, 0), (L, 0, 0), (0, 0, H), nw, nh, (0, -1, 0), pm)                          # y = L
    s.grid((0, 0, 0), (0, L, 0), (0, 0, H), nw, nh, (1, 0, 0), pm)                           # x = 0
    s.grid((L, 0, 0), (0, 0, H), (0, L, 0), nh, nw, (-1, 0, 0), pm)                          # x = L
    solids = []
    for _ in range(n_walls):
        length, t = rng.uniform(3.0, 15.0), 0.15
        if rng.uniform() < 0.5:
            x0, y0 = rng.uniform(1.0, L - 1.0 - length), rng.uniform(1.0, L - 1.0)
            b = (x0, y0, x0 + length, y0 + t)
        else:
            x0, y0 = rng.uniform(1.0, L - 1.0), rng.uniform(1.0, L - 1.0 - length)
            b = (x0, y0, x0 + t, y0 + length)
        solids.append(b)
        s.tessellated_box(b[0], b[1], 0.0, b[2], b[3], H, wall_cell, rng.randint(0, 15), top=False)
    placed = 0
    while placed < n_furniture:
        w, d, h = rng.uniform(0.4, 2.0), rng.uniform(0.4, 2.0), rng.uniform(0.4, 2.0)
        x0, y0 = rng.uniform(0.5, L - 0.5 - w), rng.uniform(0.5, L - 0.5 - d)
        b = (x0, y0, x0 + w, y0 + d)
        if any(b[0] < o[2] + 0.1 and o[0] < b[2] + 0.1 and b[1] < o[3] + 0.1 and o[1] < b[3] + 0.1 for o in solids):
            continue
        solids.append(b)
        s.tessellated_box(b[0], b[1], 0.0, b[2], b[3], h, wall_cell, rng.randint(0, 15))
        placed += 1
    while True:
        tx = (rng.uniform(2.0, L - 2.0), rng.uniform(2.0, L - 2.0), 2.0)
        if all(not (o[0] - 0.3 < tx[0] < o[2] + 0.3 and o[1] - 0.3 < tx[1] < o[3] + 0.3) for o in solids):
            break
    nx, ny = rx_grid
    px, py = L / nx, L / ny
    rx = np.array([[px / 2 + px * i, py / 2 + py * j, 1.0] for i in range(nx) for j in range(ny)])
    trace = dict(tx=tx, tess=tess, maxR=maxR, maxT=maxT, maxTotal=maxTotal, frequency=2.4e9, pt=1.0, antennaType=2,
                 receptionModel=1)
    v, n, m = s.arrays()
    return dict(name=f"indoor_{len(m)}tri_L{tess}_R{maxR}T{maxT}max{maxTotal}", soup=(v, n, m), rx=rx, materials=mats,
                trace=trace)


def urban_buildings(seed, count=40, half=250.0, fmin=10.0, fmax=40.0, hmin=10.0, hmax=60.0, gap=2.0):
    rng = SplitMix64(seed)
    boxes = []
    tries = 0
    while len(boxes) < count and tries < 100000:
        tries += 1
        w, d = rng.uniform(fmin, fmax), rng.uniform(fmin, fmax)
        h = rng.uniform(hmin, hmax)
        cx, cy = rng.uniform(-half + w / 2 + 5, half - w / 2 - 5), rng.uniform(-half + d / 2 + 5, half - d / 2 - 5)
        b = (cx - w / 2, cy - d / 2, cx + w / 2, cy + d / 2, h)
        if any(b[0] < o[2] + gap and o[0] < b[2] + gap and b[1] < o[3] + gap and o[1] < b[3] + gap for o in boxes):
            continue
        boxes.append(b)
    return boxes, rng


def urban_scene():
    """The urban-canyon scene of config2 (ground plus 40 seeded boxes) and the generator state after
    the buildings were drawn."""
    boxes, rng = urban_buildings(SEED_BASE + 2)
    s = Soup()
    h = 250.0
    s.quad((-h, -h, 0.0), (h, -h, 0.0), (h, h, 0.0), (-h, h, 0.0), (0, 0, 1))
    for (x0, y0, x1, y1, bh) in boxes:
        s.box(x0, y0, 0.0, x1, y1, bh)
    return s, boxes, rng


def street_point(rng, boxes, z, margin=3.0, half=200.0):
    """First seeded point of [-half, half]^2 at least `margin` from every building footprint."""
    while True:
        p = (rng.uniform(-half, half), rng.uniform(-half, half), z)
        if all(not (x0 - margin < p[0] < x1 + margin and y0 - margin < p[1] < y1 + margin)
               for (x0, y0, x1, y1, _) in boxes):
            return p


def config2(tess=9, rx_n=100):
    """Synthetic urban canyon at 28 GHz: 500x500 m ground (2 triangles) plus 40 seeded box buildings
    (footprints U[10,40] m, heights U[10,60] m; 482 triangles), concrete; 1 Tx at 25 m in a street
    cell; icosahedron level 9; 6 reflections plus one transmission; 100x100 receivers at z = 1.5."""
    s, boxes, rng = urban_scene()
    tx = street_point(rng, boxes, 25.0)   # Tx in a street cell
    pitch = 500.0 / rx_n
    g = -250.0 + pitch / 2 + pitch * np.arange(rx_n)
    rx = np.array([[x, y, 1.5] for x in g for y in g])
    mats = np.array([[28e9] + CONCRETE[1:]])
    trace = dict(tx=tx, tess=tess, maxR=6, maxT=1, frequency=28e9, pt=1.0, antennaType=2)
    return dict(name=f"urban_canyon_482tri_L{tess}_R6T1", soup=s.arrays(), rx=rx, materials=mats, trace=trace)


def config4(n_tx=256, rx_n=100):
    """Batched multi-transmitter run on the urban-canyon scene: n_tx transmitters in street cells
    at heights U[3, 30] m, icosahedron level U{6..9} and U{4..6} reflections each (no
    transmission), 28 GHz, 100x100 shared receivers at z = 1.5 (seed 20261203 + 4); reception radius
    alpha * L / sqrt(3)."""
    s, boxes, _ = urban_scene()
    rng = SplitMix64(SEED_BASE + 4)
    txs = []
    for _ in range(n_tx):
        z = rng.uniform(3.0, 30.0)
        tx = street_point(rng, boxes, z)
        txs.append(dict(tx=tx, tess=rng.randint(6, 9), maxR=rng.randint(4, 6), maxT=0, frequency=28e9, pt=1.0,
                        antennaType=2, receptionModel=1))
    pitch = 500.0 / rx_n
    g = -250.0 + pitch / 2 + pitch * np.arange(rx_n)
    rx = np.array([[x, y, 1.5] for x in g for y in g])
    mats = np.array([[28e9] + CONCRETE[1:]])
    return dict(name=f"urban_canyon_482tri_{n_tx}tx_L6-9_R4-6", soup=s.arrays(), rx=rx, materials=mats, txs=txs)


def config5_scene(n_buildings=2000, lot=60.0, cell=5.0):
    """Seeded synthetic city for the serving stream (seed 20261203 + 5): a square grid of lots of
    `lot` metres holding n_buildings box buildings (footprints U[15, 45] m inside their lot, heights
    U[10, 80] m), facades and roofs tessellated into cells of about `cell` metres, on a 2-triangle
    ground; concrete at 28 GHz.  About 1M triangles at the defaults.  Returns (name, soup, boxes,
    half extent, materials)."""
    rng = SplitMix64(SEED_BASE + 5)
    side = int(np.ceil(np.sqrt(n_buildings)))
    half = side * lot / 2.0
    s = Soup()
    s.quad((-half, -half, 0.0), (half, -half, 0.0), (half, half, 0.0), (-half, half, 0.0), (0, 0, 1))
    boxes = []
    for k in range(n_buildings):
        gx, gy = k % side, k // side
        w, d, h = rng.uniform(15.0, 45.0), rng.uniform(15.0, 45.0), rng.uniform(10.0, 80.0)
        x0 = -half + gx * lot + rng.uniform(5.0, lot - 5.0 - w)
        y0 = -half + gy * lot + rng.uniform(5.0, lot - 5.0 - d)
        boxes.append((x0, y0, x0 + w, y0 + d, h))
        s.tessellated_box(x0, y0, 0.0, x0 + w, y0 + d, h, cell)
    v, n, m = s.arrays()
    mats = np.array([[28e9] + CONCRETE[1:]])
    return f"city_{len(m)}tri_{n_buildings}bldg", (v, n, m), boxes, half, mats


def coverage_queries(n, boxes, seed=SEED_BASE + 5, half=250.0, tess=(6, 10), max_rx=50000, maxR=4):
    """Coverage queries for the serving loop: each a transmitter in a street cell at U[3, 30] m, an
    icosahedron level drawn from `tess`, maxR reflections, and a receiver count drawn log-uniform
    from [1, max_rx] with receivers uniform over the area at z = 1.5."""
    rng = SplitMix64(seed)
    out = []
    for _ in range(n):
        z = rng.uniform(3.0, 30.0)
        tx = street_point(rng, boxes, z, half=0.8 * half)
        level = rng.randint(tess[0], tess[1])
        nrx = max(1, min(max_rx, int(np.exp(rng.uniform(0.0, np.log(max_rx))))))
        rx = np.array([[rng.uniform(-half, half), rng.uniform(-half, half), 1.5] for _ in range(nrx)])
        out.append((dict(tx=tx, tess=level, maxR=maxR, maxT=0, frequency=28e9, pt=1.0, antennaType=2,
                         receptionModel=1), rx))
    return out


def ground_fixture():
    """The reference's committed main() scenario (SBRSolver_rev2.cpp:18-144): ground.stl, 3000
    receivers on x = 0, z = 3, y = 0..2999, Tx (0, 0, 5), txCoordSys rows (0,1,0), (0,0,1), (1,0,0),
    icosahedron level 2, 1 reflection, 1.8 GHz, metal (er = 1, sigma = 1e6)."""
    rx = np.array([[0.0, float(i), 3.0] for i in range(3000)])
    mats = np.array([[1.8e9, 221.13, 0.05, 1.0, 0.0, 1e6]])
    trace = dict(tx=(0.0, 0.0, 5.0), tess=2, maxR=1, maxT=0, frequency=1.8e9, pt=1.0,
                 txCoordSys=[[0, 1, 0], [0, 0, 1], [1, 0, 0]], antennaType=1)
    return dict(name="ground_fixture", rx=rx, materials=mats, trace=trace)
