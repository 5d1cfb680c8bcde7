"""Scene descriptions for the path tracer: flat arrays handed to
``illume_scene_create`` (include/illume_hip.h).

A scene is triangles (vertex array + index triples) and full spheres, a
material index and an emitted radiance per primitive (emissive triangles are
one-sided area lights), and a material table. Builders here produce the
Cornell-box workloads: the plain Lambertian box (256x256, 16 spp, depth 5) and
its glossy small-light variant (Trowbridge-Reitz conductor walls, a light of 1%
of the ceiling area, 512x512, depth 8).
"""
from dataclasses import dataclass, field

import numpy as np

MATTE, METAL, GLASS = 0, 1, 2
MATERIAL_DTYPE = np.dtype([("type", "<i4"), ("kd", "<f4", 3), ("eta", "<f4", 3), ("k", "<f4", 3), ("alpha", "<f4"),
                           ("ior", "<f4"), ("pad", "<f4", 4)])
CAMERA_DTYPE = np.dtype([("pos", "<f4", 3), ("look", "<f4", 3), ("up", "<f4", 3), ("fov", "<f4"), ("width", "<u4"),
                         ("height", "<u4")])
assert MATERIAL_DTYPE.itemsize == 64 and CAMERA_DTYPE.itemsize == 48

# RGB conductor constants of copper (eta, k) at roughly 650/550/450 nm.
COPPER_ETA = (0.200438, 0.924033, 1.10221)
COPPER_K = (3.91295, 2.45285, 2.14219)


def matte(kd):
    m = np.zeros(1, MATERIAL_DTYPE)
    m["type"] = MATTE
    m["kd"] = kd
    return m


def metal(alpha, eta=COPPER_ETA, k=COPPER_K):
    m = np.zeros(1, MATERIAL_DTYPE)
    m["type"] = METAL
    m["eta"] = eta
    m["k"] = k
    m["alpha"] = alpha
    return m


def glass(ior=1.5):
    m = np.zeros(1, MATERIAL_DTYPE)
    m["type"] = GLASS
    m["ior"] = ior
    return m


def camera(pos, look, up, fov, width, height):
    c = np.zeros(1, CAMERA_DTYPE)
    c["pos"], c["look"], c["up"], c["fov"], c["width"], c["height"] = pos, look, up, fov, width, height
    return c


@dataclass
class SceneDesc:
    verts: np.ndarray = field(default_factory=lambda: np.zeros((0, 3), np.float32))
    idx: np.ndarray = field(default_factory=lambda: np.zeros((0, 3), np.uint32))
    spheres: np.ndarray = field(default_factory=lambda: np.zeros((0, 4), np.float32))
    tri_material: list = field(default_factory=list)
    tri_emission: list = field(default_factory=list)
    sphere_material: list = field(default_factory=list)
    materials: list = field(default_factory=list)
    sky: tuple = None  # ((zenith rgb), (horizon rgb)) of the analytic sky, or None

    def add_material(self, m):
        self.materials.append(m)
        return len(self.materials) - 1

    def add_triangle(self, a, b, c, mat, emission=(0, 0, 0), normal=None):
        """Triangle a, b, c; if `normal` is given the vertex order is chosen so
        that the geometric normal normalize((a-c) x (b-c)) points along it."""
        a, b, c = (np.asarray(x, np.float32) for x in (a, b, c))
        if normal is not None and np.dot(np.cross(a - c, b - c), normal) < 0:
            b, c = c, b
        base = len(self.verts)
        self.verts = np.concatenate([self.verts, np.stack([a, b, c])]).astype(np.float32)
        self.idx = np.concatenate([self.idx, np.array([[base, base + 1, base + 2]], np.uint32)])
        if isinstance(self.tri_material, np.ndarray):  # per-triangle arrays of a generated mesh
            self.tri_material = np.append(self.tri_material, np.uint32(mat))
            self.tri_emission = np.concatenate([self.tri_emission, np.float32(emission).reshape(1, 3)])
        else:
            self.tri_material.append(mat)
            self.tri_emission.append(emission)

    def add_quad(self, p0, p1, p2, p3, mat, emission=(0, 0, 0), normal=None):
        """Planar quad p0 p1 p2 p3 (in order around the boundary) as two triangles."""
        self.add_triangle(p0, p1, p2, mat, emission, normal)
        self.add_triangle(p0, p2, p3, mat, emission, normal)

    def add_sphere(self, center, radius, mat):
        self.spheres = np.concatenate([self.spheres, np.array([[*center, radius]], np.float32)])
        self.sphere_material.append(mat)

    def arrays(self):
        n_t, n_s = len(self.idx), len(self.spheres)
        prim_material = np.concatenate([np.asarray(self.tri_material, np.uint32).reshape(n_t),
                                        np.asarray(self.sphere_material, np.uint32).reshape(n_s)])
        prim_emission = np.zeros((n_t + n_s, 3), np.float32)
        if n_t:
            prim_emission[:n_t] = np.asarray(self.tri_emission, np.float32).reshape(n_t, 3)
        mats = np.concatenate(self.materials) if self.materials else np.zeros(0, MATERIAL_DTYPE)
        return {
            "verts": np.ascontiguousarray(self.verts, np.float32).reshape(-1, 3),
            "idx": np.ascontiguousarray(self.idx, np.uint32).reshape(-1, 3),
            "spheres": np.ascontiguousarray(self.spheres, np.float32).reshape(-1, 4),
            "prim_material": prim_material,
            "prim_emission": prim_emission,
            "materials": np.ascontiguousarray(mats, MATERIAL_DTYPE),
            "sky": None if self.sky is None else np.asarray(self.sky, np.float32).reshape(6),
        }


def cornell_box(seed=0, glossy=False, light_size=0.25, light_L=17.0, wall_alpha_range=(0.05, 0.5), light_up=False):
    """Box [-1,1]^3 open at +z: floor, ceiling, back grey 0.75, left (x=-1)
    red, right (x=1) green; two spheres r=0.3 (albedo 0.75) with centres
    uniform in [-0.7,0.7]^3; a light_size^2 one-sided quad light just below
    the ceiling facing -y. glossy=True makes the five walls Trowbridge-Reitz
    copper with alpha uniform in wall_alpha_range (one alpha per wall).
    light_up=True turns the light over (at y = 0.9, facing +y): the room is
    then lit only by the patch of ceiling above it (all light indirect)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    s = SceneDesc()
    grey = s.add_material(matte((0.75, 0.75, 0.75)))
    red = s.add_material(matte((0.63, 0.06, 0.05)))
    green = s.add_material(matte((0.14, 0.45, 0.09)))
    centres = rng.uniform(-0.7, 0.7, (2, 3))
    if glossy:
        alphas = rng.uniform(*wall_alpha_range, 5)
        walls = [s.add_material(metal(float(a))) for a in alphas]
    else:
        walls = [grey, grey, grey, red, green]
    floor_m, ceil_m, back_m, left_m, right_m = walls
    s.add_quad((-1, -1, -1), (1, -1, -1), (1, -1, 1), (-1, -1, 1), floor_m, normal=(0, 1, 0))
    s.add_quad((-1, 1, -1), (1, 1, -1), (1, 1, 1), (-1, 1, 1), ceil_m, normal=(0, -1, 0))
    s.add_quad((-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1), back_m, normal=(0, 0, 1))
    s.add_quad((-1, -1, -1), (-1, 1, -1), (-1, 1, 1), (-1, -1, 1), left_m, normal=(1, 0, 0))
    s.add_quad((1, -1, -1), (1, 1, -1), (1, 1, 1), (1, -1, 1), right_m, normal=(-1, 0, 0))
    h = light_size / 2
    y = 0.9 if light_up else 1 - 1e-4
    light_m = s.add_material(matte((0.0, 0.0, 0.0)))
    s.add_quad((-h, y, -h), (h, y, -h), (h, y, h), (-h, y, h), light_m, emission=(light_L,) * 3,
               normal=(0, 1, 0) if light_up else (0, -1, 0))
    for c in centres:
        s.add_sphere(c, 0.3, grey)
    return s


def cornell_camera(width=256, height=256):
    return camera((0, 0, 3.75), (0, 0, 0), (0, 1, 0), 40.0, width, height)


def furnace_box(albedo=0.5, emission=1.0):
    """A closed box whose every wall is matte(albedo) AND emits `emission`
    (white furnace): without Russian roulette the radiance along any camera
    ray is emission * sum_{k=0..max_depth} albedo^k."""
    s = SceneDesc()
    m = s.add_material(matte((albedo,) * 3))
    e = (emission,) * 3
    s.add_quad((-1, -1, -1), (1, -1, -1), (1, -1, 1), (-1, -1, 1), m, e, normal=(0, 1, 0))
    s.add_quad((-1, 1, -1), (1, 1, -1), (1, 1, 1), (-1, 1, 1), m, e, normal=(0, -1, 0))
    s.add_quad((-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1), m, e, normal=(0, 0, 1))
    s.add_quad((-1, -1, 1), (1, -1, 1), (1, 1, 1), (-1, 1, 1), m, e, normal=(0, 0, -1))
    s.add_quad((-1, -1, -1), (-1, 1, -1), (-1, 1, 1), (-1, -1, 1), m, e, normal=(1, 0, 0))
    s.add_quad((1, -1, -1), (1, 1, -1), (1, 1, 1), (1, -1, 1), m, e, normal=(-1, 0, 0))
    return s


def rel_mse(img, ref, eps=0.01):
    """Mean over pixels and channels of (I - R)^2 / (R^2 + eps)."""
    img = np.asarray(img, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.mean((img - ref) ** 2 / (ref ** 2 + eps)))


# ---------------------------------------------------------------- batched procedural scenes
_ICO_CACHE = {}


def icosphere(level):
    """Unit icosphere: (verts float64 [V,3], faces int [F,3]) with 20*4^level faces."""
    if level in _ICO_CACHE:
        return _ICO_CACHE[level]
    t = (1 + 5 ** 0.5) / 2
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    verts = np.array(v, np.float64)
    verts /= np.linalg.norm(verts, axis=1, keepdims=True)
    faces = np.array(f, np.int64)
    for _ in range(level):
        edges = np.concatenate([faces[:, [0, 1]], faces[:, [1, 2]], faces[:, [2, 0]]])
        edges.sort(axis=1)
        uniq, inv = np.unique(edges, axis=0, return_inverse=True)
        mid = verts[uniq[:, 0]] + verts[uniq[:, 1]]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = len(verts)
        verts = np.concatenate([verts, mid])
        nf = len(faces)
        a, b, c = faces[:, 0], faces[:, 1], faces[:, 2]
        ab, bc, ca = base + inv[:nf], base + inv[nf:2 * nf], base + inv[2 * nf:]
        faces = np.concatenate([np.stack([a, ab, ca], 1), np.stack([b, bc, ab], 1), np.stack([c, ca, bc], 1),
                                np.stack([ab, bc, ca], 1)])
    _ICO_CACHE[level] = (verts, faces)
    return verts, faces


def procedural_scene(seed):
    """One scene of the batched workload: triangle count log-uniform in
    [1e3, 2e6] from randomly displaced icospheres on a ground quad, Lambertian
    or dielectric (IOR 1.5) objects, 1-4 quad lights, a square image of side
    log-uniform in [64, 1024], spp log-uniform in [8, 64].
    Returns (SceneDesc, camera, spp)."""
    rng = np.random.Generator(np.random.PCG64(1000 + seed))
    target = int(np.exp(rng.uniform(np.log(1e3), np.log(2e6))))
    s = SceneDesc()
    ground = s.add_material(matte((0.6, 0.6, 0.6)))
    s.add_quad((-4, 0, -4), (4, 0, -4), (4, 0, 4), (-4, 0, 4), ground, normal=(0, 1, 0))
    verts, idx = [s.verts], [s.idx]
    tri_mat = [np.asarray(s.tri_material, np.uint32)]
    n = 2
    base_v = len(s.verts)
    while n < target:
        remaining = target - n
        level = int(np.clip(np.floor(np.log(max(remaining, 20) / 20) / np.log(4)), 0, 6))
        level = min(level, int(rng.integers(3, 7)))
        v, f = icosphere(level)
        r = np.exp(rng.uniform(np.log(0.15), np.log(0.6)))
        c = np.array([rng.uniform(-3, 3), r * rng.uniform(0.8, 1.5), rng.uniform(-3, 3)])
        amp = rng.uniform(0.0, 0.15)
        disp = 1 + amp * np.sin(3 * v[:, :1] + rng.uniform(0, 6)) * np.cos(4 * v[:, 1:2] + rng.uniform(0, 6))
        vv = (c + r * v * disp).astype(np.float32)
        if rng.random() < 0.3:
            m = s.add_material(glass(1.5))
        else:
            m = s.add_material(matte(tuple(rng.uniform(0.1, 0.8, 3))))
        verts.append(vv)
        idx.append((f + base_v).astype(np.uint32))
        tri_mat.append(np.full(len(f), m, np.uint32))
        base_v += len(vv)
        n += len(f)
    s.verts = np.concatenate(verts).astype(np.float32)
    s.idx = np.concatenate(idx).astype(np.uint32)
    s.tri_material = np.concatenate(tri_mat)
    s.tri_emission = np.zeros((len(s.idx), 3), np.float32)
    n_lights = int(rng.integers(1, 5))
    lm = s.add_material(matte((0, 0, 0)))
    for _ in range(n_lights):
        h = rng.uniform(0.2, 0.6)
        cx, cz, y = rng.uniform(-2.5, 2.5), rng.uniform(-2.5, 2.5), rng.uniform(3, 5)
        L = rng.uniform(5, 20) / (h * h)
        s.add_quad((cx - h, y, cz - h), (cx + h, y, cz - h), (cx + h, y, cz + h), (cx - h, y, cz + h), lm,
                   emission=(L, L, L), normal=(0, -1, 0))
    side = int(np.exp(rng.uniform(np.log(64), np.log(1024))))
    spp = int(np.exp(rng.uniform(np.log(8), np.log(64))))
    ang = rng.uniform(0, 2 * np.pi)
    cam = camera((7 * np.cos(ang), rng.uniform(2, 4), 7 * np.sin(ang)), (0, 0.5, 0), (0, 1, 0), 45.0, side, side)
    return s, cam, spp


# ---------------------------------------------------------------- terrain + spheres + sky (combined workload)
def fbm_heightfield(n=1024, octaves=6, lacunarity=2.0, gain=0.5, base_cells=4, seed=0):
    """fBm of value noise on an n x n grid over [0,1]^2: octave k has
    base_cells * lacunarity^k lattice cells per side, amplitude gain^k, and
    smoothstep-interpolated uniform lattice values in [-1, 1]. Returns float64
    [n, n] normalised to [0, 1]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    x = np.linspace(0.0, 1.0, n)
    h = np.zeros((n, n))
    amp, cells = 1.0, float(base_cells)
    for _ in range(octaves):
        m = int(np.ceil(cells)) + 1
        lat = rng.uniform(-1.0, 1.0, (m + 1, m + 1))
        g = x * cells
        i = np.minimum(np.floor(g).astype(np.int64), m - 1)
        f = g - i
        s = f * f * (3 - 2 * f)
        i0, i1 = i[:, None], i[None, :]
        sy, sx = s[:, None], s[None, :]
        v00, v01 = lat[i0, i1], lat[i0, i1 + 1]
        v10, v11 = lat[i0 + 1, i1], lat[i0 + 1, i1 + 1]
        h += amp * ((v00 * (1 - sx) + v01 * sx) * (1 - sy) + (v10 * (1 - sx) + v11 * sx) * sy)
        amp *= gain
        cells *= lacunarity
    h -= h.min()
    return h / max(h.max(), 1e-12)


def terrain_scene(seed=0, n=1024, n_spheres=10000, height=0.35, sky=True):
    """Heightfield terrain over [-1,1]^2 (xz), 2 (n-1)^2 triangles, fBm of 6
    octaves (lacunarity 2, gain 0.5), with n_spheres spheres resting on it
    (radius lognormal, median 0.02, sigma_log 0.5); Lambertian or
    Trowbridge-Reitz conductor materials; lit by the analytic sky.
    Returns SceneDesc (with .sky set)."""
    rng = np.random.Generator(np.random.PCG64(5000 + seed))
    hf = fbm_heightfield(n, seed=seed) * height
    xs = np.linspace(-1.0, 1.0, n)
    X, Z = np.meshgrid(xs, xs, indexing="xy")
    verts = np.stack([X, hf, Z], -1).reshape(-1, 3).astype(np.float32)
    r, c = np.meshgrid(np.arange(n - 1), np.arange(n - 1), indexing="ij")
    a = (r * n + c).reshape(-1)
    b, d, e = a + 1, a + n, a + n + 1
    # counter-clockwise seen from +y
    idx = np.concatenate([np.stack([a, d, b], 1), np.stack([b, d, e], 1)]).astype(np.uint32)
    s = SceneDesc()
    mats = [s.add_material(matte((0.45, 0.42, 0.35))), s.add_material(matte((0.3, 0.45, 0.25))),
            s.add_material(metal(0.3)), s.add_material(metal(0.1))]
    s.verts, s.idx = verts, idx
    # terrain: Lambertian, two albedos by height band
    band = (hf.reshape(-1)[a] > 0.5 * height)
    tri_m = np.concatenate([np.where(band, mats[0], mats[1]), np.where(band, mats[0], mats[1])])
    s.tri_material = [int(m) for m in tri_m]
    s.tri_emission = [(0, 0, 0)] * len(idx)
    rad = np.exp(np.log(0.02) + 0.5 * rng.standard_normal(n_spheres))
    px = rng.uniform(-0.95, 0.95, n_spheres)
    pz = rng.uniform(-0.95, 0.95, n_spheres)
    gi = np.clip(((pz + 1) / 2 * (n - 1)).round().astype(np.int64), 0, n - 1)
    gj = np.clip(((px + 1) / 2 * (n - 1)).round().astype(np.int64), 0, n - 1)
    py = hf[gi, gj] + rad
    kind = rng.random(n_spheres)
    albedo = rng.uniform(0.1, 0.85, (n_spheres, 3))
    sph_mats = []
    for k in range(n_spheres):
        if kind[k] < 0.5:
            sph_mats.append(s.add_material(matte(tuple(albedo[k]))))
        else:
            sph_mats.append(mats[2] if kind[k] < 0.75 else mats[3])
    s.spheres = np.stack([px, py, pz, rad], 1).astype(np.float32)
    s.sphere_material = sph_mats
    s.sky = SKY_DEFAULT if sky else None
    return s


# zenith RGB, horizon RGB (radiance); the sky emits from the upper hemisphere (+y)
SKY_DEFAULT = ((0.35, 0.55, 1.0), (1.1, 1.05, 0.95))


def terrain_views(seed=0, n_views=16):
    """n_views cameras around and above the terrain; pixel counts
    log-uniform in [256^2, 1920*1080], aspect log-uniform in [1, 16/9]."""
    rng = np.random.Generator(np.random.PCG64(7000 + seed))
    cams = []
    for _ in range(n_views):
        npx = np.exp(rng.uniform(np.log(256 * 256), np.log(1920 * 1080)))
        asp = np.exp(rng.uniform(0.0, np.log(16 / 9)))
        w = int(round(np.sqrt(npx * asp)))
        h = int(round(w / asp))
        ang = rng.uniform(0, 2 * np.pi)
        dist = rng.uniform(1.6, 2.6)
        pos = (dist * np.cos(ang), rng.uniform(0.6, 1.4), dist * np.sin(ang))
        look = (rng.uniform(-0.3, 0.3), 0.1, rng.uniform(-0.3, 0.3))
        cams.append(camera(pos, look, (0, 1, 0), 45.0, w, h))
    return cams
