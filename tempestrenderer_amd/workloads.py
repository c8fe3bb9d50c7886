"""Synthetic workloads of the benchmark configurations, as scene descriptions.

A `SceneDesc` is plain data (numpy arrays, material dictionaries, lights,
camera, render settings).  `populate(target)` replays it through the
reference scene API method names, so the same description feeds the GPU
tracer (`RayTracerScene`) and any other object with those methods.

Workloads:
  box(...)         5 axis-aligned quads (10 triangles) forming an open box,
                   2 GGX-conductor spheres r=0.3, a 0.25x0.25 emissive quad
                   L=15 in the same mesh, Lambertian walls with albedo
                   U[0.2, 0.8]; 256x256, 16 camera samples, depth 4.
  heightfield(...) an n x n vertex grid of fBm value noise (6 octaves,
                   amplitude 0.1; 2*(n-1)^2 triangles) + 64 spheres with GGX,
                   Beckmann and SGGX-surface materials; constant environment
                   light 1.0; 1920x1080, 64 camera samples, depth 8.
  volume(...)      a unit-cube grid volume, n^3 binary16 density from 3-D fBm
                   value noise scaled to a mean extinction of 4 per unit, a
                   per-sample fibre direction from the gradient of a second
                   noise field, SGGX micro-flakes with standard deviation
                   (0.1, 1, 1) in the fibre frame, albedo 0.9; constant
                   environment light 1.0; 1024x1024, 128 camera samples,
                   depth 32 (optionally a floor and a directional light).
  anisotropic(...) the six anisotropic GGX / Beckmann models on tangent-frame
                   quads, spheres and a quad without tangents; 48x48.
  pseudo_volume(...) SGGX pseudo-volume surfaces (mapped and constant); 32x32.
Scene randomness comes from numpy's PCG64 seeded with `seed`; the path
tracer's own random numbers come from Philox keyed by the render seed.
"""
import math

import numpy as np

# IlluminationModel values (ray-tracing-common.hh:68-105)
EMISSIVE = 0
GGX = 11
GGX_CONDUCTOR = 12
GGX_DIELECTRIC = 13
SGGX_MICRO_FLAKE = 22
SGGX_SURFACE = 24
BECKMANN = 28
BECKMANN_CONDUCTOR = 29

LIGHT_DIRECTIONAL = 0
LIGHT_SPHERE_AREA = 1
LIGHT_POINT = 3


def alpha_to_specular_power(alpha):
    """SpecularPowerToGeometryFactorSq(n) = 2/(n+2) = alpha^2 (microfacet-common.hh:109-112)."""
    return 2.0 / (alpha * alpha) - 2.0


def perspective_view_proj_inv(eye, target, up, fovy_deg, aspect, near=0.1, far=100.0):
    """Inverse of a GL-style projection*view, as 16 column-major floats."""
    eye, target, up = (np.asarray(v, dtype=np.float64) for v in (eye, target, up))
    f = target - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    view = np.eye(4)
    view[0, :3], view[1, :3], view[2, :3] = s, u, -f
    view[:3, 3] = -view[:3, :3] @ eye
    t = 1.0 / math.tan(math.radians(fovy_deg) / 2.0)
    proj = np.zeros((4, 4))
    proj[0, 0] = t / aspect
    proj[1, 1] = t
    proj[2, 2] = (far + near) / (near - far)
    proj[2, 3] = 2.0 * far * near / (near - far)
    proj[3, 2] = -1.0
    inv = np.linalg.inv(proj @ view)
    return np.ascontiguousarray(inv.T.reshape(16).astype(np.float32))  # column-major


IDENTITY = np.eye(4, dtype=np.float32).reshape(16)


class SceneDesc:
    def __init__(self, name, width, height, vpi):
        self.name = name
        self.width, self.height = int(width), int(height)
        self.vpi = np.asarray(vpi, dtype=np.float32)
        self.meshes = []       # (verts float32 [V, stride/4], tris int32 [T*3], [(material, first_index, count)], two_sided)
        self.spheres = []      # (center, radius, material)
        self.lights = []       # dict(kind=..., ...)
        self.sphere_lights = []
        self.volumes = []      # (dims, min, max, density f16 [nz, ny, nx], direction f16 [nz, ny, nx, 3], material)
        self.brick_volumes = []  # (VolumeRoot, material)
        self.shapes = []       # (scene-API method, keyword arguments, material or None)
        self.textures = []     # (texels, TR_TEX_* format, sampling): handles 1, 2, ... in this order
        self.cube_faces = None  # six uint32 RGBA8 faces
        self.settings = dict(samples_camera=1, samples_gi=1, samples_light=1, max_depth=1, rr=1.0,
                             background=(0.0, 0.0, 0.0), seed=1, estimator=0)

    @property
    def pixels(self):
        return self.width * self.height

    def populate(self, target, material_factory):
        """Replays the scene into `target` (RayTracerScene-like); materials are
        made by material_factory(model, fields) so each target gets its own."""
        cache = {}

        def mat(m):
            key = id(m)
            if key not in cache:
                fields = {k: v for k, v in m.items() if k != "model"}
                cache[key] = material_factory(m["model"], fields)
            return cache[key]

        st = self.settings
        for texels, fmt, sampling in self.textures:
            target.bindTexture(texels, fmt, sampling)
        if self.cube_faces is not None:
            target.setGlobalCubeMap(self.cube_faces)
        target.setSamplesCamera(st["samples_camera"])
        target.setSamplesGlobalIllumination(st["samples_gi"])
        target.setSamplesLocalAreaLight(st["samples_light"])
        target.setMaxRayDepth(st["max_depth"])
        target.setRussianRoulette(st["rr"])
        target.setBackgroundSpectrum(st["background"])
        target.setSeed(st["seed"])
        target.setEstimator(st.get("estimator", 0))
        from .scene import Submesh
        for verts, tris, parts, two_sided in self.meshes:
            stride = verts.shape[1] * 4
            subs = [Submesh(mat(m), count, 0, first, stride) for (m, first, count) in parts]
            target.addTriangleMesh(IDENTITY, subs, tris, verts, two_sided=two_sided)
        for center, radius, m in self.spheres:
            target.addSphere(center, radius, mat(m))
        for dims, lo, hi, dens, dirs, m in self.volumes:
            target.addGridVolume(dims, lo, hi, dens, dirs, mat(m))
        for root, m in self.brick_volumes:
            target.addHierarchicalVolume(root, mat(m))
        for method, kw, m in self.shapes:
            if m is None:
                getattr(target, method)(**kw)
            else:
                getattr(target, method)(material=mat(m), **kw)
        for L in self.sphere_lights:
            target.addSphereLightSource(target.make_light(**L))
        for L in self.lights:
            target.addLightSource(target.make_light(**L))


def _quad(center, u, v):
    c, u, v = (np.asarray(x, dtype=np.float32) for x in (center, u, v))
    n = np.cross(u, v)
    n = n / np.linalg.norm(n)
    p = [c - u - v, c + u - v, c + u + v, c - u + v]
    verts = np.array([np.concatenate([q, n]) for q in p], dtype=np.float32)
    tris = np.array([0, 1, 2, 0, 2, 3], dtype=np.int32)
    return verts, tris


def box(width=256, height=256, samples_camera=16, max_depth=4, seed=1):
    """Box scene: open-front Cornell-style box, two GGX conductor spheres,
    a small emissive quad under the ceiling."""
    rng = np.random.Generator(np.random.PCG64(seed))
    vpi = perspective_view_proj_inv((0.0, 0.0, 3.5), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 40.0, width / height)
    d = SceneDesc("box", width, height, vpi)
    X, Y, Z = (1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0)
    quads = [
        ((0, -1, 0), Z, X),   # floor, normal +y
        ((0, 1, 0), X, Z),    # ceiling, normal -y
        ((0, 0, -1), X, Y),   # back, normal +z
        ((-1, 0, 0), Y, Z),   # left, normal +x
        ((1, 0, 0), Z, Y),    # right, normal -x
    ]
    verts, tris, parts = [], [], []
    for i, (c, u, v) in enumerate(quads):
        qv, qt = _quad(c, u, v)
        albedo = rng.uniform(0.2, 0.8, size=3)
        m = dict(model=GGX, diffuse=tuple(float(a) for a in albedo), specular=(0.0, 0.0, 0.0),
                 specular_power=(1.0, 1.0))
        parts.append((m, len(tris) * 6, 6))
        tris.append(qt + 4 * i)
        verts.append(qv)
    h = 0.125
    ev, et = _quad((0.0, 0.999, 0.0), (h, 0.0, 0.0), (0.0, 0.0, h))  # x cross z = -y: faces down
    emitter = dict(model=EMISSIVE, diffuse=(15.0, 15.0, 15.0))
    parts.append((emitter, len(tris) * 6, 6))
    tris.append(et + 4 * len(quads))
    verts.append(ev)
    d.meshes.append((np.concatenate(verts), np.concatenate(tris), parts, False))
    conductor = dict(model=GGX_CONDUCTOR, diffuse=(0.0, 0.0, 0.0), specular=(1.0, 1.0, 1.0),
                     specular_power=(alpha_to_specular_power(0.2),) * 2, fresnel=(0.2, 3.0))
    d.spheres.append(((-0.45, -0.7, 0.0), 0.3, conductor))
    d.spheres.append(((0.45, -0.7, 0.0), 0.3, conductor))
    d.settings.update(samples_camera=samples_camera, samples_gi=1, samples_light=1, max_depth=max_depth, rr=1.0,
                      background=(0.0, 0.0, 0.0), seed=1)
    return d


def off_emitter_mask(ref_mean):
    """Pixels of a converged image [H, W, 3] whose radiance is below 1 in
    every channel, away from any pixel that is not (two 4-neighbour
    dilations): the box image without the directly seen emitter and its
    anti-aliased silhouette, whose pixel-sampling noise no radiance cache
    can reduce."""
    bright = ref_mean.max(axis=2) >= 1.0
    for _ in range(2):
        d = bright.copy()
        d[1:] |= bright[:-1]
        d[:-1] |= bright[1:]
        d[:, 1:] |= bright[:, :-1]
        d[:, :-1] |= bright[:, 1:]
        bright = d
    return ~bright


def sphere_light_box(width=48, height=48, samples_camera=4, max_depth=4, enclosing=False):
    """The box scene with a sphere-area light (addSphereLightSource): a small
    sphere under the ceiling beside the mesh emitter, or (enclosing) a sphere
    of radius 5 around the whole box and the camera, so every shading point
    lies inside it and samples it uniformly over its area."""
    d = box(width, height, samples_camera=samples_camera, max_depth=max_depth)
    d.name = "sphere_light_box"
    if enclosing:
        d.sphere_lights.append(dict(kind=LIGHT_SPHERE_AREA, radiance=(0.6, 0.5, 0.4), center=(0.0, 0.0, 0.0),
                                    radius=5.0))
    else:
        d.sphere_lights.append(dict(kind=LIGHT_SPHERE_AREA, radiance=(6.0, 5.0, 4.0), center=(0.35, 0.55, 0.3),
                                    radius=0.15))
    return d


def shapes(width=48, height=48, samples_camera=4, max_depth=4, seed=3, tangent_map=True):
    """The user shapes of the scene API in the box (walls and emitter of
    box(), without its spheres): a rect on the back wall with a texture-
    coordinate box and (tangent_map) an RGBA8 map of tangent directions
    near +x (sampleRGB reads them as byte / 255), a disk (annulus)
    on the floor, an upright cylinder lying along z, an oblique cylinder, a
    rotated ellipsoid and a blocker under the ceiling; a directional light
    besides the mesh emitter, so shadow rays cross the shapes."""
    rng = np.random.Generator(np.random.PCG64(seed))
    d = box(width, height, samples_camera=samples_camera, max_depth=max_depth)
    d.name = "shapes"
    d.spheres = []
    if tangent_map:
        t = np.abs(rng.normal(size=(4, 4, 3)) * 0.35 + np.array([1.0, 0.0, 0.0]))
        t /= np.linalg.norm(t, axis=-1, keepdims=True)
        b = np.clip(np.round(t * 255.0), 0, 255).astype(np.uint32)
        d.textures.append((b[..., 0] | (b[..., 1] << 8) | (b[..., 2] << 16) | np.uint32(255 << 24), 2, 0))
    rough = dict(model=GGX, diffuse=(0.5, 0.45, 0.4), specular=(0.4, 0.4, 0.4),
                 specular_power=(alpha_to_specular_power(0.35),) * 2)
    metal = dict(model=GGX_CONDUCTOR, diffuse=(0.0, 0.0, 0.0), specular=(1.0, 1.0, 1.0),
                 specular_power=(alpha_to_specular_power(0.25),) * 2, fresnel=(0.3, 2.8))
    beck = dict(model=BECKMANN, diffuse=(0.3, 0.5, 0.3), specular=(0.3, 0.3, 0.3),
                specular_power=(alpha_to_specular_power(0.4),) * 2)
    matte = dict(model=GGX, diffuse=(0.7, 0.6, 0.3), specular=(0.0, 0.0, 0.0), specular_power=(1.0, 1.0))
    d.shapes.append(("addRect", dict(pos=(0.0, 0.15, -0.97), tan=(1.0, 0.0, 0.0), norm=(0.0, 0.0, 1.0),
                                     size=(0.45, 0.3), tc=(0.0, 0.0, 2.0, 1.0),
                                     tangent_map=1 if tangent_map else 0), rough))
    d.shapes.append(("addDisk", dict(pos=(0.45, -0.99, 0.35), norm=(0.0, 1.0, 0.0), inner_radius=0.08,
                                     outer_radius=0.28), beck))
    d.shapes.append(("addCylinder", dict(center=(-0.5, -0.75, 0.1), radius=0.18, half_height=0.4), metal))
    d.shapes.append(("addObliqueCylinder", dict(center=(0.05, -0.6, -0.45), radius=0.1, half_height=0.3,
                                                tilt_direction=(1.0, 0.0), tilt=0.15), matte))
    q = np.array([0.2, 0.3, 0.1, 0.9])
    q /= np.linalg.norm(q)
    d.shapes.append(("addEllipsoid", dict(center=(0.45, -0.55, 0.0), scale=(0.22, 0.12, 0.16),
                                          orientation=tuple(float(v) for v in q)), rough))
    d.shapes.append(("addBlocker", dict(pos=(-0.35, 0.6, 0.2), tan=(1.0, 0.0, 0.0), norm=(0.0, -1.0, 0.0),
                                        size=(0.2, 0.15)), None))
    ld = np.array([0.3, 0.9, 0.4]) / np.linalg.norm([0.3, 0.9, 0.4])
    d.lights.append(dict(kind=LIGHT_DIRECTIONAL, direction=tuple(float(x) for x in ld), radiance=(1.5, 1.5, 1.5)))
    return d


# samples/virtual-gonioreflectometer.cc:33-41
GONIO_FIBER_RADIUS, GONIO_FIBER_LENGTH, GONIO_YARN_RADIUS = 0.1, 20.0, 1.0


def gonio_cylinder_light(samp, samples=9):
    """The light direction of request `samp` of the reference's loop
    (virtual-gonioreflectometer.cc:159-182): tc on a samples x samples grid
    over [-1, 1]^2, Direction = (-cos theta, -sin theta tc.x, -sin theta tc.y)
    with sin^2 theta = clamp(tc.tc, 0, 1)."""
    nxt = samp + 1
    tcx = 2.0 * (nxt % samples) / (samples - 1) - 1.0
    tcy = 2.0 * (nxt // samples) / (samples - 1) - 1.0
    s2 = min(max(tcx * tcx + tcy * tcy, 0.0), 1.0)
    ct, st = math.sqrt(1.0 - s2), math.sqrt(s2)
    return (-ct, -st * tcx, -st * tcy)


def gonio_cylinders(width=90, height=90, samples_camera=64):
    """The reference's own virtual gonioreflectometer scene
    (samples/virtual-gonioreflectometer.cc:43-143): three cylinders of
    radius (YarnRadius + FiberRadius) / 3 and half height FiberLength along
    z, at -2r y, +2r x and +2r y; GGX Diffuse 0.05, Specular 0.95,
    SpecularPower 10, Fresnel ComputeReflectionCoefficient(1.56); a
    directional light of radiance 10 from normalize(0, 1, 1); 64 camera
    samples, one light and one GI sample, depth 1, Russian roulette 0.9,
    black background; an orthographic camera of half size YarnRadius +
    FiberRadius looking along +x with z up (its view * rot)."""
    r = (GONIO_YARN_RADIUS + GONIO_FIBER_RADIUS) / 3.0
    span = GONIO_YARN_RADIUS + GONIO_FIBER_RADIUS
    e = np.array([-1.0, 0.0, 0.0])
    eye = 15.0 * e
    f = -e
    up = np.array([0.0, 0.0, 1.0])
    sv = np.cross(f, up)
    sv /= np.linalg.norm(sv)
    u = np.cross(sv, f)
    view = np.eye(4)
    view[0, :3], view[1, :3], view[2, :3] = sv, u, -f
    view[:3, 3] = -view[:3, :3] @ eye
    proj = np.eye(4)
    proj[0, 0] = proj[1, 1] = 1.0 / span
    proj[2, 2] = -2.0 / (1000.0 - 0.1)
    proj[2, 3] = -(1000.0 + 0.1) / (1000.0 - 0.1)
    vpi = np.ascontiguousarray(np.linalg.inv(proj @ view).T.reshape(16).astype(np.float32))
    d = SceneDesc("gonio_cylinders", width, height, vpi)
    f0 = ((1.56 - 1.0) / (1.56 + 1.0)) ** 2          # ComputeReflectionCoefficient(1.56)
    m = dict(model=GGX, diffuse=(0.05, 0.05, 0.05), specular=(0.95, 0.95, 0.95), specular_power=(10.0, 10.0),
             fresnel=(f0, 0.0))
    for c in ((0.0, -2.0 * r, 0.0), (2.0 * r, 0.0, 0.0), (0.0, 2.0 * r, 0.0)):
        d.shapes.append(("addCylinder", dict(center=c, radius=r, half_height=GONIO_FIBER_LENGTH), m))
    ld = np.array([0.0, 1.0, 1.0]) / math.sqrt(2.0)
    d.lights.append(dict(kind=LIGHT_DIRECTIONAL, direction=tuple(float(x) for x in ld), radiance=(10.0, 10.0, 10.0)))
    d.settings.update(samples_camera=samples_camera, samples_gi=1, samples_light=1, max_depth=1, rr=0.9,
                      background=(0.0, 0.0, 0.0), seed=1)
    return d


def value_noise_fbm(n, octaves=6, amplitude=0.1, seed=1):
    """fBm of bilinear value noise on an n x n grid over [-1, 1]^2
    (lacunarity 2, gain 0.5), heights in roughly [-amplitude, amplitude]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    xs = np.linspace(0.0, 1.0, n)
    h = np.zeros((n, n), dtype=np.float64)
    amp, freq, norm = 1.0, 4, 0.0
    for _ in range(octaves):
        lattice = rng.uniform(-1.0, 1.0, size=(freq + 1, freq + 1))
        g = xs * freq
        i0 = np.minimum(g.astype(np.int64), freq - 1)
        f = g - i0
        f = f * f * (3.0 - 2.0 * f)
        a = lattice[i0][:, i0]
        b = lattice[i0 + 1][:, i0]
        c = lattice[i0][:, i0 + 1]
        e = lattice[i0 + 1][:, i0 + 1]
        fy = f[:, None]
        fx = f[None, :]
        h += amp * ((a * (1 - fy) + b * fy) * (1 - fx) + (c * (1 - fy) + e * fy) * fx)
        norm += amp
        amp *= 0.5
        freq *= 2
    return (h / norm * amplitude).astype(np.float32)


def heightfield_mesh(n, amplitude=0.1, seed=1):
    """Vertices (position + normal, stride 24) and CCW-up triangles of an
    n x n heightfield y = h(x, z) over [-1, 1]^2."""
    h = value_noise_fbm(n, amplitude=amplitude, seed=seed)
    coords = np.linspace(-1.0, 1.0, n, dtype=np.float32)
    zz, xx = np.meshgrid(coords, coords, indexing="ij")  # row = z, col = x
    step = 2.0 / (n - 1)
    dhdx = np.gradient(h, step, axis=1)
    dhdz = np.gradient(h, step, axis=0)
    nrm = np.stack([-dhdx, np.ones_like(h), -dhdz], axis=-1)
    nrm /= np.linalg.norm(nrm, axis=-1, keepdims=True)
    verts = np.concatenate([np.stack([xx, h, zz], axis=-1), nrm], axis=-1).reshape(-1, 6).astype(np.float32)
    r, c = np.meshgrid(np.arange(n - 1), np.arange(n - 1), indexing="ij")
    v00 = (r * n + c).reshape(-1)
    v01 = v00 + 1
    v10 = v00 + n
    v11 = v10 + 1
    # (v00, v10, v11): (dz, 0) x (dz, dx) -> normal +y
    tris = np.stack([v00, v10, v11, v00, v11, v01], axis=-1).reshape(-1).astype(np.int32)
    return verts, tris, h


def _random_quaternion(rng):
    q = rng.normal(size=4)
    return q / np.linalg.norm(q)


def heightfield(width=1920, height=1080, samples_camera=64, max_depth=8, grid=1024, n_spheres=64, seed=1):
    rng = np.random.Generator(np.random.PCG64(seed + 1000))
    verts, tris, h = heightfield_mesh(grid, seed=seed)
    vpi = perspective_view_proj_inv((0.0, 1.1, 2.1), (0.0, -0.1, 0.0), (0.0, 1.0, 0.0), 45.0, width / height)
    d = SceneDesc("heightfield", width, height, vpi)
    alpha = float(rng.uniform(0.05, 0.6))
    ground = dict(model=GGX, diffuse=(0.45, 0.40, 0.35), specular=(0.5, 0.5, 0.5),
                  specular_power=(alpha_to_specular_power(alpha),) * 2, fresnel=(0.04, 0.0))
    d.meshes.append((verts, tris, [(ground, 0, tris.size)], False))
    n = grid
    for i in range(n_spheres):
        r = float(rng.uniform(0.02, 0.08))
        x, z = (float(v) for v in rng.uniform(-0.9, 0.9, size=2))
        ci = min(n - 1, max(0, int(round((x + 1.0) * 0.5 * (n - 1)))))
        ri = min(n - 1, max(0, int(round((z + 1.0) * 0.5 * (n - 1)))))
        y = float(h[ri, ci]) + r
        kind = i % 3
        a = float(rng.uniform(0.05, 0.6))
        col = tuple(float(v) for v in rng.uniform(0.2, 0.8, size=3))
        if kind == 0:
            m = dict(model=BECKMANN, diffuse=tuple(0.5 * c for c in col), specular=(0.5, 0.5, 0.5),
                     specular_power=(alpha_to_specular_power(a),) * 2, fresnel=(0.04, 0.0))
        elif kind == 1:
            sig = tuple(float(v) for v in rng.uniform(0.2, 0.8, size=2))
            m = dict(model=SGGX_SURFACE, diffuse=tuple(0.6 * c for c in col), specular=(0.4, 0.4, 0.4),
                     specular_power=sig, fresnel=(1.0, 0.0),
                     sggx_basis=tuple(float(v) for v in _random_quaternion(rng)))
        else:
            m = dict(model=GGX_CONDUCTOR, diffuse=(0.0, 0.0, 0.0), specular=col,
                     specular_power=(alpha_to_specular_power(a),) * 2, fresnel=(0.2, 3.0))
        d.spheres.append(((x, y, z), r, m))
    d.settings.update(samples_camera=samples_camera, samples_gi=1, samples_light=1, max_depth=max_depth, rr=1.0,
                      background=(1.0, 1.0, 1.0), seed=1)
    return d


def value_noise_3d(n, octaves=6, base_freq=4, seed=1):
    """fBm of trilinear value noise (smoothstep weights) on an n^3 grid over
    the unit cube, lacunarity 2, gain 0.5, normalised to about [-1, 1];
    float32 [z, y, x]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    xs = np.linspace(0.0, 1.0, n)
    out = np.zeros((n, n, n), dtype=np.float32)
    amp, freq, norm = 1.0, base_freq, 0.0
    for _ in range(octaves):
        lat = rng.uniform(-1.0, 1.0, size=(freq + 1, freq + 1, freq + 1)).astype(np.float32)
        g = xs * freq
        i0 = np.minimum(g.astype(np.int64), freq - 1)
        f = g - i0
        w = (f * f * (3.0 - 2.0 * f)).astype(np.float32)
        a = lat[:, :, i0] * (1.0 - w)[None, None, :] + lat[:, :, i0 + 1] * w[None, None, :]   # x
        b = a[:, i0, :] * (1.0 - w)[None, :, None] + a[:, i0 + 1, :] * w[None, :, None]       # y
        out += amp * (b[i0] * (1.0 - w)[:, None, None] + b[i0 + 1] * w[:, None, None])       # z
        norm += amp
        amp *= 0.5
        freq *= 2
    return out / np.float32(norm)


def volume_grid(n, mean_density=4.0, seed=1):
    """Density (binary16, mean `mean_density` per unit for sigma = 1) and
    unit fibre directions (binary16 x3) of the volume workload."""
    noise = value_noise_3d(n, seed=seed)
    raw = np.maximum(noise * 1.5 + 0.35, 0.0)
    dens = (raw * np.float32(mean_density / max(float(raw.mean()), 1e-6))).astype(np.float16)
    field = value_noise_3d(n, octaves=3, base_freq=3, seed=seed + 17)
    g = np.stack(np.gradient(field), axis=-1)[..., ::-1]          # (d/dx, d/dy, d/dz)
    ln = np.linalg.norm(g, axis=-1, keepdims=True)
    g = np.where(ln > 1e-12, g / np.maximum(ln, 1e-12), np.array([0.0, 0.0, 1.0], dtype=np.float32))
    return dens, g.astype(np.float16)


def volume(width=1024, height=1024, samples_camera=128, max_depth=32, grid=256, seed=1, with_floor_light=False):
    vpi = perspective_view_proj_inv((0.3, 0.45, 2.1), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 40.0, width / height)
    d = SceneDesc("volume", width, height, vpi)
    dens, dirs = volume_grid(grid, seed=seed)
    flakes = dict(model=SGGX_MICRO_FLAKE, albedo=(0.9, 0.9, 0.9), sggx_stddev=(0.1, 1.0, 1.0))
    d.volumes.append(((grid, grid, grid), (-0.5, -0.5, -0.5), (0.5, 0.5, 0.5), dens, dirs, flakes))
    if with_floor_light:
        qv, qt = _quad((0.0, -0.6, 0.0), (0.0, 0.0, 1.5), (1.5, 0.0, 0.0))
        floor = dict(model=GGX, diffuse=(0.6, 0.6, 0.6), specular=(0.0, 0.0, 0.0), specular_power=(1.0, 1.0))
        d.meshes.append((qv, qt, [(floor, 0, 6)], False))
        ld = np.array([0.3, 0.9, 0.2]) / np.linalg.norm([0.3, 0.9, 0.2])
        d.lights.append(dict(kind=LIGHT_DIRECTIONAL, direction=tuple(float(x) for x in ld), radiance=(2.0, 2.0, 2.0)))
    d.settings.update(samples_camera=samples_camera, samples_gi=1, samples_light=1, max_depth=max_depth, rr=1.0,
                      background=(1.0, 1.0, 1.0), seed=1)
    return d


def brick_volume_root(bricks=(2, 1, 2), n=12, seed=3, empty=((1, 0, 0),), lo=(-0.5, -0.5, -0.5), hi=(0.5, 0.5, 0.5)):
    """A VolumeRoot of `bricks` bricks of n^3 voxels in the reference's brick
    layout: u8 densities from 3-D value noise (about a third of the voxels
    zero) followed by (theta, phi) byte pairs of a smooth fibre field;
    the bricks listed in `empty` have no data."""
    from .scene import VolumeRoot
    vols = []
    k = 0
    for z in range(bricks[2]):
        for y in range(bricks[1]):
            for x in range(bricks[0]):
                k += 1
                if (x, y, z) in empty:
                    vols.append(None)
                    continue
                noise = value_noise_3d(n, octaves=3, base_freq=3, seed=seed + k)
                dens = np.clip((noise - 0.2) * 500.0, 0.0, 255.0).astype(np.uint8).reshape(-1)
                f1 = value_noise_3d(n, octaves=2, base_freq=2, seed=seed + 100 + k).reshape(-1)
                f2 = value_noise_3d(n, octaves=2, base_freq=2, seed=seed + 200 + k).reshape(-1)
                theta = np.clip(f1 * 255.0, 0, 255).astype(np.uint8)
                phi = np.clip(f2 * 255.0, 0, 255).astype(np.uint8)
                data = np.concatenate([dens, np.stack([theta, phi], axis=1).reshape(-1)])
                vols.append(((n, n, n), data))
    return VolumeRoot(lo, hi, bricks, vols)


def brick_volume(width=48, height=48, samples_camera=4, max_depth=8, seed=3, with_floor_light=True):
    """A reference-format brick volume (addHierarchicalVolume) over a floor
    lit by a directional light, in a white sky."""
    vpi = perspective_view_proj_inv((0.3, 0.45, 2.1), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 40.0, width / height)
    d = SceneDesc("brick_volume", width, height, vpi)
    flakes = dict(model=SGGX_MICRO_FLAKE, albedo=(0.8, 0.85, 0.9), sggx_stddev=(0.2, 1.0, 1.0))
    d.brick_volumes.append((brick_volume_root(seed=seed), flakes))
    if with_floor_light:
        qv, qt = _quad((0.0, -0.6, 0.0), (0.0, 0.0, 1.5), (1.5, 0.0, 0.0))
        floor = dict(model=GGX, diffuse=(0.6, 0.6, 0.6), specular=(0.0, 0.0, 0.0), specular_power=(1.0, 1.0))
        d.meshes.append((qv, qt, [(floor, 0, 6)], False))
        ld = np.array([0.3, 0.9, 0.2]) / np.linalg.norm([0.3, 0.9, 0.2])
        d.lights.append(dict(kind=LIGHT_DIRECTIONAL, direction=tuple(float(x) for x in ld), radiance=(2.0, 2.0, 2.0)))
    d.settings.update(samples_camera=samples_camera, samples_gi=1, samples_light=1, max_depth=max_depth, rr=1.0,
                      background=(1.0, 1.0, 1.0), seed=1)
    return d


def _textured_quad(center, u, v, tc_scale=1.0, tangent=False):
    """A quad with texture coordinates 0..tc_scale: PTNFormat (stride 32)
    or PTTNFormat (stride 44, tangent along u)."""
    c, u, v = (np.asarray(x, dtype=np.float32) for x in (center, u, v))
    n = np.cross(u, v)
    n = n / np.linalg.norm(n)
    t = u / np.linalg.norm(u)
    pts = [c - u - v, c + u - v, c + u + v, c - u + v]
    tcs = [(0.0, 0.0), (tc_scale, 0.0), (tc_scale, tc_scale), (0.0, tc_scale)]
    rows = []
    for p, q in zip(pts, tcs):
        rows.append(np.concatenate([p, np.float32(q), t, n] if tangent else [p, np.float32(q), n]))
    return np.array(rows, dtype=np.float32), np.array([0, 1, 2, 0, 2, 3], dtype=np.int32)


def textured(width=48, height=48, samples_camera=4, max_depth=4, seed=5):
    """Texture maps and a cube-map environment: a Blinn-Phong floor with an
    sRGB checker DiffuseMap (bilinear, coordinates wrapping three times), an
    SGGX-surface wall with Diffuse (RGB32F), Specular (RGBA8),
    StandardDeviation (RG8, nearest) and Basis (RGBA32F quaternion) maps, a
    Blinn-Phong sphere with a DiffuseMap, a directional light and a cube map
    of six distinct gradient faces."""
    rng = np.random.Generator(np.random.PCG64(seed))
    vpi = perspective_view_proj_inv((0.0, 0.6, 3.2), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 45.0, width / height)
    d = SceneDesc("textured", width, height, vpi)
    checker = np.zeros((16, 16), dtype=np.uint32)
    for y in range(16):
        for x in range(16):
            a, b = (200, 40) if (x // 4 + y // 4) % 2 == 0 else (30, 180)
            checker[y, x] = a | (b << 8) | (((a + b) // 2) << 16) | (255 << 24)
    d.textures.append((checker, 2, 0))                                      # 1: RGBA8 bilinear
    grad = np.stack(np.meshgrid(np.linspace(0.1, 0.9, 8), np.linspace(0.2, 0.7, 8)), -1)
    d.textures.append((np.concatenate([grad, 0.5 * np.ones((8, 8, 1))], -1).astype(np.float32), 5, 0))  # 2: RGB32F
    spec = rng.integers(0, 256, size=(8, 8), dtype=np.uint32) * 0x010101 | 0xFF000000
    d.textures.append((spec.astype(np.uint32), 2, 0))                       # 3: RGBA8
    sd = rng.integers(20, 255, size=(8, 8, 2), dtype=np.uint8)
    d.textures.append((sd, 1, 1))                                           # 4: RG8 nearest
    q = rng.normal(size=(8, 8, 4)).astype(np.float32)
    q[..., 3] = np.abs(q[..., 3]) + 1.0
    d.textures.append((q, 6, 0))                                            # 5: RGBA32F
    floor = dict(model=1, diffuse=(0.5, 0.5, 0.5), specular=(0.2, 0.2, 0.2), specular_power=(20.0, 20.0),
                 diffuse_map=1)
    wall = dict(model=SGGX_SURFACE, diffuse=(0.3, 0.3, 0.3), specular=(0.6, 0.6, 0.6), specular_power=(0.4, 0.4),
                diffuse_map=2, specular_map=3, stddev_map=4, basis_map=5)
    fv, ft = _textured_quad((0.0, -0.5, 0.0), (0.0, 0.0, 1.2), (1.2, 0.0, 0.0), tc_scale=3.0)
    d.meshes.append((fv, ft, [(floor, 0, 6)], False))
    wv, wt = _textured_quad((0.0, 0.3, -1.0), (1.2, 0.0, 0.0), (0.0, 0.8, 0.0), tangent=True)
    d.meshes.append((wv, wt, [(wall, 0, 6)], False))
    d.spheres.append(((0.4, -0.2, 0.2), 0.3, dict(model=1, diffuse=(0.2, 0.2, 0.2), specular=(0.5, 0.5, 0.5),
                                                  specular_power=(40.0, 40.0), diffuse_map=1)))
    side = 24
    faces = []
    for f in range(6):
        yy, xx = np.mgrid[0:side, 0:side]
        r = (40 * f + 4 * xx) % 256
        g = (255 - 8 * yy) % 256
        b = (30 * f + 2 * (xx + yy)) % 256
        faces.append((r | (g << 8) | (b << 16) | (255 << 24)).astype(np.uint32))
    d.cube_faces = faces
    ld = np.array([0.4, 0.8, 0.45]) / np.linalg.norm([0.4, 0.8, 0.45])
    d.lights.append(dict(kind=LIGHT_DIRECTIONAL, direction=tuple(float(x) for x in ld), radiance=(2.0, 2.0, 2.0)))
    d.settings.update(samples_camera=samples_camera, samples_gi=1, samples_light=1, max_depth=max_depth, rr=1.0,
                      background=(0.0, 0.0, 0.0), seed=1)
    return d


GGX_ANISOTROPIC = 15
GGX_CONDUCTOR_ANISOTROPIC = 16
GGX_DIELECTRIC_ANISOTROPIC = 17
BECKMANN_ANISOTROPIC = 31
BECKMANN_CONDUCTOR_ANISOTROPIC = 32
BECKMANN_DIELECTRIC_ANISOTROPIC = 33


def anisotropic(width=48, height=48, samples_camera=4, max_depth=4):
    """The six anisotropic microfacet models: a brushed-conductor GGX floor
    and a Beckmann dielectric wall (PTTNFormat: tangent along the first
    edge), four spheres (tangent frame from the sphere normal) with the other
    variants, a PTNFormat quad (no tangent: the anisotropy collapses) with a
    Beckmann conductor, a directional light and a point light."""
    vpi = perspective_view_proj_inv((0.0, 0.7, 3.0), (0.0, -0.1, 0.0), (0.0, 1.0, 0.0), 45.0, width / height)
    d = SceneDesc("anisotropic", width, height, vpi)

    def mat(model, ax, ay, diffuse, specular, fresnel):
        return dict(model=model, diffuse=diffuse, specular=specular,
                    specular_power=(alpha_to_specular_power(ax), alpha_to_specular_power(ay)), fresnel=fresnel)

    floor = mat(GGX_CONDUCTOR_ANISOTROPIC, 0.08, 0.45, (0.05, 0.05, 0.05), (0.9, 0.8, 0.6), (0.2, 3.0))
    wall = mat(BECKMANN_DIELECTRIC_ANISOTROPIC, 0.5, 0.1, (0.4, 0.3, 0.2), (0.5, 0.5, 0.5), (1.5, 0.0))
    fv, ft = _textured_quad((0.0, -0.5, 0.0), (0.0, 0.0, 1.3), (1.3, 0.0, 0.0), tangent=True)
    d.meshes.append((fv, ft, [(floor, 0, 6)], False))
    wv, wt = _textured_quad((0.0, 0.3, -1.0), (1.3, 0.0, 0.0), (0.0, 0.8, 0.0), tangent=True)
    d.meshes.append((wv, wt, [(wall, 0, 6)], False))
    side = mat(BECKMANN_CONDUCTOR_ANISOTROPIC, 0.3, 0.05, (0.1, 0.1, 0.1), (0.7, 0.7, 0.8), (0.3, 2.5))
    sv, st = _textured_quad((-1.2, 0.0, -0.2), (0.0, 0.0, 0.8), (0.0, 0.6, 0.0))
    d.meshes.append((sv, st, [(side, 0, 6)], False))
    d.spheres.append(((-0.55, -0.2, 0.1), 0.3, mat(GGX_ANISOTROPIC, 0.1, 0.6, (0.3, 0.1, 0.1), (0.6, 0.6, 0.6),
                                                     (0.04, 0.0))))
    d.spheres.append(((0.0, -0.2, 0.3), 0.3, mat(GGX_DIELECTRIC_ANISOTROPIC, 0.4, 0.07, (0.1, 0.3, 0.1),
                                                   (0.5, 0.5, 0.5), (1.6, 0.0))))
    d.spheres.append(((0.55, -0.2, 0.1), 0.3, mat(BECKMANN_ANISOTROPIC, 0.06, 0.35, (0.1, 0.1, 0.3),
                                                    (0.6, 0.6, 0.6), (0.05, 0.0))))
    d.spheres.append(((0.3, 0.45, -0.3), 0.15, mat(GGX_CONDUCTOR_ANISOTROPIC, 0.2, 0.2, (0.0, 0.0, 0.0),
                                                     (0.9, 0.9, 0.9), (0.2, 3.5))))
    ld = np.array([0.3, 0.9, 0.4]) / np.linalg.norm([0.3, 0.9, 0.4])
    d.lights.append(dict(kind=LIGHT_DIRECTIONAL, direction=tuple(float(x) for x in ld), radiance=(1.5, 1.5, 1.5)))
    d.lights.append(dict(kind=LIGHT_POINT, position=(-0.4, 1.2, 0.8), radiance=(3.0, 2.5, 2.0)))
    d.settings.update(samples_camera=samples_camera, samples_gi=1, samples_light=1, max_depth=max_depth, rr=1.0,
                      background=(0.2, 0.25, 0.3), seed=1)
    return d


SGGX_PSEUDO_VOLUME = 25


def pseudo_volume(width=32, height=32, samples_camera=2, max_depth=3, seed=9):
    """The SGGX pseudo-volume surface (model 25): a quad (PTTNFormat, texture
    coordinates 0..1, so its BRDF's walks are seeded per point) with
    StandardDeviation (RG8) and basis (RGBA32F) maps and 256 walks of depth
    2, two spheres with constant parameters (64 walks of depth 3; 32 walks
    of depth 1 and a specular channel under the walks' 1e-3 cut), an
    anisotropic Beckmann-conductor sphere, a GGX floor and a directional
    light."""
    rng = np.random.Generator(np.random.PCG64(seed))
    vpi = perspective_view_proj_inv((0.0, 0.5, 2.8), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 45.0, width / height)
    d = SceneDesc("pseudo_volume", width, height, vpi)
    sd = rng.integers(30, 255, size=(8, 8, 2), dtype=np.uint8)
    d.textures.append((sd, 1, 0))                                           # 1: RG8 bilinear
    q = rng.normal(size=(8, 8, 4)).astype(np.float32)
    q /= np.linalg.norm(q, axis=-1, keepdims=True)
    d.textures.append((q, 6, 1))                                            # 2: RGBA32F nearest
    wall = dict(model=SGGX_PSEUDO_VOLUME, diffuse=(0.1, 0.1, 0.1), specular=(0.8, 0.7, 0.6),
                specular_power=(0.3, 0.3), stddev_map=1, basis_map=2, sggx_depth=2, sggx_sample_count=256)
    wv, wt = _textured_quad((0.0, 0.2, -0.8), (1.0, 0.0, 0.0), (0.0, 0.7, 0.0), tangent=True)
    d.meshes.append((wv, wt, [(wall, 0, 6)], False))
    floor = dict(model=GGX, diffuse=(0.4, 0.4, 0.4), specular=(0.2, 0.2, 0.2),
                 specular_power=(alpha_to_specular_power(0.3),) * 2)
    fv, ft = _quad((0.0, -0.5, 0.0), (0.0, 0.0, 1.2), (1.2, 0.0, 0.0))
    d.meshes.append((fv, ft, [(floor, 0, 6)], False))
    d.spheres.append(((-0.4, -0.2, 0.1), 0.3, dict(model=SGGX_PSEUDO_VOLUME, diffuse=(0.05, 0.1, 0.05),
                                                     specular=(0.9, 0.9, 0.9), specular_power=(0.15, 0.6),
                                                     sggx_basis=(0.0, 0.38268343, 0.0, 0.9238795),
                                                     sggx_depth=3, sggx_sample_count=64)))
    d.spheres.append(((0.45, -0.25, 0.2), 0.25, dict(model=SGGX_PSEUDO_VOLUME, diffuse=(0.2, 0.1, 0.1),
                                                       specular=(0.7, 0.0005, 0.7), specular_power=(0.5, 0.2),
                                                       sggx_depth=1, sggx_sample_count=32)))
    # an anisotropic sphere, so the texture-aware kernel instances run those models too
    d.spheres.append(((0.05, 0.35, -0.3), 0.18, dict(model=BECKMANN_CONDUCTOR_ANISOTROPIC, diffuse=(0.1, 0.1, 0.1),
                                                       specular=(0.8, 0.8, 0.8),
                                                       specular_power=(alpha_to_specular_power(0.08),
                                                                       alpha_to_specular_power(0.4)),
                                                       fresnel=(0.3, 2.5))))
    ld = np.array([0.3, 0.8, 0.5]) / np.linalg.norm([0.3, 0.8, 0.5])
    d.lights.append(dict(kind=LIGHT_DIRECTIONAL, direction=tuple(float(x) for x in ld), radiance=(2.0, 2.0, 2.0)))
    d.settings.update(samples_camera=samples_camera, samples_gi=1, samples_light=1, max_depth=max_depth, rr=1.0,
                      background=(0.3, 0.3, 0.35), seed=1)
    return d


def random_rays(desc, n, seed=7, kind="sphere"):
    """A fixed ray set: origins uniform in the scene's box [-1.2, 1.2]^3 (or
    the camera for kind='camera'), directions uniform on the sphere."""
    rng = np.random.Generator(np.random.PCG64(seed))
    org = rng.uniform(-1.2, 1.2, size=(n, 3))
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    rays = np.zeros((n, 8), dtype=np.float32)
    rays[:, 0:3] = org
    rays[:, 3] = 0.0
    rays[:, 4:7] = v
    rays[:, 7] = np.inf
    return rays


def surface_rays(desc, n, seed=11, offset=1e-3):
    """Secondary rays of a fixed ray set: origins at uniformly chosen points
    of uniformly chosen triangles of the description's first mesh, lifted by
    `offset` along the triangle's front normal, directions uniform on the
    sphere (float64, stored as float32)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    verts, tris = desc.meshes[0][0], desc.meshes[0][1].reshape(-1, 3)
    pos = verts[:, :3].astype(np.float64)
    t = rng.integers(0, tris.shape[0], size=n)
    r1, r2 = rng.random(n), rng.random(n)
    s = np.sqrt(r1)
    p0, p1, p2 = pos[tris[t, 0]], pos[tris[t, 1]], pos[tris[t, 2]]
    pt = p0 * (1.0 - s)[:, None] + p1 * (s * (1.0 - r2))[:, None] + p2 * (s * r2)[:, None]
    nrm = np.cross(p1 - p0, p2 - p0)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    rays = np.zeros((n, 8), dtype=np.float32)
    rays[:, 0:3] = pt + offset * nrm
    rays[:, 4:7] = v
    rays[:, 7] = np.inf
    return rays


def primary_rays(desc, stride=1):
    """Pixel-centre primary rays of the description's camera (one per pixel
    when stride=1), computed in float64 from the inverse view-projection."""
    W, H = desc.width, desc.height
    m = desc.vpi.astype(np.float64).reshape(4, 4).T
    ys, xs = np.mgrid[0:H:stride, 0:W:stride]
    sx = 2.0 * (xs.reshape(-1) + 0.5) / W - 1.0
    sy = 2.0 * (ys.reshape(-1) + 0.5) / H - 1.0
    p0 = m @ np.stack([sx, sy, -np.ones_like(sx), np.ones_like(sx)])
    p1 = m @ np.stack([sx, sy, np.ones_like(sx), np.ones_like(sx)])
    a = (p0[:3] / p0[3]).T
    b = (p1[:3] / p1[3]).T
    dvec = b - a
    dvec /= np.linalg.norm(dvec, axis=1, keepdims=True)
    rays = np.zeros((a.shape[0], 8), dtype=np.float32)
    rays[:, 0:3] = a
    rays[:, 4:7] = dvec
    rays[:, 7] = np.inf
    return rays


def ortho_view_proj_inv(eye_dir, half_extent, dist=5.0, near=0.1, far=20.0):
    """Inverse of an orthographic projection (GL clip z in [-1, 1], the
    reference's OrthoMatrix) times a lookAt from dist * eye_dir towards the
    origin, as 16 column-major floats; square view of half size
    half_extent."""
    e = np.asarray(eye_dir, dtype=np.float64)
    e = e / np.linalg.norm(e)
    eye = dist * e
    f = -e
    up = np.array([0.0, 1.0, 0.0]) if abs(f[1]) < 0.99 else np.array([0.0, 0.0, 1.0])
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    view = np.eye(4)
    view[0, :3], view[1, :3], view[2, :3] = s, u, -f
    view[:3, 3] = -view[:3, :3] @ eye
    proj = np.eye(4)
    proj[0, 0] = proj[1, 1] = 1.0 / half_extent
    proj[2, 2] = -2.0 / (far - near)
    proj[2, 3] = -(far + near) / (far - near)
    inv = np.linalg.inv(proj @ view)
    return np.ascontiguousarray(inv.T.reshape(16).astype(np.float32))


def hemisphere_grid(n=9):
    """n x n directions over the upper (+y) hemisphere from the square grid
    tc in [-1, 1]^2 (the reference's request grid,
    samples/virtual-gonioreflectometer.cc:167-171): (tc.x, cos, tc.y)
    normalised, cos = sqrt(max(1 - |tc|^2, 0.05))."""
    out = []
    for j in range(n):
        for i in range(n):
            tx, ty = 2.0 * i / (n - 1) - 1.0, 2.0 * j / (n - 1) - 1.0
            c = math.sqrt(max(1.0 - tx * tx - ty * ty, 0.05))
            v = np.array([tx, c, ty])
            out.append(tuple(float(x) for x in v / np.linalg.norm(v)))
    return out


def gonio(grid=64, n_dirs=9, seed=1, max_requests=None):
    """Batched virtual gonioreflectometer: a grid x grid fBm heightfield
    patch (2 (grid-1)^2 triangles; GGX alpha 0.3 with the reference
    sample's Diffuse 0.05 / Specular 0.95 / Fresnel of IOR 1.56,
    samples/virtual-gonioreflectometer.cc:99-127), a directional light of
    radiance 10, depth 1, Russian roulette 0.9; requests for every
    (light, view) pair of the n_dirs^2 hemisphere grid, an orthographic
    camera per view, image side 2^U{5..8} and camera samples from
    {4, 8, 16, 32, 64} drawn by Philox seed `seed`.  Returns
    (SceneDesc, requests)."""
    verts, tris, _ = heightfield_mesh(grid, amplitude=0.1, seed=seed)
    d = SceneDesc("gonio", 32, 32, ortho_view_proj_inv((0.0, 1.0, 0.0), 1.5))
    f0 = ((1.56 - 1.0) / (1.56 + 1.0)) ** 2          # ComputeReflectionCoefficient(1.56)
    m = dict(model=GGX, diffuse=(0.05, 0.05, 0.05), specular=(0.95, 0.95, 0.95),
             specular_power=(alpha_to_specular_power(0.3),) * 2, fresnel=(f0, 0.0))
    d.meshes.append((verts, tris, [(m, 0, tris.size)], False))
    dirs = hemisphere_grid(n_dirs)
    d.lights.append(dict(kind=LIGHT_DIRECTIONAL, direction=dirs[len(dirs) // 2], radiance=(10.0, 10.0, 10.0)))
    d.settings.update(samples_camera=1, samples_gi=1, samples_light=1, max_depth=1, rr=0.9,
                      background=(0.0, 0.0, 0.0), seed=1)
    rng = np.random.Generator(np.random.Philox(seed))
    reqs = []
    for li, ld in enumerate(dirs):
        for vi, vd in enumerate(dirs):
            side = 1 << int(rng.integers(5, 9))
            spp = int(rng.choice([4, 8, 16, 32, 64]))
            reqs.append(dict(view_proj_inv=ortho_view_proj_inv(vd, 1.5), width=side, height=side,
                             samples_camera=spp, light_direction=ld))
            if max_requests is not None and len(reqs) >= max_requests:
                return d, reqs
    return d, reqs
