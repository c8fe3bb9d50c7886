"""Host-side mirror of the reference ray-trace scene API over the C ABI.

`RayTracerScene` keeps the method names and argument meaning of the
reference classes RayTracerScene (include/tempest/graphics/ray-tracing/
ray-tracing.hh:241-395) and RayTracerCuda (include/tempest/compute/
ray-tracing-cuda.hh:80-172), so code written against them reads the same.
Every call goes through libtempest_rt.so; failures raise TempestError.
"""
import ctypes
import os

import numpy as np

from . import _abi
from ._abi import check

_f = ctypes.c_float


def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(_f))


def _rgba_arg(rgba):
    a = np.ascontiguousarray(rgba, dtype=np.uint32)
    if a.ndim != 2:
        raise ValueError("rgba must be [height, width] uint32")
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


def encode_image(rgba, fmt=_abi.TR_IMAGE_TGA):
    """File bytes of an RGBA8 picture [height, width] (row 0 at the bottom)
    as TGA (SaveTGAImage, src/image/tga-image.cc:247-309) or PNG."""
    a, p = _rgba_arg(rgba)
    lib = _abi.load()
    n = ctypes.c_size_t()
    check(lib.tr_image_encode(int(fmt), a.shape[1], a.shape[0], p, None, 0, ctypes.byref(n)))
    out = np.zeros(n.value, dtype=np.uint8)
    check(lib.tr_image_encode(int(fmt), a.shape[1], a.shape[0], p,
                              out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), n.value, ctypes.byref(n)))
    return out.tobytes()


def save_image(path, rgba):
    """SaveImage (src/image/image.cc:60-80): format from the extension."""
    a, p = _rgba_arg(rgba)
    check(_abi.load().tr_image_save(os.fsencode(path), a.shape[1], a.shape[0], p))


class Material:
    """RTMaterial / RTMicrofacetMaterial / RTSGGXSurface as one record."""

    def __init__(self, model, **fields):
        self.rec = _abi.tr_material()
        check(_abi.load().tr_material_init(ctypes.byref(self.rec), int(model)))
        for k, v in fields.items():
            self.set(k, v)

    def set(self, name, value):
        cur = getattr(self.rec, name)
        if hasattr(cur, "__len__"):
            for i, x in enumerate(value):
                cur[i] = float(x)
        else:
            setattr(self.rec, name, value)
        return self

    def setup(self):
        """RTMaterial::setup (ray-tracing.cc:80-87)."""
        check(_abi.load().tr_material_setup(ctypes.byref(self.rec)))
        return self


class Submesh:
    """RTSubmesh (ray-tracing-common.hh:240-247)."""

    def __init__(self, material, vertex_count, vertex_offset=0, base_index=0, stride=24):
        self.material = material
        self.vertex_count = int(vertex_count)
        self.vertex_offset = int(vertex_offset)
        self.base_index = int(base_index)
        self.stride = int(stride)


def make_light(kind, radiance=(0, 0, 0), direction=(0, 0, 1), position=(0, 0, 0), center=(0, 0, 0), radius=0.0):
    rec = _abi.tr_light()
    rec.type = int(kind)
    for i in range(3):
        rec.radiance[i] = radiance[i]
        rec.direction[i] = direction[i]
        rec.position[i] = position[i]
        rec.sphere_center[i] = center[i]
    rec.sphere_radius = radius
    return rec


def submesh_array(submeshes):
    arr = (_abi.tr_submesh * max(1, len(submeshes)))()
    for i, sm in enumerate(submeshes):
        arr[i].material = ctypes.pointer(sm.material.rec)
        arr[i].vertex_count = sm.vertex_count
        arr[i].vertex_offset = sm.vertex_offset
        arr[i].base_index = sm.base_index
        arr[i].stride = sm.stride
    return arr


def grid_volume_record(dims, min_corner, max_corner, density, direction, max_density=0.0):
    """A tr_grid_volume over binary16 copies of the arrays (returned too, to
    keep them alive for the call)."""
    d = np.ascontiguousarray(np.asarray(density).astype(np.float16, copy=False).reshape(-1)).view(np.uint16)
    v = np.ascontiguousarray(np.asarray(direction).astype(np.float16, copy=False).reshape(-1)).view(np.uint16)
    n = int(dims[0]) * int(dims[1]) * int(dims[2])
    if d.size != n or v.size != 3 * n:
        raise ValueError("density / direction sizes do not match dims")
    rec = _abi.tr_grid_volume()
    for a in range(3):
        rec.dims[a] = int(dims[a])
        rec.min_corner[a] = float(min_corner[a])
        rec.max_corner[a] = float(max_corner[a])
    rec.density = d.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16))
    rec.direction = v.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16))
    rec.max_density = float(max_density)
    return rec, (d, v)


class VolumeRoot:
    """VolumeRoot (volume/volume.hh:65-83): MinCorner, MaxCorner, Dimensions
    (bricks per axis) and Volumes, a list in (z, y, x) order of None or
    (dims (x, y, z), uint8 data: densities then (theta, phi) pairs)."""

    def __init__(self, min_corner, max_corner, dims, volumes):
        self.MinCorner = tuple(float(v) for v in min_corner)
        self.MaxCorner = tuple(float(v) for v in max_corner)
        self.Dimensions = tuple(int(v) for v in dims)
        self.Volumes = list(volumes)


def volume_root_record(root):
    """A tr_volume_root over the VolumeRoot's arrays (returned too, to keep
    them alive for the call)."""
    rec = _abi.tr_volume_root()
    for a in range(3):
        rec.min_corner[a] = root.MinCorner[a]
        rec.max_corner[a] = root.MaxCorner[a]
        rec.dims[a] = root.Dimensions[a]
    n = root.Dimensions[0] * root.Dimensions[1] * root.Dimensions[2]
    if len(root.Volumes) != n:
        raise ValueError("VolumeRoot.Volumes must hold one entry per brick")
    arr = (_abi.tr_volume_brick * max(1, n))()
    keep = [arr]
    for i, v in enumerate(root.Volumes):
        if v is None:
            continue
        dims, data = v
        d = np.ascontiguousarray(data, dtype=np.uint8).reshape(-1)
        if d.size != 3 * int(dims[0]) * int(dims[1]) * int(dims[2]):
            raise ValueError("a brick holds 3 bytes per voxel")
        keep.append(d)
        for a in range(3):
            arr[i].dims[a] = int(dims[a])
        arr[i].data = d.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
    rec.bricks = arr
    return rec, keep


def texture_record(texels, fmt, sampling=0):
    """A tr_texture_desc and a contiguous byte view of the texels."""
    a = np.ascontiguousarray(texels)
    if a.ndim < 2:
        raise ValueError("texels must be [height, width(, channels)]")
    desc = _abi.tr_texture_desc(int(a.shape[1]), int(a.shape[0]), int(fmt), int(sampling))
    sizes = {0: 1, 1: 2, 2: 4, 3: 4, 4: 8, 5: 12, 6: 16}
    if int(fmt) in sizes and a.nbytes != a.shape[0] * a.shape[1] * sizes[int(fmt)]:
        raise ValueError("texel array size does not match the format")
    return desc, a.view(np.uint8).reshape(-1)


def cube_faces(faces):
    arrs = [np.ascontiguousarray(f, dtype=np.uint32) for f in faces]
    if len(arrs) != 6 or any(a.shape != arrs[0].shape or a.shape[0] != a.shape[1] for a in arrs):
        raise ValueError("six square faces of one size")
    ptrs = (ctypes.c_void_p * 6)(*[a.ctypes.data for a in arrs])
    return arrs, ptrs


def _root_from_raw(raw):
    n = raw.dims[0] * raw.dims[1] * raw.dims[2]
    vols = []
    for i in range(n):
        b = raw.bricks[i]
        if not b.data:
            vols.append(None)
            continue
        dims = tuple(b.dims)
        size = 3 * dims[0] * dims[1] * dims[2]
        vols.append((dims, np.ctypeslib.as_array(b.data, shape=(size,)).copy()))
    return VolumeRoot(tuple(raw.min_corner), tuple(raw.max_corner), tuple(raw.dims), vols)


def ParseVolumeHierarchy(name):
    """ParseVolumeHierarchy (src/volume/volume.cc:361-420)."""
    lib = _abi.load()
    raw = _abi.tr_volume_root()
    check(lib.tr_parse_volume_hierarchy(os.fsencode(name), ctypes.byref(raw)))
    try:
        return _root_from_raw(raw)
    finally:
        lib.tr_volume_root_free(ctypes.byref(raw))


def ParseVolume(density, orientation):
    """ParseVolume (src/volume/volume.cc:422-453)."""
    lib = _abi.load()
    raw = _abi.tr_volume_root()
    check(lib.tr_parse_volume(os.fsencode(density), os.fsencode(orientation), ctypes.byref(raw)))
    try:
        return _root_from_raw(raw)
    finally:
        lib.tr_volume_root_free(ctypes.byref(raw))


class RayTracerScene:
    """The MI355X tracer behind the reference scene API."""

    def __init__(self, width, height, view_proj_inv, path_pool_size=0, trace_waves_per_cu=0, shade_order=0):
        self.lib = _abi.load()
        vpi = np.ascontiguousarray(np.asarray(view_proj_inv, dtype=np.float32).reshape(16))
        settings = _abi.tr_settings(32 << 20, int(path_pool_size), int(trace_waves_per_cu), int(shade_order))
        h = ctypes.c_void_p()
        check(self.lib.tr_scene_create(int(width), int(height), _fptr(vpi), ctypes.byref(settings), ctypes.byref(h)))
        self.h = h
        self.width, self.height = int(width), int(height)
        self._keep = []

    def close(self):
        if self.h:
            self.lib.tr_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    make_light = staticmethod(make_light)

    # setters (ray-tracing.hh:305-321)
    def setChunkSize(self, n): check(self.lib.tr_scene_set_chunk_size(self.h, int(n)))
    def setSamplesLocalAreaLight(self, n): check(self.lib.tr_scene_set_samples_local_area_light(self.h, int(n)))
    def setSamplesGlobalIllumination(self, n): check(self.lib.tr_scene_set_samples_global_illumination(self.h, int(n)))
    def setSamplesCamera(self, n): check(self.lib.tr_scene_set_samples_camera(self.h, int(n)))
    def setMaxRayDepth(self, n): check(self.lib.tr_scene_set_max_ray_depth(self.h, int(n)))
    def setRussianRoulette(self, r): check(self.lib.tr_scene_set_russian_roulette(self.h, float(r)))
    def setPicturePostProcess(self, m): check(self.lib.tr_scene_set_picture_post_process(self.h, int(m)))
    def setTransparentBackground(self, e): check(self.lib.tr_scene_set_transparent_background(self.h, int(bool(e))))
    def setRenderMode(self, m): check(self.lib.tr_scene_set_render_mode(self.h, int(m)))
    def repaint(self): check(self.lib.tr_scene_repaint(self.h))

    def setBackgroundSpectrum(self, rgb):
        a = np.asarray(rgb, dtype=np.float32)
        check(self.lib.tr_scene_set_background_spectrum(self.h, _fptr(a)))

    # additions beyond the reference
    def setSeed(self, seed): check(self.lib.tr_scene_set_seed(self.h, int(seed)))
    def setEstimator(self, mode): check(self.lib.tr_scene_set_estimator(self.h, int(mode)))
    def setFrameIndex(self, f): check(self.lib.tr_scene_set_frame_index(self.h, int(f)))
    def setCounting(self, on): check(self.lib.tr_scene_set_counting(self.h, int(bool(on))))
    def setTiming(self, on): check(self.lib.tr_scene_set_timing(self.h, int(bool(on))))

    # lights (ray-tracing.hh:326-331)
    def addLightSource(self, light):
        idx = ctypes.c_uint32()
        check(self.lib.tr_scene_add_light_source(self.h, ctypes.byref(light), ctypes.byref(idx)))
        return idx.value

    def updateLightSource(self, index, light):
        check(self.lib.tr_scene_update_light_source(self.h, int(index), ctypes.byref(light)))

    def clearLightSources(self):
        check(self.lib.tr_scene_clear_light_sources(self.h))

    def addSphereLightSource(self, light):
        gid = ctypes.c_uint32()
        check(self.lib.tr_scene_add_sphere_light_source(self.h, ctypes.byref(light), ctypes.byref(gid)))
        light.geometry_id = gid.value
        return gid.value

    # geometry
    def addTriangleMesh(self, world, submeshes, tris, verts, two_sided=False):
        w = np.ascontiguousarray(np.asarray(world, dtype=np.float32).reshape(16))
        t = np.ascontiguousarray(np.asarray(tris, dtype=np.int32).reshape(-1))
        v = np.ascontiguousarray(verts)
        arr = submesh_array(submeshes)
        opts = _abi.tr_mesh_options(1 if two_sided else 0)
        ids = (ctypes.c_uint64 * max(1, len(submeshes)))()
        check(self.lib.tr_scene_add_triangle_mesh(
            self.h, _fptr(w), len(submeshes), arr, t.size, t.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
            v.nbytes, v.ctypes.data_as(ctypes.c_void_p), ctypes.byref(opts), ids))
        self._keep.append(submeshes)
        return [ids[i] for i in range(len(submeshes))]

    def addSphere(self, center, radius, material):
        c = np.asarray(center, dtype=np.float32)
        gid = ctypes.c_uint64()
        check(self.lib.tr_scene_add_sphere(self.h, _fptr(c), float(radius), ctypes.byref(material.rec),
                                           ctypes.byref(gid)))
        self._keep.append(material)
        return gid.value

    # user shapes (ray-tracing.cc:1315-1446)
    def _shape_call(self, fn, material, *args):
        gid = ctypes.c_uint64()
        if material is not None:
            self._keep.append(material)
        self._shape_status(fn(self.h, *args, ctypes.byref(gid)), fn.__name__)
        return gid.value

    def addRect(self, pos, tan, norm, size, material, tc=None, tangent_map=0):
        f = [np.ascontiguousarray(np.asarray(v, dtype=np.float32)) for v in (pos, tan, norm, size)]
        self._keep.extend(f)
        tcp = None
        if tc is not None:
            tcv = np.ascontiguousarray(np.asarray(tc, dtype=np.float32))
            self._keep.append(tcv)
            tcp = _fptr(tcv)
        return self._shape_call(self.lib.tr_scene_add_rect, material, *[_fptr(v) for v in f], ctypes.byref(material.rec), tcp,
                                int(tangent_map))

    def addBlocker(self, pos, tan, norm, size):
        f = [np.ascontiguousarray(np.asarray(v, dtype=np.float32)) for v in (pos, tan, norm, size)]
        return self._shape_call(self.lib.tr_scene_add_blocker, None, *[_fptr(v) for v in f])

    def addDisk(self, pos, norm, inner_radius, outer_radius, material):
        p, n = (np.ascontiguousarray(np.asarray(v, dtype=np.float32)) for v in (pos, norm))
        return self._shape_call(self.lib.tr_scene_add_disk, material, _fptr(p), _fptr(n), float(inner_radius),
                                float(outer_radius), ctypes.byref(material.rec))

    def addCylinder(self, center, radius, half_height, material):
        c = np.ascontiguousarray(np.asarray(center, dtype=np.float32))
        return self._shape_call(self.lib.tr_scene_add_cylinder, material, _fptr(c), float(radius), float(half_height),
                                ctypes.byref(material.rec))

    def addObliqueCylinder(self, center, radius, half_height, tilt_direction, tilt, material):
        c, td = (np.ascontiguousarray(np.asarray(v, dtype=np.float32)) for v in (center, tilt_direction))
        return self._shape_call(self.lib.tr_scene_add_oblique_cylinder, material, _fptr(c), float(radius), float(half_height),
                                _fptr(td), float(tilt), ctypes.byref(material.rec))

    def addEllipsoid(self, center, scale, orientation, material):
        c, sc, q = (np.ascontiguousarray(np.asarray(v, dtype=np.float32)) for v in (center, scale, orientation))
        return self._shape_call(self.lib.tr_scene_add_ellipsoid, material, _fptr(c), _fptr(sc), _fptr(q),
                                ctypes.byref(material.rec))

    @staticmethod
    def _shape_status(st, name):
        check(st)

    def bindTexture(self, texels, fmt, sampling=0):
        """bindTexture (ray-tracing.hh:340-354): texels [height, width(, c)]
        in the layout of `fmt` (TR_TEX_*); returns the handle for the
        material map fields."""
        desc, buf = texture_record(texels, fmt, sampling)
        h = ctypes.c_uint32()
        check(self.lib.tr_scene_bind_texture(self.h, ctypes.byref(desc), buf.ctypes.data_as(ctypes.c_void_p),
                                             ctypes.byref(h)))
        return h.value

    def setGlobalCubeMap(self, faces):
        """setGlobalCubeMap: six square RGBA8 faces (+x, -x, +y, -y, +z, -z)
        as uint32 [side, side] arrays, or None."""
        if faces is None:
            check(self.lib.tr_scene_set_global_cube_map(self.h, 0, None))
            return
        arrs, ptrs = cube_faces(faces)
        check(self.lib.tr_scene_set_global_cube_map(self.h, arrs[0].shape[0], ptrs))

    def addHierarchicalVolume(self, root, material):
        """addHierarchicalVolume (ray-tracing.cc:1448-1504) of a VolumeRoot;
        returns the first non-empty brick's geometry slot."""
        rec, keep = volume_root_record(root)
        gid = ctypes.c_uint64()
        check(self.lib.tr_scene_add_hierarchical_volume(self.h, ctypes.byref(rec), ctypes.byref(material.rec),
                                                        ctypes.byref(gid)))
        self._keep.append(material)
        return gid.value

    def addGridVolume(self, dims, min_corner, max_corner, density, direction, material, max_density=0.0):
        """A grid volume (tr_scene_add_grid_volume): `density` binary16
        [nz, ny, nx] (or flat, x fastest), `direction` binary16 [..., 3];
        float arrays are converted to binary16.  Returns the geometry slot."""
        rec, keep = grid_volume_record(dims, min_corner, max_corner, density, direction, max_density)
        gid = ctypes.c_uint64()
        check(self.lib.tr_scene_add_grid_volume(self.h, ctypes.byref(rec), ctypes.byref(material.rec),
                                                ctypes.byref(gid)))
        self._keep.append(material)
        return gid.value

    def commitScene(self):
        check(self.lib.tr_scene_commit(self.h))

    def buildHostOnly(self):
        """Builds the BVH on the host only (no device); for CPU-side checks."""
        check(self.lib.tr_debug_build_scene_host(self.h))

    def info(self):
        inf = _abi.tr_scene_info()
        check(self.lib.tr_scene_get_info(self.h, ctypes.byref(inf)))
        return {k: getattr(inf, k) for k, _ in inf._fields_}

    # rendering
    def drawOnce(self):
        check(self.lib.tr_scene_draw_once(self.h))

    def draw(self, width, height, view_proj_inv):
        vpi = np.ascontiguousarray(np.asarray(view_proj_inv, dtype=np.float32).reshape(16))
        check(self.lib.tr_scene_draw(self.h, int(width), int(height), _fptr(vpi)))
        self.width, self.height = int(width), int(height)

    def frameWait(self):
        check(self.lib.tr_scene_frame_wait(self.h))

    def backbuffer(self):
        n = self.width * self.height
        out = np.zeros(n, dtype=np.uint32)
        check(self.lib.tr_scene_read_backbuffer(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), n))
        return out.reshape(self.height, self.width)

    def saveImage(self, path):
        """RayTracingSystem::saveImage (ray-tracing-system.hh:125-130): the
        backbuffer as .tga (byte-exact) or .png."""
        check(self.lib.tr_scene_save_image(self.h, os.fsencode(path)))

    def accumulation(self):
        """(per-pixel RGB sums as float64 [H, W, 3], samples per pixel)."""
        n = self.width * self.height
        out = np.zeros(n * 3, dtype=np.float64)
        spp = ctypes.c_uint64()
        check(self.lib.tr_scene_read_accum(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n,
                                           ctypes.byref(spp)))
        return out.reshape(self.height, self.width, 3), spp.value

    def traceRays(self, rays):
        """rays: float32 [N, 8] = org xyz, tnear, dir xyz, tfar -> structured hits."""
        r = np.ascontiguousarray(np.asarray(rays, dtype=np.float32).reshape(-1, 8))
        n = r.shape[0]
        hits = np.zeros(n, dtype=HIT_DTYPE)
        check(self.lib.tr_scene_trace_rays(self.h, n, r.ctypes.data_as(ctypes.POINTER(_abi.tr_ray)),
                                           hits.ctypes.data_as(ctypes.POINTER(_abi.tr_hit))))
        return hits

    def rayQuery(self, tc):
        """rayQuery (ray-tracing.cc:1932-1979): (geometry id, tr_sample_data
        or None on a miss)."""
        t = np.asarray(tc, dtype=np.float32)
        sd = _abi.tr_sample_data()
        gid = ctypes.c_uint64()
        check(self.lib.tr_scene_ray_query(self.h, _fptr(t), ctypes.byref(sd), ctypes.byref(gid)))
        return gid.value, (sd if gid.value != 0xFFFFFFFF else None)

    def traceRaysNoAlloc(self, device_rays, device_hits, count, stream=None):
        """tr_scene_trace_rays_noalloc: rays and hits are device addresses
        (tr_ray / tr_hit records) the caller allocated; `stream` a
        hipStream_t handle (None: the scene's stream, and the call waits)."""
        check(self.lib.tr_scene_trace_rays_noalloc(self.h, int(count), ctypes.c_void_p(int(device_rays)),
                                                   ctypes.c_void_p(int(device_hits)),
                                                   ctypes.c_void_p(int(stream) if stream else 0)))

    def traceRaysHost(self, rays):
        """The same closest-hit query through the host copy of the BVH8."""
        r = np.ascontiguousarray(np.asarray(rays, dtype=np.float32).reshape(-1, 8))
        n = r.shape[0]
        hits = np.zeros(n, dtype=HIT_DTYPE)
        nv, npt = ctypes.c_uint64(), ctypes.c_uint64()
        check(self.lib.tr_debug_trace_rays_host(self.h, n, r.ctypes.data_as(ctypes.POINTER(_abi.tr_ray)),
                                                hits.ctypes.data_as(ctypes.POINTER(_abi.tr_hit)),
                                                ctypes.byref(nv), ctypes.byref(npt)))
        return hits, nv.value, npt.value

    def bvhViolations(self):
        v = ctypes.c_uint64()
        check(self.lib.tr_debug_bvh_check(self.h, ctypes.byref(v)))
        return v.value

    def renderPaths(self, pixels, samples_camera, frame=None):
        """Every camera sample of the listed pixels at the scene's frame index
        (or `frame`, which then becomes the frame index): radiance [N, 3] and
        hit-sequence hash per sample.  drawOnce advances the frame index."""
        if frame is not None:
            self.setFrameIndex(frame)
        px = np.ascontiguousarray(np.asarray(pixels, dtype=np.uint32))
        n = px.size * samples_camera
        rgb = np.zeros(n * 3, dtype=np.float32)
        hsh = np.zeros(n, dtype=np.uint32)
        check(self.lib.tr_scene_render_paths(self.h, px.size, px.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                             _fptr(rgb), hsh.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))))
        return rgb.reshape(n, 3), hsh

    def renderRequests(self, requests, use_graph=True, autotune=False, pool_slots=0, trace_waves_per_cu=0,
                       images=False):
        """Batched render requests (samples/virtual-gonioreflectometer.cc:161-189
        as one call).  `requests`: sequence of dicts with view_proj_inv (16),
        width, height, samples_camera, light_direction (3).  Returns (mean
        radiance [n, 3], per-request images [h, w, 3] or None, stats dict)."""
        n = len(requests)
        arr = (_abi.tr_request * n)()
        total_px = 0
        for i, q in enumerate(requests):
            r = arr[i]
            r.view_proj_inv[:] = [float(v) for v in np.asarray(q["view_proj_inv"], dtype=np.float32).reshape(16)]
            r.width, r.height, r.samples_camera = int(q["width"]), int(q["height"]), int(q["samples_camera"])
            r.light_direction[:] = [float(v) for v in q["light_direction"]]
            total_px += r.width * r.height
        opt = _abi.tr_request_options(int(bool(use_graph)), int(bool(autotune)), int(pool_slots),
                                      int(trace_waves_per_cu))
        mean = np.zeros(n * 3, dtype=np.float32)
        img = np.zeros(total_px * 3, dtype=np.float32) if images else None
        st = _abi.tr_request_stats()
        check(self.lib.tr_scene_render_requests(self.h, n, arr, ctypes.byref(opt), _fptr(mean),
                                                _fptr(img) if images else None, ctypes.byref(st)))
        out_imgs = None
        if images:
            out_imgs, o = [], 0
            for q in requests:
                w, h = int(q["width"]), int(q["height"])
                out_imgs.append(img[o * 3:(o + w * h) * 3].reshape(h, w, 3))
                o += w * h
        return mean.reshape(n, 3), out_imgs, {k: getattr(st, k) for k, _ in st._fields_}

    # radiance cache
    def setRadianceCache(self, enabled=True, train=True, train_paths=65536, batch=16384, rough_alpha=0.2, seed=1,
                         query_nee=False):
        c = _abi.tr_cache_settings(int(bool(enabled)), int(bool(train)), int(train_paths), int(batch),
                                   float(rough_alpha), int(bool(query_nee)), int(seed))
        check(self.lib.tr_scene_set_radiance_cache(self.h, ctypes.byref(c)))

    def radianceCache(self):
        """The scene's network as a RadianceCacheMLP (owned by the scene)."""
        h = ctypes.c_void_p()
        check(self.lib.tr_scene_get_radiance_cache(self.h, ctypes.byref(h)))
        net = RadianceCacheMLP.__new__(RadianceCacheMLP)
        net.lib, net.h, net._borrowed = self.lib, h, True
        return net

    def cacheRecords(self, count):
        """Training records of the last frame: inputs [n, 10], targets [n, 3], weights [n]."""
        x = np.zeros((count, 10), dtype=np.float32)
        t = np.zeros((count, 3), dtype=np.float32)
        v = np.zeros(count, dtype=np.float32)
        check(self.lib.tr_scene_read_cache_records(self.h, count, _fptr(x), _fptr(t), _fptr(v)))
        return x, t, v

    def stats(self):
        st = _abi.tr_stats()
        check(self.lib.tr_scene_get_stats(self.h, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_}


HIT_DTYPE = np.dtype([("t", np.float32), ("u", np.float32), ("v", np.float32), ("geom", np.uint32),
                      ("prim", np.uint32)])


class RadianceCacheMLP:
    """The radiance-cache network over tr_mlp_* (host arrays in and out)."""

    N_INPUTS, N_OUTPUTS, N_PARAMS = 10, 3, 28099

    def __init__(self, seed=1):
        self.lib = _abi.load()
        self.h = ctypes.c_void_p()
        check(self.lib.tr_mlp_create(int(seed), ctypes.byref(self.h)))

    def close(self):
        if self.h and not getattr(self, "_borrowed", False):
            self.lib.tr_mlp_destroy(self.h)
        self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self):
        p = np.zeros(self.N_PARAMS, dtype=np.float32)
        check(self.lib.tr_mlp_get_params(self.h, _fptr(p), p.size))
        return p

    @params.setter
    def params(self, value):
        p = np.ascontiguousarray(np.asarray(value, dtype=np.float32))
        check(self.lib.tr_mlp_set_params(self.h, _fptr(p), p.size))

    def infer(self, inputs):
        x = np.ascontiguousarray(np.asarray(inputs, dtype=np.float32).reshape(-1, self.N_INPUTS))
        y = np.zeros((x.shape[0], self.N_OUTPUTS), dtype=np.float32)
        check(self.lib.tr_mlp_infer(self.h, x.shape[0], _fptr(x), _fptr(y)))
        return y

    def _arrays(self, inputs, targets, valid):
        x = np.ascontiguousarray(np.asarray(inputs, dtype=np.float32).reshape(-1, self.N_INPUTS))
        t = np.ascontiguousarray(np.asarray(targets, dtype=np.float32).reshape(-1, self.N_OUTPUTS))
        v = None if valid is None else np.ascontiguousarray(np.asarray(valid, dtype=np.float32).reshape(-1))
        return x, t, v

    def gradient(self, inputs, targets, valid=None):
        x, t, v = self._arrays(inputs, targets, valid)
        g = np.zeros(self.N_PARAMS, dtype=np.float32)
        loss = ctypes.c_float()
        check(self.lib.tr_mlp_gradient(self.h, x.shape[0], _fptr(x), _fptr(t), _fptr(v) if v is not None else None,
                                       _fptr(g), ctypes.byref(loss)))
        return g, loss.value

    def train_step(self, inputs, targets, valid=None):
        x, t, v = self._arrays(inputs, targets, valid)
        loss = ctypes.c_float()
        check(self.lib.tr_mlp_train_step(self.h, x.shape[0], _fptr(x), _fptr(t), _fptr(v) if v is not None else None,
                                         ctypes.byref(loss)))
        return loss.value


class RTMeshBlob:
    """RTMeshBlob (mesh/obj-loader.hh:47-59) as Python-owned copies:
    Submeshes (Submesh list whose materials are entries of Materials),
    VertexData (uint8 bytes), IndexData (int32), Materials (Material list).
    IgnoredMaps counts materials whose map_Kd was not loaded."""

    def __init__(self):
        self.Submeshes, self.Materials = [], []
        self.VertexData = np.zeros(0, dtype=np.uint8)
        self.IndexData = np.zeros(0, dtype=np.int32)
        self.Stride = 0
        self.IgnoredMaps = 0
        self.Textures = []   # TextureData: (texels, TR_TEX_* format); materials' diffuse_map = index + 1

    def bindTextures(self, scene):
        """Binds the textures to `scene` (RayTracerScene or the oracle's) and
        rewrites the materials' diffuse_map to its handles; once per blob."""
        handles = [0] + [scene.bindTexture(t, fmt, 0) for t, fmt in self.Textures]
        for m in self.Materials:
            if 0 < m.rec.diffuse_map <= len(self.Textures):
                m.rec.diffuse_map = handles[m.rec.diffuse_map]

    @property
    def SubmeshCount(self):
        return len(self.Submeshes)

    @property
    def MaterialCount(self):
        return len(self.Materials)


def _image_from_raw(im):
    if im.format == _abi.TR_TEX_R8_UNORM:
        a = np.ctypeslib.as_array(im.data, shape=(im.height, im.width)).copy()
    else:
        a = np.ctypeslib.as_array(im.data, shape=(im.height * im.width * 4,)).copy().view(np.uint32)
        a = a.reshape(im.height, im.width)
    return a, int(im.format)


def load_image(path):
    """LoadImage (src/image/image.cc:36-58): (texels, TR_TEX_* format);
    RGBA8 as uint32 [h, w] (R in the low byte), R8 as uint8 [h, w]."""
    lib = _abi.load()
    im = _abi.tr_image()
    check(lib.tr_image_load(os.fsencode(path), ctypes.byref(im)))
    try:
        return _image_from_raw(im)
    finally:
        lib.tr_image_free(ctypes.byref(im))


def LoadObjFileStaticRTGeometry(filename, flags=0):
    """LoadObjFileStaticRTGeometry (src/mesh/obj-loader.cc:635-840): the
    OBJ file and its MTL libraries as an RTMeshBlob; raises TempestError
    with file and line on any diagnostic.  Pass the result to
    RayTracerScene.addTriangleMesh(world, blob.Submeshes, blob.IndexData,
    blob.VertexData)."""
    lib = _abi.load()
    raw = _abi.tr_mesh_blob()
    check(lib.tr_load_obj_rt_geometry(os.fsencode(filename), int(flags), ctypes.byref(raw)))
    try:
        out = RTMeshBlob()
        for i in range(raw.material_count):
            m = Material.__new__(Material)
            m.rec = _abi.tr_material()
            ctypes.pointer(m.rec)[0] = raw.materials[i]
            out.Materials.append(m)
        base = ctypes.addressof(raw.materials.contents) if raw.material_count else 0
        size = ctypes.sizeof(_abi.tr_material)
        for i in range(raw.submesh_count):
            sm = raw.submeshes[i]
            mat = None
            if sm.material:
                mat = out.Materials[(ctypes.cast(sm.material, ctypes.c_void_p).value - base) // size]
            out.Submeshes.append(Submesh(mat, sm.vertex_count, sm.vertex_offset, sm.base_index, sm.stride))
        if raw.vertex_bytes:
            out.VertexData = np.ctypeslib.as_array(raw.vertex_data, shape=(raw.vertex_bytes,)).copy()
        if raw.index_count:
            out.IndexData = np.ctypeslib.as_array(raw.index_data, shape=(raw.index_count,)).copy()
        out.Stride = raw.stride
        out.IgnoredMaps = raw.ignored_maps
        for i in range(raw.texture_count):
            out.Textures.append(_image_from_raw(raw.textures[i]))
        return out
    finally:
        lib.tr_mesh_blob_free(ctypes.byref(raw))
