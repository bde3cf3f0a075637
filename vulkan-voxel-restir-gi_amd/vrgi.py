"""ctypes harness over libvrgi.so (include/vrgi.h) for tests and benchmarks.

The product is the C ABI; this module only loads it, converts numpy arrays to plain
pointers and raises on a non-zero status. It never falls back to anything: when
libvrgi.so is missing it raises ImportError.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# VRGI_LIB: another build of the same sources (the sanitizer test's host-ASan build)
LIB_PATH = os.environ.get("VRGI_LIB") or os.path.join(HERE, "libvrgi.so")

VERTEX_DTYPE = np.dtype([("position", "<f4", 3), ("t", "<u4"), ("uv", "<f4", 2)])
NODE_DTYPE = np.dtype([("left", "<u4"), ("right", "<u4"), ("min", "<f4", 3), ("power", "<f4"),
                       ("max", "<f4", 3), ("parent", "<u4")])
MATERIAL_DTYPE = np.dtype([("reflectivity", "<f4", 4), ("emissivity", "<f4", 4), ("metallicity", "<f4", 4)])
assert VERTEX_DTYPE.itemsize == 24 and NODE_DTYPE.itemsize == 40 and MATERIAL_DTYPE.itemsize == 48

(BUF_GEOM_BVH, BUF_GEOM_PRIM_ORDER, BUF_LIGHT_TL_BVH, BUF_LIGHT_GEOM_BVH, BUF_PATH_LEN,
 BUF_BOUNCE_ORIGIN, BUF_BOUNCE_DIRECTION, BUF_BOUNCE_NORMAL, BUF_BOUNCE_EMISSIVITY,
 BUF_BOUNCE_REFLECTIVITY, BUF_BOUNCE_MIS_WEIGHT, BUF_BOUNCE_BSDF_PDF, BUF_BOUNCE_NEE_PDF,
 BUF_RADIANCE, BUF_RESTIR, BUF_TEMPORAL_RESERVOIR, BUF_SPATIAL_RESERVOIR, BUF_BGRA8,
 BUF_HIT_PRIM, BUF_HIT_T, BUF_DI_LIGHT, BUF_DI_WEIGHTS, BUF_BOUNCE_ALBEDO, BUF_CACHE_RADIANCE,
 BUF_CACHE_QUERIES, BUF_CACHE_TRAIN_RECORDS, BUF_CACHE_FALLBACK) = range(27)
CACHE_FREEZE = 1
CACHE_DYNAMIC_SCALE = 2
CACHE_BF16, CACHE_FP16 = 0, 1
ABI_VERSION = 2
CACHE_WEIGHTS = 4 * 64 * 64 + 64 * 16

VRGI_OK, VRGI_E_INVALID, VRGI_E_DEVICE, VRGI_E_STATE, VRGI_E_LIMIT = 0, -1, -2, -3, -4
FLAG_REBUILD_EVERY_FRAME = 1
FLAG_KEEP_BOUNCES = 2
FLAG_SERIAL_NEE = 4
FLAG_SERIAL_BUILD = 8
FLAG_FULL_NEE_PDF = 16


class Config(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("num_bounces", C.c_uint32),
                ("rr_start_bounce", C.c_uint32), ("rr_min_survival", C.c_float),
                ("restir_mode", C.c_uint32), ("restir_spatial_radius", C.c_float),
                ("di_candidates", C.c_uint32), ("seed", C.c_uint32), ("flags", C.c_uint32),
                ("ray_sort_bits", C.c_uint32)]


class ProfileOut(C.Structure):
    _fields_ = [("frames", C.c_uint64), ("build_ms", C.c_double), ("trace_ms", C.c_double * 16),
                ("nee_ms", C.c_double * 16), ("resolve_ms", C.c_double), ("frame_ms", C.c_double),
                ("rays", C.c_uint64 * 16), ("touched_bytes", C.c_uint64 * 16), ("build_prims", C.c_uint64),
                ("build_leaf_depth_sum", C.c_uint64), ("build_nodes", C.c_uint64),
                ("nee_listed", C.c_uint64 * 16)]


PROFILE_STAGES, PROFILE_TOUCH = 1, 2


class CacheConfig(C.Structure):
    _fields_ = [("log2_table", C.c_uint32), ("res_min", C.c_uint32), ("res_max", C.c_uint32),
                ("box_lo", C.c_float * 3), ("box_hi", C.c_float * 3), ("train_fraction", C.c_float),
                ("train_bounces", C.c_uint32), ("batch", C.c_uint32), ("steps_per_frame", C.c_uint32),
                ("lr", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float),
                ("loss_scale", C.c_float), ("seed", C.c_uint32), ("flags", C.c_uint32),
                ("query_vertex", C.c_uint32), ("fallback_bound", C.c_float), ("fallback_tile", C.c_uint32),
                ("fallback_ema", C.c_float), ("precision", C.c_uint32)]


class CacheInfo(C.Structure):
    _fields_ = [("levels", C.c_uint32), ("feats", C.c_uint32), ("log2_table", C.c_uint32),
                ("res", C.c_uint32 * 16), ("box_lo", C.c_float * 3), ("box_hi", C.c_float * 3),
                ("table_floats", C.c_uint64), ("steps", C.c_uint64), ("last_loss", C.c_float),
                ("loss_scale", C.c_float), ("precision", C.c_uint32), ("skipped_steps", C.c_uint64)]


class Camera(C.Structure):
    _fields_ = [("eye", C.c_float * 3), ("front", C.c_float * 3), ("right", C.c_float * 3),
                ("up", C.c_float * 3)]


class Prefs(C.Structure):
    _fields_ = [("nee_type", C.c_uint32), ("debug_view", C.c_uint32),
                ("restir_spatial_iterations", C.c_uint32)]


class FrameOut(C.Structure):
    _fields_ = [("radiance", C.c_void_p), ("restir", C.c_void_p), ("bgra8", C.c_void_p),
                ("frame_index", C.c_uint64)]


class FrameStats(C.Structure):
    _fields_ = [("frames", C.c_uint64), ("rays_closest", C.c_uint64), ("rays_light", C.c_uint64),
                ("bvh_nodes", C.c_uint64), ("bvh_depth", C.c_uint32), ("pad", C.c_uint32)]


_lib = None


def lib():
    """Loads libvrgi.so once; raises ImportError when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        L.vrgi_last_error.restype = C.c_char_p
        L.vrgi_abi_version.restype = C.c_uint32
        _lib = L
    return _lib


class VrgiError(RuntimeError):
    pass


def check(code):
    if code != 0:
        raise VrgiError(f"vrgi status {code}: {lib().vrgi_last_error().decode()}")


def ptr(a):
    """Pointer argument that keeps `a` alive until the call returns (data_as holds a
    reference to the array, so temporaries built in the call expression stay valid)."""
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


# ------------------------------------------------------------------ caller-side scenes
def world(kind, seed=0):
    """Seeded voxel world of a benchmark workload: (blocks u16 [dx,dy,dz], n_types)."""
    L = lib()
    dims = (C.c_uint32 * 3)()
    nt = C.c_uint32()
    check(L.vrgi_world_dims(kind, dims, C.byref(nt)))
    blocks = np.zeros(tuple(dims), np.uint16)
    check(L.vrgi_world_generate(kind, C.c_uint64(seed), ptr(blocks)))
    return blocks, nt.value


def world_materials(kind, seed=0):
    L = lib()
    ns = C.c_uint32()
    check(L.vrgi_world_materials(kind, C.c_uint64(seed), None, None, 0, C.byref(ns)))
    mats = np.zeros(ns.value, MATERIAL_DTYPE)
    lum = np.zeros(ns.value, np.float32)
    check(L.vrgi_world_materials(kind, C.c_uint64(seed), ptr(mats), ptr(lum), ns.value, C.byref(ns)))
    return mats, lum


def world_translucent(kind, n_types):
    tr = np.zeros(n_types, np.uint8)
    check(lib().vrgi_world_translucent(kind, ptr(tr), n_types))
    return tr


def gen_mesh_chunk(blocks, n_types, translucent, cx, cy, cz):
    """chunk::gen_mesh of the 32^3 chunk (cx, cy, cz): structured vertex array."""
    dims = np.array(blocks.shape, np.uint32)
    blocks = np.ascontiguousarray(blocks, np.uint16)
    n = C.c_size_t()
    check_code = lib().vrgi_gen_mesh_chunk(ptr(blocks), ptr(dims), n_types, ptr(translucent), cx, cy, cz,
                                           None, C.c_size_t(0), C.byref(n))
    if check_code not in (0, -4):
        check(check_code)
    out = np.zeros(n.value, VERTEX_DTYPE)
    check(lib().vrgi_gen_mesh_chunk(ptr(blocks), ptr(dims), n_types, ptr(translucent), cx, cy, cz,
                                    ptr(out), C.c_size_t(n.value), C.byref(n)))
    return out


def world_walk(blocks, n_types, seed, walkers, n_steps):
    """Random-walk toggles of the dynamic workload: updates `blocks` and `walkers`
    (uint32 [n, 3]) in place; returns the distinct touched voxels and their final ids."""
    assert blocks.flags.c_contiguous and blocks.dtype == np.uint16
    assert walkers.flags.c_contiguous and walkers.dtype == np.uint32
    dims = np.array(blocks.shape, np.uint32)
    xyz = np.zeros((n_steps, 3), np.uint32)
    vals = np.zeros(n_steps, np.uint16)
    n = C.c_uint64()
    check(lib().vrgi_world_walk(ptr(blocks), ptr(dims), n_types, C.c_uint64(seed), len(walkers), ptr(walkers),
                                C.c_uint64(n_steps), ptr(xyz), ptr(vals), C.byref(n)))
    return xyz[:n.value], vals[:n.value]

def chunk_objects(blocks, n_types, translucent):
    """Meshes every 32^3 chunk: list of (key, transform[12], vertices); empty chunks
    are kept (they become Empty handles, as in the reference's chunk manager)."""
    nx, ny, nz = (s // 32 for s in blocks.shape)
    objs = []
    for cx in range(nx):
        for cy in range(ny):
            for cz in range(nz):
                v = gen_mesh_chunk(blocks, n_types, translucent, cx, cy, cz)
                xf = np.array([1, 0, 0, 0, 1, 0, 0, 0, 1, 32 * cx, 32 * cy, 32 * cz], np.float32)
                objs.append(((cx * ny + cy) * nz + cz, xf, v))
    return objs


def camera_look(eye, target):
    """Camera vectors as the reference's SphericalCamera builds them (camera.rs:19-33):
    world up (0, -1, 0), right = front x worldup, up = right x front."""
    eye = np.asarray(eye, np.float64)
    f = np.asarray(target, np.float64) - eye
    f /= np.linalg.norm(f)
    r = np.cross(f, [0.0, -1.0, 0.0])
    r /= np.linalg.norm(r)
    u = np.cross(r, f)
    u /= np.linalg.norm(u)
    cam = Camera()
    cam.eye[:] = eye.astype(np.float32)
    cam.front[:] = f.astype(np.float32)
    cam.right[:] = r.astype(np.float32)
    cam.up[:] = u.astype(np.float32)
    return cam


def camera_spline(kind, frame, n_frames=16, seed=0):
    """Seeded Catmull-Rom camera path over the world of `kind` (frame in [0, n_frames))."""
    rng = np.random.default_rng(1000 + seed)
    if kind == 0:
        lo, hi, ylo, yhi, span = 8.0, 56.0, 60.0, 72.0, 64.0
    else:
        lo, hi, ylo, yhi, span = 32.0, 224.0, 128.0, 150.0, 256.0
    pts = np.stack([rng.uniform(lo, hi, 6), rng.uniform(ylo, yhi, 6), rng.uniform(lo, hi, 6)], 1)

    def cr(t):
        seg = min(int(t * 3), 2)
        u = t * 3 - seg
        p0, p1, p2, p3 = pts[seg], pts[seg + 1], pts[seg + 2], pts[seg + 3]
        return 0.5 * ((2 * p1) + (-p0 + p2) * u + (2 * p0 - 5 * p1 + 4 * p2 - p3) * u * u
                      + (-p0 + 3 * p1 - 3 * p2 + p3) * u ** 3)

    t = frame / max(n_frames - 1, 1) * 0.98
    eye = cr(t)
    ahead = cr(min(t + 0.02, 1.0))
    d = ahead - eye
    d[1] = 0.0
    if np.linalg.norm(d) < 1e-6:
        d = np.array([1.0, 0.0, 0.0])
    d /= np.linalg.norm(d)
    # look at the world's interior, swinging with the path's heading
    centre = np.array([span * 0.5, 0.15 * span if kind == 1 else 16.0, span * 0.5])
    target = centre + d * span * 0.15
    return camera_look(eye, target)


# ------------------------------------------------------------------ library objects
class Context:
    """One vrgi_ctx: Scene + SceneUploader + Renderer of the reference in one handle."""

    def __init__(self, cfg, mats, lum, device=0):
        self.L = lib()
        self.h = C.c_void_p()
        self.cfg = cfg
        check(self.L.vrgi_create(C.byref(cfg), ptr(mats), ptr(lum), len(lum), device, C.byref(self.h)))
        self.objects = {}

    def close(self):
        if self.h:
            for o in self.objects.values():
                if o:
                    self.L.vrgi_object_release(o)
            self.objects = {}
            check(self.L.vrgi_destroy(self.h))
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def upload(self, verts):
        o = C.c_void_p()
        v = np.ascontiguousarray(verts, VERTEX_DTYPE)
        check(self.L.vrgi_upload_object(self.h, ptr(v), C.c_size_t(len(v)), C.byref(o)))
        return o

    def add(self, key, obj, xf):
        check(self.L.vrgi_scene_add(self.h, C.c_uint32(key), obj, ptr(np.ascontiguousarray(xf, np.float32))))

    def add_objects(self, objs):
        for key, xf, v in objs:
            o = self.upload(v)
            self.add(key, o, xf)
            if o:
                self.L.vrgi_object_release(o)  # the scene holds its own reference

    def update(self, key, xf):
        check(self.L.vrgi_scene_update(self.h, C.c_uint32(key), ptr(np.ascontiguousarray(xf, np.float32))))

    def remove(self, key):
        check(self.L.vrgi_scene_remove(self.h, C.c_uint32(key)))

    def set_textures(self, rgba8):
        """rgba8: [slots][3][h][w][4] uint8 sRGB atlas (None: constant materials)."""
        if rgba8 is None:
            check(self.L.vrgi_set_textures(self.h, 0, 0, 0, None))
            return
        a = np.ascontiguousarray(rgba8, np.uint8)
        assert a.ndim == 5 and a.shape[1] == 3 and a.shape[4] == 4
        check(self.L.vrgi_set_textures(self.h, C.c_uint32(a.shape[0]), C.c_uint32(a.shape[3]),
                                       C.c_uint32(a.shape[2]), ptr(a)))

    def render(self, cam, nee_type=0, debug_view=0, spatial=0):
        p = Prefs(nee_type, debug_view, spatial)
        out = FrameOut()
        check(self.L.vrgi_render(self.h, C.byref(cam), C.byref(p), C.byref(out)))
        return out

    def sync(self):
        check(self.L.vrgi_sync(self.h))

    def stats(self, reset=False):
        st = FrameStats()
        check(self.L.vrgi_stats(self.h, C.byref(st), int(reset)))
        return st

    # ---- device-resident voxel world
    def voxels_load(self, blocks, n_types, translucent):
        blocks = np.ascontiguousarray(blocks, np.uint16)
        dims = np.array(blocks.shape, np.uint32)
        check(self.L.vrgi_voxels_load(self.h, ptr(blocks), ptr(dims), n_types,
                                      ptr(np.ascontiguousarray(translucent, np.uint8))))
        self.vx_shape = blocks.shape

    def voxels_set(self, xyz, vals):
        xyz = np.ascontiguousarray(xyz, np.uint32).reshape(-1, 3)
        vals = np.ascontiguousarray(vals, np.uint16)
        check(self.L.vrgi_voxels_set(self.h, C.c_uint64(len(vals)), ptr(xyz), ptr(vals)))

    def voxels_commit(self):
        check(self.L.vrgi_voxels_commit(self.h))

    def voxels_read(self):
        out = np.zeros(self.vx_shape, np.uint16)
        commits, remeshed = C.c_uint64(), C.c_uint64()
        check(self.L.vrgi_voxels_read(self.h, ptr(out), C.byref(commits), C.byref(remeshed)))
        return out, commits.value, remeshed.value

    def voxels_counts(self):
        """(commits, chunks re-meshed) so far, without reading the grid back."""
        commits, remeshed = C.c_uint64(), C.c_uint64()
        check(self.L.vrgi_voxels_read(self.h, None, C.byref(commits), C.byref(remeshed)))
        return commits.value, remeshed.value

    # ---- GPU mesher
    def mesh_chunks(self, blocks, n_types, translucent, chunks):
        """chunk::gen_mesh on the GPU for chunks [(cx, cy, cz), ...]: list of vertex arrays."""
        blocks = np.ascontiguousarray(blocks, np.uint16)
        dims = np.array(blocks.shape, np.uint32)
        ch = np.ascontiguousarray(chunks, np.uint32).reshape(-1, 3)
        counts = np.zeros(len(ch), np.uint64)
        tr = np.ascontiguousarray(translucent, np.uint8)
        code = self.L.vrgi_mesh_chunks(self.h, ptr(blocks), ptr(dims), n_types, ptr(tr), len(ch), ptr(ch), None,
                                       C.c_uint64(0), ptr(counts))
        if code not in (0, -4):
            check(code)
        out = np.zeros(int(counts.sum()), VERTEX_DTYPE)
        check(self.L.vrgi_mesh_chunks(self.h, ptr(blocks), ptr(dims), n_types, ptr(tr), len(ch), ptr(ch), ptr(out),
                                      C.c_uint64(len(out)), ptr(counts)))
        res, o = [], 0
        for k in counts:
            res.append(out[o:o + int(k)])
            o += int(k)
        return res

    # ---- measurement
    def profile(self, flags):
        check(self.L.vrgi_profile(self.h, flags))

    def profile_read(self, reset=True):
        p = ProfileOut()
        check(self.L.vrgi_profile_read(self.h, C.byref(p), int(reset)))
        return p

    def frame_set(self, frame):
        check(self.L.vrgi_frame_set(self.h, C.c_uint64(frame)))

    # ---- FP64 accumulation of the radiance (convergence studies)
    def accum_enable(self, on=True):
        check(self.L.vrgi_accum_enable(self.h, int(on)))

    def accum_read(self):
        P = self.cfg.width * self.cfg.height
        out = np.zeros((P, 3), np.float64)
        n = C.c_uint64()
        check(self.L.vrgi_accum_read(self.h, ptr(out), C.byref(n)))
        return out, n.value

    # ---- learned radiance cache
    def cache_enable(self, **kw):
        cfg = CacheConfig()
        self.L.vrgi_cache_config_default(C.byref(cfg))
        for k, v in kw.items():
            if k in ("box_lo", "box_hi"):
                setattr(cfg, k, (C.c_float * 3)(*v))
            else:
                setattr(cfg, k, v)
        check(self.L.vrgi_cache_enable(self.h, C.byref(cfg)))
        return cfg

    def cache_disable(self):
        check(self.L.vrgi_cache_disable(self.h))

    def cache_info(self):
        info = CacheInfo()
        check(self.L.vrgi_cache_info_get(self.h, C.byref(info)))
        return info

    def cache_get(self):
        info = self.cache_info()
        w = np.zeros(CACHE_WEIGHTS, np.float32)
        t = np.zeros(info.table_floats, np.float32)
        check(self.L.vrgi_cache_get(self.h, ptr(w), ptr(t)))
        return w, t

    def cache_set(self, weights=None, table=None):
        w = None if weights is None else np.ascontiguousarray(weights, np.float32)
        t = None if table is None else np.ascontiguousarray(table, np.float32)
        check(self.L.vrgi_cache_set(self.h, None if w is None else ptr(w), None if t is None else ptr(t)))

    @staticmethod
    def _samples(pos, dir, nrm, alb):
        return np.ascontiguousarray(np.concatenate([np.asarray(a, np.float32).reshape(-1, 3)
                                                    for a in (pos, dir, nrm, alb)], axis=1))

    def cache_infer(self, pos, dir, nrm, alb):
        sm = self._samples(pos, dir, nrm, alb)
        out = np.zeros((len(sm), 3), np.float32)
        check(self.L.vrgi_cache_infer(self.h, C.c_uint64(len(sm)), ptr(sm), ptr(out)))
        return out

    def cache_grads(self, pos, dir, nrm, alb, target):
        sm = self._samples(pos, dir, nrm, alb)
        t = np.ascontiguousarray(target, np.float32).reshape(-1, 3)
        gw = np.zeros(CACHE_WEIGHTS, np.float32)
        gt = np.zeros(self.cache_info().table_floats, np.float32)
        loss = C.c_float()
        check(self.L.vrgi_cache_grads(self.h, C.c_uint64(len(sm)), ptr(sm), ptr(t), ptr(gw), ptr(gt), C.byref(loss)))
        return loss.value, gw, gt

    def cache_train(self, pos, dir, nrm, alb, target):
        sm = self._samples(pos, dir, nrm, alb)
        t = np.ascontiguousarray(target, np.float32).reshape(-1, 3)
        loss = C.c_float()
        check(self.L.vrgi_cache_train(self.h, C.c_uint64(len(sm)), ptr(sm), ptr(t), C.byref(loss)))
        return loss.value

    def read(self, which, dtype=np.float32):
        n = C.c_size_t()
        check(self.L.vrgi_buffer_size(self.h, which, C.byref(n)))
        if n.value == 0:
            return np.zeros(0, dtype)
        out = np.zeros(n.value // np.dtype(dtype).itemsize, dtype)
        check(self.L.vrgi_read_buffer(self.h, which, ptr(out), C.c_size_t(out.nbytes)))
        return out

    def trace_closest(self, o, d, tmin=1e-3, tmax=1000.0):
        o = np.ascontiguousarray(o, np.float32)
        d = np.ascontiguousarray(d, np.float32)
        n = len(o)
        hit = np.zeros(n, np.uint32)
        t = np.zeros(n, np.float32)
        uv = np.zeros((n, 2), np.float32)
        check(self.L.vrgi_trace_closest(self.h, ptr(o), ptr(d), C.c_uint64(n), C.c_float(tmin),
                                        C.c_float(tmax), ptr(hit), ptr(t), ptr(uv)))
        return hit, t, uv

    def build_bl_bvh(self, tri, lum, ids):
        tri = np.ascontiguousarray(tri, np.float32).reshape(-1, 9)
        n = len(tri)
        nodes = np.zeros(2 * n - 1, NODE_DTYPE)
        order = np.zeros(n, np.uint32)
        box = np.zeros(6, np.float32)
        pw = C.c_float()
        check(self.L.vrgi_build_bl_bvh(self.h, ptr(tri), ptr(np.ascontiguousarray(lum, np.float32)),
                                       ptr(np.ascontiguousarray(ids, np.uint32)), C.c_uint64(n),
                                       ptr(nodes), ptr(order), ptr(box), C.byref(pw)))
        return nodes, order, box, pw.value

    def build_tl_bvh(self, xf, aabbs, powers, ids):
        m = len(powers)
        nodes = np.zeros(2 * m - 1, NODE_DTYPE)
        check(self.L.vrgi_build_tl_bvh(self.h, ptr(np.ascontiguousarray(xf, np.float32)),
                                       ptr(np.ascontiguousarray(aabbs, np.float32)),
                                       ptr(np.ascontiguousarray(powers, np.float32)),
                                       ptr(np.ascontiguousarray(ids, np.uint32)), C.c_uint64(m), ptr(nodes)))
        return nodes


# ------------------------------------------------------------------ batched scenes
class BatchConfig(C.Structure):
    _fields_ = [("spp", C.c_uint32), ("rays_per_block", C.c_uint32), ("pixels_per_block", C.c_uint32),
                ("flags", C.c_uint32)]


class BatchInfo(C.Structure):
    _fields_ = [("n_scenes", C.c_uint32), ("rays_per_block", C.c_uint32), ("pixels_per_block", C.c_uint32),
                ("tuned", C.c_uint32), ("graph", C.c_uint32), ("pad", C.c_uint32), ("triangles", C.c_uint64),
                ("light_triangles", C.c_uint64), ("pixels", C.c_uint64), ("build_ms", C.c_float),
                ("passes_ms", C.c_float), ("rays", C.c_uint64)]


BATCH_NO_GRAPH = 1


class Batch:
    """vrgi_batch over Context objects: one multi-tree build and one launch per stage for
    all scenes; each of the `spp` passes is one frame of every scene."""

    def __init__(self, contexts, spp=1, rays_per_block=0, pixels_per_block=0, flags=0):
        self.L = lib()
        self.contexts = list(contexts)
        cfg = BatchConfig(spp, rays_per_block, pixels_per_block, flags)
        hs = (C.c_void_p * len(self.contexts))(*[c.h for c in self.contexts])
        self.h = C.c_void_p()
        check(self.L.vrgi_batch_create(hs, len(self.contexts), C.byref(cfg), C.byref(self.h)))

    def close(self):
        if self.h:
            check(self.L.vrgi_batch_destroy(self.h))
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def render(self, cams, nee_type=0, debug_view=0, spatial=0):
        arr = (Camera * len(cams))(*cams)
        p = Prefs(nee_type, debug_view, spatial)
        check(self.L.vrgi_batch_render(self.h, arr, C.byref(p)))

    def sync(self):
        check(self.L.vrgi_batch_sync(self.h))

    def mean(self, k):
        c = self.contexts[k]
        out = np.zeros((c.cfg.width * c.cfg.height, 3), np.float32)
        check(self.L.vrgi_batch_read_mean(self.h, C.c_uint32(k), ptr(out)))
        return out

    def info(self):
        i = BatchInfo()
        check(self.L.vrgi_batch_info_get(self.h, C.byref(i)))
        return i


def batch_scene(i):
    """Scene i of the batched workload: (objects, mats, lum, (W, H), camera). The world is
    kind 100 + i (sides 32..160 voxels); the square resolution is 128, 256, 384 or 512
    pixels, drawn per scene; the camera looks from above one corner to the far side."""
    blocks, nt = world(100 + i, i)
    mats, lum = world_materials(100 + i, i)
    tr = world_translucent(100 + i, nt)
    objs = chunk_objects(blocks, nt, tr)
    h = (i * 2654435761 + 12345) & 0xFFFFFFFF
    res = 128 * (1 + (h >> 13) % 4)
    dx, dy, dz = blocks.shape
    cam = camera_look([0.1 * dx, 0.95 * dy, 0.1 * dz], [0.6 * dx, 0.2 * dy, 0.6 * dz])
    return objs, mats, lum, (res, res), cam

def config(width, height, bounces=3, rr_start=0xFFFFFFFF, rr_min=0.1, seed=0, flags=0, radius=15.0,
           mode=0, di_candidates=32, sort_bits=0):
    cfg = Config()
    lib().vrgi_config_reference(C.byref(cfg), width, height)
    cfg.restir_mode = mode
    cfg.di_candidates = di_candidates
    cfg.num_bounces = bounces
    cfg.rr_start_bounce = rr_start
    cfg.rr_min_survival = rr_min
    cfg.seed = seed
    cfg.flags = flags
    cfg.restir_spatial_radius = radius
    cfg.ray_sort_bits = sort_bits
    return cfg
