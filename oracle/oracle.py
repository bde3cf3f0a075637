"""ctypes harness over liboracle.so — the CPU restatement of the reference's hot path.

TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg load this module, and only as the checker / the CPU baseline.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# ORACLE_LIB: another build of the same sources (the sanitizer test's ASan/UBSan build)
LIB_PATH = os.environ.get("ORACLE_LIB") or os.path.join(HERE, "liboracle.so")

NODE_DTYPE = np.dtype([("left", "<u4"), ("right", "<u4"), ("min", "<f4", 3), ("power", "<f4"),
                       ("max", "<f4", 3), ("parent", "<u4")])
VERTEX_DTYPE = np.dtype([("position", "<f4", 3), ("t", "<u4"), ("uv", "<f4", 2)])


def build():
    """Compiles liboracle.so with gcc (no GPU code)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.or_murmur3_combine.restype = C.c_uint32
        L.or_murmur3_finalize.restype = C.c_uint32
        L.or_float_construct.restype = C.c_float
        L.or_scene_create.restype = C.c_void_p
        L.or_scene_count.restype = C.c_uint64
        L.or_renderer_create.restype = C.c_void_p
        L.or_reverse_traverse_light_bvh.restype = C.c_float
        L.or_nrc_create.restype = C.c_void_p
        L.or_nrc_hash.restype = C.c_uint32
        L.or_nrc_grads.restype = C.c_double
        L.or_nrc_train_step.restype = C.c_double
        L.or_restir_gi_jacobian.restype = C.c_float
        L.or_nrc_round16.restype = C.c_double
        L.or_srgb_decode.restype = C.c_float
        L.or_texel_index.restype = C.c_uint32
        L.or_texel_index.argtypes = [C.c_float, C.c_uint32]
        _lib = L
    return _lib


def ptr(a):
    """Pointer argument that keeps `a` alive until the call returns (data_as holds a
    reference to the array, so temporaries built in the call expression stay valid)."""
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


# planted mistakes (oracle.h OR_MUT_*): used only to show that a pinning check fails
MUTATIONS = {"nee_area": 1, "nee_dist": 2, "nee_bary": 3, "nee_cos": 4, "gi_ucw_no_m": 5, "gi_merge_no_m": 6,
             "gi_z_all": 7, "gi_final_cos": 8, "gi_jac_sq": 9, "gi_jac_swap": 10, "di_area": 11, "di_cosy": 12,
             "di_ucw_no_m": 13, "di_spatial_m": 14}


class mutation:
    """Context manager planting one mistake in the restatement for the duration of a
    block (always removed on exit)."""

    def __init__(self, name):
        self.m = MUTATIONS[name] if name else 0

    def __enter__(self):
        lib().or_set_mutation(self.m)
        return self

    def __exit__(self, *exc):
        lib().or_set_mutation(0)
        return False


def gi_jacobian(q_res, r_res):
    """computeJacobianQToR of two 20-word reservoirs (q's sample reused at r)."""
    q = np.ascontiguousarray(q_res, np.float32)
    r = np.ascontiguousarray(r_res, np.float32)
    assert q.size == 20 and r.size == 20
    return lib().or_restir_gi_jacobian(ptr(q), ptr(r))


def combine(h, k):
    return lib().or_murmur3_combine(C.c_uint32(h), C.c_uint32(k))


def finalize(h):
    return lib().or_murmur3_finalize(C.c_uint32(h))


def float_construct(m):
    return lib().or_float_construct(C.c_uint32(m))


def sincos(x):
    s, c = C.c_float(), C.c_float()
    lib().or_sincos(C.c_float(x), C.byref(s), C.byref(c))
    return s.value, c.value


def build_bl_bvh(tri, lum, ids):
    tri = np.ascontiguousarray(tri, np.float32).reshape(-1, 9)
    n = len(tri)
    nodes = np.zeros(2 * n - 1, NODE_DTYPE)
    order = np.zeros(n, np.uint32)
    box = np.zeros(6, np.float32)
    pw = C.c_float()
    depth = C.c_uint32()
    r = lib().or_build_bl_bvh(C.c_uint64(n), ptr(tri), ptr(np.ascontiguousarray(lum, np.float32)),
                              ptr(np.ascontiguousarray(ids, np.uint32)), ptr(nodes), ptr(order), ptr(box),
                              C.byref(pw), C.byref(depth))
    assert r == 0
    return nodes, order, box, pw.value, depth.value


def build_tl_bvh(xf, aabbs, powers, ids):
    m = len(powers)
    nodes = np.zeros(2 * m - 1, NODE_DTYPE)
    r = lib().or_build_tl_bvh(C.c_uint64(m), ptr(np.ascontiguousarray(xf, np.float32)),
                              ptr(np.ascontiguousarray(aabbs, np.float32)),
                              ptr(np.ascontiguousarray(powers, np.float32)),
                              ptr(np.ascontiguousarray(ids, np.uint32)), ptr(nodes))
    assert r == 0
    return nodes


def find_best_plane(tri):
    tri = np.ascontiguousarray(tri, np.float32).reshape(-1, 9)
    d, s = C.c_uint32(), C.c_float()
    lib().or_find_best_plane(C.c_uint64(len(tri)), ptr(tri), C.byref(d), C.byref(s))
    return d.value, s.value


def gen_mesh_chunk(blocks, n_types, translucent, cx, cy, cz):
    dims = np.array(blocks.shape, np.uint32)
    blocks = np.ascontiguousarray(blocks, np.uint16)
    n = C.c_uint64()
    lib().or_gen_mesh_chunk(ptr(blocks), ptr(dims), n_types, ptr(translucent), cx, cy, cz, None,
                            C.c_uint64(0), C.byref(n))
    out = np.zeros(n.value, VERTEX_DTYPE)
    r = lib().or_gen_mesh_chunk(ptr(blocks), ptr(dims), n_types, ptr(translucent), cx, cy, cz,
                                ptr(out), C.c_uint64(n.value), C.byref(n))
    assert r == 0
    return out


class Config(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("num_bounces", C.c_uint32),
                ("rr_start_bounce", C.c_uint32), ("rr_min_survival", C.c_float),
                ("restir_mode", C.c_uint32), ("restir_spatial_radius", C.c_float),
                ("di_candidates", C.c_uint32), ("seed", C.c_uint32), ("flags", C.c_uint32)]


class Prefs(C.Structure):
    _fields_ = [("nee_type", C.c_uint32), ("debug_view", C.c_uint32),
                ("restir_spatial_iterations", C.c_uint32)]


FRAME_FIELDS = ["origin", "direction", "normal", "emissivity", "reflectivity", "mis_weight", "bsdf_pdf",
                "nee_pdf", "l_out", "q_omega", "hit_prim", "hit_t", "path_len", "temporal", "spatial",
                "restir", "debug", "bgra8", "di_light", "di_weights"]


class CacheFrame(C.Structure):
    _fields_ = [("train_fraction", C.c_float), ("train_seed", C.c_uint32), ("query_vertex", C.c_uint32),
                ("fallback", C.c_void_p), ("cache_rad", C.c_void_p),
                ("albedo", C.c_void_p), ("qrec", C.c_void_p), ("n_qrec", C.c_uint64), ("trec", C.c_void_p),
                ("trec_cap", C.c_uint64), ("n_trec", C.c_uint64)]


def frame_seed(seed, frame, salt):
    """Counter-based per-frame seed: finalize(combine(combine(seed, salt), frame))."""
    L, u = lib(), C.c_uint32
    return L.or_murmur3_finalize(u(L.or_murmur3_combine(u(L.or_murmur3_combine(u(seed), u(salt))), u(frame))))


def srgb_decode(v):
    """Linear value of an 8-bit sRGB channel as the texture path reads it."""
    return float(lib().or_srgb_decode(C.c_uint32(v)))


def texel_index(c, n):
    """Texel of normalised coordinate c over n texels (nearest, clamp to edge)."""
    return int(lib().or_texel_index(c, n))


def texture_luminances(rgba8):
    """utils.rs:223-235 over an atlas [slots][3][h][w][4]."""
    a = np.ascontiguousarray(rgba8, np.uint8)
    out = np.zeros(a.shape[0], np.float32)
    lib().or_texture_luminances(C.c_uint32(a.shape[0]), C.c_uint32(a.shape[3]), C.c_uint32(a.shape[2]), ptr(a),
                                ptr(out))
    return out


class Frame(C.Structure):
    _fields_ = [(f, C.c_void_p) for f in FRAME_FIELDS]


class Scene:
    def __init__(self, mats, lum):
        self.L = lib()
        self.h = C.c_void_p(self.L.or_scene_create(ptr(mats), ptr(np.ascontiguousarray(lum, np.float32)),
                                                   len(lum)))

    def close(self):
        if self.h:
            self.L.or_scene_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, key, verts, xf):
        v = np.ascontiguousarray(verts, VERTEX_DTYPE)
        assert self.L.or_scene_add(self.h, key, ptr(v), C.c_uint64(len(v)),
                                   ptr(np.ascontiguousarray(xf, np.float32))) == 0

    def add_objects(self, objs):
        for key, xf, v in objs:
            self.add(key, v, xf)

    def remove(self, key):
        self.L.or_scene_remove(self.h, key)

    def set_textures(self, rgba8):
        """rgba8: [slots][3][h][w][4] uint8 (reflectivity, emissivity, metallicity images of
        every texture slot, sRGB); None: constant materials."""
        if rgba8 is None:
            assert self.L.or_scene_set_textures(self.h, 0, 0, 0, None) == 0
            return
        a = np.ascontiguousarray(rgba8, np.uint8)
        assert a.ndim == 5 and a.shape[1] == 3 and a.shape[4] == 4
        assert self.L.or_scene_set_textures(self.h, a.shape[0], a.shape[3], a.shape[2], ptr(a)) == 0

    def build(self):
        self.L.or_scene_build(self.h)

    def count(self, what):
        return int(self.L.or_scene_count(self.h, what))

    def read(self, what, dtype, n):
        out = np.zeros(n, dtype)
        assert self.L.or_scene_read(self.h, what, ptr(out), C.c_uint64(out.nbytes)) == 0
        return out

    def geom_bvh(self):
        return self.read(0, NODE_DTYPE, self.count(5))

    def geom_order(self):
        return self.read(1, np.uint32, self.count(0))

    def tl_bvh(self):
        return self.read(2, NODE_DTYPE, self.count(3))

    def light_geom_bvh(self):
        return self.read(3, NODE_DTYPE, self.count(4))

    def triangles(self):
        return self.read(4, np.float32, 9 * self.count(0)).reshape(-1, 9)

    def light_bvh(self, key):
        n = C.c_uint64()
        self.L.or_scene_light_bvh(self.h, key, None, C.c_uint64(0), C.byref(n))
        out = np.zeros(n.value, NODE_DTYPE)
        if n.value:
            self.L.or_scene_light_bvh(self.h, key, ptr(out), C.c_uint64(n.value), C.byref(n))
        return out

    def trace_closest(self, o, d, tmin=1e-3, tmax=1000.0, brute=False):
        o = np.ascontiguousarray(o, np.float32)
        d = np.ascontiguousarray(d, np.float32)
        n = len(o)
        hit = np.zeros(n, np.uint32)
        t = np.zeros(n, np.float32)
        uv = np.zeros((n, 2), np.float32)
        vis = C.c_uint64()
        self.L.or_trace_closest(self.h, C.c_uint64(n), ptr(o), ptr(d), C.c_float(tmin), C.c_float(tmax),
                                int(brute), ptr(hit), ptr(t), ptr(uv), C.byref(vis))
        return hit, t, uv

    def trace_lights(self, o, d, tmin=1e-4, tmax=1000.0, cap=64):
        hits = np.zeros(cap, np.uint32)
        uvs = np.zeros((cap, 2), np.float32)
        n = C.c_uint32()
        self.L.or_trace_lights(self.h, ptr(np.asarray(o, np.float32)), ptr(np.asarray(d, np.float32)),
                               C.c_float(tmin), C.c_float(tmax), ptr(hits), cap, C.byref(n), ptr(uvs))
        return hits[:n.value], uvs[:n.value]

    def nee_pdf(self, o, nrm, d):
        """Per ray: the sum of computeNeePdf over every light triangle it crosses."""
        o = np.ascontiguousarray(o, np.float32).reshape(-1, 3)
        nrm = np.ascontiguousarray(np.broadcast_to(np.asarray(nrm, np.float32), o.shape))
        d = np.ascontiguousarray(d, np.float32).reshape(-1, 3)
        out = np.zeros(len(o), np.float32)
        assert self.L.or_nee_pdf(self.h, C.c_uint64(len(o)), ptr(o), ptr(nrm), ptr(d), ptr(out)) == 0
        return out

    def traverse_light_bvh(self, p, n, seed):
        inst, prim = C.c_uint32(), C.c_uint32()
        prob, imp = C.c_float(), C.c_float()
        ok = self.L.or_traverse_light_bvh(self.h, ptr(np.asarray(p, np.float32)), ptr(np.asarray(n, np.float32)),
                                          C.c_uint32(seed), C.byref(inst), C.byref(prim), C.byref(prob),
                                          C.byref(imp))
        return bool(ok), inst.value, prim.value, prob.value, imp.value

    def reverse_traverse_light_bvh(self, p, n, inst, bl_node):
        return self.L.or_reverse_traverse_light_bvh(self.h, ptr(np.asarray(p, np.float32)),
                                                    ptr(np.asarray(n, np.float32)), inst, bl_node)


class Renderer:
    def __init__(self, width, height, bounces=3, rr_start=0xFFFFFFFF, rr_min=0.1, seed=0, radius=15.0,
                 mode=0, di_candidates=32):
        self.L = lib()
        self.cfg = Config(width, height, bounces, rr_start, rr_min, mode, radius, di_candidates, seed, 0)
        self.h = C.c_void_p(self.L.or_renderer_create(C.byref(self.cfg)))
        self.W, self.H, self.B = width, height, bounces

    def close(self):
        if self.h:
            self.L.or_renderer_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def render(self, scene, cam, frame, nee_type=0, debug_view=0, spatial=0, threads=0):
        P, B = self.W * self.H, self.B
        out = {
            "origin": np.zeros((B + 1, P, 3), np.float32), "direction": np.zeros((B + 1, P, 3), np.float32),
            "normal": np.zeros((B, P, 3), np.float32), "emissivity": np.zeros((B, P, 3), np.float32),
            "reflectivity": np.zeros((B, P, 3), np.float32), "mis_weight": np.zeros((B, P), np.float32),
            "bsdf_pdf": np.zeros((B, P), np.float32), "nee_pdf": np.zeros((B, P), np.float32),
            "l_out": np.zeros((B, P, 3), np.float32), "q_omega": np.zeros((B, P), np.float32),
            "hit_prim": np.zeros((B, P), np.uint32), "hit_t": np.zeros((B, P), np.float32),
            "path_len": np.zeros(P, np.uint32), "temporal": np.zeros((P, 20), np.float32),
            "spatial": np.zeros((P, 20), np.float32), "restir": np.zeros((P, 3), np.float32),
            "debug": np.zeros((P, 3), np.float32), "bgra8": np.zeros((P, 4), np.uint8),
            "di_light": np.zeros(P, np.uint32), "di_weights": np.zeros((P, 4), np.float32),
        }
        fr = Frame(**{k: C.c_void_p(v.ctypes.data) for k, v in out.items()})
        p = Prefs(nee_type, debug_view, spatial)
        r = self.L.or_render(self.h, scene.h, C.byref(cam), C.byref(p), C.c_uint64(frame), C.byref(fr), threads)
        assert r == 0
        return out

    def render_cache(self, scene, cam, frame, cache_rad, train_fraction, train_seed, nee_type=0, spatial=0,
                     threads=0, trec_cap=1 << 20, query_vertex=0, fallback=None):
        """A learned-radiance-cache frame (see or_cache_frame): cache_rad [P][3] is the
        cache's reflected radiance at the bounce-1 vertex of every queried pixel."""
        P, B = self.W * self.H, self.B
        out = {
            "origin": np.zeros((B + 1, P, 3), np.float32), "direction": np.zeros((B + 1, P, 3), np.float32),
            "normal": np.zeros((B, P, 3), np.float32), "emissivity": np.zeros((B, P, 3), np.float32),
            "reflectivity": np.zeros((B, P, 3), np.float32), "mis_weight": np.zeros((B, P), np.float32),
            "bsdf_pdf": np.zeros((B, P), np.float32), "nee_pdf": np.zeros((B, P), np.float32),
            "l_out": np.zeros((B, P, 3), np.float32), "q_omega": np.zeros((B, P), np.float32),
            "hit_prim": np.zeros((B, P), np.uint32), "hit_t": np.zeros((B, P), np.float32),
            "path_len": np.zeros(P, np.uint32), "temporal": np.zeros((P, 20), np.float32),
            "spatial": np.zeros((P, 20), np.float32), "restir": np.zeros((P, 3), np.float32),
            "debug": np.zeros((P, 3), np.float32), "bgra8": np.zeros((P, 4), np.uint8),
            "di_light": np.zeros(P, np.uint32), "di_weights": np.zeros((P, 4), np.float32),
        }
        fr = Frame(**{k: C.c_void_p(v.ctypes.data) for k, v in out.items()})
        crad = np.ascontiguousarray(cache_rad, np.float32).reshape(P, 3)
        alb = np.zeros((B, P, 3), np.float32)
        qrec = np.zeros((P, 16), np.float32)
        trec = np.zeros((trec_cap, 16), np.float32)
        fb = None if fallback is None else np.ascontiguousarray(fallback, np.uint8).reshape(P)
        cf = CacheFrame(train_fraction, train_seed, query_vertex, None if fb is None else fb.ctypes.data, crad.ctypes.data, alb.ctypes.data, qrec.ctypes.data, 0,
                        trec.ctypes.data, trec_cap, 0)
        p = Prefs(nee_type, 0, spatial)
        r = self.L.or_render_cache(self.h, scene.h, C.byref(cam), C.byref(p), C.c_uint64(frame), C.byref(fr),
                                   threads, C.byref(cf))
        assert r == 0
        out["albedo"] = alb
        out["qrec"] = qrec[:cf.n_qrec].copy()
        out["trec"] = trec[:min(cf.n_trec, trec_cap)].copy()
        out["n_trec"] = cf.n_trec
        return out

    def render_rows(self, scene, cam, frame, y0, y1, nee_type=0, threads=0):
        rad = np.zeros(((y1 - y0) * self.W, 3), np.float32)
        r = self.L.or_render_rows(self.h, scene.h, C.byref(cam), C.byref(Prefs(nee_type, 0, 0)),
                                  C.c_uint64(frame), y0, y1, ptr(rad), threads)
        assert r == 0
        return rad


# ---------------------------------------------------------------- learned radiance cache
N_IN, N_WIDTH, N_HIDDEN, N_OUT_PAD = 64, 64, 4, 16
N_WEIGHTS = N_HIDDEN * N_IN * N_WIDTH + N_WIDTH * N_OUT_PAD


class NrcConfig(C.Structure):
    _fields_ = [("levels", C.c_uint32), ("feats", C.c_uint32), ("log2_T", C.c_uint32),
                ("res", C.c_uint32 * 16), ("box_lo", C.c_float * 3), ("box_hi", C.c_float * 3),
                ("lr", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float),
                ("loss_scale", C.c_float), ("loss", C.c_uint32), ("emulate_bf16", C.c_uint32)]


def nrc_hash(x, y, z, log2_T):
    return lib().or_nrc_hash(C.c_uint32(x), C.c_uint32(y), C.c_uint32(z), C.c_uint32(log2_T))


def nrc_round16(x, mode):
    return lib().or_nrc_round16(C.c_double(x), C.c_uint32(mode))


def nrc_sh16(v):
    out = np.zeros(16, np.float64)
    lib().or_nrc_sh16(ptr(np.ascontiguousarray(v, np.float64)), ptr(out))
    return out


class Nrc:
    """FP64 learned radiance cache (encoding, MLP, relative L2 loss, backprop, Adam)."""

    def __init__(self, levels, feats, log2_T, res, box_lo, box_hi, weights, table, lr=1e-2, beta1=0.9,
                 beta2=0.99, eps=1e-10, loss_scale=128.0, loss=0, emulate_bf16=False, emulate_fp16=False):
        self.L = lib()
        r = (C.c_uint32 * 16)(*list(res))
        self.cfg = NrcConfig(levels, feats, log2_T, r, (C.c_float * 3)(*box_lo), (C.c_float * 3)(*box_hi),
                             lr, beta1, beta2, eps, loss_scale, loss, 2 if emulate_fp16 else int(bool(emulate_bf16)))
        self.n_table = levels * (1 << log2_T) * feats
        w = np.ascontiguousarray(weights, np.float32)
        t = np.ascontiguousarray(table, np.float32)
        assert w.size == N_WEIGHTS and t.size == self.n_table
        self.h = C.c_void_p(self.L.or_nrc_create(C.byref(self.cfg), ptr(w), ptr(t)))

    def close(self):
        if self.h:
            self.L.or_nrc_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _in(pos, dir, nrm, alb):
        f = lambda a: np.ascontiguousarray(a, np.float32).reshape(-1, 3)
        return f(pos), f(dir), f(nrm), f(alb)

    def encode(self, pos, dir, nrm, alb):
        p, d, n, a = self._in(pos, dir, nrm, alb)
        out = np.zeros((len(p), N_IN), np.float64)
        self.L.or_nrc_encode(self.h, C.c_uint64(len(p)), ptr(p), ptr(d), ptr(n), ptr(a), ptr(out))
        return out

    def forward(self, pos, dir, nrm, alb):
        p, d, n, a = self._in(pos, dir, nrm, alb)
        out = np.zeros((len(p), 3), np.float64)
        self.L.or_nrc_forward(self.h, C.c_uint64(len(p)), ptr(p), ptr(d), ptr(n), ptr(a), ptr(out))
        return out

    def grads(self, pos, dir, nrm, alb, target):
        p, d, n, a = self._in(pos, dir, nrm, alb)
        t = np.ascontiguousarray(target, np.float32).reshape(-1, 3)
        gw = np.zeros(N_WEIGHTS, np.float64)
        gt = np.zeros(self.n_table, np.float64)
        loss = self.L.or_nrc_grads(self.h, C.c_uint64(len(p)), ptr(p), ptr(d), ptr(n), ptr(a), ptr(t), ptr(gw), ptr(gt))
        return loss, gw, gt

    def train_step(self, pos, dir, nrm, alb, target):
        p, d, n, a = self._in(pos, dir, nrm, alb)
        t = np.ascontiguousarray(target, np.float32).reshape(-1, 3)
        return self.L.or_nrc_train_step(self.h, C.c_uint64(len(p)), ptr(p), ptr(d), ptr(n), ptr(a), ptr(t))

    def get(self):
        w = np.zeros(N_WEIGHTS, np.float64)
        t = np.zeros(self.n_table, np.float64)
        self.L.or_nrc_get(self.h, ptr(w), ptr(t))
        return w, t

    def set(self, weights=None, table=None):
        w = None if weights is None else np.ascontiguousarray(weights, np.float64)
        t = None if table is None else np.ascontiguousarray(table, np.float64)
        self.L.or_nrc_set(self.h, None if w is None else ptr(w), None if t is None else ptr(t))
