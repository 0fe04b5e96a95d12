"""Btrc's JSON scene files, read into this package's scene description.

    python -m btrc_amd.scene_json scene.json      # render it like BtrcCLI

Follows the reference's front end: `$include{file}` text includes and
`$reference{path}` node references (factory/node/parser.cpp:59-252), JSON
with comments, `$env{scene_directory}` in paths (cli/main.cpp:29-30,
factory/path_resolver.cpp), typed values (factory/node/node.cpp:50-200:
vectors and spectra from one or three numbers, angles as {"deg"} or
{"rad"}, transforms as translate / rotate / rotate_x / rotate_y / rotate_z /
scale or arrays of them applied first to last), and the creators of the
registered object types (builtin/register.cpp) that the wavefront path
tracer renders: triangle_mesh geometry (OBJ), diffuse / disney / invisible /
black materials, mesh_light, gradient_sky and ibl lights, homogeneous /
hetergeneous / transform media, volume primitives, constant and array
textures, the pinhole camera, box / gaussian film filters, the wfpt and pt
renderers, and the aces_tonemap / save_to_image post processors. Other
types raise an error that names them.
"""
import json
import math
import os
import re
import struct
import sys

import numpy as np

from . import _abi, scenes


class SceneFileError(ValueError):
    pass


# ------------------------------------------------------------ text front end

def _strip_comments(src):
    """JSON with // and /* */ comments (nlohmann's ignore_comments)"""
    out, i, n = [], 0, len(src)
    in_str = False
    while i < n:
        c = src[i]
        if in_str:
            out.append(c)
            if c == "\\" and i + 1 < n:
                out.append(src[i + 1])
                i += 2
                continue
            if c == '"':
                in_str = False
            i += 1
        elif c == '"':
            in_str = True
            out.append(c)
            i += 1
        elif src.startswith("//", i):
            j = src.find("\n", i)
            i = n if j < 0 else j
        elif src.startswith("/*", i):
            j = src.find("*/", i + 2)
            if j < 0:
                raise SceneFileError("unterminated comment")
            i = j + 2
        else:
            out.append(c)
            i += 1
    return "".join(out)


def _process_includes(src, include_dirs):
    # JSONParser::process_includes: "$include{file}" (with the quotes) is
    # replaced by the file's text
    while True:
        beg = src.find('"$include{')
        if beg < 0:
            return src
        end = src.find('}"', beg + 10)
        if end < 0:
            raise SceneFileError("failed to find the enclosing '}' of a '$include{'")
        name = src[beg + 10:end]
        path = name if os.path.isabs(name) else next(
            (os.path.join(d, name) for d in include_dirs if os.path.exists(os.path.join(d, name))), None)
        if path is None:
            raise SceneFileError(f"included file '{name}' is not found")
        with open(path) as f:
            src = src[:beg] + f.read() + src[end + 2:]


def _find_node(path_nodes, ref):
    # JSONParser::find_node: "$root/a/b", or a path starting at the nearest
    # enclosing group (or array) that has its first section
    secs = ref.split("/")
    if not secs or not secs[0]:
        raise SceneFileError("invalid node path: " + ref)
    if secs[0] == "$root":
        start = 0
    else:
        start = len(path_nodes) - 1
        while start >= 0:
            n = path_nodes[start]
            if isinstance(n, dict) and secs[0] in n:
                break
            if isinstance(n, list) and secs[0].isdigit() and int(secs[0]) < len(n):
                break
            start -= 1
        if start < 0:
            raise SceneFileError("invalid node path: " + ref)
    node = path_nodes[start]
    for i, sec in enumerate(secs):
        if i == 0 and sec == "$root":
            continue
        if isinstance(node, dict) and sec in node:
            node = node[sec]
        elif isinstance(node, list) and sec.isdigit() and int(sec) < len(node):
            node = node[int(sec)]
        else:
            raise SceneFileError("invalid node path: " + ref)
    if isinstance(node, str) and node.startswith("$reference{"):
        node = _find_node(path_nodes, _ref_target(node))
    return node


def _ref_target(s):
    if not s.endswith("}"):
        raise SceneFileError("'}' for '$reference{' is not found")
    return s[len("$reference{"):-1]


def _resolve_references(path_nodes, node):
    if isinstance(node, dict):
        path_nodes.append(node)
        for k in list(node):
            node[k] = _resolve_references(path_nodes, node[k])
        path_nodes.pop()
        return node
    if isinstance(node, list):
        path_nodes.append(node)
        for k in range(len(node)):
            node[k] = _resolve_references(path_nodes, node[k])
        path_nodes.pop()
        return node
    if isinstance(node, str) and node.startswith("$reference{"):
        return _find_node(path_nodes, _ref_target(node))
    return node


def parse_text(src, include_dirs=()):
    """the node tree of a scene file's text (includes and references resolved)"""
    src = _process_includes(src, list(include_dirs))
    root = json.loads(_strip_comments(src))
    if not isinstance(root, dict):
        raise SceneFileError("root node is not a group")
    return _resolve_references([root], root)


# ------------------------------------------------------------ typed values

def vec3(node):
    """parse_vec3f (node.cpp:50-74): one number, or one or three in an array"""
    if isinstance(node, list):
        if len(node) == 1:
            return [float(node[0])] * 3
        if len(node) == 3:
            return [float(v) for v in node]
        raise SceneFileError("vec3 expects 1 or 3 numbers")
    if isinstance(node, (int, float, str)):
        return [float(node)] * 3
    raise SceneFileError("vec3 expected")


spectrum = vec3   # parse_spectrum (node.cpp:100-124) reads the same shapes


def radian(node):
    if isinstance(node, dict):
        if "deg" in node:
            return math.pi / 180.0 * float(node["deg"])
        if "rad" in node:
            return float(node["rad"])
    raise SceneFileError("invalid angle node")


def degree(node):
    if isinstance(node, dict):
        if "deg" in node:
            return float(node["deg"])
        if "rad" in node:
            return 180.0 / math.pi * float(node["rad"])
    raise SceneFileError("invalid angle node")


def _rot(c, s, axis):
    # Mat4::rotate_x / y / z (utils/math/mat4.cpp:147-175), row-major
    m = np.eye(4)
    i, j = {"x": (1, 2), "y": (2, 0), "z": (0, 1)}[axis]
    m[i, i], m[i, j], m[j, i], m[j, j] = c, -s, s, c
    return m


def _rot_axis(axis, rad):
    # Mat4::rotate(axis, rad) (mat4.cpp:117-145) fills data[column][row] with
    # the Rodrigues terms, i.e. the transpose of the textbook matrix
    a = np.asarray(axis, dtype=np.float64)
    a = a / np.linalg.norm(a)
    s, c = math.sin(rad), math.cos(rad)
    x, y, z = a
    r = np.array([[x * x + (1 - x * x) * c, x * y * (1 - c) - z * s, x * z * (1 - c) + y * s],
                  [x * y * (1 - c) + z * s, y * y + (1 - y * y) * c, y * z * (1 - c) - x * s],
                  [x * z * (1 - c) - y * s, y * z * (1 - c) + x * s, z * z + (1 - z * z) * c]])
    m = np.eye(4)
    m[:3, :3] = r.T
    return m


def transform3d(node):
    """parse_transform3d (node.cpp:166-200) as (mat, inv); an array applies
    its elements first to last. Transform3D::rotate_z keeps the reference's
    pair (rotate_z(-rad), rotate_z(-rad)) (utils/math/transform.cpp:111-114)."""
    if isinstance(node, list):
        mat, inv = np.eye(4), np.eye(4)
        for e in node:
            m, i = transform3d(e)
            mat, inv = m @ mat, inv @ i
        return mat, inv
    if not isinstance(node, dict):
        raise SceneFileError("transform expected")
    if "translate" in node:
        t = vec3(node["translate"])
        m, i = np.eye(4), np.eye(4)
        m[:3, 3] = t
        i[:3, 3] = [-v for v in t]
        return m, i
    if "rotate" in node:
        r = node["rotate"]
        rad = radian(r["angle"])
        return _rot_axis(vec3(r["axis"]), rad), _rot_axis(vec3(r["axis"]), -rad)
    for ax in ("x", "y"):
        if "rotate_" + ax in node:
            rad = radian(node["rotate_" + ax])
            return _rot(math.cos(rad), math.sin(rad), ax), _rot(math.cos(-rad), math.sin(-rad), ax)
    if "rotate_z" in node:
        rad = radian(node["rotate_z"])
        m = _rot(math.cos(-rad), math.sin(-rad), "z")
        return m, m.copy()
    if "scale" in node:
        s = vec3(node["scale"])
        return np.diag(s + [1.0]), np.diag([1.0 / v for v in s] + [1.0])
    raise SceneFileError("unknown transform3d type")


def _bool(v):
    if isinstance(v, str):
        return v == "true"
    return bool(v)


# ------------------------------------------------------------ files

def load_obj(path, unit_cube=False):
    """TriangleMeshLoader (utils/triangle_mesh_loader.cpp:12-146) over an OBJ
    file: polygons fanned into triangles, per-corner normals (a zero normal
    or none: the face normal) and texture coordinates (none: 0); with
    unit_cube, positions centred and scaled by 1 / the largest extent
    (:197-213). Returns positions (n, 9), normals (n, 9), uvs (n, 6)."""
    vs, vns, vts, faces = [], [], [], []
    with open(path) as f:
        for line in f:
            p = line.split()
            if not p or p[0].startswith("#"):
                continue
            if p[0] == "v":
                vs.append([float(x) for x in p[1:4]])
            elif p[0] == "vn":
                vns.append([float(x) for x in p[1:4]])
            elif p[0] == "vt":
                vts.append([float(x) for x in p[1:3]] + [0.0] * (2 - len(p[1:3])))
            elif p[0] == "f":
                corner = []
                for tok in p[1:]:
                    idx = (tok.split("/") + ["", ""])[:3]
                    def fix(s, n):
                        if not s:
                            return -1
                        k = int(s)
                        return k - 1 if k > 0 else n + k
                    corner.append((fix(idx[0], len(vs)), fix(idx[1], len(vts)), fix(idx[2], len(vns))))
                for k in range(1, len(corner) - 1):
                    faces.append((corner[0], corner[k], corner[k + 1]))
    V = np.asarray(vs, dtype=np.float32).reshape(-1, 3)
    if unit_cube and len(V):
        lo, hi = V.min(axis=0), V.max(axis=0)
        ext = float((hi - lo).max())
        scale = np.float32(1.0 / ext) if ext > 0 else np.float32(1.0)
        V = (scale * (V - np.float32(0.5) * (lo + hi))).astype(np.float32)
    n = len(faces)
    pos = np.zeros((n, 9), np.float32)
    nor = np.zeros((n, 9), np.float32)
    uvs = np.zeros((n, 6), np.float32)
    VN = np.asarray(vns, dtype=np.float32).reshape(-1, 3)
    VT = np.asarray(vts, dtype=np.float32).reshape(-1, 2)
    for t, face in enumerate(faces):
        for k, (vi, ti, ni) in enumerate(face):
            if not 0 <= vi < len(V):
                raise SceneFileError("invalid obj vertex index: out of range")
            pos[t, 3 * k:3 * k + 3] = V[vi]
            if ni >= 0:
                if ni >= len(VN):
                    raise SceneFileError("invalid obj normal index: out of range")
                nor[t, 3 * k:3 * k + 3] = VN[ni]    # zero normals become the face normal on upload
            if ti >= 0:
                if ti >= len(VT):
                    raise SceneFileError("invalid obj tex coord index: out of range")
                uvs[t, 2 * k:2 * k + 2] = VT[ti]
    return pos, nor, uvs


def _load_hdr(path):
    """a Radiance RGBE file as (H, W, 4) float32 with alpha 1: flat or
    run-length scanlines, each texel m * 2^(e - 136) (utils/image.cpp:150-174
    over stb_image's HDR reader)"""
    data = open(path, "rb").read()
    end = data.find(b"\n\n")
    if not data.startswith(b"#?") or end < 0:
        raise SceneFileError(f"{path}: not a Radiance HDR file")
    eol = data.find(b"\n", end + 2)
    res = data[end + 2:eol].split()
    if len(res) != 4 or res[0] != b"-Y" or res[2] != b"+X":
        raise SceneFileError(f"{path}: unsupported HDR orientation")
    h, w = int(res[1]), int(res[3])
    pos = eol + 1
    out = np.zeros((h, w, 4), np.uint8)
    for y in range(h):
        head = data[pos:pos + 4]
        if 8 <= w < 32768 and len(head) == 4 and head[0] == 2 and head[1] == 2 and (head[2] << 8 | head[3]) == w:
            pos += 4
            for c in range(4):
                x = 0
                while x < w:
                    n = data[pos]
                    pos += 1
                    if n > 128:
                        out[y, x:x + n - 128, c] = data[pos]
                        pos += 1
                        x += n - 128
                    else:
                        out[y, x:x + n, c] = np.frombuffer(data, np.uint8, n, pos)
                        pos += n
                        x += n
        else:
            out[y] = np.frombuffer(data, np.uint8, 4 * w, pos).reshape(w, 4)
            pos += 4 * w
    e = out[..., 3].astype(np.int32)
    f = np.where(e > 0, np.ldexp(np.float32(1.0), e - 136).astype(np.float32), np.float32(0.0))
    rgba = np.ones((h, w, 4), np.float32)
    rgba[..., :3] = out[..., :3].astype(np.float32) * f[..., None]
    return rgba


def load_image(path):
    """the Array2D texels of an image file (ImageDynamic::load,
    utils/image.cpp:299-348): PNG / JPG as (H, W, 4) uint8 (RGB gets alpha
    255), HDR as (H, W, 4) float32"""
    if os.path.splitext(path)[1].lower() == ".hdr":
        return _load_hdr(path)
    if os.path.splitext(path)[1].lower() == ".exr":
        raise SceneFileError(f"{path}: EXR images cannot be loaded (as in the reference)")
    from PIL import Image
    im = Image.open(path)
    if im.mode not in ("RGB", "RGBA", "L", "LA", "P"):
        raise SceneFileError(f"{path}: only 8-bit images can be textures here (mode {im.mode})")
    return np.asarray(im.convert("RGBA"), dtype=np.uint8).copy()


# ------------------------------------------------------------ objects

class SceneFile:
    """what BtrcCLI's run() builds from a file (cli/main.cpp:17-81)"""

    def __init__(self):
        self.scene = scenes.Scene()
        self.camera = None
        self.width = self.height = 0
        self.renderer = {}
        self.filter = None          # None: box; ("gaussian", radius, alpha)
        self.post_processors = []   # dicts
        self.directory = "."


class _Builder:
    def __init__(self, root, directory):
        self.root = root
        self.dir = directory
        self.out = SceneFile()
        self.out.directory = directory
        self.s = self.out.scene
        self.cache = {}             # Context::object_pool_: one object per node

    def path(self, p):
        # PathResolver::resolve with "$env{scene_directory}"
        p = p.replace("$env{scene_directory}", self.dir)
        return os.path.normpath(p if os.path.isabs(p) else os.path.join(os.getcwd(), p))

    def _cached(self, kind, node, make):
        key = (kind, id(node))
        if key not in self.cache:
            self.cache[key] = (node, make(node))
        return self.cache[key][1]

    # --- textures
    def texture2d(self, node):
        """('const', rgb) or ('tex', index)"""
        if not isinstance(node, dict):
            return ("const", spectrum(node))
        t = node.get("type")
        if t == "constant":
            return ("const", spectrum(node["value"]))
        if t == "array":
            def make(n):
                texels = load_image(self.path(n["filename"]))
                return self.s.add_texture(
                    texels, address_u=n.get("address_mode_u", "clamp"), address_v=n.get("address_mode_v", "clamp"),
                    filter=n.get("filter", "linear"), srgb_to_linear=_bool(n.get("srgb_to_linear", False)),
                    border=self._border(n))
            return ("tex", self._cached("tex2d", node, make))
        raise SceneFileError(f"texture2d type '{t}' is not supported by this renderer")

    @staticmethod
    def _border(n):
        b = n.get("border")
        if b is None:
            return (0.0, 0.0, 0.0, 0.0)
        if isinstance(b, list) and len(b) == 3:
            return (float(b[0]), float(b[1]), float(b[2]), 0.0)
        v = float(b[0] if isinstance(b, list) else b)
        return (v, v, v, v)

    def texture3d(self, node):
        """('const', rgb) or ('grid', array) with Array3D's text layout"""
        if not isinstance(node, dict):
            return ("const", spectrum(node))
        t = node.get("type")
        if t == "constant":
            return ("const", spectrum(node["value"]))
        if t == "array3d":
            if "text" not in node:
                raise SceneFileError("array3d: only text grids are supported")
            return ("grid", self._cached("tex3d", node, lambda n: scenes.load_grid_text(self.path(n["text"]))))
        raise SceneFileError(f"texture3d type '{t}' is not supported by this renderer")

    # --- materials
    def material(self, node):
        return self._cached("mat", node, self._make_material)

    def _scalar(self, node, name, default):
        """(constant, None) or (default, texture index) of a Disney parameter"""
        if name not in node:
            return default, None
        kind, v = self.texture2d(node[name])
        if kind == "const":
            return v[0], None
        return default, v

    def _normal_map(self, node):
        """NormalMap::load (material/utils/normal_map.cpp:5-11): the texture
        under "normal", if any, as a texture index"""
        if "normal" not in node:
            return None
        kind, v = self.texture2d(node["normal"])
        if kind == "const":
            # a constant normal map: a one-texel texture of that colour
            v = self.s.add_texture(np.array([[[*v[:3], 1.0]]], np.float32), filter="point")
        return v

    def _make_material(self, node):
        t = node.get("type")
        st = _bool(node.get("shadow_terminator_term", True))
        if t == "diffuse":
            kind, v = self.texture2d(node["color"])
            nmap = self._normal_map(node)
            if kind == "const":
                return self.s.add_diffuse(v, shadow_terminator=st, normal_texture=nmap)
            return self.s.add_diffuse([1.0, 1.0, 1.0], shadow_terminator=st, color_texture=v, normal_texture=nmap)
        if t == "black":
            return self.s.add_diffuse([0.0, 0.0, 0.0], shadow_terminator=st)
        if t == "invisible":
            return self.s.add_invisible()
        if t == "disney":
            # DisneyMaterialCreator, material/disney.cpp:844-894
            kind, col = self.texture2d(node["color"])
            textures = {}
            color = col
            if kind != "const":
                color, textures["color"] = [1.0, 1.0, 1.0], col
            kw = {}
            for name, d in (("metallic", 0.0), ("roughness", 0.0), ("transmission", 0.0), ("ior", 1.5),
                            ("specular", 1.0), ("specular_tint", 0.0), ("anisotropic", 0.0), ("sheen", 0.0),
                            ("sheen_tint", 0.0), ("clearcoat", 0.0), ("clearcoat_gloss", 1.0),
                            ("transmission_roughness", None)):
                kw[name], tex = self._scalar(node, name, d)
                if tex is not None:
                    textures[name] = tex
            if "transmission_roughness" in textures and kw["transmission_roughness"] is None:
                kw["transmission_roughness"] = 0.0
            return self.s.add_disney(color, shadow_terminator=st, textures=textures,
                                     normal_texture=self._normal_map(node), **kw)
        raise SceneFileError(f"material type '{t}' is not supported by this renderer")

    # --- media
    def medium(self, node):
        return self._cached("med", node, self._make_medium)

    def _make_medium(self, node, xf=None):
        t = node.get("type")
        if t == "transform":
            # TransformMediumCreator (factory/context.h:146-164)
            if "world_to_local" in node:
                w2l = transform3d(node["world_to_local"])[0]
            else:
                w2l = transform3d(node["local_to_world"])[1]
            return self._make_medium(node["medium"], w2l)
        pr = float(node.get("priority", 0.0))
        if t == "homogeneous":
            return self.s.add_homogeneous(float(node["sigma_t"]), spectrum(node["albedo"]),
                                          g=float(node.get("g", 0.0)), priority=pr)
        if t == "hetergeneous":
            kind, sig = self.texture3d(node["sigma_t"])
            grid = np.full((1, 1, 1), sig[0], np.float32) if kind == "const" else np.asarray(sig, np.float32)
            if grid.ndim != 3:
                raise SceneFileError("hetergeneous: sigma_t must have one channel")
            akind, alb = self.texture3d(node["albedo"])
            g = 0.0
            if "g" in node:
                gk, gv = self.texture3d(node["g"])
                if gk != "const":
                    raise SceneFileError("hetergeneous: a grid of g is not supported by this renderer")
                g = gv[0]
            if akind == "const":
                return self.s.add_heterogeneous(grid, alb, g=g, priority=pr, world_to_local=xf)
            alb = np.asarray(alb, np.float32)
            if alb.ndim != 4 or alb.shape[3] < 3:
                raise SceneFileError("hetergeneous: an albedo grid needs three channels")
            return self.s.add_heterogeneous(grid, (1.0, 1.0, 1.0), g=g, priority=pr, world_to_local=xf,
                                            albedo_grid=np.ascontiguousarray(alb[..., :3]))
        raise SceneFileError(f"medium type '{t}' is not supported by this renderer")

    # --- scene
    def build(self):
        root = self.root
        scene = root["scene"]
        for e in scene.get("entities", []):
            g = e["geometry"]
            if g.get("type") != "triangle_mesh":
                raise SceneFileError(f"geometry type '{g.get('type')}' is not supported by this renderer")
            geo = self._cached("geo", g, lambda n: self.s.add_geometry(
                *load_obj(self.path(n["filename"]), _bool(n.get("transform_to_unit_cube", False)))))
            mat = self.material(e["material"])
            l2w, w2l = transform3d(e["local_to_world"]) if "local_to_world" in e else (np.eye(4), np.eye(4))
            light = None
            if "light" in e:
                ln = e["light"]
                if ln.get("type") != "mesh_light":
                    raise SceneFileError(f"area light type '{ln.get('type')}' is not supported by this renderer")
                light = tuple(spectrum(ln["intensity"]))
            inner = self.medium(e["inner_medium"]) if "inner_medium" in e else -1
            outer = self.medium(e["outer_medium"]) if "outer_medium" in e else -1
            k = self.s.add_instance(geo, mat, l2w.astype(np.float32), light=light, inner_medium=inner,
                                    outer_medium=outer)
            self.s.instances[k]["inv"] = w2l.astype(np.float32)
        for v in scene.get("volumes", []):
            kind, sig = self.texture3d(v["sigma_t"])
            sigma = sig[0] if kind == "const" else np.asarray(sig, np.float32)
            akind, alb = self.texture3d(v["albedo"])
            albedo = alb if akind == "const" else np.ascontiguousarray(np.asarray(alb, np.float32)[..., :3])
            self.s.add_volume(vec3(v["o"]), vec3(v["x"]), vec3(v["y"]), vec3(v["z"]), sigma, albedo)
        if "envir_light" in scene:
            el = scene["envir_light"]
            if el.get("type") == "gradient_sky":
                self.s.set_gradient_sky(vec3(el["up"]), spectrum(el["lower"]), spectrum(el["upper"]))
            elif el.get("type") == "ibl":
                # IBLCreator, light/ibl.cpp:80-91
                kind, tex = self.texture2d(el["tex"])
                if kind == "const":
                    texel = np.ones((1, 1, 4), np.float32)
                    texel[0, 0, :3] = tex
                    tex = self.s.add_texture(texel, filter="point")
                up = tuple(vec3(el["up"])) if "up" in el else (0.0, 0.0, 1.0)
                self.s.set_ibl(tex, up=up, lut_width=int(el.get("lut_width", 400)),
                               lut_height=int(el.get("lut_height", 200)))
            else:
                raise SceneFileError(f"environment light type '{el.get('type')}' is not supported by this renderer")

        out = self.out
        out.width, out.height = int(root["width"]), int(root["height"])
        cam = root["camera"]
        if cam.get("type") != "pinhole":
            raise SceneFileError(f"camera type '{cam.get('type')}' is not supported by this renderer")
        out.camera = scenes.Camera(tuple(vec3(cam["eye"])), tuple(vec3(cam["dst"])), tuple(vec3(cam["up"])),
                                   degree(cam["fov_y"]))
        r = root["renderer"]
        if r.get("type") not in ("wfpt", "pt"):
            raise SceneFileError(f"renderer type '{r.get('type')}' is not supported")
        # WavefrontPathTracer::Params defaults (renderer/wavefront.h:13-27)
        defaults = dict(tile=False, spp=128, min_depth=5, max_depth=10, rr_threshold=0.1, rr_cont_prob=0.5,
                        state_count=1000000, albedo=False, normal=False)
        out.renderer = {k: (_bool(r[k]) if isinstance(d, bool) else type(d)(r[k])) if k in r else d
                        for k, d in defaults.items()}
        if "filter" in r:
            f = r["filter"]
            if f.get("type") == "gaussian":
                out.filter = ("gaussian", float(f.get("radius", 1.3)), float(f.get("alpha", 1.6)))
            elif f.get("type") != "box":
                raise SceneFileError(f"film filter type '{f.get('type')}' is not supported")
        for pp in root.get("post_processors", []) or []:
            t = pp.get("type")
            if t == "aces_tonemap":
                out.post_processors.append(dict(type=t, exposure=float(pp.get("exposure", 1.0))))
            elif t == "save_to_image":
                out.post_processors.append(dict(type=t, gamma=float(pp.get("gamma", 1.0)),
                                                **{k: self.path(pp[k]) for k in ("color", "albedo", "normal")
                                                   if pp.get(k)}))
            elif t == "optix_denoiser":
                raise SceneFileError("the OptiX denoiser has no counterpart in this renderer")
            else:
                raise SceneFileError(f"post processor type '{t}' is not supported")
        return out


def load(path):
    """a SceneFile from a Btrc scene file"""
    directory = os.path.dirname(os.path.abspath(path))
    with open(path) as f:
        root = parse_text(f.read(), [directory])
    return _Builder(root, directory).build()


# ------------------------------------------------------------ post processing

def save_image(path, rgba):
    """Image::save (utils/image.cpp, image.h:495-540) by extension: PNG / JPG
    as bytes min(int(256 x), 255) (negative values as 0), HDR as RGBE, EXR
    as uncompressed FP32 B, G, R"""
    ext = os.path.splitext(path)[1].lower()
    rgb = np.asarray(rgba, np.float32)[..., :3]
    if ext in (".png", ".jpg", ".jpeg"):
        from PIL import Image
        b = np.clip(np.floor(np.maximum(rgb, 0.0) * 256.0), 0, 255).astype(np.uint8)
        Image.fromarray(b, "RGB").save(path)
    elif ext == ".hdr":
        _save_hdr(path, rgb)
    elif ext == ".exr":
        _save_exr(path, rgb)
    else:
        raise SceneFileError(f"unknown image format: {path}")


def _save_hdr(path, rgb):
    h, w = rgb.shape[:2]
    m = rgb.max(axis=2)
    e = np.zeros_like(m, dtype=np.int32)
    mant = np.zeros_like(m)
    ok = m > 1e-32
    mant[ok], e[ok] = np.frexp(m[ok])
    scale = np.where(ok, mant * 256.0 / np.where(ok, m, 1.0), 0.0)
    out = np.zeros((h, w, 4), np.uint8)
    out[..., :3] = np.clip(rgb * scale[..., None], 0, 255).astype(np.uint8)
    out[..., 3] = np.where(ok, e + 128, 0).astype(np.uint8)
    with open(path, "wb") as f:
        f.write(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n")
        f.write(f"-Y {h} +X {w}\n".encode())
        f.write(out.tobytes())


def _save_exr(path, rgb):
    """OpenEXR 2.0 single-part scanline file, no compression, FP32 channels"""
    h, w = rgb.shape[:2]

    def attr(name, typ, data):
        return name.encode() + b"\0" + typ.encode() + b"\0" + struct.pack("<i", len(data)) + data

    chans = b"".join(c.encode() + b"\0" + struct.pack("<iBBBBii", 2, 0, 0, 0, 0, 1, 1) for c in "BGR") + b"\0"
    box = struct.pack("<iiii", 0, 0, w - 1, h - 1)
    header = (struct.pack("<ii", 20000630, 2) + attr("channels", "chlist", chans) +
              attr("compression", "compression", b"\0") + attr("dataWindow", "box2i", box) +
              attr("displayWindow", "box2i", box) + attr("lineOrder", "lineOrder", b"\0") +
              attr("pixelAspectRatio", "float", struct.pack("<f", 1.0)) +
              attr("screenWindowCenter", "v2f", struct.pack("<ff", 0.0, 0.0)) +
              attr("screenWindowWidth", "float", struct.pack("<f", 1.0)) + b"\0")
    line_bytes = 3 * 4 * w
    offsets = len(header) + 8 * h + np.arange(h, dtype=np.int64) * (8 + line_bytes)
    with open(path, "wb") as f:
        f.write(header)
        f.write(offsets.astype("<u8").tobytes())
        for y in range(h):
            f.write(struct.pack("<ii", y, line_bytes))
            for c in (2, 1, 0):
                f.write(rgb[y, :, c].astype("<f4").tobytes())


def render_file(path, ctx=None, log=print):
    """BtrcCLI's run(): load, render with the wavefront path tracer, run the
    post processors; returns the RenderResult"""
    import btrc_amd
    sf = load(path)
    r = btrc_amd.WavefrontPathTracer(ctx, **sf.renderer)
    r.set_scene(sf.scene)
    r.set_camera(sf.camera)
    r.set_film(sf.width, sf.height)
    if sf.filter is not None:
        r.set_film_filter(*sf.filter)
    log(f"render {sf.width}x{sf.height}, {sf.renderer['spp']} spp")
    result = r.render()
    color = result.color
    tonemapped = False
    for pp in sf.post_processors:
        if pp["type"] == "aces_tonemap":
            # on the device, from the film (ACESToneMap::process)
            if tonemapped:
                raise SceneFileError("only one aces_tonemap post processor is supported")
            color = r.ctx.tonemapped(sf.width, sf.height, pp["exposure"])
            tonemapped = True
        elif pp["type"] == "save_to_image":
            if "color" in pp:
                save_image(pp["color"], np.power(np.maximum(color, 0.0), 1.0 / pp["gamma"]))
            if "albedo" in pp and result.albedo is not None:
                save_image(pp["albedo"], result.albedo)
            if "normal" in pp and result.normal is not None:
                save_image(pp["normal"], result.normal)
    result.color = color
    return result


if __name__ == "__main__":
    if len(sys.argv) != 2:
        print("usage: python -m btrc_amd.scene_json config.json")
        sys.exit(0)
    try:
        render_file(sys.argv[1])
    except (SceneFileError, OSError) as e:
        print(e, file=sys.stderr)
        sys.exit(1)
