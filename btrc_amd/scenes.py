"""Scene description on the host and the seeded synthetic scenes the tracer
is measured on.

A `Scene` holds what Btrc's JSON scene factory produces for the render path
(factory/scene.cpp:5-94): triangle-mesh geometries, materials, entities
(geometry + material + local_to_world + optional mesh light) and an optional
environment light. `Scene.to_desc()` packs it into the C structs of
include/btrc_hip.h.

Every random draw of the synthetic scenes comes from PCG32 (O'Neill 2014,
XSH-RR output, 64-bit LCG state) with a fixed seed, so a scene is the same
bits on every host.
"""
import ctypes as C
import math

import numpy as np

from . import _abi

# --------------------------------------------------------------------- PCG32


class PCG32:
    """PCG-XSH-RR 64/32 (M. E. O'Neill, "PCG: A Family of Simple Fast
    Space-Efficient Statistically Good Algorithms for Random Number
    Generation", 2014)."""

    MULT = 6364136223846793005
    MASK = (1 << 64) - 1

    def __init__(self, seed, seq=54):
        self.inc = ((seq << 1) | 1) & self.MASK
        self.state = 0
        self.next_u32()
        self.state = (self.state + seed) & self.MASK
        self.next_u32()

    def next_u32(self):
        old = self.state
        self.state = (old * self.MULT + self.inc) & self.MASK
        xorshifted = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        return ((xorshifted >> rot) | (xorshifted << ((-rot) & 31))) & 0xFFFFFFFF

    def uniform(self):
        """float32 in [0, 1): the top 24 bits of one draw"""
        return np.float32((self.next_u32() >> 8) * (1.0 / 16777216.0))

    def range(self, lo, hi):
        return np.float32(lo + (hi - lo) * float(self.uniform()))


# ------------------------------------------------------------ transforms


def translate(x, y, z):
    m = np.eye(4)
    m[:3, 3] = (x, y, z)
    return m


def scale(s):
    return np.diag([s, s, s, 1.0])


def rotate_y(rad):
    c, s = math.cos(rad), math.sin(rad)
    m = np.eye(4)
    m[0, 0], m[0, 2], m[2, 0], m[2, 2] = c, s, -s, c
    return m


def quat_matrix(q):
    w, x, y, z = q
    m = np.eye(4)
    m[:3, :3] = [[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                 [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                 [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]]
    return m


def uniform_rotation(rng):
    """Shoemake's uniform random unit quaternion ("Uniform random rotations",
    Graphics Gems III, 1992) as a 4x4 matrix."""
    u1, u2, u3 = float(rng.uniform()), float(rng.uniform()), float(rng.uniform())
    a, b = math.sqrt(1 - u1), math.sqrt(u1)
    q = (b * math.cos(2 * math.pi * u3), a * math.sin(2 * math.pi * u2),
         a * math.cos(2 * math.pi * u2), b * math.sin(2 * math.pi * u3))
    return quat_matrix(q)


# ------------------------------------------------------------ scene model


class Scene:
    def __init__(self):
        self.geometries = []   # dicts: positions (n,9) f32, normals (n,9) or None, uvs (n,6) or None
        self.materials = []    # MaterialDesc kwargs
        self.instances = []    # dicts
        self.media = []        # dicts
        self.envir = None      # ("sky", up, lower, upper) or ("ibl", texture, up, lut_width, lut_height)
        self.accel = _abi.ACCEL_FLAT
        self.textures = []     # dicts
        self.volumes = []      # dicts

    def add_geometry(self, positions, normals=None, uvs=None):
        positions = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 9)
        if normals is not None:
            normals = np.ascontiguousarray(normals, dtype=np.float32).reshape(-1, 9)
        if uvs is not None:
            uvs = np.ascontiguousarray(uvs, dtype=np.float32).reshape(-1, 6)
        self.geometries.append(dict(positions=positions, normals=normals, uvs=uvs))
        return len(self.geometries) - 1

    def add_texture(self, texels, address_u="clamp", address_v="clamp", filter="linear", srgb_to_linear=False,
                    border=(0.0, 0.0, 0.0, 0.0)):
        """Array2D texture (texture2d/array2d.cpp, description.cpp:32-93);
        texels (H, W, 4) uint8 (RGBA8) or float32 (RGBA32F), row 0 at v = 0"""
        texels = np.ascontiguousarray(texels)
        assert texels.ndim == 3 and texels.shape[2] == 4
        fmt = _abi.TEXTURE_RGBA8 if texels.dtype == np.uint8 else _abi.TEXTURE_RGBA32F
        if fmt == _abi.TEXTURE_RGBA32F:
            texels = np.ascontiguousarray(texels, dtype=np.float32)
        modes = dict(clamp=_abi.ADDRESS_CLAMP, wrap=_abi.ADDRESS_WRAP, mirror=_abi.ADDRESS_MIRROR,
                     border=_abi.ADDRESS_BORDER)
        filters = dict(linear=_abi.TEX_FILTER_LINEAR, point=_abi.TEX_FILTER_POINT)
        self.textures.append(dict(texels=texels, format=fmt, address_u=modes[address_u], address_v=modes[address_v],
                                  filter=filters[filter], srgb=bool(srgb_to_linear), border=tuple(border)))
        return len(self.textures) - 1

    def add_diffuse(self, color, shadow_terminator=True, color_texture=None, normal_texture=None):
        """Diffuse (material/diffuse.cpp): albedo `color` or a texture index,
        optional normal map texture"""
        self.materials.append(dict(type=_abi.MATERIAL_DIFFUSE, color=color,
                                   shadow_terminator=shadow_terminator, color_texture=color_texture,
                                   normal_texture=normal_texture))
        return len(self.materials) - 1

    def add_disney(self, color, metallic, roughness, specular=1.0, specular_tint=0.0, anisotropic=0.0,
                   sheen=0.0, sheen_tint=0.0, clearcoat=0.0, clearcoat_gloss=1.0, transmission=0.0,
                   transmission_roughness=None, ior=1.5, shadow_terminator=True, color_texture=None,
                   metallic_texture=None, roughness_texture=None, textures=None, normal_texture=None):
        """DisneyMaterial with the creator's defaults (builtin/material/
        disney.cpp:844-894). Any parameter may come from a texture: the
        `<name>_texture` arguments or `textures` = {name: texture index};
        transmission_roughness follows roughness (value and texture) unless
        given. normal_texture: a normal map (normal_map.cpp)."""
        tex = dict(textures or {})
        for name, t in (("color", color_texture), ("metallic", metallic_texture), ("roughness", roughness_texture)):
            if t is not None:
                tex[name] = t
        unknown = set(tex) - set(_abi.MATERIAL_TEXTURE_SLOTS[:-1])
        if unknown:
            raise ValueError(f"add_disney: no parameter named {sorted(unknown)}")
        if transmission_roughness is None:
            transmission_roughness = roughness
            if "roughness" in tex and "transmission_roughness" not in tex:
                tex["transmission_roughness"] = tex["roughness"]
        m = dict(type=_abi.MATERIAL_DISNEY, color=color, metallic=metallic,
                 roughness=roughness, specular=specular, specular_tint=specular_tint,
                 anisotropic=anisotropic, sheen=sheen, sheen_tint=sheen_tint,
                 clearcoat=clearcoat, clearcoat_gloss=clearcoat_gloss,
                 transmission=transmission, transmission_roughness=transmission_roughness,
                 ior=ior, shadow_terminator=shadow_terminator, normal_texture=normal_texture)
        for name, t in tex.items():
            m[name + "_texture"] = t
        self.materials.append(m)
        return len(self.materials) - 1

    def add_invisible(self):
        """InvisibleSurface (builtin/material/invisible.cpp): a boundary rays pass through"""
        self.materials.append(dict(type=_abi.MATERIAL_INVISIBLE, color=(1.0, 1.0, 1.0)))
        return len(self.materials) - 1

    def add_homogeneous(self, sigma_t, albedo, g=0.0, priority=0.0):
        """HomogeneousMedium (builtin/medium/homogeneous.cpp:75-88)"""
        self.media.append(dict(type=_abi.MEDIUM_HOMOGENEOUS, sigma_t=sigma_t, albedo=albedo, g=g,
                               priority=priority))
        return len(self.media) - 1

    def add_heterogeneous(self, sigma_t_grid, albedo, g=0.0, priority=0.0, world_to_local=None, albedo_grid=None):
        """HetergeneousMedium (builtin/medium/hetergeneous.cpp:108-133) over a
        density grid indexed [z, y, x], sampled at 0.5 + 0.5 * uvw; with
        world_to_local, uvw = world_to_local(p) (TransformMedium)."""
        grid = np.ascontiguousarray(sigma_t_grid, dtype=np.float32)
        assert grid.ndim == 3
        ag = None if albedo_grid is None else np.ascontiguousarray(albedo_grid, dtype=np.float32)
        if ag is not None:
            assert ag.shape == grid.shape + (3,)
        self.media.append(dict(type=_abi.MEDIUM_HETEROGENEOUS, grid=grid, albedo=albedo, g=g, priority=priority,
                               xf=None if world_to_local is None else np.asarray(world_to_local, dtype=np.float64),
                               albedo_grid=ag))
        return len(self.media) - 1

    def add_volume(self, o, x, y, z, sigma_t, albedo):
        """VolumePrimitive (core/volume.cpp:21-79): the box o + u x + v y + w z;
        sigma_t a float or a (D, H, W) grid, albedo an RGB triple or a
        (D, H, W, 3) grid; all volumes form the volume-primitive medium"""
        v = dict(o=tuple(map(float, o)), x=tuple(map(float, x)), y=tuple(map(float, y)), z=tuple(map(float, z)))
        if np.ndim(sigma_t) == 0:
            v.update(sigma_t=float(sigma_t), sigma_grid=None)
        else:
            v.update(sigma_t=0.0, sigma_grid=np.ascontiguousarray(sigma_t, dtype=np.float32))
        if np.ndim(albedo) == 1:
            v.update(albedo=tuple(map(float, albedo)), albedo_grid=None)
        else:
            v.update(albedo=(0.0, 0.0, 0.0), albedo_grid=np.ascontiguousarray(albedo, dtype=np.float32))
        self.volumes.append(v)
        return len(self.volumes) - 1

    def add_instance(self, geometry, material, local_to_world=None, light=None, inner_medium=-1, outer_medium=-1):
        m = np.eye(4) if local_to_world is None else np.asarray(local_to_world, dtype=np.float64)
        self.instances.append(dict(geometry=geometry, material=material, mat=m,
                                   inv=np.linalg.inv(m), light=light, inner=inner_medium, outer=outer_medium))
        return len(self.instances) - 1

    def set_gradient_sky(self, up, lower, upper):
        self.envir = ("sky", up, lower, upper)

    def set_ibl(self, texture, up=(0.0, 0.0, 1.0), lut_width=400, lut_height=200):
        """an image-based environment light (IBLCreator, light/ibl.cpp:80-91):
        `texture` is an index from add_texture, looked up at (phi / 2pi,
        theta / pi) of each direction in the frame whose z axis is `up`"""
        if not 0 <= texture < len(self.textures):
            raise ValueError("set_ibl: unknown texture")
        self.envir = ("ibl", int(texture), up, int(lut_width), int(lut_height))

    def world_box(self):
        """(lo, 1 / (hi - lo)) over the world-space triangle vertices in
        float32, as the device computes it for the radiance cache"""
        lo = np.full(3, np.float32(3.4028234663852886e38))
        hi = -lo
        for inst in self.instances:
            p = self.geometries[inst["geometry"]]["positions"].reshape(-1, 3).astype(np.float32)
            if len(p) == 0:
                continue
            m = inst["mat"].astype(np.float32)
            w = ((m[:, 0] * p[:, :1] + m[:, 1] * p[:, 1:2]) + m[:, 2] * p[:, 2:3]) + m[:, 3]
            lo = np.minimum(lo, w[:, :3].min(axis=0))
            hi = np.maximum(hi, w[:, :3].max(axis=0))
        ext = np.ones(3, np.float32)
        ok = hi > lo
        ext[ok] = hi[ok] - lo[ok]
        return np.concatenate([lo, (np.float32(1.0) / ext).astype(np.float32)]).astype(np.float32)

    @property
    def triangle_count(self):
        return sum(len(self.geometries[i["geometry"]]["positions"]) for i in self.instances)

    def to_desc(self):
        """Pack into _abi.SceneDesc; the returned holder keeps the arrays alive."""
        holder = {}
        gd = (_abi.GeometryDesc * max(1, len(self.geometries)))()
        for k, g in enumerate(self.geometries):
            fp = C.POINTER(C.c_float)
            gd[k].triangle_count = len(g["positions"])
            gd[k].positions = g["positions"].ctypes.data_as(fp)
            gd[k].normals = g["normals"].ctypes.data_as(fp) if g["normals"] is not None else None
            gd[k].tex_coords = g["uvs"].ctypes.data_as(fp) if g["uvs"] is not None else None
        md = (_abi.MaterialDesc * max(1, len(self.materials)))()
        for k, m in enumerate(self.materials):
            d = md[k]
            d.type = m["type"]
            d.shadow_terminator = 1 if m.get("shadow_terminator", True) else 0
            d.color = _abi.F3(*[float(v) for v in m["color"]])
            for key in ("metallic", "roughness", "specular", "specular_tint", "anisotropic", "sheen",
                        "sheen_tint", "clearcoat", "clearcoat_gloss", "transmission",
                        "transmission_roughness", "ior"):
                setattr(d, key, float(m.get(key, 0.0)))
            for slot in _abi.MATERIAL_TEXTURE_SLOTS:
                t = m.get(slot + "_texture")
                setattr(d, slot + "_texture", 0 if t is None else int(t) + 1)
        idesc = (_abi.InstanceDesc * max(1, len(self.instances)))()
        for k, inst in enumerate(self.instances):
            d = idesc[k]
            d.geometry = inst["geometry"]
            d.material = inst["material"]
            d.local_to_world = _abi.F16(*inst["mat"].astype(np.float32).reshape(-1).tolist())
            d.world_to_local = _abi.F16(*inst["inv"].astype(np.float32).reshape(-1).tolist())
            d.has_light = 1 if inst["light"] is not None else 0
            if inst["light"] is not None:
                d.light_intensity = _abi.F3(*[float(v) for v in inst["light"]])
            d.inner_medium = inst.get("inner", -1)
            d.outer_medium = inst.get("outer", -1)
        med = (_abi.MediumDesc * max(1, len(self.media)))()
        fp = C.POINTER(C.c_float)
        for k, m in enumerate(self.media):
            d = med[k]
            d.type = m["type"]
            d.priority = float(m["priority"])
            d.albedo = _abi.F3(*[float(v) for v in m["albedo"]])
            d.g = float(m["g"])
            if m["type"] == _abi.MEDIUM_HOMOGENEOUS:
                d.sigma_t = float(m["sigma_t"])
            else:
                nz, ny, nx = m["grid"].shape
                d.grid_size = (C.c_int32 * 3)(nx, ny, nz)
                d.sigma_t_grid = m["grid"].ctypes.data_as(fp)
                d.albedo_grid = m["albedo_grid"].ctypes.data_as(fp) if m["albedo_grid"] is not None else None
                if m["xf"] is not None:
                    d.has_transform = 1
                    d.world_to_local = _abi.F16(*m["xf"].astype(np.float32).reshape(-1).tolist())
        sd = _abi.SceneDesc()
        sd.geometry_count = len(self.geometries)
        sd.geometries = gd
        sd.material_count = len(self.materials)
        sd.materials = md
        sd.instance_count = len(self.instances)
        sd.instances = idesc
        sd.medium_count = len(self.media)
        sd.media = med
        sd.accel = self.accel
        td = (_abi.TextureDesc * max(1, len(self.textures)))()
        for k, t in enumerate(self.textures):
            h, w = t["texels"].shape[:2]
            td[k].width, td[k].height, td[k].format = w, h, t["format"]
            td[k].texels = t["texels"].ctypes.data
            td[k].address_u, td[k].address_v = t["address_u"], t["address_v"]
            td[k].filter = t["filter"]
            td[k].srgb_to_linear = 1 if t["srgb"] else 0
            td[k].border = (C.c_float * 4)(*map(float, t["border"]))
        sd.texture_count = len(self.textures)
        sd.textures = td
        vd = (_abi.VolumeDesc * max(1, len(self.volumes)))()
        for k, v in enumerate(self.volumes):
            d = vd[k]
            d.o, d.x, d.y, d.z = (_abi.F3(*v[key]) for key in ("o", "x", "y", "z"))
            d.sigma_t = v["sigma_t"]
            if v["sigma_grid"] is not None:
                nz, ny, nx = v["sigma_grid"].shape
                d.sigma_t_size = (C.c_int32 * 3)(nx, ny, nz)
                d.sigma_t_grid = v["sigma_grid"].ctypes.data_as(fp)
            d.albedo = _abi.F3(*v["albedo"])
            if v["albedo_grid"] is not None:
                nz, ny, nx, _ = v["albedo_grid"].shape
                d.albedo_size = (C.c_int32 * 3)(nx, ny, nz)
                d.albedo_grid = v["albedo_grid"].ctypes.data_as(fp)
        sd.volume_count = len(self.volumes)
        sd.volumes = vd
        if self.envir is not None and self.envir[0] == "sky":
            _, up, lower, upper = self.envir
            sd.envir_light.type = _abi.ENVIR_GRADIENT_SKY
            sd.envir_light.up = _abi.F3(*map(float, up))
            sd.envir_light.lower = _abi.F3(*map(float, lower))
            sd.envir_light.upper = _abi.F3(*map(float, upper))
        elif self.envir is not None:
            _, tex, up, lw, lh = self.envir
            sd.envir_light.type = _abi.ENVIR_IBL
            sd.envir_light.up = _abi.F3(*map(float, up))
            sd.envir_light.texture = tex
            sd.envir_light.lut_width = lw
            sd.envir_light.lut_height = lh
        else:
            sd.envir_light.type = _abi.ENVIR_NONE
        holder.update(gd=gd, md=md, idesc=idesc, med=med, td=td, vd=vd, geoms=self.geometries, media=self.media,
                      textures=self.textures, volumes=self.volumes)
        return sd, holder


class Camera:
    """PinholeCamera parameters (builtin/camera/pinhole.cpp:101-116)."""

    def __init__(self, eye, dst, up, fov_y_deg):
        self.eye, self.dst, self.up, self.fov_y_deg = eye, dst, up, fov_y_deg

    def to_desc(self, width, height):
        d = _abi.CameraDesc()
        d.eye = _abi.F3(*map(float, self.eye))
        d.dst = _abi.F3(*map(float, self.dst))
        d.up = _abi.F3(*map(float, self.up))
        d.fov_y_deg = float(self.fov_y_deg)
        # main.cpp:52: camera->set_w_over_h(width / height)
        d.w_over_h = float(np.float32(width) / np.float32(height))
        return d


# ------------------------------------------------------------ meshes


def quad(p0, p1, p2, p3):
    """two triangles (p0, p1, p2), (p0, p2, p3); counter-clockwise seen from the front"""
    return np.array([[*p0, *p1, *p2], [*p0, *p2, *p3]], dtype=np.float32)


def box_faces(sx, sy, sz):
    """the six faces of an axis-aligned box centred at the origin, outward normals"""
    x, y, z = sx / 2, sy / 2, sz / 2
    c = lambda a, b, d: (a * x, b * y, d * z)
    return [
        quad(c(-1, -1, 1), c(1, -1, 1), c(1, 1, 1), c(-1, 1, 1)),       # +z
        quad(c(1, -1, -1), c(-1, -1, -1), c(-1, 1, -1), c(1, 1, -1)),    # -z
        quad(c(1, -1, 1), c(1, -1, -1), c(1, 1, -1), c(1, 1, 1)),        # +x
        quad(c(-1, -1, -1), c(-1, -1, 1), c(-1, 1, 1), c(-1, 1, -1)),    # -x
        quad(c(-1, 1, 1), c(1, 1, 1), c(1, 1, -1), c(-1, 1, -1)),        # +y
        quad(c(-1, -1, -1), c(1, -1, -1), c(1, -1, 1), c(-1, -1, 1)),    # -y
    ]


def uv_sphere(rng=None, slices=16, stacks=9, displacement=0.0):
    """UV sphere of 2*slices + 2*slices*(stacks-2) triangles (256 for 16 x 9),
    radius 1, optionally displaced radially by U[-displacement, displacement]
    per vertex."""
    verts = {}

    def v(i, j):
        key = (i % slices if 0 < j < stacks else 0, j)
        if key not in verts:
            th = math.pi * j / stacks
            ph = 2 * math.pi * key[0] / slices
            r = 1.0
            if rng is not None and displacement > 0:
                r += displacement * (2.0 * float(rng.uniform()) - 1.0)
            verts[key] = (r * math.sin(th) * math.cos(ph), r * math.cos(th), r * math.sin(th) * math.sin(ph))
        return verts[key]

    tris = []
    for j in range(stacks):
        for i in range(slices):
            a, b, c, d = v(i, j), v(i + 1, j), v(i + 1, j + 1), v(i, j + 1)
            # counter-clockwise seen from outside: outward geometric normals
            if j == 0:
                tris.append([*a, *c, *d])       # top cap
            elif j == stacks - 1:
                tris.append([*a, *b, *d])       # bottom cap
            else:
                tris.append([*a, *c, *d])
                tris.append([*a, *b, *c])
    return np.array(tris, dtype=np.float32)


def uv_sphere_uvs(slices=16, stacks=9):
    """texture coordinates of uv_sphere's triangles, in its triangle order:
    u = slice / slices, v = stack / stacks at each corner"""
    def t(i, j):
        return (i / slices, j / stacks)

    uvs = []
    for j in range(stacks):
        for i in range(slices):
            a, b, c, d = t(i, j), t(i + 1, j), t(i + 1, j + 1), t(i, j + 1)
            if j == 0:
                uvs.append([*a, *c, *d])
            elif j == stacks - 1:
                uvs.append([*a, *b, *d])
            else:
                uvs.append([*a, *c, *d])
                uvs.append([*a, *b, *c])
    return np.array(uvs, dtype=np.float32)


def noise_texture(n=1024, seed=5, octaves=4, base=8):
    """(n, n, 4) RGBA8 texture of seeded value noise: per channel, octave k
    has a (base * 2^k + 1)^2 lattice of PCG32 values, smoothstep-weighted
    bilinear interpolation, amplitude 2^-k, normalised to [0, 255]; alpha 255"""
    rng = PCG32(seed)
    c = (np.arange(n) + 0.5) / n
    out = np.empty((n, n, 4), dtype=np.uint8)
    out[..., 3] = 255
    for ch in range(3):
        total = np.zeros((n, n), dtype=np.float64)
        for k in range(octaves):
            f = base * 2 ** k
            lat = np.array([rng.uniform() for _ in range((f + 1) ** 2)], dtype=np.float64).reshape(f + 1, f + 1)
            x = c * f
            i = np.minimum(np.floor(x).astype(np.int64), f - 1)
            t = x - i
            w = t * t * (3 - 2 * t)
            iy, ix = np.ix_(i, i)
            wy, wx = np.ix_(w, w)
            v = (lat[iy, ix] * (1 - wy) * (1 - wx) + lat[iy, ix + 1] * (1 - wy) * wx
                 + lat[iy + 1, ix] * wy * (1 - wx) + lat[iy + 1, ix + 1] * wy * wx)
            total += v * 0.5 ** k
        lo, hi = total.min(), total.max()
        out[..., ch] = np.clip(np.round((total - lo) / max(hi - lo, 1e-12) * 255.0), 0, 255).astype(np.uint8)
    return out


# ------------------------------------------------------------ synthetic scenes


def cornell(seed=1):
    """Cornell-box-style scene: 5 axis-aligned wall quads + 2 rotated boxes
    (34 triangles) + 1 quad mesh light (17, 12, 4); Lambert albedo per face
    from U[0.1, 0.8]^3 (PCG32, seed 1)."""
    rng = PCG32(seed)
    s = Scene()

    def albedo():
        return [float(rng.range(0.1, 0.8)) for _ in range(3)]

    walls = [
        quad((-1, 0, 1), (1, 0, 1), (1, 0, -1), (-1, 0, -1)),        # floor, +y
        quad((-1, 2, -1), (1, 2, -1), (1, 2, 1), (-1, 2, 1)),        # ceiling, -y
        quad((-1, 0, -1), (1, 0, -1), (1, 2, -1), (-1, 2, -1)),      # back, +z
        quad((-1, 0, 1), (-1, 0, -1), (-1, 2, -1), (-1, 2, 1)),      # left, +x
        quad((1, 0, -1), (1, 0, 1), (1, 2, 1), (1, 2, -1)),          # right, -x
    ]
    for w in walls:
        s.add_instance(s.add_geometry(w), s.add_diffuse(albedo()))
    boxes = [((0.6, 0.6, 0.6), translate(0.35, 0.3, 0.3) @ rotate_y(math.radians(-18))),
             ((0.6, 1.2, 0.6), translate(-0.35, 0.6, -0.3) @ rotate_y(math.radians(15)))]
    for size, xf in boxes:
        for f in box_faces(*size):
            s.add_instance(s.add_geometry(f), s.add_diffuse(albedo()), xf)
    light = quad((-0.25, 1.99, -0.25), (0.25, 1.99, -0.25), (0.25, 1.99, 0.25), (-0.25, 1.99, 0.25))
    s.add_instance(s.add_geometry(light), s.add_diffuse([0.0, 0.0, 0.0]), light=(17.0, 12.0, 4.0))
    cam = Camera((0.0, 1.0, 3.4), (0.0, 1.0, 0.0), (0.0, 1.0, 0.0), 40.0)
    return s, cam


def fbm_grid(n=128, octaves=5, seed=3, max_sigma=2.0, base=4):
    """n^3 density grid [z, y, x] of seeded fBm value noise: octave k has a
    (base * 2^k + 1)^3 lattice of PCG32 values in [0, 1), smoothstep-weighted
    trilinear interpolation, amplitude 2^-k; the sum is scaled so that its
    maximum is max_sigma."""
    rng = PCG32(seed)
    c = (np.arange(n) + 0.5) / n
    total = np.zeros((n, n, n), dtype=np.float64)
    for k in range(octaves):
        f = base * 2 ** k
        lat = np.array([rng.uniform() for _ in range((f + 1) ** 3)], dtype=np.float64).reshape(f + 1, f + 1, f + 1)
        x = c * f
        i = np.minimum(np.floor(x).astype(np.int64), f - 1)
        t = x - i
        w = t * t * (3 - 2 * t)
        iz, iy, ix = np.ix_(i, i, i)
        wz, wy, wx = np.ix_(w, w, w)
        v = 0.0
        for dz in (0, 1):
            for dy in (0, 1):
                for dx in (0, 1):
                    v = v + (lat[iz + dz, iy + dy, ix + dx] * (wz if dz else 1 - wz) * (wy if dy else 1 - wy)
                             * (wx if dx else 1 - wx))
        total += v * 0.5 ** k
    return (total * (max_sigma / total.max())).astype(np.float32)


def save_grid_text(path, grid):
    """Array::load_from_text layout (utils/cuda/array.cpp:313-385): header
    "float C W H D", then the values with x fastest, then y, then z"""
    g = np.asarray(grid, dtype=np.float32)
    ch = 1 if g.ndim == 3 else g.shape[3]
    d, h, w = g.shape[:3]
    with open(path, "w") as f:
        f.write(f"float {ch} {w} {h} {d}\n")
        np.savetxt(f, g.reshape(-1, ch), fmt="%.9g")


def load_grid_text(path):
    """inverse of save_grid_text; also reads "byte" grids (values / 255 as
    the SNorm8 texture would not: bytes are returned as integers)"""
    with open(path) as f:
        comp, ch, w, h, d = f.readline().split()
        ch, w, h, d = int(ch), int(w), int(h), int(d)
        vals = np.loadtxt(f, dtype=np.float32 if comp == "float" else np.int64).reshape(-1)
    if comp not in ("float", "byte"):
        raise ValueError(f"unknown component type {comp}")
    if ch not in (1, 3, 4):
        raise ValueError(f"unsupported channel count {ch}")
    g = vals.reshape(d, h, w, ch)
    return g[..., 0] if ch == 1 else g


def cornell_media(scale=30.0, grid=None, g=0.6, albedo=0.8, grid_n=128, seed=3):
    """The Cornell-style scene scaled by `scale` (room 60^3 at 30) holding a
    40^3 invisible box centred in the room, filled with a heterogeneous
    medium: `grid` (default: 128^3 seeded 5-octave fBm, max sigma_t 2.0),
    albedo 0.8, Henyey-Greenstein g = 0.6."""
    base, cam = cornell()
    s = Scene()
    s.geometries = base.geometries
    s.materials = base.materials
    sm = np.diag([scale, scale, scale, 1.0])
    for inst in base.instances:
        s.add_instance(inst["geometry"], inst["material"], sm @ inst["mat"], light=inst["light"])
    if grid is None:
        grid = fbm_grid(grid_n, seed=seed)
    half = 20.0
    centre = (0.0, scale, 0.0)
    w2l = np.diag([1 / half, 1 / half, 1 / half, 1.0]) @ translate(-centre[0], -centre[1], -centre[2])
    med = s.add_heterogeneous(grid, (albedo, albedo, albedo), g=g, world_to_local=w2l)
    inv = s.add_invisible()
    box = np.concatenate(box_faces(2 * half, 2 * half, 2 * half))
    s.add_instance(s.add_geometry(box), inv, translate(*centre), inner_medium=med)
    e = np.array(cam.eye) * scale
    t = np.array(cam.dst) * scale
    return s, Camera(tuple(e), tuple(t), cam.up, cam.fov_y_deg)


def sky_texture(width=128, height=64, seed=0, sun=(0.3, 0.2), sun_radius=0.04, sun_radiance=40.0):
    """(height, width, 4) float32 latitude-longitude radiance map for an IBL:
    a blue-to-white sky over a darker ground, seeded value noise, and a
    bright disc (the sun) centred at texture coordinate `sun`"""
    rng = np.random.default_rng(seed)
    v = (np.arange(height) + 0.5) / height
    u = (np.arange(width) + 0.5) / width
    uu, vv = np.meshgrid(u, v)
    sky = np.where(vv[..., None] < 0.5,
                   np.array([0.35, 0.55, 1.0]) * (0.6 + 0.8 * vv[..., None]),
                   np.array([0.25, 0.2, 0.15]) * (1.4 - vv[..., None]))
    noise = rng.uniform(0.8, 1.2, (height // 4 + 1, width // 4 + 1))
    sky = sky * np.kron(noise, np.ones((4, 4)))[:height, :width, None]
    du = np.minimum(np.abs(uu - sun[0]), 1.0 - np.abs(uu - sun[0]))
    disc = du ** 2 + (vv - sun[1]) ** 2 < sun_radius ** 2
    sky[disc] = sun_radiance * np.array([1.0, 0.9, 0.7])
    out = np.ones((height, width, 4), np.float32)
    out[..., :3] = sky
    return out


def sphere_field(instances=4096, lights=64, seed=2, extent=100.0, envir=(0.5, 0.5, 0.5)):
    """Procedural field: `instances` seeded placements (position U[0, extent]^3,
    uniform rotation, scale U[0.5, 2]) of one 256-triangle displaced sphere,
    each with its own Disney material (roughness U[0.05, 1], metallic
    Bernoulli(0.3), specular 0.5, base colour U[0, 1]^3); `lights` sphere mesh
    lights; a constant environment (gradient sky with upper == lower)."""
    rng = PCG32(seed)
    s = Scene()
    g = s.add_geometry(uv_sphere(rng, displacement=0.15))
    for _ in range(instances):
        pos = [float(rng.range(0.0, extent)) for _ in range(3)]
        rot = uniform_rotation(rng)
        sc = float(rng.range(0.5, 2.0))
        mat = s.add_disney(color=[float(rng.uniform()) for _ in range(3)],
                           metallic=1.0 if float(rng.uniform()) < 0.3 else 0.0,
                           roughness=float(rng.range(0.05, 1.0)), specular=0.5)
        s.add_instance(g, mat, translate(*pos) @ rot @ scale(sc))
    lg = s.add_geometry(uv_sphere())
    black = s.add_diffuse([0.0, 0.0, 0.0])
    for _ in range(lights):
        pos = [float(rng.range(0.0, extent)) for _ in range(3)]
        power = float(rng.range(5.0, 20.0))
        tint = [float(rng.range(0.5, 1.0)) for _ in range(3)]
        s.add_instance(lg, black, translate(*pos) @ scale(1.0), light=tuple(power * t for t in tint))
    s.set_gradient_sky((0.0, 1.0, 0.0), envir, envir)
    c = extent / 2
    cam = Camera((c, c, -0.6 * extent), (c, c, c), (0.0, 1.0, 0.0), 60.0)
    return s, cam


def instanced_field(instances=65536, lights=64, seed=2, extent=250.0, geometries=4, textures=8, texture_size=1024,
                    envir=(0.5, 0.5, 0.5)):
    """Textured instanced field: `geometries` 256-triangle displaced spheres
    with texture coordinates, `instances` seeded placements (position
    U[0, extent]^3, uniform rotation, scale U[0.5, 2]) each with its own
    Disney material (roughness U[0.05, 1], metallic Bernoulli(0.3), specular
    0.5) whose base colour is one of `textures` seeded RGBA8 noise textures
    (texture_size^2, wrapped in u); `lights` sphere mesh lights; a constant
    environment. Traced through the two-level structure. Returns the scene
    and eight views of 512x512 up to 1920x1080 looking into the field."""
    rng = PCG32(seed)
    s = Scene()
    s.accel = _abi.ACCEL_TWO_LEVEL
    uvs = uv_sphere_uvs()
    geos = [s.add_geometry(uv_sphere(rng, displacement=0.15), uvs=uvs) for _ in range(geometries)]
    texs = [s.add_texture(noise_texture(texture_size, seed=seed * 100 + k), address_u="wrap") for k in range(textures)]
    for k in range(instances):
        pos = [float(rng.range(0.0, extent)) for _ in range(3)]
        rot = uniform_rotation(rng)
        sc = float(rng.range(0.5, 2.0))
        mat = s.add_disney(color=[1.0, 1.0, 1.0], metallic=1.0 if float(rng.uniform()) < 0.3 else 0.0,
                           roughness=float(rng.range(0.05, 1.0)), specular=0.5, color_texture=texs[k % textures])
        s.add_instance(geos[k % geometries], mat, translate(*pos) @ rot @ scale(sc))
    lg = s.add_geometry(uv_sphere())
    black = s.add_diffuse([0.0, 0.0, 0.0])
    for _ in range(lights):
        pos = [float(rng.range(0.0, extent)) for _ in range(3)]
        power = float(rng.range(5.0, 20.0))
        tint = [float(rng.range(0.5, 1.0)) for _ in range(3)]
        s.add_instance(lg, black, translate(*pos) @ scale(1.0), light=tuple(power * t for t in tint))
    s.set_gradient_sky((0.0, 1.0, 0.0), envir, envir)
    c = extent / 2
    sizes = [(512, 512), (640, 480), (800, 600), (1024, 768), (1280, 720), (1280, 1024), (1600, 900), (1920, 1080)]
    views = []
    for k, (w, h) in enumerate(sizes):
        a = 2 * math.pi * k / len(sizes)
        eye = (c + 0.8 * extent * math.cos(a), c + 0.1 * extent * (k % 3 - 1), c + 0.8 * extent * math.sin(a))
        views.append((Camera(eye, (c, c, c), (0.0, 1.0, 0.0), 60.0), w, h))
    return s, views


def load_obj_triangles(path):
    """Triangles of an OBJ file the way lib/bvh's test loader reads it
    (lib/bvh/test/obj.hpp:62-96: fan triangulation (p0, p1, v), p1 <- v),
    with strtof-exact vertex parsing."""
    libc = C.CDLL(None)
    libc.strtof.restype = C.c_float
    libc.strtof.argtypes = [C.c_char_p, C.c_void_p]
    verts, tris = [], []
    with open(path, "r") as f:
        for line in f:
            t = line.strip()
            if not t or t.startswith("#"):
                continue
            parts = t.split()
            if parts[0] == "v":
                verts.append([libc.strtof(p.encode(), None) for p in parts[1:4]])
            elif parts[0] == "f":
                idx = []
                for p in parts[1:]:
                    i = int(p.split("/")[0])
                    idx.append(len(verts) + i if i < 0 else i - 1)
                p0, p1 = verts[idx[0]], verts[idx[1]]
                for k in idx[2:]:
                    tris.append([*p0, *p1, *verts[k]])
                    p1 = verts[k]
    return np.array(tris, dtype=np.float32)
