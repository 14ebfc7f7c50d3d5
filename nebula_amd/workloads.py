"""Synthetic inputs of the tracking path: triangle geometries, primary beams and
material tables, plus writers for Nebula's file formats.

Formats (reference: io/load_tri_file.cpp:8-60, io/load_pri_file.cpp:72-121):
  .tri  text, one triangle per line: ``in out x0 y0 z0 x1 y1 z1 x2 y2 z2`` (nm),
        ``#`` comments allowed.
  .pri  binary little-endian records of 7 float32 (pos, dir, E in eV) and 2
        int32 (pixel x, y): 36 bytes per primary electron.

Every generator is deterministic for a given seed.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# Special material indices (core/material_manager.h:77-86).
NOP, TERMINATOR, DETECTOR, DETECTOR_LT50, DETECTOR_GE50, VACUUM, MIRROR = (
    -128, -127, -126, -125, -124, -123, -122)

PRI_DTYPE = np.dtype([("pos", "<f4", 3), ("dir", "<f4", 3), ("E", "<f4"),
                      ("px", "<i4"), ("py", "<i4")])
assert PRI_DTYPE.itemsize == 36


# --------------------------------------------------------------- geometry --
@dataclass
class Mesh:
    """Triangles as vertex arrays: r0, r1, r2 (n x 3 float32) and materials."""
    r0: np.ndarray
    r1: np.ndarray
    r2: np.ndarray
    mat_in: np.ndarray
    mat_out: np.ndarray

    @property
    def n(self) -> int:
        return len(self.mat_in)

    def vertices(self) -> np.ndarray:
        """n x 9 float32 array (r0, r1, r2)."""
        return np.concatenate([self.r0, self.r1, self.r2], axis=1).astype(np.float32)

    @staticmethod
    def concat(meshes) -> "Mesh":
        return Mesh(*[np.concatenate([getattr(m, k) for m in meshes])
                      for k in ("r0", "r1", "r2", "mat_in", "mat_out")])


class _MeshBuilder:
    def __init__(self):
        self.tris = []

    def tri(self, a, b, c, mat_in, mat_out):
        """Triangle a-b-c; its normal (b-a)x(c-a) points from mat_in to mat_out."""
        self.tris.append((a, b, c, mat_in, mat_out))

    def quad(self, a, b, c, d, mat_in, mat_out):
        """Planar quad a-b-c-d (counter-clockwise seen from the mat_out side)."""
        self.tri(a, b, c, mat_in, mat_out)
        self.tri(a, c, d, mat_in, mat_out)

    def mesh(self) -> Mesh:
        a = np.array([t[0] for t in self.tris], dtype=np.float64)
        b = np.array([t[1] for t in self.tris], dtype=np.float64)
        c = np.array([t[2] for t in self.tris], dtype=np.float64)
        return Mesh(a.astype(np.float32), b.astype(np.float32), c.astype(np.float32),
                    np.array([t[3] for t in self.tris], dtype=np.int32),
                    np.array([t[4] for t in self.tris], dtype=np.int32))


def _box(mb: _MeshBuilder, lo, hi, mat_in, mat_out):
    """Axis-aligned box with outward normals, 12 triangles."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    p = lambda x, y, z: (x, y, z)  # noqa: E731
    # -z, +z, -x, +x, -y, +y faces (counter-clockwise seen from outside)
    mb.quad(p(x0, y0, z0), p(x0, y1, z0), p(x1, y1, z0), p(x1, y0, z0), mat_in, mat_out)
    mb.quad(p(x0, y0, z1), p(x1, y0, z1), p(x1, y1, z1), p(x0, y1, z1), mat_in, mat_out)
    mb.quad(p(x0, y0, z0), p(x0, y0, z1), p(x0, y1, z1), p(x0, y1, z0), mat_in, mat_out)
    mb.quad(p(x1, y0, z0), p(x1, y1, z0), p(x1, y1, z1), p(x1, y0, z1), mat_in, mat_out)
    mb.quad(p(x0, y0, z0), p(x1, y0, z0), p(x1, y0, z1), p(x0, y0, z1), mat_in, mat_out)
    mb.quad(p(x0, y1, z0), p(x0, y1, z1), p(x1, y1, z1), p(x1, y1, z0), mat_in, mat_out)


def _detector_plane(mb: _MeshBuilder, x0, x1, y0, y1, z):
    mb.quad((x0, y0, z), (x1, y0, z), (x1, y1, z), (x0, y1, z), DETECTOR, DETECTOR)


def slab_geometry(size_nm: float = 1000.0, detector_z: float = 10.0) -> Mesh:
    """Flat slab: a size^3 box of material 0 with its top face at z=0
    (x, y in [-size/2, size/2], z in [-size, 0]), vacuum around it, and a
    two-triangle detector plane at z=detector_z covering the top."""
    mb = _MeshBuilder()
    h = size_nm / 2
    _box(mb, (-h, -h, -size_nm), (h, h, 0.0), 0, VACUUM)
    _detector_plane(mb, -h, h, -h, h, detector_z)
    return mb.mesh()


def special_geometry(half: float = 40.0, depth: float = 60.0) -> Mesh:
    """A block of material 0 (x, y in [-half, half], z in [-depth, 0]) under
    planes of every special outgoing material of core/material_manager.h:
    77-86, so that tracking meets each branch of boundary_intersect.h:77-100:

    - NOP planes inside the block (z = -4) and above it (z = +3, x > 0):
      the electron stops on them and goes on unchanged;
    - a DETECTOR_LT50 plane at z = +6 (detects electrons below 50 eV, lets
      the others through);
    - at z = +12 a DETECTOR_GE50 plane over x < 0 (detects 50 eV and above)
      and a MIRROR over x > 0 (reflects everything back down);
    - TERMINATOR walls at x = +-half and y = +-half between z = 0 and +20;
    - a DETECTOR plane at z = +20 over everything.

    Every plane has the same special material on both sides."""
    mb = _MeshBuilder()
    h = half
    _box(mb, (-h, -h, -depth), (h, h, 0.0), 0, VACUUM)

    def plane(x0, x1, y0, y1, z, m):
        mb.quad((x0, y0, z), (x1, y0, z), (x1, y1, z), (x0, y1, z), m, m)

    plane(-h, h, -h, h, -4.0, NOP)
    plane(0.0, h, -h, h, 3.0, NOP)
    plane(-h, h, -h, h, 6.0, DETECTOR_LT50)
    plane(-h, 0.0, -h, h, 12.0, DETECTOR_GE50)
    plane(0.0, h, -h, h, 12.0, MIRROR)
    plane(-h, h, -h, h, 20.0, DETECTOR)
    T = TERMINATOR
    mb.quad((-h, -h, 0.0), (-h, h, 0.0), (-h, h, 20.0), (-h, -h, 20.0), T, T)
    mb.quad((h, -h, 0.0), (h, -h, 20.0), (h, h, 20.0), (h, h, 0.0), T, T)
    mb.quad((-h, -h, 0.0), (-h, -h, 20.0), (h, -h, 20.0), (h, -h, 0.0), T, T)
    mb.quad((-h, h, 0.0), (h, h, 0.0), (h, h, 20.0), (-h, h, 20.0), T, T)
    return mb.mesh()


def gaussian_random_field(shape, spacing, sigma, corr_len, rng):
    """Zero-mean Gaussian field with standard deviation `sigma` and Gaussian
    autocorrelation exp(-r^2/corr_len^2), by filtering white noise in Fourier
    space.  `shape` and `spacing` are per axis."""
    noise = rng.standard_normal(shape)
    freqs = np.meshgrid(*[np.fft.fftfreq(n, d) for n, d in zip(shape, spacing)], indexing="ij")
    k2 = sum((2 * np.pi * f) ** 2 for f in freqs)
    filt = np.exp(-k2 * corr_len ** 2 / 8.0)
    field = np.real(np.fft.ifftn(np.fft.fftn(noise) * filt))
    std = field.std()
    return field * (sigma / std) if std > 0 else field


def grating_geometry(pitch=32.0, cd=16.0, height=32.0, sidewall_deg=88.0,
                     n_lines=10, length=320.0, substrate_depth=200.0,
                     rough_sigma=1.0, rough_corr=10.0, dy=2.0, nz=4,
                     detector_gap=10.0, seed=1,
                     line_mat=1, substrate_mat=0) -> Mesh:
    """Trapezoid lines (material `line_mat`) along y on a substrate
    (`substrate_mat`) whose top face is z=0; sidewalls carry Gaussian roughness
    (x displacement, seeded); a detector plane sits `detector_gap` nm above the
    line tops.  The scan field is x in [-n_lines*pitch/2, +], y in
    [-length/2, +]."""
    rng = np.random.default_rng(seed)
    mb = _MeshBuilder()
    X = n_lines * pitch / 2
    Y = length / 2
    ny = int(round(length / dy))
    ys = np.linspace(-Y, Y, ny + 1)
    zs = np.linspace(0.0, height, nz + 1)
    dx_side = height / math.tan(math.radians(sidewall_deg))
    centers = -X + pitch * (np.arange(n_lines) + 0.5)
    # wall x positions: [line][side][iy][iz]
    walls = np.zeros((n_lines, 2, ny + 1, nz + 1))
    for i, xc in enumerate(centers):
        for side, sgn in ((0, -1.0), (1, +1.0)):
            rough = gaussian_random_field((ny + 1, nz + 1), (dy, height / nz),
                                          rough_sigma, rough_corr, rng)
            # CD at mid-height; the bottom is wider by dx_side/2 per side
            base = xc + sgn * (cd / 2 + dx_side * (0.5 - zs / height))
            walls[i, side] = base[None, :] + rough
    P = lambda x, y, z: (float(x), float(y), float(z))  # noqa: E731
    L, S, V = line_mat, substrate_mat, VACUUM
    for i in range(n_lines):
        for j in range(ny):
            y0, y1 = ys[j], ys[j + 1]
            for k in range(nz):
                z0, z1 = zs[k], zs[k + 1]
                # left wall: normal towards -x (outside), mat_in = line
                a = P(walls[i, 0, j, k], y0, z0)
                b = P(walls[i, 0, j, k + 1], y0, z1)
                c = P(walls[i, 0, j + 1, k + 1], y1, z1)
                d = P(walls[i, 0, j + 1, k], y1, z0)
                mb.quad(a, b, c, d, L, V)
                # right wall: normal towards +x
                a = P(walls[i, 1, j, k], y0, z0)
                b = P(walls[i, 1, j + 1, k], y1, z0)
                c = P(walls[i, 1, j + 1, k + 1], y1, z1)
                d = P(walls[i, 1, j, k + 1], y0, z1)
                mb.quad(a, b, c, d, L, V)
            # top face of the line
            a = P(walls[i, 0, j, nz], y0, height)
            b = P(walls[i, 1, j, nz], y0, height)
            c = P(walls[i, 1, j + 1, nz], y1, height)
            d = P(walls[i, 0, j + 1, nz], y1, height)
            mb.quad(a, b, c, d, L, V)
            # interface under the line: normal towards -z (into substrate)
            a = P(walls[i, 0, j, 0], y0, 0.0)
            b = P(walls[i, 0, j + 1, 0], y1, 0.0)
            c = P(walls[i, 1, j + 1, 0], y1, 0.0)
            d = P(walls[i, 1, j, 0], y0, 0.0)
            mb.quad(a, b, c, d, L, S)
            # substrate top between this line and the next (or the field edge)
            xr0, xr1 = walls[i, 1, j, 0], walls[i, 1, j + 1, 0]
            if i + 1 < n_lines:
                xl0, xl1 = walls[i + 1, 0, j, 0], walls[i + 1, 0, j + 1, 0]
            else:
                xl0 = xl1 = X
            mb.quad(P(xr0, y0, 0), P(xl0, y0, 0), P(xl1, y1, 0), P(xr1, y1, 0), S, V)
            if i == 0:
                xl0, xl1 = walls[0, 0, j, 0], walls[0, 0, j + 1, 0]
                mb.quad(P(-X, y0, 0), P(xl0, y0, 0), P(xl1, y1, 0), P(-X, y1, 0), S, V)
        # end caps of the line at y = -Y and y = +Y (fans over the wall profile)
        for side_y, jy, sgn in ((-Y, 0, -1.0), (Y, ny, 1.0)):
            poly = [P(walls[i, 0, jy, k], side_y, zs[k]) for k in range(nz + 1)]
            poly += [P(walls[i, 1, jy, k], side_y, zs[k]) for k in range(nz, -1, -1)]
            # polygon order: up the left wall, down the right wall
            for k in range(1, len(poly) - 1):
                if sgn > 0:
                    mb.tri(poly[0], poly[k], poly[k + 1], L, V)
                else:
                    mb.tri(poly[0], poly[k + 1], poly[k], L, V)
    # substrate: bottom and four sides (top is built above)
    D = substrate_depth
    mb.quad(P(-X, -Y, -D), P(-X, Y, -D), P(X, Y, -D), P(X, -Y, -D), S, V)
    mb.quad(P(-X, -Y, -D), P(-X, -Y, 0), P(-X, Y, 0), P(-X, Y, -D), S, V)
    mb.quad(P(X, -Y, -D), P(X, Y, -D), P(X, Y, 0), P(X, -Y, 0), S, V)
    mb.quad(P(-X, -Y, -D), P(X, -Y, -D), P(X, -Y, 0), P(-X, -Y, 0), S, V)
    mb.quad(P(-X, Y, -D), P(-X, Y, 0), P(X, Y, 0), P(X, Y, -D), S, V)
    _detector_plane(mb, -X, X, -Y, Y, height + detector_gap)
    return mb.mesh()


def write_tri(path: str, mesh: Mesh) -> None:
    """Write a .tri text file; coordinates as shortest round-trip float32."""
    v = mesh.vertices()
    with open(path, "w") as f:
        f.write("# in out x0 y0 z0 x1 y1 z1 x2 y2 z2 (nm)\n")
        for i in range(mesh.n):
            vals = " ".join(repr(float(np.float32(x))) for x in v[i])
            f.write(f"{int(mesh.mat_in[i])} {int(mesh.mat_out[i])} {vals}\n")


def read_tri(path: str) -> Mesh:
    """Parse a .tri file with the reference's rules (io/load_tri_file.cpp:8-60):
    blank and '#' lines skipped, exactly 11 columns else ValueError."""
    rows = []
    with open(path) as f:
        for ln, line in enumerate(f, start=2):
            s = line.strip()
            if not s or s.startswith("#"):
                continue
            cols = s.split()
            if len(cols) != 11:
                raise ValueError(f"invalid number of columns in line '{ln}'")
            rows.append([int(cols[0]), int(cols[1])] + [float(c) for c in cols[2:]])
    a = np.array(rows, dtype=np.float64).reshape(-1, 11)
    v = a[:, 2:].astype(np.float32)
    return Mesh(v[:, 0:3].copy(), v[:, 3:6].copy(), v[:, 6:9].copy(),
                a[:, 0].astype(np.int32), a[:, 1].astype(np.int32))


def contact_hole_geometry(nx: int = 4, ny: int = 4, pitch: float = 80.0, diameter: float = 40.0,
                          layers=((80.0, 1),), substrate_mat: int = 0, substrate_depth: float = 200.0,
                          n_seg: int = 32, nz: int = 1, radius_sigma: float = 0.0,
                          detector_gap: float = 10.0, seed: int = 1) -> Mesh:
    """An nx x ny array of cylindrical holes (vacuum) through a stack of
    layers [(thickness, material), ...] from z=0 down, on a substrate.  Each
    hole is an n_seg-gon (n_seg a multiple of 4) whose vertex radii carry
    Gaussian jitter `radius_sigma` (seeded), straight through the stack; each
    layer's sidewall is split into `nz` bands.  Every cell of side `pitch`
    meshes its faces as a ring between the square cell edge and the hole, so
    neighbouring cells share edges exactly."""
    if n_seg % 4:
        raise ValueError("n_seg must be a multiple of 4")
    rng = np.random.default_rng(seed)
    mb = _MeshBuilder()
    X, Y = nx * pitch / 2, ny * pitch / 2
    V, S = VACUUM, substrate_mat
    # layer boundaries: z_top[0] = 0 > ... > z_bottom[-1] = -depth
    tops, mats_l = [], []
    z = 0.0
    for t, m in layers:
        tops.append((z, z - t))
        mats_l.append(m)
        z -= t
    depth = -z
    ang = np.pi / 4 + 2 * np.pi * np.arange(n_seg) / n_seg     # includes the cell corners
    c, sn = np.cos(ang), np.sin(ang)
    edge = (pitch / 2) / np.maximum(np.abs(c), np.abs(sn))      # ray to the cell edge
    P = lambda x, y, z: (float(x), float(y), float(z))  # noqa: E731
    for ix in range(nx):
        for iy in range(ny):
            cx, cy = -X + pitch * (ix + 0.5), -Y + pitch * (iy + 0.5)
            r = diameter / 2 + radius_sigma * rng.standard_normal(n_seg)
            r = np.clip(r, 0.05 * diameter, 0.45 * pitch)
            hx, hy = cx + r * c, cy + r * sn
            bx, by = cx + edge * c, cy + edge * sn
            k1 = np.roll(np.arange(n_seg), -1)
            # rings: top face, then each interface below a layer (normal +z,
            # from the lower material to the upper one)
            ring_mats = [(mats_l[0], V, 0.0)]
            for li in range(len(layers)):
                below = mats_l[li + 1] if li + 1 < len(layers) else S
                ring_mats.append((below, mats_l[li], tops[li][1]))
            for m_in, m_out, zz in ring_mats:
                for k in range(n_seg):
                    j = k1[k]
                    mb.quad(P(hx[k], hy[k], zz), P(bx[k], by[k], zz), P(bx[j], by[j], zz),
                            P(hx[j], hy[j], zz), m_in, m_out)
            # sidewalls, normal into the hole
            for li, (z_hi, z_lo) in enumerate(tops):
                zs = np.linspace(z_lo, z_hi, nz + 1)
                for b in range(nz):
                    z0, z1 = zs[b], zs[b + 1]
                    for k in range(n_seg):
                        j = k1[k]
                        mb.quad(P(hx[k], hy[k], z0), P(hx[k], hy[k], z1), P(hx[j], hy[j], z1),
                                P(hx[j], hy[j], z0), mats_l[li], V)
            # hole bottom: substrate under vacuum
            for k in range(n_seg):
                j = k1[k]
                mb.tri(P(cx, cy, -depth), P(hx[k], hy[k], -depth), P(hx[j], hy[j], -depth), S, V)
    # outer faces of every layer and of the substrate, and the substrate bottom
    D = depth + substrate_depth
    for (z_hi, z_lo), m in list(zip(tops, mats_l)) + [((-depth, -D), S)]:
        mb.quad(P(-X, -Y, z_lo), P(-X, -Y, z_hi), P(-X, Y, z_hi), P(-X, Y, z_lo), m, V)
        mb.quad(P(X, -Y, z_lo), P(X, Y, z_lo), P(X, Y, z_hi), P(X, -Y, z_hi), m, V)
        mb.quad(P(-X, -Y, z_lo), P(X, -Y, z_lo), P(X, -Y, z_hi), P(-X, -Y, z_hi), m, V)
        mb.quad(P(-X, Y, z_lo), P(-X, Y, z_hi), P(X, Y, z_hi), P(X, Y, z_lo), m, V)
    mb.quad(P(-X, -Y, -D), P(-X, Y, -D), P(X, Y, -D), P(X, -Y, -D), S, V)
    _detector_plane(mb, -X, X, -Y, Y, detector_gap)
    return mb.mesh()


# -------------------------------------------------------------- primaries --
def pencil_beam(n: int, energy: float = 1000.0, z0: float = 5.0) -> np.ndarray:
    """n primaries at (0, 0, z0) moving along -z, pixel (0, 0)."""
    pri = np.zeros(n, dtype=PRI_DTYPE)
    pri["pos"][:, 2] = z0
    pri["dir"][:, 2] = -1.0
    pri["E"] = energy
    return pri


def raster_scan(nx: int, ny: int, per_pixel: int, energy: float, x0: float, x1: float,
                y0: float, y1: float, z: float, spot_sigma: float = 1.0,
                seed: int = 1) -> np.ndarray:
    """nx*ny pixel raster over [x0,x1]x[y0,y1] (pixel centres), `per_pixel`
    primaries per pixel drawn from a Gaussian spot of std `spot_sigma`,
    starting at height z and moving along -z.  Pixel-major order."""
    rng = np.random.default_rng(seed)
    xs = x0 + (np.arange(nx) + 0.5) * (x1 - x0) / nx
    ys = y0 + (np.arange(ny) + 0.5) * (y1 - y0) / ny
    n = nx * ny * per_pixel
    pri = np.zeros(n, dtype=PRI_DTYPE)
    ix = np.repeat(np.tile(np.arange(nx), ny), per_pixel)
    iy = np.repeat(np.repeat(np.arange(ny), nx), per_pixel)
    off = rng.standard_normal((n, 2)) * spot_sigma
    pri["pos"][:, 0] = xs[ix] + off[:, 0]
    pri["pos"][:, 1] = ys[iy] + off[:, 1]
    pri["pos"][:, 2] = z
    pri["dir"][:, 2] = -1.0
    pri["E"] = energy
    pri["px"] = ix
    pri["py"] = iy
    return pri


def write_pri(path: str, pri: np.ndarray) -> None:
    pri.astype(PRI_DTYPE).tofile(path)


def read_pri(path: str) -> np.ndarray:
    """Read a .pri file; a size that is not a multiple of 36 bytes raises
    ValueError("Unexpected end of primaries file") (io/load_pri_file.cpp:107)."""
    raw = np.fromfile(path, dtype=np.uint8)
    if raw.size % 36:
        raise ValueError("Unexpected end of primaries file")
    return raw.view(PRI_DTYPE)


def batch_scans(n_scans: int = 64, seed: int = 1, per_pixel_scale: float = 1.0):
    """Independent scans of uneven size: each a seeded random geometry (a line
    grating or a contact-hole array, 1e3 to 2e5 triangles, log-uniform), a
    beam energy log-uniform in 200 eV - 20 keV, a raster of 32x32 to 256x256
    pixels and 10 to 200 primaries per pixel (times `per_pixel_scale`).
    Materials: 0 = Si substrate, 1 = PMMA (lines or hole layer).  Returns a
    list of dicts with kind, mesh, primaries, energy, raster side, per_pixel
    and target triangle count."""
    rng = np.random.default_rng(seed)
    scans = []
    for i in range(n_scans):
        target = float(np.exp(rng.uniform(np.log(1e3), np.log(2e5))))
        energy = float(np.exp(rng.uniform(np.log(200.0), np.log(20000.0))))
        side = int(rng.integers(32, 257))
        per_px = max(1, int(round(int(rng.integers(10, 201)) * per_pixel_scale)))
        gseed = int(rng.integers(1 << 30))
        if rng.random() < 0.5:
            kind = "grating"
            n_lines, nz = 10, 4
            ny = max(2, int(round(target / (n_lines * (4 * nz + 6)))))
            mesh = grating_geometry(n_lines=n_lines, nz=nz, dy=320.0 / ny, seed=gseed)
            half, top = 160.0, 32.0
        else:
            kind = "holes"
            nh = int(rng.integers(2, 11))
            n_seg = max(8, 4 * int(round(target / (9.0 * nh * nh) / 4)))
            mesh = contact_hole_geometry(nh, nh, pitch=80.0, diameter=40.0, layers=((80.0, 1),),
                                         n_seg=n_seg, nz=2, radius_sigma=1.0, seed=gseed)
            half, top = 40.0 * nh, 0.0
        f = 0.8 * half
        pri = raster_scan(side, side, per_px, energy, -f, f, -f, f, top + 5.0, spot_sigma=1.0, seed=gseed)
        scans.append({"kind": kind, "mesh": mesh, "primaries": pri, "energy": energy, "side": side,
                      "per_pixel": per_px, "target_triangles": int(target)})
    return scans


# -------------------------------------------------------------- materials --
HARTREE_EV = 27.211386
BOHR_NM = 0.0529177
E2_EVNM = 1.439964       # e^2 / (4 pi eps0) in eV nm
HBAR_SQRT2M = 0.19519    # hbar / sqrt(2 m_e) in nm eV^1/2
GRAM = 6.2415e15         # the reference's gram in internal units
AVOGADRO = 6.02214076e23


@dataclass
class MaterialSpec:
    """Parameters of a synthetic material.  Elastic: screened Rutherford for
    atomic number Z at `atom_density` (nm^-3).  Inelastic: single-pole Drude
    energy-loss function with plasmon energy `plasmon` and width `width` (eV),
    with inner shells (binding, fraction)."""
    name: str
    Z: float
    atom_density: float
    atomic_mass_u: float
    work_function: float
    fermi: float
    band_gap: float
    plasmon: float
    width: float
    phonon_loss: float
    inner_shells: tuple = ()
    outer_shells: tuple = ()


SI = MaterialSpec("silicon", Z=14, atom_density=49.94, atomic_mass_u=28.0855,
                  work_function=4.79, fermi=7.8, band_gap=1.1, plasmon=16.7, width=4.0,
                  phonon_loss=0.0159, inner_shells=((99.8, 0.30), (1839.0, 0.05)),
                  outer_shells=(13.5, 99.8, 149.7, 1839.0))
PMMA = MaterialSpec("pmma", Z=6, atom_density=50.1, atomic_mass_u=6.675,
                    work_function=2.5, fermi=3.0, band_gap=5.6, plasmon=19.9, width=8.0,
                    phonon_loss=0.1, inner_shells=((284.0, 0.25), (532.0, 0.10)),
                    outer_shells=(11.0, 14.0, 284.0, 532.0))
# a silicon-dioxide-like layer material (mean per atom: Z 10, 20.03 u; 2.2 g/cm3)
SIO2 = MaterialSpec("sio2", Z=10, atom_density=66.2, atomic_mass_u=20.03,
                    work_function=0.9, fermi=10.0, band_gap=9.0, plasmon=22.4, width=10.0,
                    phonon_loss=0.1, inner_shells=((99.8, 0.10), (532.0, 0.20)),
                    outer_shells=(12.0, 99.8, 150.0, 532.0))


@dataclass
class RawDataset:
    """A dataset as stored in a .mat file: float32 data, unit string and one
    optional dimension scale (float64 values + unit string) per axis."""
    data: np.ndarray
    unit: str
    scales: list = field(default_factory=list)   # [(values, unit) or None]


@dataclass
class RawMaterial:
    """The contents of a .mat file (io/hdf5_file.cpp conventions): quantity
    attributes {name: (value, unit)} and datasets {path: RawDataset}."""
    attrs: dict
    datasets: dict
    inel_model: int   # 0 full Penn, 1 Kieft & Bosch


def _elf(spec: MaterialSpec, w: np.ndarray) -> np.ndarray:
    wp, g = spec.plasmon, spec.width
    return wp * wp * g * w / ((w * w - wp * wp) ** 2 + (g * w) ** 2)


def _inelastic_tables(spec: MaterialSpec, K: np.ndarray, P: np.ndarray, nw: int = 2048):
    """Inverse IMFP (nm^-1) and omega ICDF (eV) in the optical approximation:
    d(imfp)/d(omega) = Im(-1/eps(omega)) ln[(sqrt(E)+sqrt(E-w))/(sqrt(E)-sqrt(E-w))]
    / (pi E) in atomic units, omega in (w_min, E - E_F]."""
    imfp = np.zeros(len(K))
    icdf = np.zeros((len(K), len(P)))
    for i, E in enumerate(K):
        wmax = E - spec.fermi
        wmin = 1e-2
        if wmax <= wmin * 1.01:
            imfp[i] = 1e-6
            icdf[i, :] = wmin
            continue
        w = np.geomspace(wmin, wmax, nw)
        sE, sEw = np.sqrt(E), np.sqrt(np.maximum(E - w, 0.0))
        lg = np.log((sE + sEw) / np.maximum(sE - sEw, 1e-30))
        dens = _elf(spec, w) * lg / (np.pi * (E / HARTREE_EV)) / HARTREE_EV / BOHR_NM
        cdf = np.concatenate([[0.0], np.cumsum(0.5 * (dens[1:] + dens[:-1]) * np.diff(w))])
        imfp[i] = max(cdf[-1], 1e-6)
        icdf[i] = np.interp(P, cdf / cdf[-1], w) if cdf[-1] > 0 else wmin
    return imfp, icdf


def _binding_icdf(spec: MaterialSpec, omega: np.ndarray, P: np.ndarray) -> np.ndarray:
    """Binding-energy ICDF for energy loss `omega` (any shape): a Fermi-sea
    fraction (binding 0, read as 'none'), then inner shells in increasing
    binding, each with weight frac*(1 - binding/omega) above its edge."""
    out = np.zeros(omega.shape + (len(P),))
    cum = np.zeros(omega.shape)
    weights = []
    for b, frac in spec.inner_shells:
        wgt = np.where(omega > b, frac * (1.0 - b / np.maximum(omega, 1e-30)), 0.0)
        weights.append((b, wgt))
        cum = cum + wgt
    sea = 1.0 - cum
    edge = sea.copy()
    for b, wgt in weights:
        out = np.where(P >= edge[..., None], b, out)
        edge = edge + wgt
    return out


def make_material(spec: MaterialSpec, model: str = "penn", nK: int = 128, nP: int = 1024,
                  nW: int = 64, nOmega: int = 32, Kmin: float = 10.0, Kmax: float = 50e3
                  ) -> RawMaterial:
    """Synthetic material tables on nK log-spaced energies x nP probabilities.

    model "penn": /elastic, /full_penn and a 3-D /ionization/binding_icdf
    (physics/full_penn.h:294-406).  model "kieft": /elastic, /inelastic_kieft,
    a 2-D /ionization/binding_icdf and /ionization/outer_shells
    (physics/kieft/inelastic.h:212-327)."""
    K = np.geomspace(Kmin, Kmax, nK)
    P = np.linspace(0.0, 1.0, nP)
    # elastic: screened Rutherford
    eta = 4.34 * spec.Z ** (2.0 / 3.0) / K
    sigma = np.pi * E2_EVNM ** 2 * spec.Z * (spec.Z + 1) / (4 * K * K * eta * (1 + eta))
    el_imfp = spec.atom_density * sigma
    U = P[None, :]
    e = eta[:, None]
    el_icdf = 1.0 - 2.0 * e * U / (1.0 + e - U)
    imfp, w_icdf = _inelastic_tables(spec, K, P)
    kscale = (K, "eV")
    pscale = (P, "")
    ds = {
        "/elastic/imfp": RawDataset(el_imfp.astype(np.float32), "nm^-1", [kscale]),
        "/elastic/costheta_icdf": RawDataset(el_icdf.astype(np.float32), "", [kscale, pscale]),
    }
    if model == "penn":
        Wr = np.linspace(0.0, 1.0, nW)
        Orat = np.geomspace(1e-4, 1.0, nOmega)
        # q ICDF: p(q) ~ 1/q between q- and q+ (units eV^1/2), stored in nm^-1
        Eg = K[:, None, None]
        w = np.maximum(Wr[None, :, None] * Eg, 1e-3)
        w = np.minimum(w, Eg)
        qm = np.sqrt(Eg) - np.sqrt(np.maximum(Eg - w, 0.0))
        qp = np.sqrt(Eg) + np.sqrt(np.maximum(Eg - w, 0.0))
        qm = np.maximum(qm, 1e-6)
        q = qm * (qp / qm) ** P[None, None, :]
        ion = _binding_icdf(spec, K[:, None] * Orat[None, :], P)
        ds.update({
            "/full_penn/imfp": RawDataset(imfp.astype(np.float32), "nm^-1", [kscale]),
            "/full_penn/omega_icdf": RawDataset(w_icdf.astype(np.float32), "eV", [kscale, pscale]),
            "/full_penn/q_icdf": RawDataset((q / HBAR_SQRT2M).astype(np.float32), "nm^-1",
                                            [kscale, (Wr, ""), pscale]),
            "/ionization/binding_icdf": RawDataset(ion.astype(np.float32), "eV",
                                                   [kscale, (Orat, ""), pscale]),
        })
        inel = 0
    elif model == "kieft":
        ion = _binding_icdf(spec, K, P)
        ds.update({
            "/inelastic_kieft/imfp": RawDataset(imfp.astype(np.float32), "nm^-1", [kscale]),
            "/inelastic_kieft/w0_icdf": RawDataset(w_icdf.astype(np.float32), "eV", [kscale, pscale]),
            "/ionization/binding_icdf": RawDataset(ion.astype(np.float32), "eV", [kscale, pscale]),
            "/ionization/outer_shells": RawDataset(np.array(spec.outer_shells, dtype=np.float32), "eV", [None]),
        })
        inel = 1
    else:
        raise ValueError(f"unknown inelastic model {model!r}")
    attrs = {
        "barrier": (spec.work_function + spec.fermi, "eV"),
        "fermi": (spec.fermi, "eV"),
        "band_gap": (spec.band_gap, "eV"),
        "phonon_loss": (spec.phonon_loss, "eV"),
        "effective_A": (spec.atomic_mass_u / AVOGADRO, "g"),
    }
    return RawMaterial(attrs, ds, inel)
