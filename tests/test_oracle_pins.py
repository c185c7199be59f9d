"""Pins of the oracle that do not come from the oracle itself.

Each solid below gets an independent description: a signed margin m(p), positive inside,
negative outside, written from the solid's geometry (planes, cylinders, cones, spheres,
tori, polygons) and never from the oracle's code. Every margin is 1-Lipschitz (a minimum /
maximum of distances to planes, cylinders, cones, spheres and circles), so |m(p)| never
exceeds the distance to the surface by construction of each term, and a ray can only cross
the surface where the sampled margin leaves room for it.

Checks, for the shapes of the batched workload and the reference's test shapes:

- Inside and Contains against the sign of the margin for points at least 1e-6 mm from the
  surface (Contains == Inside == kInside there);
- DistanceToIn / DistanceToOut against the first crossing found by marching the margin
  along the ray and bisecting (a bisection of the implicit equation, for the torus the
  quartic's roots); rays whose march cannot exclude a hidden crossing pair (two samples
  both closer to the surface than their spacing) or whose crossing grazes the surface
  (inside for less than 1e-6 mm) are left out, and at least 90 % of the rays must remain;
- the conventions of the reference's ShapeTester (test/shape_tester/ConventionChecker.cpp:
  129-399): inside points give DistanceToIn < 0, finite DistanceToOut, SafetyToIn < 0,
  SafetyToOut > 0; outside points give DistanceToIn > 0, DistanceToOut < 0, SafetyToIn > 0,
  SafetyToOut < 0; surface points (the bisected crossings) are kSurface with both safeties
  within the tolerance;
- safeties bound distances (ShapeTester.cpp TestSafety*: SafetyToIn <= DistanceToIn and
  SafetyToOut <= DistanceToOut along every sampled direction);
- the Orb's closed forms: DistanceToOut = -p.d + sqrt((p.d)^2 - |p|^2 + R^2), safeties |R - |p||.

The margins themselves are pinned by a plausible-mistake check at the end: a margin with a
flipped phi wedge or a dropped inner surface fails the Inside comparison.
"""
import math

import numpy as np
import pytest

from oracle.pyoracle import OracleShape

INF = 1e300
BAND = 1e-6


# ----------------------------------------------------------------------------------------- margins
def _rho(p):
    return np.hypot(p[0], p[1])


def phi_margin(p, sphi, dphi):
    """Signed distance-like margin of the wedge sphi <= phi <= sphi + dphi (planes through z)."""
    if dphi >= 2 * math.pi - 1e-12:
        return np.full(p.shape[1], np.inf)
    e = sphi + dphi
    d1 = -math.sin(sphi) * p[0] + math.cos(sphi) * p[1]   # > 0 on the wedge side of the start plane
    d2 = math.sin(e) * p[0] - math.cos(e) * p[1]          # > 0 on the wedge side of the end plane
    if dphi <= math.pi:  # the intersection of the two half-spaces (one half-space at exactly pi)
        return np.minimum(d1, d2)
    return np.maximum(d1, d2)


def cone_margin(p, r1, r2, dz, outer):
    """rho <= r(z) (outer) or rho >= r(z) (inner), r linear from r1 at -dz to r2 at +dz."""
    b = (r2 - r1) / (2 * dz)
    rz = r1 + b * (p[2] + dz)
    s = math.sqrt(1 + b * b)
    return (rz - _rho(p)) / s if outer else (_rho(p) - rz) / s


def theta_margin(p, t1, t2):
    r = np.linalg.norm(p, axis=0)
    th = np.arccos(np.clip(p[2] / np.maximum(r, 1e-300), -1, 1))
    m = np.full(p.shape[1], np.inf)
    if t1 > 0:
        m = np.minimum(m, r * np.sin(np.clip(th - t1, -math.pi / 2, math.pi / 2)))
    if t2 < math.pi:
        m = np.minimum(m, r * np.sin(np.clip(t2 - th, -math.pi / 2, math.pi / 2)))
    return m


def margin_box(params, p):
    h = np.asarray(params[:3])[:, None]
    return np.min(h - np.abs(p), axis=0)


def margin_orb(params, p):
    return params[0] - np.linalg.norm(p, axis=0)


def margin_tube(params, p):
    rmin, rmax, dz, sphi, dphi = params
    m = np.minimum(dz - np.abs(p[2]), rmax - _rho(p))
    if rmin > 0:
        m = np.minimum(m, _rho(p) - rmin)
    return np.minimum(m, phi_margin(p, sphi, dphi))


def margin_cone(params, p):
    rmin1, rmax1, rmin2, rmax2, dz, sphi, dphi = params
    m = np.minimum(dz - np.abs(p[2]), cone_margin(p, rmax1, rmax2, dz, True))
    if rmin1 > 0 or rmin2 > 0:
        m = np.minimum(m, cone_margin(p, rmin1, rmin2, dz, False))
    return np.minimum(m, phi_margin(p, sphi, dphi))


def margin_sphere(params, p):
    rmin, rmax, sphi, dphi, stheta, dtheta = params
    r = np.linalg.norm(p, axis=0)
    m = rmax - r
    if rmin > 0:
        m = np.minimum(m, r - rmin)
    m = np.minimum(m, phi_margin(p, sphi, dphi))
    return np.minimum(m, theta_margin(p, stheta, min(math.pi, stheta + dtheta)))


def margin_torus(params, p):
    rmin, rmax, rtor, sphi, dphi = params
    dc = np.hypot(_rho(p) - rtor, p[2])  # distance to the tube's centre circle
    m = rmax - dc
    if rmin > 0:
        m = np.minimum(m, dc - rmin)
    return np.minimum(m, phi_margin(p, sphi, dphi))


def profile_margin(u, z, zs, rin, rout):
    """Signed distance (positive inside) of (u, z) to the polygon of a z-plane profile: outer
    radii going up, inner radii coming back down. Edges lying on the axis u = 0 are not surfaces
    and are left out of the distance. 1-Lipschitz in (u, z)."""
    verts = [(rout[k], zs[k]) for k in range(len(zs))] + [(rin[k], zs[k]) for k in reversed(range(len(zs)))]
    V = np.array(verts)
    A, B = V, np.roll(V, -1, axis=0)
    dist = np.full(u.shape, np.inf)
    inside = np.zeros(u.shape, dtype=bool)
    for (au, az), (bu, bz) in zip(A, B):
        if (au, az) == (bu, bz):
            continue
        if (az > z).any() or (bz > z).any():  # even-odd crossing test along +u
            cross = (az > z) != (bz > z)
            with np.errstate(divide="ignore", invalid="ignore"):
                ux = au + (z - az) * (bu - au) / (bz - az)
            inside ^= cross & (u < ux)
        if au == 0.0 and bu == 0.0:
            continue
        eu, ez = bu - au, bz - az
        tt = np.clip(((u - au) * eu + (z - az) * ez) / (eu * eu + ez * ez), 0.0, 1.0)
        dist = np.minimum(dist, np.hypot(u - (au + tt * eu), z - (az + tt * ez)))
    return np.where(inside, dist, -dist)


def margin_polycone(params, p):
    sphi, dphi, nz = params[0], params[1], int(params[2])
    z = params[3:3 + nz]
    rmin = params[3 + nz:3 + 2 * nz]
    rmax = params[3 + 2 * nz:3 + 3 * nz]
    return np.minimum(profile_margin(_rho(p), p[2], z, rmin, rmax), phi_margin(p, sphi, dphi))


def margin_polyhedron(params, p):
    """Regular polygon cross-sections whose radii are the distances to the sides (apothems),
    side j spanning phi in [sphi + j*dside, sphi + (j+1)*dside]: the (r, z) profile applies to
    the projection on the normal of the point's own side, which inside the wedge is the largest
    projection over the sides (a maximum of unit-normal projections, so 1-Lipschitz)."""
    sphi, dphi, nsides, nz = params[0], params[1], int(params[2]), int(params[3])
    z = params[4:4 + nz]
    rmin = params[4 + nz:4 + 2 * nz]
    rmax = params[4 + 2 * nz:4 + 3 * nz]
    dside = dphi / nsides
    centres = sphi + (np.arange(nsides) + 0.5) * dside
    proj = np.max(np.cos(centres)[:, None] * p[0][None, :] + np.sin(centres)[:, None] * p[1][None, :], axis=0)
    return np.minimum(profile_margin(proj, p[2], z, rmin, rmax), phi_margin(p, sphi, dphi))


def margin_polyhedron_ref(params, p):
    """The solid the reference's Inside / Contains describe: the phi quadrilaterals are tested as two
    half-spaces (Planes::Inside / Contains on segment.phi, PolyhedronImplementation.h:728,826), an
    intersection also when the phi range exceeds pi (where the wedge is their union)."""
    sphi, dphi, nsides, nz = params[0], params[1], int(params[2]), int(params[3])
    z = params[4:4 + nz]
    dside = dphi / nsides
    centres = sphi + (np.arange(nsides) + 0.5) * dside
    proj = np.max(np.cos(centres)[:, None] * p[0][None, :] + np.sin(centres)[:, None] * p[1][None, :], axis=0)
    m = profile_margin(proj, p[2], z, params[4 + nz:4 + 2 * nz], params[4 + 2 * nz:4 + 3 * nz])
    e = sphi + dphi
    d1 = -math.sin(sphi) * p[0] + math.cos(sphi) * p[1]
    d2 = math.sin(e) * p[0] - math.cos(e) * p[1]
    return np.minimum(m, np.minimum(d1, d2))


def trapezoid_corners(params):
    """The eight corners of G4Trap parameters (dz theta phi dy1 dx1 dx2 alpha1 dy2 dx3 dx4 alpha2)."""
    dz, th, ph, dy1, dx1, dx2, a1, dy2, dx3, dx4, a2 = params
    tx, ty = math.tan(th) * math.cos(ph), math.tan(th) * math.sin(ph)
    out = []
    for zs, dy, dxl, dxh, ta in ((-1, dy1, dx1, dx2, math.tan(a1)), (1, dy2, dx3, dx4, math.tan(a2))):
        cz = zs * dz
        for ys, dx in ((-1, dxl), (1, dxh)):
            for xs in (-1, 1):
                out.append((tx * cz + ta * ys * dy + xs * dx, ty * cz + ys * dy, cz))
    return np.array(out)


def margin_convex(corners):
    from scipy.spatial import ConvexHull
    eq = ConvexHull(corners).equations  # unit outward normals, offsets: n.x + c <= 0 inside

    def f(params, p):
        return -np.max(eq[:, :3] @ p + eq[:, 3:4], axis=0)
    return f


def trd_corners(params):
    """The eight corners of a Trd (dx1 dx2 dy1 dy2 dz): half-widths dx1, dy1 at -dz and dx2, dy2 at +dz."""
    dx1, dx2, dy1, dy2, dz = params
    return np.array([(xs * dx, ys * dy, zs * dz) for zs, dx, dy in ((-1, dx1, dy1), (1, dx2, dy2))
                     for ys in (-1, 1) for xs in (-1, 1)])


def margin_paraboloid(params, p):
    """|z| <= dz and rho <= r(z) = sqrt(k1 z + k2), the circle radii rlo at -dz and rhi at +dz on a
    paraboloid of revolution. h = r(clamp(z)) - rho is L-Lipschitz with L = sqrt(1 + r'^2) and
    r' = k1 / (2 r) largest where r is smallest on [-dz, dz], so h / L is 1-Lipschitz; both terms
    vanish on the surface."""
    rlo, rhi, dz = params
    k1 = (rhi * rhi - rlo * rlo) / (2 * dz)
    k2 = (rhi * rhi + rlo * rlo) / 2
    lip = math.sqrt(1 + (k1 / (2 * min(rlo, rhi))) ** 2)
    zc = np.clip(p[2], -dz, dz)
    h = np.sqrt(k1 * zc + k2) - _rho(p)
    return np.minimum(dz - np.abs(p[2]), h / lip)


# ------------------------------------------------------------------------------- shapes under test
TWO_PI = 2 * math.pi
CASES = [
    ("box", [10.0, 20.0, 30.0], margin_box),
    ("orb", [9.0], margin_orb),
    ("tube", [5.0, 10.0, 15.0, 0.0, TWO_PI], margin_tube),
    ("tube", [2.0, 8.0, 10.0, math.radians(30), math.radians(240)], margin_tube),
    ("tube", [0.0, 8.0, 10.0, math.radians(10), math.radians(100)], margin_tube),
    ("cone", [2.0, 6.0, 4.0, 10.0, 12.0, 0.0, TWO_PI], margin_cone),
    ("cone", [0.0, 5.0, 3.0, 8.0, 10.0, math.radians(20), math.radians(270)], margin_cone),
    ("sphere", [5.0, 12.0, 0.0, TWO_PI, 0.0, math.pi], margin_sphere),
    ("sphere", [4.0, 12.0, math.radians(20), math.radians(300), math.radians(30), math.radians(100)],
     margin_sphere),
    ("sphere", [0.0, 10.0, 0.0, math.radians(120), math.radians(100), math.radians(80)], margin_sphere),
    ("torus2", [0.0, 10.0, 50.0, 0.0, TWO_PI], margin_torus),
    ("torus2", [3.0, 8.0, 30.0, math.radians(40), math.radians(250)], margin_torus),
    ("polycone", [0.0, TWO_PI, 4, -10.0, -2.0, 3.0, 10.0, 1.0, 2.0, 0.0, 1.5, 5.0, 8.0, 6.0, 9.0],
     margin_polycone),
    ("polycone", [math.radians(30), math.radians(200), 3, -8.0, 0.0, 8.0, 2.0, 1.0, 3.0, 9.0, 6.0, 10.0],
     margin_polycone),
    ("polyhedron", [0.0, TWO_PI, 8, 4, -12.0, -4.0, 4.0, 12.0, 0.0, 2.0, 3.0, 0.0, 8.0, 10.0, 11.0, 7.0],
     margin_polyhedron),
    ("polyhedron", [math.radians(15), math.radians(120), 3, 2, -10.0, 10.0, 3.0, 4.0, 9.0, 12.0],
     margin_polyhedron),
]
# Inside / Contains only: a polyhedron with a phi range above pi, whose reference answer is the
# solid of margin_polyhedron_ref (its distances follow the faces themselves)
INSIDE_ONLY = [
    ("polyhedron", [0.5, 4.0, 8, 4, -20.0, -10.0, 10.0, 20.0, 0.0, 5.0, 3.0, 0.0, 10.0, 15.0, 12.0, 8.0],
     margin_polyhedron_ref),
]
# planar side faces need (dx2 - dx1) / dy1 + 2 tan(alpha1) == (dx4 - dx3) / dy2 + 2 tan(alpha2)
_TRAPS = [
    [15.0, 0.0, 0.0, 10.0, 6.0, 8.0, 0.0, 12.0, 9.0, 11.4, 0.0],
    [20.0, math.radians(15), math.radians(30), 8.0, 5.0, 7.0, math.radians(10), 10.0, 6.0, 8.5, math.radians(10)],
]
for _t in _TRAPS:
    CASES.append(("trapezoid", _t, margin_convex(trapezoid_corners(_t))))

def _hull_case():
    from vecgeom_amd.workloads import convex_hull_sphere
    params = convex_hull_sphere()
    verts = np.array(params[1:]).reshape(-1, 3)
    return ("tessellated", params, margin_convex(verts))


CASES.append(_hull_case())
for _t in ([10.0, 30.0, 20.0, 40.0, 40.0], [5.0, 12.0, 9.0, 9.0, 15.0], [20.0, 20.0, 30.0, 30.0, 40.0],
           [5.0, 12.0, 9.0, 9.0, 15.0, 1.0]):  # the last: the Trd1 kind
    CASES.append(("trd", _t, margin_convex(trd_corners(_t[:5]))))
CASES.append(("paraboloid", [6.0, 9.0, 10.0], margin_paraboloid))
# the reference's root selection takes the quadratic's leading coefficient as positive, i.e. rhi > rlo
# (the G4Paraboloid convention R2 > R1); rlo > rhi is not pinned
CASES.append(("paraboloid", [1.0, 12.0, 8.0], margin_paraboloid))


def parallelepiped_corners(params):
    """The eight corners of a parallelepiped (dx dy dz alpha theta phi): the oblique-frame corners
    (+-dx, +-dy, +-dz) with y = y' + tan(theta) sin(phi) z and x = x' + tan(theta) cos(phi) z + tan(alpha) y'."""
    dx, dy, dz, al, th, ph = params
    tc, ts, ta = math.tan(th) * math.cos(ph), math.tan(th) * math.sin(ph), math.tan(al)
    return np.array([(xs * dx + tc * zs * dz + ta * ys * dy, ys * dy + ts * zs * dz, zs * dz)
                     for zs in (-1, 1) for ys in (-1, 1) for xs in (-1, 1)])


for _t in ([20.0, 30.0, 40.0, math.radians(30), math.radians(30), math.radians(45)],
           [8.0, 5.0, 12.0, math.radians(-20), math.radians(15), math.radians(200)]):
    CASES.append(("parallelepiped", _t, margin_convex(parallelepiped_corners(_t))))
for _t in ([0.0, 0.0, 2.0, 0.0, 0.0, 0.0, 2.0, 0.0, 0.0, 0.0, 2.0, 0.0],
           [6.0, 4.0, 5.0, -5.0, -6.0, 4.0, -4.0, 5.0, -6.0, 5.0, -5.0, -4.0]):
    CASES.append(("tet", _t, margin_convex(np.array(_t).reshape(4, 3))))


def margin_ellipsoid(params, p):
    """x^2/a^2 + y^2/b^2 + z^2/c^2 <= 1 between the z cuts: min(a, b, c) (1 - |(x/a, y/b, z/c)|) is
    1-Lipschitz (the scaled norm's gradient is at most 1 / min(a, b, c)) and vanishes on the surface;
    the cuts clamped to [-c, c] as the solid's constructor does (both 0: no cut)."""
    a, b, c = params[:3]
    zb, zt = (params[3], params[4]) if len(params) == 5 else (0.0, 0.0)
    if (zb == 0 and zt == 0) or zb >= c or zt <= -c or zb >= zt:
        zb, zt = -c, c
    zb, zt = max(zb, -c), min(zt, c)
    q = np.sqrt((p[0] / a) ** 2 + (p[1] / b) ** 2 + (p[2] / c) ** 2)
    return np.minimum(min(a, b, c) * (1 - q), np.minimum(zt - p[2], p[2] - zb))


def margin_eltube(params, p):
    """An elliptical cylinder |z| <= dz: min(dx, dy) (1 - |(x/dx, y/dy)|), 1-Lipschitz as above."""
    a, b, dz = params
    q = np.sqrt((p[0] / a) ** 2 + (p[1] / b) ** 2)
    return np.minimum(min(a, b) * (1 - q), dz - np.abs(p[2]))


for _t in ([5.0, 10.0, 20.0], [30.0, 12.0, 8.0]):
    CASES.append(("ellipticaltube", _t, margin_eltube))


def margin_cuttube(params, p):
    """A tube of unbounded length cut by the planes through (0, 0, -dz) and (0, 0, dz) with the given
    outward normals: the tube margin without its z term, and each plane's signed distance."""
    rmin, rmax, dz, sphi, dphi = params[:5]
    m = margin_tube([rmin, rmax, np.inf, sphi, dphi], p)
    for nrm, z0 in ((params[5:8], -dz), (params[8:11], dz)):
        u = np.asarray(nrm, float) / np.linalg.norm(nrm)
        m = np.minimum(m, -(u[0] * p[0] + u[1] * p[1] + u[2] * (p[2] - z0)))
    return m


for _t in ([0.0, 10.0, 20.0, 0.0, TWO_PI, 0.0, -0.7, -0.71, 0.7, 0.0, 0.71],
           [3.0, 8.0, 15.0, math.radians(30), math.radians(240), 0.3, 0.2, -1.0, -0.4, 0.1, 1.0],
           [1.0, 5.0, 10.0, 1.0, 4.0, -1.0, -1.0, -1.0, 1.0, 1.0, 1.0]):
    CASES.append(("cuttube", _t, margin_cuttube))


def margin_elcone(params, p):
    """An elliptical cone sqrt((x/a)^2 + (y/b)^2) + z <= h, |z| <= zcut: the cone term scaled by
    min(a, b) / sqrt(1 + min(a, b)^2), which makes it 1-Lipschitz (its gradient is at most
    sqrt(1 + 1 / min(a, b)^2))."""
    a, b, h, zc = params
    m = min(a, b)
    cone = (h - np.sqrt((p[0] / a) ** 2 + (p[1] / b) ** 2) - p[2]) * m / math.sqrt(1 + m * m)
    return np.minimum(cone, zc - np.abs(p[2]))


for _t in ([0.5, 0.4, 10.0, 5.0], [1.0, 2.0, 3.0, 4.0 - 1.0], [0.3, 0.9, 20.0, 8.0]):
    CASES.append(("ellipticalcone", _t, margin_elcone))


def gentrap_corners(params):
    """The eight corners of a GenTrap: (x[i], y[i], -dz) for i < 4, (x[i], y[i], +dz) above."""
    x, y, dz = params[:8], params[8:16], params[16]
    return np.unique(np.array([(x[i], y[i], -dz if i < 4 else dz) for i in range(8)]), axis=0)


# planar traps (twisted ones are not built): the reference test's frustum (clockwise), an anticlockwise
# skewed one (the constructor reverts it), one whose top is a segment (two collapsed edges)
for _t in ([-3.0, -3.0, 3.0, 3.0, -2.0, -2.0, 2.0, 2.0, -3.0, 3.0, 3.0, -3.0, -2.0, 2.0, 2.0, -2.0, 5.0],
           [-4.0, 5.0, 6.0, -3.0, -1.4, 4.0, 4.6, -0.8, -5.0, -4.0, 4.0, 3.0, -2.5, -1.9, 2.9, 2.3, 8.0],
           [-3.0, -3.0, 3.0, 3.0, -2.0, -2.0, 2.0, 2.0, -3.0, 3.0, 3.0, -3.0, 0.0, 0.0, 0.0, 0.0, 6.0]):
    CASES.append(("gentrap", _t, margin_convex(gentrap_corners(_t))))
for _t in ([3.0, 4.0, 5.0, -4.5, 3.5], [10.0, 7.0, 12.0], [6.0, 9.0, 4.0, 1.0, 4.0], [6.0, 6.0, 6.0, -2.0, 5.0]):
    CASES.append(("ellipsoid", _t, margin_ellipsoid))
IDS = ["%s-%d" % (c[0], i) for i, c in enumerate(CASES)]


def _bbox(s):
    lo, hi = s.extent()
    return lo, hi


def _sample(s, n, seed):
    lo, hi = _bbox(s)
    c, h = 0.5 * (lo + hi), 0.5 * (hi - lo)
    g = np.random.default_rng(seed)
    p = c[:, None] + g.uniform(-1.3, 1.3, size=(3, n)) * h[:, None]
    d = g.normal(size=(3, n))
    d /= np.linalg.norm(d, axis=0)
    return np.ascontiguousarray(p), np.ascontiguousarray(d)


def first_crossing(margin, params, p, d, tmax, entering, max_iter=4000):
    """First t in (0, tmax] where the margin's sign turns positive (entering) or negative
    (exiting); inf when there is none. Sphere tracing: the margin is 1-Lipschitz, so a step of
    |m| cannot pass a crossing; once |m| < 1e-9 the contact is bracketed with a probe 1e-6
    further on and bisected. Returns (t, certain): a ray is certain unless it does not converge
    within max_iter steps or its contact is not a crossing that lasts 1e-6 (a grazing touch)."""
    n = p.shape[1]
    t = np.zeros(n)
    out = np.full(n, np.inf)
    certain = np.ones(n, dtype=bool)
    act = np.ones(n, dtype=bool)
    lo = np.zeros(n)
    hi = np.zeros(n)
    bracketed = np.zeros(n, dtype=bool)
    for _ in range(max_iter):
        idx = np.nonzero(act)[0]
        if idx.size == 0:
            break
        m = margin(params, p[:, idx] + d[:, idx] * t[idx][None, :])
        m = -m if entering else m  # positive while the ray has not crossed
        crossed = m <= 0
        near = ~crossed & (m < 1e-9)
        far = ~crossed & ~near
        t[idx[far]] += m[far]
        beyond = idx[far][t[idx[far]] > tmax[idx[far]]]
        act[beyond] = False  # no crossing up to tmax
        # a contact: probe 1e-6 further, then bisect [t, probe]
        for group, start_inside in ((idx[near], False), (idx[crossed], True)):
            if group.size == 0:
                continue
            if start_inside:  # landed exactly on or across the surface (m <= 0 at t)
                lo[group], hi[group] = np.maximum(t[group] - 1e-9, 0.0), t[group]
            else:
                lo[group], hi[group] = t[group], t[group] + 1e-6
            bracketed[group] = True
            act[group] = False
    certain &= ~act  # did not converge
    rows = np.nonzero(bracketed)[0]
    if rows.size:
        def state(tt):
            mm = margin(params, p[:, rows] + d[:, rows] * tt[None, :])
            return (mm > 0) if entering else (mm < 0)
        a, b = lo[rows], hi[rows]
        ok = state(b)
        ok &= ~state(a) | (a == 0)
        for _ in range(64):
            mid = 0.5 * (a + b)
            w = state(mid)
            b = np.where(w, mid, b)
            a = np.where(w, a, mid)
        grazing = ~ok | ~state(b + 1e-6)
        out[rows] = np.where(ok, b, np.inf)
        certain[rows] &= ~grazing
    return out, certain


@pytest.mark.parametrize("case", CASES + INSIDE_ONLY, ids=IDS + ["%s-ref-%d" % (c[0], i) for i, c in enumerate(INSIDE_ONLY)])
def test_inside_and_contains_match_margin(case):
    kind, params, margin = case
    s = OracleShape(kind, params)
    p, _ = _sample(s, 20000, 11)
    m = margin(params, p)
    keep = np.abs(m) > BAND
    p, m = np.ascontiguousarray(p[:, keep]), m[keep]
    assert (m > 0).mean() > 0.05 and (m < 0).mean() > 0.05
    ins = s.query("Inside", p)
    con = s.query("Contains", p).astype(bool)
    assert np.array_equal(ins, np.where(m > 0, 1, 3)), int((ins != np.where(m > 0, 1, 3)).sum())
    if kind == "torus2" and params[4] < TWO_PI:
        # the reference's own exception: TorusImplementation2::Contains (TorusImplementation2.h:520-527)
        # runs GenericKernelForContainsAndInside<Real_v, true, false>, whose notForDisk = false skips
        # the phi wedge, so Contains of a phi-cut torus is the Contains of the full torus
        full = margin(params[:3] + [0.0, TWO_PI], p)
        assert np.array_equal(con[np.abs(full) > BAND], full[np.abs(full) > BAND] > 0)
        assert not np.array_equal(con, m > 0)
        return
    assert np.array_equal(con, m > 0)


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_distances_match_bisection(case):
    kind, params, margin = case
    s = OracleShape(kind, params)
    p, d = _sample(s, 3000, 12)
    m = margin(params, p)
    keep = np.abs(m) > BAND
    p, d, m = np.ascontiguousarray(p[:, keep]), np.ascontiguousarray(d[:, keep]), m[keep]
    lo, hi = _bbox(s)
    diag = float(np.linalg.norm(hi - lo))
    c = 0.5 * (lo + hi)
    tmax = np.linalg.norm(p - c[:, None], axis=0) + diag
    for entering in (True, False):
        sel = (m < 0) if entering else (m > 0)
        ps, ds = np.ascontiguousarray(p[:, sel]), np.ascontiguousarray(d[:, sel])
        exp, certain = first_crossing(margin, params, ps, ds, tmax[sel], entering)
        got = s.query("DistanceToIn" if entering else "DistanceToOut", ps, ds)
        both_inf = (got >= INF) & ~np.isfinite(exp)
        ok = both_inf | (np.abs(got - exp) <= 1e-9 * np.maximum(1.0, np.abs(exp)))
        if kind == "torus2" and entering and params[4] < TWO_PI:
            # the reference's own exception: TorusImplementation2::DistanceToIn keeps a hit on a phi
            # face only when it lies on the bounding tube's entry (|distance| < kTolerance) and
            # otherwise resets it to kInfLength (TorusImplementation2.h:610-640), so a ray entering
            # through a phi face gets the next toroidal entry or a miss: a later (or no) hit
            q = ps + ds * np.where(np.isfinite(exp), exp, 0.0)[None, :]
            on_phi = np.isfinite(exp) & (np.abs(phi_margin(q, params[3], params[4])) < 1e-7)
            assert on_phi.any()
            assert np.all(got[on_phi & certain] >= exp[on_phi & certain] - 1e-9)
            ok |= on_phi
        bad = ~ok & certain
        assert not bad.any(), (entering, int(bad.sum()), ps[:, bad][:, :3].T.tolist(), ds[:, bad][:, :3].T.tolist(),
                               got[bad][:3].tolist(), exp[bad][:3].tolist())
        assert certain.mean() > 0.9, certain.mean()
        if entering:
            assert np.isfinite(exp[certain]).any()


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_conventions(case):
    """ConventionChecker.cpp:129-399 (ShapeConventionSurfacePoint / InsidePoint / OutsidePoint)."""
    kind, params, margin = case
    s = OracleShape(kind, params)
    p, d = _sample(s, 6000, 13)
    m = margin(params, p)
    keep = np.abs(m) > BAND
    p, d, m = np.ascontiguousarray(p[:, keep]), np.ascontiguousarray(d[:, keep]), m[keep]
    ins = m > 0
    pi, di = np.ascontiguousarray(p[:, ins]), np.ascontiguousarray(d[:, ins])
    po, do = np.ascontiguousarray(p[:, ~ins]), np.ascontiguousarray(d[:, ~ins])
    assert np.all(s.query("DistanceToIn", pi, di) < 0)        # 6
    assert np.all(s.query("DistanceToOut", pi, di) < INF)     # 7
    assert np.all(s.query("SafetyToIn", pi) < 0)              # 8
    assert np.all(s.query("SafetyToOut", pi) > 0)             # 9
    assert np.all(s.query("DistanceToIn", po, do) > 0)        # 10
    dout = s.query("DistanceToOut", po, do)
    if kind == "cone" and params[6] < TWO_PI:
        # the reference's own exception: ConeImplementation::DistanceToOut runs the sector test only
        # when the point is already outside (`if (checkPhiTreatment && !MaskEmpty(outside))`,
        # ConeImplementation.h:293-297), so a point outside the phi wedge alone is not given -1
        only_phi = phi_margin(po, params[5], params[6]) < -BAND
        only_phi &= margin_cone(params[:5] + [0.0, TWO_PI], po) > BAND
        assert only_phi.any() and (dout[only_phi] >= 0).any()
        dout = dout[~only_phi]
    assert np.all(dout < 0)                                    # 11
    assert np.all(s.query("SafetyToIn", po) > 0)              # 12
    assert np.all(s.query("SafetyToOut", po) < 0)             # 13
    # surface points: crossings of inside points' rays, bisected to 1e-13 (0-5)
    lo, hi = _bbox(s)
    tmax = np.full(pi.shape[1], float(np.linalg.norm(hi - lo)) * 2)
    t, certain = first_crossing(margin, params, pi, di, tmax, entering=False)
    ok = certain & np.isfinite(t)
    surf = np.ascontiguousarray(pi[:, ok] + di[:, ok] * t[ok][None, :])
    assert surf.shape[1] > 100
    assert np.all(s.query("Inside", surf) == 2)
    assert np.all(s.query("SafetyToIn", surf) <= 1e-9)
    assert np.all(s.query("SafetyToOut", surf) <= 1e-9)


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_safeties_bound_distances(case):
    kind, params, margin = case
    s = OracleShape(kind, params)
    p, _ = _sample(s, 1500, 14)
    m = margin(params, p)
    keep = np.abs(m) > BAND
    p, m = np.ascontiguousarray(p[:, keep]), m[keep]
    ins = m > 0
    sin_, sout = s.query("SafetyToIn", p), s.query("SafetyToOut", p)
    g = np.random.default_rng(15)
    for _ in range(24):
        d = g.normal(size=p.shape)
        d = np.ascontiguousarray(d / np.linalg.norm(d, axis=0))
        di = s.query("DistanceToIn", p, d)
        do = s.query("DistanceToOut", p, d)
        assert np.all(sin_[~ins] <= di[~ins] + 1e-9)
        assert np.all(sout[ins] <= do[ins] + 1e-9)
    # the margin never exceeds the distance to the surface; an exact safety is at least as large
    # only where the reference's safety is exact (orb: |R - |p||, box: inside)
    if kind == "orb":
        assert np.allclose(np.where(ins, sout, sin_), np.abs(m), rtol=0, atol=1e-12)


def test_orb_closed_forms():
    """OrbImplementation.h:95-190 as closed forms of the sphere |x| = R."""
    R = 9.0
    s = OracleShape("orb", [R])
    g = np.random.default_rng(16)
    p = g.uniform(-8.9, 8.9, size=(3, 4000))
    p = np.ascontiguousarray(p[:, np.linalg.norm(p, axis=0) < R - 1e-6])
    d = g.normal(size=p.shape)
    d = np.ascontiguousarray(d / np.linalg.norm(d, axis=0))
    b = np.sum(p * d, axis=0)
    c = np.sum(p * p, axis=0) - R * R
    exp = -b + np.sqrt(b * b - c)
    got = s.query("DistanceToOut", p, d)
    assert np.allclose(got, exp, rtol=0, atol=1e-12)
    assert np.allclose(s.query("SafetyToOut", p), R - np.linalg.norm(p, axis=0), rtol=0, atol=1e-12)
    q = np.ascontiguousarray(p * 3.0)
    q = np.ascontiguousarray(q[:, np.linalg.norm(q, axis=0) > R + 1e-6])
    assert np.allclose(s.query("SafetyToIn", q), np.linalg.norm(q, axis=0) - R, rtol=0, atol=1e-12)
    # DistanceToIn from outside: the nearer root when the ray meets the sphere ahead
    e = g.normal(size=q.shape)
    e = np.ascontiguousarray(e / np.linalg.norm(e, axis=0))
    b = np.sum(q * e, axis=0)
    c = np.sum(q * q, axis=0) - R * R
    disc = b * b - c
    hit = (disc > 1e-6) & (b < 0)
    got = s.query("DistanceToIn", q, e)
    assert np.allclose(got[hit], (-b - np.sqrt(np.maximum(disc, 0)))[hit], rtol=0, atol=1e-10)
    assert np.all(got[(disc < -1e-6) | (b > 0)] >= INF)


def test_margins_detect_planted_mistakes():
    """The margin check has teeth: an oracle stand-in with a plausible mistake (the phi wedge
    measured from the wrong plane, the inner radius dropped) disagrees with Inside."""
    params = [2.0, 8.0, 10.0, math.radians(30), math.radians(240)]
    s = OracleShape("tube", params)
    p, _ = _sample(s, 20000, 17)
    ins = s.query("Inside", p)
    m_ok = margin_tube(params, p)
    wrong_phi = margin_tube([2.0, 8.0, 10.0, -math.radians(30), math.radians(240)], p)
    no_inner = margin_tube([0.0, 8.0, 10.0, math.radians(30), math.radians(240)], p)
    keep = np.abs(m_ok) > BAND
    assert np.array_equal(ins[keep], np.where(m_ok[keep] > 0, 1, 3))
    assert not np.array_equal(ins[keep], np.where(wrong_phi[keep] > 0, 1, 3))
    assert not np.array_equal(ins[keep], np.where(no_inner[keep] > 0, 1, 3))
