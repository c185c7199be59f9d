"""GDML text for tests: a GeometryDesc written as a GDML document (box, tube, cone, trap, trd, orb,
sphere, torus, paraboloid solids), and the description the reader must return for it (logical
volumes in document order, placements in the order of their <physvol> elements)."""
import math

import numpy as np

from vecgeom_amd import shapes as S

_NAMES = {S.BOX: "box", S.TUBE: "tube", S.CONE: "cone", S.TRAPEZOID: "trap", S.TRD: "trd", S.ORB: "orb",
          S.SPHERE: "sphere", S.TORUS2: "torus", S.PARABOLOID: "paraboloid"}


def _solid_xml(name, t, p):
    r = repr
    if t == S.BOX:
        a = 'x="%s" y="%s" z="%s"' % (r(2 * p[0]), r(2 * p[1]), r(2 * p[2]))
    elif t == S.TUBE:
        a = 'rmin="%s" rmax="%s" z="%s" startphi="%s" deltaphi="%s" aunit="rad"' % (
            r(p[0]), r(p[1]), r(2 * p[2]), r(p[3]), r(p[4]))
    elif t == S.CONE:
        a = 'rmin1="%s" rmax1="%s" rmin2="%s" rmax2="%s" z="%s" startphi="%s" deltaphi="%s" aunit="rad"' % (
            r(p[0]), r(p[1]), r(p[2]), r(p[3]), r(2 * p[4]), r(p[5]), r(p[6]))
    elif t == S.TRAPEZOID:
        dz, th, ph, dy1, dx1, dx2, a1, dy2, dx3, dx4, a2 = p
        a = ('z="%s" theta="%s" phi="%s" y1="%s" x1="%s" x2="%s" alpha1="%s" y2="%s" x3="%s" x4="%s" '
             'alpha2="%s" aunit="rad"' % (r(2 * dz), r(th), r(ph), r(2 * dy1), r(2 * dx1), r(2 * dx2), r(a1),
                                          r(2 * dy2), r(2 * dx3), r(2 * dx4), r(a2)))
    elif t == S.TRD:
        a = 'x1="%s" x2="%s" y1="%s" y2="%s" z="%s"' % tuple(r(2 * v) for v in p[:5])
    elif t == S.ORB:
        a = 'r="%s"' % r(p[0])
    elif t == S.SPHERE:
        a = ('rmin="%s" rmax="%s" startphi="%s" deltaphi="%s" starttheta="%s" deltatheta="%s" aunit="rad"'
             % tuple(r(v) for v in p))
    elif t == S.TORUS2:
        a = 'rmin="%s" rmax="%s" rtor="%s" startphi="%s" deltaphi="%s" aunit="rad"' % tuple(r(v) for v in p)
    elif t == S.PARABOLOID:
        a = 'rlo="%s" rhi="%s" dz="%s"' % tuple(r(v) for v in p)
    else:
        raise ValueError("no GDML writer for shape type %d" % t)
    return '    <%s name="%s" lunit="mm" %s/>' % (_NAMES[t], name, a)


def z_angle(t):
    """The GDML z rotation angle whose matrix is the description's rotation about z (or None)."""
    r = np.array(t[3:]).reshape(3, 3)
    if not (r[2, 2] == 1.0 and r[0, 2] == 0 and r[1, 2] == 0 and r[2, 0] == 0 and r[2, 1] == 0):
        return None
    return -math.atan2(r[1, 0], r[0, 0])


def write(g, path):
    """Writes g; every rotation must be about z and every logical volume reachable from the world.
    Returns the order the volumes were written in (a list of logical indices)."""
    daughters = {}
    for k, (m, d, t) in enumerate(g.placements):
        daughters.setdefault(m, []).append(k)
    order, seen = [], set()

    def visit(lv):  # daughters first: a volume is defined before it is referenced
        if lv in seen:
            return
        seen.add(lv)
        for k in daughters.get(lv, []):
            visit(g.placements[k][1])
        order.append(lv)
    visit(g.world)
    out = ['<?xml version="1.0"?>', "<!-- written by tests/gdml_util.py -->", "<gdml>", "  <define/>", "  <solids>"]
    used = sorted({g.logicals[lv] for lv in order})
    for s in used:
        t, p = g.shapes[s]
        out.append(_solid_xml("s%d" % s, t, p))
    out += ["  </solids>", "  <structure>"]
    for lv in order:
        out.append('    <volume name="v%d">' % lv)
        out.append('      <solidref ref="s%d"/>' % g.logicals[lv])
        for k in daughters.get(lv, []):
            _, d, t = g.placements[k]
            ang = z_angle(t)
            if ang is None:
                raise ValueError("only rotations about z are written")
            out.append('      <physvol name="p%d">' % k)
            out.append('        <volumeref ref="v%d"/>' % d)
            out.append('        <position name="pos%d" unit="mm" x="%r" y="%r" z="%r"/>' % (k, t[0], t[1], t[2]))
            out.append('        <rotation name="rot%d" unit="rad" x="0" y="0" z="%r"/>' % (k, ang))
            out.append("      </physvol>")
        out.append("    </volume>")
    out += ["  </structure>", '  <setup name="default" version="1.0">', '    <world ref="v%d"/>' % g.world,
            "  </setup>", "</gdml>"]
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")
    return order


def expected(g, order, scale=1.0):
    """The description the reader returns for write(g): shapes in written order, logical volumes in
    document order, placements in <physvol> order (volume by volume), lengths times `scale`."""
    used = sorted({g.logicals[lv] for lv in order})
    shape_index = {s: i for i, s in enumerate(used)}
    lv_index = {lv: i for i, lv in enumerate(order)}
    shapes = []
    for s in used:
        t, p = g.shapes[s]
        shapes.append((t, p))
    logicals = [shape_index[g.logicals[lv]] for lv in order]
    placements = []
    for lv in order:
        for k, (m, d, t) in enumerate(g.placements):
            if m == lv:
                # a GDML rotation by a about z is the passive one: rows (cos a, sin a, 0), (-sin a, cos a, 0)
                a = float(repr(z_angle(t)))
                c, s = math.cos(a), math.sin(a)
                tr = [t[0] * scale, t[1] * scale, t[2] * scale, c, s, 0.0, -s, c, 0.0, 0.0, 0.0, 1.0]
                placements.append((lv_index[m], lv_index[d], tr))
    return shapes, logicals, placements, lv_index[g.world]
