"""Geometry descriptions (shapes, logical volumes, placements) and the synthetic
calorimeter used for the navigation workloads.

A description is plain data. Placement k becomes placed volume id k; the world
placed volume gets id len(placements). Transformations are 12 doubles
(tx, ty, tz, r0..r8) with the row-major rotation of
VecGeom/base/Transformation3D.h:59-63 (master = R * local + t).
"""
import math

import numpy as np

from .shapes import BOX, TRAPEZOID, TUBE

IDENTITY = (0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0)


class GeometryDesc:
    def __init__(self):
        self.shapes = []       # (type, [params])
        self.logicals = []     # shape index
        self.placements = []   # (mother_lv, daughter_lv, 12 doubles)
        self.world = None
        self.names = []

    def add_shape(self, type_id, params):
        self.shapes.append((int(type_id), [float(x) for x in params]))
        return len(self.shapes) - 1

    def add_logical(self, shape, name=""):
        self.logicals.append(int(shape))
        self.names.append(name)
        return len(self.logicals) - 1

    def place(self, mother, daughter, transform=IDENTITY):
        t = tuple(float(x) for x in transform)
        assert len(t) == 12
        self.placements.append((int(mother), int(daughter), t))
        return len(self.placements) - 1

    def set_world(self, lv):
        self.world = int(lv)

    # flattened arrays for the C ABIs
    def arrays(self):
        stype = np.array([s[0] for s in self.shapes], dtype=np.int32)
        soff = np.zeros(len(self.shapes) + 1, dtype=np.int32)
        for i, s in enumerate(self.shapes):
            soff[i + 1] = soff[i] + len(s[1])
        sparams = np.array([x for s in self.shapes for x in s[1]] or [0.0], dtype=np.float64)
        lvshape = np.array(self.logicals, dtype=np.int32)
        plm = np.array([p[0] for p in self.placements], dtype=np.int32)
        pll = np.array([p[1] for p in self.placements], dtype=np.int32)
        plt = np.array([x for p in self.placements for x in p[2]] or [0.0], dtype=np.float64)
        return stype, soff, sparams, lvshape, plm, pll, plt

    def touchable_count(self):
        daughters = {}
        for m, d, _ in self.placements:
            daughters.setdefault(m, []).append(d)

        def count(lv):
            return 1 + sum(count(d) for d in daughters.get(lv, []))

        return count(self.world)


def rot_z(phi):
    """Rotation by phi about z: local x axis at angle phi in the mother frame."""
    c, s = math.cos(phi), math.sin(phi)
    return (c, -s, 0.0, s, c, 0.0, 0.0, 0.0, 1.0)


def calorimeter(n_layers=40, n_cells=64, n_plates=8, r_min=1000.0, r_max=2000.0, half_z=2500.0,
                world=(2000.0, 2000.0, 3000.0)):
    """World box 4x4x6 m holding concentric tube layers (25 mm thick for the defaults);
    each layer holds n_cells trapezoid cells in phi whose side faces lie on the sector
    planes (neighbouring cells touch), and each cell holds n_plates box absorber plates
    stacked radially. Depth 4: world / layer / cell / plate."""
    g = GeometryDesc()
    wlv = g.add_logical(g.add_shape(BOX, world), "world")
    g.set_world(wlv)
    thick = (r_max - r_min) / n_layers
    alpha = math.pi / n_cells  # half opening angle of a cell
    margin = 0.5
    for L in range(n_layers):
        rin = r_min + L * thick
        rout = rin + thick
        layer = g.add_logical(g.add_shape(TUBE, [rin, rout, half_z, 0.0, 2.0 * math.pi]), "layer%d" % L)
        h_in = rin + margin
        h_out = (rout - margin) * math.cos(alpha)
        dy = 0.5 * (h_out - h_in)
        rc = 0.5 * (h_out + h_in)
        dx1 = h_in * math.tan(alpha)
        dx2 = (rout - margin) * math.sin(alpha)
        cz = half_z - margin
        cell = g.add_logical(g.add_shape(TRAPEZOID, [cz, 0.0, 0.0, dy, dx1, dx2, 0.0, dy, dx1, dx2, 0.0]),
                             "cell%d" % L)
        pitch = 2.0 * dy / n_plates
        half_t = 0.35 * pitch
        for k in range(n_plates):
            yc = -dy + pitch * (k + 0.5)
            y_lo = yc - half_t
            w = dx1 + (dx2 - dx1) * (y_lo + dy) / (2.0 * dy) - margin
            plate = g.add_logical(g.add_shape(BOX, [w, half_t, cz - margin]), "plate%d_%d" % (L, k))
            g.place(cell, plate, (0.0, yc, 0.0) + IDENTITY[3:])
        for c in range(n_cells):
            phi = (c + 0.5) * 2.0 * math.pi / n_cells
            # local y (the radial axis of the trapezoid) points along phi: rotate by phi - 90 deg
            t = (rc * math.cos(phi), rc * math.sin(phi), 0.0) + rot_z(phi - 0.5 * math.pi)
            g.place(layer, cell, t)
        g.place(wlv, layer, IDENTITY)
    return g


def small_calorimeter():
    """A reduced calorimeter with the same structure for fast parity tests."""
    return calorimeter(n_layers=4, n_cells=16, n_plates=4, r_min=100.0, r_max=200.0, half_z=250.0,
                       world=(200.0, 200.0, 300.0))


def world_points(g, n, seed):
    """Uniform points inside the world box, isotropic directions (Philox streams)."""
    from . import workloads
    hx, hy, hz = g.shapes[g.logicals[g.world]][1]
    rng = workloads.rng(seed)
    p = workloads.uniform_in_box(rng, n, [-hx, -hy, -hz], [hx, hy, hz], scale=1.0)
    d = workloads.isotropic_dirs(rng, n)
    return p, d
