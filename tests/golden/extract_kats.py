#!/usr/bin/env python3
"""Extract known-answer vectors from VecGeom's unit tests into JSON fixtures.

Reads the reference's test/unit_tests/Test<Shape>.cpp files as text (nothing
is compiled, imported or executed: numeric expressions are evaluated by a
whitelisted AST walker) and writes tests/golden/kat_<shape>.json with one
record per recognised assertion:

    {"line": N, "shape": [params...], "method": "DistanceToIn", "p": [..],
     "d": [..] | null, "expect": {"kind": "approx", "value": v}}

Expectation kinds: "approx" (ApproxEqual, test/unit_tests/ApproxEqual.h),
"anyof" (a disjunction of ApproxEqual), "inside" (EInside enum),
"negative" (assert(Dist < 0)), "nonpositive" (assert(Dist <= 0),
TestWrongSidePoint), "absdiff" (assert(fabs(s->M(..) - v) < tol)), "bool"
(assert(s->Contains(p)) / assert(!s->Contains(p))), "extent" (s.Extent(lo, hi) then
assert(ApproxEqual(lo, Vec_t(..))), method "Extent").

Run in a container that holds the reference tree:
    python tests/golden/extract_kats.py /root/reference
"""
import ast
import json
import math
import os
import re
import sys

DBL_MAX = sys.float_info.max
DBL_MIN = sys.float_info.min

CONSTS = {
    "kPi": math.pi, "kTwoPi": 2 * math.pi, "kHalfPi": 0.5 * math.pi,
    "kInfLength": DBL_MAX, "kInfinity": DBL_MAX, "kTolerance": 1e-9,
    "kHalfTolerance": 0.5e-9, "kCarTolerance": 1e-9, "kDegToRad": math.pi / 180.0,
    "kRadToDeg": 180.0 / math.pi, "M_PI": math.pi, "deg": math.pi / 180.0,
    "kRadTolerance": 1e-9, "kAngTolerance": 1e-9, "PI": math.pi,
}
FUNCS = {"sqrt": math.sqrt, "cos": math.cos, "sin": math.sin, "tan": math.tan,
         "atan": math.atan, "atan2": math.atan2, "fabs": math.fabs, "abs": abs,
         "Sqrt": math.sqrt, "Cos": math.cos, "Sin": math.sin}


class Unsupported(Exception):
    pass


def _norm_expr(s):
    s = s.strip()
    s = re.sub(r"\bstd::|\bvecgeom::|\bvecCore::math::|\bvecgeom::cxx::|\bcxx::|\bVECGEOM_NAMESPACE::", "", s)
    s = re.sub(r"InfinityLength<\w+>\(\)", "kInfLength", s)
    s = re.sub(r"(\d)[fF]\b", r"\1", s)
    s = re.sub(r"\(\s*(?:double|Precision|float)\s*\)", "", s)
    return s


CURRENT_VECS = {}


def eval_scalar(expr, scalars):
    expr = _norm_expr(expr)
    # vector components of known vectors: name.x() / name.y() / name.z()
    def comp(m):
        v = CURRENT_VECS.get(m.group(1))
        if v is None:
            raise Unsupported(m.group(0))
        return repr(v["xyz".index(m.group(2))])
    expr = re.sub(r"\b(\w+)\.([xyz])\(\)", comp, expr)
    try:
        tree = ast.parse(expr, mode="eval")
    except SyntaxError as e:
        raise Unsupported(expr) from e

    def ev(n):
        if isinstance(n, ast.Expression):
            return ev(n.body)
        if isinstance(n, ast.Constant) and isinstance(n.value, (int, float)):
            return float(n.value)
        if isinstance(n, ast.UnaryOp) and isinstance(n.op, (ast.USub, ast.UAdd)):
            v = ev(n.operand)
            return -v if isinstance(n.op, ast.USub) else v
        if isinstance(n, ast.BinOp):
            a, b = ev(n.left), ev(n.right)
            if isinstance(n.op, ast.Add):
                return a + b
            if isinstance(n.op, ast.Sub):
                return a - b
            if isinstance(n.op, ast.Mult):
                return a * b
            if isinstance(n.op, ast.Div):
                return a / b
            raise Unsupported(expr)
        if isinstance(n, ast.Name):
            if n.id in CONSTS:
                return CONSTS[n.id]
            if n.id in scalars:
                return scalars[n.id]
            raise Unsupported(n.id)
        if isinstance(n, ast.Call) and isinstance(n.func, ast.Name) and n.func.id in FUNCS:
            return float(FUNCS[n.func.id](*[ev(a) for a in n.args]))
        raise Unsupported(expr)

    return ev(tree)


def split_args(s):
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch in "([{":
            depth += 1
        elif ch in ")]}":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur)
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur)
    return [a.strip() for a in out]


def unit(v):
    m2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2]
    s = math.sqrt(m2 + DBL_MIN)  # Vector3D::Unit, base/Vector3D.h:284-290
    return [v[0] / s, v[1] / s, v[2] / s]


def eval_vec(expr, vecs, scalars, vec_type):
    e = expr.strip()
    m = re.fullmatch(r"\((.*)\)", e)
    if m and e.count("(") == 1:
        e = m.group(1).strip()
    if e.endswith(".Unit()"):
        return unit(eval_vec(e[: -len(".Unit()")], vecs, scalars, vec_type))
    if e.startswith("-") and re.fullmatch(r"-\w+", e) and e[1:] in vecs:
        return [-c for c in vecs[e[1:]]]
    if e in vecs:
        return list(vecs[e])
    m = re.fullmatch(r"(?:%s|Vector3D<\w+>|Vec3D_t)\s*\((.*)\)" % vec_type, e)
    if m:
        args = split_args(m.group(1))
        if len(args) != 3:
            raise Unsupported(e)
        return [eval_scalar(a, scalars) for a in args]
    raise Unsupported(e)


def statements(text):
    """Yield (line_number, statement) with comments removed, joined up to ';'."""
    text = re.sub(r"/\*.*?\*/", lambda m: "\n" * m.group(0).count("\n"), text, flags=re.S)
    cur, start = "", None
    for i, raw in enumerate(text.split("\n"), 1):
        line = raw.split("//")[0].strip()
        if not line:
            continue
        if start is None:
            start = i
        cur += " " + line
        while ";" in cur:
            stmt, cur = cur.split(";", 1)
            yield start, stmt.strip()
            start = i if cur.strip() else None
        if not cur.strip():
            cur, start = "", None


METHODS = "DistanceToIn|DistanceToOut|SafetyToIn|SafetyToOut|Inside|Contains"


def trd_like_triangles(x1, x2, y1, y2, z):
    """Triangles of test/unit_tests/TestTessellated.cpp:18-38 CreateTrdLikeTessellated: six
    quadrilateral facets, each added as the triangles (v0, v1, v2) and (v0, v2, v3)
    (TessellatedStruct::AddQuadrilateralFacet). Returns [ntri, 9 doubles per triangle]."""
    quads = [
        [(-x2, y2, z), (-x2, -y2, z), (x2, -y2, z), (x2, y2, z)],
        [(-x1, y1, -z), (x1, y1, -z), (x1, -y1, -z), (-x1, -y1, -z)],
        [(-x2, -y2, z), (-x1, -y1, -z), (x1, -y1, -z), (x2, -y2, z)],
        [(x2, -y2, z), (x1, -y1, -z), (x1, y1, -z), (x2, y2, z)],
        [(x2, y2, z), (x1, y1, -z), (-x1, y1, -z), (-x2, y2, z)],
        [(-x2, y2, z), (-x1, y1, -z), (-x1, -y1, -z), (-x2, -y2, z)],
    ]
    tris = []
    for q in quads:
        tris.append([q[0], q[1], q[2]])
        tris.append([q[0], q[2], q[3]])
    return [float(len(tris))] + [float(c) for t in tris for v in t for c in v]


def extract(path, shape_t, vec_type="Vec_t", nparams=None, fixed_arrays=None):
    text = open(path).read()
    vecs, scalars, shapes, pending, arrays, extents = {}, {}, {}, {}, {}, {}
    global CURRENT_VECS
    CURRENT_VECS = vecs
    records = []
    fname = os.path.relpath(path, REF_ROOT)

    def shape_params(argstr):
        args = split_args(argstr)
        if args and args[0].startswith('"'):
            args = args[1:]
        vals = []
        for a in args:
            a = a.strip()
            if a in arrays:
                if any(v is None for v in arrays[a]):
                    raise Unsupported(a)
                vals.extend(arrays[a])
            else:
                vals.append(eval_scalar(a, scalars))
        return vals

    for ln, st in statements(text):
        # drop block braces and function headers joined in front of a statement
        while True:
            st2 = re.sub(r"^[^=;]*?[{}]\s*(?=[A-Za-z_!])", "", st)
            st2 = re.sub(r"^.*?\)\s*\{\s*(?=[A-Za-z_!])", "", st2)
            if st2 == st:
                break
            st = st2
        st = re.sub(r"^\s*(?:const\s+)?", "", st)
        # tessellated solids made by the test's CreateTrdLikeTessellated helper
        m = re.match(r"(?:vecgeom::)?SimpleTessellated\s*&\s*(\w+)\s*=\s*\*\s*CreateTrdLikeTessellated\s*<\s*\w+\s*>"
                     r"\s*\((.*)\)$", st)
        if m and shape_t == "SimpleTessellated":
            try:
                shapes[m.group(1)] = trd_like_triangles(*shape_params(m.group(2)))
            except Unsupported:
                shapes.pop(m.group(1), None)
            continue
        # shape definitions
        m = re.match(r"(?:%s)\s*\*?\s*(\w+)\s*(?:=\s*new\s+(?:%s)\s*)?\((.*)\)$" % (shape_t, shape_t), st)
        if m:
            # one or several declarations: T a(...), b(...)
            decls = split_args(st[re.match(r"(?:%s)\s*\*?\s*" % shape_t, st).end():])
            for decl in decls:
                dm = re.fullmatch(r"\*?\s*(\w+)\s*(?:=\s*new\s+(?:%s)\s*)?\((.*)\)" % shape_t, decl.strip(), re.S)
                if not dm:
                    continue
                try:
                    shapes[dm.group(1)] = shape_params(dm.group(2))
                except Unsupported:
                    shapes.pop(dm.group(1), None)
            continue
        # array definitions: T name[] = {a, b, ...} (flattened into shape parameter lists)
        m = re.match(r"(?:Precision|double)\s+(\w+)\s*\[\s*\w*\s*\]\s*=\s*\{(.*)\}$", st)
        if m:
            try:
                arrays[m.group(1)] = [eval_scalar(a, scalars) for a in split_args(m.group(2))]
            except Unsupported:
                arrays.pop(m.group(1), None)
            continue
        # arrays declared with a size and filled element by element: T name[N]; name[i] = v;
        m = re.match(r"(?:Precision|double)\s+(\w+)\s*\[\s*(\w+)\s*\]$", st)
        if m:
            try:
                arrays[m.group(1)] = [None] * int(eval_scalar(m.group(2), scalars))
            except (Unsupported, ValueError):
                arrays.pop(m.group(1), None)
            continue
        m = re.match(r"(\w+)\s*\[\s*(\d+)\s*\]\s*=\s*([^=].*)$", st)
        if m and m.group(1) in arrays:
            try:
                arrays[m.group(1)][int(m.group(2))] = eval_scalar(m.group(3), scalars)
            except (Unsupported, IndexError):
                arrays.pop(m.group(1), None)
            continue
        # integer scalars
        m = re.match(r"(?:int|unsigned int|size_t)\s+(\w+)\s*=\s*([^=]+)$", st)
        if m:
            try:
                scalars[m.group(1)] = eval_scalar(m.group(2), scalars)
            except Unsupported:
                scalars.pop(m.group(1), None)
            continue
        # several scalar definitions in one statement: double a = 1, b = 2 * PI
        m = re.match(r"(?:Precision|double)\s+(\w+\s*=.*,\s*\w+\s*=.*)$", st)
        if m and not re.search(r"\.(%s)\(" % METHODS, st):
            for decl in split_args(m.group(1)):
                dm = re.fullmatch(r"(\w+)\s*=\s*(.+)", decl.strip(), re.S)
                if not dm:
                    continue
                try:
                    scalars[dm.group(1)] = eval_scalar(dm.group(2), scalars)
                except Unsupported:
                    scalars.pop(dm.group(1), None)
            continue
        # scalar definitions
        m = re.match(r"(?:Precision|double)\s+(\w+)\s*=\s*([^=]+)$", st)
        if m and not re.search(r"\.(%s)\(" % METHODS, st):
            try:
                scalars[m.group(1)] = eval_scalar(m.group(2), scalars)
            except Unsupported:
                scalars.pop(m.group(1), None)
            continue
        # vector definitions (possibly several per statement)
        m = re.match(r"(?:%s)\s+(.*)$" % vec_type, st)
        if m and "(" in st and not re.search(r"\.(%s)\(" % METHODS, st):
            for decl in split_args(m.group(1)):
                d1 = re.fullmatch(r"(\w+)\s*\((.*)\)", decl)
                d2 = re.fullmatch(r"(\w+)\s*=\s*(.*)", decl)
                try:
                    if d1:
                        vecs[d1.group(1)] = eval_vec("%s(%s)" % (vec_type, d1.group(2)), vecs, scalars, vec_type)
                    elif d2:
                        vecs[d2.group(1)] = eval_vec(d2.group(2), vecs, scalars, vec_type)
                except Unsupported:
                    name = (d1 or d2).group(1) if (d1 or d2) else None
                    vecs.pop(name, None)
            continue
        # queries stored into a variable
        m = re.match(r"(?:Precision\s+|double\s+|auto\s+|vecgeom::EnumInside\s+|EnumInside\s+|Inside_t\s+)?(\w+)\s*=\s*"
                     r"(\w+)\s*(?:\.|->)\s*(%s)\s*\((.*)\)$" % METHODS, st)
        if m:
            var, sh, meth, args = m.groups()
            pending.pop(var, None)
            if sh in shapes:
                a = split_args(args)
                try:
                    p = eval_vec(a[0], vecs, scalars, vec_type)
                    d = eval_vec(a[1], vecs, scalars, vec_type) if meth in ("DistanceToIn", "DistanceToOut") else None
                    pending[var] = (sh, meth, p, d, ln)
                except (Unsupported, IndexError):
                    pass
            continue
        # shape->Extent(lo, hi): remember which vector holds which corner
        m = re.fullmatch(r"(\w+)\s*(?:\.|->)\s*Extent\s*\(\s*(\w+)\s*,\s*(\w+)\s*\)", st)
        if m:
            if m.group(1) in shapes:
                extents[m.group(2)] = (m.group(1), "lo")
                extents[m.group(3)] = (m.group(1), "hi")
            else:
                extents.pop(m.group(2), None)
                extents.pop(m.group(3), None)
            continue
        # assertions on a stored result
        m = re.match(r"assert\s*\((.*)\)$", st)
        if m:
            body = m.group(1).strip()
            rec = None
            me = re.fullmatch(r"ApproxEqual(?:<\w+>)?\s*\(\s*(\w+)\s*,\s*(.+)\)", body)
            if me and me.group(1) in extents:
                try:
                    sh, which = extents[me.group(1)]
                    v = eval_vec(me.group(2), vecs, scalars, vec_type)
                    records.append({"line": "%s:%d" % (fname, ln), "shape": shapes[sh], "method": "Extent",
                                    "p": None, "d": None, "expect": {"kind": "extent", "which": which, "value": v}})
                except Unsupported:
                    pass
                continue
            # fabs(shape->Method(args) - value) < tol, fabs(shape->Method(args)) < tol
            mq = re.fullmatch(r"(?:std::)?fabs\s*\(\s*\(?\s*(\w+)\s*(?:\.|->)\s*(%s)\s*\(([^()]*(?:\([^()]*\))?[^()]*)\)"
                              r"\s*(?:-\s*([-+\w.()*/ ]+?))?\s*\)?\s*\)\s*<\s*(\w+)" % METHODS, body)
            # shape->Contains(p) / !shape->Contains(p)
            mc = re.fullmatch(r"(!?)\s*(\w+)\s*(?:\.|->)\s*Contains\s*\((.*)\)", body)
            mm = re.fullmatch(r"(\w+)\s*(?:\.|->)\s*Inside\s*\((.*)\)\s*==\s*(?:vecgeom::)?(?:EInside|EnumInside)::[ke](\w+)", body)
            if mq and mq.group(1) in shapes and "||" not in body:
                try:
                    a = split_args(mq.group(3))
                    p = eval_vec(a[0], vecs, scalars, vec_type)
                    meth = mq.group(2)
                    d = eval_vec(a[1], vecs, scalars, vec_type) if meth in ("DistanceToIn", "DistanceToOut") else None
                    val = eval_scalar(mq.group(4), scalars) if mq.group(4) else 0.0
                    tol = eval_scalar(mq.group(5), scalars)
                    rec = (mq.group(1), meth, p, d, ln, {"kind": "absdiff", "value": val, "tol": tol})
                except (Unsupported, IndexError):
                    rec = None
            elif mc and mc.group(2) in shapes:
                try:
                    p = eval_vec(mc.group(3), vecs, scalars, vec_type)
                    rec = (mc.group(2), "Contains", p, None, ln, {"kind": "bool", "value": mc.group(1) != "!"})
                except Unsupported:
                    rec = None
            elif mm and mm.group(1) in shapes:
                try:
                    p = eval_vec(mm.group(2), vecs, scalars, vec_type)
                    rec = (mm.group(1), "Inside", p, None, ln, {"kind": "inside", "value": mm.group(3)})
                except Unsupported:
                    rec = None
            else:
                if "&&" in body and "||" not in body:
                    # a conjunction: every scalar ApproxEqual(var, value) term on a stored result is
                    # a requirement of its own (the other terms, e.g. on normals, are skipped)
                    terms = [t.strip() for t in body.split("&&")]
                    body = next((t for t in terms
                                 if re.fullmatch(r"ApproxEqual(?:<\w+>)?\s*\(\s*(\w+)\s*,.*\)", t)
                                 and re.match(r"ApproxEqual(?:<\w+>)?\s*\(\s*(\w+)", t).group(1) in pending), body)
                alts = re.findall(r"ApproxEqual(?:<\w+>)?\s*\(\s*(\w+)\s*,\s*((?:[^()]|\([^()]*(?:\([^()]*\))?[^()]*\))+)\)", body)
                if alts and all(a[0] == alts[0][0] for a in alts) and alts[0][0] in pending and "&&" not in body:
                    var = alts[0][0]
                    try:
                        vals = [eval_scalar(a[1], scalars) for a in alts]
                        exp = {"kind": "approx", "value": vals[0]} if len(vals) == 1 else {"kind": "anyof", "values": vals}
                        sh, meth, p, d, _ = pending[var]
                        rec = (sh, meth, p, d, ln, exp)
                    except Unsupported:
                        rec = None
                else:
                    mm = re.fullmatch(r"(\w+)\s*==\s*(?:vecgeom::)?(?:EInside|EnumInside)::[ke](\w+)", body)
                    if mm and mm.group(1) in pending and pending[mm.group(1)][1] == "Inside":
                        sh, meth, p, d, _ = pending[mm.group(1)]
                        rec = (sh, meth, p, d, ln, {"kind": "inside", "value": mm.group(2)})
                    mm = re.fullmatch(r"(\w+)\s*(<|<=)\s*0(?:\.0*)?", body)
                    if mm and mm.group(1) in pending:
                        sh, meth, p, d, _ = pending[mm.group(1)]
                        rec = (sh, meth, p, d, ln, {"kind": "negative" if mm.group(2) == "<" else "nonpositive"})
                    # TestWrongSidePoint(Dist, msg) is Dist <= 0 (test/unit_tests/TestSphere.cpp:24-29)
                    mm = re.fullmatch(r"TestWrongSidePoint\s*\(\s*(\w+)\s*,.*\)", body)
                    if mm and mm.group(1) in pending:
                        sh, meth, p, d, _ = pending[mm.group(1)]
                        rec = (sh, meth, p, d, ln, {"kind": "nonpositive"})
            if rec:
                sh, meth, p, d, line, exp = rec
                records.append({"line": "%s:%d" % (fname, line), "shape": shapes[sh], "method": meth,
                                "p": p, "d": d, "expect": exp})
            continue
        # plain re-assignment of a scalar or vector: name = expr
        m = re.fullmatch(r"(\w+)\s*=\s*([^=].*)", st)
        if m and m.group(1) not in shapes:
            name, expr = m.groups()
            pending.pop(name, None)
            try:
                scalars[name] = eval_scalar(expr, scalars)
                vecs.pop(name, None)
                continue
            except Unsupported:
                scalars.pop(name, None)
            try:
                vecs[name] = eval_vec(expr, vecs, scalars, vec_type)
            except Unsupported:
                vecs.pop(name, None)
            continue
        # anything else that mutates a known vector or scalar makes it unknown
        m = re.match(r"(\w+)\s*(?:\.\s*\w+\(\))?\s*(?:[-+*/]?=)", st)
        if m:
            vecs.pop(m.group(1), None)
            scalars.pop(m.group(1), None)
            pending.pop(m.group(1), None)
    return records


SHAPES = {
    # fixture name: (file, shape class pattern, accepted ctor parameter counts)
    "box": ("TestBox.cpp", r"Box_t", (3,)),
    "tube": ("TestTube.cpp", r"Tube_t", (5,)),
    "trapezoid": ("TestTrap.cpp", r"Trap_t", (11,)),
    "cone": ("TestCons.cpp", r"Cone_t", (7,)),
    "orb": ("TestOrb.cpp", r"Orb_t", (1,)),
    "polycone": ("TestPolycone.cpp", r"Polycone_t", None),
    "sphere": ("TestSphere.cpp", r"Sphere_t", (6,)),
    "polyhedron": ("TestPolyhedra.cpp", r"Polyhedra_t", "polyhedron"),
    "torus2": ("TestTorus2.cpp", r"Torus_t", (5,)),
    "tessellated": ("TestTessellated.cpp", r"SimpleTessellated", "tessellated"),
    "hype": ("TestHype.cpp", r"Hype_t", (5,)),
}

if __name__ == "__main__":
    REF_ROOT = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    out_dir = os.path.dirname(os.path.abspath(__file__))
    only = sys.argv[2:] or list(SHAPES)
    for name in only:
        fn, pat, counts = SHAPES[name]
        recs = extract(os.path.join(REF_ROOT, "test", "unit_tests", fn), pat)
        if counts is None:  # sphi dphi nz z[nz] rmin[nz] rmax[nz]
            recs = [r for r in recs if len(r["shape"]) >= 3 and len(r["shape"]) == 3 + 3 * int(r["shape"][2])]
        elif counts == "tessellated":  # ntri, 9 doubles per triangle
            recs = [r for r in recs if len(r["shape"]) == 1 + 9 * int(r["shape"][0])]
        elif counts == "polyhedron":  # sphi dphi nsides nz z[nz] rmin[nz] rmax[nz]
            recs = [r for r in recs if len(r["shape"]) >= 4 and len(r["shape"]) == 4 + 3 * int(r["shape"][3])]
        else:
            recs = [r for r in recs if len(r["shape"]) in counts]
        with open(os.path.join(out_dir, "kat_%s.json" % name), "w") as f:
            json.dump({"source": "test/unit_tests/%s" % fn, "records": recs}, f, indent=0)
        print("%-12s %4d records" % (name, len(recs)))
