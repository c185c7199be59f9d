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
assert(ApproxEqual(lo, Vec_t(..))), method "Extent"); for method "Normal" (valid = s.Normal(p, normal)):
"normal" (ApproxEqual(normal, Vec_t(..))), "normal_valid" (assert([!]valid)), "normal_mag" /
"normal_mag2" (ApproxEqual(normal.Mag[2](), v)). A Normal point written p0 + Dist * v, with Dist a stored
distance query, is recorded as "chain": {"base": p0, "dir": v, "dist": the query} ("p" null).

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
# shapes of the file being read (name -> constructor parameters) and the getters of its shape class
# that the evaluator may call (method name -> function of the parameters), e.g. a test that builds
# its points from para.GetTanThetaCosPhi()
CURRENT_SHAPES = {}
CURRENT_GETTERS = {}
_VEC_CTOR = re.compile(r"(?:Vec_t|Vector3D\s*<\s*\w+\s*>|Vec3D_t)\s*\(")


def _dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]  # Vector3D::Dot, base/Vector3D.h


def eval_expr(expr, scalars, vecs=None):
    """Value of a C++ arithmetic expression over scalars and Vector3D values (lists of 3).
    Vector operations follow base/Vector3D.h: componentwise +, -, scalar * and /; Mag2 = Dot(v, v);
    Mag = sqrt(Mag2); Unit = v / sqrt(Mag2 + kMinimum); Normalized = v * (1 / Mag)."""
    vecs = CURRENT_VECS if vecs is None else vecs
    expr = _norm_expr(expr)
    expr = _VEC_CTOR.sub("__V(", expr)
    try:
        tree = ast.parse(expr, mode="eval")
    except SyntaxError as e:
        raise Unsupported(expr) from e

    def isvec(v):
        return isinstance(v, list)

    def ev(n):
        if isinstance(n, ast.Expression):
            return ev(n.body)
        if isinstance(n, ast.Constant) and isinstance(n.value, (int, float)) and not isinstance(n.value, bool):
            return float(n.value)
        if isinstance(n, ast.UnaryOp) and isinstance(n.op, (ast.USub, ast.UAdd)):
            v = ev(n.operand)
            if isinstance(n.op, ast.UAdd):
                return v
            return [-c for c in v] if isvec(v) else -v
        if isinstance(n, ast.BinOp):
            a, b = ev(n.left), ev(n.right)
            if isinstance(n.op, (ast.Add, ast.Sub)):
                if isvec(a) != isvec(b):
                    raise Unsupported(expr)
                sgn = 1.0 if isinstance(n.op, ast.Add) else -1.0
                if isvec(a):
                    return [a[i] + b[i] if sgn > 0 else a[i] - b[i] for i in range(3)]
                return a + b if sgn > 0 else a - b
            if isinstance(n.op, ast.Mult):
                if isvec(a) and isvec(b):
                    raise Unsupported(expr)
                if isvec(a):
                    return [c * b for c in a]
                if isvec(b):
                    return [a * c for c in b]
                return a * b
            if isinstance(n.op, ast.Div):
                if isvec(b):
                    raise Unsupported(expr)
                return [c / b for c in a] if isvec(a) else a / b
            raise Unsupported(expr)
        if isinstance(n, ast.Name):
            if n.id in CONSTS:
                return CONSTS[n.id]
            if n.id in scalars:
                return scalars[n.id]
            if n.id in vecs:
                return list(vecs[n.id])
            raise Unsupported(n.id)
        if isinstance(n, ast.Call) and isinstance(n.func, ast.Name):
            if n.func.id == "__V":
                if len(n.args) != 3:
                    raise Unsupported(expr)
                vals = [ev(a) for a in n.args]
                if any(isvec(v) for v in vals):
                    raise Unsupported(expr)
                return vals
            if n.func.id in FUNCS:
                args = [ev(a) for a in n.args]
                if any(isvec(v) for v in args):
                    raise Unsupported(expr)
                return float(FUNCS[n.func.id](*args))
        if (isinstance(n, ast.Call) and isinstance(n.func, ast.Attribute) and not n.args
                and isinstance(n.func.value, ast.Name) and n.func.value.id in CURRENT_SHAPES
                and n.func.attr in CURRENT_GETTERS):
            return float(CURRENT_GETTERS[n.func.attr](CURRENT_SHAPES[n.func.value.id]))
        if isinstance(n, ast.Call) and isinstance(n.func, ast.Attribute) and not n.args:
            v = ev(n.func.value)
            if not isvec(v):
                raise Unsupported(expr)
            at = n.func.attr
            if at in ("x", "y", "z"):
                return v["xyz".index(at)]
            if at in ("Mag2", "Length2"):
                return _dot(v, v)
            if at in ("Mag", "Length"):
                return math.sqrt(_dot(v, v))
            if at == "Unit":
                return unit(v)
            if at == "Normalized":
                f = 1.0 / math.sqrt(_dot(v, v))
                return [c * f for c in v]
        raise Unsupported(expr)

    return ev(tree)


def eval_scalar(expr, scalars):
    v = eval_expr(expr, scalars)
    if isinstance(v, list):
        raise Unsupported(expr)
    return v


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
    v = eval_expr(expr, scalars, vecs)
    if not isinstance(v, list):
        raise Unsupported(expr)
    return v


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


def _strip_parens(e):
    e = e.strip()
    while e.startswith("(") and _match_paren(e, 0) == len(e) - 1:
        e = e[1:-1].strip()
    return e


def _match_paren(e, i):
    """Index of the parenthesis closing the one at e[i], or -1."""
    depth = 0
    for j in range(i, len(e)):
        if e[j] == "(":
            depth += 1
        elif e[j] == ")":
            depth -= 1
            if depth == 0:
                return j
    return -1


def split_call(e):
    """'obj->Method(args) rest' -> (obj, Method, args, rest), else None."""
    m = re.match(r"\s*(\w+)\s*(?:\.|->)\s*(%s)\s*\(" % METHODS, e)
    if not m:
        return None
    j = _match_paren(e, m.end() - 1)
    if j < 0:
        return None
    return m.group(1), m.group(2), e[m.end():j], e[j + 1:].strip()


def _query_point_dir(meth, argstr, vecs, scalars, vec_type):
    a = split_args(argstr)
    p = eval_vec(a[0], vecs, scalars, vec_type)
    d = eval_vec(a[1], vecs, scalars, vec_type) if meth in ("DistanceToIn", "DistanceToOut") else None
    return p, d


def direct_query_assert(body, shapes, vecs, scalars, vec_type, ln):
    """Asserts that call the shape inside the assert:
    fabs(s->M(args) [- v]) < tol  -> absdiff
    s->M(args) == v               -> exact (Contains: == true/false -> bool)
    s->M(args) <,<=,>,>= 0        -> negative / nonpositive / positive / nonnegative
    [!]s->Contains(p)             -> bool
    s->Inside(p) == EInside::kX   -> inside
    Returns a record tuple, or None when the body has none of these forms (or the shape is unknown)."""
    b = _strip_parens(body)
    if "||" in b or "&&" in b:
        return None
    m = re.match(r"(?:std::)?fabs\s*\(", b)
    if m:
        j = _match_paren(b, m.end() - 1)
        if j < 0:
            return None
        inner, rest = _strip_parens(b[m.end():j]), b[j + 1:].strip()
        c = split_call(inner)
        mt = re.fullmatch(r"<\s*(.+)", rest)
        if not c or not mt or c[0] not in shapes:
            return None
        sh, meth, argstr, tail = c
        p, d = _query_point_dir(meth, argstr, vecs, scalars, vec_type)
        val = 0.0
        if tail:
            mv = re.fullmatch(r"-\s*(.+)", tail)
            if not mv:
                return None
            val = eval_scalar(mv.group(1), scalars)
        return (sh, meth, p, d, ln, {"kind": "absdiff", "value": val, "tol": eval_scalar(mt.group(1), scalars)})
    neg = b.startswith("!")
    c = split_call(b[1:] if neg else b)
    if not c or c[0] not in shapes:
        return None
    sh, meth, argstr, rest = c
    p, d = _query_point_dir(meth, argstr, vecs, scalars, vec_type)
    if meth == "Contains":
        if rest == "":
            return (sh, meth, p, None, ln, {"kind": "bool", "value": not neg})
        mb = re.fullmatch(r"==\s*(true|false)", rest)
        if mb and not neg:
            return (sh, meth, p, None, ln, {"kind": "bool", "value": mb.group(1) == "true"})
        return None
    if neg:
        return None
    if meth == "Inside":
        mi = re.fullmatch(r"==\s*(?:vecgeom::)?(?:(?:EInside|EnumInside)::)?[ke](Inside|Outside|Surface)", rest)
        return (sh, meth, p, None, ln, {"kind": "inside", "value": mi.group(1)}) if mi else None
    mc = re.fullmatch(r"(==|<=|<|>=|>)\s*(.+)", rest)
    if not mc:
        return None
    op, rhs = mc.groups()
    if op == "==":
        return (sh, meth, p, d, ln, {"kind": "exact", "value": eval_scalar(rhs, scalars)})
    if eval_scalar(rhs, scalars) != 0.0:
        return None
    kind = {"<": "negative", "<=": "nonpositive", ">": "positive", ">=": "nonnegative"}[op]
    return (sh, meth, p, d, ln, {"kind": kind})


def skip_reason(body, shapes, pending, unknown_results):
    """Why an assert produced no record (for kat_stats.json)."""
    c = None
    for mcall in re.finditer(r"(\w+)\s*(?:\.|->)\s*(%s)\s*\(" % METHODS, body):
        c = mcall
        break
    names = set(re.findall(r"\b(\w+)\b", body))
    if c is not None:
        if c.group(1) not in shapes:
            return "shape not reconstructed by the extractor"
        return "query form not recognised"
    res = [n for n in names if n in pending or n in unknown_results]
    if res:
        why = unknown_results.get(res[0])
        if why:
            return why
        return "comparison on a query result not recognised"
    if re.search(r"\b(?:normal|Normal|valid|norm\w*)\b", body):
        return "Normal() of a shape or point the extractor does not reconstruct (loops, unknown shapes)"
    return "not a query assert (capacity, surface area, getters, bounding box)"


def extract(path, shape_t, vec_type="Vec_t", unplaced=None, stats=None, getters=None, setters=None):
    text = open(path).read()
    vecs, scalars, shapes, pending, arrays, extents, logicals = {}, {}, {}, {}, {}, {}, {}
    global CURRENT_SHAPES, CURRENT_GETTERS
    CURRENT_SHAPES, CURRENT_GETTERS = shapes, dict(getters or {})
    unknown_results = {}
    normals, valids = {}, {}  # Normal() output vector / validity variables -> (shape, p, chain, line)
    vec_arrays = set()
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
                ma = re.fullmatch(r"(\w+)\s*=\s*(.+)", a)  # an argument that also assigns: Shape(n, a = 3., ..)
                if ma:
                    scalars[ma.group(1)] = eval_scalar(ma.group(2), scalars)
                    vals.append(scalars[ma.group(1)])
                    continue
                v = eval_expr(a, scalars, vecs)  # a vector argument (Tet's corners) adds its components
                vals.extend(v if isinstance(v, list) else [v])
        return vals

    def normal_point(expr):
        """The point argument of a Normal call: (p, None), or (None, chain) for p0 + Dist * v where Dist
        holds a stored distance query (the test derives the surface point from the query's result)."""
        try:
            return eval_vec(expr, vecs, scalars, vec_type), None
        except (Unsupported, IndexError, ValueError):
            pass
        e = _strip_parens(expr.strip())
        m = (re.fullmatch(r"(\w+)\s*([-+])\s*(\w+)\s*\*\s*(\w+)", e))
        if m:
            a, sign, x, y = m.groups()
            dvar, vvar = (x, y) if x in pending else (y, x)
            if a in vecs and vvar in vecs and dvar in pending and pending[dvar][1] in ("DistanceToIn",
                                                                                        "DistanceToOut"):
                sh2, meth2, p2, d2, _ = pending[dvar]
                # p0 - Dist * v is p0 + Dist * (-v) in floating point as well
                v = [-c for c in vecs[vvar]] if sign == "-" else list(vecs[vvar])
                return None, {"base": list(vecs[a]), "dir": v,
                              "dist": {"shape": shapes[sh2], "method": meth2, "p": p2, "d": d2}}
        raise Unsupported(expr)

    def normal_assert_term(t):
        """One term of an assert on a Normal() result: an expectation dict with its call, or None."""
        t = _strip_parens(t.strip())
        mv = re.fullmatch(r"(!?)\s*(\w+)(?:\s*==\s*(true|false))?", t)
        if mv and mv.group(2) in valids:
            val = (mv.group(3) or "true") == "true"
            return valids[mv.group(2)], {"kind": "normal_valid", "value": val != bool(mv.group(1))}
        mn = re.fullmatch(r"ApproxEqual(?:<\w+>)?\s*\(\s*(\w+)\s*,\s*(.+)\)", t, re.S)
        if mn and mn.group(1) in normals:
            return normals[mn.group(1)], {"kind": "normal", "value": eval_vec(mn.group(2), vecs, scalars,
                                                                               vec_type)}
        mm = re.fullmatch(r"ApproxEqual(?:<\w+>)?\s*\(\s*(\w+)\s*\.\s*(Mag2?)\s*\(\s*\)\s*,\s*(.+)\)", t, re.S)
        if mm and mm.group(1) in normals:
            kind = "normal_mag2" if mm.group(2) == "Mag2" else "normal_mag"
            return normals[mm.group(1)], {"kind": kind, "value": eval_scalar(mm.group(3), scalars)}
        mc = re.fullmatch(r"(\w+)\s*(?:\.|->)\s*Normal\s*\((.*)\)", t, re.S)
        if mc and mc.group(1) in shapes:
            a = split_args(mc.group(2))
            pt, chain = normal_point(a[0])
            call = (mc.group(1), pt, chain, ln)
            normals[a[1].strip()] = call
            vecs.pop(a[1].strip(), None)
            return call, {"kind": "normal_valid", "value": True}
        return None

    def normal_records(body):
        """Records of an assert that tests a Normal() result, alone or in a conjunction with distance
        terms (assert(ApproxEqual(Dist, 20) && ApproxEqual(normal, vx))). None when the body names no
        Normal() result."""
        names = set(re.findall(r"\b(\w+)\b", body))
        if not (names & (set(normals) | set(valids))) and not re.search(r"(?:\.|->)\s*Normal\s*\(", body):
            return None
        if "||" in body:
            return []
        out = []
        for t in body.split("&&"):
            try:
                r = normal_assert_term(t)
            except (Unsupported, IndexError, ValueError, ZeroDivisionError):
                return []
            if r is not None:
                (sh, pt, chain, line), exp = r
                out.append({"line": "%s:%d" % (fname, ln), "shape": shapes[sh], "method": "Normal", "p": pt,
                            "d": None, "chain": chain, "expect": exp})
                continue
            rec, _ = query_assert(t.strip())
            if rec is None:
                return []
            sh, meth, p, d, line, exp = rec
            out.append({"line": "%s:%d" % (fname, line), "shape": shapes[sh], "method": meth, "p": p, "d": d,
                        "expect": exp})
        return out

    def query_assert(body):
        """One assert body on the shape queries: (record tuple or None, reason or None)."""
        rec, why = None, None
        try:
            rec = direct_query_assert(body, shapes, vecs, scalars, vec_type, ln)
        except (Unsupported, IndexError, ValueError, ZeroDivisionError):
            rec, why = None, "argument or expected value not evaluable"
        if rec is None and why is None:
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
                mm = re.fullmatch(r"(\w+)\s*==\s*(?:vecgeom::)?(?:(?:EInside|EnumInside)::)?[ke](Inside|Outside|Surface)", body)
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
                mm = re.fullmatch(r"(\w+)\s*(>|>=)\s*0(?:\.0*)?", body)
                if mm and mm.group(1) in pending:
                    sh, meth, p, d, _ = pending[mm.group(1)]
                    rec = (sh, meth, p, d, ln, {"kind": "positive" if mm.group(2) == ">" else "nonnegative"})
                # d == value (exact), d >= kInfLength (a miss)
                mm = re.fullmatch(r"(\w+)\s*(==|>=)\s*(.+)", body)
                if rec is None and mm and mm.group(1) in pending and pending[mm.group(1)][1] != "Inside":
                    try:
                        v = eval_scalar(mm.group(3), scalars)
                        sh, meth, p, d, _ = pending[mm.group(1)]
                        if mm.group(2) == "==":
                            rec = (sh, meth, p, d, ln, {"kind": "exact", "value": v})
                        elif v == DBL_MAX:
                            rec = (sh, meth, p, d, ln, {"kind": "exact", "value": v})
                    except Unsupported:
                        why = "argument or expected value not evaluable"
        return rec, why

    for ln, st in statements(text):
        # drop block braces and function headers joined in front of a statement
        while True:
            st2 = re.sub(r"^[^=;]*?[{}]\s*(?=[A-Za-z_!])", "", st)
            st2 = re.sub(r"^.*?\)\s*\{\s*(?=[A-Za-z_!])", "", st2)
            if st2 == st:
                break
            st = st2
        st = re.sub(r"^\s*(?:const\s+)?", "", st)
        st = re.sub(r"^(?:vecgeom::(?:cxx::)?|VECGEOM_NAMESPACE::)(?=[A-Za-z_])", "", st)
        # shapes made by GeoManager::MakeInstance<UnplacedX>(args)
        m = re.match(r"(?:auto|\w+)\s*\*?\s*(\w+)\s*=\s*(?:vecgeom::)?GeoManager::MakeInstance\s*<\s*"
                     r"(?:vecgeom::)?(?:cxx::)?(\w+)\s*>\s*\((.*)\)$", st)
        if m:
            if unplaced and re.fullmatch(unplaced, m.group(2)):
                try:
                    shapes[m.group(1)] = shape_params(m.group(3))
                except Unsupported:
                    shapes.pop(m.group(1), None)
            else:
                shapes.pop(m.group(1), None)
            continue
        # identity placements: (new LogicalVolume("n", s))->Place([new Transformation3D()]),
        # or LogicalVolume l("n", s) followed by auto p = l.Place()
        m = re.match(r"LogicalVolume\s+(\w+)\s*\(\s*\"[^\"]*\"\s*,\s*&?\s*(\w+)\s*\)$", st)
        if m:
            logicals[m.group(1)] = m.group(2)
            continue
        m = re.match(r"(?:VPlacedVolume\s+const\s*\*|VPlacedVolume\s*\*|auto\s*\*?)\s*(\w+)\s*=\s*"
                     r"(?:\(\s*new\s+LogicalVolume\s*\(\s*\"[^\"]*\"\s*,\s*&?\s*(\w+)\s*\)\s*\)|(\w+))\s*(?:->|\.)\s*"
                     r"Place\s*\(\s*(?:new\s+Transformation3D\s*\(\s*\))?\s*\)$", st)
        if m:
            src = m.group(2) or logicals.get(m.group(3))
            if src in shapes:
                shapes[m.group(1)] = shapes[src]
            else:
                shapes.pop(m.group(1), None)
            continue
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
        # arrays of vectors (trapezoid corners): Vec_t name[8] = {Vec_t(..), ...}; Vec_t name[8]; name[i] = ..
        m = re.match(r"(?:%s)\s+(\w+)\s*\[\s*(\w+)\s*\]\s*(?:=\s*\{(.*)\})?$" % vec_type, st)
        if m:
            try:
                if m.group(3) is not None:
                    arrays[m.group(1)] = [c for e in split_args(m.group(3))
                                          for c in eval_vec(e, vecs, scalars, vec_type)]
                else:
                    arrays[m.group(1)] = [None] * (3 * int(eval_scalar(m.group(2), scalars)))
                vec_arrays.add(m.group(1))
            except (Unsupported, ValueError):
                arrays.pop(m.group(1), None)
            continue
        m = re.match(r"(\w+)\s*\[\s*(\d+)\s*\]\s*=\s*([^=].*)$", st)
        if m and m.group(1) in vec_arrays and m.group(1) in arrays:
            try:
                i = int(m.group(2))
                arrays[m.group(1)][3 * i:3 * i + 3] = eval_vec(m.group(3), vecs, scalars, vec_type)
            except (Unsupported, IndexError):
                arrays.pop(m.group(1), None)
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
        # Normal calls: [bool] valid = s.Normal(p, normal) or s.Normal(p, normal)
        m = re.match(r"(?:bool\s+)?(?:(\w+)\s*=\s*)?(\w+)\s*(?:\.|->)\s*Normal\s*\((.*)\)$", st)
        if m:
            vvar, sh, args = m.groups()
            a = split_args(args)
            nvar = a[1].strip() if len(a) == 2 else None
            for v in (vvar, nvar):
                if v:
                    normals.pop(v, None), valids.pop(v, None), pending.pop(v, None), vecs.pop(v, None)
                    scalars.pop(v, None)
            if sh in shapes and nvar:
                try:
                    pt, chain = normal_point(a[0])
                    call = (sh, pt, chain, ln)
                    normals[nvar] = call
                    if vvar:
                        valids[vvar] = call
                except Unsupported:
                    pass
            continue
        # queries stored into a variable
        m = re.match(r"(?:Precision\s+|double\s+|auto\s+|vecgeom::EnumInside\s+|EnumInside\s+|Inside_t\s+)?(\w+)\s*=\s*"
                     r"(\w+)\s*(?:\.|->)\s*(%s)\s*\((.*)\)$" % METHODS, st)
        if m:
            var, sh, meth, args = m.groups()
            pending.pop(var, None)
            unknown_results.pop(var, None)
            if sh in shapes:
                a = split_args(args)
                try:
                    p = eval_vec(a[0], vecs, scalars, vec_type)
                    d = eval_vec(a[1], vecs, scalars, vec_type) if meth in ("DistanceToIn", "DistanceToOut") else None
                    pending[var] = (sh, meth, p, d, ln)
                except (Unsupported, IndexError):
                    unknown_results[var] = "argument or expected value not evaluable"
            else:
                unknown_results[var] = "shape not reconstructed by the extractor"
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
                ok = False
                try:
                    sh, which = extents[me.group(1)]
                    v = eval_vec(me.group(2), vecs, scalars, vec_type)
                    records.append({"line": "%s:%d" % (fname, ln), "shape": shapes[sh], "method": "Extent",
                                    "p": None, "d": None, "expect": {"kind": "extent", "which": which, "value": v}})
                    ok = True
                except Unsupported:
                    pass
                if stats is not None:
                    stats["asserts"] += 1
                    if ok:
                        stats["extracted"] += 1
                    else:
                        stats["skipped"].setdefault(
                            "not a query assert (capacity, surface area, getters, bounding box)", []).append(ln)
                continue
            nrecs = normal_records(body)
            if nrecs is not None:
                records.extend(nrecs)
                if stats is not None:
                    stats["asserts"] += 1
                    if nrecs:
                        stats["extracted"] += 1
                    else:
                        stats["skipped"].setdefault("Normal() at a point the extractor cannot evaluate (sampled surface points)",
                                                    []).append(ln)
                continue
            rec, why = query_assert(body)
            if rec:
                sh, meth, p, d, line, exp = rec
                records.append({"line": "%s:%d" % (fname, line), "shape": shapes[sh], "method": meth,
                                "p": p, "d": d, "expect": exp})
            if stats is not None:
                stats["asserts"] += 1
                if rec:
                    stats["extracted"] += 1
                else:
                    why = why or skip_reason(body, shapes, pending, unknown_results)
                    stats["skipped"].setdefault(why, []).append(ln)
            continue
        # plain re-assignment of a scalar or vector: name = expr
        m = re.fullmatch(r"(\w+)\s*=\s*([^=].*)", st)
        if m and m.group(1) not in shapes:
            name, expr = m.groups()
            pending.pop(name, None)
            normals.pop(name, None)
            valids.pop(name, None)
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
        # v.Set(x, y, z) / v.Set(a) (Vector3D::Set, base/Vector3D.h)
        m = re.fullmatch(r"(\w+)\s*\.\s*Set\s*\((.*)\)", st)
        if m and m.group(1) not in shapes:
            name = m.group(1)
            pending.pop(name, None), normals.pop(name, None), valids.pop(name, None)
            try:
                a = [eval_scalar(x, scalars) for x in split_args(m.group(2))]
                if len(a) not in (1, 3):
                    raise Unsupported(st)
                vecs[name] = a * 3 if len(a) == 1 else a
            except Unsupported:
                vecs.pop(name, None)
            continue
        # in-place vector updates: v.Normalize(); v op= expr; v.x() = expr
        m = re.fullmatch(r"(\w+)\s*\.\s*Normalize\s*\(\s*\)", st)
        if m and m.group(1) in vecs:
            v = vecs[m.group(1)]
            f = 1.0 / math.sqrt(_dot(v, v))  # Vector3D::Normalize, base/Vector3D.h:198
            vecs[m.group(1)] = [c * f for c in v]
            continue
        m = re.fullmatch(r"(\w+)\s*([-+*/])=\s*(.+)", st)
        if m and m.group(1) in vecs:
            try:
                rhs = eval_expr(m.group(3), scalars, vecs)
                v, op = vecs[m.group(1)], m.group(2)
                if op in "+-" and isinstance(rhs, list):
                    vecs[m.group(1)] = [v[i] + rhs[i] if op == "+" else v[i] - rhs[i] for i in range(3)]
                elif op in "*/" and not isinstance(rhs, list):
                    vecs[m.group(1)] = [c * rhs if op == "*" else c / rhs for c in v]
                else:
                    vecs.pop(m.group(1), None)
            except Unsupported:
                vecs.pop(m.group(1), None)
            continue
        m = re.fullmatch(r"(\w+)\s*\.\s*([xyz])\s*\(\s*\)\s*=\s*([^=].*)", st)
        if m and m.group(1) in vecs:
            try:
                vecs[m.group(1)]["xyz".index(m.group(2))] = eval_scalar(m.group(3), scalars)
            except Unsupported:
                vecs.pop(m.group(1), None)
            continue
        # a setter of a known shape changes its parameters (a setter the fixture does not describe, or
        # arguments the evaluator cannot reach, make the shape unknown)
        m = re.match(r"(\w+)\s*(?:\.|->)\s*(Set\w*)\s*\((.*)\)$", st)
        if m and m.group(1) in shapes:
            f = (setters or {}).get(m.group(2))
            try:
                if f is None:
                    raise Unsupported(m.group(2))
                shapes[m.group(1)] = f(list(shapes[m.group(1)]), shape_params(m.group(3)))
            except Unsupported:
                shapes.pop(m.group(1), None)
            continue
        # anything else that mutates a known vector or scalar makes it unknown
        m = re.match(r"(\w+)\s*(?:\.\s*\w+\(\))?\s*(?:[-+*/]?=)", st)
        if m:
            vecs.pop(m.group(1), None)
            scalars.pop(m.group(1), None)
            pending.pop(m.group(1), None)
    return records


SHAPES = {
    # fixture name: (file, shape class pattern, accepted ctor parameter counts, unplaced class pattern)
    "box": ("TestBox.cpp", r"Box_t", (3,), r"UnplacedBox"),
    "tube": ("TestTube.cpp", r"Tube_t", (5,), r"(?:S?Unplaced|G)Tube\w*"),
    "trapezoid": ("TestTrap.cpp", r"Trap_t", (3, 4, 5, 6, 11, 24), r"UnplacedTrapezoid"),
    "cone": ("TestCons.cpp", r"Cone_t", (7,), r"S?UnplacedCone\w*"),
    "orb": ("TestOrb.cpp", r"Orb_t", (1,), r"UnplacedOrb"),
    "polycone": ("TestPolycone.cpp", r"Polycone_t", None, r"UnplacedPolycone"),
    "sphere": ("TestSphere.cpp", r"Sphere_t", (6,), r"UnplacedSphere"),
    "polyhedron": ("TestPolyhedra.cpp", r"Polyhedra_t", "polyhedron", r"UnplacedPolyhedron"),
    "torus2": ("TestTorus2.cpp", r"Torus_t", (5,), r"UnplacedTorus2"),
    "tessellated": ("TestTessellated.cpp", r"SimpleTessellated", "tessellated", None),
    "hype": ("TestHype.cpp", r"Hype_t", (5,), r"S?UnplacedHype\w*"),
    "trd": ("TestTrd.cpp", r"Trd_t", (3, 4, 5), r"UnplacedTrd"),
    "paraboloid": ("TestParaboloid.cpp", r"Paraboloid_t", (3,), r"UnplacedParaboloid"),
    "parallelepiped": ("TestParallelepiped.cpp", r"Parallelepiped_t", (6,), r"UnplacedParallelepiped"),
    "tet": ("TestTet.cpp", r"Tet_t", (12,), r"UnplacedTet"),
    "ellipticaltube": ("TestEllipticalTube.cpp", r"EllipticalTube_t", (3,), r"UnplacedEllipticalTube"),
    "ellipsoid": ("TestEllipsoid.cpp", r"Ellipsoid_t", (5,), r"UnplacedEllipsoid"),
    "gentrap": ("TestGenTrap.cpp", r"SimpleGenTrap", (17,), r"UnplacedGenTrap"),
    "ellipticalcone": ("TestEllipticalCone.cpp", r"EllipticalCone_t", (4,), r"UnplacedEllipticalCone"),
}


# Setters of a shape class a test calls between queries: the constructor parameters they replace
# (UnplacedEllipsoid::SetSemiAxes / SetZCuts, UnplacedEllipticalTube::SetParameters / SetDx / SetDy / SetDz).
SETTERS = {
    "ellipsoid": {"SetSemiAxes": lambda p, a: a[:3] + p[3:], "SetZCuts": lambda p, a: p[:3] + a[:2]},
    "ellipticalcone": {"SetParameters": lambda p, a: a[:4]},
    "ellipticaltube": {"SetParameters": lambda p, a: a[:3], "SetDx": lambda p, a: [a[0]] + p[1:],
                       "SetDy": lambda p, a: p[:1] + [a[0]] + p[2:], "SetDz": lambda p, a: p[:2] + [a[0]]},
}

# Getters of a shape class a test calls while building its points: ParallelepipedStruct's cached
# tangents (VecGeom/volumes/ParallelepipedStruct.h:80-88), computed as the reference computes them.
GETTERS = {
    "parallelepiped": {
        "GetTanThetaCosPhi": lambda a: math.tan(a[4]) * math.cos(a[5]),
        "GetTanThetaSinPhi": lambda a: math.tan(a[4]) * math.sin(a[5]),
        "GetTanAlpha": lambda a: math.tan(a[3]),
        "GetX": lambda a: a[0], "GetY": lambda a: a[1], "GetZ": lambda a: a[2],
    },
}


def params_ok(counts, shape):
    if counts is None:  # sphi dphi nz z[nz] rmin[nz] rmax[nz]
        return len(shape) >= 3 and len(shape) == 3 + 3 * int(shape[2])
    if counts == "tessellated":  # ntri, 9 doubles per triangle
        return len(shape) == 1 + 9 * int(shape[0])
    if counts == "polyhedron":  # sphi dphi nsides nz z[nz] rmin[nz] rmax[nz]
        return len(shape) >= 4 and len(shape) == 4 + 3 * int(shape[3])
    return len(shape) in counts


if __name__ == "__main__":
    REF_ROOT = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    out_dir = os.path.dirname(os.path.abspath(__file__))
    only = sys.argv[2:] or list(SHAPES)
    stats_path = os.path.join(out_dir, "kat_stats.json")
    all_stats = json.load(open(stats_path)) if os.path.exists(stats_path) else {}
    for name in only:
        fn, pat, counts, unplaced = SHAPES[name]
        st = {"source": "test/unit_tests/%s" % fn, "asserts": 0, "extracted": 0, "skipped": {}}
        recs = extract(os.path.join(REF_ROOT, "test", "unit_tests", fn), pat, unplaced=unplaced, stats=st,
                       getters=GETTERS.get(name), setters=SETTERS.get(name))
        kept = []
        for r in recs:
            if params_ok(counts, r["shape"]):
                kept.append(r)
            else:
                st["extracted"] -= 1
                st["skipped"].setdefault("constructor form not supported (%d parameters)" % len(r["shape"]),
                                         []).append(int(r["line"].rsplit(":", 1)[1]))
        st["records"] = len(kept)
        st["skipped"] = {k: sorted(v) for k, v in sorted(st["skipped"].items())}
        all_stats[name] = st
        with open(os.path.join(out_dir, "kat_%s.json" % name), "w") as f:
            json.dump({"source": "test/unit_tests/%s" % fn, "records": kept}, f, indent=0)
        print("%-12s %4d records  %4d asserts, %4d extracted, %4d skipped" % (
            name, len(kept), st["asserts"], st["extracted"], sum(len(v) for v in st["skipped"].values())))
    with open(stats_path, "w") as f:
        json.dump(all_stats, f, indent=1, sort_keys=True)
