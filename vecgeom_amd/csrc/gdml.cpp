// GDML reader: a flat geometry description (vgmi_geom_desc) from a GDML file, following what the
// reference's GDML middleware builds (persistency/gdml/source/src/Middleware.cpp) without Xerces:
// a small XML parser of its own reads the document into an element tree, which is walked the way
// Middleware::processNode walks the DOM (:168-216).
//
// Conventions kept from the middleware, so that a file yields the same solids, volumes and
// placements as the reference:
//  - a numeric attribute is the number at the start of its text (std::stod, Helper.cpp:59-68), else
//    the value of the <constant> of that name, 0 for an unknown name (:144-153); <variable> and
//    <quantity> are not read (processNode handles neither);
//  - lengths carry the element's lunit ("unit" on <position>): mm (the default), m, km, um, nm, cm,
//    times the internal value of one millimetre (GeoManager's default 0.1, i.e. centimetres); angles
//    carry aunit ("unit" on <rotation>): deg, else radians (:674-704);
//  - full lengths are halved where the middleware halves them (box, tube, cone, trd, trap, hype z);
//  - each solid goes through GeoManager::MakeInstance, whose substitutions are applied here
//    (make_instance below);
//  - logical volumes are numbered in document order, placements in the order of their <physvol>
//    elements, the world's placement last; a placement's rotation is the one of
//    makeRotationMatrixFromCartesianAngles (:114-132).
#include "gdml_reader.h"

#include "shape_data.h"
#include "vgmi_internal.h"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

namespace vgmi {
namespace gdml {

// ------------------------------------------------------------------------------------ XML
struct Element {
  std::string name;
  std::vector<std::pair<std::string, std::string>> attrs;
  std::vector<Element> kids;
  // the attribute's text, "" when absent (Helper::GetAttribute)
  std::string attr(const char *key) const
  {
    for (auto const &a : attrs)
      if (a.first == key) return a.second;
    return std::string();
  }
};

class XmlParser {
public:
  explicit XmlParser(std::string const &text) : s_(text) {}

  bool parse(Element &root, std::string &err)
  {
    skip_misc();
    if (!parse_element(root, 0)) {
      err = error_;
      return false;
    }
    skip_misc();
    if (i_ != s_.size()) {
      err = where("content after the root element");
      return false;
    }
    return true;
  }

private:
  std::string const &s_;
  size_t i_ = 0;
  std::string error_;

  std::string where(std::string const &what) const
  {
    size_t line = 1;
    for (size_t k = 0; k < i_ && k < s_.size(); ++k) line += s_[k] == '\n';
    std::ostringstream o;
    o << "XML, line " << line << ": " << what;
    return o.str();
  }
  bool fail(std::string const &what)
  {
    if (error_.empty()) error_ = where(what);
    return false;
  }
  bool starts(const char *t) const { return s_.compare(i_, std::strlen(t), t) == 0; }
  void skip_ws()
  {
    while (i_ < s_.size() && (s_[i_] == ' ' || s_[i_] == '\t' || s_[i_] == '\n' || s_[i_] == '\r')) ++i_;
  }
  // comments, processing instructions, a DOCTYPE and whitespace between elements
  void skip_misc()
  {
    for (;;) {
      skip_ws();
      if (starts("<!--")) {
        const size_t e = s_.find("-->", i_ + 4);
        i_             = e == std::string::npos ? s_.size() : e + 3;
      } else if (starts("<?")) {
        const size_t e = s_.find("?>", i_ + 2);
        i_             = e == std::string::npos ? s_.size() : e + 2;
      } else if (starts("<!DOCTYPE")) {
        int depth = 0;
        for (; i_ < s_.size(); ++i_) {
          if (s_[i_] == '[') ++depth;
          if (s_[i_] == ']') --depth;
          if (s_[i_] == '>' && depth <= 0) {
            ++i_;
            break;
          }
        }
      } else {
        return;
      }
    }
  }
  static bool name_char(char c)
  {
    return std::isalnum(static_cast<unsigned char>(c)) || c == '_' || c == ':' || c == '-' || c == '.';
  }
  bool parse_name(std::string &out)
  {
    const size_t b = i_;
    while (i_ < s_.size() && name_char(s_[i_])) ++i_;
    if (i_ == b) return fail("expected a name");
    out.assign(s_, b, i_ - b);
    return true;
  }
  static void decode(std::string &v)
  {
    static const std::pair<const char *, char> ents[] = {{"&lt;", '<'},   {"&gt;", '>'},   {"&quot;", '"'},
                                                         {"&apos;", '\''}, {"&amp;", '&'}};
    std::string out;
    for (size_t k = 0; k < v.size();) {
      bool hit = false;
      if (v[k] == '&')
        for (auto const &e : ents)
          if (v.compare(k, std::strlen(e.first), e.first) == 0) {
            out += e.second, k += std::strlen(e.first), hit = true;
            break;
          }
      if (!hit) out += v[k++];
    }
    v.swap(out);
  }
  bool parse_element(Element &el, int depth)
  {
    if (depth > 256) return fail("elements nested too deeply");
    if (i_ >= s_.size() || s_[i_] != '<') return fail("expected '<'");
    ++i_;
    if (!parse_name(el.name)) return false;
    for (;;) {
      skip_ws();
      if (i_ >= s_.size()) return fail("unterminated start tag <" + el.name + ">");
      if (starts("/>")) {
        i_ += 2;
        return true;
      }
      if (s_[i_] == '>') {
        ++i_;
        break;
      }
      std::string key, val;
      if (!parse_name(key)) return false;
      skip_ws();
      if (i_ >= s_.size() || s_[i_] != '=') return fail("expected '=' after attribute " + key);
      ++i_;
      skip_ws();
      if (i_ >= s_.size() || (s_[i_] != '"' && s_[i_] != '\'')) return fail("expected a quoted value for " + key);
      const char q   = s_[i_++];
      const size_t e = s_.find(q, i_);
      if (e == std::string::npos) return fail("unterminated value of attribute " + key);
      val.assign(s_, i_, e - i_);
      i_ = e + 1;
      decode(val);
      el.attrs.emplace_back(std::move(key), std::move(val));
    }
    // content: child elements; text, comments and CDATA are skipped
    for (;;) {
      while (i_ < s_.size() && s_[i_] != '<') ++i_;
      if (i_ >= s_.size()) return fail("missing </" + el.name + ">");
      if (starts("<!--")) {
        const size_t e = s_.find("-->", i_ + 4);
        if (e == std::string::npos) return fail("unterminated comment");
        i_ = e + 3;
      } else if (starts("<![CDATA[")) {
        const size_t e = s_.find("]]>", i_ + 9);
        if (e == std::string::npos) return fail("unterminated CDATA section");
        i_ = e + 3;
      } else if (starts("<?")) {
        const size_t e = s_.find("?>", i_ + 2);
        if (e == std::string::npos) return fail("unterminated processing instruction");
        i_ = e + 2;
      } else if (starts("</")) {
        i_ += 2;
        std::string close;
        if (!parse_name(close)) return false;
        if (close != el.name) return fail("</" + close + "> closes <" + el.name + ">");
        skip_ws();
        if (i_ >= s_.size() || s_[i_] != '>') return fail("expected '>'");
        ++i_;
        return true;
      } else {
        el.kids.emplace_back();
        if (!parse_element(el.kids.back(), depth + 1)) return false;
      }
    }
  }
};

// ------------------------------------------------------------------------------------ middleware
constexpr double kPiGdml = 3.14159265358979323846;

struct V3 {
  double x = 0, y = 0, z = 0;
};

class Reader {
public:
  Reader(double mm, bool specialized) : mm_(mm), specialized_(specialized) {}
  Model model;
  std::string err;

  bool node(Element const &e)
  {
    const std::string &n = e.name;
    if (is_solid(n)) return solid(e);
    if (n == "constant") return constant(e);
    if (n == "position") return position(e);
    if (n == "rotation") return rotation(e);
    if (n == "scale") return scale(e);
    if (n == "isotope" || n == "element" || n == "material" || n == "userinfo") return true; // no materials here
    if (n == "volume") return volume(e);
    if (n == "world") return world(e);
    for (auto const &k : e.kids)
      if (!node(k)) return false;
    return true;
  }

private:
  double mm_;
  bool specialized_;
  std::map<std::string, double> constants_;
  std::map<std::string, V3> positions_, rotations_, scales_;
  std::map<std::string, int> solids_;  // solid name -> shape index
  std::map<std::string, int> volumes_; // logical volume name -> index (first of that name)

  bool fail(std::string const &what)
  {
    if (err.empty()) err = what;
    return false;
  }

  static bool is_solid(std::string const &n)
  {
    static const char *names[] = {"multiUnion", "reflectedSolid", "scaledSolid", "union", "subtraction",
                                  "intersection", "box", "twistedbox", "twistedtrap", "twistedtrd", "paraboloid",
                                  "sphere", "ellipsoid", "tube", "twistedtubs", "cutTube", "cone", "elcone",
                                  "polycone", "genericPolycone", "para", "trd", "trap", "torus", "orb", "polyhedra",
                                  "genericPolyhedra", "xtru", "hype", "eltube", "tet", "arb8", "tessellated"};
    for (auto *s : names)
      if (n == s) return true;
    return false;
  }

  // Middleware::GetDoubleAttribute: the leading number of the text, else a constant (0 when unknown)
  double num(Element const &e, const char *key)
  {
    const std::string t = e.attr(key);
    const char *b       = t.c_str();
    char *end           = nullptr;
    const double v      = std::strtod(b, &end);
    if (end != b) return v;
    return constants_[t];
  }
  // Helper::GetAttribute<int>: 0 when absent
  bool integer(Element const &e, const char *key, int &out)
  {
    const std::string t = e.attr(key);
    if (t.empty()) return out = 0, true;
    const char *b = t.c_str();
    char *end     = nullptr;
    const long v  = std::strtol(b, &end, 10);
    if (end == b) return fail("attribute " + std::string(key) + "=\"" + t + "\" of <" + e.name + "> is not an integer");
    out = int(v);
    return true;
  }
  double length_unit(Element const &e) const
  {
    const std::string u = e.attr(e.name == "position" ? "unit" : (e.name == "auxiliary" ? "auxunit" : "lunit"));
    if (u == "m") return 1e3 * mm_;
    if (u == "km") return 1e6 * mm_;
    if (u == "um") return 1e-3 * mm_;
    if (u == "nm") return 1e-6 * mm_;
    if (u == "cm") return 10 * mm_;
    return mm_;
  }
  static double angle_unit(Element const &e)
  {
    return e.attr(e.name == "rotation" ? "unit" : "aunit") == "deg" ? kPiGdml / 180. : 1.;
  }

  bool constant(Element const &e)
  {
    const std::string name = e.attr("name");
    const std::string t    = e.attr("value");
    char *end              = nullptr;
    double v               = std::strtod(t.c_str(), &end);
    if (end == t.c_str()) v = std::nan(""); // GetAttribute<double> of a non-number
    if (!constants_.emplace(name, v).second) return fail("constant " + name + " defined twice");
    return true;
  }
  bool position(Element const &e)
  {
    const double l = length_unit(e);
    const V3 v{l * num(e, "x"), l * num(e, "y"), l * num(e, "z")};
    if (!positions_.emplace(e.attr("name"), v).second) return fail("position \"" + e.attr("name") + "\" defined twice");
    return true;
  }
  bool rotation(Element const &e)
  {
    const double a = angle_unit(e);
    const V3 v{a * num(e, "x"), a * num(e, "y"), a * num(e, "z")};
    if (!rotations_.emplace(e.attr("name"), v).second) return fail("rotation \"" + e.attr("name") + "\" defined twice");
    return true;
  }
  bool scale(Element const &e)
  {
    const V3 v{num(e, "x"), num(e, "y"), num(e, "z")};
    if (!scales_.emplace(e.attr("name"), v).second) return fail("scale \"" + e.attr("name") + "\" defined twice");
    return true;
  }

  int add_shape(int type, std::vector<double> const &p)
  {
    model.shape_type.push_back(type);
    model.shape_params.insert(model.shape_params.end(), p.begin(), p.end());
    model.shape_offset.push_back(int(model.shape_params.size()));
    return int(model.shape_type.size()) - 1;
  }

  // GeoManager::MakeInstance: the solid that the factory returns for these parameters. A build with
  // VECGEOM_NO_SPECIALIZATION=ON (the reference's default, CMakeLists.txt:155) returns every solid as
  // written (each Maker's #else branch); the substitutions below are those of a build with it OFF,
  // applied when the caller asks for them (specialized_).
  //  cone -> tube when the radii allow it (source/UnplacedCone.cpp:449-540; a full phi range is set
  //          to exactly 2 pi); hype -> tube or the solid hype (source/UnplacedHype.cpp:40-60);
  //  polycone of two planes -> cone of half-height |z1 - z0| / 2 centred at z = 0, as the factory
  //          builds it (source/UnplacedPolycone.cpp:50-60); sphere -> orb (source/UnplacedSphere.cpp:
  //          285-296); trapezoid -> box, trd or parallelepiped (source/UnplacedTrapezoid.cpp:
  //          100-125); trd -> box or the Trd1 kind (source/UnplacedTrd.cpp:48-85).
  int make_instance(int type, std::vector<double> p)
  {
    if (!specialized_) return add_shape(type, p);
    const double twoPi = 2 * kPiGdml;
    switch (type) {
    case kCone: { // rmin1 rmax1 rmin2 rmax2 dz sphi dphi
      const double rmin1 = p[0], rmax1 = p[1], rmin2 = p[2], rmax2 = p[3], dz = p[4], sphi = p[5], dphi = p[6];
      if (rmin1 <= 0 && rmin2 <= 0) {
        const double d = dphi >= twoPi ? twoPi : dphi;
        if (rmax1 == rmax2) return add_shape(kTube, {rmin1, rmax1, dz, sphi, d});
        return add_shape(kCone, {rmin1, rmax1, rmin2, rmax2, dz, sphi, d});
      }
      if (rmin1 > 0 || rmin2 > 0) {
        if (rmin1 == rmin2 && rmax1 == rmax2) return add_shape(kTube, {rmin1, rmax1, dz, sphi, dphi});
        return add_shape(kCone, p);
      }
      return add_shape(kCone, p); // NaN radii: the factory returns none; the builder rejects them
    }
    case kHype: { // rmin rmax stIn stOut dz
      const double rMin = p[0], rMax = p[1], stIn = p[2], stOut = p[3], dz = p[4];
      if (rMin <= 0) {
        if (stIn <= 0. && stOut > 0.) return add_shape(kHype, {0., rMax, 0., stOut, dz});
        if (stIn <= 0. && stOut <= 0.) return add_shape(kTube, {0., rMax, dz, 0., twoPi});
      } else {
        if (stIn <= 0. && stOut <= 0.) return add_shape(kTube, {rMin, rMax, dz, 0., twoPi});
      }
      return add_shape(kHype, p);
    }
    case kPolycone: { // sphi dphi nz z[] rmin[] rmax[]
      const int nz = int(p[2]);
      if (nz == 2) {
        const double dz = std::fabs(p[4] - p[3]) / 2.;
        return add_shape(kCone, {p[5], p[7], p[6], p[8], dz, p[0], p[1]});
      }
      return add_shape(kPolycone, p);
    }
    case kSphere: // rmin rmax sphi dphi stheta dtheta
      if (p[0] == 0. && p[3] >= 2 * kPiGdml && p[5] == kPiGdml) return add_shape(kOrb, {p[1]});
      return add_shape(kSphere, p);
    case kTrapezoid: { // dz theta phi dy1 dx1 dx2 alpha1 dy2 dx3 dx4 alpha2
      const double dz = p[0], th = p[1], ph = p[2], dy1 = p[3], dx1 = p[4], dx2 = p[5], a1 = p[6], dy2 = p[7];
      const double dx3 = p[8], dx4 = p[9], a2 = p[10];
      const bool square = th == 0. && ph == 0. && a1 == 0. && a2 == 0;
      if (square && dx1 == dx2 && dx2 == dx3 && dx3 == dx4 && dy1 == dy2) return add_shape(kBox, {dx1, dy1, dz});
      if (square && dx1 == dx2 && dx3 == dx4 && dx2 != dx3 && dy1 == dy2)
        return add_shape(kTrd, {dx1, dx3, dy1, dy1, dz, 1.});
      if (square && dx1 == dx2 && dx3 == dx4 && dx2 != dx3 && dy1 != dy2) return add_shape(kTrd, {dx1, dx3, dy1, dy2, dz});
      if (a1 == a2 && dx1 == dx2 && dx2 == dx3 && dx3 == dx4 && dy1 == dy2)
        return add_shape(kParallelepiped, {dx1, dy1, dz, a1, th, ph});
      return add_shape(kTrapezoid, p);
    }
    case kTrd: { // x1 x2 y1 y2 z
      if (p[0] == p[1] && p[2] == p[3]) return add_shape(kBox, {p[0], p[2], p[4]});
      if (p[2] == p[3]) return add_shape(kTrd, {p[0], p[1], p[2], p[3], p[4], 1.});
      return add_shape(kTrd, p);
    }
    default:
      return add_shape(type, p);
    }
  }

  // the z planes of a polycone or polyhedra: every child element (Middleware.cpp:836-906)
  void planes(Element const &e, double l, std::vector<double> &z, std::vector<double> &rmin, std::vector<double> &rmax)
  {
    for (auto const &k : e.kids) {
      rmax.push_back(l * num(k, "rmax"));
      rmin.push_back(l * num(k, "rmin"));
      z.push_back(l * num(k, "z"));
    }
  }

  bool tessellated(Element const &e, std::vector<double> &p)
  {
    std::vector<double> tri;
    int ntri    = 0;
    auto vertex = [&](Element const &f, int k) { return positions_[f.attr(("vertex" + std::to_string(k)).c_str())]; };
    auto push   = [&](V3 a, V3 b, V3 c) {
      for (V3 const &v : {a, b, c}) tri.insert(tri.end(), {v.x, v.y, v.z});
      ++ntri;
    };
    auto add = [](V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; };
    for (auto const &f : e.kids) {
      // UnplacedTessellated::AddTriangularFacet / AddQuadrilateralFacet (VecGeom/volumes/TessellatedStruct.h:
      // 596-660): RELATIVE vertices are offsets, accumulated from the first one; a quadrilateral is
      // the two triangles (0 1 2) and (0 2 3)
      const bool absolute = f.attr("type") == "ABSOLUTE";
      const V3 v0 = vertex(f, 1), v1 = vertex(f, 2), v2 = vertex(f, 3);
      if (f.name == "triangular") {
        if (absolute)
          push(v0, v1, v2);
        else
          push(v0, add(v0, v1), add(add(v0, v1), v2));
      } else {
        const V3 v3 = vertex(f, 4);
        if (absolute) {
          push(v0, v1, v2);
          push(v0, v2, v3);
        } else {
          push(v0, add(v0, v1), add(add(v0, v1), v2));
          push(v0, add(v0, v2), add(add(v0, v2), v3));
        }
      }
    }
    p.assign(1, double(ntri));
    p.insert(p.end(), tri.begin(), tri.end());
    return true;
  }

  bool solid(Element const &e)
  {
    const std::string &n = e.name;
    const std::string name = e.attr("name");
    const double l = length_unit(e), a = angle_unit(e);
    std::vector<double> p;
    int type = -1;
    if (n == "box") {
      type = kBox, p = {l * num(e, "x") / 2., l * num(e, "y") / 2., l * num(e, "z") / 2.};
    } else if (n == "orb") {
      type = kOrb, p = {l * num(e, "r")};
    } else if (n == "tube") {
      type = kTube;
      p    = {l * num(e, "rmin"), l * num(e, "rmax"), l * num(e, "z") / 2., a * num(e, "startphi"),
           a * num(e, "deltaphi")};
    } else if (n == "cone") {
      type = kCone;
      p    = {l * num(e, "rmin1"), l * num(e, "rmax1"), l * num(e, "rmin2"), l * num(e, "rmax2"),
           l * num(e, "z") / 2., a * num(e, "startphi"), a * num(e, "deltaphi")};
    } else if (n == "polycone" || n == "polyhedra") {
      std::vector<double> z, rmin, rmax;
      planes(e, l, z, rmin, rmax);
      p = {a * num(e, "startphi"), a * num(e, "deltaphi")};
      if (n == "polyhedra") {
        int sides = 0;
        if (!integer(e, "numsides", sides)) return false;
        type = kPolyhedron, p.push_back(double(sides));
      } else {
        type = kPolycone;
      }
      p.push_back(double(z.size()));
      p.insert(p.end(), z.begin(), z.end());
      p.insert(p.end(), rmin.begin(), rmin.end());
      p.insert(p.end(), rmax.begin(), rmax.end());
    } else if (n == "torus") {
      type = kTorus2;
      p    = {l * num(e, "rmin"), l * num(e, "rmax"), l * num(e, "rtor"), a * num(e, "startphi"),
           a * num(e, "deltaphi")};
    } else if (n == "sphere") {
      type = kSphere;
      p    = {l * num(e, "rmin"),      l * num(e, "rmax"),        a * num(e, "startphi"),
           a * num(e, "deltaphi"), a * num(e, "starttheta"), a * num(e, "deltatheta")};
    } else if (n == "trd") {
      type = kTrd;
      p    = {l * num(e, "x1") / 2., l * num(e, "x2") / 2., l * num(e, "y1") / 2., l * num(e, "y2") / 2.,
           l * num(e, "z") / 2.};
    } else if (n == "trap") {
      type = kTrapezoid;
      p    = {l * num(e, "z") / 2.,  a * num(e, "theta"),    a * num(e, "phi"),     l * num(e, "y1") / 2.,
           l * num(e, "x1") / 2., l * num(e, "x2") / 2.,  a * num(e, "alpha1"),  l * num(e, "y2") / 2.,
           l * num(e, "x3") / 2., l * num(e, "x4") / 2.,  a * num(e, "alpha2")};
    } else if (n == "paraboloid") {
      type = kParaboloid, p = {l * num(e, "rlo"), l * num(e, "rhi"), l * num(e, "dz")};
    } else if (n == "hype") {
      type = kHype;
      p    = {l * num(e, "rmin"), l * num(e, "rmax"), num(e, "inst"), num(e, "outst"), l * num(e, "z") / 2.};
    } else if (n == "tessellated") {
      type = kTessellated;
      if (!tessellated(e, p)) return false;
    } else if (n == "elcone") { // processElCone (Middleware.cpp:818-834): every value in lengths
      type = kEllipticalCone, p = {l * num(e, "dx"), l * num(e, "dy"), l * num(e, "zmax"), l * num(e, "zcut")};
    } else if (n == "cutTube") { // processCutTube (Middleware.cpp:763-794): z halved, the normals in lengths too
      type = kCutTube;
      p    = {l * num(e, "rmin"), l * num(e, "rmax"), l * num(e, "z") / 2., a * num(e, "startphi"),
           a * num(e, "deltaphi"), l * num(e, "lowX"), l * num(e, "lowY"), l * num(e, "lowZ"),
           l * num(e, "highX"), l * num(e, "highY"), l * num(e, "highZ")};
    } else if (n == "eltube") { // processElTube (Middleware.cpp:747-761): semi-axes, not halved
      type = kEllipticalTube, p = {l * num(e, "dx"), l * num(e, "dy"), l * num(e, "dz")};
    } else if (n == "ellipsoid") { // processEllipsoid (Middleware.cpp:948-965)
      type = kEllipsoid;
      p    = {l * num(e, "ax"), l * num(e, "by"), l * num(e, "cz"), l * num(e, "zcut1"), l * num(e, "zcut2")};
    } else if (n == "arb8") { // processGenTrap (Middleware.cpp:1047-1096): the vertices without the length unit
      type = kGenTrap;
      for (const char *c : {"v1x", "v2x", "v3x", "v4x", "v5x", "v6x", "v7x", "v8x"}) p.push_back(num(e, c));
      for (const char *c : {"v1y", "v2y", "v3y", "v4y", "v5y", "v6y", "v7y", "v8y"}) p.push_back(num(e, c));
      p.push_back(l * num(e, "dz"));
    } else if (n == "para") { // processParallelepiped (Middleware.cpp:967-988): x y z halved
      type = kParallelepiped;
      p    = {l * num(e, "x") / 2., l * num(e, "y") / 2., l * num(e, "z") / 2., a * num(e, "alpha"),
           a * num(e, "theta"), a * num(e, "phi")};
    } else if (n == "tet") { // processTet (Middleware.cpp:1155-1172): the corners are named positions
      type = kTet;
      for (const char *v : {"vertex1", "vertex2", "vertex3", "vertex4"}) {
        V3 const q = positions_[e.attr(v)]; // an unknown name reads (0, 0, 0), as the middleware's map does
        p.insert(p.end(), {q.x, q.y, q.z});
      }
    } else {
      return fail("solid <" + n + "> (\"" + name + "\") is not supported");
    }
    const int idx = make_instance(type, p);
    if (!solids_.emplace(name, idx).second) return fail("solid \"" + name + "\" defined twice");
    return true;
  }

  // makeRotationMatrixFromCartesianAngles (Middleware.cpp:114-132): the rotation by the three
  // angles with each sine negated, as rows xx xy xz / yx yy yz / zx zy zz.
  static void rotation_matrix(V3 const &r, double m[9])
  {
    const double s1 = -std::sin(r.x), c1 = std::cos(r.x);
    const double s2 = -std::sin(r.y), c2 = std::cos(r.y);
    const double s3 = -std::sin(r.z), c3 = std::cos(r.z);
    m[0] = c2 * c3, m[1] = -c2 * s3, m[2] = s2;
    m[3] = c1 * s3 + c3 * s1 * s2, m[4] = c1 * c3 - s1 * s2 * s3, m[5] = -c2 * s1;
    m[6] = s1 * s3 - c1 * c3 * s2, m[7] = c3 * s1 + c1 * s2 * s3, m[8] = c1 * c2;
  }

  bool physvol(Element const &e, int mother)
  {
    int lv = -1;
    V3 pos, rot, scl{1, 1, 1};
    for (auto const &k : e.kids) {
      const std::string &n = k.name;
      if (n == "volumeref") {
        auto it = volumes_.find(k.attr("ref"));
        if (it == volumes_.end()) return fail("physvol refers to unknown volume \"" + k.attr("ref") + "\"");
        lv = it->second;
      } else if (n == "position") {
        if (!position(k)) return false;
        pos = positions_[k.attr("name")];
      } else if (n == "rotation") {
        if (!rotation(k)) return false;
        rot = rotations_[k.attr("name")];
      } else if (n == "scale") {
        if (!scale(k)) return false;
        scl = scales_[k.attr("name")];
      } else if (n == "positionref") {
        pos = positions_[k.attr("ref")];
      } else if (n == "rotationref") {
        rot = rotations_[k.attr("ref")];
      } else if (n == "scaleref") {
        scl = scales_[k.attr("ref")];
      } else {
        return fail("unexpected <" + n + "> in <physvol>");
      }
    }
    if (lv < 0) return fail("physvol without a volumeref");
    // ReflFactory::Place (persistency/gdml/source/src/ReflFactory.cpp:20-50): a scale without a
    // reflection is ignored; reflections need reflected solids, which are not supported here
    if (scl.x * scl.y * scl.z < 0) return fail("reflected placement (negative scale) is not supported");
    double m[9];
    rotation_matrix(rot, m);
    model.placed_mother.push_back(mother);
    model.placed_logical.push_back(lv);
    model.placed_transform.insert(model.placed_transform.end(), {pos.x, pos.y, pos.z});
    model.placed_transform.insert(model.placed_transform.end(), m, m + 9);
    model.placed_names.push_back(e.attr("name"));
    return true;
  }

  // Middleware::processLogicVolume (:1283-1364): the volume exists from its <solidref> on
  bool volume(Element const &e)
  {
    int lv                 = -1;
    const std::string name = e.attr("name");
    for (auto const &k : e.kids) {
      if (k.name == "solidref") {
        auto it = solids_.find(k.attr("ref"));
        if (it == solids_.end()) return fail("volume \"" + name + "\" refers to unknown solid \"" + k.attr("ref") + "\"");
        lv = int(model.logical_shape.size());
        model.logical_shape.push_back(it->second);
        model.logical_names.push_back(name);
        volumes_.emplace(name, lv);
      } else if (k.name == "physvol") {
        if (lv < 0) return fail("physvol before the solidref of volume \"" + name + "\"");
        if (!physvol(k, lv)) return false;
      }
    }
    if (lv < 0) return fail("volume \"" + name + "\" has no solidref");
    return true;
  }

  bool world(Element const &e)
  {
    auto it = volumes_.find(e.attr("ref"));
    if (it == volumes_.end()) return fail("world refers to unknown volume \"" + e.attr("ref") + "\"");
    model.world = it->second;
    return true;
  }
};

bool Read(std::string const &path, double mm, bool specialized, Model &out, std::string &err)
{
  std::ifstream f(path, std::ios::binary);
  if (!f) return err = "cannot open " + path, false;
  std::ostringstream ss;
  ss << f.rdbuf();
  const std::string text = ss.str();
  Element root;
  XmlParser parser(text);
  if (!parser.parse(root, err)) return false;
  if (root.name != "gdml") return err = "root element is <" + root.name + ">, not <gdml>", false;
  Reader r(mm, specialized);
  r.model.shape_offset.push_back(0);
  if (!r.node(root)) return err = r.err, false;
  if (r.model.world < 0) return err = "no <world> in the <setup>", false;
  out = std::move(r.model);
  return true;
}

} // namespace gdml
} // namespace vgmi

using namespace vgmi;

struct vgmi_gdml {
  gdml::Model m;
};

extern "C" {

vgmi_status vgmi_gdml_read(const char *path, double mm_unit, int flags, vgmi_gdml **out)
{
  if (!out) return set_error(VGMI_ERR_INVALID, "vgmi_gdml_read: null output handle");
  *out = nullptr;
  if (!path) return set_error(VGMI_ERR_INVALID, "vgmi_gdml_read: null path");
  if (!(mm_unit > 0) || !std::isfinite(mm_unit))
    return set_error(VGMI_ERR_INVALID, "vgmi_gdml_read: the millimetre unit must be positive");
  if (flags & ~VGMI_GDML_SPECIALIZED) return set_error(VGMI_ERR_INVALID, "vgmi_gdml_read: unknown flags %d", flags);
  auto *g = new (std::nothrow) vgmi_gdml();
  if (!g) return set_error(VGMI_ERR_INVALID, "vgmi_gdml_read: out of host memory");
  std::string err;
  if (!gdml::Read(path, mm_unit, (flags & VGMI_GDML_SPECIALIZED) != 0, g->m, err)) {
    delete g;
    return set_error(VGMI_ERR_INVALID, "vgmi_gdml_read: %s", err.c_str());
  }
  *out = g;
  return VGMI_OK;
}

vgmi_status vgmi_gdml_desc(const vgmi_gdml *g, vgmi_geom_desc *d)
{
  if (!g || !d) return set_error(VGMI_ERR_INVALID, "vgmi_gdml_desc: null argument");
  gdml::Model const &m = g->m;
  d->nshapes            = int(m.shape_type.size());
  d->shape_type         = m.shape_type.data();
  d->shape_param_offset = m.shape_offset.data();
  d->shape_params       = m.shape_params.data();
  d->nlogical           = int(m.logical_shape.size());
  d->logical_shape      = m.logical_shape.data();
  d->nplaced            = int(m.placed_mother.size());
  d->placed_mother      = m.placed_mother.data();
  d->placed_logical     = m.placed_logical.data();
  d->placed_transform   = m.placed_transform.data();
  d->world_logical      = m.world;
  return VGMI_OK;
}

vgmi_status vgmi_gdml_name(const vgmi_gdml *g, int kind, int index, const char **name)
{
  if (!g || !name) return set_error(VGMI_ERR_INVALID, "vgmi_gdml_name: null argument");
  std::vector<std::string> const &v = kind == 0 ? g->m.logical_names : g->m.placed_names;
  if (kind != 0 && kind != 1) return set_error(VGMI_ERR_INVALID, "vgmi_gdml_name: kind is 0 (logical) or 1 (placed)");
  if (index < 0 || index >= int(v.size()))
    return set_error(VGMI_ERR_INVALID, "vgmi_gdml_name: index %d not in [0, %d)", index, int(v.size()));
  *name = v[size_t(index)].c_str();
  return VGMI_OK;
}

vgmi_status vgmi_gdml_destroy(vgmi_gdml *g)
{
  delete g;
  return VGMI_OK;
}

} // extern "C"
