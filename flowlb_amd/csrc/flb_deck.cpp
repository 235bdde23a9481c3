// flb_deck.cpp -- host reader of FlowLB case decks and the flag set-up that
// follows it.  The decisions reproduce the reference's reader and shapes so
// that the same deck gives the same flags bit for bit:
//   line handling  Reader::splitString / extractString (src/io/reader.cpp:23-120)
//   Geometry.txt   GeometryReader (reader.cpp:770-1217)
//   mesh size      Mesh::calculateNumberPointsMesh (src/geometry/mesh.cpp:45-52)
//   shapes         src/geometry/shapes.cpp (rectangle, circle, CSV outline, CDWs)
//   OptionsCalc    CalculationReader (reader.cpp:138-768)
//   flags          initData (src/initData.h:220-333) + initFreeSurfaceFlags
// Only 2D decks are accepted (the 3D branches of the reference are unfinished).
#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/flb_deck.h"

namespace {

constexpr double kEps = 1e-12;  // Shape::m_Epsilon, CDW::m_Epsilon, initData epsilon
constexpr uint8_t kNone = 100;  // TypesNodes::NONE (src/utils.h:107)

struct DeckError {
  int status;
  std::string msg;
};

[[noreturn]] void fail(int status, const std::string& msg) { throw DeckError{status, msg}; }
[[noreturn]] void fail_line(const char* file, unsigned line, const std::string& what) {
  fail(FLB_ERR_ARGUMENT, std::string(file) + " line " + std::to_string(line) + ": " + what);
}

bool is_space(char c) {
  return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r';
}
std::string strip_all(const std::string& s) {  // std::regex_replace(s, "\\s+", "")
  std::string o;
  for (char c : s)
    if (!is_space(c)) o.push_back(c);
  return o;
}

// Reader::splitString: text after '#' dropped, at most two '='-separated fields
// (std::getline semantics: a trailing '=' adds no field), whitespace removed.
bool split_option(std::string line, std::string out[2]) {
  const size_t hash = line.find('#');
  if (hash != std::string::npos) line = line.substr(0, hash);
  std::stringstream ss(line);
  std::string part;
  int i = 0;
  while (std::getline(ss, part, '=')) {
    if (i > 1) return false;
    out[i++] = strip_all(part);
  }
  return true;
}

// extractString(option, delimiter, toEnd): the text before the delimiter (all
// of it when absent), or after "<delimiter>." when toEnd.
std::string before(const std::string& s, char d) {
  const size_t p = s.find(d);
  return p == std::string::npos ? s : s.substr(0, p);
}
std::string attribute(const std::string& s) {
  const size_t p = s.find(')');
  if (p == std::string::npos || p + 2 > s.size()) return std::string();
  return s.substr(p + 2);
}
std::string between(const std::string& s, char a, char b) {
  const size_t i = s.find(a), j = s.find(b);
  if (i == std::string::npos || j == std::string::npos || j < i) return std::string();
  return s.substr(i + 1, j - i - 1);
}

double to_d(const std::string& s) { return std::stod(s); }  // throws like the reference
int to_i(const std::string& s) { return std::stoi(s); }

// "(x,y,z)" after whitespace removal -> three numbers, as the reference's
// three extractString calls read them (x up to the first comma, y between the
// commas, z from the second comma on, parsed as a number prefix).
void triple(const std::string& v, double* out) {
  const size_t c1 = v.find(','), c2 = c1 == std::string::npos ? c1 : v.find(',', c1 + 1);
  const size_t p = v.find('(');
  if (p == std::string::npos || c1 == std::string::npos || c2 == std::string::npos)
    throw std::invalid_argument("triple");
  out[0] = to_d(v.substr(p + 1, c1 - p - 1));
  out[1] = to_d(v.substr(c1 + 1, c2 - c1 - 1));
  out[2] = to_d(v.substr(c2 + 1));
}

struct Point3 {
  double x, y, z;
};

// ---------------------------------------------------------------- shapes
enum class Kind { Rectangle, Circle, Outline };

struct Shape {
  Kind kind;
  double x0 = 0, y0 = 0, z0 = 0;
  double rx = 0, ry = 0, rz = 0;  // rotations (rad)
  // rectangle
  double height = 0, xwidth = 0, zwidth = 0;
  double xmin = 0, xmax = 0, ymin = 0, ymax = 0, zmin = 0, zmax = 0;
  double rot_x[2] = {0, 0}, rot_y[2] = {0, 0};
  bool rotate = false;
  // circle
  double radius = 0, thickness = 0;
  // CSV outline
  std::string file;
  std::vector<double> px, py;

  explicit Shape(Kind k) : kind(k) {}

  // RectangleShape::initShape (shapes.cpp:64-136), 2D part.
  void init_rectangle() {
    rotate = !(std::abs(rz - 0.0) <= 1e-6);
    rot_x[0] = std::cos(-rz);
    rot_x[1] = std::sin(-rz);
    rot_y[0] = std::sin(-rz);
    rot_y[1] = std::cos(-rz);
    xmax = x0 + xwidth / 2.0 + kEps;
    xmin = x0 - xwidth / 2.0 - kEps;
    ymax = y0 + height / 2.0 + kEps;
    ymin = y0 - height / 2.0 - kEps;
    zmax = z0 + zwidth / 2.0 + kEps;
    zmin = z0 - zwidth / 2.0 - kEps;
  }

  // Imported2DShape::initShape + calculateAreaCentroid + moveRotateCoordinates
  // (shapes.cpp:185-236, 298-345).
  void init_outline(const std::string& dir) {
    const std::string path = dir + "/" + file;
    std::ifstream in(path);
    if (!in.is_open()) fail(FLB_ERR_ARGUMENT, "cannot open " + path);
    std::vector<std::string> lines;
    std::string line;
    size_t non_empty = 0;
    while (std::getline(in, line)) {
      lines.push_back(line);
      if (!line.empty()) ++non_empty;
    }
    if (non_empty < 3) fail(FLB_ERR_ARGUMENT, path + ": an outline needs at least 3 lines");
    // CSVReader's row iterator: rows split at ';' after removing spaces; an
    // empty line repeats the previous row (a zero-length edge, no effect).
    size_t rows = 0;
    for (size_t k = 0; k < lines.size() && rows < non_empty; ++k) {
      std::string l = lines[k];
      l.erase(std::remove(l.begin(), l.end(), ' '), l.end());
      if (l.empty()) {
        if (!px.empty()) {
          px.push_back(px.back());
          py.push_back(py.back());
        }
        continue;
      }
      const size_t sc = l.find(';');
      if (sc == std::string::npos) fail(FLB_ERR_ARGUMENT, path + ": a row needs x;y");
      const std::string xs = l.substr(0, sc);
      std::string ys = l.substr(sc + 1);
      const size_t sc2 = ys.find(';');
      if (sc2 != std::string::npos) ys = ys.substr(0, sc2);
      double x = 0, y = 0;
      auto rx_ = std::from_chars(xs.data(), xs.data() + xs.size(), x, std::chars_format::general);
      auto ry_ = std::from_chars(ys.data(), ys.data() + ys.size(), y, std::chars_format::general);
      if (rx_.ec != std::errc() || ry_.ec != std::errc())
        fail(FLB_ERR_ARGUMENT, path + " row " + std::to_string(rows + 1) + ": not a number");
      px.push_back(x);
      py.push_back(y);
      ++rows;
    }
    // Shoelace area and centroid; the centroid is taken in absolute value as
    // the reference does.
    const size_t n = px.size();
    double area = 0, cx = 0, cy = 0;
    for (size_t i = 0; i + 1 < n; ++i) {
      const double pr = px[i] * py[i + 1] - px[i + 1] * py[i];
      area += pr;
      cx += (px[i] + px[i + 1]) * pr;
      cy += (py[i] + py[i + 1]) * pr;
    }
    const double pr = px[n - 1] * py[0] - px[0] * py[n - 1];
    area += pr;
    cx += (px[n - 1] + px[0]) * pr;
    cy += (py[n - 1] + py[0]) * pr;
    area *= 0.5;
    area = area < 0 ? -area : area;
    cx /= 6.0 * area;
    cx = cx < 0 ? -cx : cx;
    cy /= 6.0 * area;
    cy = cy < 0 ? -cy : cy;
    const bool rot = !(std::abs(rz - 0.0) <= 1e-6);
    const double c = std::cos(rz), s = std::sin(rz);
    xmin = xmax = ymin = ymax = 0.0;  // the reference's box starts at the origin
    for (size_t i = 0; i < n; ++i) {
      double x = px[i] - cx, y = py[i] - cy;
      if (rot) {
        const double xp = x;
        x = x * c - y * s;
        y = xp * s + y * c;
      }
      x += x0;
      y += y0;
      px[i] = x;
      py[i] = y;
      if (x > xmax) xmax = x;
      else if (x < xmin) xmin = x;
      if (y > ymax) ymax = y;
      else if (y < ymin) ymin = y;
    }
    xmax += kEps;
    xmin -= kEps;
    ymax += kEps;
    ymin -= kEps;
  }

  bool inside(double xs, double ys, double zs) const {
    switch (kind) {
      case Kind::Circle: {  // CircleShape::isNodeInside, 2D (shapes.cpp:27-45)
        const double d = (ys - y0) * (ys - y0) + (xs - x0) * (xs - x0);
        return !(d > radius * radius + kEps);
      }
      case Kind::Rectangle: {  // RectangleShape::isNodeInside (shapes.cpp:138-160)
        if (rotate) {
          const double ax = xs - x0, ay = ys - y0;
          xs = ax * rot_x[0] - ay * rot_x[1] + x0;
          ys = ax * rot_y[0] + ay * rot_y[1] + y0;
        }
        return !(xs > xmax || xs < xmin || ys > ymax || ys < ymin || zs > zmax || zs < zmin);
      }
      case Kind::Outline: {  // Imported2DShape::isNodeInside: even-odd crossing test
        if (xs > xmax || xs < xmin || ys > ymax || ys < ymin) return false;
        bool in = false;
        const size_t n = px.size();
        for (size_t i = 0, j = n - 1; i < n; j = i++) {
          if (((py[i] > ys) != (py[j] > ys)) &&
              (xs < (px[j] - px[i]) * (ys - py[i]) / (py[j] - py[i]) + px[i]))
            in = !in;
        }
        return in;
      }
    }
    return false;
  }
};

// Cross drainage work: a circular barrel (its centre line runs radius above the
// base polyline) or a rectangle that, in 2D, only holds the base line itself
// (CircularCDW / RectangularCDW::isNodeInside, shapes.cpp:395-450).
struct Cdw {
  bool circle = true;
  double radius = 0, height = 0, width = 0;
};

}  // namespace

struct flb_deck {
  std::string dir;
  bool is3d = false;
  std::vector<Point3> points;
  double size = 0.0;
  int idx_init = 0, idx_end = 0;
  std::vector<Cdw> cdws;
  std::vector<Shape> obstacles, initial;
  double x_min = 0, x_max = 0, y_min = 0, y_max = 0, z_min = 0, z_max = 0;
  uint32_t nx = 0, ny = 0;
  flb_deck_info opt{};
};

namespace {

// ------------------------------------------------------------- Geometry.txt
void geometry_line(flb_deck& d, const std::string& key, const std::string& value, unsigned ln) {
  const char* F = "Geometry.txt";
  const std::string type = before(key, '(');
  // getMeshElements (reader.cpp:785-831)
  if (type == "POINT") {
    const int num = to_i(between(key, '(', ')'));
    if (int(d.points.size()) >= num) fail_line(F, ln, "point numbering is not correct");
    double c[3];
    triple(value, c);
    if (c[0] > d.x_max) d.x_max = c[0];
    if (c[0] < d.x_min) d.x_min = c[0];
    if (c[1] > d.y_max) d.y_max = c[1];
    if (c[1] < d.y_min) d.y_min = c[1];
    if (c[2] > d.z_max) d.z_max = c[2];
    if (c[2] < d.z_min) d.z_min = c[2];
    d.points.push_back(Point3{c[0], c[1], c[2]});
  }
  if (type == "SIZE") d.size = to_d(value);
  if (type == "TYPE_DOMAIN") {
    if (value == "2D") d.is3d = false;
    else if (value == "3D") d.is3d = true;
  }
  // getCDWData (reader.cpp:833-955)
  if (type == "INITIAL_POSITION_SHAPES") {
    d.idx_init = to_i(value) - 1;
    if (d.idx_init >= int(d.points.size()))
      fail_line(F, ln, "the initial point does not exist or is the last point");
    if (d.idx_init < 0) fail_line(F, ln, "point numbers start at 1");
  } else if (type == "END_POSITION_SHAPES") {
    d.idx_end = to_i(value) - 1;
    // The reference accepts idx_end == number of points and then reads past
    // the point list; that case is an error here.
    if (d.idx_end >= int(d.points.size())) fail_line(F, ln, "the end point does not exist");
    if (d.idx_end < 0) fail_line(F, ln, "point numbers start at 1");
  } else if (type == "CDW") {
    const int num = to_i(between(key, '(', ')'));
    const bool exists = num <= int(d.cdws.size());
    const std::string a = attribute(key);
    if (a == "type") {
      if (exists) fail_line(F, ln, "the CDW number already exists");
      if (value == "CIRCLE") d.cdws.push_back(Cdw{true});
      else if (value == "RECTANGLE") d.cdws.push_back(Cdw{false});
      else fail_line(F, ln, "the shape of the CDW is not correct");
    } else if (a == "height" || a == "width" || a == "radius") {
      if (!exists || num < 1) fail_line(F, ln, "the type of the CDW must come first");
      Cdw& c = d.cdws[num - 1];
      if (a == "height" && !c.circle) c.height = to_d(value);
      if (a == "width" && !c.circle) c.width = to_d(value);
      if (a == "radius" && c.circle) c.radius = to_d(value);
    }
  }
  // getObstaclesData (reader.cpp:957-1120)
  if (type == "OBSTACLE") {
    const int num = to_i(between(key, '(', ')'));
    const bool exists = num <= int(d.obstacles.size());
    const std::string a = attribute(key);
    if (a == "type") {
      if (exists) fail_line(F, ln, "the obstacle number already exists");
      if (value == "WALL") d.obstacles.emplace_back(Kind::Rectangle);
      else if (value == "CIRCLE") d.obstacles.emplace_back(Kind::Circle);
      else if (value == "CSV_2D") d.obstacles.emplace_back(Kind::Outline);
      return;
    }
    if (a != "name" && a != "position" && a != "rotation" && a != "height" && a != "xWidth" &&
        a != "zWidth" && a != "radius" && a != "thickness")
      return;
    if (!exists || num < 1) fail_line(F, ln, "the type of the obstacle must come first");
    Shape& s = d.obstacles[num - 1];
    if (a == "name") {
      if (s.kind == Kind::Outline) s.file = value + ".csv";
    } else if (a == "position") {
      double c[3];
      triple(value, c);
      s.x0 = c[0];
      s.y0 = c[1];
      s.z0 = c[2];
    } else if (a == "rotation") {
      double c[3];
      triple(value, c);
      constexpr double pi = 3.141592653589793;
      // CircleShape ignores rotations; an outline keeps only the z rotation.
      if (s.kind == Kind::Rectangle) {
        s.rx = c[0] * pi / 180.0;
        s.ry = c[1] * pi / 180.0;
      }
      if (s.kind != Kind::Circle) s.rz = c[2] * pi / 180.0;
    } else if (a == "height") {
      if (s.kind == Kind::Rectangle) s.height = to_d(value);
    } else if (a == "xWidth") {
      if (s.kind == Kind::Rectangle) s.xwidth = to_d(value);
    } else if (a == "zWidth") {
      if (s.kind == Kind::Rectangle) s.zwidth = to_d(value);
    } else if (a == "radius") {
      if (s.kind == Kind::Circle) s.radius = to_d(value);
    } else if (a == "thickness") {
      if (s.kind == Kind::Circle) s.thickness = to_d(value);
    }
  }
}

// modifyGeometryData (reader.cpp:1122-1178), 2D part, and checkOptionsGeometry.
void finish_geometry(flb_deck& d) {
  if (d.is3d) fail(FLB_ERR_UNSUPPORTED, "3D decks are not supported");
  if (d.points.size() < 2) fail(FLB_ERR_ARGUMENT, "Geometry.txt: at least two points needed");
  if (d.idx_init >= d.idx_end)
    fail(FLB_ERR_ARGUMENT, "Geometry.txt: the initial point must come before the end point");
  const double h0 = d.points[d.idx_init].y, h1 = d.points[d.idx_end].y;
  for (const Cdw& c : d.cdws) {
    const double a = c.circle ? h0 + 2 * c.radius : h0 + c.height;
    const double b = c.circle ? h1 + 2 * c.radius : h1 + c.height;
    if (a > d.y_max) d.y_max = a;
    if (b > d.y_max) d.y_max = b;
  }
  for (Shape& s : d.obstacles) {
    if (s.kind == Kind::Rectangle) s.init_rectangle();
    else if (s.kind == Kind::Outline) s.init_outline(d.dir);
  }
  if (!(d.size > 0.0)) fail(FLB_ERR_ARGUMENT, "Geometry.txt: SIZE must be positive");
  // Mesh::calculateNumberPointsMesh
  const double fx = std::floor((d.x_max - d.x_min) / d.size) + 1;
  const double fy = std::floor((d.y_max - d.y_min) / d.size) + 1;
  if (!(fx >= 2 && fy >= 2 && fx * fy < 4.0e9))
    fail(FLB_ERR_ARGUMENT, "Geometry.txt: the mesh must have at least 2 x 2 nodes");
  d.nx = uint32_t(fx);
  d.ny = uint32_t(fy);
}

// ---------------------------------------------------------- OptionsCalc.txt
uint8_t node_type(const std::string& v, uint8_t cur) {  // getTypeNode
  if (v == "INPUT") return FLB_INLET;
  if (v == "OUTPUT") return FLB_OUTLET;
  if (v == "WALL") return FLB_WALL;
  if (v == "PERIODIC") return FLB_FLUID;
  return cur;
}

bool flag_value(const std::string& v, int32_t* out) {
  if (v == "FALSE") *out = 0;
  else if (v == "TRUE") *out = 1;
  else return false;
  return true;
}

// getOptionCalculation (reader.cpp:267-590).  Returns false for an invalid
// value or an unknown key (keys starting with INITIAL are read elsewhere).
bool option_line(flb_deck_info& o, const std::string& key, const std::string& v) {
  double c[3];
  if (key == "TYPE_ANALYSIS") {
    if (v != "2D" && v != "3D") return false;
    if (v == "3D") fail(FLB_ERR_UNSUPPORTED, "3D analyses are not supported");
  } else if (key == "CALCULATION_API") {
    if (v != "CUDA") return false;
  } else if (key == "CUDA_BLOCK_SIZE") {
    const size_t p = v.find('('), c1 = v.find(',');
    const size_t c2 = c1 == std::string::npos ? c1 : v.find(',', c1 + 1);
    if (p == std::string::npos || c2 == std::string::npos) return false;
    o.block_x = to_i(v.substr(p + 1, c1 - p - 1));
    o.block_y = to_i(v.substr(c1 + 1, c2 - c1 - 1));
    o.block_z = to_i(v.substr(c2 + 1));
  } else if (key == "GRAPHICS_API") {
    if (v == "NONE") o.graphics = 0;
    else if (v == "OPENGL") o.graphics = 1;
    else return false;
  } else if (key == "TYPE_PROBLEM") {
    if (v == "MONOFLUID") o.problem = FLB_MONOFLUID;
    else if (v == "FREE_SURFACE") o.problem = FLB_FREE_SURFACE;
    else return false;
  } else if (key == "USE_VOLUMETRIC_FORCE") {
    return flag_value(v, &o.use_force);
  } else if (key == "ACCELERATION") {
    triple(v, o.acceleration);
  } else if (key == "VOLUMETRIC_FORCE") {
    triple(v, o.volumetric_force);
  } else if (key == "COLLISION_OPERATOR") {
    if (v == "SRT") o.collision = FLB_SRT;
    else if (v == "TRT") o.collision = FLB_TRT;
    else return false;
  } else if (key == "SUBGRID_TURBULENCE_MODEL") {
    return flag_value(v, &o.smagorinsky);
  } else if (key == "SURFACE_TENSION") {
    return flag_value(v, &o.surface_tension);
  } else if (key == "PRECISION") {
    if (v == "16") o.precision = 16;
    else if (v == "32") o.precision = 32;
    else if (v == "64") o.precision = 64;
    else return false;
  } else if (key == "REFERENCE_VELOCITY_LB" || key == "REFERENCE_VELOCITY_SI") {
    const double x = to_d(v);
    if (key == "REFERENCE_VELOCITY_LB") o.ref_velocity_lb = x;
    else o.ref_velocity_si = x;
    if (x < 1e-6) return false;
  } else if (key == "FLOW") {
    o.flow = to_d(v);
    if (o.flow < 0) return false;
  } else if (key == "VELOCITY") {
    triple(v, o.velocity_si);
    if (o.velocity_si[0] < 0 || o.velocity_si[1] < 0 || o.velocity_si[2] < 0) return false;
  } else if (key == "RELAXATION_TIME") {
    o.relaxation_time = to_d(v);
    if (o.relaxation_time <= 0.5) return false;
  } else if (key == "KINEMATIC_VISCOSITY") {
    o.kinematic_viscosity = to_d(v);
    if (o.kinematic_viscosity < 0) return false;
  } else if (key == "DENSITY") {
    o.density = to_d(v);
  } else if (key == "LEFT_BOUNDARY") {
    o.bc_left = node_type(v, o.bc_left);
  } else if (key == "RIGHT_BOUNDARY") {
    o.bc_right = node_type(v, o.bc_right);
  } else if (key == "UP_BOUNDARY") {
    o.bc_up = node_type(v, o.bc_up);
  } else if (key == "DOWN_BOUNDARY") {
    o.bc_down = node_type(v, o.bc_down);
  } else if (key == "TIME_SIMULATION") {
    o.time_simulation = std::stof(v);
    if (o.time_simulation < 0) return false;
  } else if (key == "TIME_SAVE") {
    o.time_save = std::stof(v);
    if (o.time_save < 0) return false;
  } else if (key.substr(0, 7) != "INITIAL") {
    return false;
  }
  return true;
}

// getInitialConditions (reader.cpp:592-661): CUBE and CSV_2D shapes that set
// FLUID (no rotation attribute).
void initial_line(flb_deck& d, const std::string& key, const std::string& value, unsigned ln) {
  const char* F = "OptionsCalc.txt";
  if (before(key, '(') != "INITIAL_CONDITION_FLAG") return;
  const int num = to_i(between(key, '(', ')'));
  const bool exists = num <= int(d.initial.size());
  const std::string a = attribute(key);
  if (a == "type") {
    if (exists) fail_line(F, ln, "the initial condition number already exists");
    if (value == "CUBE") d.initial.emplace_back(Kind::Rectangle);
    else if (value == "CSV_2D") d.initial.emplace_back(Kind::Outline);
    return;
  }
  if (a != "name" && a != "position" && a != "height" && a != "xWidth" && a != "zWidth") return;
  if (!exists || num < 1) fail_line(F, ln, "the type of the initial condition must come first");
  Shape& s = d.initial[num - 1];
  if (a == "name") {
    if (s.kind == Kind::Outline) s.file = value + ".csv";
  } else if (a == "position") {
    double c[3];
    triple(value, c);
    s.x0 = c[0];
    s.y0 = c[1];
    s.z0 = c[2];
  } else if (s.kind == Kind::Rectangle) {
    if (a == "height") s.height = to_d(value);
    else if (a == "xWidth") s.xwidth = to_d(value);
    else s.zwidth = to_d(value);
  }
}

// checkOptionsCalculation (reader.cpp:138-265).
void check_options(const flb_deck_info& o, bool all) {
  auto bad = [](const char* m) { fail(FLB_ERR_ARGUMENT, std::string("OptionsCalc.txt: ") + m); };
  const auto zero = [](double a) { return std::abs(a - 0.0) <= 1e-6; };
  if (all && o.use_force &&
      zero(std::abs(o.volumetric_force[0]) + std::abs(o.volumetric_force[1]) +
           std::abs(o.volumetric_force[2])) &&
      zero(std::abs(o.acceleration[0]) + std::abs(o.acceleration[1]) + std::abs(o.acceleration[2])))
    bad("a volumetric force or an acceleration is needed");
  const double vs = o.velocity_si[0] + o.velocity_si[1] + o.velocity_si[2];
  if (all && o.flow > 1e-10 && vs > 1e-10) bad("give a flow or a velocity, not both");
  if (all && o.flow < 1e-10 && vs < 1e-10) bad("a flow or a velocity is needed");
  if (all && o.kinematic_viscosity < 1e-12) bad("the kinematic viscosity must be positive");
  if (all && o.density < 1e-12) bad("the density must be positive");
  if (o.bc_left == kNone || o.bc_right == kNone || o.bc_up == kNone || o.bc_down == kNone)
    bad("every boundary needs a type");
  if (o.bc_left != o.bc_right && (o.bc_left & ~(FLB_INLET | FLB_WALL)))
    bad("a periodic or outlet boundary on the left needs the same type on the right");
  if (all) {
    if (o.block_x <= 0 || o.block_y <= 0) bad("block sizes must be positive");
    if (int64_t(o.block_x) * o.block_y > 1024) bad("more than 1024 threads per block");
  }
  if (all && o.graphics == 0) {
    if (std::abs(o.time_save - 0.0f) <= 1e-6f) bad("TIME_SAVE must be positive");
    if (std::abs(o.time_simulation - 0.0f) <= 1e-6f) bad("TIME_SIMULATION must be positive");
    if (o.time_save >= o.time_simulation) bad("TIME_SAVE must be below TIME_SIMULATION");
  }
}

void read_options(flb_deck& d, const std::string& path, int mode) {
  std::ifstream in(path);
  if (!in.is_open()) fail(FLB_ERR_ARGUMENT, "cannot open " + path);
  flb_deck_info& o = d.opt;
  std::string line;
  unsigned ln = 0;
  static const char* kDomainKeys[] = {"TYPE_PROBLEM", "LEFT_BOUNDARY", "RIGHT_BOUNDARY",
                                      "UP_BOUNDARY", "DOWN_BOUNDARY"};
  while (std::getline(in, line)) {
    ++ln;
    std::string compact = line;
    compact.erase(std::remove(compact.begin(), compact.end(), ' '), compact.end());
    std::string kv[2];
    if (mode == FLB_DECK_DOMAIN) {
      // readSomeOptionsCalculation works on the line with spaces removed.
      if (compact.empty() || compact[0] == '#') continue;
      if (!split_option(compact, kv)) fail_line("OptionsCalc.txt", ln, "more than one '='");
      initial_line(d, kv[0], kv[1], ln);
      if (std::find(std::begin(kDomainKeys), std::end(kDomainKeys), kv[0]) == std::end(kDomainKeys))
        continue;
      if (!option_line(o, kv[0], kv[1])) fail_line("OptionsCalc.txt", ln, "invalid option");
    } else {
      // As in readOptionsCalculation a line of blanks is not skipped: its empty
      // key is an unknown option.
      if ((!compact.empty() && compact[0] == '#') || line.empty()) continue;
      if (!split_option(line, kv)) fail_line("OptionsCalc.txt", ln, "more than one '='");
      if (!option_line(o, kv[0], kv[1]))
        fail_line("OptionsCalc.txt", ln, "invalid option '" + kv[0] + "'");
      initial_line(d, kv[0], kv[1], ln);
    }
  }
  // modifyCalculationData: initial-condition shapes are set up now.
  for (Shape& s : d.initial) {
    if (s.kind == Kind::Rectangle) s.init_rectangle();
    else if (s.kind == Kind::Outline) s.init_outline(d.dir);
  }
  check_options(o, mode == FLB_DECK_CALCULATION);
}

void read_geometry(flb_deck& d, const std::string& path) {
  std::ifstream in(path);
  if (!in.is_open()) fail(FLB_ERR_ARGUMENT, "cannot open " + path);
  std::string line;
  unsigned ln = 0;
  while (std::getline(in, line)) {
    ++ln;
    std::string compact = line;
    compact.erase(std::remove(compact.begin(), compact.end(), ' '), compact.end());
    if ((!compact.empty() && compact[0] == '#') || line.empty()) continue;
    std::string kv[2];
    if (!split_option(line, kv)) fail_line("Geometry.txt", ln, "more than one '='");
    geometry_line(d, kv[0], kv[1], ln);
  }
  finish_geometry(d);
}

// getEndPointBaseCoordinate / getBaseCoordinatey (src/initData.h:42-66).
uint32_t end_point(double xs, const std::vector<Point3>& pts) {
  uint32_t k = 1;
  for (size_t i = 1; i < pts.size(); ++i, ++k)
    if (pts[i].x >= xs - 1e-12) return k;
  return uint32_t(pts.size() - 1);  // past the last point: keep the last segment
}
double base_y(double xs, uint32_t k, const std::vector<Point3>& pts) {
  const double slope = (pts[k].y - pts[k - 1].y) / (pts[k].x - pts[k - 1].x);
  return pts[k - 1].y + (xs - pts[k - 1].x) * slope;
}

// distanceSquaredPointLine (src/math/math.cpp:89-103): squared distance from
// (x, y, z) to the line through points k-1 and k raised by `lift`.
double dist2_line(double x, double y, double z, const std::vector<Point3>& p, uint32_t k,
                  double lift) {
  const double qx = x - p[k].x, qy = y - (p[k].y + lift), qz = z - p[k].z;
  const double vx = p[k - 1].x - p[k].x, vy = p[k - 1].y - p[k].y, vz = p[k - 1].z - p[k].z;
  const double cx = qy * vz - qz * vy, cy = qz * vx - qx * vz, cz = qx * vy - qy * vx;
  return (cx * cx + cy * cy + cz * cz) / (vx * vx + vy * vy + vz * vz);
}

bool cdw_inside(const Cdw& c, double xs, double ys, double zs, const std::vector<Point3>& p,
                uint32_t k) {
  if (c.circle) return dist2_line(xs, ys, zs, p, k, c.radius) <= c.radius * c.radius + kEps;
  const double slope = (p[k].y - p[k - 1].y) / (p[k].x - p[k - 1].x);
  const double y_down = p[k - 1].y + (xs - p[k - 1].x) * slope;
  return ys >= y_down - kEps && ys <= y_down + kEps;
}

int catch_all(flb_deck* d, std::string* msg, int (*body)(flb_deck*, void*), void* arg) {
  try {
    return body(d, arg);
  } catch (const DeckError& e) {
    if (msg) *msg = e.msg;
    return e.status;
  } catch (const std::exception& e) {
    if (msg) *msg = std::string("invalid number in deck: ") + e.what();
    return FLB_ERR_ARGUMENT;
  }
}

}  // namespace

extern "C" int flb_deck_load(const char* dir, int mode, flb_deck** out, char* err,
                             size_t err_len) {
  if (err && err_len) err[0] = '\0';
  if (!dir || !out || (mode != FLB_DECK_DOMAIN && mode != FLB_DECK_CALCULATION))
    return FLB_ERR_ARGUMENT;
  *out = nullptr;
  std::unique_ptr<flb_deck> d(new flb_deck());
  d->dir = dir;
  flb_deck_info& o = d->opt;
  o.problem = FLB_MONOFLUID;
  o.bc_left = o.bc_right = o.bc_up = o.bc_down = kNone;
  o.precision = 32;
  o.ref_velocity_lb = 0.1;
  o.ref_velocity_si = 1.0;
  std::string msg;
  struct Arg {
    int mode;
  } arg{mode};
  const int rc = catch_all(d.get(), &msg, [](flb_deck* dd, void* a) -> int {
    read_geometry(*dd, dd->dir + "/Geometry.txt");
    read_options(*dd, dd->dir + "/OptionsCalc.txt", static_cast<Arg*>(a)->mode);
    return FLB_OK;
  }, &arg);
  if (rc != FLB_OK) {
    if (err && err_len) std::snprintf(err, err_len, "%s", msg.c_str());
    return rc;
  }
  flb_deck_info& i = d->opt;
  i.nx = d->nx;
  i.ny = d->ny;
  i.x_min = d->x_min;
  i.x_max = d->x_max;
  i.y_min = d->y_min;
  i.y_max = d->y_max;
  i.size = d->size;
  i.n_points = uint32_t(d->points.size());
  i.n_cdws = uint32_t(d->cdws.size());
  i.n_obstacles = uint32_t(d->obstacles.size());
  i.n_initial = uint32_t(d->initial.size());
  *out = d.release();
  return FLB_OK;
}

extern "C" int flb_deck_get_info(const flb_deck* d, flb_deck_info* info) {
  if (!d || !info) return FLB_ERR_ARGUMENT;
  *info = d->opt;
  return FLB_OK;
}

extern "C" int flb_deck_build_flags(const flb_deck* d, uint8_t* flags) {
  if (!d || !flags) return FLB_ERR_ARGUMENT;
  const flb_deck_info& o = d->opt;
  const auto& pts = d->points;
  const double x0 = pts[d->idx_init].x - kEps, x1 = pts[d->idx_end].x + kEps;
  const uint8_t node = o.problem == FLB_FREE_SURFACE ? FLB_GAS : FLB_FLUID;
  const uint32_t nx = d->nx, ny = d->ny;
  for (uint32_t y = 0; y < ny; ++y) {
    for (uint32_t x = 0; x < nx; ++x) {
      const size_t idx = size_t(x) + size_t(y) * nx;
      // The reference pushes the index itself (truncated to 8 bits) and
      // overwrites it below; a node outside the CDW span that is neither on
      // the bed nor on the top keeps that value.
      uint8_t v = uint8_t(idx);
      const double xs = d->x_min + x * d->size;
      const double ys = d->y_min + y * d->size;
      const double zs = d->z_min + 0 * d->size;
      const uint32_t k = end_point(xs, pts);
      const double yb = base_y(xs, k, pts);
      if (xs >= x0 && xs <= x1) {
        for (const Cdw& c : d->cdws) v = cdw_inside(c, xs, ys, zs, pts, k) ? node : uint8_t(FLB_WALL);
      } else {
        if (ys <= yb + kEps) v = o.bc_down;
        else if (ys >= d->y_max - kEps) v = o.bc_up;
      }
      if (ys <= yb + kEps && v == node) v = o.bc_down;
      else if (ys >= d->y_max - kEps && v == node) v = o.bc_up;
      else if (x == 0 && v == node) v = o.bc_left;
      else if (x == nx - 1 && v == node) v = o.bc_right;
      for (const Shape& s : d->obstacles)
        if (s.inside(xs, ys, zs)) v = FLB_WALL;
      for (const Shape& s : d->initial)
        if (s.inside(xs, ys, zs)) v = FLB_FLUID;
      flags[idx] = v;
    }
  }
  if (o.problem == FLB_FREE_SURFACE) return flb_seed_free_surface(flags, nx, ny);
  return FLB_OK;
}

extern "C" int flb_deck_units_in(const flb_deck* d, flb_units_in* in) {
  if (!d || !in) return FLB_ERR_ARGUMENT;
  const flb_deck_info& o = d->opt;
  std::memset(in, 0, sizeof(*in));
  in->length_si = d->x_max - d->x_min;
  in->nx = d->nx;
  in->ref_velocity_si = o.ref_velocity_si;
  in->ref_velocity_lb = o.ref_velocity_lb;
  in->density_si = o.density;
  in->relaxation_time = o.relaxation_time;
  in->kinematic_viscosity = o.kinematic_viscosity;
  in->use_force = o.use_force;
  in->acceleration_x = o.acceleration[0];
  in->acceleration_y = o.acceleration[1];
  in->volumetric_force_x = o.volumetric_force[0];
  in->volumetric_force_y = o.volumetric_force[1];
  in->velocity_si_x = o.velocity_si[0];
  in->velocity_si_y = o.velocity_si[1];
  return FLB_OK;
}

extern "C" void flb_deck_free(flb_deck* d) { delete d; }
