// Host-side construction of the serial-partition extruded cubed-sphere function spaces:
// panel topology, local vertex numbering, 3D entity connectivity, column-contiguous dofmaps
// and the closed-form 4-colouring.  Runs once at set-up, like the reference's mesh and
// function-space constructors; the result is uploaded once and stays on the device.
//
// The numbering must equal the reference's bit for bit (it fixes which dof is which), so the
// steps follow the reference algorithms:
//   panel adjacency and vertex numbering  mesh_tools/source/support/gencube_ps_mod.F90:336-720
//   panel orientation (PANEL_ROTATIONS)   gencube_ps_mod.F90:69, 1561-1601
//   local vertex ids (first encounter)    infrastructure/source/mesh/local_mesh_mod.f90:381-406
//   3D face / edge ids                    infrastructure/source/mesh/mesh_constructor_helper_functions_mod.F90:266-438
//   dofmap                                infrastructure/source/function_space/function_space_constructor_helper_functions_mod.F90:1715-2300
//   colouring                             infrastructure/source/mesh/mesh_colouring_mod.F90:151-308, 703-837
//   vertex longitude / latitude           gencube_ps_mod.F90:887-1113 (calc_coords, before orientation)
//   extrusion eta                         infrastructure/source/mesh/extrusion_mod.f90:292-305, 341-354, 390-409, 618-629
//   Cartesian coordinate fields           mesh_constructor_helper_functions_mod.F90:120-252 (mesh_extruder),
//                                         :450-477 (set_dz), components/driver/source/driver_coordinates_mod.F90:357-454
#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>
#include <cstring>

#include "internal.h"

namespace lfric {
namespace {

enum Dir { DW = 0, DS = 1, DE = 2, DN = 3 };       // cell_next slots
enum Corner { CSW = 0, CSE = 1, CNE = 2, CNW = 3 };  // vertex slots

struct Panelled {
  int n, cpp, nc;
  std::vector<int> next;  // [cell-1][dir], 1-based cells
  std::vector<int> vert;  // [cell-1][corner], 1-based global vertex ids
  int &nb(int cell, int d) { return next[(cell - 1) * 4 + d]; }
  int &vx(int cell, int c) { return vert[(cell - 1) * 4 + c]; }
  // first cell of row/column along a panel side, i = 1..n
  int side_cell(int panel, int side, int i) const {
    int b = (panel - 1) * cpp;
    switch (side) {
      case DW: return b + n * (i - 1) + 1;
      case DS: return b + cpp - n + i;
      case DE: return b + n * i;
      default: return b + i;
    }
  }
};

// Neighbour panel across each side, the side of that panel it meets and whether the running
// index reverses.  Panels: 1-4 around the equator, 5 on top, 6 below; cells numbered from the
// NW corner of each panel, row by row.
struct PanelLink {
  int panel, side, reversed;
};
constexpr PanelLink kLinks[6][4] = {
    // W              S              E              N
    {{4, DE, 0}, {6, DW, 1}, {2, DW, 0}, {5, DW, 0}},
    {{1, DE, 0}, {6, DN, 0}, {3, DW, 0}, {5, DS, 0}},
    {{2, DE, 0}, {6, DE, 0}, {4, DW, 0}, {5, DE, 1}},
    {{3, DE, 0}, {6, DS, 1}, {1, DW, 0}, {5, DN, 1}},
    {{1, DN, 0}, {2, DN, 0}, {3, DN, 1}, {4, DN, 1}},
    {{1, DS, 1}, {4, DS, 1}, {3, DS, 0}, {2, DS, 0}},
};

void panel_adjacency(Panelled &P) {
  const int n = P.n;
  for (int c = 1; c <= P.nc; ++c) {
    P.nb(c, DW) = c - 1;
    P.nb(c, DS) = c + n;
    P.nb(c, DE) = c + 1;
    P.nb(c, DN) = c - n;
  }
  for (int p = 1; p <= 6; ++p)
    for (int s = 0; s < 4; ++s) {
      const PanelLink &L = kLinks[p - 1][s];
      for (int i = 1; i <= n; ++i)
        P.nb(P.side_cell(p, s, i), s) = P.side_cell(L.panel, L.side, L.reversed ? n + 1 - i : i);
    }
}

void vertex_numbering(Panelled &P) {
  const int n = P.n, cpp = P.cpp;
  std::fill(P.vert.begin(), P.vert.end(), 0);
  auto nb = [&](int c, int d) { return P.nb(c, d); };
  int id = 1;
  // equatorial panels: NW corner of every cell takes the cell's id, NE from the E neighbour
  for (int c = 1; c <= 4 * cpp; ++c) P.vx(c, CNW) = id++;
  for (int c = 1; c <= 4 * cpp; ++c) P.vx(c, CNE) = P.vx(nb(c, DE), CNW);
  for (int p = 1; p <= 4; ++p)
    for (int c = (p - 1) * cpp + 1; c <= p * cpp - n; ++c) {
      P.vx(c, CSW) = P.vx(nb(c, DS), CNW);
      P.vx(c, CSE) = P.vx(nb(c, DS), CNE);
    }
  // interior vertices of the top panel
  for (int r = 0; r <= n - 2; ++r)
    for (int c = 4 * cpp + r * n + 1; c <= 4 * cpp + r * n + n - 1; ++c) {
      int v = id++;
      P.vx(c, CSE) = v;
      P.vx(nb(c, DE), CSW) = v;
      P.vx(nb(c, DS), CNE) = v;
      P.vx(nb(nb(c, DS), DE), CNW) = v;
    }
  // bottom panel: NW of every cell, then SW below the first rows, then the last row
  for (int c = 5 * cpp + 1; c <= 6 * cpp; ++c) P.vx(c, CNW) = id++;
  for (int c = 5 * cpp + 1; c <= 6 * cpp - n; ++c) P.vx(c, CSW) = P.vx(nb(c, DS), CNW);
  for (int c = 6 * cpp - n + 1; c <= 6 * cpp; ++c) P.vx(c, CSW) = id++;
  // last row of panel 3 and its SE corner
  for (int c = 3 * cpp - n + 1; c <= 3 * cpp; ++c) {
    P.vx(c, CSW) = id;
    P.vx(nb(c, DW), CSE) = id;
    ++id;
  }
  P.vx(3 * cpp, CSE) = id;
  P.vx(nb(3 * cpp, DE), CSW) = id;
  // seams between panels
  for (int c = cpp - n + 1; c <= cpp; ++c) {
    P.vx(c, CSW) = P.vx(nb(c, DS), CSW);
    P.vx(c, CSE) = P.vx(nb(c, DS), CNW);
  }
  for (int c = 1; c <= n; ++c) {
    P.vx(nb(c, DN), CNW) = P.vx(c, CNW);
    P.vx(nb(c, DN), CSW) = P.vx(c, CNE);
  }
  for (int c = 2 * cpp + 1; c <= 2 * cpp + n; ++c) {
    P.vx(nb(c, DN), CSE) = P.vx(c, CNW);
    P.vx(nb(c, DN), CNE) = P.vx(c, CNE);
  }
  for (int c = 3 * cpp - n + 1; c <= 3 * cpp; ++c) {
    P.vx(nb(c, DS), CNE) = P.vx(c, CSW);
    P.vx(nb(c, DS), CSE) = P.vx(c, CSE);
  }
  for (int c = 3 * cpp + 1; c <= 3 * cpp + n; ++c) {
    P.vx(nb(c, DN), CNE) = P.vx(c, CNW);
    P.vx(nb(c, DN), CNW) = P.vx(c, CNE);
  }
  for (int r = 0; r <= n - 1; ++r)
    for (int c = 5 * cpp + 1 + r * n; c <= 5 * cpp + r * n + n - 1; ++c) {
      P.vx(c, CNE) = P.vx(nb(c, DE), CNW);
      P.vx(c, CSE) = P.vx(nb(c, DE), CSW);
    }
  for (int c = 6 * cpp - n + 1; c <= 6 * cpp; ++c) {
    P.vx(nb(c, DS), CSW) = P.vx(c, CSE);
    P.vx(nb(c, DS), CSE) = P.vx(c, CSW);
  }
  for (int c = 2 * cpp - n + 1; c <= 2 * cpp; ++c) {
    P.vx(c, CSW) = P.vx(nb(c, DS), CNW);
    P.vx(c, CSE) = P.vx(nb(c, DS), CNE);
  }
  for (int c = cpp + 1; c <= cpp + n; ++c) {
    P.vx(nb(c, DN), CSW) = P.vx(c, CNW);
    P.vx(nb(c, DN), CSE) = P.vx(c, CNE);
  }
}

// xyz2ll of coord_transform_mod (infrastructure/source/utilities/coord_transform_mod.F90:226-283):
// longitude from atan(y/x) moved to the x < 0 half-plane by +-pi, latitude atan(z / sqrt(x^2+y^2))
constexpr double kPi = 3.141592653589793115997963468544185161590576171875;  // 4 atan(1) in FP64
void xyz_to_lonlat(double x, double y, double z, double &lon, double &lat) {
  if (x == 0.0) {
    lon = (y >= 0.0 ? 0.5 : -0.5) * kPi;
  } else {
    lon = std::atan(y / x);
    if (x < 0.0) lon = y > 0.0 ? lon + kPi : lon - kPi;
  }
  const double r = std::sqrt(x * x + y * y);
  if (r <= DBL_EPSILON) {
    lat = (z > 0.0 ? 0.5 : -0.5) * kPi;
    lon = 0.0;
  } else {
    lat = std::atan(z / r);
  }
}

// Longitude and negated latitude (radians) of every global vertex, on the equiangular grid of
// each panel: grid point (i, j) of a panel is the unit cube-face point (1, tan(l1), tan(l2))
// normalised, l = -pi/4 + (0..n) pi/(2n), turned onto the panel's face.  Panels 1-4 give their
// vertices ids in grid order; the top and bottom panels' vertices are found through the corners
// of the unoriented cells.  P.vert must hold the numbering before orient_panels.
void vertex_lonlat(const Panelled &P, std::vector<double> &lon, std::vector<double> &lat) {
  const int n = P.n, cpp = P.cpp;
  lon.assign(static_cast<size_t>(6) * cpp + 2, 0.0);
  lat.assign(lon.size(), 0.0);
  const double dl = 0.5 * kPi / n, pio4 = kPi / 4.0;
  auto unit = [&](int i, int j, double &xs, double &ys, double &zs) {
    const double t1 = std::tan(i * dl - pio4), t2 = std::tan(j * dl - pio4);
    xs = 1.0 / std::sqrt(1.0 + t1 * t1 + t2 * t2);
    ys = xs * t1;
    zs = xs * t2;
  };
  auto put = [&](int gid, double x, double y, double z) {
    double lo, la;
    xyz_to_lonlat(x, y, z, lo, la);
    lon[gid - 1] = lo;
    lat[gid - 1] = -la;
  };
  auto V = [&](int cell, int corner) { return P.vert[(cell - 1) * 4 + corner]; };
  double xs, ys, zs;
  for (int y = 1; y <= n; ++y)
    for (int x = 1; x <= n; ++x) {
      unit(x - 1, y - 1, xs, ys, zs);
      const int v = (y - 1) * n + x;
      put(v, xs, ys, zs);                  // panel 1
      put(v + cpp, -ys, xs, zs);           // panel 2
      put(v + 2 * cpp, -xs, -ys, zs);      // panel 3
      put(v + 3 * cpp, ys, -xs, zs);       // panel 4
    }
  for (int y = 1; y <= n - 1; ++y)        // top panel: SE corners of its interior cells
    for (int x = 1; x <= n - 1; ++x) {
      unit(x, y, xs, ys, zs);
      put(V(4 * cpp + (y - 1) * n + x, CSE), -ys, zs, -xs);
    }
  for (int y = 1; y <= n; ++y)            // bottom panel: NW corners
    for (int x = 1; x <= n; ++x) {
      unit(x - 1, y - 1, xs, ys, zs);
      put(V(5 * cpp + (y - 1) * n + x, CNW), -ys, -zs, xs);
    }
  for (int x = 1; x <= n; ++x) {          // bottom panel: SW corners of its last row
    unit(x - 1, n, xs, ys, zs);
    put(V(6 * cpp - n + x, CSW), -ys, -zs, xs);
  }
  for (int y = 1; y <= n; ++y) {          // bottom panel: NE corners of its last column
    unit(n, y - 1, xs, ys, zs);
    put(V(5 * cpp + y * n, CNE), -ys, -zs, xs);
  }
  unit(n, n, xs, ys, zs);                 // the last vertex
  put(6 * cpp + 2, -ys, -zs, xs);
}

// Panels 3 and 4 are turned by one slot (cshift +1), panel 5 by one slot the other way.
void orient_panels(Panelled &P) {
  constexpr int rot[6] = {0, 0, 1, 1, -1, 0};
  for (int p = 1; p <= 6; ++p) {
    if (!rot[p - 1]) continue;
    for (int c = (p - 1) * P.cpp + 1; c <= p * P.cpp; ++c)
      for (std::vector<int> *arr : {&P.vert, &P.next}) {
        int *a = &(*arr)[(c - 1) * 4];
        std::array<int, 4> t{a[0], a[1], a[2], a[3]};
        for (int i = 0; i < 4; ++i) a[i] = t[(i + (rot[p - 1] > 0 ? 1 : 3)) % 4];
      }
  }
}

int closed_form_colour(int n, int cell) {
  // cell_colour_order_per_panel(column,row,panel), even n (mesh_colouring_mod.F90:242-247)
  static const int even[6][2][2] = {{{1, 3}, {2, 4}}, {{3, 1}, {2, 4}}, {{3, 1}, {4, 2}},
                                    {{1, 3}, {4, 2}}, {{3, 4}, {2, 1}}, {{2, 1}, {3, 4}}};
  static const int odd[6][2][2] = {{{1, 3}, {2, 4}}, {{4, 2}, {1, 3}}, {{1, 3}, {2, 4}},
                                   {{4, 2}, {1, 3}}, {{3, 4}, {2, 1}}, {{3, 4}, {2, 1}}};
  const int cpp = n * n;
  int panel = (cell - 1) / cpp;
  int rel = cell - cpp * panel;
  int column = 1 - (rel % n) % 2;        // 0-based column slot
  int row = ((rel - 1) / n) % 2;         // 0-based row slot
  return (n % 2 == 0 ? even : odd)[panel][column][row];
}

// No two cells of one colour may share an edge or a vertex (mesh_colouring_mod.F90:765-837,
// checked here directly on shared vertices).
bool colouring_valid(const Topology &t) {
  const int K = t.ncolours;
  std::vector<int> seen(static_cast<size_t>(t.nverts + 1) * K, 0);
  for (int c = 1; c <= t.ncells; ++c) {
    int col = t.colour_of_cell[c - 1] - 1;
    if (col < 0 || col >= K) return false;
    for (int v = 0; v < 4; ++v) {
      int &s = seen[static_cast<size_t>(t.vert_on_cell[(c - 1) * 4 + v]) * K + col];
      if (s) return false;
      s = c;
    }
  }
  return true;
}

// set_colours_generic (mesh_colouring_mod.F90:514-650) with choose_colour_greedy (:848-864),
// the reference's fallback when the closed form does not validate (:100-113): cells in id order
// take the lowest colour (of at most 50) not yet held by a cell sharing an edge, or a vertex as
// found through the N and S neighbours.  Returns the number of colours, 0 on failure.
int greedy_colour(Topology &t) {
  constexpr int kMaxColours = 50;
  enum { kW = 0, kS = 1, kE = 2, kN = 3 };
  const int nc = t.ncells;
  auto nx = [&](int d, int c) { return c > 0 ? t.cell_next[static_cast<size_t>(c - 1) * 4 + d] : 0; };
  auto cm = [&](int c) { return c > 0 ? t.colour_of_cell[c - 1] : 0; };
  std::fill(t.colour_of_cell.begin(), t.colour_of_cell.end(), 0);
  std::vector<int> count(kMaxColours + 1, 0);
  for (int c = 1; c <= nc; ++c) {
    bool used[kMaxColours + 1] = {};
    auto mark = [&](int cell) { used[cm(cell)] = true; };
    // N and S neighbours, then the two cells across each of them that share only a vertex;
    // which two depends on the neighbour's orientation (its S or N leads back to c)
    for (int d : {kN, kS}) {
      const int u = nx(d, c);
      if (u <= 0) continue;
      mark(u);
      if (c == nx(d == kN ? kS : kN, u)) {
        mark(nx(kE, u));
        mark(nx(kW, u));
      } else {
        mark(nx(kN, u));
        mark(nx(kS, u));
      }
    }
    mark(nx(kE, c));
    mark(nx(kW, c));
    int colour = 0;
    for (int i = 1; i <= kMaxColours; ++i)
      if (!used[i]) {
        colour = i;
        break;
      }
    if (colour == 0) return 0;
    t.colour_of_cell[c - 1] = colour;
    count[colour]++;
  }
  int ncol = kMaxColours;
  for (int i = 1; i <= kMaxColours; ++i)
    if (count[i] == 0) {
      ncol = i - 1;
      break;
    }
  return ncol;
}

// test_mesh_planar_quadrilateral (mesh_colouring_mod.F90:663-691): every neighbour across a
// side leads back across the opposite side.
bool planar_quadrilateral(const Topology &t) {
  static const int opposite[4] = {2, 3, 0, 1};  // W<->E, S<->N
  for (int c = 1; c <= t.ncells; ++c)
    for (int d = 0; d < 4; ++d) {
      const int o = t.cell_next[static_cast<size_t>(c - 1) * 4 + d];
      if (o > 0 && o <= t.ncells && t.cell_next[static_cast<size_t>(o - 1) * 4 + opposite[d]] != c) return false;
    }
  return true;
}

// set_colours_planar_quadrilateral (mesh_colouring_mod.F90:348-495): breadth-first from each
// uncoloured cell (the first, then the next uncoloured one by id), visiting the N, E, S, W
// neighbours of each cell of the current front, a neighbour taking the colour that the
// direction and the cell's colour give in cell_next_colour.
void planar_colour(Topology &t) {
  // next_colour[direction W,S,E,N][colour 1..4]
  static const int next_colour[4][4] = {{2, 1, 4, 3}, {3, 4, 1, 2}, {2, 1, 4, 3}, {3, 4, 1, 2}};
  static const int order[4] = {3, 2, 1, 0};  // N, E, S, W
  const int nc = t.ncells;
  std::fill(t.colour_of_cell.begin(), t.colour_of_cell.end(), 0);
  std::vector<int> todo(static_cast<size_t>(nc) + 1, 0);
  int cell = 1, start = 1, end = 1;
  todo[1] = 1;
  while (cell <= nc) {
    t.colour_of_cell[cell - 1] = 1;
    while (start <= end) {
      const int last = end;
      for (int i = start; i <= last; ++i) {
        const int c = todo[i];
        for (int k = 0; k < 4; ++k) {
          const int d = order[k];
          const int o = t.cell_next[static_cast<size_t>(c - 1) * 4 + d];
          if (o > 0 && o <= nc && t.colour_of_cell[o - 1] == 0) {
            t.colour_of_cell[o - 1] = next_colour[d][t.colour_of_cell[c - 1] - 1];
            todo[++end] = o;
          }
        }
      }
      start = last + 1;
    }
    while (t.colour_of_cell[cell - 1] != 0) {
      ++cell;
      if (cell > nc) break;
    }
  }
  t.ncolours = 4;
}

}  // namespace

int build_cubedsphere_topology(int n, Topology &t) {
  if (n < 2) return fail(LFRIC_E_ARG, "cubed sphere needs edge_cells >= 2");
  Panelled P{n, n * n, 6 * n * n, {}, {}};
  P.next.assign(static_cast<size_t>(P.nc) * 4, 0);
  P.vert.assign(static_cast<size_t>(P.nc) * 4, 0);
  panel_adjacency(P);
  vertex_numbering(P);
  std::vector<double> lon, lat;
  vertex_lonlat(P, lon, lat);
  orient_panels(P);
  if (int rc = build_topology(P.nc, P.next, P.vert, 6, t)) return rc;
  // coordinates by local vertex id (orientation only permutes the corner slots)
  t.vert_lon.assign(t.nverts, 0.0);
  t.vert_lat.assign(t.nverts, 0.0);
  for (size_t i = 0; i < P.vert.size(); ++i) {
    t.vert_lon[t.vert_on_cell[i] - 1] = lon[P.vert[i] - 1];
    t.vert_lat[t.vert_on_cell[i] - 1] = lat[P.vert[i] - 1];
  }
  return LFRIC_OK;
}

// eta(0:nl) of the uniform (k/nl), quadratic ((k/nl)^2) and geometric (stretching 1.03)
// extrusions
int extrusion_eta(int kind, int nl, std::vector<double> &eta) {
  LF_CHECK(nl >= 1, LFRIC_E_ARG, "extrusion needs at least one layer");
  eta.assign(static_cast<size_t>(nl) + 1, 0.0);
  switch (kind) {
    case LFRIC_EXTRUSION_UNIFORM:
      for (int k = 0; k <= nl; ++k) eta[k] = static_cast<double>(k) / static_cast<double>(nl);
      return LFRIC_OK;
    case LFRIC_EXTRUSION_QUADRATIC:
      for (int k = 0; k <= nl; ++k) {
        const double r = static_cast<double>(k) / static_cast<double>(nl);
        eta[k] = r * r;
      }
      return LFRIC_OK;
    case LFRIC_EXTRUSION_GEOMETRIC: {
      // stretching_factor**number_of_layers is an integer power (repeated products here);
      // geometric_func's stretching_factor**real(k) a real one
      const double sf = 1.03;
      double sfn = 1.0;
      for (int k = 0; k < nl; ++k) sfn *= sf;
      const double deta = (sf - 1.0) / (sfn - 1.0);
      for (int k = 1; k <= nl; ++k) eta[k] = deta * (std::pow(sf, static_cast<double>(k)) - 1.0) / (sf - 1.0);
      return LFRIC_OK;
    }
  }
  return fail(LFRIC_E_ARG, "unknown extrusion");
}

// Cartesian coordinates of the lowest-order W0 dofs of a generated cubed-sphere space, as
// assign_coordinate_field's xyz branch leaves them: dz = (eta(k) - eta(k-1)) * depth; vertex at
// level j: (x, y, z) = llr2xyz(lon, lat, 1 + sum(dz(1:j))) (the mesh extruder's radius over a
// zero surface height), projected onto the shell of radius scaled_radius + sum(dz(1:j)); the
// top level's radius is (scaled_radius + sum(dz(1:nl-1))) + dz(nl), the top-node form the last
// layer writes.
int cubedsphere_chi(const FunctionSpace &f, int kind, double domain_top, double scaled_radius, double *chi1,
                    double *chi2, double *chi3) {
  LF_CHECK(f.topo && !f.topo->vert_lon.empty(), LFRIC_E_UNSUPPORTED,
           "coordinates are known for generated cubed-sphere spaces only");
  LF_CHECK(f.space == LFRIC_W0 && f.order == 0, LFRIC_E_UNSUPPORTED, "coordinate fields live on lowest-order W0");
  LF_CHECK(domain_top > 0.0 && scaled_radius > 0.0, LFRIC_E_ARG, "domain top and radius must be positive");
  const int nl = f.nlayers;
  std::vector<double> eta;
  if (int rc = extrusion_eta(kind, nl, eta)) return rc;
  std::vector<double> dz(nl), sdz(static_cast<size_t>(nl) + 1, 0.0);
  for (int k = 1; k <= nl; ++k) dz[k - 1] = (eta[k] - eta[k - 1]) * domain_top;
  for (int k = 1; k <= nl; ++k) sdz[k] = sdz[k - 1] + dz[k - 1];
  const Topology &t = *f.topo;
  std::vector<char> done(t.nverts, 0);
  for (int c = 0; c < f.ncells; ++c)
    for (int i = 0; i < 4; ++i) {
      const int v = t.vert_on_cell[static_cast<size_t>(c) * 4 + i] - 1;
      if (done[v]) continue;
      done[v] = 1;
      const int base = f.dofmap[static_cast<size_t>(c) * 8 + i] - 1;
      const double sl = std::sin(t.vert_lon[v]), cl = std::cos(t.vert_lon[v]);
      const double sa = std::sin(t.vert_lat[v]), ca = std::cos(t.vert_lat[v]);
      for (int j = 0; j <= nl; ++j) {
        const double r = 1.0 + sdz[j];
        const double x = r * cl * ca, y = r * sl * ca, z = r * sa;
        double rc = j < nl ? (scaled_radius + sdz[j]) + 0.0 * dz[j] : (scaled_radius + sdz[nl - 1]) + 1.0 * dz[nl - 1];
        rc = rc / std::sqrt(x * x + y * y + z * z);
        chi1[base + j] = x * rc;
        chi2[base + j] = y * rc;
        chi3[base + j] = z * rc;
      }
    }
  return LFRIC_OK;
}

// Everything below the global mesh: local vertex ids, the 3D connectivity of the base layer
// and the colouring (mesh_colouring_mod.F90:49-128: the cubed-sphere closed form for a six-panel
// mesh of 6 n^2 cells, the planar breadth-first form for a one-panel quadrilateral mesh, the
// generic greedy form otherwise or when the first does not validate).
int build_topology(int ncells, const std::vector<int> &cell_next, const std::vector<int> &vert_global, int npanels,
                   Topology &t) {
  LF_CHECK(ncells > 0 && cell_next.size() == static_cast<size_t>(ncells) * 4 && vert_global.size() == cell_next.size(),
           LFRIC_E_ARG, "mesh: connectivity sizes");
  for (int v : vert_global) LF_CHECK(v > 0, LFRIC_E_ARG, "mesh: a face without four corner nodes");
  for (int c : cell_next) LF_CHECK(c >= 0 && c <= ncells, LFRIC_E_ARG, "mesh: a neighbour id out of range");
  const int n = static_cast<int>(std::lround(std::sqrt(ncells / 6.0)));
  const bool cubed = npanels == 6 && 6 * n * n == ncells;
  struct {
    std::vector<int> next, vert;
    int nc;
  } P{cell_next, vert_global, ncells};
  t.edge_cells = cubed ? n : 0;
  t.ncells = P.nc;
  t.cell_next = P.next;
  // local vertex ids in order of first appearance
  {
    int maxg = *std::max_element(P.vert.begin(), P.vert.end());
    std::vector<int> lid(static_cast<size_t>(maxg) + 1, 0);
    t.vert_on_cell.resize(P.vert.size());
    int nv = 0;
    for (size_t i = 0; i < P.vert.size(); ++i) {
      int &l = lid[P.vert[i]];
      if (!l) l = ++nv;
      t.vert_on_cell[i] = l;
    }
    t.nverts = nv;
  }
  // 3D connectivity of the base layer: side faces and horizontal edges are numbered when
  // first met and shared with the neighbour across that side; B/T faces and bottom/top edge
  // pairs take consecutive ids; vertical edges are shared round their vertex.
  const int nc = t.ncells;
  t.face_on_cell.assign(static_cast<size_t>(nc) * 6, 0);
  t.edge_on_cell.assign(static_cast<size_t>(nc) * 12, 0);
  auto NX = [&](int c, int s) { return t.cell_next[(c - 1) * 4 + s]; };
  auto F = [&](int c, int f) -> int & { return t.face_on_cell[(c - 1) * 6 + f]; };
  auto E = [&](int c, int e) -> int & { return t.edge_on_cell[(c - 1) * 12 + e]; };
  auto V = [&](int c, int v) { return t.vert_on_cell[(c - 1) * 4 + v]; };
  int fid = 1, nside = 0;
  for (int c = 1; c <= nc; ++c) {
    for (int s = 0; s < 4; ++s) {
      if (F(c, s)) continue;
      F(c, s) = fid;
      int o = NX(c, s);
      if (o)
        for (int so = 0; so < 4; ++so)
          if (NX(o, so) == c) F(o, so) = fid;
      ++fid;
      ++nside;
    }
    F(c, 4) = fid++;
    F(c, 5) = fid++;
  }
  int eid = 1, nh = 0;
  for (int c = 1; c <= nc; ++c) {
    for (int s = 0; s < 4; ++s) {
      if (E(c, s)) continue;
      E(c, s) = eid;
      E(c, s + 8) = eid + 1;
      int o = NX(c, s);
      if (o)
        for (int so = 0; so < 4; ++so)
          if (NX(o, so) == c) {
            E(o, so) = eid;
            E(o, so + 8) = eid + 1;
          }
      eid += 2;
      ++nh;
    }
    for (int v = 0; v < 4; ++v) {
      if (E(c, 4 + v)) continue;
      E(c, 4 + v) = eid;
      const int vert = V(c, v);
      for (int s = 0; s < 4; ++s) {
        int o = NX(c, s);
        if (!o) continue;
        for (int vo = 0; vo < 4; ++vo) {
          if (V(o, vo) != vert) continue;
          E(o, 4 + vo) = eid;
          for (int s2 = 0; s2 < 4; ++s2) {
            int o2 = NX(o, s2);
            if (!o2) continue;
            for (int v2 = 0; v2 < 4; ++v2)
              if (V(o2, v2) == vert) E(o2, 4 + v2) = eid;
          }
        }
      }
      ++eid;
    }
  }
  t.nedges = nh;
  (void)nside;
  t.colour_of_cell.assign(nc, 0);
  bool ok = false;
  if (cubed) {
    for (int c = 1; c <= nc; ++c) t.colour_of_cell[c - 1] = closed_form_colour(n, c);
    t.ncolours = 4;
    ok = colouring_valid(t);
  } else if (npanels == 1 && planar_quadrilateral(t)) {
    planar_colour(t);
    ok = colouring_valid(t);
  }
  if (!ok) {
    t.ncolours = greedy_colour(t);
    if (t.ncolours == 0 || !colouring_valid(t))
      return fail(LFRIC_E_UNSUPPORTED, "no valid colouring found for a mesh of " + std::to_string(nc) + " cells");
  }
  return LFRIC_OK;
}

// Entity dof counts for the spaces built here (function_space_constructor_helper_functions_mod.F90:
// 258-266 W0, 284-295 W2, 391-400 W3).
namespace {
struct EntityDofs {
  int vert, edge_h, edge_v, face_h, face_v, vol, cell;
};
bool entity_dofs(int space, int k, EntityDofs &d) {
  switch (space) {
    case LFRIC_W0: d = {1, k, k, k * k, k * k, k * k * k, (k + 2) * (k + 2) * (k + 2)}; return true;
    case LFRIC_W2:
      d = {0, 0, 0, (k + 1) * (k + 1), (k + 1) * (k + 1), 3 * k * (k + 1) * (k + 1),
           2 * (k + 2) * (k + 1) * (k + 1) + (k + 1) * (k + 1) * (k + 2)};
      return true;
    case LFRIC_W3: d = {0, 0, 0, 0, 0, (k + 1) * (k + 1) * (k + 1), (k + 1) * (k + 1) * (k + 1)}; return true;
  }
  return false;
}
}  // namespace

int build_dofmap(const Topology &t, int nl, int space, int order, std::vector<int> &dofmap,
                 int &ndf, int &undf) {
  EntityDofs d{};
  if (!entity_dofs(space, order, d) || order < 0 || order > 1 || (space != LFRIC_W0 && order != 0))
    return fail(LFRIC_E_UNSUPPORTED, "function space / element order not supported");
  const int nc = t.ncells, nv2 = t.nverts, ne2 = t.nedges;
  // per-entity column bases (0 = unassigned); top entities are bottom + 1
  std::vector<int> vcol(static_cast<size_t>(2 * nv2 + 1) * std::max(d.vert, 1), 0);
  std::vector<int> ecol(static_cast<size_t>(2 * ne2 + nv2 + 1) * std::max({d.edge_h, d.edge_v, 1}), 0);
  std::vector<int> fcol(static_cast<size_t>(ne2 + 2 * nc + 1) * std::max({d.face_h, d.face_v, 1}), 0);
  std::vector<int> ccol(static_cast<size_t>(nc + 1) * std::max(d.vol, 1), 0);
  const int sv = std::max(d.vert, 1), se = std::max({d.edge_h, d.edge_v, 1});
  const int sf = std::max({d.face_h, d.face_v, 1}), sc = std::max(d.vol, 1);
  auto VC = [&](int e, int i) -> int & { return vcol[static_cast<size_t>(e) * sv + i]; };
  auto EC = [&](int e, int i) -> int & { return ecol[static_cast<size_t>(e) * se + i]; };
  auto FC = [&](int e, int i) -> int & { return fcol[static_cast<size_t>(e) * sf + i]; };
  auto CC = [&](int e, int i) -> int & { return ccol[static_cast<size_t>(e) * sc + i]; };
  auto face = [&](int c, int f) { return t.face_on_cell[(c - 1) * 6 + f]; };
  auto edge = [&](int c, int e) { return t.edge_on_cell[(c - 1) * 12 + e]; };
  auto vert3 = [&](int c, int v) { return v < 4 ? t.vert_on_cell[(c - 1) * 4 + v] : t.vert_on_cell[(c - 1) * 4 + v - 4] + nv2; };

  int next = 1;
  for (int c = 1; c <= nc; ++c) {
    for (int i = 0; i < d.vol; ++i) { CC(c, i) = next; next += nl; }
    for (int f = 0; f < 4; ++f)
      if (!FC(face(c, f), 0))
        for (int i = 0; i < d.face_h; ++i) { FC(face(c, f), i) = next; next += nl; }
    // bottom/top faces share one column of nl+1 dofs per face dof
    for (int i = 0; i < d.face_v; ++i) {
      FC(face(c, 4), i) = next;
      FC(face(c, 5), i) = next + 1;
      next += nl + 1;
    }
    for (int e = 0; e < 4; ++e)
      if (!EC(edge(c, e), 0))
        for (int i = 0; i < d.edge_h; ++i) {
          EC(edge(c, e), i) = next;
          EC(edge(c, e + 8), i) = next + 1;
          next += nl + 1;
        }
    for (int e = 4; e < 8; ++e)
      if (!EC(edge(c, e), 0))
        for (int i = 0; i < d.edge_v; ++i) { EC(edge(c, e), i) = next; next += nl; }
    for (int v = 0; v < 4; ++v)
      if (!VC(vert3(c, v), 0))
        for (int i = 0; i < d.vert; ++i) {
          VC(vert3(c, v), i) = next;
          VC(vert3(c, v + 4), i) = next + 1;
          next += nl + 1;
        }
  }
  undf = next - 1;
  ndf = d.cell;
  dofmap.assign(static_cast<size_t>(nc) * ndf, 0);
  for (int c = 1; c <= nc; ++c) {
    int *row = &dofmap[static_cast<size_t>(c - 1) * ndf];
    int k = 0;
    for (int i = 0; i < d.vol; ++i) row[k++] = CC(c, i);
    for (int f = 0; f < 6; ++f)
      for (int i = 0; i < (f < 4 ? d.face_h : d.face_v); ++i) row[k++] = FC(face(c, f), i);
    for (int e = 0; e < 12; ++e)
      for (int i = 0; i < ((e < 4 || e >= 8) ? d.edge_h : d.edge_v); ++i) row[k++] = EC(edge(c, e), i);
    for (int v = 0; v < 8; ++v)
      for (int i = 0; i < d.vert; ++i) row[k++] = VC(vert3(c, v), i);
    if (k != ndf) return fail(LFRIC_E_STATE, "dofmap entity count mismatch");
  }
  return LFRIC_OK;
}

}  // namespace lfric
