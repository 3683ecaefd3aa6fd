// Host-side mesh tables: owning Topology, the reference mesh reader's table
// conventions, a Gmsh 4.1 reader, StructTriangMesh and a seeded jittered
// Delaunay generator with Morton cell order.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/swe_fvm.h"
#include "swe_internal.h"

namespace swe {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

// Table builder: boundary edges first in line order (ep0 < ep1, et = (t, -1)),
// the rest by first appearance over cells, k = 0..2 of (tp[k], tp[k+1]);
// te[k] = that edge; interior et = (later cell, earlier cell); tt[k] the
// neighbour across te[k] or -1.  These are the conventions of the
// reference's mesh reader as recorded in notebooks/topology.dat.
int build_topology(int64_t V, int64_t T, const int64_t* tri, int64_t nb, const int64_t* bnd,
                   Topo& m) {
  m.V = V;
  m.T = T;
  m.ep.clear();
  m.et.clear();
  std::unordered_map<uint64_t, int64_t> map;
  map.reserve(size_t(2 * (T + nb) + 16));
  auto key = [](int64_t a, int64_t b) {
    int64_t lo = std::min(a, b), hi = std::max(a, b);
    return (uint64_t(lo) << 32) | uint64_t(hi);
  };
  for (int64_t k = 0; k < nb; ++k) {
    int64_t a = bnd[2 * k], b = bnd[2 * k + 1];
    if (a < 0 || b < 0 || a >= V || b >= V || a == b)
      return fail(SWE_ERR_ARG, "boundary line with invalid node index");
    uint64_t kk = key(a, b);
    if (map.count(kk)) continue;
    map.emplace(kk, int64_t(m.ep.size() / 2));
    m.ep.push_back(std::min(a, b));
    m.ep.push_back(std::max(a, b));
    m.et.push_back(-1);
    m.et.push_back(-1);
  }
  m.tp.assign(tri, tri + 3 * T);
  m.te.assign(size_t(3 * T), -1);
  m.tt.assign(size_t(3 * T), -1);
  for (int64_t t = 0; t < T; ++t) {
    for (int k = 0; k < 3; ++k) {
      int64_t a = tri[3 * t + k], b = tri[3 * t + (k + 1) % 3];
      if (a < 0 || b < 0 || a >= V || b >= V || a == b)
        return fail(SWE_ERR_ARG, "triangle with invalid node index");
      uint64_t kk = key(a, b);
      auto it = map.find(kk);
      int64_t e;
      if (it == map.end()) {
        e = int64_t(m.ep.size() / 2);
        map.emplace(kk, e);
        m.ep.push_back(std::min(a, b));
        m.ep.push_back(std::max(a, b));
        m.et.push_back(-1);
        m.et.push_back(-1);
      } else {
        e = it->second;
      }
      m.te[3 * t + k] = e;
      if (m.et[2 * e] == -1) {
        m.et[2 * e] = t;
      } else {
        if (m.et[2 * e + 1] != -1) return fail(SWE_ERR_DOMAIN, "edge shared by more than two triangles");
        m.et[2 * e + 1] = m.et[2 * e];
        m.et[2 * e] = t;
      }
    }
  }
  m.E = int64_t(m.ep.size() / 2);
  for (int64_t e = 0; e < m.E; ++e)  // a boundary line is an edge of some triangle
    if (m.et[2 * e] < 0) return fail(SWE_ERR_DOMAIN, "boundary line that is not an edge of any triangle");
  for (int64_t t = 0; t < T; ++t)
    for (int k = 0; k < 3; ++k) {
      int64_t e = m.te[3 * t + k];
      int64_t a = m.et[2 * e], b = m.et[2 * e + 1];
      m.tt[3 * t + k] = (a == t) ? b : a;
    }
  return SWE_OK;
}

// ------------------------------------------------------------- generators --

static uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static double uniform01(uint64_t& s) { return double(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

// Positive when d lies inside the circumcircle of the CCW triangle abc.
static double incircle(const double* a, const double* b, const double* c, const double* d) {
  double adx = a[0] - d[0], ady = a[1] - d[1];
  double bdx = b[0] - d[0], bdy = b[1] - d[1];
  double cdx = c[0] - d[0], cdy = c[1] - d[1];
  double ad = adx * adx + ady * ady, bd = bdx * bdx + bdy * bdy, cd = cdx * cdx + cdy * cdy;
  return ad * (bdx * cdy - cdx * bdy) + bd * (cdx * ady - adx * cdy) + cd * (adx * bdy - bdx * ady);
}
static double orient(const double* a, const double* b, const double* c) {
  return (b[0] - a[0]) * (c[1] - a[1]) - (c[0] - a[0]) * (b[1] - a[1]);
}

static uint64_t morton2(uint32_t x, uint32_t y) {
  uint64_t key = 0;
  for (int k = 0; k < 21; ++k) {
    key |= uint64_t((x >> k) & 1u) << (2 * k);
    key |= uint64_t((y >> k) & 1u) << (2 * k + 1);
  }
  return key;
}

// Reorders cells along the Morton curve of their centroids and renumbers
// nodes by first appearance; boundary lines are remapped.
void morton_reorder(std::vector<double>& xyz, std::vector<int64_t>& tri, std::vector<int64_t>& bnd) {
  int64_t T = int64_t(tri.size() / 3), V = int64_t(xyz.size() / 3);
  double xmin = 1e300, xmax = -1e300, ymin = 1e300, ymax = -1e300;
  for (int64_t n = 0; n < V; ++n) {
    xmin = std::min(xmin, xyz[3 * n]); xmax = std::max(xmax, xyz[3 * n]);
    ymin = std::min(ymin, xyz[3 * n + 1]); ymax = std::max(ymax, xyz[3 * n + 1]);
  }
  double sx = (xmax > xmin) ? 2097151.0 / (xmax - xmin) : 0.0;
  double sy = (ymax > ymin) ? 2097151.0 / (ymax - ymin) : 0.0;
  std::vector<std::pair<uint64_t, int64_t>> key(static_cast<size_t>(T));
  for (int64_t t = 0; t < T; ++t) {
    double cx = 0, cy = 0;
    for (int k = 0; k < 3; ++k) { cx += xyz[3 * tri[3 * t + k]]; cy += xyz[3 * tri[3 * t + k] + 1]; }
    cx /= 3.0; cy /= 3.0;
    uint32_t qx = uint32_t(std::min(2097151.0, std::max(0.0, (cx - xmin) * sx)));
    uint32_t qy = uint32_t(std::min(2097151.0, std::max(0.0, (cy - ymin) * sy)));
    key[size_t(t)] = {morton2(qx, qy), t};
  }
  std::sort(key.begin(), key.end());
  std::vector<int64_t> ntri(size_t(3 * T));
  std::vector<int64_t> nid(size_t(V), -1);
  std::vector<double> nxyz(size_t(3 * V));
  int64_t next = 0;
  for (int64_t t = 0; t < T; ++t) {
    int64_t o = key[size_t(t)].second;
    for (int k = 0; k < 3; ++k) {
      int64_t n = tri[3 * o + k];
      if (nid[size_t(n)] < 0) {
        nid[size_t(n)] = next;
        for (int d = 0; d < 3; ++d) nxyz[size_t(3 * next + d)] = xyz[size_t(3 * n + d)];
        ++next;
      }
      ntri[size_t(3 * t + k)] = nid[size_t(n)];
    }
  }
  for (int64_t n = 0; n < V; ++n)  // unreferenced nodes keep a slot at the end
    if (nid[size_t(n)] < 0) {
      nid[size_t(n)] = next;
      for (int d = 0; d < 3; ++d) nxyz[size_t(3 * next + d)] = xyz[size_t(3 * n + d)];
      ++next;
    }
  for (auto& b : bnd) b = nid[size_t(b)];
  tri.swap(ntri);
  xyz.swap(nxyz);
}

// Jittered-grid Delaunay: per-square diagonal by the empty-circle test, then
// Lawson flips until every interior edge is locally Delaunay.
void jittered_delaunay(int64_t n, double jitter, uint64_t seed, double x0, double y0, double L,
                       std::vector<double>& xyz, std::vector<int64_t>& tri, std::vector<int64_t>& bnd) {
  int64_t nv = (n + 1) * (n + 1);
  double h = L / double(n);
  xyz.assign(size_t(3 * nv), 0.0);
  uint64_t s = seed;
  for (int64_t j = 0; j <= n; ++j)
    for (int64_t i = 0; i <= n; ++i) {
      int64_t id = j * (n + 1) + i;
      double dx = 0, dy = 0;
      double r1 = uniform01(s), r2 = uniform01(s);
      if (i > 0 && i < n && j > 0 && j < n) {
        dx = (2.0 * r1 - 1.0) * jitter * h;
        dy = (2.0 * r2 - 1.0) * jitter * h;
      }
      xyz[size_t(3 * id)] = x0 + double(i) * h + dx;
      xyz[size_t(3 * id + 1)] = y0 + double(j) * h + dy;
    }
  auto P = [&](int64_t id) { return &xyz[size_t(3 * id)]; };
  int64_t T = 2 * n * n;
  tri.assign(size_t(3 * T), 0);
  for (int64_t j = 0; j < n; ++j)
    for (int64_t i = 0; i < n; ++i) {
      int64_t a = j * (n + 1) + i, b = a + 1, c = a + n + 2, d = a + n + 1;  // CCW quad
      int64_t t = 2 * (j * n + i);
      // diagonal a-c unless d lies inside circle(a, b, c)
      bool ac = !(incircle(P(a), P(b), P(c), P(d)) > 0.0);
      if (ac) {
        int64_t v[6] = {a, b, c, a, c, d};
        std::copy(v, v + 6, &tri[size_t(3 * t)]);
      } else {
        int64_t v[6] = {a, b, d, b, c, d};
        std::copy(v, v + 6, &tri[size_t(3 * t)]);
      }
    }
  // neighbour table: nb[3t+k] = triangle across edge opposite vertex k
  std::vector<int64_t> nb(size_t(3 * T), -1);
  {
    std::unordered_map<uint64_t, int64_t> half;
    half.reserve(size_t(4 * T));
    for (int64_t t = 0; t < T; ++t)
      for (int k = 0; k < 3; ++k) {
        int64_t a = tri[size_t(3 * t + (k + 1) % 3)], b = tri[size_t(3 * t + (k + 2) % 3)];
        uint64_t key = (uint64_t(std::min(a, b)) << 32) | uint64_t(std::max(a, b));
        auto it = half.find(key);
        if (it == half.end()) {
          half.emplace(key, 3 * t + k);
        } else {
          int64_t o = it->second;
          nb[size_t(3 * t + k)] = o / 3;
          nb[size_t(o)] = t;
        }
      }
  }
  std::vector<int64_t> stack;
  stack.reserve(size_t(3 * T));
  for (int64_t t = 0; t < T; ++t)
    for (int k = 0; k < 3; ++k) stack.push_back(3 * t + k);
  auto opp_index = [&](int64_t u, int64_t t) {
    for (int k = 0; k < 3; ++k)
      if (nb[size_t(3 * u + k)] == t) return k;
    return -1;
  };
  int64_t guard = 0;
  while (!stack.empty() && guard < 64 * T) {
    ++guard;
    int64_t he = stack.back();
    stack.pop_back();
    int64_t t = he / 3;
    int k = int(he % 3);
    int64_t u = nb[size_t(he)];
    if (u < 0) continue;
    int ku = opp_index(u, t);
    if (ku < 0) continue;
    int64_t a = tri[size_t(3 * t + k)];
    int64_t b = tri[size_t(3 * t + (k + 1) % 3)];
    int64_t c = tri[size_t(3 * t + (k + 2) % 3)];
    int64_t d = tri[size_t(3 * u + ku)];
    double scale = h * h * h * h;
    if (!(incircle(P(a), P(b), P(c), P(d)) > 1e-12 * scale)) continue;
    // the flipped pair must stay convex
    if (!(orient(P(a), P(b), P(d)) > 0.0) || !(orient(P(a), P(d), P(c)) > 0.0)) continue;
    // neighbours of the quad a-b-d-c (t = a b c, u = c b d seen from u)
    int64_t n_ab = nb[size_t(3 * t + (k + 2) % 3)];   // across edge a-b (opposite c)
    int64_t n_ca = nb[size_t(3 * t + (k + 1) % 3)];   // across edge c-a (opposite b)
    // in u, vertices around: d = tri[u][ku], then ku+1, ku+2 are c, b in CCW order
    int64_t n_bd = -1, n_dc = -1;
    for (int q = 0; q < 3; ++q) {
      int64_t vq = tri[size_t(3 * u + q)];
      if (vq == c) n_bd = nb[size_t(3 * u + q)];  // edge opposite c in u is b-d
      if (vq == b) n_dc = nb[size_t(3 * u + q)];  // edge opposite b in u is d-c
    }
    // new t = (a, b, d), new u = (a, d, c)
    tri[size_t(3 * t)] = a; tri[size_t(3 * t + 1)] = b; tri[size_t(3 * t + 2)] = d;
    tri[size_t(3 * u)] = a; tri[size_t(3 * u + 1)] = d; tri[size_t(3 * u + 2)] = c;
    nb[size_t(3 * t)] = n_bd;      // opposite a: b-d
    nb[size_t(3 * t + 1)] = u;     // opposite b: d-a
    nb[size_t(3 * t + 2)] = n_ab;  // opposite d: a-b
    nb[size_t(3 * u)] = n_dc;      // opposite a: d-c
    nb[size_t(3 * u + 1)] = n_ca;  // opposite d: c-a
    nb[size_t(3 * u + 2)] = t;     // opposite c: a-d
    auto relink = [&](int64_t w, int64_t from, int64_t to) {
      if (w < 0) return;
      for (int q = 0; q < 3; ++q)
        if (nb[size_t(3 * w + q)] == from) { nb[size_t(3 * w + q)] = to; return; }
    };
    relink(n_bd, u, t);
    relink(n_ca, t, u);
    stack.push_back(3 * t);
    stack.push_back(3 * t + 2);
    stack.push_back(3 * u);
    stack.push_back(3 * u + 1);
  }
  // walls: bottom, right, top, left, counter-clockwise
  bnd.clear();
  for (int64_t i = 0; i < n; ++i) { bnd.push_back(i); bnd.push_back(i + 1); }
  for (int64_t j = 0; j < n; ++j) { bnd.push_back(j * (n + 1) + n); bnd.push_back((j + 1) * (n + 1) + n); }
  for (int64_t i = n; i > 0; --i) { bnd.push_back(n * (n + 1) + i); bnd.push_back(n * (n + 1) + i - 1); }
  for (int64_t j = n; j > 0; --j) { bnd.push_back(j * (n + 1)); bnd.push_back((j - 1) * (n + 1)); }
}

void struct_mesh(int64_t ni, int64_t nj, double h, double x0, double y0, std::vector<double>& xyz,
                 std::vector<int64_t>& tri, std::vector<int64_t>& bnd) {
  int64_t ng = (ni + 1) * (nj + 1);
  int64_t nv = ng + ni * nj;
  xyz.assign(size_t(3 * nv), 0.0);
  auto g = [&](int64_t i, int64_t j) { return j * (ni + 1) + i; };
  for (int64_t j = 0; j <= nj; ++j)
    for (int64_t i = 0; i <= ni; ++i) {
      xyz[size_t(3 * g(i, j))] = x0 + double(i) * h;
      xyz[size_t(3 * g(i, j) + 1)] = y0 + double(j) * h;
    }
  for (int64_t j = 0; j < nj; ++j)
    for (int64_t i = 0; i < ni; ++i) {
      int64_t c = ng + j * ni + i;
      xyz[size_t(3 * c)] = x0 + (double(i) + 0.5) * h;
      xyz[size_t(3 * c + 1)] = y0 + (double(j) + 0.5) * h;
    }
  tri.clear();
  tri.reserve(size_t(12 * ni * nj));
  for (int64_t j = 0; j < nj; ++j)
    for (int64_t i = 0; i < ni; ++i) {
      int64_t c = ng + j * ni + i;
      int64_t v[12] = {g(i, j),         g(i + 1, j),     c,   // bottom
                       g(i + 1, j),     g(i + 1, j + 1), c,   // right
                       g(i + 1, j + 1), g(i, j + 1),     c,   // top
                       g(i, j + 1),     g(i, j),         c};  // left
      tri.insert(tri.end(), v, v + 12);
    }
  bnd.clear();
  for (int64_t i = 0; i < ni; ++i) { bnd.push_back(g(i, 0)); bnd.push_back(g(i + 1, 0)); }
  for (int64_t j = 0; j < nj; ++j) { bnd.push_back(g(ni, j)); bnd.push_back(g(ni, j + 1)); }
  for (int64_t i = ni; i > 0; --i) { bnd.push_back(g(i, nj)); bnd.push_back(g(i - 1, nj)); }
  for (int64_t j = nj; j > 0; --j) { bnd.push_back(g(0, j)); bnd.push_back(g(0, j - 1)); }
}

// Gmsh 4.1 ASCII: $Nodes blocks (tags then coordinates), $Elements blocks
// (type 1 lines, type 2 triangles).  Node index = rank of its tag.  Every
// header and block count is checked against the section's totals and each
// section must end with its $End marker, so a truncated or malformed file is
// an error, not a partial mesh.
static bool end_marker(std::istream& in, const char* want) {
  std::string s;
  in >> s;
  return bool(in) && s == want;
}
static int read_msh_sections(std::istream& in, std::map<int64_t, std::array<double, 3>>& nodes,
                             std::vector<std::pair<int64_t, int64_t>>& lines,
                             std::vector<std::array<int64_t, 3>>& tris) {
  std::string line;
  bool fmt_ok = false, have_nodes = false, have_elements = false;
  while (std::getline(in, line)) {
    if (line.rfind("$MeshFormat", 0) == 0) {
      std::getline(in, line);
      double ver = std::atof(line.c_str());
      if (ver < 4.0 || ver >= 5.0) return fail(SWE_ERR_ARG, "only Gmsh 4.x ASCII meshes are read");
      if (line.find(" 1 ") != std::string::npos && line.find(" 0 ") == std::string::npos)
        return fail(SWE_ERR_ARG, "binary Gmsh files are not read");
      fmt_ok = true;
    } else if (line.rfind("$Nodes", 0) == 0) {
      int64_t nblocks = -1, nn = -1, mn = 0, mx = 0;
      in >> nblocks >> nn >> mn >> mx;
      if (!in || nblocks < 0 || nn < 0) return fail(SWE_ERR_ARG, "malformed $Nodes header");
      int64_t got = 0;
      for (int64_t b = 0; b < nblocks; ++b) {
        int64_t ed = -1, et = 0, par = -1, cnt = -1;
        in >> ed >> et >> par >> cnt;
        if (!in || ed < 0 || ed > 3 || par < 0 || cnt < 0 || cnt > nn - got)
          return fail(SWE_ERR_ARG, "malformed $Nodes block");
        std::vector<int64_t> tags(static_cast<size_t>(cnt));
        for (auto& t : tags) in >> t;
        for (int64_t k = 0; k < cnt; ++k) {
          std::array<double, 3> p;
          in >> p[0] >> p[1] >> p[2];
          if (par) {  // skip parametric coordinates
            double skip;
            for (int64_t q = 0; q < ed; ++q) in >> skip;
          }
          nodes[tags[size_t(k)]] = p;
        }
        if (!in) return fail(SWE_ERR_ARG, "truncated $Nodes section");
        got += cnt;
      }
      if (got != nn) return fail(SWE_ERR_ARG, "$Nodes blocks do not add up to the declared node count");
      if (!end_marker(in, "$EndNodes")) return fail(SWE_ERR_ARG, "truncated $Nodes section (no $EndNodes)");
      have_nodes = true;
    } else if (line.rfind("$Elements", 0) == 0) {
      int64_t nblocks = -1, ne = -1, mn = 0, mx = 0;
      in >> nblocks >> ne >> mn >> mx;
      if (!in || nblocks < 0 || ne < 0) return fail(SWE_ERR_ARG, "malformed $Elements header");
      int64_t got = 0;
      for (int64_t b = 0; b < nblocks; ++b) {
        int64_t ed = -1, et = 0, typ = -1, cnt = -1;
        in >> ed >> et >> typ >> cnt;
        if (!in || ed < 0 || ed > 3 || cnt < 0 || cnt > ne - got)
          return fail(SWE_ERR_ARG, "malformed $Elements block");
        int nper = (typ == 1) ? 2 : (typ == 2) ? 3 : (typ == 15) ? 1 : (typ == 3) ? 4 : (typ == 4) ? 4 : -1;
        if (nper < 0) return fail(SWE_ERR_ARG, "unsupported Gmsh element type");
        for (int64_t k = 0; k < cnt; ++k) {
          int64_t tag, v[4];
          in >> tag;
          for (int q = 0; q < nper; ++q) in >> v[q];
          if (typ == 1) lines.push_back({v[0], v[1]});
          else if (typ == 2) tris.push_back({v[0], v[1], v[2]});
        }
        if (!in) return fail(SWE_ERR_ARG, "truncated $Elements section");
        got += cnt;
      }
      if (got != ne) return fail(SWE_ERR_ARG, "$Elements blocks do not add up to the declared element count");
      if (!end_marker(in, "$EndElements"))
        return fail(SWE_ERR_ARG, "truncated $Elements section (no $EndElements)");
      have_elements = true;
    }
    if (!in) break;
  }
  if (!fmt_ok) return fail(SWE_ERR_ARG, "missing $MeshFormat");
  if (!have_nodes || !have_elements) return fail(SWE_ERR_ARG, "missing $Nodes or $Elements section");
  return SWE_OK;
}

int read_msh(const char* path, std::vector<double>& xyz, std::vector<int64_t>& tri,
             std::vector<int64_t>& bnd) {
  std::ifstream in(path);
  if (!in) return fail(SWE_ERR_ARG, std::string("cannot open ") + path);
  std::map<int64_t, std::array<double, 3>> nodes;
  std::vector<std::pair<int64_t, int64_t>> lines;
  std::vector<std::array<int64_t, 3>> tris;
  try {
    if (int rc = read_msh_sections(in, nodes, lines, tris)) return rc;
  } catch (const std::exception& e) {  // e.g. an allocation for counts no memory holds
    return fail(SWE_ERR_ARG, std::string("malformed Gmsh file: ") + e.what());
  }
  std::unordered_map<int64_t, int64_t> idx;
  xyz.clear();
  int64_t k = 0;
  for (auto& kv : nodes) {
    idx[kv.first] = k++;
    xyz.insert(xyz.end(), kv.second.begin(), kv.second.end());
  }
  tri.clear();
  bnd.clear();
  for (auto& t : tris)
    for (int q = 0; q < 3; ++q) {
      auto it = idx.find(t[size_t(q)]);
      if (it == idx.end()) return fail(SWE_ERR_ARG, "triangle references an unknown node");
      tri.push_back(it->second);
    }
  for (auto& l : lines) {
    auto a = idx.find(l.first), b = idx.find(l.second);
    if (a == idx.end() || b == idx.end()) return fail(SWE_ERR_ARG, "line references an unknown node");
    bnd.push_back(a->second);
    bnd.push_back(b->second);
  }
  return SWE_OK;
}

}  // namespace swe

using namespace swe;

extern "C" {

const char* swe_last_error(void) { return g_last_error.c_str(); }

void swe_free(void* p) { std::free(p); }

int swe_topology_create(int64_t num_nodes, int64_t num_edges, const int64_t* edge_nodes,
                        const int64_t* edge_elements, int64_t num_elements,
                        const int64_t* element_nodes, const int64_t* element_edges,
                        const int64_t* element_neighbours, swe_topology** out) {
  if (!out || num_nodes < 0 || num_edges < 0 || num_elements < 0)
    return fail(SWE_ERR_ARG, "invalid sizes");
  if ((num_edges && (!edge_nodes || !edge_elements)) ||
      (num_elements && (!element_nodes || !element_edges || !element_neighbours)))
    return fail(SWE_ERR_ARG, "null table");
  // indices in range; negative element ids are boundary tags (only
  // SOLID_WALL = -1 is accepted later by swe_create)
  for (int64_t k = 0; k < 2 * num_edges; ++k) {
    if (edge_nodes[k] < 0 || edge_nodes[k] >= num_nodes) return fail(SWE_ERR_DOMAIN, "edge node out of range");
    if (edge_elements[k] >= num_elements || (k % 2 == 0 && edge_elements[k] < 0))
      return fail(SWE_ERR_DOMAIN, "edge element out of range");
  }
  for (int64_t k = 0; k < 3 * num_elements; ++k) {
    if (element_nodes[k] < 0 || element_nodes[k] >= num_nodes) return fail(SWE_ERR_DOMAIN, "element node out of range");
    if (element_edges[k] < 0 || element_edges[k] >= num_edges) return fail(SWE_ERR_DOMAIN, "element edge out of range");
    if (element_neighbours[k] >= num_elements) return fail(SWE_ERR_DOMAIN, "element neighbour out of range");
  }
  auto* t = new swe_topology;
  Topo& m = t->m;
  m.V = num_nodes; m.E = num_edges; m.T = num_elements;
  m.ep.assign(edge_nodes, edge_nodes + 2 * num_edges);
  m.et.assign(edge_elements, edge_elements + 2 * num_edges);
  m.tp.assign(element_nodes, element_nodes + 3 * num_elements);
  m.te.assign(element_edges, element_edges + 3 * num_elements);
  m.tt.assign(element_neighbours, element_neighbours + 3 * num_elements);
  *out = t;
  return SWE_OK;
}

int swe_topology_build(int64_t num_nodes, int64_t num_triangles, const int64_t* triangle_nodes,
                       int64_t num_boundary_lines, const int64_t* boundary_lines,
                       swe_topology** out) {
  if (!out || num_nodes < 0 || num_triangles < 0 || num_boundary_lines < 0)
    return fail(SWE_ERR_ARG, "invalid sizes");
  auto* t = new swe_topology;
  int rc = build_topology(num_nodes, num_triangles, triangle_nodes, num_boundary_lines,
                          boundary_lines, t->m);
  if (rc != SWE_OK) { delete t; return rc; }
  *out = t;
  return SWE_OK;
}

static int emit_mesh(std::vector<double>& xyz, std::vector<int64_t>& tri, std::vector<int64_t>& bnd,
                     double** pxyz, int64_t* nn, swe_topology** out) {
  int64_t V = int64_t(xyz.size() / 3);
  auto* t = new swe_topology;
  int rc = build_topology(V, int64_t(tri.size() / 3), tri.data(), int64_t(bnd.size() / 2), bnd.data(), t->m);
  if (rc != SWE_OK) { delete t; return rc; }
  double* buf = (double*)std::malloc(sizeof(double) * xyz.size());
  if (!buf) { delete t; return fail(SWE_ERR_ARG, "out of host memory"); }
  std::memcpy(buf, xyz.data(), sizeof(double) * xyz.size());
  *pxyz = buf;
  *nn = V;
  *out = t;
  return SWE_OK;
}

int swe_topology_read_msh(const char* path, double** xyz, int64_t* num_nodes, swe_topology** out) {
  if (!path || !xyz || !num_nodes || !out) return fail(SWE_ERR_ARG, "null argument");
  std::vector<double> p;
  std::vector<int64_t> tri, bnd;
  int rc = read_msh(path, p, tri, bnd);
  if (rc != SWE_OK) return rc;
  return emit_mesh(p, tri, bnd, xyz, num_nodes, out);
}

int swe_mesh_struct(int64_t ni, int64_t nj, double h, double x0, double y0, double** xyz,
                    int64_t* num_nodes, swe_topology** out) {
  if (ni < 1 || nj < 1 || !(h > 0) || !xyz || !num_nodes || !out)
    return fail(SWE_ERR_ARG, "invalid structured mesh arguments");
  std::vector<double> p;
  std::vector<int64_t> tri, bnd;
  struct_mesh(ni, nj, h, x0, y0, p, tri, bnd);
  return emit_mesh(p, tri, bnd, xyz, num_nodes, out);
}

int swe_mesh_jittered(int64_t n, double jitter, uint64_t seed, double x0, double y0, double L,
                      double** xyz, int64_t* num_nodes, swe_topology** out) {
  if (n < 1 || jitter < 0 || jitter > 0.3 || !(L > 0) || !xyz || !num_nodes || !out)
    return fail(SWE_ERR_ARG, "invalid jittered mesh arguments (jitter in [0, 0.3])");
  if (2 * n * n > 0x7fffffffll) return fail(SWE_ERR_ARG, "mesh too large for int32 indices");
  std::vector<double> p;
  std::vector<int64_t> tri, bnd;
  jittered_delaunay(n, jitter, seed, x0, y0, L, p, tri, bnd);
  morton_reorder(p, tri, bnd);
  return emit_mesh(p, tri, bnd, xyz, num_nodes, out);
}

int64_t swe_topology_num_nodes(const swe_topology* t) { return t ? t->m.V : -1; }
int64_t swe_topology_num_edges(const swe_topology* t) { return t ? t->m.E : -1; }
int64_t swe_topology_num_elements(const swe_topology* t) { return t ? t->m.T : -1; }

int swe_topology_copy(const swe_topology* t, int64_t* edge_nodes, int64_t* edge_elements,
                      int64_t* element_nodes, int64_t* element_edges, int64_t* element_neighbours) {
  if (!t) return fail(SWE_ERR_ARG, "null topology");
  const Topo& m = t->m;
  if (edge_nodes) std::memcpy(edge_nodes, m.ep.data(), sizeof(int64_t) * m.ep.size());
  if (edge_elements) std::memcpy(edge_elements, m.et.data(), sizeof(int64_t) * m.et.size());
  if (element_nodes) std::memcpy(element_nodes, m.tp.data(), sizeof(int64_t) * m.tp.size());
  if (element_edges) std::memcpy(element_edges, m.te.data(), sizeof(int64_t) * m.te.size());
  if (element_neighbours) std::memcpy(element_neighbours, m.tt.data(), sizeof(int64_t) * m.tt.size());
  return SWE_OK;
}

// Pairs boundary edges a_k <-> b_k as PERIODIC (tag -3, include/Includes.h:21):
// the outside state of each is the inside state of its partner.
int swe_topology_set_periodic(swe_topology* t, int64_t n, const int64_t* pairs) {
  if (!t) return fail(SWE_ERR_ARG, "null topology");
  if (n < 0 || (n > 0 && !pairs)) return fail(SWE_ERR_ARG, "invalid periodic pairs");
  Topo& m = t->m;
  std::vector<int64_t> pe = m.pe.empty() ? std::vector<int64_t>(size_t(m.E), -1) : m.pe;
  std::vector<int64_t> et = m.et;
  for (int64_t k = 0; k < n; ++k) {
    const int64_t a = pairs[2 * k], b = pairs[2 * k + 1];
    if (a < 0 || a >= m.E || b < 0 || b >= m.E || a == b) return fail(SWE_ERR_ARG, "periodic pair out of range");
    if (et[size_t(2 * a + 1)] >= 0 || et[size_t(2 * b + 1)] >= 0)
      return fail(SWE_ERR_DOMAIN, "only boundary edges can be periodic");
    if (pe[size_t(a)] >= 0 || pe[size_t(b)] >= 0) return fail(SWE_ERR_DOMAIN, "edge paired twice");
    pe[size_t(a)] = b;
    pe[size_t(b)] = a;
    et[size_t(2 * a + 1)] = -3;
    et[size_t(2 * b + 1)] = -3;
  }
  m.pe.swap(pe);
  m.et.swap(et);
  return SWE_OK;
}

int swe_topology_periodic(const swe_topology* t, int64_t* partner) {
  if (!t || !partner) return fail(SWE_ERR_ARG, "null argument");
  const Topo& m = t->m;
  for (int64_t e = 0; e < m.E; ++e) partner[e] = m.pe.empty() ? -1 : m.pe[size_t(e)];
  return SWE_OK;
}

void swe_topology_destroy(swe_topology* t) { delete t; }

}  // extern "C"
