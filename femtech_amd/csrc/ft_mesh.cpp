// ft_mesh.cpp -- host-side mesh tools: structured block generator, Abaqus
// .inp writer, .inp / .k readers with the one-rank node numbering of
// PartitionMesh.cpp:485-536, and the materials.dat reader.
#include <ctype.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>

#include <algorithm>
#include <random>
#include <string>
#include <vector>

#include "../../include/femtech_amd.h"

extern "C" void ft_internal_set_error(const char *msg);

namespace {

int mfail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  ft_internal_set_error(buf);
  return code;
}

std::vector<std::string> split_tokens(const char *line, const char *delim) {
  std::vector<std::string> out;
  std::string cur;
  for (const char *p = line; *p; ++p) {
    if (strchr(delim, *p) || *p == '\n' || *p == '\r') {
      if (!cur.empty()) out.push_back(cur);
      cur.clear();
    } else {
      cur.push_back(*p);
    }
  }
  if (!cur.empty()) out.push_back(cur);
  return out;
}

// Numeric tokens of a data line; false if a token is not a number
// (LineToArray, ReadInputFile.cpp:84-174, stops at the first non-number).
bool numbers(const char *line, const char *delim, std::vector<double> &v) {
  v.clear();
  for (auto &t : split_tokens(line, delim)) {
    char *end = nullptr;
    double x = strtod(t.c_str(), &end);
    if (end == t.c_str()) return false;
    v.push_back(x);
  }
  return !v.empty();
}

std::string upper(std::string s) {
  for (auto &c : s) c = (char)toupper((unsigned char)c);
  return s;
}

struct RawMesh {
  std::vector<long> node_id;
  std::vector<double> node_xyz;
  std::vector<long> elem_nodes;  // 8 global ids per element
  std::vector<int> pid;
};

int read_inp(FILE *f, RawMesh &m) {
  char line[4096];
  enum { NONE, NODE, ELEM } sec = NONE;
  std::vector<std::string> elsets;
  int cur_pid = 0;
  bool node_done = false;
  std::vector<double> v;
  while (fgets(line, sizeof(line), f)) {
    if (line[0] == '*') {
      if (line[1] == '*') continue;  // comment
      std::string hdr = upper(line);
      if (sec == NODE) node_done = true;
      sec = NONE;
      auto toks = split_tokens(hdr.c_str(), ", \t=");
      if (toks.empty()) continue;
      // The keyword is the whole word before the first comma: "*Node" and
      // "*Element, type=..." open sections, "*Node Output" and "*Element
      // Output" do not (ReadAbaqus.cpp:359-412).
      const std::string kw = toks[0];
      const size_t after = hdr.find_first_not_of(" \t", kw.size());
      const bool bare = after == std::string::npos || hdr[after] == '\n' || hdr[after] == '\r';
      const bool comma = after != std::string::npos && hdr[after] == ',';
      if (kw == "*NODE" && (bare || comma)) {
        // only the first *Node section is read (ReadAbaqus.cpp:41-49)
        if (!node_done) sec = NODE;
      } else if (kw == "*ELEMENT" && comma) {
        std::string type, elset;
        for (size_t i = 1; i + 1 < toks.size(); ++i) {
          if (toks[i] == "TYPE") type = toks[i + 1];
          if (toks[i] == "ELSET") elset = toks[i + 1];
        }
        if (type.empty()) continue;  // not an element section without type=
        if (type != "C3D8")
          return mfail(FT_ERR_INPUT, "element type %s not supported (C3D8 only)", type.c_str());
        // pid = order of first appearance of the ELSET name (ReadAbaqus.cpp:172-179)
        auto it = std::find(elsets.begin(), elsets.end(), elset);
        if (it == elsets.end()) {
          elsets.push_back(elset);
          cur_pid = (int)elsets.size() - 1;
        } else {
          cur_pid = (int)(it - elsets.begin());
        }
        sec = ELEM;
      }
      continue;
    }
    if (sec == NODE) {
      if (!numbers(line, ", \t", v)) continue;
      if (v.size() < 4) return mfail(FT_ERR_INPUT, "bad node line: %s", line);
      m.node_id.push_back((long)v[0]);
      m.node_xyz.insert(m.node_xyz.end(), v.begin() + 1, v.begin() + 4);
    } else if (sec == ELEM) {
      if (!numbers(line, ", \t", v)) continue;
      if (v.size() != 9) return mfail(FT_ERR_INPUT, "C3D8 line needs 9 integers: %s", line);
      for (int k = 1; k < 9; ++k) m.elem_nodes.push_back((long)v[k]);
      m.pid.push_back(cur_pid);
    }
  }
  return 0;
}

int read_k(FILE *f, RawMesh &m) {
  char line[4096];
  enum { NONE, NODE, ELEM } sec = NONE;
  std::vector<double> v;
  while (fgets(line, sizeof(line), f)) {
    if (line[0] == '$') continue;
    if (line[0] == '*') {
      std::string hdr = upper(line);
      auto toks = split_tokens(hdr.c_str(), ", \t");
      sec = NONE;
      if (!toks.empty() && toks[0] == "*NODE") sec = NODE;
      if (!toks.empty() && toks[0] == "*ELEMENT_SOLID") sec = ELEM;
      continue;
    }
    if (!numbers(line, ", \t", v)) continue;
    if (sec == NODE) {
      if (v.size() < 4) return mfail(FT_ERR_INPUT, "bad *NODE line: %s", line);
      m.node_id.push_back((long)v[0]);
      m.node_xyz.insert(m.node_xyz.end(), v.begin() + 1, v.begin() + 4);
    } else if (sec == ELEM) {
      // eid pid n1..n8 (ReadLsDyna.cpp:205-245): pid = PART - 1
      if (v.size() < 10) return mfail(FT_ERR_INPUT, "*ELEMENT_SOLID needs 8 nodes: %s", line);
      m.pid.push_back((int)v[1] - 1);
      for (int k = 2; k < 10; ++k) m.elem_nodes.push_back((long)v[k]);
    }
  }
  return 0;
}

}  // namespace

extern "C" {

int ft_mesh_block(int nx, int ny, int nz, double hx, double hy, double hz, double jitter,
                  unsigned long long seed, double *coords, int *conn) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || !coords || !conn) return mfail(FT_ERR_ARG, "bad block arguments");
  const long sx = nx + 1, sy = ny + 1, sz = nz + 1;
  const long nn = sx * sy * sz;
  if (nn > 2000000000L / 3) return mfail(FT_ERR_ARG, "block too large");
  std::mt19937_64 rng(seed);
  const double inv53 = 1.0 / 9007199254740992.0;
  for (long k = 0; k < sz; ++k)
    for (long j = 0; j < sy; ++j)
      for (long i = 0; i < sx; ++i) {
        const long n = i + sx * (j + sy * k);
        double x = i * hx, y = j * hy, z = k * hz;
        const bool interior = i > 0 && i < nx && j > 0 && j < ny && k > 0 && k < nz;
        if (interior && jitter != 0.0) {
          double r0 = (double)(rng() >> 11) * inv53;
          double r1 = (double)(rng() >> 11) * inv53;
          double r2 = (double)(rng() >> 11) * inv53;
          x += (2.0 * r0 - 1.0) * jitter * hx;
          y += (2.0 * r1 - 1.0) * jitter * hy;
          z += (2.0 * r2 - 1.0) * jitter * hz;
        }
        coords[3 * n] = x;
        coords[3 * n + 1] = y;
        coords[3 * n + 2] = z;
      }
  for (long k = 0; k < nz; ++k)
    for (long j = 0; j < ny; ++j)
      for (long i = 0; i < nx; ++i) {
        const long e = i + nx * (j + (long)ny * k);
        const long b = i + sx * (j + sy * k);
        int *c = conn + 8 * e;
        c[0] = (int)b;
        c[1] = (int)(b + 1);
        c[2] = (int)(b + 1 + sx);
        c[3] = (int)(b + sx);
        c[4] = (int)(b + sx * sy);
        c[5] = (int)(b + 1 + sx * sy);
        c[6] = (int)(b + 1 + sx + sx * sy);
        c[7] = (int)(b + sx + sx * sy);
      }
  return 0;
}

// Ellipsoid by a mapped cube: the n^3 block on [-1, 1]^3 (numbering as
// ft_mesh_block) is mapped onto the unit ball, then scaled by the semi-axes.
// A node at cube "radius" r = max|x_i| goes to the blend (1 - s) x + s r d
// with s = r^0.3 and d the unit direction of tan(pi/4 x_i / r) (the
// equiangular cubed-sphere coordinates of its concentric cube): the cube's
// faces land exactly on the sphere, concentric cubes become near-spheres
// toward the surface, and the centre stays a regular cube.  Unlike a direct
// cube-to-ball formula, which collapses the elements at the 8 cube corners,
// this keeps every element valid (smallest element volume about half the
// mean at n = 126).  radius (per element, nullable): the ball radius of the
// mapped element centre, the normalised radius used to assign regions.
namespace {
void ball_map(double x, double y, double z, double out[3]) {
  const double r = std::max(fabs(x), std::max(fabs(y), fabs(z)));
  if (r == 0.0) {
    out[0] = out[1] = out[2] = 0.0;
    return;
  }
  const double q = 0.78539816339744831 / r;  // pi/4 / r
  const double d[3] = {tan(q * x), tan(q * y), tan(q * z)};
  const double nd = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  const double s = pow(r, 0.3);
  const double c[3] = {x, y, z};
  for (int k = 0; k < 3; ++k) out[k] = (1.0 - s) * c[k] + s * r * d[k] / nd;
}
}  // namespace

int ft_mesh_ellipsoid(int n, double a, double b, double c, double *coords, int *conn, double *radius) {
  if (n <= 0 || !(a > 0) || !(b > 0) || !(c > 0) || !coords || !conn)
    return mfail(FT_ERR_ARG, "bad ellipsoid arguments");
  const double h = 2.0 / n;
  int rc = ft_mesh_block(n, n, n, h, h, h, 0.0, 0, coords, conn);
  if (rc) return rc;
  const long s = n + 1, nn = s * s * s;
  for (long i = 0; i < nn; ++i) {
    double q[3];
    ball_map(coords[3 * i] - 1.0, coords[3 * i + 1] - 1.0, coords[3 * i + 2] - 1.0, q);
    coords[3 * i] = a * q[0];
    coords[3 * i + 1] = b * q[1];
    coords[3 * i + 2] = c * q[2];
  }
  if (radius) {
    const long ne = (long)n * n * n;
    for (long e = 0; e < ne; ++e) {
      const long i = e % n, j = (e / n) % n, k = e / ((long)n * n);
      double q[3];
      ball_map(-1.0 + (i + 0.5) * h, -1.0 + (j + 0.5) * h, -1.0 + (k + 0.5) * h, q);
      radius[e] = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
    }
  }
  return 0;
}

int ft_write_inp(const char *path, int n_nodes, int n_elems, const double *coords, const int *conn,
                 const int *pid) {
  FILE *f = fopen(path, "w");
  if (!f) return mfail(FT_ERR_INPUT, "cannot open %s", path);
  fprintf(f, "*Heading\n** written by femtech_amd\n*Node\n");
  for (int n = 0; n < n_nodes; ++n)
    fprintf(f, "%d, %.17g, %.17g, %.17g\n", n + 1, coords[3 * n], coords[3 * n + 1], coords[3 * n + 2]);
  // One *Element section per run of equal part id, in element order.  The
  // reader numbers parts by first appearance of their ELSET name
  // (ReadAbaqus.cpp:172-179), so part ids must first appear in increasing
  // order for the ids to survive the round trip.
  int next_new = 0;
  std::vector<char> seen;
  for (int e = 0; e < n_elems; ++e) {
    int p = pid ? pid[e] : 0;
    if (p < 0) {
      fclose(f);
      return mfail(FT_ERR_ARG, "negative part id");
    }
    if ((int)seen.size() <= p) seen.resize(p + 1, 0);
    if (!seen[p]) {
      if (p != next_new) {
        fclose(f);
        return mfail(FT_ERR_ARG, "part ids must first appear in increasing order");
      }
      seen[p] = 1;
      ++next_new;
    }
  }
  int prev = -1;
  for (int e = 0; e < n_elems; ++e) {
    int p = pid ? pid[e] : 0;
    if (p != prev) fprintf(f, "*Element, type=C3D8, elset=PART%d\n", p);
    prev = p;
    const int *c = conn + 8 * (size_t)e;
    fprintf(f, "%d, %d, %d, %d, %d, %d, %d, %d, %d\n", e + 1, c[0] + 1, c[1] + 1, c[2] + 1, c[3] + 1,
            c[4] + 1, c[5] + 1, c[6] + 1, c[7] + 1);
  }
  fprintf(f, "*End\n");
  fclose(f);
  return 0;
}

int ft_read_mesh(const char *path, int *n_nodes, int *n_elems, double *coords, int *conn, int *pid,
                 int *global_node_id) {
  if (!path) return mfail(FT_ERR_ARG, "null path");
  const char *dot = strrchr(path, '.');
  if (!dot) return mfail(FT_ERR_INPUT, "Input file type is unknown");
  FILE *f = fopen(path, "rb");
  if (!f) return mfail(FT_ERR_INPUT, "Cannot open input mesh file %s", path);
  RawMesh m;
  int rc;
  if (strcasecmp(dot, ".inp") == 0) rc = read_inp(f, m);
  else if (strcasecmp(dot, ".k") == 0) rc = read_k(f, m);
  else rc = mfail(FT_ERR_INPUT, "Input file type is unknown");
  fclose(f);
  if (rc) return rc;
  if (m.pid.empty()) return mfail(FT_ERR_INPUT, "No element found");
  // local node i = i-th smallest global id used by the connectivity
  std::vector<long> used(m.elem_nodes);
  std::sort(used.begin(), used.end());
  used.erase(std::unique(used.begin(), used.end()), used.end());
  const int nn = (int)used.size(), ne = (int)m.pid.size();
  if (n_nodes) *n_nodes = nn;
  if (n_elems) *n_elems = ne;
  if (!coords && !conn && !pid && !global_node_id) return 0;
  std::vector<long> order(m.node_id.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = (long)i;
  std::sort(order.begin(), order.end(), [&](long a, long b) { return m.node_id[a] < m.node_id[b]; });
  std::vector<long> sorted_ids(order.size());
  for (size_t i = 0; i < order.size(); ++i) sorted_ids[i] = m.node_id[order[i]];
  if (coords) {
    for (int i = 0; i < nn; ++i) {
      auto it = std::lower_bound(sorted_ids.begin(), sorted_ids.end(), used[i]);
      if (it == sorted_ids.end() || *it != used[i])
        return mfail(FT_ERR_INPUT, "node %ld used by an element is not defined", used[i]);
      const long src = order[it - sorted_ids.begin()];
      for (int d = 0; d < 3; ++d) coords[3 * i + d] = m.node_xyz[3 * src + d];
    }
  }
  if (conn)
    for (size_t k = 0; k < m.elem_nodes.size(); ++k)
      conn[k] = (int)(std::lower_bound(used.begin(), used.end(), m.elem_nodes[k]) - used.begin());
  if (pid)
    for (int e = 0; e < ne; ++e) pid[e] = m.pid[e];
  if (global_node_id)
    for (int i = 0; i < nn; ++i) global_node_id[i] = (int)used[i];
  return 0;
}

int ft_read_materials(const char *path, int n_parts, int *material_id, double *props) {
  FILE *f = fopen(path, "r");
  if (!f) return mfail(FT_ERR_INPUT, "Cannot open materials.dat file");
  // ReadMaterials.cpp:43-122 for 0-5; this engine's laws 6-8: Ogden (rho, mu,
  // lambda, alpha), Ogden + Prony (rho, mu, lambda, alpha, unused, g1, tau1,
  // g2, tau2), anisotropic HGO (rho, mu, lambda, k1, k2, kappa)
  const int nprops[9] = {1, 3, 3, 3, 5, 9, 4, 9, 6};
  std::vector<char> seen(n_parts, 0);
  for (int p = 0; p < n_parts; ++p)
    for (int k = 0; k < 9; ++k) props[9 * p + k] = 0.0;
  for (int i = 0; i < n_parts; ++i) {
    int part = -1, mat = -1;
    if (fscanf(f, "%d", &part) != 1 || fscanf(f, "%d", &mat) != 1) {
      fclose(f);
      return mfail(FT_ERR_INPUT, "materials.dat: expected %d part lines", n_parts);
    }
    if (part < 0 || part >= n_parts) {
      fclose(f);
      return mfail(FT_ERR_INPUT, "materials.dat: part %d out of range", part);
    }
    if (mat < 0 || mat > 8) {
      fclose(f);
      return mfail(FT_ERR_INPUT, "Material ID for Part %d not found", part);
    }
    material_id[part] = mat;
    for (int k = 0; k < nprops[mat]; ++k)
      if (fscanf(f, "%lf", &props[9 * part + k]) != 1) {
        fclose(f);
        return mfail(FT_ERR_INPUT, "materials.dat: part %d needs %d properties", part, nprops[mat]);
      }
    seen[part] = 1;
  }
  fclose(f);
  for (int p = 0; p < n_parts; ++p)
    if (!seen[p]) return mfail(FT_ERR_INPUT, "Material properties of partID %d missing", p);
  return 0;
}

}  // extern "C"
