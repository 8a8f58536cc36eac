// bh_host.cpp -- host-side setup of the assembly path: reference tables, dof
// numbering, CSR pattern, cell colouring (vector assembly) and the gather
// structure of the deterministic row reduction (matrix assembly).
//
// Everything here runs once per mesh/space/pattern; the per-assembly work is
// on the GPU (bh_kernels.hip).  Reference citations are relative to the
// Bcube.jl source tree.
#include <algorithm>
#include <array>
#include <atomic>
#include <climits>
#include <cmath>
#include <cstring>
#include <functional>
#include <thread>
#include <unordered_map>

#include "bh_internal.h"

namespace bh {

// ---------------------------------------------------------------------------
// Small helpers
// ---------------------------------------------------------------------------
void parallel_for(i64 n, const std::function<void(i64, i64)>& body) {
  if (n <= 0) return;
  unsigned hw = std::thread::hardware_concurrency();
  int nt = (int)std::min<i64>(std::max(1u, std::min(hw, 16u)), (n + 4095) / 4096);
  if (nt <= 1) { body(0, n); return; }
  std::vector<std::thread> th;
  i64 chunk = (n + nt - 1) / nt;
  for (int t = 0; t < nt; ++t) {
    i64 b = t * chunk, e = std::min(n, b + chunk);
    if (b >= e) break;
    th.emplace_back([&body, b, e] { body(b, e); });
  }
  for (auto& t : th) t.join();
}

int shape_of(int ct) {  // src/mesh/shape.jl:25-32
  switch (ct) {
    case BH_QUAD4: case BH_QUAD9: return SQUARE;
    case BH_TRI3: return TRIANGLE;
    case BH_HEXA8: case BH_HEXA27: return CUBE;
    case BH_BAR2: return LINE;
  }
  return -1;
}
int dim_of_shape(int s) { return s == CUBE ? 3 : (s == LINE ? 1 : 2); }
int nnodes_of(int ct) {
  switch (ct) {
    case BH_QUAD4: return 4; case BH_QUAD9: return 9; case BH_TRI3: return 3;
    case BH_HEXA8: return 8; case BH_HEXA27: return 27; case BH_BAR2: return 2;
  }
  return 0;
}
int ndofs_lagrange(int shape, int order) {  // hypercube.jl:122-133, triangle.jl:67-70
  if (order < 0) return 0;
  if (shape == TRIANGLE) return (order + 1) * (order + 2) / 2;
  int n = order + 1;
  return shape == LINE ? n : (shape == SQUARE ? n * n : n * n * n);
}

// ---------------------------------------------------------------------------
// Tables: Gauss-Legendre rule, tensor Lagrange basis on uniform nodes, and the
// Quad4/Hexa8 geometric basis.
// ---------------------------------------------------------------------------
int gauss_points_per_axis(int quad_degree) {  // quadrature.jl:238-240
  if (quad_degree <= 0) quad_degree = 1;     // degree 0 uses the degree-1 rule
  return (quad_degree + 2) / 2;
}

// Gauss-Legendre nodes/weights by Newton iteration on the three-term
// recurrence of P_n, in long double.
void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  for (int k = 0; k < (n + 1) / 2; ++k) {
    long double z = cosl(3.141592653589793238462643383279502884L * (k + 0.75L) / (n + 0.5L));
    long double dp = 0;
    for (int it = 0; it < 64; ++it) {
      long double p0 = 1, p1 = z;
      for (int j = 2; j <= n; ++j) {
        long double p2 = ((2 * j - 1) * z * p1 - (j - 1) * p0) / j;
        p0 = p1; p1 = p2;
      }
      if (n == 1) { p1 = z; p0 = 1; }
      dp = n * (z * p1 - p0) / (z * z - 1);
      long double dz = p1 / dp;
      z -= dz;
      if (fabsl(dz) < 1e-20L) break;
    }
    long double p0 = 1, p1 = z;
    for (int j = 2; j <= n; ++j) {
      long double p2 = ((2 * j - 1) * z * p1 - (j - 1) * p0) / j;
      p0 = p1; p1 = p2;
    }
    if (n == 1) { p1 = z; p0 = 1; }
    dp = n * (z * p1 - p0) / (z * z - 1);
    double wk = (double)(2.0L / ((1 - z * z) * dp * dp));
    x[n - 1 - k] = (double)z;
    x[k] = -(double)z;
    w[k] = w[n - 1 - k] = wk;
  }
  if (n % 2 == 1) x[n / 2] = 0.0;
}

// 1D Lagrange basis on uniform nodes of [-1,1] (Lagrange{:Uniform},
// lagrange.jl:16) and its derivative.
void lagrange_1d(int order, double t, double* v, double* dv) {
  int n = order + 1;
  if (order == 0) { v[0] = 1.0; dv[0] = 0.0; return; }
  std::vector<double> nodes(n);
  for (int j = 0; j < n; ++j) nodes[j] = -1.0 + 2.0 * j / order;
  for (int j = 0; j < n; ++j) {
    double den = 1.0, num = 1.0, der = 0.0;
    for (int k = 0; k < n; ++k) {
      if (k == j) continue;
      den *= nodes[j] - nodes[k];
      der = der * (t - nodes[k]) + num;  // product rule, accumulated
      num *= t - nodes[k];
    }
    v[j] = num / den;
    dv[j] = der / den;
  }
}

// Triangle rules on the reference triangle (0,0),(1,0),(0,1).  Degrees 1..3:
// the GLTRI1/GLTRI3/GLTRI4 rules of the FEMQuad 0.3 dependency
// (quadrature.jl:280-288): the centroid, the three interior points
// (2/3,1/6),(1/6,2/3),(1/6,1/6), and Strang & Fix's four-point rule with its
// negative centroid weight.  Degrees 4..8: the GLTRI6/7/12/16 tables of the
// reference (quadrature.jl:289-377, Witherden & Vincent).
static void triangle_rule(int degree, std::vector<double>& w, std::vector<double>& x) {
  w.clear();
  x.clear();
  auto pt = [&](double a, double b) { x.push_back(a); x.push_back(b); };
  if (degree == 1) {
    w = {0.5};
    pt(1.0 / 3, 1.0 / 3);
  } else if (degree == 2) {
    w = {1.0 / 6, 1.0 / 6, 1.0 / 6};
    pt(2.0 / 3, 1.0 / 6); pt(1.0 / 6, 2.0 / 3); pt(1.0 / 6, 1.0 / 6);
  } else if (degree == 3) {
    w = {-27.0 / 96, 25.0 / 96, 25.0 / 96, 25.0 / 96};
    pt(1.0 / 3, 1.0 / 3); pt(0.6, 0.2); pt(0.2, 0.6); pt(0.2, 0.2);
  } else if (degree == 4) {
    const double P1 = 0.22338158967801146569500700843312 / 2;
    const double P2 = 0.10995174365532186763832632490021 / 2;
    const double A = 0.44594849091596488631832925388305;
    const double B = 0.09157621350977074345957146340220;
    w = {P2, P2, P2, P1, P1, P1};
    pt(B, B); pt(1.0 - 2.0 * B, B); pt(B, 1.0 - 2.0 * B);
    pt(A, 1.0 - 2 * A); pt(A, A); pt(1.0 - 2.0 * A, A);
  } else if (degree == 5) {
    const double A = 0.47014206410511508977044120951345;
    const double B = 0.10128650732345633880098736191512;
    const double P1 = 0.13239415278850618073764938783315 / 2;
    const double P2 = 0.12593918054482715259568394550018 / 2;
    w = {9.0 / 80, P1, P1, P1, P2, P2, P2};
    pt(1.0 / 3, 1.0 / 3); pt(A, A); pt(1.0 - 2.0 * A, A); pt(A, 1.0 - 2.0 * A);
    pt(B, B); pt(1.0 - 2.0 * B, B); pt(B, 1.0 - 2.0 * B);
  } else if (degree == 6) {
    const double A = (-0.87382197101699554331933679425836168532 + 1) / 2;
    const double B = (-0.50142650965817915741672289378596184782 + 1) / 2;
    const double C = (-0.37929509793243118916678453208689569359 + 1) / 2;
    const double D = (-0.89370990031036610529350065673720370601 + 1) / 2;
    const double P1 = 0.10168981274041363384187361821373796809 / 4;
    const double P2 = 0.23357255145275873205057922277115888265 / 4;
    const double P3 = 0.16570215123674715038710691284088490796 / 4;
    w = {P1, P1, P1, P2, P2, P2, P3, P3, P3, P3, P3, P3};
    pt(A, A); pt(1.0 - 2.0 * A, A); pt(A, 1.0 - 2.0 * A);
    pt(B, B); pt(1.0 - 2.0 * B, B); pt(B, 1.0 - 2.0 * B);
    pt(C, D); pt(D, C); pt(1.0 - C - D, C); pt(1.0 - C - D, D); pt(C, 1.0 - C - D);
    pt(D, 1.0 - C - D);
  } else if (degree == 7 || degree == 8) {
    // (x, y, w) on [-1,1]^2, rescaled by (x+1)/2, (y+1)/2, w/4
    static const double T[16][3] = {
        {-0.33333333333333333333333333333333333333, -0.33333333333333333333333333333333333333, 0.2886312153555743365021822209781292496},
        {-0.081414823414553687942368971011661355879, -0.83717035317089262411526205797667728824, 0.1901832685345692495877922087771686332},
        {-0.83717035317089262411526205797667728824, -0.081414823414553687942368971011661355879, 0.1901832685345692495877922087771686332},
        {-0.081414823414553687942368971011661355879, -0.081414823414553687942368971011661355879, 0.1901832685345692495877922087771686332},
        {-0.65886138449647958675541299701707099796, 0.31772276899295917351082599403414199593, 0.20643474106943650056358310058425806003},
        {0.31772276899295917351082599403414199593, -0.65886138449647958675541299701707099796, 0.20643474106943650056358310058425806003},
        {-0.65886138449647958675541299701707099796, -0.65886138449647958675541299701707099796, 0.20643474106943650056358310058425806003},
        {-0.89890554336593804908315289880680210631, 0.79781108673187609816630579761360421262, 0.064916995246396160621851856683561193593},
        {0.79781108673187609816630579761360421262, -0.89890554336593804908315289880680210631, 0.064916995246396160621851856683561193593},
        {-0.89890554336593804908315289880680210631, -0.89890554336593804908315289880680210631, 0.064916995246396160621851856683561193593},
        {-0.98321044518008478932557233092141110162, 0.45698478591080856248200075835212392604, 0.05446062834886998852968938014781784832},
        {0.45698478591080856248200075835212392604, -0.98321044518008478932557233092141110162, 0.05446062834886998852968938014781784832},
        {-0.47377434073072377315642842743071282442, 0.45698478591080856248200075835212392604, 0.05446062834886998852968938014781784832},
        {0.45698478591080856248200075835212392604, -0.47377434073072377315642842743071282442, 0.05446062834886998852968938014781784832},
        {-0.47377434073072377315642842743071282442, -0.98321044518008478932557233092141110162, 0.05446062834886998852968938014781784832},
        {-0.98321044518008478932557233092141110162, -0.47377434073072377315642842743071282442, 0.05446062834886998852968938014781784832}};
    for (int k = 0; k < 16; ++k) {
      pt(0.5 * (T[k][0] + 1), 0.5 * (T[k][1] + 1));
      w.push_back(T[k][2] / 4);
    }
  }
}

int triangle_rule_points(int quad_degree) {
  std::vector<double> w, x;
  triangle_rule(quad_degree, w, x);
  return (int)w.size();
}

// Triangle Lagrange nodes as barycentric lattice points (i, j, k), i+j+k = p,
// in the reference's dof order: vertices, edge dofs from each edge's first
// vertex to its second, interior (triangle.jl:1-45, 120-145; P4 extends P3's
// layout with interior nodes (2,1,1), (1,2,1), (1,1,2)).
static const int kTriLattice[5][15][3] = {
    {{0, 0, 0}},
    {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}},
    {{2, 0, 0}, {0, 2, 0}, {0, 0, 2}, {1, 1, 0}, {0, 1, 1}, {1, 0, 1}},
    {{3, 0, 0}, {0, 3, 0}, {0, 0, 3}, {2, 1, 0}, {1, 2, 0}, {0, 2, 1}, {0, 1, 2}, {1, 0, 2},
     {2, 0, 1}, {1, 1, 1}},
    {{4, 0, 0}, {0, 4, 0}, {0, 0, 4}, {3, 1, 0}, {2, 2, 0}, {1, 3, 0}, {0, 3, 1}, {0, 2, 2},
     {0, 1, 3}, {1, 0, 3}, {2, 0, 2}, {3, 0, 1}, {2, 1, 1}, {1, 2, 1}, {1, 1, 2}}};

// Π_{m<n} (p λ - m)/(m+1) and its derivative in λ.
static void lattice_factor(int p, int n, double lam, double& f, double& df) {
  f = 1.0;
  df = 0.0;
  for (int m = 0; m < n; ++m) {
    double t = (p * lam - m) / (m + 1);
    df = df * t + f * (double)p / (m + 1);
    f *= t;
  }
}

static void triangle_basis(int p, double x, double y, double* val, double* grad) {
  const double l[3] = {1.0 - x - y, x, y};
  const int nl = (p + 1) * (p + 2) / 2;
  for (int n = 0; n < nl; ++n) {
    double f[3], df[3];
    for (int a = 0; a < 3; ++a) lattice_factor(p, kTriLattice[p][n][a], l[a], f[a], df[a]);
    val[n] = f[0] * f[1] * f[2];
    const double d1 = df[0] * f[1] * f[2], d2 = f[0] * df[1] * f[2], d3 = f[0] * f[1] * df[2];
    grad[2 * n] = d2 - d1;      // ∂/∂ξ1: λ2 = ξ1, λ1 = 1 - ξ1 - ξ2
    grad[2 * n + 1] = d3 - d1;  // ∂/∂ξ2: λ3 = ξ2
  }
}

static Tables make_triangle_tables(int order, int quad_degree) {
  Tables T;
  T.dim = 2;
  T.order = order;
  T.nl = (order + 1) * (order + 2) / 2;
  triangle_rule(quad_degree, T.w, T.xi);
  T.nq = (int)T.w.size();
  T.phi.assign((size_t)T.nq * T.nl, 0.0);
  T.dphi.assign((size_t)T.nq * T.nl * 2, 0.0);
  for (int q = 0; q < T.nq; ++q)
    triangle_basis(order, T.xi[2 * q], T.xi[2 * q + 1], &T.phi[(size_t)q * T.nl],
                   &T.dphi[(size_t)q * T.nl * 2]);
  // Tri3 geometry (mapping.jl:395-416): N = (1-ξ1-ξ2, ξ1, ξ2)
  T.ngeo = 3;
  for (int q = 0; q < T.nq; ++q) {
    const double x = T.xi[2 * q], y = T.xi[2 * q + 1];
    T.gphi.insert(T.gphi.end(), {1.0 - x - y, x, y});
    T.dgphi.insert(T.dgphi.end(), {-1.0, -1.0, 1.0, 0.0, 0.0, 1.0});
  }
  return T;
}

Tables make_tables(int shape, int order, int quad_degree, int geo_ct) {
  if (shape == TRIANGLE) return make_triangle_tables(order, quad_degree);
  Tables T;
  T.dim = dim_of_shape(shape);
  T.order = order;
  T.nl = ndofs_lagrange(shape, order);
  int d = T.dim;
  int n1 = gauss_points_per_axis(quad_degree);
  std::vector<double> gx, gw;
  gauss_legendre(n1, gx, gw);
  T.nq = d == 1 ? n1 : (d == 2 ? n1 * n1 : n1 * n1 * n1);
  // tensor rule, first axis fastest (quadrature.jl:455-487); the Line rule
  // is the 1D Gauss-Legendre rule itself
  for (int q = 0; q < T.nq; ++q) {
    int i = q % n1, j = (q / n1) % n1, k = q / (n1 * n1);
    double w = gw[i];
    if (d >= 2) w *= gw[j];
    if (d == 3) w *= gw[k];
    T.w.push_back(w);
    T.xi.push_back(gx[i]);
    if (d >= 2) T.xi.push_back(gx[j]);
    if (d == 3) T.xi.push_back(gx[k]);
  }
  // tensor Lagrange basis, dof index first-axis fastest (hypercube.jl:56-92)
  int n = order + 1;
  T.phi.assign((size_t)T.nq * T.nl, 0.0);
  T.dphi.assign((size_t)T.nq * T.nl * d, 0.0);
  std::vector<double> v[3], dv[3];
  for (int a = 0; a < 3; ++a) { v[a].resize(n); dv[a].resize(n); }
  for (int q = 0; q < T.nq; ++q) {
    for (int a = 0; a < d; ++a) lagrange_1d(order, T.xi[q * d + a], v[a].data(), dv[a].data());
    for (int l = 0; l < T.nl; ++l) {
      int i = l % n, j = (l / n) % n, k = l / (n * n);
      double* g = &T.dphi[((size_t)q * T.nl + l) * d];
      if (d == 1) {  // Line: the 1D Lagrange basis (hypercube.jl:51-54)
        T.phi[(size_t)q * T.nl + l] = v[0][i];
        g[0] = dv[0][i];
      } else if (d == 2) {
        T.phi[(size_t)q * T.nl + l] = v[0][i] * v[1][j];
        g[0] = dv[0][i] * v[1][j];
        g[1] = v[0][i] * dv[1][j];
      } else {
        T.phi[(size_t)q * T.nl + l] = v[0][i] * v[1][j] * v[2][k];
        g[0] = dv[0][i] * v[1][j] * v[2][k];
        g[1] = v[0][i] * dv[1][j] * v[2][k];
        g[2] = v[0][i] * v[1][j] * dv[2][k];
      }
    }
  }
  // geometric basis in CGNS node order: bilinear / trilinear on Quad4 / Hexa8
  // (mapping.jl:460-465, 618-639), products of the 1D quadratics on {-1, 0, 1}
  // on Quad9 / Hexa27 (mapping.jl:544-570, 664-697); each node's reference
  // coordinates are entity.jl's (Quad9: vertices, edge midpoints 12 23 34 41,
  // centre; Hexa27: vertices, 12 edges, 6 face centres, centre)
  static const int8_t q4[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
  static const int8_t h8[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                  {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
  static const int8_t nq9[9][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}, {0, -1},
                                   {1, 0},   {0, 1},  {-1, 0}, {0, 0}};
  static const int8_t nh27[27][3] = {
      {-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1}, {-1, -1, 1}, {1, -1, 1}, {1, 1, 1},
      {-1, 1, 1},   {0, -1, -1}, {1, 0, -1}, {0, 1, -1},  {-1, 0, -1}, {-1, -1, 0}, {1, -1, 0},
      {1, 1, 0},    {-1, 1, 0},  {0, -1, 1}, {1, 0, 1},   {0, 1, 1},   {-1, 0, 1},  {0, 0, -1},
      {0, -1, 0},   {1, 0, 0},   {0, 1, 0},  {-1, 0, 0},  {0, 0, 1},   {0, 0, 0}};
  const bool quadratic = geo_ct == BH_QUAD9 || geo_ct == BH_HEXA27;
  T.ngeo = geo_ct == BH_HEXA27 ? 27 : geo_ct == BH_QUAD9 ? 9
           : geo_ct == BH_HEXA8 ? 8 : geo_ct == BH_QUAD4 ? 4 : (geo_ct == BH_BAR2 ? 2 : 0);
  // Bar2: nodes at ξ = -1, 1 (mapping.jl's Bar2 map (1-ξ)/2 x1 + (1+ξ)/2 x2)
  static const int8_t b2[2] = {-1, 1};
  T.gphi.assign((size_t)T.nq * T.ngeo, 0.0);
  T.dgphi.assign((size_t)T.nq * T.ngeo * d, 0.0);
  // 1D geometric basis at t for the node coordinate s in {-1, 0, 1}
  auto basis1 = [&](int s, double t, double& v, double& dv) {
    if (!quadratic) {
      v = 0.5 * (1.0 + s * t);
      dv = 0.5 * s;
    } else if (s == -1) {
      v = 0.5 * t * (t - 1.0);
      dv = t - 0.5;
    } else if (s == 0) {
      v = 1.0 - t * t;
      dv = -2.0 * t;
    } else {
      v = 0.5 * t * (t + 1.0);
      dv = t + 0.5;
    }
  };
  for (int q = 0; q < T.nq; ++q) {
    const double* x = &T.xi[q * d];
    for (int k = 0; k < T.ngeo; ++k) {
      double f[3], df[3];
      for (int a = 0; a < d; ++a) {
        const int s = geo_ct == BH_HEXA27 ? nh27[k][a]
                      : geo_ct == BH_QUAD9 ? nq9[k][a]
                      : d == 1 ? b2[k] : d == 3 ? h8[k][a] : q4[k][a];
        basis1(s, x[a], f[a], df[a]);
      }
      double val = 1.0;
      for (int a = 0; a < d; ++a) val *= f[a];
      T.gphi[(size_t)q * T.ngeo + k] = val;
      for (int a = 0; a < d; ++a) {
        double g = df[a];
        for (int b = 0; b < d; ++b)
          if (b != a) g *= f[b];
        T.dgphi[((size_t)q * T.ngeo + k) * d + a] = g;
      }
    }
  }
  return T;
}

// ---------------------------------------------------------------------------
// Entity tables (src/mesh/entity.jl, CGNS order) and Lagrange dof locations
// (hypercube.jl:136-331, triangle.jl:72-98); all 0-based.
// ---------------------------------------------------------------------------
typedef std::vector<std::vector<int>> Lists;

static Lists edges_of(int ct) {
  switch (ct) {
    case BH_QUAD4: return {{0, 1}, {1, 2}, {2, 3}, {3, 0}};
    case BH_QUAD9: return {{0, 1, 4}, {1, 2, 5}, {2, 3, 6}, {3, 0, 7}};
    case BH_TRI3: return {{0, 1}, {1, 2}, {2, 0}};
    case BH_HEXA8: return {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {0, 4}, {1, 5},
                           {2, 6}, {3, 7}, {4, 5}, {5, 6}, {6, 7}, {7, 4}};
    case BH_HEXA27: return {{0, 1, 8}, {1, 2, 9}, {2, 3, 10}, {3, 0, 11},
                            {0, 4, 12}, {1, 5, 13}, {2, 6, 14}, {3, 7, 15},
                            {4, 5, 16}, {5, 6, 17}, {6, 7, 18}, {7, 4, 19}};
  }
  return {};
}
static Lists faces_of(int ct) {
  switch (ct) {
    case BH_QUAD4: case BH_QUAD9: case BH_TRI3: return edges_of(ct);
    case BH_HEXA8: return {{0, 3, 2, 1}, {0, 1, 5, 4}, {1, 2, 6, 5},
                           {2, 3, 7, 6}, {0, 4, 7, 3}, {4, 5, 6, 7}};
    case BH_HEXA27: return {{0, 3, 2, 1, 11, 10, 9, 8, 20}, {0, 1, 5, 4, 8, 13, 16, 12, 21},
                            {1, 2, 6, 5, 9, 14, 17, 13, 22}, {2, 3, 7, 6, 10, 15, 18, 14, 23},
                            {0, 4, 7, 3, 12, 19, 15, 11, 24}, {4, 5, 6, 7, 16, 17, 18, 19, 25}};
  }
  return {};
}

static Lists vertex_dofs(int shape, int order) {
  int nv = shape == CUBE ? 8 : (shape == SQUARE ? 4 : (shape == LINE ? 2 : 3));
  if (order == 0) return Lists(nv);
  if (shape == TRIANGLE) return {{0}, {1}, {2}};
  int n = order + 1, m = n - 1;
  if (shape == LINE) return {{0}, {m}};  // _idof_by_vertex(::Line), hypercube.jl:136-141
  if (shape == SQUARE) return {{0}, {m}, {m + n * m}, {n * m}};
  int nn = n * n;
  return {{0}, {m}, {m + n * m}, {n * m},
          {nn * m}, {m + nn * m}, {m + n * m + nn * m}, {n * m + nn * m}};
}

static Lists edge_dofs(int shape, int order) {
  if (shape == TRIANGLE) {
    if (order <= 1) return Lists(3);
    if (order == 2) return {{3}, {4}, {5}};
    if (order == 3) return {{3, 4}, {5, 6}, {7, 8}};
    // P4 (beyond the reference's P3): three dofs per edge, first vertex to second
    if (order == 4) return {{3, 4, 5}, {6, 7, 8}, {9, 10, 11}};
    return Lists();
  }
  int ne = shape == CUBE ? 12 : 4;
  if (order == 0) return Lists(ne);
  int n = order + 1;
  auto I = [&](int i, int j, int k) { return i + n * j + n * n * k; };  // 0-based axes
  Lists e(ne);
  int L = n - 1;
  for (int t = 1; t <= n - 2; ++t) {
    int r = t, rr = n - 1 - t;  // forward and reversed interior index
    if (shape == SQUARE) {
      e[0].push_back(I(r, 0, 0));
      e[1].push_back(I(L, r, 0));
      e[2].push_back(I(rr, L, 0));
      e[3].push_back(I(0, rr, 0));
    } else {
      e[0].push_back(I(r, 0, 0));
      e[1].push_back(I(L, r, 0));
      e[2].push_back(I(rr, L, 0));
      e[3].push_back(I(0, rr, 0));
      e[4].push_back(I(0, 0, r));
      e[5].push_back(I(L, 0, r));
      e[6].push_back(I(L, L, r));
      e[7].push_back(I(0, L, r));
      e[8].push_back(I(r, 0, L));
      e[9].push_back(I(L, r, L));
      e[10].push_back(I(rr, L, L));
      e[11].push_back(I(0, rr, L));
    }
  }
  return e;
}

static Lists face_dofs(int shape, int order) {
  if (shape != CUBE) return Lists();
  if (order == 0) return Lists(6);
  int n = order + 1, L = n - 1;
  auto I = [&](int i, int j, int k) { return i + n * j + n * n * k; };
  Lists f(6);
  for (int b = 1; b <= n - 2; ++b)
    for (int a = 1; a <= n - 2; ++a) {  // column-major vec(): first range fastest
      f[0].push_back(I(a, b, 0));
      f[1].push_back(I(a, 0, b));
      f[2].push_back(I(L, a, b));
      f[3].push_back(I(a, L, b));
      f[4].push_back(I(0, a, b));
      f[5].push_back(I(a, b, L));
    }
  return f;
}

// oriented_cell_side (entity.jl:487-502): index of the face of cell (ct, nodes)
// whose node set is contained in `fn`, negative if traversed in the other
// direction, 0 if none.
static int oriented_cell_side(int ct, const i64* nodes, const std::vector<i64>& fn) {
  Lists fl = faces_of(ct);
  for (size_t s = 0; s < fl.size(); ++s) {
    bool sub = true;
    for (int ln : fl[s])
      if (std::find(fn.begin(), fn.end(), nodes[ln]) == fn.end()) { sub = false; break; }
    if (!sub) continue;
    int side = (int)s + 1, n1 = 0;
    for (size_t q = 0; q < fl[s].size(); ++q)
      if (nodes[fl[s][q]] == fn[0]) { n1 = (int)q + 1; break; }
    if (fn.size() == 2) return n1 == 2 ? -side : side;
    int n2 = n1 + 1 > (int)fl[s].size() ? 1 : n1 + 1;
    if (fn[1] != nodes[fl[s][n2 - 1]]) side = -side;
    return side;
  }
  return 0;
}

struct Key9 {
  std::array<i64, 9> k;
  bool operator==(const Key9& o) const { return k == o.k; }
};
struct Key9Hash {
  size_t operator()(const Key9& x) const {
    uint64_t h = 1469598103934665603ull;
    for (i64 v : x.k) { h ^= (uint64_t)v + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2); }
    return (size_t)h;
  }
};
static Key9 make_key(const std::vector<i64>& nodes) {
  Key9 key;
  key.k.fill(-1);
  std::vector<i64> s = nodes;
  std::sort(s.begin(), s.end());
  for (size_t i = 0; i < s.size() && i < 9; ++i) key.k[i] = s[i];
  return key;
}

bool dof_table(int ct, int order, DofTab& t) {
  t = DofTab();
  const int s = shape_of(ct);
  if (s < 0 || order < 0 || order > 4) return false;
  auto fill = [](auto& a) { for (auto& row : a) for (auto& v : row) v = -1; };
  for (int& v : t.vdof) v = -1;
  fill(t.enode); fill(t.edof); fill(t.fnode); fill(t.fdof); fill(t.snode);
  t.nd = ndofs_lagrange(s, order);
  t.dim = dim_of_shape(s);
  const Lists vd = vertex_dofs(s, order);
  t.nv = std::min((int)vd.size(), nnodes_of(ct));
  for (int v = 0; v < t.nv; ++v) {
    if (vd[v].size() > 1) return false;
    t.vdof[v] = vd[v].empty() ? -1 : vd[v][0];
  }
  if (t.dim > 1) {
    const Lists el = edges_of(ct), ed = edge_dofs(s, order);
    if (order > 0 && s == TRIANGLE && ed.size() != el.size()) return false;
    t.ne = (int)std::min(el.size(), ed.size());
    for (int e = 0; e < t.ne; ++e) {
      t.elen = (int)el[e].size();
      for (int k = 0; k < t.elen; ++k) t.enode[e][k] = el[e][k];
      t.ned = (int)ed[e].size();
      if (t.ned > 3) return false;
      for (int k = 0; k < t.ned; ++k) t.edof[e][k] = ed[e][k];
    }
  }
  if (t.dim > 2) {
    const Lists fl = faces_of(ct), fd = face_dofs(s, order);
    t.nf = (int)std::min(fl.size(), fd.size());
    for (int f = 0; f < t.nf; ++f) {
      t.flen = (int)fl[f].size();
      for (int k = 0; k < t.flen; ++k) t.fnode[f][k] = fl[f][k];
      t.nfd = (int)fd[f].size();
      if (t.nfd > 9) return false;
      for (int k = 0; k < t.nfd; ++k) t.fdof[f][k] = fd[f][k];
    }
  }
  const Lists sl = faces_of(ct);
  t.nside = (int)sl.size();
  for (int q = 0; q < t.nside; ++q) {
    t.slen = (int)sl[q].size();
    for (int k = 0; k < t.slen; ++k) t.snode[q][k] = sl[q][k];
  }
  return true;
}

// Continuous Lagrange space of per-cell order on a 2D Quad4/Tri3 mesh (no
// reference counterpart: the reference's DofHandler takes one FunctionSpace,
// dofhandler.jl:37-42).  An edge carries the lowest order p_e of its cells and
// p_e - 1 dofs, oriented like the lowest-numbered cell that has it (the
// reference orients edge dofs by their first cell, dofhandler.jl:239-290).
// Dofs are numbered by first appearance over cells in mesh order and each
// cell's local dofs in Lagrange order, which with one order everywhere is the
// reference's DofHandler + densify.  A cell's own nodes on an edge of lower
// order are not dofs: of its p_c - 1 slots on that edge, slot k <= p_e - 1
// holds the edge's dof j = k counted in the cell's direction and the others
// are -1; eord[c*4 + e] = p_e for such an edge (0 when the edge has the cell's
// order).  Edges are found by sorting (lower node, higher node) keys, so the
// cost is one sort of the cell-edge list and two passes over the cells.
i64 dofhandler_orders(const HostMesh& m, const int32_t* order, int nl, std::vector<int32_t>& dofs,
                      std::vector<uint8_t>& eord, std::string& err) {
  const i64 nc = m.ncells;
  if (m.dim != 2) { err = "per-cell orders need a 2D mesh"; return -1; }
  if (m.nnodes >= ((i64)1 << 31)) { err = "too many nodes"; return -1; }
  // local roles per (cell type, order): kind 0 vertex / 1 edge / 2 interior
  struct Role { int8_t kind[25], ent[25], k[25]; int nd; DofTab t; };
  std::vector<Role> roles(10);  // [Quad4: 0, Tri3: 1][order]
  auto role_of = [&](int ct, int p) -> Role& { return roles[(ct == BH_TRI3 ? 5 : 0) + p]; };
  for (int ct : {BH_QUAD4, BH_TRI3})
    for (int p = 1; p <= 4; ++p) {
      Role& r = role_of(ct, p);
      if (!dof_table(ct, p, r.t)) { err = "no dof table"; return -1; }
      r.nd = r.t.nd;
      for (int i = 0; i < r.nd; ++i) { r.kind[i] = 2; r.ent[i] = -1; r.k[i] = 0; }
      for (int v = 0; v < r.t.nv; ++v) { r.kind[r.t.vdof[v]] = 0; r.ent[r.t.vdof[v]] = (int8_t)v; }
      for (int e = 0; e < r.t.ne; ++e)
        for (int k = 0; k < r.t.ned; ++k) {
          const int i = r.t.edof[e][k];
          r.kind[i] = 1;
          r.ent[i] = (int8_t)e;
          r.k[i] = (int8_t)(k + 1);
        }
    }
  int need = 0;
  for (i64 c = 0; c < nc; ++c) {
    const int ct = m.ctype[c], p = order[c];
    if (ct != BH_QUAD4 && ct != BH_TRI3) { err = "per-cell orders need Quad4 / Tri3 cells"; return -1; }
    if (p < 1 || p > 4) { err = "cell orders must be 1..4"; return -1; }
    need = std::max(need, role_of(ct, p).nd);
  }
  if (need != nl) { err = "slot count is not the largest cell's dof count"; return -1; }
  // edges: sorted (key, cell*4 + local edge)
  std::vector<std::pair<uint64_t, int64_t>> er;
  er.reserve((size_t)nc * 4);
  for (i64 c = 0; c < nc; ++c) {
    const Role& r = role_of(m.ctype[c], order[c]);
    const i64* nodes = &m.c2n[m.c2n_ptr[c]];
    for (int e = 0; e < r.t.ne; ++e) {
      const uint64_t a = (uint64_t)nodes[r.t.enode[e][0]], b = (uint64_t)nodes[r.t.enode[e][1]];
      er.push_back({a < b ? (a << 32) | b : (b << 32) | a, c * 4 + e});
    }
  }
  std::sort(er.begin(), er.end());
  std::vector<int32_t> eid((size_t)nc * 4, -1);
  std::vector<int8_t> pe;
  std::vector<int32_t> efirst;  // first node of the edge's dof direction
  for (size_t a = 0; a < er.size();) {
    size_t b = a;
    int q = 5;
    while (b < er.size() && er[b].first == er[a].first) {
      q = std::min(q, (int)order[er[b].second / 4]);
      ++b;
    }
    const i64 c0 = er[a].second / 4;
    const int e0 = (int)(er[a].second % 4);
    const Role& r0 = role_of(m.ctype[c0], order[c0]);
    efirst.push_back((int32_t)m.c2n[m.c2n_ptr[c0] + r0.t.enode[e0][0]]);
    for (size_t k = a; k < b; ++k) eid[er[k].second] = (int32_t)pe.size();
    pe.push_back((int8_t)q);
    a = b;
  }
  std::vector<i64> eoff(pe.size() + 1, 0);
  for (size_t E = 0; E < pe.size(); ++E) eoff[E + 1] = eoff[E] + pe[E] - 1;
  std::vector<int32_t> pool(eoff.back(), -1), vid(m.nnodes, -1);
  dofs.assign((size_t)nc * nl, -1);
  eord.assign((size_t)nc * 4, 0);
  i64 next = 0;
  for (i64 c = 0; c < nc; ++c) {
    const int p = order[c];
    const Role& r = role_of(m.ctype[c], p);
    const i64* nodes = &m.c2n[m.c2n_ptr[c]];
    for (int e = 0; e < r.t.ne; ++e) {
      const int q = pe[eid[c * 4 + e]];
      if (q < p) eord[c * 4 + e] = (uint8_t)q;
    }
    int32_t* slot = &dofs[(size_t)c * nl];
    for (int i = 0; i < r.nd; ++i) {
      if (r.kind[i] == 0) {
        int32_t& id = vid[nodes[r.ent[i]]];
        if (id < 0) id = (int32_t)next++;
        slot[i] = id;
      } else if (r.kind[i] == 1) {
        const int e = r.ent[i], k = r.k[i];
        const int32_t E = eid[c * 4 + e];
        const int q = pe[E];
        if (k > q - 1) continue;
        const int j = efirst[E] == nodes[r.t.enode[e][0]] ? k : q - k;
        int32_t& id = pool[eoff[E] + j - 1];
        if (id < 0) id = (int32_t)next++;
        slot[i] = id;
      } else {
        slot[i] = (int32_t)next++;
      }
      if (next >= (i64)INT32_MAX) { err = "too many dofs"; return -1; }
    }
  }
  return next;
}

i64 dofhandler(const HostMesh& m, int order, bool continuous, std::vector<i64>& iglob,
               std::vector<i64>& offset, std::string& err) {
  i64 nc = m.ncells;
  offset.assign(nc, 0);
  i64 tot = 0;
  for (i64 c = 0; c < nc; ++c) {
    offset[c] = tot;
    tot += ndofs_lagrange(shape_of(m.ctype[c]), order);
  }
  iglob.resize(tot);
  for (i64 i = 0; i < tot; ++i) iglob[i] = i + 1;
  if (continuous) {
    // vertex: one record per node (first cell's global ids)
    std::vector<i64> vfirst_ptr(m.nnodes, -1);  // index into vstore
    std::vector<i64> vstore;
    std::unordered_map<Key9, std::pair<i64, std::vector<i64>>, Key9Hash> dict_e, dict_f;
    dict_e.reserve((size_t)nc * 4);
    if (m.dim == 3) dict_f.reserve((size_t)nc * 4);
    for (i64 c = 0; c < nc; ++c) {
      int ct = m.ctype[c];
      int s = shape_of(ct);
      const i64* nodes = &m.c2n[m.c2n_ptr[c]];
      int nn = (int)(m.c2n_ptr[c + 1] - m.c2n_ptr[c]);
      i64* ig = &iglob[offset[c]];
      Lists vd = vertex_dofs(s, order);
      for (size_t v = 0; v < vd.size() && (int)v < nn; ++v) {
        if (vd[v].empty()) continue;
        i64 node = nodes[v];
        if (vfirst_ptr[node] >= 0) {
          for (size_t d = 0; d < vd[v].size(); ++d) ig[vd[v][d]] = vstore[vfirst_ptr[node] + d];
        } else {
          vfirst_ptr[node] = (i64)vstore.size();
          for (int ld : vd[v]) vstore.push_back(ig[ld]);
        }
      }
      if (dim_of_shape(s) > 1) {
        Lists el = edges_of(ct);
        Lists ed = edge_dofs(s, order);
        if (order > 0 && ed.size() != el.size() && s == TRIANGLE) {
          err = "triangle Lagrange order > 4 is not supported";
          return -1;
        }
        for (size_t e = 0; e < el.size() && e < ed.size(); ++e) {
          if (ed[e].empty()) continue;
          std::vector<i64> en;
          for (int ln : el[e]) en.push_back(nodes[ln]);
          Key9 key = make_key(en);
          auto it = dict_e.find(key);
          if (it != dict_e.end()) {
            i64 jc = it->second.first;
            const std::vector<i64>& jd = it->second.second;
            int jside = oriented_cell_side(m.ctype[jc], &m.c2n[m.c2n_ptr[jc]], en);
            for (size_t d = 0; d < jd.size(); ++d)
              ig[ed[e][d]] = jside > 0 ? jd[d] : jd[jd.size() - 1 - d];
          } else {
            std::vector<i64> g;
            for (int ld : ed[e]) g.push_back(ig[ld]);
            dict_e.emplace(key, std::make_pair(c, std::move(g)));
          }
        }
      }
      if (dim_of_shape(s) > 2) {
        Lists fl = faces_of(ct);
        Lists fd = face_dofs(s, order);
        for (size_t f = 0; f < fl.size() && f < fd.size(); ++f) {
          if (fd[f].empty()) continue;
          std::vector<i64> fn;
          for (int ln : fl[f]) fn.push_back(nodes[ln]);
          Key9 key = make_key(fn);
          auto it = dict_f.find(key);
          if (it != dict_f.end()) {
            const std::vector<i64>& jd = it->second.second;
            if (jd.size() / 12 > 0) {
              err = "continuous hexa Lagrange order >= 4: face-dof permutation not supported";
              return -1;
            }
            for (size_t d = 0; d < jd.size(); ++d) ig[fd[f][d]] = jd[d];
          } else {
            std::vector<i64> g;
            for (int ld : fd[f]) g.push_back(ig[ld]);
            dict_f.emplace(key, std::make_pair(c, std::move(g)));
          }
        }
      }
    }
  }
  // densify: first-appearance renumbering (utils.jl:7-16)
  std::vector<i64> remap(tot + 1, 0);
  i64 next = 0;
  for (i64 i = 0; i < tot; ++i) {
    i64 v = iglob[i];
    if (remap[v] == 0) remap[v] = ++next;
    iglob[i] = remap[v];
  }
  return next;
}

// ---------------------------------------------------------------------------
// Cell colouring: greedy, cells in index order, smallest colour not used by a
// cell sharing a row dof.  On a structured hexa grid this is the 8-colour
// parity pattern.
// ---------------------------------------------------------------------------
int colour_cells(const std::vector<int32_t>& rdofs, int nl, i64 nrows, i64 ncells,
                 std::vector<int32_t>& color_ptr, std::vector<int32_t>& cells_by_color) {
  std::vector<i64> aptr(nrows + 1, 0);
  for (i64 c = 0; c < ncells; ++c)
    for (int i = 0; i < nl; ++i)
      if (rdofs[c * nl + i] >= 0) aptr[rdofs[c * nl + i] + 1]++;
  for (i64 r = 0; r < nrows; ++r) aptr[r + 1] += aptr[r];
  std::vector<int32_t> adj(aptr[nrows]);
  std::vector<i64> fill(aptr.begin(), aptr.end() - 1);
  for (i64 c = 0; c < ncells; ++c)
    for (int i = 0; i < nl; ++i)
      if (rdofs[c * nl + i] >= 0) adj[fill[rdofs[c * nl + i]]++] = (int32_t)c;
  std::vector<int32_t> color(ncells, -1);
  int ncol = 0;
  std::vector<uint8_t> used(256, 0);
  for (i64 c = 0; c < ncells; ++c) {
    std::fill(used.begin(), used.begin() + std::max(ncol, 1), 0);
    for (int i = 0; i < nl; ++i) {
      i64 r = rdofs[c * nl + i];
      if (r < 0) continue;
      for (i64 p = aptr[r]; p < aptr[r + 1]; ++p) {
        int32_t o = adj[p];
        if (o >= c) break;  // adjacency lists are in cell order
        used[color[o]] = 1;
      }
    }
    int k = 0;
    while (k < 255 && used[k]) ++k;
    if (k >= 255) return -1;
    color[c] = k;
    ncol = std::max(ncol, k + 1);
  }
  color_ptr.assign(ncol + 1, 0);
  for (i64 c = 0; c < ncells; ++c) color_ptr[color[c] + 1]++;
  for (int k = 0; k < ncol; ++k) color_ptr[k + 1] += color_ptr[k];
  cells_by_color.resize(ncells);
  std::vector<int32_t> pos(color_ptr.begin(), color_ptr.end() - 1);
  for (i64 c = 0; c < ncells; ++c) cells_by_color[pos[color[c]]++] = (int32_t)c;
  return ncol;
}

// ---------------------------------------------------------------------------
// CSR pattern of the element dof-pair graph, rows = test dofs, columns = trial
// dofs, columns sorted and unique per row (the structure sparse() produces,
// assembler.jl:185), plus the row -> element-row gather lists.
// ---------------------------------------------------------------------------
bool build_pattern(const std::vector<int32_t>& rdofs, int nl_row, i64 nrows,
                   const std::vector<int32_t>& cdofs, int nl_col, i64 ncols, i64 ncells,
                   HostPattern& p, std::string& err) {
  // dof lists are nl per cell; -1 marks a padding slot (cells with fewer dofs
  // than the stride on mixed meshes)
  p.nrows = nrows; p.ncols = ncols; p.nl_row = nl_row; p.nl_col = nl_col;
  for (i64 k = 0; k < ncells * nl_row; ++k)
    if (rdofs[k] < -1 || rdofs[k] >= nrows) { err = "row dof out of range"; return false; }
  for (i64 k = 0; k < ncells * nl_col; ++k)
    if (cdofs[k] < -1 || cdofs[k] >= ncols) { err = "column dof out of range"; return false; }
  // row -> cells adjacency
  std::vector<i64> aptr(nrows + 1, 0);
  for (i64 c = 0; c < ncells; ++c)
    for (int i = 0; i < nl_row; ++i)
      if (rdofs[c * nl_row + i] >= 0) aptr[rdofs[c * nl_row + i] + 1]++;
  for (i64 r = 0; r < nrows; ++r) aptr[r + 1] += aptr[r];
  std::vector<int32_t> adj(aptr[nrows]);
  {
    std::vector<i64> fill(aptr.begin(), aptr.end() - 1);
    for (i64 c = 0; c < ncells; ++c)
      for (int i = 0; i < nl_row; ++i)
        if (rdofs[c * nl_row + i] >= 0) adj[fill[rdofs[c * nl_row + i]]++] = (int32_t)c;
  }
  auto row_columns = [&](i64 r, std::vector<int32_t>& buf) {
    buf.clear();
    for (i64 q = aptr[r]; q < aptr[r + 1]; ++q) {
      i64 c = adj[q];
      for (int j = 0; j < nl_col; ++j)
        if (cdofs[c * nl_col + j] >= 0) buf.push_back(cdofs[c * nl_col + j]);
    }
    std::sort(buf.begin(), buf.end());
    return (i64)(std::unique(buf.begin(), buf.end()) - buf.begin());
  };
  // pass 1: row lengths
  std::vector<i64> len(nrows, 0);
  parallel_for(nrows, [&](i64 b, i64 e) {
    std::vector<int32_t> buf;
    for (i64 r = b; r < e; ++r) len[r] = row_columns(r, buf);
  });
  p.rowptr.assign(nrows + 1, 0);
  for (i64 r = 0; r < nrows; ++r) p.rowptr[r + 1] = p.rowptr[r] + len[r];
  p.nnz = p.rowptr[nrows];
  if (p.nnz >= (i64)INT32_MAX) { err = "nnz exceeds 2^31-1"; return false; }
  p.colval.resize(p.nnz);
  p.max_row = 0;
  for (i64 r = 0; r < nrows; ++r) p.max_row = std::max<int>(p.max_row, (int)len[r]);
  parallel_for(nrows, [&](i64 b, i64 e) {
    std::vector<int32_t> buf;
    for (i64 r = b; r < e; ++r) {
      i64 n = row_columns(r, buf);
      std::copy(buf.begin(), buf.begin() + n, p.colval.begin() + p.rowptr[r]);
    }
  });
  // gather structure: row -> (cell, i) in cell order, and per-(cell, i) column offsets
  if (ncells * nl_row >= (i64)INT32_MAX) { err = "ncells * nl exceeds 2^31-1"; return false; }
  p.pptr.assign(nrows + 1, 0);
  for (i64 r = 0; r <= nrows; ++r) p.pptr[r] = (int32_t)aptr[r];
  p.pairs.resize(aptr[nrows]);
  parallel_for(nrows, [&](i64 b, i64 e) {
    for (i64 r = b; r < e; ++r)
      for (i64 q = aptr[r]; q < aptr[r + 1]; ++q) {
        i64 c = adj[q];
        int li = -1;
        for (int i = 0; i < nl_row; ++i)
          if (rdofs[c * nl_row + i] == r) { li = i; break; }
        p.pairs[q] = (int32_t)(c * nl_row + li);
      }
  });
  p.max_adj = 0;
  for (i64 r = 0; r < nrows; ++r) p.max_adj = std::max<int>(p.max_adj, p.pptr[r + 1] - p.pptr[r]);
  if (p.max_adj <= 16) {
    p.pad = p.max_adj <= 4 ? 4 : (p.max_adj <= 8 ? 8 : 16);
    p.pairs_padded.assign((size_t)nrows * p.pad, -1);
    for (i64 r = 0; r < nrows; ++r)
      for (int q = p.pptr[r]; q < p.pptr[r + 1]; ++q) p.pairs_padded[r * p.pad + (q - p.pptr[r])] = p.pairs[q];
  }
  // offsets: the largest value of the type marks a padding column (never a
  // valid offset: rows are shorter than that)
  if (p.max_row <= 255) p.offset_bytes = 1;
  else if (p.max_row <= 65535) p.offset_bytes = 2;
  else { err = "row longer than 65535 entries"; return false; }
  i64 per = (i64)nl_row * nl_col;
  p.goff.assign((size_t)(ncells * per * p.offset_bytes), 0);
  parallel_for(ncells, [&](i64 b, i64 e) {
    for (i64 c = b; c < e; ++c)
      for (int i = 0; i < nl_row; ++i) {
        i64 r = rdofs[c * nl_row + i];
        for (int j = 0; j < nl_col; ++j) {
          size_t idx = (size_t)(c * per + (i64)i * nl_col + j);
          i64 cd = cdofs[c * nl_col + j];
          i64 off;
          if (r < 0 || cd < 0) {
            off = p.offset_bytes == 1 ? 0xFF : 0xFFFF;
          } else {
            const int32_t* row = &p.colval[p.rowptr[r]];
            i64 rl = p.rowptr[r + 1] - p.rowptr[r];
            off = std::lower_bound(row, row + rl, (int32_t)cd) - row;
          }
          if (p.offset_bytes == 1) {
            p.goff[idx] = (uint8_t)off;
          } else {
            uint16_t v = (uint16_t)off;
            std::memcpy(&p.goff[idx * 2], &v, 2);
          }
        }
      }
  });
  return true;
}

// ---------------------------------------------------------------------------
// DG faces (_build_faces! mesh.jl:328-374, inner_faces mesh.jl:552-555,
// FaceInfo/cell_side domain.jl:555-593 + entity.jl:478-483, side-p
// permutation eval_point.jl:174-206, mapping_face mapping.jl:324-342).
// ---------------------------------------------------------------------------
namespace {

// Reference cube: vertex coordinates and the vertex lists of its faces
// (shape.jl:212-287).
const int kCubeX[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                          {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
const int kCubeF[6][4] = {{0, 3, 2, 1}, {0, 1, 5, 4}, {1, 2, 6, 5},
                          {2, 3, 7, 6}, {0, 4, 7, 3}, {4, 5, 6, 7}};
const int kSideDim[6] = {2, 1, 0, 1, 0, 2};

// Code of the affine face map (s,t) -> cell reference point built from the
// Quad4 map of vertices kCubeF[side][perm[k]]; -1 if it is not a symmetry of
// the square.
int face_map_code(int side, const int perm[4]) {
  int X[4][3];
  for (int k = 0; k < 4; ++k)
    for (int a = 0; a < 3; ++a) X[k][a] = kCubeX[kCubeF[side][perm[k]]][a];
  int ds[3], dt[3], sk[3], ce[3];
  for (int a = 0; a < 3; ++a) {
    ds[a] = -X[0][a] + X[1][a] + X[2][a] - X[3][a];  // 4 * d xi / ds
    dt[a] = -X[0][a] - X[1][a] + X[2][a] + X[3][a];  // 4 * d xi / dt
    sk[a] = X[0][a] - X[1][a] + X[2][a] - X[3][a];   // bilinear part
    ce[a] = X[0][a] + X[1][a] + X[2][a] + X[3][a];
  }
  int d = kSideDim[side];
  int fa = d == 0 ? 1 : 0, fb = d == 2 ? 1 : 2;
  for (int a = 0; a < 3; ++a)
    if (sk[a] != 0) return -1;
  if (ds[d] != 0 || dt[d] != 0 || ce[fa] != 0 || ce[fb] != 0) return -1;
  if (std::abs(ds[fa]) == 4 && std::abs(dt[fb]) == 4)
    return 0 | (ds[fa] < 0 ? 2 : 0) | (dt[fb] < 0 ? 4 : 0);
  if (std::abs(dt[fa]) == 4 && std::abs(ds[fb]) == 4)
    return 1 | (dt[fa] < 0 ? 2 : 0) | (ds[fb] < 0 ? 4 : 0);
  return -1;
}

struct Key4 {
  i64 v[4];
  bool operator==(const Key4& o) const {
    return v[0] == o.v[0] && v[1] == o.v[1] && v[2] == o.v[2] && v[3] == o.v[3];
  }
};
struct Key4Hash {
  size_t operator()(const Key4& k) const {
    uint64_t h = 1469598103934665603ull;
    for (int i = 0; i < 4; ++i) h = (h ^ (uint64_t)k.v[i]) * 1099511628211ull;
    return (size_t)h;
  }
};

}  // namespace

bool build_dg_faces(const HostMesh& m, HostDGFaces& F, std::string& err) {
  if (!m.uniform || (m.celltype != BH_HEXA8 && m.celltype != BH_HEXA27)) {
    err = "DG faces need a uniform Hexa8 or Hexa27 mesh";
    return false;
  }
  i64 nc = m.ncells;
  if (nc > INT32_MAX / 6) { err = "too many cells"; return false; }
  // all faces in creation order: (creator, side), then the second cell
  std::unordered_map<Key4, int32_t, Key4Hash> dict;
  dict.reserve((size_t)nc * 4);
  std::vector<int32_t> fc_n, fc_p;
  std::vector<int8_t> fs_n, fs_p;
  for (i64 c = 0; c < nc; ++c) {
    const i64* cn = &m.c2n[m.c2n_ptr[c]];
    for (int s = 0; s < 6; ++s) {
      Key4 k;
      for (int j = 0; j < 4; ++j) k.v[j] = cn[kCubeF[s][j]];
      std::sort(k.v, k.v + 4);
      auto it = dict.find(k);
      if (it == dict.end()) {
        dict.emplace(k, (int32_t)fc_n.size());
        fc_n.push_back((int32_t)c);
        fs_n.push_back((int8_t)s);
        fc_p.push_back(-1);
        fs_p.push_back(-1);
      } else {
        int32_t f = it->second;
        if (fc_p[f] >= 0) { err = "a face is shared by more than two cells"; return false; }
        fc_p[f] = (int32_t)c;
        fs_p[f] = (int8_t)s;
      }
    }
  }
  F.cell_n.clear(); F.cell_p.clear(); F.side_n.clear(); F.side_p.clear();
  F.cs.assign((size_t)nc * 6 * 4, 0);
  for (i64 k = 0; k < nc * 6; ++k) F.cs[k * 4] = -1;
  const int ident[4] = {0, 1, 2, 3};
  for (size_t f = 0; f < fc_n.size(); ++f) {
    if (fc_p[f] < 0) continue;  // boundary face: not in the interior-face domain
    int32_t cn = fc_n[f], cp = fc_p[f];
    int sn = fs_n[f], sp = fs_p[f];
    const i64* nn = &m.c2n[m.c2n_ptr[cn]];
    const i64* pn = &m.c2n[m.c2n_ptr[cp]];
    // permut[k] = position in side n's vertex list of side p's k-th vertex
    int perm[4];
    for (int k = 0; k < 4; ++k) {
      perm[k] = -1;
      for (int q = 0; q < 4; ++q)
        if (nn[kCubeF[sn][q]] == pn[kCubeF[sp][k]]) perm[k] = q;
      if (perm[k] < 0) { err = "non-conforming face"; return false; }
    }
    int code_n = face_map_code(sn, ident), code_p = face_map_code(sp, perm);
    if (code_n < 0 || code_p < 0) { err = "face orientation is not a square symmetry"; return false; }
    int32_t id = (int32_t)F.cell_n.size();
    F.cell_n.push_back(cn); F.cell_p.push_back(cp);
    F.side_n.push_back((int8_t)sn); F.side_p.push_back((int8_t)sp);
    int32_t* rn = &F.cs[((size_t)cn * 6 + sn) * 4];
    int32_t* rp = &F.cs[((size_t)cp * 6 + sp) * 4];
    rn[0] = id; rn[1] = cp; rn[2] = 0 | (code_n << 1) | (sp << 4) | (code_p << 7);
    rp[0] = id; rp[1] = cn; rp[2] = 1 | (code_p << 1) | (sn << 4) | (code_n << 7);
  }
  F.nfaces = (i64)F.cell_n.size();
  return true;
}

}  // namespace bh

// ---------------------------------------------------------------------------
// Schedule of the fused element + assembly kernel (see bh_internal.h).
// ---------------------------------------------------------------------------
namespace bh {

bool build_fused_schedule(const HostMesh& m, int nl, int64_t nrows,
                          const std::vector<int64_t>& rowptr, const std::vector<int32_t>& pptr,
                          const std::vector<int32_t>& pairs, const std::vector<uint8_t>& goff,
                          int prog_cap, int min_blocks, FusedSchedule& s, std::string& err) {
  const int d = m.dim;
  const int64_t nc = m.ncells;
  if (nc <= 0 || nl <= 0 || nl > 32 || (d != 2 && d != 3)) {
    err = "fused schedule: unsupported mesh";
    return false;
  }
  if ((int64_t)pptr.size() != nrows + 1 || (int64_t)rowptr.size() != nrows + 1 ||
      (int64_t)goff.size() < nc * nl * nl) {
    err = "fused schedule: inconsistent pattern arrays";
    return false;
  }
  // 1. cell centroids and a cell size estimate from the bounding box
  std::vector<double> cen((size_t)nc * d, 0.0);
  parallel_for(nc, [&](int64_t b, int64_t e) {
    for (int64_t c = b; c < e; ++c) {
      const int64_t k0 = m.c2n_ptr[c], k1 = m.c2n_ptr[c + 1];
      for (int64_t k = k0; k < k1; ++k)
        for (int a = 0; a < d; ++a) cen[c * d + a] += m.coords[m.c2n[k] * d + a];
      for (int a = 0; a < d; ++a) cen[c * d + a] /= (double)(k1 - k0);
    }
  });
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int64_t c = 0; c < nc; ++c)
    for (int a = 0; a < d; ++a) {
      lo[a] = std::min(lo[a], cen[c * d + a]);
      hi[a] = std::max(hi[a], cen[c * d + a]);
    }
  double vol = 1.0;
  double ext[3];
  for (int a = 0; a < d; ++a) {
    ext[a] = std::max(hi[a] - lo[a], 1e-300);
    vol *= ext[a];
  }
  // centroids span (n - 1) cells per axis on a structured box
  const double h = std::pow(vol / (double)nc, 1.0 / d);
  // 2. pencils: tiles of the first d-1 axes, layers along the last axis
  const double tw = d == 3 ? 4.0 : 16.0;  // tile width in cells
  int nt[2] = {1, 1};
  for (int a = 0; a < d - 1; ++a)
    nt[a] = std::max(1, (int)std::floor((ext[a] + h) / (tw * h) + 0.5));
  auto tile_of = [&](int64_t c, int a) {
    const double u = (cen[c * d + a] - lo[a] + 0.5 * h) / (ext[a] + h);
    return std::min(nt[a] - 1, std::max(0, (int)std::floor(u * nt[a])));
  };
  const int la = d - 1;  // layer axis
  auto layer_of = [&](int64_t c) { return (int64_t)std::floor((cen[c * d + la] - lo[la]) / h + 0.5); };
  std::vector<int64_t> order(nc);
  std::vector<int64_t> pen(nc), lay(nc);
  for (int64_t c = 0; c < nc; ++c) {
    order[c] = c;
    pen[c] = d == 3 ? (int64_t)tile_of(c, 1) * nt[0] + tile_of(c, 0) : tile_of(c, 0);
    lay[c] = layer_of(c);
  }
  std::sort(order.begin(), order.end(), [&](int64_t x, int64_t y) {
    if (pen[x] != pen[y]) return pen[x] < pen[y];
    if (lay[x] != lay[y]) return lay[x] < lay[y];
    for (int a = d - 2; a >= 0; --a)
      if (cen[x * d + a] != cen[y * d + a]) return cen[x * d + a] < cen[y * d + a];
    return x < y;
  });
  // 3. steps of 16 cells per pencil; with fewer pencils than min_blocks each
  //    pencil is cut into runs of at least 8 steps, one workgroup per run
  //    (rows across a cut are spill rows)
  s = FusedSchedule();
  s.nl = nl;
  std::vector<int64_t> pstart;
  for (int64_t k = 0; k < nc; ++k)
    if (k == 0 || pen[order[k]] != pen[order[k - 1]]) pstart.push_back(k);
  pstart.push_back(nc);
  const int64_t npen = (int64_t)pstart.size() - 1;
  const int64_t cuts = std::max<int64_t>(1, (min_blocks + npen - 1) / std::max<int64_t>(npen, 1));
  std::vector<int32_t> cstep(nc), clane(nc), cblk(nc);
  s.bstep.push_back(0);
  int64_t nsteps = 0;
  for (int64_t pp = 0; pp < npen; ++pp) {
    const int64_t k = pstart[pp], e = pstart[pp + 1];
    const int64_t psteps = (e - k + 15) / 16;
    const int64_t nrun = std::max<int64_t>(1, std::min(cuts, psteps / 8));  // runs of >= 8 steps
    for (int64_t ru = 0; ru < nrun; ++ru) {
      const int64_t q0 = k + 16 * (psteps * ru / nrun), q1 = std::min(e, k + 16 * (psteps * (ru + 1) / nrun));
      for (int64_t q = q0; q < q1; q += 16) {
        for (int l = 0; l < 16; ++l) {
          if (q + l < q1) {
            const int64_t c = order[q + l];
            cstep[c] = (int32_t)nsteps;
            clane[c] = l;
            cblk[c] = s.nblocks;
            s.cells.push_back((int32_t)c);
          } else {
            s.cells.push_back(-1);
          }
        }
        ++nsteps;
      }
      ++s.nblocks;
      s.bstep.push_back((int32_t)nsteps);
    }
  }
  // 4. local rows: all element rows in one block, within two consecutive steps
  std::vector<int32_t> row_step(nrows, -1);
  parallel_for(nrows, [&](int64_t b, int64_t e) {
    for (int64_t r = b; r < e; ++r) {
      const int32_t p0 = pptr[r], p1 = pptr[r + 1];
      if (p1 <= p0 || rowptr[r + 1] - rowptr[r] > 128) continue;
      const int64_t c0 = pairs[p0] / nl;
      int32_t smin = cstep[c0], smax = cstep[c0];
      bool ok = true;
      for (int32_t q = p0; q < p1 && ok; ++q) {
        const int64_t c = pairs[q] / nl;
        ok = cblk[c] == cblk[c0];
        smin = std::min(smin, cstep[c]);
        smax = std::max(smax, cstep[c]);
      }
      if (ok && smax - smin <= 1) row_step[r] = smax;
    }
  });
  // 5. step programs (rows ascending); rows that do not fit spill.  Layout:
  //    header {u32 rows, u32 pairs, 8 bytes 0}, then per row 32 bytes
  //    {i64 first value, u16 length, u16 element rows, u32 first pair, u16
  //    codes[8]}, then per pair 32 column offsets (0xFF: none); code =
  //    (step - cell's step) * 16 + cell's lane) * nl + element row
  std::vector<std::vector<int32_t>> rows_of(nsteps);
  for (int64_t r = 0; r < nrows; ++r)
    if (row_step[r] >= 0) {
      if (pptr[r + 1] - pptr[r] > kFusedRowCodes) row_step[r] = -1;
      else rows_of[row_step[r]].push_back((int32_t)r);
    }
  s.prog_off.assign(nsteps + 1, 0);
  std::vector<int64_t> psize(nsteps, 0);
  for (int64_t st = 0; st < nsteps; ++st) {
    int64_t nr = 0, np = 0;
    for (int32_t r : rows_of[st]) {
      const int64_t np2 = np + (pptr[r + 1] - pptr[r]);
      const int64_t sz = 16 + 32 * (nr + 1) + 32 * np2;
      if (sz > prog_cap) {
        row_step[r] = -1;  // spills
        continue;
      }
      ++nr;
      np = np2;
    }
    psize[st] = nr ? 16 + 32 * nr + 32 * np : 0;
    s.prog_off[st + 1] = s.prog_off[st] + psize[st];
  }
  s.prog.assign((size_t)s.prog_off[nsteps], 0);
  parallel_for(nsteps, [&](int64_t b, int64_t e) {
    for (int64_t st = b; st < e; ++st) {
      if (!psize[st]) continue;
      uint8_t* P = s.prog.data() + s.prog_off[st];
      uint32_t nr = 0, np = 0;
      for (int32_t r : rows_of[st])
        if (row_step[r] == st) { ++nr; np += (uint32_t)(pptr[r + 1] - pptr[r]); }
      std::memcpy(P, &nr, 4);
      std::memcpy(P + 4, &np, 4);
      uint8_t* R = P + 16;
      uint8_t* G = P + 16 + 32 * (int64_t)nr;
      std::memset(G, 0xFF, (size_t)32 * np);
      uint32_t k = 0, q = 0;
      for (int32_t r : rows_of[st]) {
        if (row_step[r] != st) continue;
        const int64_t vb = rowptr[r];
        const uint16_t len = (uint16_t)(rowptr[r + 1] - rowptr[r]);
        const uint16_t cnt = (uint16_t)(pptr[r + 1] - pptr[r]);
        uint8_t* rec = R + 32 * (int64_t)k;
        std::memcpy(rec, &vb, 8);
        std::memcpy(rec + 8, &len, 2);
        std::memcpy(rec + 10, &cnt, 2);
        std::memcpy(rec + 12, &q, 4);
        int u = 0;
        for (int32_t pq = pptr[r]; pq < pptr[r + 1]; ++pq, ++q, ++u) {
          const int64_t c = pairs[pq] / nl;
          const int i = (int)(pairs[pq] - c * nl);
          const int rel = (int)(st - cstep[c]);
          const uint16_t code = (uint16_t)((rel * 16 + clane[c]) * nl + i);
          std::memcpy(rec + 16 + 2 * u, &code, 2);
          std::memcpy(G + 32 * (int64_t)q, &goff[(size_t)pairs[pq] * nl], (size_t)nl);
        }
        ++k;
      }
    }
  });
  for (int64_t st = 0; st < nsteps; ++st) s.max_prog = std::max<int64_t>(s.max_prog, psize[st]);
  // 6. spill rows and the element rows they read
  s.spill_mask.assign(nc, 0u);
  for (int64_t r = 0; r < nrows; ++r) {
    if (row_step[r] >= 0) {
      ++s.local_rows;
      s.local_pairs += pptr[r + 1] - pptr[r];
      continue;
    }
    s.spill_rows.push_back((int32_t)r);
    for (int32_t q = pptr[r]; q < pptr[r + 1]; ++q) {
      const int64_t c = pairs[q] / nl;
      s.spill_mask[c] |= 1u << (pairs[q] - c * nl);
    }
  }
  return true;
}


bool build_sf_structure(const std::vector<int32_t>& dofs, i64 ncells, i64 ndofs,
                        const std::vector<i64>& rowptr, const std::vector<int32_t>& colval,
                        SfStructure& S, std::string& why) {
  S = SfStructure();
  if (ncells < 1 || (i64)dofs.size() != ncells * 27 || (i64)rowptr.size() != ndofs + 1) {
    why = "not a Q2 hexahedral space";
    return false;
  }
  auto f = [&](i64 c, int i1, int i2, int i3) { return dofs[c * 27 + i1 + 3 * i2 + 9 * i3]; };
  auto xadj = [&](i64 a, i64 b) {
    for (int i3 = 0; i3 < 3; ++i3)
      for (int i2 = 0; i2 < 3; ++i2)
        if (f(a, 2, i2, i3) != f(b, 0, i2, i3)) return false;
    return true;
  };
  auto yadj = [&](i64 a, i64 b) {
    for (int i3 = 0; i3 < 3; ++i3)
      for (int i1 = 0; i1 < 3; ++i1)
        if (f(a, i1, 2, i3) != f(b, i1, 0, i3)) return false;
    return true;
  };
  i64 nx = 1;
  while (nx < ncells && xadj(nx - 1, nx)) ++nx;
  i64 ny = 1;
  while ((ny + 1) * nx <= ncells && yadj((ny - 1) * nx, ny * nx)) ++ny;
  if (ncells % (nx * ny) != 0) {
    why = "cells do not form an nx x ny x nz grid";
    return false;
  }
  const i64 nz = ncells / (nx * ny);
  const i64 LX = 2 * nx + 1, LY = 2 * ny + 1, LZ = 2 * nz + 1;
  if (nx > 32767 || ny > 32767 || nz > 32767 || LX * LY * LZ != ndofs ||
      LX * LY * LZ > (i64)INT32_MAX) {
    why = "the dofs are not the grid's Q2 lattice";
    return false;
  }
  S.lat2dof.assign((size_t)ndofs, -1);
  std::vector<i64> dof2lat((size_t)ndofs, -1);
  for (i64 cz = 0, c = 0; cz < nz; ++cz)
    for (i64 cy = 0; cy < ny; ++cy)
      for (i64 cx = 0; cx < nx; ++cx, ++c)
        for (int i3 = 0; i3 < 3; ++i3)
          for (int i2 = 0; i2 < 3; ++i2)
            for (int i1 = 0; i1 < 3; ++i1) {
              const i64 L = ((2 * cz + i3) * LY + 2 * cy + i2) * LX + 2 * cx + i1;
              const int32_t d = f(c, i1, i2, i3);
              if (d < 0 || d >= ndofs) {
                why = "dof id out of range";
                return false;
              }
              if (S.lat2dof[L] < 0) {
                if (dof2lat[d] >= 0) {
                  why = "a dof sits on two lattice points";
                  return false;
                }
                S.lat2dof[L] = d;
                dof2lat[d] = L;
              } else if (S.lat2dof[L] != d) {
                why = "neighbouring cells disagree on a shared dof";
                return false;
              }
            }
  // row templates: per row, the box entry of each CSR position; rows hashed in
  // parallel, then one id per distinct template, then every row checked
  // against its template's bytes (a hash collision disqualifies the space)
  auto row_key = [&](i64 r, uint8_t* key, int* len) -> bool {
    const i64 L = dof2lat[r];
    const i64 X = L % LX, Y = (L / LX) % LY, Z = L / (LX * LY);
    const int ex = (X & 1) ? 3 : 5, ey = (Y & 1) ? 3 : 5, ez = (Z & 1) ? 3 : 5;
    const int hx = ex / 2, hy = ey / 2, hz = ez / 2;
    auto span = [](i64 p, int h, i64 n) {
      return (i64)(std::min(p + h, n - 1) - std::max<i64>(p - h, 0) + 1);
    };
    const i64 n = rowptr[r + 1] - rowptr[r];
    if (n != span(X, hx, LX) * span(Y, hy, LY) * span(Z, hz, LZ) || n > 125) return false;
    for (i64 k = 0; k < n; ++k) {
      const i64 Lc = dof2lat[colval[rowptr[r] + k]];
      const i64 dx = Lc % LX - X, dy = (Lc / LX) % LY - Y, dz = Lc / (LX * LY) - Z;
      if (dx < -hx || dx > hx || dy < -hy || dy > hy || dz < -hz || dz > hz) return false;
      key[k] = (uint8_t)(((dz + hz) * ey + dy + hy) * ex + dx + hx);
    }
    *len = (int)n;
    return true;
  };
  auto hash = [](const uint8_t* key, int len) {
    uint64_t h = 1469598103934665603ull ^ (uint64_t)len;
    for (int k = 0; k < len; ++k) h = (h ^ key[k]) * 1099511628211ull;
    return h;
  };
  std::vector<uint64_t> rh((size_t)ndofs);
  std::atomic<bool> fail_rows{false};
  parallel_for(ndofs, [&](i64 b, i64 e) {
    uint8_t key[128];
    for (i64 r = b; r < e; ++r) {
      int len = 0;
      if (!row_key(r, key, &len)) {
        fail_rows = true;
        return;
      }
      rh[r] = hash(key, len);
    }
  });
  if (fail_rows) {
    why = "a CSR row is not the box of its lattice point";
    return false;
  }
  std::unordered_map<uint64_t, int> ids;
  S.tplid.resize((size_t)ndofs);
  for (i64 r = 0; r < ndofs; ++r) {
    auto it = ids.find(rh[r]);
    if (it == ids.end()) {
      if (S.ntpl >= 65535) {
        why = "too many row shapes";
        return false;
      }
      uint8_t key[128] = {0};
      int len = 0;
      row_key(r, key, &len);
      S.tpl.insert(S.tpl.end(), key, key + 128);
      it = ids.emplace(rh[r], S.ntpl++).first;
    }
    S.tplid[r] = (uint16_t)it->second;
  }
  parallel_for(ndofs, [&](i64 b, i64 e) {
    uint8_t key[128];
    for (i64 r = b; r < e; ++r) {
      int len = 0;
      row_key(r, key, &len);
      if (std::memcmp(key, &S.tpl[(size_t)S.tplid[r] * 128], (size_t)len) != 0) {
        fail_rows = true;
        return;
      }
    }
  });
  if (fail_rows) {
    why = "row template hash collision";
    return false;
  }
  S.nx = (int)nx;
  S.ny = (int)ny;
  S.nz = (int)nz;
  return true;
}


// ---------------------------------------------------------------------------
// Offset templates of the row gather (see bh_internal.h).  Each thread keeps
// the distinct vectors of its run of element rows (by a 64-bit FNV-1a hash,
// bytes compared on a hit), the runs are merged in thread order, then every
// element row gets its template id.
int offset_templates(const uint8_t* goff, int nl, i64 nrow_el, const int32_t* pairs_padded,
                     i64 nslots, std::vector<uint8_t>& tpl, std::vector<uint32_t>& words,
                     int& shift) {
  if (nrow_el >= ((i64)1 << 31) || nl <= 0) return -1;
  auto hash = [nl](const uint8_t* k) {
    uint64_t h = 1469598103934665603ull;
    for (int j = 0; j < nl; ++j) h = (h ^ k[j]) * 1099511628211ull;
    return h;
  };
  constexpr int kMax = 255;
  const int T = std::max(1, std::min(64, (int)(nrow_el / 65536) + 1));
  std::vector<std::vector<i64>> firsts(T);  // per run: element rows that first showed a vector
  std::vector<char> over(T, 0);
  parallel_for(T, [&](i64 b, i64 e) {
    for (i64 t = b; t < e; ++t) {
      const i64 r0 = nrow_el * t / T, r1 = nrow_el * (t + 1) / T;
      std::unordered_map<uint64_t, i64> seen;
      for (i64 r = r0; r < r1 && !over[t]; ++r) {
        const uint8_t* k = goff + r * nl;
        auto it = seen.find(hash(k));
        if (it == seen.end()) {
          seen.emplace(hash(k), r);
          firsts[t].push_back(r);
          if ((int)firsts[t].size() > kMax) over[t] = 1;
        } else if (std::memcmp(goff + it->second * nl, k, nl) != 0) {
          over[t] = 1;  // a hash collision: no templates
        }
      }
    }
  });
  for (int t = 0; t < T; ++t)
    if (over[t]) return -1;
  std::unordered_map<uint64_t, int> id;
  tpl.clear();
  for (int t = 0; t < T; ++t)
    for (i64 r : firsts[t]) {
      const uint8_t* k = goff + r * nl;
      auto it = id.find(hash(k));
      if (it != id.end()) {
        if (std::memcmp(&tpl[(size_t)it->second * nl], k, nl) != 0) return -1;
        continue;
      }
      if ((int)id.size() == kMax) return -1;
      id.emplace(hash(k), (int)id.size());
      tpl.insert(tpl.end(), k, k + nl);
    }
  // the template id takes the top bits, as few as its count needs
  int tbits = 1;
  while ((1 << tbits) < (int)id.size()) ++tbits;
  shift = 32 - tbits;
  if (nrow_el >= ((i64)1 << shift)) return -1;
  std::vector<uint8_t> rid((size_t)nrow_el);
  parallel_for(nrow_el, [&](i64 b, i64 e) {
    for (i64 r = b; r < e; ++r) rid[r] = (uint8_t)id.find(hash(goff + r * nl))->second;
  });
  words.assign((size_t)nslots, 0xFFFFFFFFu);
  std::atomic<bool> bad{false};
  parallel_for(nslots, [&](i64 b, i64 e) {
    for (i64 q = b; q < e; ++q) {
      const int32_t v = pairs_padded[q];
      if (v < 0) continue;
      if ((i64)v >= nrow_el) { bad.store(true); continue; }
      words[q] = (uint32_t)v | ((uint32_t)rid[v] << shift);
    }
  });
  return bad.load() ? -1 : (int)id.size();
}

}  // namespace bh
