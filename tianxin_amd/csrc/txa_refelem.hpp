// Reference-element tables (basis values / gradients at the integration points,
// cubature), computed on the host in FP64 and uploaded once per topology.
//
// Bases: Intrepid2 HGRAD_HEX_C1, HGRAD_TET_C1, HGRAD_WEDGE_C1 and HGRAD_HEX_C2 as selected by
// disc-fe/src/Panzer_IntrepidBasisFactory.hpp:159-184; basis functions are
// listed in element-DOF (subcell) order: vertices in shards order, then edge
// midpoints, face centres and the cell centre (dof-mgr/src/Panzer_FieldAggPattern.cpp:202-246).
// Cubature: tensor Gauss-Legendre (2 or 3 points per axis) on hexes, the
// 4-point degree-2 rule on tets, and on wedges the 3-point degree-2 triangle rule
// times 2-point Gauss (disc-fe/src/Panzer_IntegrationValues2.cpp:69-94).
#pragma once

#include <cmath>
#include <cstring>

namespace txa {

constexpr int kMaxNB = 27;
constexpr int kMaxNQ = 27;

struct RefElem {
  int topo = 0;
  int nb = 0;   // field basis functions
  int nv = 0;   // geometry vertices
  int nq = 0;   // integration points
  double qw[kMaxNQ];
  double N[kMaxNB][kMaxNQ];
  double dN[kMaxNB][kMaxNQ][3];
  double gdN[8][kMaxNQ][3];   // geometry (vertex) basis gradients
};

inline void refelem_build(RefElem &r, int topo) {
  std::memset(&r, 0, sizeof(r));
  r.topo = topo;
  static const int hexv[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                 {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  if (topo == 4 /* WEDGE6 */) {
    // shards Wedge<6>: triangle (0,0) (1,0) (0,1) at zeta = -1 (nodes 0-2) and +1 (nodes 3-5);
    // N = lam_i (1 -+ zeta) / 2, lam = (1 - xi - eta, xi, eta); triangle point fastest
    static const double tq[3][2] = {{1.0 / 6.0, 1.0 / 6.0}, {2.0 / 3.0, 1.0 / 6.0}, {1.0 / 6.0, 2.0 / 3.0}};
    static const double dlam[3][2] = {{-1.0, -1.0}, {1.0, 0.0}, {0.0, 1.0}};
    r.nb = 6; r.nv = 6; r.nq = 6;
    for (int k = 0; k < 2; ++k)
      for (int t = 0; t < 3; ++t) {
        const int q = 3 * k + t;
        const double x = tq[t][0], y = tq[t][1], z = (k == 0 ? -1.0 : 1.0) / std::sqrt(3.0);
        const double lam[3] = {1.0 - x - y, x, y};
        r.qw[q] = 1.0 / 6.0;
        for (int l = 0; l < 2; ++l) {
          const double h = l == 0 ? 0.5 * (1.0 - z) : 0.5 * (1.0 + z), dh = l == 0 ? -0.5 : 0.5;
          for (int i = 0; i < 3; ++i) {
            const int v = 3 * l + i;
            r.N[v][q] = lam[i] * h;
            r.dN[v][q][0] = dlam[i][0] * h;
            r.dN[v][q][1] = dlam[i][1] * h;
            r.dN[v][q][2] = lam[i] * dh;
            for (int d = 0; d < 3; ++d) r.gdN[v][q][d] = r.dN[v][q][d];
          }
        }
      }
    return;
  }
  if (topo == 2 /* TET4 */) {
    r.nb = 4; r.nv = 4; r.nq = 4;
    const double a = 0.5854101966249685, b = 0.1381966011250105;
    for (int q = 0; q < 4; ++q) {
      double p[3] = {b, b, b};
      if (q > 0) p[q - 1] = a;
      r.qw[q] = 1.0 / 24.0;
      double n4[4] = {1.0 - p[0] - p[1] - p[2], p[0], p[1], p[2]};
      const double g4[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
      for (int v = 0; v < 4; ++v) {
        r.N[v][q] = n4[v];
        for (int d = 0; d < 3; ++d) { r.dN[v][q][d] = g4[v][d]; r.gdN[v][q][d] = g4[v][d]; }
      }
    }
    return;
  }
  const int ng = topo == 3 ? 3 : 2;
  double g1[3], w1[3];
  if (ng == 2) {
    g1[0] = -1.0 / std::sqrt(3.0); g1[1] = -g1[0]; w1[0] = w1[1] = 1.0;
  } else {
    g1[0] = -std::sqrt(0.6); g1[1] = 0.0; g1[2] = std::sqrt(0.6);
    w1[0] = w1[2] = 5.0 / 9.0; w1[1] = 8.0 / 9.0;
  }
  r.nv = 8;
  int q = 0;
  double qp[kMaxNQ][3];
  for (int k = 0; k < ng; ++k)
    for (int j = 0; j < ng; ++j)
      for (int i = 0; i < ng; ++i) {
        qp[q][0] = g1[i]; qp[q][1] = g1[j]; qp[q][2] = g1[k];
        r.qw[q] = w1[i] * w1[j] * w1[k];
        ++q;
      }
  r.nq = q;
  for (int qq = 0; qq < q; ++qq)
    for (int v = 0; v < 8; ++v) {
      double s[3], f[3];
      for (int d = 0; d < 3; ++d) { s[d] = 2.0 * hexv[v][d] - 1.0; f[d] = 0.5 * (1.0 + s[d] * qp[qq][d]); }
      for (int d = 0; d < 3; ++d) {
        double g = 0.5 * s[d];
        for (int e = 0; e < 3; ++e) if (e != d) g *= f[e];
        r.gdN[v][qq][d] = g;
      }
      if (topo == 1) {
        r.N[v][qq] = f[0] * f[1] * f[2];
        for (int d = 0; d < 3; ++d) r.dN[v][qq][d] = r.gdN[v][qq][d];
      }
    }
  if (topo == 1) { r.nb = 8; return; }
  // HEX27: Q2 nodes at {-1,0,1}^3 in subcell order
  static const int edge[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                                  {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
  static const int face[6][4] = {{0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6},
                                 {0, 4, 7, 3}, {0, 3, 2, 1}, {4, 5, 6, 7}};
  int t[27][3];
  for (int v = 0; v < 8; ++v) for (int d = 0; d < 3; ++d) t[v][d] = 2 * hexv[v][d] - 1;
  for (int e = 0; e < 12; ++e) for (int d = 0; d < 3; ++d) t[8 + e][d] = (t[edge[e][0]][d] + t[edge[e][1]][d]) / 2;
  for (int f = 0; f < 6; ++f)
    for (int d = 0; d < 3; ++d) {
      int sum = 0;
      for (int k = 0; k < 4; ++k) sum += t[face[f][k]][d];
      t[20 + f][d] = sum / 4;
    }
  for (int d = 0; d < 3; ++d) t[26][d] = 0;
  auto L = [](int node, double x) {
    return node < 0 ? 0.5 * x * (x - 1.0) : (node == 0 ? 1.0 - x * x : 0.5 * x * (x + 1.0));
  };
  auto dL = [](int node, double x) { return node < 0 ? x - 0.5 : (node == 0 ? -2.0 * x : x + 0.5); };
  r.nb = 27;
  for (int qq = 0; qq < q; ++qq)
    for (int b = 0; b < 27; ++b) {
      double l[3], dl[3];
      for (int d = 0; d < 3; ++d) { l[d] = L(t[b][d], qp[qq][d]); dl[d] = dL(t[b][d], qp[qq][d]); }
      r.N[b][qq] = l[0] * l[1] * l[2];
      r.dN[b][qq][0] = dl[0] * l[1] * l[2];
      r.dN[b][qq][1] = l[0] * dl[1] * l[2];
      r.dN[b][qq][2] = l[0] * l[1] * dl[2];
    }
}

}  // namespace txa
