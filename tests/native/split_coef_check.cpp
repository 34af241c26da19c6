// Brute-force check of the nodal PML pull coefficients (hr_split.h): for every
// PML node of small boxes, the side-scaled stencils must equal the element-by-
// element sums over the PML elements holding the node of
//   u rows: W(point) DM(p, l) / det J   at the offset of point p on each axis line,
//   s rows: W(node)  DM(l, p) / det J,
// with W = det J w(a) w(b) w(c) of the 3-point GLL rule, and the node's PML
// weight sum against the sum of w(a) w(b) w(c) over those elements.  Also the
// list-free enumeration of the PML nodes and the rule by which the nodal pass
// finds a node's x-neighbours among the list entries.  Prints the largest
// difference and exits 1 above 1e-14.
#include <cmath>
#include <cstdio>
#include <vector>
#include "../../hirensi_amd/csrc/hr_split.h"

using namespace hr;

static double gw(int a) { return a == 1 ? 4.0 / 3.0 : 1.0 / 3.0; }
// derivative of the quadratic Lagrange basis c at the GLL point a (xi = -1, 0, 1)
static double gdm(int a, int c) {
  const double x = a - 1.0;
  if (c == 0) return x - 0.5;
  if (c == 1) return -2.0 * x;
  return x + 0.5;
}

int main() {
  struct Box { int ne[3]; int rlo[3]; int rhi[2]; };
  // the last box has a regular domain one element wide: its rows hold PML nodes
  // on both sides of a one-node gap
  const Box boxes[] = {{{6, 5, 4}, {2, 1, 1}, {5, 4}}, {{7, 6, 5}, {1, 2, 2}, {6, 4}}, {{4, 4, 3}, {1, 1, 1}, {3, 3}},
                       {{5, 5, 3}, {2, 2, 1}, {3, 3}}};
  double worst = 0.0;
  long checked = 0, ifaces = 0;
  for (const Box& B : boxes) {
    const int N[3] = {2 * B.ne[0] + 1, 2 * B.ne[1] + 1, 2 * B.ne[2] + 1};
    for (int k = 0; k < N[2]; ++k)
      for (int j = 0; j < N[1]; ++j)
        for (int i = 0; i < N[0]; ++i) {
          const int c[3] = {i, j, k};
          // PML node: touched by a PML element
          bool pmln = false, rdn = false;
          double bu[3][5] = {}, bs[3][5] = {}, bw = 0.0;
          for (int ez = (k - 2) / 2; ez <= k / 2; ++ez)
            for (int ey = (j - 2) / 2; ey <= j / 2; ++ey)
              for (int ex = (i - 2) / 2; ex <= i / 2; ++ex) {
                const int e[3] = {ex, ey, ez};
                bool in = true;
                int l[3];
                for (int a = 0; a < 3; ++a) {
                  if (e[a] < 0 || e[a] >= B.ne[a]) in = false;
                  l[a] = c[a] - 2 * e[a];
                  if (l[a] < 0 || l[a] > 2) in = false;
                }
                if (!in) continue;
                const bool rd = e[0] >= B.rlo[0] && e[0] < B.rhi[0] && e[1] >= B.rlo[1] && e[1] < B.rhi[1] && e[2] >= B.rlo[2];
                if (rd) { rdn = true; continue; }
                pmln = true;
                bw += gw(l[0]) * gw(l[1]) * gw(l[2]);
                for (int ax = 0; ax < 3; ++ax) {
                  const int o1 = ax == 0 ? 1 : 0, o2 = ax == 2 ? 1 : 2;
                  const double wo = gw(l[o1]) * gw(l[o2]);
                  for (int p = 0; p < 3; ++p) {
                    const int d = p - l[ax] + 2;
                    bu[ax][d] += gw(p) * wo * gdm(p, l[ax]);
                    bs[ax][d] += gw(l[ax]) * wo * gdm(l[ax], p);
                  }
                }
              }
          if (!pmln) continue;
          const SpAxis X = sp_axis(i, B.ne[0], B.rlo[0], B.rhi[0]);
          const SpAxis Y = sp_axis(j, B.ne[1], B.rlo[1], B.rhi[1]);
          const SpAxis Z = sp_axis(k, B.ne[2], B.rlo[2], B.ne[2]);
          const SpAxis ax3[3] = {X, Y, Z};
          if (sp_iface(X, Y, Z) != rdn) { std::printf("interface flag differs at %d %d %d\n", i, j, k); return 1; }
          ifaces += rdn ? 1 : 0;
          worst = std::fmax(worst, std::fabs(sp_pml_weight(X, Y, Z) - bw));
          for (int ax = 0; ax < 3; ++ax) {
            double s0 = 0.0, s1 = 0.0, cu[5];
            sp_side_scales(ax3[ax], ax3[ax == 0 ? 1 : 0], ax3[ax == 2 ? 1 : 2], s0, s1);
            sp_stencil(ax3[ax].odd, s0, s1, cu);
            for (int d = 0; d < 5; ++d) {
              const double cs = d == 2 ? cu[d] : -cu[d];
              worst = std::fmax(worst, std::fabs(cu[d] - bu[ax][d]));
              worst = std::fmax(worst, std::fabs(cs - bs[ax][d]));
            }
          }
          ++checked;
        }
  }
  std::printf("nodes %ld interface %ld max diff %.3e\n", checked, ifaces, worst);
  // list-free enumeration of the PML nodes in lattice order; the nodal pass
  // finds the x-neighbour at offset di (|di| <= 2) of entry t at entry t + di,
  // or (|di| = 2) at t + di / 2, or it is no PML node
  long listed = 0, xpairs = 0;
  for (const Box& B : boxes) {
    const int N[3] = {2 * B.ne[0] + 1, 2 * B.ne[1] + 1, 2 * B.ne[2] + 1};
    std::vector<long> pos((size_t)N[0] * N[1] * N[2], -1);
    {
      long q = 0;
      for (int k = 0; k < N[2]; ++k)
        for (int j = 0; j < N[1]; ++j)
          for (int i = 0; i < N[0]; ++i) {
            const bool rdi = i > 2 * B.rlo[0] && i < 2 * B.rhi[0] && j > 2 * B.rlo[1] && j < 2 * B.rhi[1] && k > 2 * B.rlo[2];
            if (!rdi) pos[(size_t)i + (size_t)N[0] * ((size_t)j + (size_t)N[1] * k)] = q++;
          }
      for (int k = 0; k < N[2]; ++k)
        for (int j = 0; j < N[1]; ++j)
          for (int i = 0; i < N[0]; ++i) {
            const long t0 = pos[(size_t)i + (size_t)N[0] * ((size_t)j + (size_t)N[1] * k)];
            if (t0 < 0) continue;
            for (int di = -2; di <= 2; ++di) {
              if (di == 0 || i + di < 0 || i + di >= N[0]) continue;
              const long t1 = pos[(size_t)(i + di) + (size_t)N[0] * ((size_t)j + (size_t)N[1] * k)];
              if (t1 < 0) continue;
              const bool ok = t1 == t0 + di || ((di == 2 || di == -2) && t1 == t0 + di / 2);
              if (!ok) { std::printf("x-neighbour %d of (%d %d %d) at entry %ld, not found\n", di, i, j, k, t1); return 1; }
              ++xpairs;
            }
          }
    }
    unsigned t = 0;
    for (int k = 0; k < N[2]; ++k)
      for (int j = 0; j < N[1]; ++j)
        for (int i = 0; i < N[0]; ++i) {
          const bool rdi = i > 2 * B.rlo[0] && i < 2 * B.rhi[0] && j > 2 * B.rlo[1] && j < 2 * B.rhi[1] && k > 2 * B.rlo[2];
          if (rdi) continue;
          int a = -1, b = -1, q = -1;
          if (!sp_pml_node(t, N[0], N[1], N[2], B.rlo, B.rhi, a, b, q) || a != i || b != j || q != k) {
            std::printf("PML node %u: got (%d %d %d), expected (%d %d %d)\n", t, a, b, q, i, j, k);
            return 1;
          }
          ++t;
          ++listed;
        }
    int a, b, q;
    if (sp_pml_node(t, N[0], N[1], N[2], B.rlo, B.rhi, a, b, q) || t != sp_pml_count(N[0], N[1], N[2], B.rlo, B.rhi)) {
      std::printf("PML node count differs\n");
      return 1;
    }
  }
  std::printf("PML nodes enumerated %ld, x-neighbour pairs %ld\n", listed, xpairs);
  return (worst <= 1e-14 && checked > 0 && ifaces > 0) ? 0 : 1;
}
