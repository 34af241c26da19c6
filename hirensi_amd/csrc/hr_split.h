// Coefficients of the nodal PML pull (split stage path), shared by the HIP
// kernels and the host-side checks.  Plain constexpr functions: the HIP
// compiler treats them as callable on both sides.
//
// A lattice node n sits in one element per axis (odd lattice coordinate, local
// point 1) or in up to two (even coordinate: the lower element, local point 2,
// and the upper one, local point 0).  Every PML element e holding n adds, per
// axis k, the transposed derivative of the nodal flux G_ik along the k-line
// through n (u rows) and, weighted by W(n in e), the forward derivative of the
// F-weighted input w_k along the same line (s rows).  The weights factor, W =
// det J w(a) w(b) w(c), so along the x-line the elements on one x-side differ
// only by w(l_y) w(l_z), a constant per axis parity: their x terms combine into
// one stencil scaled by
//   s_x[side] = w_y w_z * (number of PML elements among the y-z sides),
// where regular-domain elements drop out (they contribute through sigma).  The
// element box of the regular domain is a product of per-axis ranges, so that
// count is (n_y n_z - r_x[side] R_y R_z) with n the sides per axis and R the
// regular-domain sides.
//
// With side scales s0 (lower or only element) and s1 (upper), the u-row
// stencil of offsets -2..2 is, from w(p) DM(p, l) of the 3-point GLL rule,
//   even: (-s0/6, 2 s0/3, (s0 - s1)/2, -2 s1/3, s1/6),   odd: (0, 2 s0/3, 0, -2 s0/3, 0),
// and by summation by parts (w_a DM(a, b) + w_b DM(b, a) = 0 off the end
// points) the s-row stencil w(l) DM(l, p) is the same with the sign of every
// offset but 0 flipped.
#pragma once

namespace hr {

struct SpAxis {
  int odd;      // odd lattice coordinate: one element
  int v[2];     // side exists
  int r[2];     // side exists and is inside the regular-domain element range
};

// lattice coordinate c of an axis with nel elements; regular-domain element range [lo, hi)
constexpr SpAxis sp_axis(int c, int nel, int lo, int hi) {
  SpAxis a{c & 1, {0, 0}, {0, 0}};
  int e0 = 0, e1 = 0;
  if (a.odd) {
    a.v[0] = 1; e0 = (c - 1) >> 1;
  } else {
    a.v[0] = c >= 2 ? 1 : 0; e0 = (c >> 1) - 1;
    a.v[1] = (c >> 1) < nel ? 1 : 0; e1 = c >> 1;
  }
  a.r[0] = (a.v[0] && e0 >= lo && e0 < hi) ? 1 : 0;
  a.r[1] = (a.v[1] && e1 >= lo && e1 < hi) ? 1 : 0;
  return a;
}

constexpr double sp_wn(const SpAxis& a) { return a.odd ? 4.0 / 3.0 : 1.0 / 3.0; }

// side scales of axis A, B and C being the other two axes
constexpr void sp_side_scales(const SpAxis& A, const SpAxis& B, const SpAxis& C, double& s0, double& s1) {
  const int nb = B.v[0] + B.v[1], nc = C.v[0] + C.v[1];
  const int rb = B.r[0] + B.r[1], rc = C.r[0] + C.r[1];
  const double w = sp_wn(B) * sp_wn(C);
  s0 = A.v[0] ? w * (double)(nb * nc - A.r[0] * rb * rc) : 0.0;
  s1 = A.v[1] ? w * (double)(nb * nc - A.r[1] * rb * rc) : 0.0;
}

// u-row stencil of offsets -2..2 (index d + 2); the s-row stencil is
// cs[2] = cu[2], cs[d] = -cu[d] otherwise
constexpr void sp_stencil(int odd, double s0, double s1, double cu[5]) {
  cu[0] = odd ? 0.0 : -s0 / 6.0;
  cu[1] = 2.0 * s0 / 3.0;
  cu[2] = odd ? 0.0 : 0.5 * (s0 - s1);
  cu[3] = odd ? -2.0 * s0 / 3.0 : -2.0 * s1 / 3.0;
  cu[4] = odd ? 0.0 : s1 / 6.0;
}

// sum of w(a) w(b) w(c) over the PML elements holding the node (times det J:
// the node's PML weight sum): the elements holding it are the products of its
// per-axis sides, the regular-domain ones those of its regular-domain sides
constexpr double sp_pml_weight(const SpAxis& X, const SpAxis& Y, const SpAxis& Z) {
  const int n = (X.v[0] + X.v[1]) * (Y.v[0] + Y.v[1]) * (Z.v[0] + Z.v[1]);
  const int r = (X.r[0] + X.r[1]) * (Y.r[0] + Y.r[1]) * (Z.r[0] + Z.r[1]);
  return sp_wn(X) * sp_wn(Y) * sp_wn(Z) * (double)(n - r);
}

// whether a regular-domain element holds the node (an interface node)
constexpr bool sp_iface(const SpAxis& X, const SpAxis& Y, const SpAxis& Z) {
  return (X.r[0] + X.r[1]) * (Y.r[0] + Y.r[1]) * (Z.r[0] + Z.r[1]) > 0;
}

// The PML nodes in lattice order (i fastest, then j, then k) without a list:
// the planes k <= 2 lo[2] are whole PML planes; every higher plane lacks the
// regular-domain interior i in (2 lo[0], 2 hi[0]), j in (2 lo[1], 2 hi[1]), so its
// rows with j in that range hold the two runs i <= 2 lo[0] and i >= 2 hi[0].
// Entry t of that order -> (i, j, k); returns false past the last PML node.
constexpr bool sp_pml_node(unsigned t, int NX, int NY, int NZ, const int lo[3], const int hi[2], int& i, int& j,
                           int& k) {
  const unsigned nxy = (unsigned)NX * (unsigned)NY;
  const unsigned ax = (unsigned)(2 * (hi[0] - lo[0]) - 1), ay = (unsigned)(2 * (hi[1] - lo[1]) - 1);
  const unsigned P = nxy - ax * ay;                       // PML nodes of a holed plane
  const unsigned F = (unsigned)(2 * lo[2] + 1) * nxy;     // nodes of the whole planes
  if (t < F) {
    k = (int)(t / nxy);
    const unsigned r = t - (unsigned)k * nxy;
    j = (int)(r / (unsigned)NX);
    i = (int)(r - (unsigned)j * (unsigned)NX);
    return true;
  }
  const unsigned u = t - F;
  const unsigned kk = u / P;
  if ((int)kk >= NZ - 2 * lo[2] - 1) return false;
  k = 2 * lo[2] + 1 + (int)kk;
  const unsigned r = u - kk * P;
  const unsigned R0 = (unsigned)(2 * lo[1] + 1) * (unsigned)NX;   // whole rows below the interior
  const unsigned W = (unsigned)NX - ax;                            // nodes of a holed row
  const unsigned H = ay * W;
  if (r < R0) {
    j = (int)(r / (unsigned)NX);
    i = (int)(r - (unsigned)j * (unsigned)NX);
  } else if (r - R0 < H) {
    const unsigned r2 = r - R0, jr = r2 / W, q = r2 - jr * W;
    j = 2 * lo[1] + 1 + (int)jr;
    i = (int)(q <= (unsigned)(2 * lo[0]) ? q : q + ax);
  } else {
    const unsigned r3 = r - R0 - H, jr = r3 / (unsigned)NX;
    j = 2 * hi[1] + (int)jr;
    i = (int)(r3 - jr * (unsigned)NX);
  }
  return true;
}

// number of PML nodes
constexpr unsigned sp_pml_count(int NX, int NY, int NZ, const int lo[3], const int hi[2]) {
  const unsigned nxy = (unsigned)NX * (unsigned)NY;
  const unsigned ax = (unsigned)(2 * (hi[0] - lo[0]) - 1), ay = (unsigned)(2 * (hi[1] - lo[1]) - 1);
  return (unsigned)(2 * lo[2] + 1) * nxy + (unsigned)(NZ - 2 * lo[2] - 1) * (nxy - ax * ay);
}

}  // namespace hr
