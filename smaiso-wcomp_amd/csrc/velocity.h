// The velocity solve at one cosine mode (a11 tail + a12, smaiso-wcomp.cpp:
// 1479-1974) and the terms of the spectral divergence of its irrotational
// part (S:2042-2063), shared by the solve's own kernel (stepper.hip) and by
// the z pass that computes it on load (fused.hip), so both give the same
// values bit for bit.
//
// Cosine mode (i, j, k) carries the x component at (i-1, j, k), the y
// component at (i, j-1, k) and the z component at (i, j, k-1): on their
// staggered lattices these three points have the wavenumbers of cosine mode
// (i, j, k) (vs_x[i-1] = vq_x[i], likewise on y and z).  The shifted reads of
// the three solves (x: S_y(i+1, j-1, k), S_z(i+1, j, k-1), fgrad^(i+1, j, k);
// y: S_x(i-1, j+1, k), S_z(i, j+1, k-1), fgrad^(i, j+1, k); z: S_x(i-1, j,
// k+1), S_y(i, j-1, k+1), fgrad^(i, j, k+1)) then all land on the same values
// S_x(i-1, j, k), S_y(i, j-1, k), S_z(i, j, k-1), fgrad^(i, j, k) and
// erho^(i, j, k), so the three components share one projection q.S^, one
// denominator |q|^2 and one Gaussian G(q) (a component's factor exp(-1.57^2
// 2 |q_a|^2) on its own lattice is G at cosine mode (i, j, k), which is also
// the G of fgrad^ += G erho^).  virr_x(i-1, j, k), virr_y(i, j-1, k) and
// virr_z(i, j, k-1) are exactly the terms of div v^(i, j, k) = sum_a Vq_a
// virr_a(n - e_a).
#pragma once

#include "stepper.h"

namespace smaiso {

template <typename A>
struct VelMode {
  A tx, ty, tz;  // vsol^ + virr^ of the three components
  A rx, ry, rz;  // virr^ of the three components
  A div;         // div v^ at (i, j, k)
};

// Sx = S_x(i-1, j, k), Sy = S_y(i, j-1, k), Sz = S_z(i, j, k-1) (0 where the
// shifted index leaves the grid), fg = fgrad^(i, j, k), er = erho^(i, j, k).
// A component that does not exist or that the reference leaves at 0 (quirks
// Q1, Q2) is 0.
template <typename A>
__device__ __forceinline__ VelMode<A> velocity_mode(const StepConsts& c, const MemberConsts& mm, int i, int j, int k,
                                                    A Sx, A Sy, A Sz, A fgq, A erq) {
  // wavenumbers from the spacings (no table loads)
  const A qs0 = (A)c.qs[0], qs1 = (A)c.qs[1], qs2 = (A)c.qs[2], scale = (A)c.scale;
  const A qx = i * qs0, qy = j * qs1, qz = k * qs2;
  const A rnu = mm.rnu, rlam = mm.rlam;
  const A q2 = qx * qx + qy * qy + qz * qz;
  // |q|^2 = 0 only at (0, 0, 0), where no component is computed
  const A inv = A(1) / (q2 > A(0) ? q2 : A(1));
  const A cm1 = scale * inv;
  const A G = exp(A(-1.57 * 1.57 * 2) * q2);
  const A fg = fgq + erq * G;
  const A dot = qx * Sx + qy * Sy + qz * Sz;
  const A sn = cm1 * rnu, sl = cm1 * rlam;
  VelMode<A> r;
  // x component at (i-1, j, k) (S:1583-1709); the plane i = Nx-1 is zero
  // (quirk Q2), which cosine plane i = 0 stands for
  A ox = 0, rx = 0;
  if (i > 0) {
    const A cm2 = qx * inv;
    ox = sn * (Sx - cm2 * dot);
    rx = sl * (qx * fg + cm2 * dot * G);
  }
  // y component at (i, j-1, k) and z component at (i, j, k-1) (S:1759-1974):
  // computed only for j <= Ny-2 and k <= Nz-2 at their own points (quirk Q1,
  // S:1760-1761); the line j = Ny-1 of v_y and the column k = Nz-1 of v_z are
  // zero, which cosine line j = 0 and cosine mode k = 0 stand for
  A oy = 0, ry = 0, oz = 0, rz = 0;
  if (j > 0 && k < c.nz - 1) {
    const A cm2 = qy * inv;
    oy = sn * (Sy - cm2 * dot);
    ry = sl * (qy * fg + cm2 * dot * G);
  }
  if (k > 0 && j < c.ny - 1) {
    const A cm2 = qz * inv;
    // quirk Q3 (S:1932): at k = 0, j >= 1, except on the first plane of each
    // reference slab, the dot term takes CM1y in its exponent: the y
    // component's factor at the same point, G at cosine mode (i, j+1, 0)
    const int slab = c.slab_n0 > 0 ? c.slab_n0 : c.nx;
    A gz = G;
    if (k == 1 && j > 0 && (i % slab) != 0) {
      const A wy = (j + 1) * qs1;
      gz = exp(A(-1.57 * 1.57 * 2) * (qx * qx + wy * wy));
    }
    oz = sn * (Sz - cm2 * dot);
    rz = sl * (qz * fg + cm2 * dot * gz);
  }
  r.tx = ox + rx; r.ty = oy + ry; r.tz = oz + rz;
  r.rx = rx; r.ry = ry; r.rz = rz;
  // div v^(i, j, k) = sum_a Vq_a[n_a] virr_a(n - e_a), a term being 0 at n_a = 0
  r.div = qx * rx + qy * ry + qz * rz;
  return r;
}

}  // namespace smaiso
