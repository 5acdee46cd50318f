// One row of the sp3d5s* Slater-Koster block, for the SK-compressed
// operator (the block is rebuilt from a bond's direction cosines and its 21
// strain-scaled integrals on every product instead of being stored).
//
// sk_row<a>(l, m, n, v, row) writes row a (anion orbital a) of
// T[a][c] = <a on anion | H | c on cation>; the formulas are those of
// sk_table in nh_sk.h (which the assembly uses), split by row so that a
// lane evaluates only what it multiplies next.  The row index is a
// template argument: every lane of a wavefront runs the same code.
#pragma once

#include "nh_sk.h"

namespace nh {

template <int A, typename R>
__device__ __forceinline__ void sk_row(R l, R m, R n, const R *v, R row[10]) {
  const R s3 = R(kSqrt3Ref);
  const R l2 = l * l, m2 = m * m, n2 = n * n;
  const R sd[5] = {s3 * l * m, s3 * m * n, s3 * n * l, R(0.5) * s3 * (l2 - m2),
                   n2 - R(0.5) * (l2 + m2)};
  const R q[3] = {l, m, n};
  if constexpr (A == 0 || A == 1) {
    // s* row: ststs, stass, q*stapcs, sd*stadcs; s row: sasts, sss, q*sapcs, sd*sadcs
    const R v0 = A == 0 ? v[1] : v[3];   // s*-s* / s-s*
    const R v1 = A == 0 ? v[2] : v[0];   // s*-s  / s-s
    const R vp = A == 0 ? v[6] : v[4];   // stapcs / sapcs
    const R vd = A == 0 ? v[10] : v[8];  // stadcs / sadcs
    row[0] = v0;
    row[1] = v1;
#pragma unroll
    for (int c = 0; c < 3; ++c) row[2 + c] = q[c] * vp;
#pragma unroll
    for (int c = 0; c < 5; ++c) row[5 + c] = sd[c] * vd;
  } else if constexpr (A >= 2 && A <= 4) {
    constexpr int a = A - 2;
    const R qa = q[a];
    const R pps = v[12], ppp = v[13], padcs = v[14], padcp = v[16];
    row[0] = qa * (-v[7]);  // stcpas
    row[1] = qa * (-v[5]);  // scpas
#pragma unroll
    for (int c = 0; c < 3; ++c)
      row[2 + c] = a == c ? qa * qa * pps + (R(1) - qa * qa) * ppp
                          : qa * q[c] * pps - qa * q[c] * ppp;
    const R lmn = l * m * n;
    R ps[5], pp[5];
    if constexpr (a == 0) {
      ps[0] = s3 * l2 * m;            pp[0] = m * (R(1) - R(2) * l2);
      ps[1] = s3 * lmn;               pp[1] = R(-2) * lmn;
      ps[2] = s3 * l2 * n;            pp[2] = n * (R(1) - R(2) * l2);
      ps[3] = R(0.5) * s3 * l * (l2 - m2); pp[3] = l * (R(1) - l2 + m2);
      ps[4] = l * (n2 - R(0.5) * (l2 + m2)); pp[4] = -s3 * l * n2;
    } else if constexpr (a == 1) {
      ps[0] = s3 * m2 * l;            pp[0] = l * (R(1) - R(2) * m2);
      ps[1] = s3 * m2 * n;            pp[1] = n * (R(1) - R(2) * m2);
      ps[2] = s3 * lmn;               pp[2] = R(-2) * lmn;
      ps[3] = R(0.5) * s3 * m * (l2 - m2); pp[3] = -m * (R(1) + l2 - m2);
      ps[4] = m * (n2 - R(0.5) * (l2 + m2)); pp[4] = -s3 * m * n2;
    } else {
      ps[0] = s3 * lmn;               pp[0] = R(-2) * lmn;
      ps[1] = s3 * n2 * m;            pp[1] = m * (R(1) - R(2) * n2);
      ps[2] = s3 * n2 * l;            pp[2] = l * (R(1) - R(2) * n2);
      ps[3] = R(0.5) * s3 * n * (l2 - m2); pp[3] = -n * (l2 - m2);
      ps[4] = n * (n2 - R(0.5) * (l2 + m2)); pp[4] = s3 * n * (l2 + m2);
    }
#pragma unroll
    for (int c = 0; c < 5; ++c) row[5 + c] = ps[c] * padcs + pp[c] * padcp;
  } else {
    constexpr int d = A - 5;
    // d row: s-type columns with stcdas / scdas
    row[0] = sd[d] * v[11];
    row[1] = sd[d] * v[9];
    // p columns: E_{p,d}(c, d) with pcdas, pcdap (negated integrals)
    const R pcdas = -v[15], pcdap = -v[17];
    const R lmn = l * m * n;
    R ps[3], pp[3];
    if constexpr (d == 0) {
      ps[0] = s3 * l2 * m; pp[0] = m * (R(1) - R(2) * l2);
      ps[1] = s3 * m2 * l; pp[1] = l * (R(1) - R(2) * m2);
      ps[2] = s3 * lmn;    pp[2] = R(-2) * lmn;
    } else if constexpr (d == 1) {
      ps[0] = s3 * lmn;    pp[0] = R(-2) * lmn;
      ps[1] = s3 * m2 * n; pp[1] = n * (R(1) - R(2) * m2);
      ps[2] = s3 * n2 * m; pp[2] = m * (R(1) - R(2) * n2);
    } else if constexpr (d == 2) {
      ps[0] = s3 * l2 * n; pp[0] = n * (R(1) - R(2) * l2);
      ps[1] = s3 * lmn;    pp[1] = R(-2) * lmn;
      ps[2] = s3 * n2 * l; pp[2] = l * (R(1) - R(2) * n2);
    } else if constexpr (d == 3) {
      ps[0] = R(0.5) * s3 * l * (l2 - m2); pp[0] = l * (R(1) - l2 + m2);
      ps[1] = R(0.5) * s3 * m * (l2 - m2); pp[1] = -m * (R(1) + l2 - m2);
      ps[2] = R(0.5) * s3 * n * (l2 - m2); pp[2] = -n * (l2 - m2);
    } else {
      ps[0] = l * (n2 - R(0.5) * (l2 + m2)); pp[0] = -s3 * l * n2;
      ps[1] = m * (n2 - R(0.5) * (l2 + m2)); pp[1] = -s3 * m * n2;
      ps[2] = n * (n2 - R(0.5) * (l2 + m2)); pp[2] = s3 * n * (l2 + m2);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) row[2 + c] = ps[c] * pcdas + pp[c] * pcdap;
    // d-d
    const R dds = v[18], ddp = v[19], ddd = v[20];
    const R lm = l * m, mn = m * n, nl = n * l;
    const R dl = l2 - m2, z = n2 - R(0.5) * (l2 + m2), lpm = l2 + m2;
    const R e00 = R(3) * l2 * m2 * dds + (l2 + m2 - R(4) * l2 * m2) * ddp +
                  (n2 + l2 * m2) * ddd;
    const R e01 = R(3) * l * m2 * n * dds + nl * (R(1) - R(4) * m2) * ddp +
                  nl * (m2 - R(1)) * ddd;
    const R e02 = R(3) * l2 * mn * dds + mn * (R(1) - R(4) * l2) * ddp +
                  mn * (l2 - R(1)) * ddd;
    const R e03 = R(1.5) * lm * dl * dds + R(2) * lm * (m2 - l2) * ddp +
                  R(0.5) * lm * dl * ddd;
    const R e04 = s3 * lm * z * dds - s3 * R(2) * lm * n2 * ddp +
                  R(0.5) * s3 * lm * (R(1) + n2) * ddd;
    if constexpr (d == 0) {
      row[5] = e00; row[6] = e01; row[7] = e02; row[8] = e03; row[9] = e04;
    } else {
      const R e11 = R(3) * m2 * n2 * dds + (m2 + n2 - R(4) * m2 * n2) * ddp +
                    (l2 + m2 * n2) * ddd;
      const R e12 = R(3) * m * n2 * l * dds + l * m * (R(1) - R(4) * n2) * ddp +
                    l * m * (n2 - R(1)) * ddd;
      const R e13 = R(1.5) * mn * dl * dds - mn * (R(1) + R(2) * dl) * ddp +
                    mn * (R(1) + R(0.5) * dl) * ddd;
      const R e14 = s3 * mn * z * dds + s3 * mn * (lpm - n2) * ddp -
                    R(0.5) * s3 * mn * lpm * ddd;
      if constexpr (d == 1) {
        row[5] = e01; row[6] = e11; row[7] = e12; row[8] = e13; row[9] = e14;
      } else {
        const R e22 = R(3) * n2 * l2 * dds + (n2 + l2 - R(4) * n2 * l2) * ddp +
                      (m2 + n2 * l2) * ddd;
        const R e23 = R(1.5) * nl * dl * dds + nl * (R(1) - R(2) * dl) * ddp -
                      nl * (R(1) - R(0.5) * dl) * ddd;
        const R e24 = s3 * nl * z * dds + s3 * nl * (lpm - n2) * ddp -
                      R(0.5) * s3 * nl * lpm * ddd;
        if constexpr (d == 2) {
          row[5] = e02; row[6] = e12; row[7] = e22; row[8] = e23; row[9] = e24;
        } else {
          const R e33 = R(0.75) * dl * dl * dds + (lpm - dl * dl) * ddp +
                        (n2 + R(0.25) * dl * dl) * ddd;
          const R e34 = R(0.5) * s3 * dl * z * dds + s3 * n2 * (m2 - l2) * ddp +
                        R(0.25) * s3 * (R(1) + n2) * dl * ddd;
          if constexpr (d == 3) {
            row[5] = e03; row[6] = e13; row[7] = e23; row[8] = e33; row[9] = e34;
          } else {
            const R e44 = z * z * dds + R(3) * n2 * lpm * ddp +
                          R(0.75) * lpm * lpm * ddd;
            row[5] = e04; row[6] = e14; row[7] = e24; row[8] = e34; row[9] = e44;
          }
        }
      }
    }
  }
}

}  // namespace nh
