// LSDA exchange-correlation at one point: Dirac exchange + PW92 correlation with the spin interpolation, f64
// (qm/src/xc.rs:120-210). Shared by the XC kernel and the host free-atom solver.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>

namespace lira {

__host__ __device__ inline void pw92_g(double rs, double a, double a1, double b1, double b2, double b3, double b4, double &g,
                              double &dg) {
    const double srs = sqrt(rs);
    const double q0 = -2.0 * a * (1.0 + a1 * rs);
    const double q1 = 2.0 * a * (b1 * srs + b2 * rs + b3 * rs * srs + b4 * rs * rs);
    const double q1p = a * (b1 / srs + 2.0 * b2 + 3.0 * b3 * srs + 4.0 * b4 * rs);
    const double ln = log(1.0 + 1.0 / q1);
    g = q0 * ln;
    dg = -2.0 * a * a1 * ln - q0 * q1p / (q1 * q1 + q1);
}

__host__ __device__ inline void exc_point(double nu, double nd, double &e, double &vu, double &vd) {
    nu = nu > 0.0 ? nu : 0.0;
    nd = nd > 0.0 ? nd : 0.0;
    const double n = nu + nd;
    if (n < 1e-20) {
        e = vu = vd = 0.0;
        return;
    }
    const double third = 1.0 / 3.0;
    const double cx = 0.75 * cbrt(3.0 / M_PI) * cbrt(2.0);
    const double cnu = cbrt(nu), cnd = cbrt(nd);
    const double ex = -cx * (nu * cnu + nd * cnd);
    const double vxu = -cbrt(6.0 / M_PI) * cnu, vxd = -cbrt(6.0 / M_PI) * cnd;
    // correlation
    const double rs = cbrt(3.0 / (4.0 * M_PI * n));
    double z = (nu - nd) / n;
    z = z < -1.0 ? -1.0 : (z > 1.0 ? 1.0 : z);
    double ec0, dec0, ec1, dec1, mac, dmac;
    pw92_g(rs, 0.031091, 0.21370, 7.5957, 3.5876, 1.6382, 0.49294, ec0, dec0);
    pw92_g(rs, 0.015545, 0.20548, 14.1189, 6.1977, 3.3662, 0.62517, ec1, dec1);
    pw92_g(rs, 0.016887, 0.11125, 10.357, 3.6231, 0.88026, 0.49671, mac, dmac);
    const double ac = -mac, dac = -dmac;
    const double denom = cbrt(16.0) - 2.0;  // 2^{4/3} - 2
    const double opz = cbrt(1.0 + z), omz = cbrt(1.0 - z);
    const double f = ((1.0 + z) * opz + (1.0 - z) * omz - 2.0) / denom;
    const double fp = (4.0 * third) * (opz - omz) / denom;
    const double FPP0 = 1.709921;
    const double z3 = z * z * z, z4 = z3 * z;
    const double ec = ec0 + ac * f / FPP0 * (1.0 - z4) + (ec1 - ec0) * f * z4;
    const double decdrs = dec0 + dac * f / FPP0 * (1.0 - z4) + (dec1 - dec0) * f * z4;
    const double decdz = ac / FPP0 * (fp * (1.0 - z4) - 4.0 * z3 * f) + (ec1 - ec0) * (fp * z4 + 4.0 * z3 * f);
    const double common = ec - rs * third * decdrs;
    e = ex + n * ec;
    vu = vxu + common - (z - 1.0) * decdz;
    vd = vxd + common - (z + 1.0) * decdz;
}

}  // namespace lira
