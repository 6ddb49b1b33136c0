// devcov.h - covariance lookups shared by the solve kernels: the box table
// (precalculated_covariance.h:13-75, 0 outside the box) and the analytic
// model (cov_model.h:63-181) evaluated on the device.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.h"

namespace hpgl_amd {

__device__ inline double cov_eval(const DevCov &c, int dx, int dy, int dz) {
    if (c.table) {
        if (dx < -c.tr[0] || dx > c.tr[0] || dy < -c.tr[1] || dy > c.tr[1] || dz < -c.tr[2] || dz > c.tr[2])
            return 0.0;
        return c.table[((size_t)(dz + c.tr[2]) * c.td[1] + (dy + c.tr[1])) * c.td[0] + (dx + c.tr[0])];
    }
    // cov_model.h:63-77 transfrom_and_norm then :144-181
    const double v0 = dx, v1 = dy, v2 = dz;
    double r = 0.0;
    {
        double d = 0.0;
        d += v0 * c.t[0];
        d += v1 * c.t[1];
        d += v2 * c.t[2];
        r += d * d;
    }
    {
        double d = 0.0;
        d += v0 * c.t[3];
        d += v1 * c.t[4];
        d += v2 * c.t[5];
        r += d * d;
    }
    {
        double d = 0.0;
        d += v0 * c.t[6];
        d += v1 * c.t[7];
        d += v2 * c.t[8];
        r += d * d;
    }
    const double h = sqrt(r);
    if (h < 0.0001) return c.c0;
    const double a = c.range0;
    if (c.type == 0) {
        if (h > a) return 0.0;
        return c.c0 * (1 - 1.5 * h / a + 0.5 * pow(h / a, 3.0));
    }
    if (c.type == 1) return c.c0 * exp(-3 * h / a);
    return c.c0 * exp(-3 * pow(h / a, 2.0));
}

__device__ inline void cell_xyz(int cell, int X, int Y, int &x, int &y, int &z) {
    x = cell % X;
    y = cell % (X * Y) / X;
    z = cell / (X * Y);
}

// Covariance from the box table, branch-free (the entry for out-of-box
// offsets is read from index 0 and replaced by 0).
__device__ inline double cov_tab(const DevCov &c, int dx, int dy, int dz) {
    const bool in = dx >= -c.tr[0] && dx <= c.tr[0] && dy >= -c.tr[1] && dy <= c.tr[1] && dz >= -c.tr[2] &&
                    dz <= c.tr[2];
    const int idx = ((dz + c.tr[2]) * c.td[1] + (dy + c.tr[1])) * c.td[0] + (dx + c.tr[0]);
    const double v = c.table[in ? idx : 0];
    return in ? v : 0.0;
}

// The same from the table's octant (DevCov::fold), for tables that equal it
// bit for bit: offsets outside the box clamp into the octant's zero layer.
__device__ inline double cov_fold(const DevCov &c, int dx, int dy, int dz) {
    const int ax = min(abs(dx), c.tr[0] + 1), ay = min(abs(dy), c.tr[1] + 1), az = min(abs(dz), c.tr[2] + 1);
    return c.fold[(az * (c.tr[1] + 2) + ay) * (c.tr[0] + 2) + ax];
}

}  // namespace hpgl_amd
