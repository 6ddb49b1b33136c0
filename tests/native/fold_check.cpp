// fold_check.cpp - checks BoxTable::fold (hpgl_amd/csrc/model.cpp), the
// octant copy of a covariance table that the solve kernels read when the
// table is symmetric in every axis:
//
//   * models without rotation (all three types, anisotropic ranges, with and
//     without nugget) fold, and the octant entry (ax, ay, az) at
//     (az (ry + 2) + ay) (rx + 2) + ax, with ax = min(|dx|, rx + 1) and so
//     on, equals the full table's entry (0 outside the box) for every offset
//     of the box grown by two in each axis, bit for bit: the octant's
//     appended layer holds zeros, so offsets past the box read 0;
//   * rotated models do not fold (their table is not symmetric per axis);
//   * a table with one entry changed (a planted asymmetry) does not fold.
//
// Exit status 0 and "ok <folded> <refused>" on success; prints the first
// failure otherwise.
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <vector>

#include "../../hpgl_amd/csrc/model.h"

using hpgl_amd::BoxTable;
using hpgl_amd::CovModel;
using hpgl_amd::CovSpec;

static bool same(double a, double b) { return std::memcmp(&a, &b, sizeof(double)) == 0; }

int main() {
    int folded = 0, refused = 0;
    const int radii[][3] = {{5, 4, 2}, {12, 12, 4}, {7, 9, 3}, {1, 1, 1}};
    const double ranges[][3] = {{5, 4, 2}, {12, 12, 4}, {20, 9, 3.5}, {1, 1, 1}};
    for (int type = 0; type < 3; ++type)
        for (int k = 0; k < 4; ++k)
            for (int rot = 0; rot < 2; ++rot) {
                CovSpec s;
                s.type = type;
                for (int i = 0; i < 3; ++i) s.ranges[i] = ranges[k][i];
                s.angles[0] = rot ? 30.0 : 0.0;
                s.angles[1] = rot ? 10.0 : 0.0;
                s.angles[2] = 0.0;
                s.sill = 1.0;
                s.nugget = (k & 1) ? 0.1 : 0.0;
                CovModel m(s);
                BoxTable t;
                t.build(m, radii[k]);
                std::vector<double> oct;
                const bool f = t.fold(oct);
                if (rot) {
                    // a rotated model's table is not symmetric per axis
                    // (radius 1 boxes are too small to tell)
                    if (f && k != 3) {
                        std::printf("rotated model folded: type %d case %d\n", type, k);
                        return 1;
                    }
                    refused += !f;
                    continue;
                }
                if (!f) {
                    std::printf("unrotated model did not fold: type %d case %d\n", type, k);
                    return 1;
                }
                ++folded;
                const int rx = t.r[0], ry = t.r[1], rz = t.r[2];
                if (oct.size() != (size_t)(rx + 2) * (ry + 2) * (rz + 2)) {
                    std::printf("octant size %zu, type %d case %d\n", oct.size(), type, k);
                    return 1;
                }
                auto cl = [](int v, int r) { return std::min(std::abs(v), r + 1); };
                for (int z = -rz - 2; z <= rz + 2; ++z)
                    for (int y = -ry - 2; y <= ry + 2; ++y)
                        for (int x = -rx - 2; x <= rx + 2; ++x) {
                            const size_t i = ((size_t)cl(z, rz) * (ry + 2) + cl(y, ry)) * (rx + 2) + cl(x, rx);
                            if (!same(oct[i], t.at(x, y, z))) {
                                std::printf("octant entry differs at (%d,%d,%d), type %d case %d\n", x, y, z, type,
                                            k);
                                return 1;
                            }
                        }
                // planted asymmetry: one entry of a negative offset changed
                BoxTable u = t;
                const int x = -1, y = 0, z = 0;
                u.v[((size_t)(z + rz) * u.d[1] + (y + ry)) * u.d[0] + (x + rx)] += 1e-12;
                std::vector<double> o2;
                if (u.fold(o2)) {
                    std::printf("asymmetric table folded: type %d case %d\n", type, k);
                    return 1;
                }
            }
    std::printf("ok %d %d\n", folded, refused);
    return 0;
}
