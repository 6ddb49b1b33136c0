// model.cpp - host-side job set-up (see model.h).
#include "model.h"

#include <cstdlib>
#include <cstring>

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace hpgl_amd {

namespace {

const double kPi = 3.14159265358979323846264338327950288419;

double deg2rad(double g) { return (g * kPi) / 180; }  // cov_model.h:11-14

// 3x3 row-major product accumulated as sum = sum + a*b from 0, the order
// of TNT's matmult (tnt_126/tnt_cmat.h:471-497).
void mul3(const double a[9], const double b[9], double c[9]) {
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) {
            double sum = 0;
            for (int j = 0; j < 3; ++j) sum = sum + a[3 * i + j] * b[3 * j + k];
            c[3 * i + k] = sum;
        }
}

}  // namespace

void validate_cov(const CovSpec &s) {
    for (int i = 0; i < 3; ++i)
        if (s.ranges[i] == 0) throw std::runtime_error("create_transform: All covariance ranges should be non-zero");
    if (s.type < 0 || s.type > 2) throw std::runtime_error("covariance type must be 0, 1 or 2");
}

CovModel::CovModel(const CovSpec &spec) : s(spec) {
    validate_cov(s);
    // T = S * ((Rx * Ry) * Rz), cov_model.h:28-55
    double S[9] = {0}, Rz[9] = {0}, Ry[9] = {0}, Rx[9] = {0};
    S[0] = 1.0;
    S[4] = s.ranges[0] / s.ranges[1];
    S[8] = s.ranges[0] / s.ranges[2];
    const double a0 = deg2rad(s.angles[0]), a1 = deg2rad(s.angles[1]), a2 = deg2rad(s.angles[2]);
    Rz[0] = cos(a0);
    Rz[4] = cos(a0);
    Rz[1] = sin(a0);
    Rz[3] = -sin(a0);
    Rz[8] = 1;
    Ry[0] = cos(a1);
    Ry[8] = cos(a1);
    Ry[2] = -sin(a1);
    Ry[6] = sin(a1);
    Ry[4] = 1;
    Rx[4] = cos(a2);
    Rx[8] = cos(a2);
    Rx[5] = sin(a2);
    Rx[7] = -sin(a2);
    Rx[0] = 1;
    double xy[9], xyz[9];
    mul3(Rx, Ry, xy);
    mul3(xy, Rz, xyz);
    mul3(S, xyz, t);
}

double CovModel::norm(const double v[3]) const {
    double result = 0.0;
    for (int i = 0; i < 3; ++i) {
        double d = 0.0;
        for (int j = 0; j < 3; ++j) d += v[j] * t[3 * i + j];
        result += d * d;
    }
    return sqrt(result);
}

double CovModel::of_h(double h) const {
    const double c = s.sill - s.nugget;
    if (h < 0.0001) return c;
    const double a = s.ranges[0];
    switch (s.type) {
        case 0:
            if (h > a) return 0;
            return c * (1 - 1.5 * h / a + 0.5 * pow(h / a, 3));
        case 1:
            return c * exp(-3 * h / a);
        default:
            return c * exp(-3 * pow(h / a, 2));
    }
}

double CovModel::at(int dx, int dy, int dz) const {
    double v[3] = {(double)dx, (double)dy, (double)dz};
    return of_h(norm(v));
}

void BoxTable::build(const CovModel &m, const int radii[3]) {
    for (int i = 0; i < 3; ++i) {
        r[i] = radii[i];
        d[i] = 2 * radii[i] + 1;
    }
    v.resize((size_t)d[0] * d[1] * d[2]);
    // The reference stores cov(0, (x,y,z)), i.e. the model at v = -(x,y,z);
    // norm(-v) == norm(v) exactly, so C(v) is stored.
    for (int z = 0; z < d[2]; ++z)
        for (int y = 0; y < d[1]; ++y)
            for (int x = 0; x < d[0]; ++x)
                v[((size_t)z * d[1] + y) * d[0] + x] = m.at(-(x - r[0]), -(y - r[1]), -(z - r[2]));
}

bool BoxTable::fold(std::vector<double> &out) const {
    const size_t f0 = (size_t)r[0] + 2, f1 = (size_t)r[1] + 2, f2 = (size_t)r[2] + 2;
    for (int z = -r[2]; z <= r[2]; ++z)
        for (int y = -r[1]; y <= r[1]; ++y)
            for (int x = -r[0]; x <= r[0]; ++x) {
                const double a = at(x, y, z), b = at(std::abs(x), std::abs(y), std::abs(z));
                if (std::memcmp(&a, &b, sizeof(double)) != 0) return false;
            }
    out.resize(f0 * f1 * f2);
    for (size_t z = 0; z < f2; ++z)
        for (size_t y = 0; y < f1; ++y)
            for (size_t x = 0; x < f0; ++x) out[(z * f1 + y) * f0 + x] = at((int)x, (int)y, (int)z);
    return true;
}

void build_template(const CovModel &m, const BoxTable *table, const int radii[3], Template &out) {
    const int rx = radii[0], ry = radii[1], rz = radii[2];
    auto key = [&](int x, int y, int z) { return table ? table->at(x, y, z) : m.at(-x, -y, -z); };
    const double threshold = key(0, 0, 0) / 100;
    const int xd = 2 * rx + 1, yd = 2 * ry + 1;
    std::vector<double> data((size_t)xd * yd * (2 * rz + 1));
    struct Off {
        int c[3];
    };
    std::vector<Off> offs;
    size_t n = 0;
    // generation order z, y, x; stored offset = -(x, y, z) (covariance_field.h:74-89)
    for (int z = -rz; z <= rz; ++z)
        for (int y = -ry; y <= ry; ++y)
            for (int x = -rx; x <= rx; ++x) {
                double value = key(x, y, z);
                data[n++] = value;
                if (value > threshold) offs.push_back(Off{{-x, -y, -z}});
            }
    auto k = [&](const Off &o) {
        return data[(size_t)(o.c[2] + rz) * yd * xd + (size_t)(o.c[1] + ry) * xd + (o.c[0] + rx)];
    };
    // predicate_2 (covariance_field.h:13-34), unstable libstdc++ introsort
    std::sort(offs.begin(), offs.end(), [&](const Off &a, const Off &b) { return k(a) > k(b); });
    for (size_t i = 0; i < offs.size(); ++i)
        if (key(offs[i].c[0], offs[i].c[1], offs[i].c[2]) < threshold) {
            offs.resize(i);
            break;
        }
    out.xyz.resize(offs.size() * 4);
    for (size_t i = 0; i < offs.size(); ++i) {
        out.xyz[4 * i + 0] = offs[i].c[0];
        out.xyz[4 * i + 1] = offs[i].c[1];
        out.xyz[4 * i + 2] = offs[i].c[2];
        out.xyz[4 * i + 3] = 0;
    }
}

void RankTable::build(const CovModel &m, const BoxTable *table, const int span_xyz[3]) {
    for (int i = 0; i < 3; ++i) {
        span[i] = span_xyz[i];
        d[i] = 2 * span_xyz[i] + 1;
    }
    const size_t n = (size_t)d[0] * d[1] * d[2];
    std::vector<double> key(n);
    // Key of a candidate at offset (cand - centre): the reference evaluates
    // cov(centre, cand); the table looks up cand - centre, the model uses
    // centre - cand.  Both equal C(cand - centre) bit for bit.
    for (int z = 0; z < d[2]; ++z)
        for (int y = 0; y < d[1]; ++y)
            for (int x = 0; x < d[0]; ++x) {
                int ox = x - span[0], oy = y - span[1], oz = z - span[2];
                key[((size_t)z * d[1] + y) * d[0] + x] = table ? table->at(ox, oy, oz) : m.at(-ox, -oy, -oz);
            }
    const double c00 = table ? table->at(0, 0, 0) : m.at(0, 0, 0);
    const double threshold = c00 / 100;
    std::vector<double> vals;
    vals.reserve(n);
    for (double k : key)
        if (k > threshold) vals.push_back(k);
    std::sort(vals.begin(), vals.end());
    vals.erase(std::unique(vals.begin(), vals.end()), vals.end());
    rank.resize(n);
    for (size_t i = 0; i < n; ++i) {
        if (key[i] > threshold)
            rank[i] = (uint32_t)(std::lower_bound(vals.begin(), vals.end(), key[i]) - vals.begin()) + 1;
        else
            rank[i] = 0;
    }
}

void BinGeometry::init(const int grid[3], const int radii[3], int template_size) {
    nbins = 1;
    for (int i = 0; i < 3; ++i) {
        r[i] = radii[i];
        // search_area_t holds radii as double: bins per axis = int(g / r + 2)
        m[i] = (int)(grid[i] / (double)radii[i] + 2);
        nbins *= m[i];
    }
    limit = template_size / 250;
}

}  // namespace hpgl_amd
