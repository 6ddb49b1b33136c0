// model.h - host-side job set-up for the kriging kernels: covariance model,
// box covariance tables, search templates and cell-bin geometry.
//
// Everything here runs once per job on the host and is uploaded to HBM.
// Quantities that decide neighbour ORDER (template order, cluster-path sort
// keys) are computed here with the host libm and libstdc++ std::sort, so
// they are bit-identical to the reference's CPU computation; the GPU only
// compares them.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace hpgl_amd {

struct CovSpec {
    int type;            // 0 spherical, 1 exponential, 2 gaussian
    double ranges[3];
    double angles[3];    // degrees
    double sill;
    double nugget;
};

// Analytic model of cov_model.h:79-182 with the anisotropy/rotation
// transform of create_transform (cov_model.h:16-61).
struct CovModel {
    CovSpec s;
    double t[9];          // row-major 3x3 transform
    explicit CovModel(const CovSpec &spec);
    double norm(const double v[3]) const;          // transfrom_and_norm, cov_model.h:63-77
    double of_h(double h) const;                   // cov_model.h:144-181
    double at(int dx, int dy, int dz) const;       // C at integer offset v = (dx,dy,dz)
    double c0() const { return s.sill - s.nugget; }
};

// Covariance over the search box [-r, r]^3 of integer offsets, 0 outside
// (precalculated_covariance.h:13-75, covariance_field.h:108-169).
struct BoxTable {
    int r[3] = {0, 0, 0};
    int d[3] = {1, 1, 1};  // 2r+1
    std::vector<double> v;  // index ((z+rz)*dy + (y+ry))*dx + (x+rx)
    void build(const CovModel &m, const int radii[3]);
    // The non-negative octant with one layer of zeros appended in each axis
    // (x = rx+1, y = ry+1, z = rz+1), index (z*(ry+2) + y)*(rx+2) + x, into
    // out when every entry equals its octant entry bit for bit (true for
    // models without rotation); false (out untouched) otherwise.  An offset
    // whose components are clamped to r+1 then reads its covariance, 0
    // outside the box, with no test.
    bool fold(std::vector<double> &out) const;
    double at(int dx, int dy, int dz) const {
        if (dx < -r[0] || dx > r[0] || dy < -r[1] || dy > r[1] || dz < -r[2] || dz > r[2]) return 0;
        return v[((size_t)(dz + r[2]) * d[1] + (dy + r[1])) * d[0] + (dx + r[0])];
    }
};

// Search template: offsets in search order (calc_cov_field,
// covariance_field.h:63-106), packed for the device as (dx, dy, dz, 0).
struct Template {
    std::vector<int32_t> xyz;  // 4 ints per offset
    int size() const { return (int)(xyz.size() / 4); }
};

// key(dx,dy,dz) -> covariance used to build and order the template; the
// caller passes either the analytic model (OK) or the box table.
void build_template(const CovModel &m, const BoxTable *table, const int radii[3], Template &out);

// Sort keys of the cell-binned search path, as dense ranks: for every
// integer offset within [-span, span]^3 the rank of its covariance among
// the distinct covariance values (higher covariance = higher rank), with
// rank 0 for offsets at or below the C0/100 threshold.  Comparing ranks is
// the same as comparing the double keys of
// sugarbox_indexed_neighbour_lookup.h:84-99.
struct RankTable {
    int span[3] = {0, 0, 0};
    int d[3] = {1, 1, 1};
    std::vector<uint32_t> rank;
    void build(const CovModel &m, const BoxTable *table, const int span_xyz[3]);
};

// Cell-bin geometry of clusterizer_t (clusterizer.cpp:255-278):
// bin size = radii, bin coordinate = loc / r, (g / r + 2) bins per axis.
struct BinGeometry {
    int r[3];
    int m[3];
    long long nbins;
    int limit;  // template size / 250 (sugarbox_indexed_neighbour_lookup.h:58)
    void init(const int grid[3], const int radii[3], int template_size);
};

// Throws std::runtime_error with the reference's message on invalid input.
void validate_cov(const CovSpec &s);

}  // namespace hpgl_amd
