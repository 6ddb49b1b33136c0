/* cvariogram.h - C ABI of hpgl_amd/_cvariogram.so, a drop-in for the
 * reference's experimental-variogram module (src/geo_bsd/_cvariogram/api.h),
 * which geo_bsd/cvariogram.py loads with
 * numpy.ctypeslib.load_library('_cvariogram', ...).
 *
 * Struct layouts are those of _cvariogram/api.h:23-90 (LP64).  Errors never
 * cross the ABI: a failing call leaves its outputs untouched and sets the
 * message returned by cvar_amd_last_error (no reference counterpart). */
#ifndef HPGL_AMD_CVARIOGRAM_H
#define HPGL_AMD_CVARIOGRAM_H

#ifdef __cplusplus
extern "C" {
#endif

struct vector_t {                      /* api.h:23 */
    double m_data[3];
};

struct ellipsoid_t {                   /* api.h:28, 96 bytes */
    struct vector_t m_direction1;
    struct vector_t m_direction2;
    struct vector_t m_direction3;
    double m_R1;
    double m_R2;
    double m_R3;
};

struct variogram_search_template_t {   /* api.h:38, 136 bytes */
    double m_lag_width;
    double m_lag_separation;
    double m_tol_distance;
    int m_num_lags;
    double m_first_lag_distance;
    struct ellipsoid_t m_ellipsoid;
};

struct cont_point_set_t {              /* api.h:64, 40 bytes */
    float *xs;
    float *ys;
    float *zs;
    float *values;
    int size;
};

struct hard_data_t {                   /* api.h:73, 64 bytes; strides in elements */
    float *m_data;
    unsigned char *m_mask;
    int m_data_shape[3];
    int m_data_strides[3];
    int m_mask_shape[3];
    int m_mask_strides[3];
};

struct float_data_t {                  /* api.h:83, 32 bytes */
    float *m_data;
    int m_data_shape[3];
    int m_data_strides[3];
};

/* api.h:92-97, ellipsoid.cpp:39-71: the three axes of the search ellipsoid
 * for azimuth, dip and rotation in degrees. */
void fill_ellipsoid_directions(struct ellipsoid_t *ell, double azimuth, double dip, double rotation);

/* api.h:117-122, variograms.cpp:362-446: semivariogram of gridded hard data
 * for the first min(num_lags, result_length) lags (NaN where a lag has no
 * pair).  percentToUse < 100 keeps each pair with that probability by a
 * deterministic hash of (offset, cell) instead of the C library's rand().
 * Only the informed cells cross PCIe; the device buffers (a dense copy of
 * the grid, about 5 bytes per cell) are kept per calling thread and device
 * between calls and held until the process ends. */
void calc_variograms(struct variogram_search_template_t *templ, struct hard_data_t *data,
                     float *result_covariations, int result_length, int percentToUse);

/* api.h:124-128, variograms.cpp:448-502: semivariogram of a point set (0
 * where a lag has no pair). */
void calc_variograms_from_point_set(struct variogram_search_template_t *templ, struct cont_point_set_t *point_set,
                                    float *result_covariations, int result_length);

/* api.h:130-138, stack_layers.cpp:7-21 (stack_layers.h:6-63): fills the
 * cube `result` with layer markers from layer thickness maps.  Cells below
 * the cube (k < 0, reached when erosions make the running depth negative),
 * which the reference writes out of bounds, are skipped. */
void cvar_stack_layers(struct float_data_t *thick_layers, int *layer_markers, int layers_count, int nz,
                       float scalez, int blank_value, struct float_data_t *result);

/* ---- extensions (no counterpart in the reference) ---- */

/* Message of the calling thread's last failed call ("" if it succeeded). */
const char *cvar_amd_last_error(void);

/* Statistics of the calling thread's last calc_variograms call: offsets in
 * the search tunnel, candidate pairs tested (informed cell x offset) and
 * device time in milliseconds. */
void cvar_amd_last_stats(long long *offsets, long long *pairs_tested, double *device_ms);

#ifdef __cplusplus
}
#endif

#endif /* HPGL_AMD_CVARIOGRAM_H */
