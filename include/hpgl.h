/*
 * hpgl.h - C ABI of the MI355X kriging library (hpgl_amd/hpgl.so).
 *
 * Part 1 declares the drop-in boundary: the same structs (same field order,
 * same LP64 layout) and the same extern "C" entry points as the reference's
 * C ABI, so that its ctypes front end (geo_bsd/hpgl_wrap.py) binds to this
 * library unchanged.  Each declaration cites the reference declaration it
 * replaces (geo_bsd/hpgl/api.h:LINE) and its implementation (api.cpp:LINE).
 *
 * Part 2 declares entry points with no counterpart in the reference, all
 * prefixed hpgl_amd_: precision selection, per-call statistics, a neighbour
 * dump for parity checks, and device-pointer variants of the kriging calls
 * for callers whose data already lives in HBM.
 *
 * All buffers are dense Fortran order (x fastest): cell index =
 * x + X*y + X*Y*z.  The reference ignores the strides field, and so does
 * this library.
 *
 * Errors: the reference lets C++ exceptions escape through extern "C"
 * (api.cpp:31-35 is never called).  This library catches every exception
 * at the boundary, stores its text for hpgl_get_last_exception_message()
 * and returns; hpgl_amd_get_last_error() returns non-zero after a failed call.
 */
#ifndef HPGL_AMD_HPGL_H
#define HPGL_AMD_HPGL_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Part 1: structs, identical layout to geo_bsd/hpgl/api.h:18-153 ---- */

struct hpgl_shape_t {               /* api.h:18, 24 bytes */
    int m_data[3];
    int m_strides[3];
};

struct hpgl_cont_masked_array_t {   /* api.h:24, 40 bytes */
    float *m_data;
    unsigned char *m_mask;
    struct hpgl_shape_t m_shape;
};

struct hpgl_ind_masked_array_t {    /* api.h:31, 48 bytes */
    unsigned char *m_data;
    unsigned char *m_mask;
    struct hpgl_shape_t m_shape;
    int m_indicator_count;
};

struct hpgl_float_array_t {         /* api.h:39, 32 bytes */
    float *m_data;
    struct hpgl_shape_t m_shape;
};

struct hpgl_ubyte_array_t {         /* api.h:45, 32 bytes */
    unsigned char *m_data;
    struct hpgl_shape_t m_shape;
};

struct hpgl_ok_params_t {           /* api.h:51, 88 bytes */
    int m_covariance_type;          /* 0 spherical, 1 exponential, 2 gaussian */
    double m_ranges[3];
    double m_angles[3];             /* degrees */
    double m_sill;
    double m_nugget;
    int m_radiuses[3];
    int m_max_neighbours;
};

struct hpgl_sk_params_t {           /* api.h:62, 104 bytes */
    int m_covariance_type;
    double m_ranges[3];
    double m_angles[3];
    double m_sill;
    double m_nugget;
    int m_radiuses[3];
    int m_max_neighbours;
    unsigned char m_automatic_mean;
    double m_mean;
};

struct hpgl_cov_params_t {          /* api.h:75, 72 bytes */
    int m_covariance_type;
    double m_ranges[3];
    double m_angles[3];
    double m_sill;
    double m_nugget;
};

struct hpgl_sgs_params_t {          /* api.h:84, 112 bytes */
    int m_covariance_type;
    double m_ranges[3];
    double m_angles[3];
    double m_sill;
    double m_nugget;
    int m_radiuses[3];
    int m_max_neighbours;
    int m_kriging_kind;             /* 0 ordinary, 1 simple (sgs_params.h:10) */
    long m_seed;
    int m_min_neighbours;           /* accepted and ignored, as in the reference */
};

struct hpgl_median_ik_params_t {    /* api.h:98, 104 bytes */
    int m_covariance_type;
    double m_ranges[3];
    double m_angles[3];
    double m_sill;
    double m_nugget;
    int m_radiuses[3];
    int m_max_neighbours;
    double m_marginal_probs[2];
};

struct hpgl_ik_params_t {           /* api.h:110, 96 bytes */
    int m_covariance_type;
    double m_ranges[3];
    double m_angles[3];
    double m_sill;
    double m_nugget;
    int m_radiuses[3];
    int m_max_neighbours;
    double m_marginal_prob;
};

struct hpgl_cokriging_m1_params_t { /* api.h:122, 120 bytes */
    int m_covariance_type;
    double m_ranges[3];
    double m_angles[3];
    double m_sill;
    double m_nugget;
    int m_radiuses[3];
    int m_max_neighbours;
    double m_primary_mean;
    double m_secondary_mean;
    double m_secondary_variance;
    double m_correlation_coef;
};

struct hpgl_cokriging_m2_params_t { /* api.h:137, 184 bytes */
    struct hpgl_cov_params_t m_primary_cov_params;
    struct hpgl_cov_params_t m_secondary_cov_params;
    int m_radiuses[3];
    int m_max_neighbours;
    double m_primary_mean;
    double m_secondary_mean;
    double m_correlation_coef;
};

struct hpgl_non_parametric_cdf_t {  /* api.h:148, 24 bytes */
    float *m_values;
    float *m_probs;
    long long m_size;
};

/* ---- Part 1: entry points (drop-in for geo_bsd/hpgl/api.h:155-313) ---- */

/* api.h:159-160, api.cpp:39 */
char *hpgl_get_last_exception_message(void);
/* api.h:162-163 (output.cpp:62) */
void hpgl_set_output_handler(int (*handler)(char *str, void *param), void *param);
/* api.h:165-166 (output.cpp:68) */
void hpgl_set_progress_handler(int (*handler)(char *stage, int percentage, void *param), void *param);
/* api.h:168-169, api.cpp:44-52: kept for ABI compatibility; the thread
 * count sizes the host worker pool used for per-job set-up. */
void hpgl_set_thread_num(int n_threads);
int hpgl_get_thread_num(void);

/* api.h:171-214, api.cpp:54-193: property files (host code).  INC files are
 * read as read_inc_file.cpp:8-200 reads them (byte values then mapped to
 * their index in values[]); INC and GSLIB files are written as
 * property_writer.cpp writes them.  Errors set the last error message
 * (the reference throws). */
void hpgl_read_inc_file_float(char *filename, float undefined_value, int size,
                              float *data, unsigned char *mask);
void hpgl_read_inc_file_byte(char *filename, int undefined_value, int size,
                             unsigned char *data, unsigned char *mask,
                             unsigned char *values, int values_count);
void hpgl_write_inc_file_float(char *filename, struct hpgl_cont_masked_array_t *arr,
                               float undefined_value, char *name);
void hpgl_write_inc_file_byte(char *filename, struct hpgl_ind_masked_array_t *arr,
                              int undefined_value, char *name,
                              unsigned char *values, int values_count);
void hpgl_write_gslib_cont_property(struct hpgl_cont_masked_array_t *data,
                                    const char *filename, const char *name,
                                    double undefined_value);
void hpgl_write_gslib_byte_property(struct hpgl_ind_masked_array_t *data,
                                    const char *filename, const char *name,
                                    double undefined_value, unsigned char *values,
                                    int values_count);

/* api.h:218-221, api.cpp:195-231 (ordinary_kriging.cpp:14-46).
 * output_data arrives holding a copy of the input (geo.py:383); informed
 * cells are copied, kriged cells get value+mask=1, failed cells are left
 * untouched (undefined_on_failure). */
void hpgl_ordinary_kriging(struct hpgl_cont_masked_array_t *input_data,
                           struct hpgl_ok_params_t *params,
                           struct hpgl_cont_masked_array_t *output_data);

/* api.h:223-230, api.cpp:261-290 (simple_kriging.cpp:87-120).
 * Failed cells get the mean (mean_on_failure). */
void hpgl_simple_kriging(float *input_data, unsigned char *input_mask,
                         struct hpgl_shape_t *input_data_shape,
                         struct hpgl_sk_params_t *params,
                         float *output_data, unsigned char *output_mask,
                         struct hpgl_shape_t *output_data_shape);

/* api.h:232-240, api.cpp:292-342 (simple_kriging.cpp:65-85).  Returns 0.
 * Weights are left unwritten when the system is singular, as in the
 * reference, except that the reference then throws from std::vector::at. */
int hpgl_simple_kriging_weights(float *center_coords, float *neighbours_x,
                                float *neighbours_y, float *neighbours_z,
                                int neighbours_count, struct hpgl_cov_params_t *params,
                                float *weights);

/* No counterpart in the reference: a variant of hpgl_simple_kriging_weights
 * that allocates nothing, for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops).  center (3 floats),
 * xs, ys, zs (n floats) and weights (n floats) are device pointers; the
 * work is queued on `stream` and the call returns after it completes.  The
 * system and its factorisation use a static device array of about 2 MiB per
 * device: n is limited to 512 (larger n sets the last error), and calls
 * that overlap on one GPU race.  Returns 0; a singular system leaves the
 * weights unwritten and sets the last error, as hpgl_simple_kriging_weights
 * does. */
int hpgl_amd_simple_kriging_weights_noalloc(const float *center, const float *xs, const float *ys,
                                            const float *zs, int n, const struct hpgl_cov_params_t *params,
                                            float *weights, void *stream);

/* api.h:242-251, api.cpp:344-383 (lvm_kriging.cpp:16-37). */
void hpgl_lvm_kriging(float *input_data, unsigned char *input_mask,
                      struct hpgl_shape_t *input_data_shape,
                      float *mean_data, struct hpgl_shape_t *mean_data_shape,
                      struct hpgl_ok_params_t *params,
                      float *output_data, unsigned char *output_mask,
                      struct hpgl_shape_t *output_data_shape);

/* api.h:253-258, api.cpp:408-438 (sequential_gaussian_simulation.cpp:18-75).
 * In place on prop.  cdf, mean and simulation_mask may be NULL. */
void hpgl_sgs_simulation(struct hpgl_cont_masked_array_t *prop,
                         struct hpgl_sgs_params_t *params,
                         struct hpgl_non_parametric_cdf_t *cdf,
                         double *mean,
                         struct hpgl_ubyte_array_t *simulation_mask);

/* api.h:260-265, api.cpp:440-467 (sequential_gaussian_simulation.cpp:77-128). */
void hpgl_sgs_lvm_simulation(struct hpgl_cont_masked_array_t *data,
                             struct hpgl_sgs_params_t *params,
                             struct hpgl_non_parametric_cdf_t *cdf,
                             struct hpgl_float_array_t *means,
                             struct hpgl_ubyte_array_t *simulation_mask);

/* api.h:267-272, api.cpp:385-406 (indicator_kriging.h:58-124).
 * params points to indicator_count consecutive structs. */
void hpgl_indicator_kriging(struct hpgl_ind_masked_array_t *in_data,
                            struct hpgl_ind_masked_array_t *out_data,
                            struct hpgl_ik_params_t *params,
                            int indicator_count);

/* api.h:274-277, api.cpp:469-515 (median_ik.cpp:24-93). */
void hpgl_median_ik(struct hpgl_ind_masked_array_t *in_data,
                    struct hpgl_median_ik_params_t *params,
                    struct hpgl_ind_masked_array_t *out_data);

/* api.h:279-285, api.cpp:517-545 (sequential_indicator_simulation.cpp:33-157).
 * In place on data.  params points to indicator_count consecutive structs
 * (m_marginal_prob of each category); indicator_count == 2 is median SIS.
 * simulation_mask may be NULL. */
void hpgl_sis_simulation(struct hpgl_ind_masked_array_t *data,
                         struct hpgl_ik_params_t *params,
                         int indicator_count,
                         int seed,
                         struct hpgl_ubyte_array_t *simulation_mask);

/* api.h:287-295, api.cpp:547-585 (sequential_indicator_simulation.cpp:159-187).
 * mean_data: indicator_count per-cell mean arrays.  use_correlograms != 0
 * selects the correlogram weights (corellogramed_weights_3,
 * my_kriging_weights.h:258-362); those systems are solved in FP64. */
void hpgl_sis_simulation_lvm(struct hpgl_ind_masked_array_t *data,
                             struct hpgl_ik_params_t *params,
                             struct hpgl_float_array_t *mean_data,
                             int indicator_count,
                             int seed,
                             struct hpgl_ubyte_array_t *simulation_mask,
                             int use_correlograms);

/* api.h:297-302, api.cpp:587-652 (simple_cokriging_markI.cpp:185-242).
 * Simple colocated cokriging, Markov model I: every cell of output_data is
 * written (informed primary cells copy the input).  The systems are solved
 * in FP64 in gauss_solve's operation order whatever the precision setting. */
void hpgl_simple_cokriging_mark1(struct hpgl_cont_masked_array_t *input_data,
                                 struct hpgl_cont_masked_array_t *secondary_data,
                                 struct hpgl_cokriging_m1_params_t *params,
                                 struct hpgl_cont_masked_array_t *output_data);

/* api.h:304-309, api.cpp:654-712 (simple_cokriging_markI.cpp:244-305).
 * Markov model II: the cross covariance follows the secondary model. */
void hpgl_simple_cokriging_mark2(struct hpgl_cont_masked_array_t *primary_data,
                                 struct hpgl_cont_masked_array_t *secondary_data,
                                 struct hpgl_cokriging_m2_params_t *params,
                                 struct hpgl_cont_masked_array_t *output_data);

/* ---- Part 2: hpgl_amd_* extensions (no counterpart in the reference) ---- */

/* Precision of the per-node solves.  MIXED (default): FP32 Cholesky, one or
 * two FP64-residual refinement steps, FP64 Cholesky fallback on breakdown,
 * small pivots or non-converged refinement.  FP64: every system factorised
 * in FP64 in the reference's dpotf2/dpotrs operation order. */
#define HPGL_AMD_PRECISION_MIXED 0
#define HPGL_AMD_PRECISION_FP64 1
void hpgl_amd_set_precision(int mode);
int hpgl_amd_get_precision(void);

/* Number of HIP devices visible to the library (0 without a GPU). */
int hpgl_amd_device_count(void);

/* Diagnostic with no counterpart in the reference: checks that a device
 * block the library frees while kernels that use it are still queued on one
 * stream is not handed to another stream before they complete.  Returns 0
 * when that holds, 1 when the check was inconclusive (the probe kernel ended
 * first), negative when the ordering is broken, -100 on an error. */
int hpgl_amd_pool_selftest(void);

/* Non-zero when the last call on this thread failed; the message is
 * available from hpgl_get_last_exception_message(). */
int hpgl_amd_get_last_error(void);

/* Statistics of the last kriging/simulation call. */
struct hpgl_amd_stats_t {
    long long nodes_target;         /* cells submitted to the search */
    long long nodes_kriged;         /* successful solves */
    long long nodes_no_neighbours;
    long long nodes_singular;
    long long solves_fp32;          /* accepted after FP32 + refinement */
    long long solves_fp64;          /* solved by the FP64 path */
    long long nodes_template_path;  /* searched by template scan */
    long long nodes_cluster_path;   /* searched through the cell-binned index */
    double ms_setup_host;           /* host set-up: tables, template sort */
    double ms_search;               /* device time, search kernels */
    double ms_solve;                /* device time, build+solve+combine kernels */
    double ms_device_total;         /* device time of the whole call */
    double ms_wall;                 /* wall time of the call */
    double ms_k_template;           /* device time of the template-scan kernel launches */
    double ms_k_cluster;            /* ... cell-binned search kernel launches */
    double ms_k_solve_mixed;        /* ... mixed-precision solve kernel launches */
    double ms_k_solve_fp64;         /* ... FP64 solve kernel launches */
    long long n_k_template;         /* launch counts of the same kernels */
    long long n_k_cluster;
    long long n_k_solve_mixed;
    long long n_k_solve_fp64;
    long long cluster_candidates;   /* candidate cells gathered by the cell-binned search */
    long long neighbours_total;     /* neighbours over all searched targets */
    long long cluster_lists_spilled; /* cell-binned candidate lists sorted in global scratch */
    long long solves_fp32_two_steps; /* FP32 solves accepted after a second refinement step */
};
void hpgl_amd_get_last_stats(struct hpgl_amd_stats_t *out);

/* When indices != NULL, the next kriging call (OK/SK/LVM/median IK, or IK
 * with category 0..K-1 laid out one after another) writes each cell's
 * neighbour list, in search order, to indices[cell*max_n ...] and its count
 * to counts[cell] (-1 for cells that are not estimated).  n_cells and max_n
 * describe the caller's buffers; the dump is cleared after one call. */
void hpgl_amd_set_neighbour_dump(int *indices, int *counts, long long n_cells, int max_n);

/* Device-pointer variants: all array arguments are device (HBM) pointers,
 * the work is enqueued on `stream` (a hipStream_t, NULL = default stream)
 * and the call returns after the stream has finished.  out arrays must be
 * pre-filled by the caller exactly as for the host variants. */
void hpgl_amd_ordinary_kriging_dev(const float *d_in, const unsigned char *d_in_mask,
                                   const struct hpgl_shape_t *shape,
                                   const struct hpgl_ok_params_t *params,
                                   float *d_out, unsigned char *d_out_mask,
                                   void *stream);
void hpgl_amd_simple_kriging_dev(const float *d_in, const unsigned char *d_in_mask,
                                 const struct hpgl_shape_t *shape,
                                 const struct hpgl_sk_params_t *params,
                                 float *d_out, unsigned char *d_out_mask,
                                 void *stream);
void hpgl_amd_lvm_kriging_dev(const float *d_in, const unsigned char *d_in_mask,
                              const struct hpgl_shape_t *shape, const float *d_mean,
                              const struct hpgl_ok_params_t *params,
                              float *d_out, unsigned char *d_out_mask,
                              void *stream);
void hpgl_amd_indicator_kriging_dev(const unsigned char *d_in, const unsigned char *d_in_mask,
                                    const struct hpgl_shape_t *shape,
                                    const struct hpgl_ik_params_t *params, int indicator_count,
                                    unsigned char *d_out, unsigned char *d_out_mask,
                                    void *stream);

/* R realisations of sequential Gaussian simulation in one call.  Each
 * realisation r is what hpgl_sgs_simulation (lvm_means == NULL) or
 * hpgl_sgs_lvm_simulation would return with params->m_seed = seeds[r];
 * prop is not modified, realisation r goes to out[r*N ...] and
 * out_mask[r*N ...] (N cells each).  Realisations share one set of device
 * launches: path, search and solve kernels run over all of them together
 * and one ordered sweep writes all their values.  No reference counterpart:
 * the reference runs one realisation per call. */
void hpgl_amd_sgs_simulation_batch(const struct hpgl_cont_masked_array_t *prop,
                                   const struct hpgl_sgs_params_t *params,
                                   const struct hpgl_non_parametric_cdf_t *cdf,
                                   const struct hpgl_float_array_t *lvm_means,
                                   const struct hpgl_ubyte_array_t *simulation_mask,
                                   const long *seeds, int R,
                                   float *out, unsigned char *out_mask);
/* Device-pointer variant of hpgl_amd_sgs_simulation_batch; cdf arrays and
 * seeds stay on the host.  d_out and d_out_mask hold R*N entries. */
void hpgl_amd_sgs_simulation_batch_dev(const float *d_in, const unsigned char *d_in_mask,
                                       const struct hpgl_shape_t *shape,
                                       const struct hpgl_sgs_params_t *params,
                                       const struct hpgl_non_parametric_cdf_t *cdf,
                                       const float *d_lvm_mean, const unsigned char *d_sim_mask,
                                       const long *seeds, int R,
                                       float *d_out, unsigned char *d_out_mask, void *stream);

/* R realisations of sequential indicator simulation in one call.
 * Realisation r is what hpgl_sis_simulation (mean_data == NULL) or
 * hpgl_sis_simulation_lvm would return with seed = seeds[r]; data is not
 * modified, realisation r goes to out[r*N ...] and out_mask[r*N ...].  No
 * reference counterpart: the reference runs one realisation per call. */
void hpgl_amd_sis_simulation_batch(const struct hpgl_ind_masked_array_t *data,
                                   const struct hpgl_ik_params_t *params,
                                   int indicator_count,
                                   const struct hpgl_float_array_t *mean_data,
                                   int use_correlograms,
                                   const struct hpgl_ubyte_array_t *simulation_mask,
                                   const int *seeds, int R,
                                   unsigned char *out, unsigned char *out_mask);

/* Device-pointer variant of hpgl_amd_sis_simulation_batch: d_lvm holds
 * indicator_count consecutive per-cell mean arrays (or NULL); seeds stay on
 * the host.  d_out and d_out_mask hold R*N entries. */
void hpgl_amd_sis_simulation_batch_dev(const unsigned char *d_in, const unsigned char *d_in_mask,
                                       const struct hpgl_shape_t *shape,
                                       const struct hpgl_ik_params_t *params, int indicator_count,
                                       const float *d_lvm, int use_correlograms,
                                       const unsigned char *d_sim_mask, const int *seeds, int R,
                                       unsigned char *d_out, unsigned char *d_out_mask, void *stream);

/* Device-pointer variants of the cokriging entry points (arrays of N
 * cells already in HBM; the work is queued on `stream` and the call returns
 * when it has finished).  d_out and d_out_mask are written for every cell. */
void hpgl_amd_simple_cokriging_mark1_dev(const float *d_in, const unsigned char *d_in_mask,
                                         const float *d_sec, const unsigned char *d_sec_mask,
                                         const struct hpgl_shape_t *shape,
                                         const struct hpgl_cokriging_m1_params_t *params,
                                         float *d_out, unsigned char *d_out_mask, void *stream);
void hpgl_amd_simple_cokriging_mark2_dev(const float *d_in, const unsigned char *d_in_mask,
                                         const float *d_sec, const unsigned char *d_sec_mask,
                                         const struct hpgl_shape_t *shape,
                                         const struct hpgl_cokriging_m2_params_t *params,
                                         float *d_out, unsigned char *d_out_mask, void *stream);

/* ---- Request loop with continuous batching (no reference counterpart:
 * the reference runs one blocking call at a time).  Jobs wait in arrival
 * order; a free worker slot (own HIP stream) starts the first waiting job
 * whose estimated HBM footprint fits the unused budget, so jobs overlap on
 * the device; a job larger than the budget runs alone.  A job's result is
 * the one its entry point returns alone.  The caller keeps every buffer of
 * a job alive until hpgl_amd_server_wait returns for it. ---- */
#define HPGL_AMD_JOB_OK 0          /* hpgl_ordinary_kriging;  params: hpgl_ok_params_t */
#define HPGL_AMD_JOB_SK 1          /* hpgl_simple_kriging;    params: hpgl_sk_params_t */
#define HPGL_AMD_JOB_LVM 2         /* hpgl_lvm_kriging;       params: hpgl_ok_params_t, mean */
#define HPGL_AMD_JOB_IK 3          /* hpgl_indicator_kriging; params: hpgl_ik_params_t[indicator_count] */
#define HPGL_AMD_JOB_MEDIAN_IK 4   /* hpgl_median_ik;         params: hpgl_median_ik_params_t */
#define HPGL_AMD_JOB_SGS 5         /* hpgl_sgs_simulation / _lvm_ (mean != NULL), in place on cont_* */

struct hpgl_amd_job_t {
    int kind;
    struct hpgl_shape_t shape;
    float *cont_data;               /* OK, SK, LVM input; SGS property (in place) */
    unsigned char *cont_mask;
    unsigned char *ind_data;        /* IK, median IK input */
    unsigned char *ind_mask;
    int indicator_count;
    float *mean;                    /* LVM and LVM SGS per-cell mean, else NULL */
    void *params;
    struct hpgl_non_parametric_cdf_t *cdf;  /* SGS, may be NULL */
    unsigned char *sim_mask;        /* SGS, may be NULL */
    float *out_cont;                /* OK, SK, LVM output (prefilled as for the entry point) */
    unsigned char *out_cont_mask;
    unsigned char *out_ind;         /* IK, median IK output */
    unsigned char *out_ind_mask;
};

struct hpgl_amd_job_result_t {
    int status;                     /* 0 = success, else the call failed (message) */
    double queue_ms;                /* submit to start */
    double latency_ms;              /* submit to completion */
    double run_ms;                  /* start to completion */
    long long nodes;                /* cells estimated (x categories for IK) */
    long long hbm_bytes;            /* footprint the job was admitted with */
    double start_ms;                /* start and completion on the server's clock */
    double end_ms;                  /* (milliseconds since the server was created) */
    char message[256];
};

/* slots = worker threads; hbm_budget_bytes <= 0 takes half the free HBM.
 * Returns NULL on failure (message via hpgl_get_last_exception_message). */
void *hpgl_amd_server_create(int slots, long long hbm_budget_bytes);
/* Returns the job id, or -1 on failure. */
long long hpgl_amd_server_submit(void *server, const struct hpgl_amd_job_t *job);
/* Blocks until the job has finished; 0 on success, -1 for an unknown id. */
int hpgl_amd_server_wait(void *server, long long id, struct hpgl_amd_job_result_t *result);
/* Waits for every submitted job, then stops the workers. */
void hpgl_amd_server_destroy(void *server);

#ifdef __cplusplus
}
#endif
#endif
