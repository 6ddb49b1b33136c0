// propio.h - property files (INC read/write, GSLIB write), host code.
#pragma once

#include <cstdint>
#include <vector>

namespace hpgl_amd {

// Errors are thrown as std::runtime_error("<where>: <message>").
void read_inc_float(const char *file_name, float undefined_value, long long size, float *data, uint8_t *mask);
void read_inc_byte(const char *file_name, int undefined_value, long long size, uint8_t *data, uint8_t *mask);
// api.cpp:117-136: values when there are at least as many as indicators
// (the first indicator_count), else 0, 1, ..., indicator_count - 1.
std::vector<uint8_t> remap_table(const uint8_t *values, int values_count, int indicator_count);
void write_inc_float(const char *file_name, const char *name, const float *data, const uint8_t *mask, long long n,
                     double undefined_value);
void write_inc_byte(const char *file_name, const char *name, const uint8_t *data, const uint8_t *mask, long long n,
                    uint8_t undefined_value, const std::vector<uint8_t> &remap);
void write_gslib_float(const char *file_name, const char *name, const float *data, const uint8_t *mask, long long n,
                       double undefined_value);
void write_gslib_byte(const char *file_name, const char *name, const uint8_t *data, const uint8_t *mask, long long n,
                      uint8_t undefined_value, const std::vector<uint8_t> &remap);

}  // namespace hpgl_amd
