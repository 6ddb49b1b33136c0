// propio.cpp - property files: the Eclipse-style INC format (read and
// write) and the GSLIB single-variable format (write), as the reference's
// read_inc_file.cpp:8-200, property_writer.cpp:10-240 and
// api.cpp:54-193 define them.  Host code: the file is read whole and
// tokenised in memory instead of by one fscanf per token.
//
// INC reading (read_inc_file.cpp):
//   * the property name is the first line whose first character is a
//     letter; each line read is followed by skipping all white space (the
//     "\n" directive of "%1023[^\n]\n"), and lines longer than 1023
//     characters are read in 1023-character pieces;
//   * values are white-space separated tokens of at most 255 characters; a
//     token starting with "--" skips the rest of its line, a token starting
//     with "/" before all values are read is an error, any other token is
//     converted with strtof / strtol (the conversions behind scanf's %f and
//     %d: the longest valid prefix, an error when there is none);
//   * float mask = value != undefined; byte mask = int value != undefined,
//     data = the value as unsigned char, then (api.cpp:85-98) every
//     informed byte equal to values[j] becomes j.
// Two inputs the reference does not survive are errors here: an empty first
// line (its fscanf loop never advances) and a file with fewer values than
// requested (it re-parses its last token for the rest).
//
// Writing (property_writer.cpp): the name line, one value per line -
// "%E" of the value as a double, or "%d" of the byte as a signed char - the
// undefined value for cells whose mask is 0, and a closing "/" line (INC);
// GSLIB files start with "HPGL saved GSLIB file", "1" and the name.  Byte
// values go through the remap table of api.cpp:117-136.
#include "propio.h"

#include <cctype>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace hpgl_amd {

namespace {

[[noreturn]] void io_error(const char *where, const std::string &what) {
    throw std::runtime_error(std::string(where) + ": " + what);
}

std::vector<char> read_all(const char *where, const char *file_name) {
    std::unique_ptr<FILE, int (*)(FILE *)> f(std::fopen(file_name, "rb"), std::fclose);
    if (!f) io_error(where, std::string("Error opening file:") + file_name + ".");
    std::vector<char> buf;
    char chunk[1 << 16];
    size_t got;
    while ((got = std::fread(chunk, 1, sizeof(chunk), f.get())) > 0) buf.insert(buf.end(), chunk, chunk + got);
    if (std::ferror(f.get())) io_error(where, "Error reading file.");
    return buf;
}

struct Cursor {
    const std::vector<char> &b;
    size_t pos = 0;
    bool at_end() const { return pos >= b.size(); }
    void skip_space() {
        while (pos < b.size() && std::isspace(static_cast<unsigned char>(b[pos]))) ++pos;
    }
    // "%1023[^\n]\n": up to 1023 characters before a newline, then white
    // space; returns the characters read (empty when none matched)
    std::string line_piece() {
        const size_t start = pos;
        while (pos < b.size() && b[pos] != '\n' && pos - start < 1023) ++pos;
        std::string s(b.data() + start, pos - start);
        if (!s.empty()) skip_space();
        return s;
    }
    // "%255s"; false at the end of the data
    bool token(char *out) {
        skip_space();
        if (at_end()) return false;
        size_t n = 0;
        while (pos < b.size() && !std::isspace(static_cast<unsigned char>(b[pos])) && n < 255) out[n++] = b[pos++];
        out[n] = 0;
        return true;
    }
    // "%*[^\n]"
    void skip_line() {
        while (pos < b.size() && b[pos] != '\n') ++pos;
    }
};

std::string read_prop_name(Cursor &c) {
    for (;;) {
        std::string line = c.line_piece();
        if (line.empty()) {
            if (c.at_end()) io_error("read_prop_name", "Property name not found.");
            ++c.pos;  // an empty line (the reference would not advance past it)
            continue;
        }
        const bool named = std::isalpha(static_cast<unsigned char>(line[0])) != 0;
        std::string piece = line;
        while (piece.size() == 1023) {
            piece = c.line_piece();
            if (piece.empty()) break;
            if (named) line += piece;
        }
        if (named) return line;
    }
}

// The next value token, skipping "--" comments.
const char *next_value(Cursor &c, char *tok, const char *where) {
    for (;;) {
        if (!c.token(tok)) io_error(where, "Unexpected end of file.");
        const size_t len = std::strlen(tok);
        if (len >= 2 && tok[0] == '-' && tok[1] == '-') {
            c.skip_line();
            continue;
        }
        if (len >= 1 && tok[0] == '/') io_error(where, "Unexpected end of data.");
        return tok;
    }
}

void put(std::string &out, FILE *f) {
    if (out.size() >= (1u << 20)) {
        if (std::fwrite(out.data(), 1, out.size(), f) != out.size()) io_error("property_writer", "Error writing file.");
        out.clear();
    }
}

void flush(std::string &out, FILE *f) {
    if (!out.empty() && std::fwrite(out.data(), 1, out.size(), f) != out.size())
        io_error("property_writer", "Error writing file.");
    out.clear();
}

void append_double(std::string &out, double v) {
    char s[64];
    const int n = std::snprintf(s, sizeof(s), "%E\n", v);
    out.append(s, n);
}

void append_byte(std::string &out, unsigned char v) {
    char s[16];
    const int n = std::snprintf(s, sizeof(s), "%d\n", (int)(signed char)v);
    out.append(s, n);
}

std::unique_ptr<FILE, int (*)(FILE *)> open_out(const char *file_name) {
    std::unique_ptr<FILE, int (*)(FILE *)> f(std::fopen(file_name, "w+"), std::fclose);
    if (!f) io_error("open_file_checked", std::string("Can't open file '") + file_name + "'.");
    return f;
}

template <typename Emit>
void write_file(const char *file_name, const std::string &header, long long n, Emit emit, bool inc) {
    auto f = open_out(file_name);
    std::string out = header;
    for (long long i = 0; i < n; ++i) {
        emit(out, i);
        put(out, f.get());
    }
    if (inc) out += "/\n";
    flush(out, f.get());
    if (std::fflush(f.get()) != 0) io_error("property_writer", "Error writing file.");
}

}  // namespace

void read_inc_float(const char *file_name, float undefined_value, long long size, float *data, uint8_t *mask) {
    if (size < 0) io_error("read_inc_file_float", "negative size");
    const std::vector<char> buf = read_all("read_inc_file_float", file_name);
    Cursor c{buf};
    read_prop_name(c);
    char tok[256];
    for (long long i = 0; i < size; ++i) {
        next_value(c, tok, "load_floats_into_vector");
        char *end = nullptr;
        const float v = std::strtof(tok, &end);
        if (end == tok) io_error("load_floats_into_vector", std::string("Error parsing '") + tok + "' string.");
        data[i] = v;
    }
    if (mask)
        for (long long i = 0; i < size; ++i) mask[i] = data[i] == undefined_value ? 0 : 1;
}

void read_inc_byte(const char *file_name, int undefined_value, long long size, uint8_t *data, uint8_t *mask) {
    if (size < 0) io_error("read_inc_file_byte", "negative size");
    const std::vector<char> buf = read_all("read_inc_file_float", file_name);
    Cursor c{buf};
    read_prop_name(c);
    char tok[256];
    for (long long i = 0; i < size; ++i) {
        next_value(c, tok, "read_bytes");
        char *end = nullptr;
        const long v = std::strtol(tok, &end, 10);
        if (end == tok) io_error("load_floats_into_vector", std::string("Error parsing '") + tok + "' string.");
        const int value = (int)v;
        data[i] = (uint8_t)value;
        mask[i] = value == undefined_value ? 0 : 1;
    }
}

std::vector<uint8_t> remap_table(const uint8_t *values, int values_count, int indicator_count) {
    std::vector<uint8_t> t;
    if (values_count >= indicator_count && values_count > 0)
        t.assign(values, values + indicator_count);
    else
        for (int i = 0; i < indicator_count; ++i) t.push_back((uint8_t)i);
    return t;
}

void write_inc_float(const char *file_name, const char *name, const float *data, const uint8_t *mask, long long n,
                     double undefined_value) {
    write_file(
        file_name, std::string(name) + "\n", n,
        [&](std::string &out, long long i) { append_double(out, mask[i] ? (double)data[i] : undefined_value); }, true);
}

void write_inc_byte(const char *file_name, const char *name, const uint8_t *data, const uint8_t *mask, long long n,
                    uint8_t undefined_value, const std::vector<uint8_t> &remap) {
    write_file(
        file_name, std::string(name) + "\n", n,
        [&](std::string &out, long long i) {
            if (!mask[i]) {
                append_byte(out, undefined_value);
                return;
            }
            if (data[i] >= remap.size()) io_error("write_property_ind", "indicator value outside the remap table");
            append_byte(out, remap[data[i]]);
        },
        true);
}

void write_gslib_float(const char *file_name, const char *name, const float *data, const uint8_t *mask, long long n,
                       double undefined_value) {
    write_file(
        file_name, std::string("HPGL saved GSLIB file\n1\n") + name + "\n", n,
        [&](std::string &out, long long i) { append_double(out, mask[i] ? (double)data[i] : undefined_value); }, false);
}

void write_gslib_byte(const char *file_name, const char *name, const uint8_t *data, const uint8_t *mask, long long n,
                      uint8_t undefined_value, const std::vector<uint8_t> &remap) {
    write_file(
        file_name, std::string("HPGL saved GSLIB file\n1\n") + name + "\n", n,
        [&](std::string &out, long long i) {
            if (!mask[i]) {
                append_byte(out, undefined_value);
                return;
            }
            if (data[i] >= remap.size()) io_error("write_gslib_property_ind", "indicator value outside the remap table");
            append_byte(out, remap[data[i]]);
        },
        false);
}

}  // namespace hpgl_amd
