// NetCDF classic-format files written and read from host memory.
//
// The format is the published "NetCDF Classic and 64-bit Offset Format" (CDF-2, magic
// "CDF\x02": 64-bit variable offsets, at most 4 GiB per fixed-size variable) and its
// 64-bit-data extension (CDF-5, magic "CDF\x05": 64-bit counts and sizes), which a file
// switches to when one variable is larger than CDF-2 allows.  Every value is big-endian.
// Only fixed-size variables are written (no record dimension): the reference's output
// file has its time axis N_t as an ordinary dimension (MainSrc/OUTPUTS.f08:51).
#pragma once

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace trt {
namespace nc {

enum Type : int { NC_BYTE = 1, NC_CHAR = 2, NC_SHORT = 3, NC_INT = 4, NC_FLOAT = 5, NC_DOUBLE = 6 };
size_t type_size(int type);

struct Att {
    std::string name;
    int type = NC_CHAR;
    size_t n = 0;                     // number of values
    std::vector<unsigned char> raw;   // the values in host byte order
};

struct Dim {
    std::string name;
    uint64_t len = 0;
};

struct Var {
    std::string name;
    int type = NC_DOUBLE;
    std::vector<int> dims;            // dimension ids, slowest-varying first (C order)
    std::vector<Att> atts;
    uint64_t begin = 0;               // file offset of the data
};

struct Header {
    int version = 2;
    std::vector<Dim> dims;
    std::vector<Att> gatts;
    std::vector<Var> vars;
    uint64_t nelem(const Var &v) const;
    uint64_t vbytes(const Var &v) const { return nelem(v) * type_size(v.type); }
    int find_var(const std::string &name) const;
    int find_dim(const std::string &name) const;
};

// Serialised header of h (begins as they are set).
std::vector<unsigned char> encode_header(const Header &h);
// Parses the header at the start of f; throws trt::Error on anything that is not a
// CDF-1, CDF-2 or CDF-5 header.
Header decode_header(FILE *f);

class Writer {
public:
    Writer() = default;
    ~Writer();
    Writer(const Writer &) = delete;
    Writer &operator=(const Writer &) = delete;
    void create(const std::string &path);
    int def_dim(const std::string &name, uint64_t len);
    int def_var(const std::string &name, int type, const std::vector<int> &dims);
    // varid -1: global attribute.  vals are n values of type in host byte order.
    void put_att(int varid, const std::string &name, int type, const void *vals, size_t n);
    void put_att_text(int varid, const std::string &name, const std::string &s);
    // Leaves define mode: lays the variables out one after the other behind the header and
    // `room` spare bytes, writes the header and, with fill, every variable's fill value.
    void enddef(bool fill, size_t room);
    // Defines more dimensions and variables after enddef (define mode again); the next
    // enddef places the new variables behind all existing data and rewrites the header in
    // place, which must still fit in front of the first variable.
    void redef();
    // Writes count[] values at start[] (C order) of the variable, given in its type and
    // host byte order.
    void put(int varid, const uint64_t *start, const uint64_t *count, const void *vals);
    void close();
    // 5: always CDF-5; 0: CDF-2 unless a variable exceeds its 4 GiB limit
    void set_format(int v) { force5_ = v == 5; }
    bool is_open() const { return f_ != nullptr; }
    const Header &header() const { return h_; }

private:
    void write_at(uint64_t off, const void *p, size_t n);
    void fill_var(const Var &v);
    void check_define(const char *what) const;
    FILE *f_ = nullptr;
    std::string path_;
    Header h_;
    bool define_ = true;
    bool force5_ = false;
    bool laid_out_ = false;
    size_t nlaid_ = 0;         // variables with a place in the file
    uint64_t data_begin_ = 0;  // offset of the first variable
    uint64_t end_ = 0;         // end of the data
};

class Reader {
public:
    ~Reader();
    void open(const std::string &path);
    const Header &header() const { return h_; }
    // count[] values at start[] of the variable as FP64 (NC_BYTE/SHORT/INT/FLOAT/DOUBLE)
    void get(int varid, const uint64_t *start, const uint64_t *count, double *out);
    void close();

private:
    FILE *f_ = nullptr;
    Header h_;
};

}  // namespace nc
}  // namespace trt
