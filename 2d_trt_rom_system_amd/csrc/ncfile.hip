// NetCDF classic-format files (CDF-2; CDF-5 when a variable outgrows CDF-2): header
// encoding and decoding, hyperslab writes and reads.  Host code only.
//
// Header grammar (the published classic format specification):
//   header   = magic numrecs dim_list gatt_list var_list
//   dim_list = ABSENT | NC_DIMENSION nelems dim*        dim  = name dim_length
//   att_list = ABSENT | NC_ATTRIBUTE nelems attr*       attr = name nc_type nelems values
//   var_list = ABSENT | NC_VARIABLE  nelems var*        var  = name nelems dimid* vatt_list
//                                                                nc_type vsize begin
// Tags and nc_type are 32-bit; nelems, lengths, dimids and vsize are 32-bit in CDF-2 and
// 64-bit in CDF-5; begin is 64-bit in both; names and values are padded to 4 bytes.
// Fixed-size data: each variable at its begin, row-major (C order), padded to 4 bytes.
#include "ncfile.h"

#include <fcntl.h>
#include <unistd.h>

#include <algorithm>
#include <cstring>

#include "trt_internal.h"

namespace trt {
namespace nc {

namespace {

constexpr uint32_t TAG_DIMENSION = 0x0A, TAG_VARIABLE = 0x0B, TAG_ATTRIBUTE = 0x0C;
constexpr uint64_t CDF2_VAR_MAX = 4294967292ull;   // 2^32 - 4

uint64_t round4(uint64_t n) { return (n + 3) & ~uint64_t(3); }

// big-endian copy of n values of size es
void to_be(unsigned char *dst, const void *src, size_t n, size_t es)
{
    const unsigned char *s = (const unsigned char *)src;
    if (es == 1) { memcpy(dst, s, n); return; }
    for (size_t i = 0; i < n; ++i)
        for (size_t b = 0; b < es; ++b) dst[i * es + b] = s[i * es + (es - 1 - b)];
}

struct Enc {
    std::vector<unsigned char> b;
    bool wide = false;   // CDF-5 counts
    void u32(uint32_t v) { for (int s = 24; s >= 0; s -= 8) b.push_back((unsigned char)(v >> s)); }
    void u64(uint64_t v) { for (int s = 56; s >= 0; s -= 8) b.push_back((unsigned char)(v >> s)); }
    void count(uint64_t v)
    {
        if (wide) u64(v);
        else {
            if (v > 0xFFFFFFFFull) throw Error("netCDF: count too large for CDF-2");
            u32((uint32_t)v);
        }
    }
    void pad() { while (b.size() % 4) b.push_back(0); }
    void name(const std::string &s)
    {
        count(s.size());
        b.insert(b.end(), s.begin(), s.end());
        pad();
    }
    void absent() { u32(0); count(0); }
    void atts(const std::vector<Att> &a)
    {
        if (a.empty()) { absent(); return; }
        u32(TAG_ATTRIBUTE);
        count(a.size());
        for (const Att &t : a) {
            name(t.name);
            u32((uint32_t)t.type);
            count(t.n);
            const size_t es = type_size(t.type), at = b.size();
            b.resize(at + t.n * es);
            to_be(b.data() + at, t.raw.data(), t.n, es);
            pad();
        }
    }
};

struct Dec {
    FILE *f;
    bool wide = false;
    uint32_t u32()
    {
        unsigned char c[4];
        if (fread(c, 1, 4, f) != 4) throw Error("netCDF: truncated header");
        return (uint32_t)c[0] << 24 | (uint32_t)c[1] << 16 | (uint32_t)c[2] << 8 | c[3];
    }
    uint64_t u64() { uint64_t hi = u32(); return hi << 32 | u32(); }
    uint64_t count() { return wide ? u64() : u32(); }
    void skip_pad(uint64_t n)
    {
        const uint64_t p = round4(n) - n;
        unsigned char c[4];
        if (p && fread(c, 1, p, f) != p) throw Error("netCDF: truncated header");
    }
    std::string name()
    {
        const uint64_t n = count();
        if (n > (1u << 20)) throw Error("netCDF: implausible name length");
        std::string s(n, '\0');
        if (n && fread(&s[0], 1, n, f) != n) throw Error("netCDF: truncated header");
        skip_pad(n);
        return s;
    }
    std::vector<Att> atts()
    {
        std::vector<Att> out;
        const uint32_t tag = u32();
        const uint64_t n = count();
        if (tag == 0 && n == 0) return out;
        if (tag != TAG_ATTRIBUTE) throw Error("netCDF: bad attribute list tag");
        for (uint64_t i = 0; i < n; ++i) {
            Att a;
            a.name = name();
            a.type = (int)u32();
            a.n = count();
            const size_t es = type_size(a.type);
            std::vector<unsigned char> be(a.n * es);
            if (!be.empty() && fread(be.data(), 1, be.size(), f) != be.size()) throw Error("netCDF: truncated header");
            skip_pad(be.size());
            a.raw.resize(be.size());
            to_be(a.raw.data(), be.data(), a.n, es);   // the swap is its own inverse
            out.push_back(std::move(a));
        }
        return out;
    }
};

template <typename F>
void for_each_run(const Header &h, const Var &v, const uint64_t *start, const uint64_t *count, F fn)
{
    // fn(element offset in the variable, element offset in the caller's buffer, run length)
    const int nd = (int)v.dims.size();
    std::vector<uint64_t> len(nd);
    for (int k = 0; k < nd; ++k) {
        len[k] = h.dims[v.dims[k]].len;
        if (start[k] + count[k] > len[k]) throw Error("netCDF: hyperslab beyond the variable " + v.name);
        if (count[k] == 0) return;
    }
    if (nd == 0) { fn(0, 0, 1); return; }
    // the longest contiguous run: the innermost dimension, extended over the dimensions
    // outside it while the ones inside are whole
    int j = nd - 1;
    uint64_t run = count[j];
    while (j > 0 && count[j] == len[j]) { --j; run *= count[j]; }
    std::vector<uint64_t> idx(j, 0);
    uint64_t src = 0;
    for (;;) {
        uint64_t off = 0;
        for (int k = 0; k < nd; ++k) off = off * len[k] + (k < j ? start[k] + idx[k] : start[k]);
        fn(off, src, run);
        src += run;
        int k = j - 1;
        while (k >= 0 && ++idx[k] == count[k]) { idx[k] = 0; --k; }
        if (k < 0) break;
    }
}

}  // namespace

size_t type_size(int type)
{
    switch (type) {
    case NC_BYTE: case NC_CHAR: return 1;
    case NC_SHORT: return 2;
    case NC_INT: case NC_FLOAT: return 4;
    case NC_DOUBLE: return 8;
    default: throw Error("netCDF: unsupported type " + std::to_string(type));
    }
}

uint64_t Header::nelem(const Var &v) const
{
    uint64_t n = 1;
    for (int d : v.dims) n *= dims[d].len;
    return n;
}

int Header::find_var(const std::string &name) const
{
    for (size_t i = 0; i < vars.size(); ++i) if (vars[i].name == name) return (int)i;
    return -1;
}

int Header::find_dim(const std::string &name) const
{
    for (size_t i = 0; i < dims.size(); ++i) if (dims[i].name == name) return (int)i;
    return -1;
}

std::vector<unsigned char> encode_header(const Header &h)
{
    Enc e;
    e.wide = h.version == 5;
    e.b = {'C', 'D', 'F', (unsigned char)h.version};
    e.count(0);   // numrecs: no record variables
    if (h.dims.empty()) e.absent();
    else {
        e.u32(TAG_DIMENSION);
        e.count(h.dims.size());
        for (const Dim &d : h.dims) { e.name(d.name); e.count(d.len); }
    }
    e.atts(h.gatts);
    if (h.vars.empty()) e.absent();
    else {
        e.u32(TAG_VARIABLE);
        e.count(h.vars.size());
        for (const Var &v : h.vars) {
            e.name(v.name);
            e.count(v.dims.size());
            for (int d : v.dims) e.count((uint64_t)d);
            e.atts(v.atts);
            e.u32((uint32_t)v.type);
            const uint64_t vs = round4(h.vbytes(v));
            e.count(e.wide ? vs : std::min<uint64_t>(vs, 0xFFFFFFFFull));
            if (h.version == 1) e.u32((uint32_t)v.begin);
            else e.u64(v.begin);
        }
    }
    return e.b;
}

Header decode_header(FILE *f)
{
    Header h;
    unsigned char m[4];
    if (fread(m, 1, 4, f) != 4 || m[0] != 'C' || m[1] != 'D' || m[2] != 'F' || (m[3] != 1 && m[3] != 2 && m[3] != 5))
        throw Error("netCDF: not a classic-format file (bad magic)");
    h.version = m[3];
    Dec d{f};
    d.wide = h.version == 5;
    const uint64_t numrecs = d.count();
    (void)numrecs;
    uint32_t tag = d.u32();
    uint64_t n = d.count();
    if (tag == TAG_DIMENSION) {
        for (uint64_t i = 0; i < n; ++i) { Dim x; x.name = d.name(); x.len = d.count(); h.dims.push_back(x); }
    } else if (tag != 0 || n != 0) throw Error("netCDF: bad dimension list tag");
    h.gatts = d.atts();
    tag = d.u32();
    n = d.count();
    if (tag == TAG_VARIABLE) {
        for (uint64_t i = 0; i < n; ++i) {
            Var v;
            v.name = d.name();
            const uint64_t nd = d.count();
            for (uint64_t k = 0; k < nd; ++k) {
                const uint64_t id = d.count();
                if (id >= h.dims.size()) throw Error("netCDF: bad dimension id in " + v.name);
                v.dims.push_back((int)id);
            }
            v.atts = d.atts();
            v.type = (int)d.u32();
            type_size(v.type);
            (void)d.count();   // vsize: recomputed from the shape
            v.begin = h.version == 1 ? d.u32() : d.u64();
            h.vars.push_back(std::move(v));
        }
    } else if (tag != 0 || n != 0) throw Error("netCDF: bad variable list tag");
    for (const Var &v : h.vars)
        for (int dd : v.dims)
            if (h.dims[dd].len == 0) throw Error("netCDF: record variables are not supported (" + v.name + ")");
    return h;
}

// ------------------------------------------------------------------------- Writer

Writer::~Writer()
{
    if (f_) fclose(f_);
}

void Writer::create(const std::string &path)
{
    f_ = fopen(path.c_str(), "w+b");
    if (!f_) throw Error("netCDF: cannot create " + path);
    path_ = path;
}

void Writer::check_define(const char *what) const
{
    if (!f_) throw Error(std::string("netCDF: ") + what + " on a closed file");
    if (!define_) throw Error(std::string("netCDF: ") + what + " outside define mode");
}

int Writer::def_dim(const std::string &name, uint64_t len)
{
    check_define("def_dim");
    if (len == 0) throw Error("netCDF: dimension " + name + " of length 0 (no record dimension)");
    if (h_.find_dim(name) >= 0) throw Error("netCDF: dimension " + name + " defined twice");
    h_.dims.push_back({name, len});
    return (int)h_.dims.size() - 1;
}

int Writer::def_var(const std::string &name, int type, const std::vector<int> &dims)
{
    check_define("def_var");
    type_size(type);
    if (h_.find_var(name) >= 0) throw Error("netCDF: variable " + name + " defined twice");
    for (int d : dims)
        if (d < 0 || d >= (int)h_.dims.size()) throw Error("netCDF: bad dimension id for " + name);
    Var v;
    v.name = name;
    v.type = type;
    v.dims = dims;
    h_.vars.push_back(std::move(v));
    return (int)h_.vars.size() - 1;
}

void Writer::put_att(int varid, const std::string &name, int type, const void *vals, size_t n)
{
    check_define("put_att");
    std::vector<Att> &list = varid < 0 ? h_.gatts : h_.vars.at(varid).atts;
    Att a;
    a.name = name;
    a.type = type;
    a.n = n;
    a.raw.assign((const unsigned char *)vals, (const unsigned char *)vals + n * type_size(type));
    for (Att &x : list)
        if (x.name == name) { x = std::move(a); return; }   // a second put replaces the value
    list.push_back(std::move(a));
}

void Writer::put_att_text(int varid, const std::string &name, const std::string &s)
{
    put_att(varid, name, NC_CHAR, s.data(), s.size());
}

void Writer::write_at(uint64_t off, const void *p, size_t n)
{
    if (fseeko(f_, (off_t)off, SEEK_SET) != 0 || fwrite(p, 1, n, f_) != n)
        throw Error("netCDF: write failed on " + path_);
}

void Writer::fill_var(const Var &v)
{
    const size_t es = type_size(v.type);
    unsigned char one[8];
    switch (v.type) {
    case NC_DOUBLE: { double x = 9.9692099683868690e+36; to_be(one, &x, 1, 8); break; }
    case NC_FLOAT: { float x = 9.9692099683868690e+36f; to_be(one, &x, 1, 4); break; }
    case NC_INT: { int32_t x = -2147483647; to_be(one, &x, 1, 4); break; }
    case NC_SHORT: { int16_t x = -32767; to_be(one, &x, 1, 2); break; }
    case NC_BYTE: one[0] = (unsigned char)(signed char)-127; break;
    default: one[0] = 0;
    }
    const uint64_t total = h_.vbytes(v);
    const size_t chunk = std::min<uint64_t>(total, (uint64_t)8 << 20) / es * es;
    if (chunk == 0) return;
    std::vector<unsigned char> buf(chunk);
    for (size_t i = 0; i < chunk; i += es) memcpy(&buf[i], one, es);
    for (uint64_t off = 0; off < total; off += chunk)
        write_at(v.begin + off, buf.data(), (size_t)std::min<uint64_t>(chunk, total - off));
}

void Writer::enddef(bool fill, size_t room)
{
    check_define("enddef");
    bool big = false;
    for (const Var &v : h_.vars) big = big || h_.vbytes(v) > CDF2_VAR_MAX;
    h_.version = big || force5_ ? 5 : 2;
    for (size_t i = nlaid_; i < h_.vars.size(); ++i) h_.vars[i].begin = 1;   // placeholder for the size
    const size_t hsz = encode_header(h_).size();
    if (!laid_out_) {
        data_begin_ = (hsz + room + 511) & ~uint64_t(511);
        end_ = data_begin_;
        laid_out_ = true;
    } else if (hsz > data_begin_) {
        throw Error("netCDF: the header outgrew the room reserved in front of the data of " + path_);
    }
    for (size_t i = nlaid_; i < h_.vars.size(); ++i) {
        Var &v = h_.vars[i];
        v.begin = (end_ + 7) & ~uint64_t(7);
        end_ = v.begin + round4(h_.vbytes(v));
    }
    const std::vector<unsigned char> hb = encode_header(h_);
    write_at(0, hb.data(), hb.size());
    if (fill)
        for (size_t i = nlaid_; i < h_.vars.size(); ++i) fill_var(h_.vars[i]);
    if (end_ > 0) {   // the file reaches the end of the last variable even without fill
        unsigned char z = 0;
        fseeko(f_, 0, SEEK_END);
        if ((uint64_t)ftello(f_) < end_) write_at(end_ - 1, &z, 1);
    }
    nlaid_ = h_.vars.size();
    define_ = false;
}

void Writer::redef()
{
    if (!f_) throw Error("netCDF: redef on a closed file");
    define_ = true;
}

void Writer::put(int varid, const uint64_t *start, const uint64_t *count, const void *vals)
{
    if (!f_) throw Error("netCDF: put on a closed file");
    if (define_ || varid < 0 || (size_t)varid >= nlaid_) throw Error("netCDF: put before the variable is laid out");
    const Var &v = h_.vars[varid];
    const size_t es = type_size(v.type);
    const unsigned char *src = (const unsigned char *)vals;
    std::vector<unsigned char> buf;
    for_each_run(h_, v, start, count, [&](uint64_t off, uint64_t from, uint64_t n) {
        const uint64_t chunk = ((uint64_t)8 << 20) / es;
        for (uint64_t k = 0; k < n; k += chunk) {
            const uint64_t m = std::min(chunk, n - k);
            buf.resize(m * es);
            to_be(buf.data(), src + (from + k) * es, m, es);
            write_at(v.begin + (off + k) * es, buf.data(), buf.size());
        }
    });
}

void Writer::close()
{
    if (!f_) return;
    if (define_) enddef(true, 0);
    const int rc = fclose(f_);
    f_ = nullptr;
    if (rc != 0) throw Error("netCDF: close failed on " + path_);
}

// ------------------------------------------------------------------------- Reader

Reader::~Reader()
{
    if (f_) fclose(f_);
}

void Reader::open(const std::string &path)
{
    f_ = fopen(path.c_str(), "rb");
    if (!f_) throw Error("netCDF: cannot open " + path);
    h_ = decode_header(f_);
}

void Reader::get(int varid, const uint64_t *start, const uint64_t *count, double *out)
{
    if (!f_) throw Error("netCDF: get on a closed file");
    const Var &v = h_.vars.at(varid);
    const size_t es = type_size(v.type);
    if (v.type == NC_CHAR) throw Error("netCDF: " + v.name + " holds text");
    std::vector<unsigned char> be, ne;
    for_each_run(h_, v, start, count, [&](uint64_t off, uint64_t to, uint64_t n) {
        be.resize(n * es);
        ne.resize(n * es);
        if (fseeko(f_, (off_t)(v.begin + off * es), SEEK_SET) != 0 || fread(be.data(), 1, be.size(), f_) != be.size())
            throw Error("netCDF: short read of " + v.name);
        to_be(ne.data(), be.data(), n, es);
        for (uint64_t i = 0; i < n; ++i) {
            const unsigned char *p = &ne[i * es];
            double x;
            switch (v.type) {
            case NC_BYTE: x = (double)*(const signed char *)p; break;
            case NC_SHORT: { int16_t y; memcpy(&y, p, 2); x = y; break; }
            case NC_INT: { int32_t y; memcpy(&y, p, 4); x = y; break; }
            case NC_FLOAT: { float y; memcpy(&y, p, 4); x = y; break; }
            default: memcpy(&x, p, 8);
            }
            out[to + i] = x;
        }
    });
}

void Reader::close()
{
    if (f_) fclose(f_);
    f_ = nullptr;
}

}  // namespace nc
}  // namespace trt
