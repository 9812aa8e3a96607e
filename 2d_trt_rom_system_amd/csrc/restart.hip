// Checkpoint and resume of a context's state in the reference's restart file
// format.  RESTART_OUT (MainSrc/OUTPUTS.f08:1624-1674) writes 25 Fortran
// unformatted sequential records in a fixed order; RESTART_IN
// (MainSrc/INPUTS.f08:1734-1783) reads them back; RESTART_TIME_IN (:1785-1803)
// reads the first.  The reference calls them at AL:234-237 (resume before the
// time loop) and AL:889-894 (every restart_freq-th step).
//
// Record framing (the Fortran runtime's unformatted sequential layout, as
// documented for gfortran; flang writes the same for records below 2 GiB):
// a record is one or more subrecords, each a 4-byte little-endian length
// marker, the data, and the marker again.  A subrecord holds at most
// 2147483639 bytes; a longer record is split, the leading marker of every
// subrecord but the last is negated (another one follows) and so is the
// trailing marker of every subrecord but the first (one precedes).
// TRT_SUBRECORD_LENGTH lowers the split size, as GFORTRAN_SUBRECORD_LENGTH does.
//
// Host code only: the arrays move through pull/push of the context's device
// state; nothing here launches a kernel of its own.
#include "trt_internal.h"

#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace trt {
namespace {

constexpr size_t kMaxSub = 2147483639;   // data bytes of one subrecord with 4-byte markers

size_t subrecord_limit()
{
    const char *s = getenv("TRT_SUBRECORD_LENGTH");
    if (!s || !*s) return kMaxSub;
    char *end = nullptr;
    const long long v = strtoll(s, &end, 10);
    if (*end || v <= 0 || (unsigned long long)v > kMaxSub)
        throw Error(std::string("TRT_SUBRECORD_LENGTH must be an integer in 1..2147483639, got ") + s);
    return (size_t)v;
}

class File {
public:
    File(const char *path, const char *mode) : path_(path ? path : "")
    {
        if (path_.empty()) throw Error("empty file name");
        f_ = fopen(path_.c_str(), mode);
        if (!f_) throw Error("the file " + path_ + " could not be opened: " + strerror(errno));
    }
    ~File()
    {
        if (f_) fclose(f_);
    }
    File(const File &) = delete;
    File &operator=(const File &) = delete;
    void write(const void *p, size_t n)
    {
        if (n && fwrite(p, 1, n, f_) != n) throw Error("writing " + path_ + " failed: " + strerror(errno));
    }
    bool read(void *p, size_t n) { return n == 0 || fread(p, 1, n, f_) == n; }
    void close()
    {
        FILE *f = f_;
        f_ = nullptr;
        if (f && fclose(f) != 0) throw Error("closing " + path_ + " failed: " + strerror(errno));
    }
    // closes without reporting errors (the file is being discarded)
    void abandon()
    {
        if (f_) fclose(f_);
        f_ = nullptr;
    }
    const std::string &path() const { return path_; }

private:
    std::string path_;
    FILE *f_ = nullptr;
};

// A new file is written under path + ".tmp" and renamed over path only once it is
// complete, so a failed write (a full disk, a bad array) leaves an existing file
// at path as it was; appending writes to path itself.
class RecordWriter {
public:
    RecordWriter(const char *path, bool append)
        : final_(path ? path : ""), tmp_(append || final_.empty() ? final_ : final_ + ".tmp"),
          max_(subrecord_limit()), file_(tmp_.c_str(), append ? "ab" : "wb")
    {
    }
    ~RecordWriter()
    {
        if (!done_ && tmp_ != final_) {
            file_.abandon();
            remove(tmp_.c_str());
        }
    }
    RecordWriter(const RecordWriter &) = delete;
    RecordWriter &operator=(const RecordWriter &) = delete;
    // Starts a record of `bytes` data bytes; put() must then supply exactly that many.
    void begin(size_t bytes)
    {
        if (open_) throw Error("record started before the previous one was complete");
        left_ = bytes;
        first_ = true;
        open_sub();
    }
    void put(const void *p, size_t n)
    {
        const char *b = static_cast<const char *>(p);
        while (n) {
            if (!open_) throw Error("more data than the record holds");
            const size_t k = std::min(n, sub_left_);
            file_.write(b, k);
            b += k;
            n -= k;
            sub_left_ -= k;
            if (sub_left_ == 0) close_sub();
        }
    }
    void finish()
    {
        if (open_) throw Error("file closed inside a record");
        file_.close();
        if (tmp_ != final_ && rename(tmp_.c_str(), final_.c_str()) != 0)
            throw Error("renaming " + tmp_ + " to " + final_ + " failed: " + strerror(errno));
        done_ = true;
    }

private:
    void marker(int32_t m) { file_.write(&m, 4); }
    void open_sub()
    {
        sub_ = std::min(left_, max_);
        left_ -= sub_;
        marker(left_ > 0 ? -(int32_t)sub_ : (int32_t)sub_);
        sub_left_ = sub_;
        open_ = true;
        if (sub_ == 0) close_sub();
    }
    void close_sub()
    {
        marker(first_ ? (int32_t)sub_ : -(int32_t)sub_);
        open_ = false;
        first_ = false;
        if (left_ > 0) open_sub();
    }
    std::string final_, tmp_;
    size_t max_;       // checked before the file is created
    File file_;
    size_t left_ = 0, sub_ = 0, sub_left_ = 0;
    bool first_ = true, open_ = false, done_ = false;
};

class RecordReader {
public:
    explicit RecordReader(const char *path) : file_(path, "rb") {}
    // Starts record k (named for messages), which must hold exactly `bytes` data bytes.
    void begin(int k, const char *name, size_t bytes)
    {
        k_ = k;
        name_ = name;
        want_ = bytes;
        got_ = 0;
        first_ = true;
        open_sub();
    }
    void get(void *p, size_t n)
    {
        char *b = static_cast<char *>(p);
        while (n) {
            if (sub_left_ == 0) fail("reading past the end of the record");
            const size_t k = std::min(n, sub_left_);
            if (!file_.read(b, k)) fail("the file ends inside the record");
            b += k;
            n -= k;
            sub_left_ -= k;
            if (sub_left_ == 0) close_sub();
        }
    }

private:
    [[noreturn]] void fail(const std::string &what) const
    {
        throw Error(file_.path() + ": record " + std::to_string(k_ + 1) + " (" + name_ + "): " + what);
    }
    void mismatch(size_t holds, bool at_least) const
    {
        fail("holds " + std::string(at_least ? "at least " : "") + std::to_string(holds) + " bytes, expected " +
             std::to_string(want_) + " (a file of another grid, quadrature or group count?)");
    }
    void open_sub()
    {
        int32_t m;
        if (!file_.read(&m, 4)) fail(got_ ? "the file ends inside the record" : "the file ends before the record");
        more_ = m < 0;
        sub_ = (size_t)(m < 0 ? -(int64_t)m : (int64_t)m);
        if (got_ + sub_ > want_) mismatch(got_ + sub_, more_);
        sub_left_ = sub_;
        if (sub_ == 0) close_sub();
    }
    void close_sub()
    {
        int32_t t;
        if (!file_.read(&t, 4)) fail("the file ends inside the record");
        const size_t tl = (size_t)(t < 0 ? -(int64_t)t : (int64_t)t);
        if (tl != sub_ || (sub_ && (t < 0) == first_)) fail("length markers disagree (not a Fortran unformatted sequential file?)");
        got_ += sub_;
        first_ = false;
        if (more_) open_sub();
        else if (got_ != want_) mismatch(got_, false);
    }
    File file_;
    int k_ = 0;
    const char *name_ = "";
    size_t want_ = 0, got_ = 0, sub_ = 0, sub_left_ = 0;
    bool first_ = true, more_ = false;
};

// RESTART_OUT's record order (OUTPUTS.f08:1648-1672)
const char *const kRecords[] = {"Time",     "Temp",    "I_crn",    "fg_avg_xx", "fg_avg_yy", "fg_edgV_xx", "fg_edgV_xy",
                                "fg_edgH_yy", "fg_edgH_xy", "RT_Src", "MGQD_Src", "GQD_Src",   "E_avg",     "Eg_avg",
                                "Eg_edgV",  "Eg_edgH", "Fxg_edgV", "Fyg_edgH",  "KapE_Bar",  "KapE",       "KapR",
                                "Cg_L",     "Cg_B",    "Cg_R",     "Cg_T"};
constexpr int kNRec = sizeof(kRecords) / sizeof(kRecords[0]);

void to_host(Ctx &c, const char *name, std::vector<double> &h)
{
    const DevArray &d = c.arrays.at(name);
    h.resize(d.size());
    TRT_CHECK(hipMemcpyAsync(h.data(), d.p, h.size() * 8, hipMemcpyDeviceToHost, c.stream));
    TRT_CHECK(hipStreamSynchronize(c.stream));
}

}  // namespace

void restart_out(Ctx &c, const char *path)
{
    const int nx = c.nx, ny = c.ny, nm = c.nm, ng = c.ng;
    const size_t nI = (size_t)4 * nx * ny * nm * ng;
    TRT_CHECK(hipStreamSynchronize(c.stream));
    RecordWriter w(path, false);
    std::vector<double> h;
    for (int k = 0; k < kNRec; ++k) {
        const char *name = kRecords[k];
        if (k == 0) {
            w.begin(8);
            w.put(&c.time, 8);
        } else if (!strcmp(name, "I_crn")) {
            h.assign(nI, 0.0);
            pull_icrn(c, c.icur, h.data());
            w.begin(nI * 8);
            w.put(h.data(), nI * 8);
            std::vector<double>().swap(h);
        } else if (!strcmp(name, "RT_Src")) {
            // RT_Src(2N_x, 2N_y, N_m, N_g): the (cell, group) value at all four
            // corners and every direction (UPDATES.f08:141-150)
            to_host(c, name, h);
            std::vector<double> plane((size_t)4 * nx * ny);
            w.begin(nI * 8);
            for (int g = 0; g < ng; ++g) {
                const double *src = h.data() + (size_t)nx * ny * g;
                for (int jj = 0; jj < 2 * ny; ++jj)
                    for (int ii = 0; ii < 2 * nx; ++ii) plane[ii + (size_t)2 * nx * jj] = src[ii / 2 + (size_t)nx * (jj / 2)];
                for (int m = 0; m < nm; ++m) w.put(plane.data(), plane.size() * 8);
            }
        } else {
            to_host(c, name, h);
            w.begin(h.size() * 8);
            w.put(h.data(), h.size() * 8);
        }
    }
    w.finish();
}

void restart_in(Ctx &c, const char *path)
{
    const int nx = c.nx, ny = c.ny, nm = c.nm, ng = c.ng;
    const size_t nI = (size_t)4 * nx * ny * nm * ng;
    // read and check the whole file before any device array changes
    RecordReader r(path);
    double time = 0.0;
    std::vector<std::vector<double>> vals(kNRec);
    for (int k = 0; k < kNRec; ++k) {
        const char *name = kRecords[k];
        std::vector<double> &v = vals[k];
        if (k == 0) {
            r.begin(k, name, 8);
            r.get(&time, 8);
        } else if (!strcmp(name, "I_crn")) {
            v.resize(nI);
            r.begin(k, name, nI * 8);
            r.get(v.data(), nI * 8);
        } else if (!strcmp(name, "RT_Src")) {
            // kept per (cell, group): every corner and direction must hold the same value
            v.resize((size_t)nx * ny * ng);
            std::vector<double> plane((size_t)4 * nx * ny), first(plane.size());
            r.begin(k, name, nI * 8);
            for (int g = 0; g < ng; ++g) {
                for (int m = 0; m < nm; ++m) {
                    r.get(m == 0 ? first.data() : plane.data(), plane.size() * 8);
                    if (m > 0 && memcmp(plane.data(), first.data(), plane.size() * 8) != 0)
                        throw Error(std::string(path) + ": RT_Src of group " + std::to_string(g + 1) + " differs between directions 1 and " +
                                    std::to_string(m + 1) + "; this build keeps RT_Src per (cell, group) as MATERIAL_UPDATE writes it");
                }
                for (int j = 0; j < ny; ++j)
                    for (int i = 0; i < nx; ++i) {
                        const size_t b = (size_t)2 * i + (size_t)2 * nx * 2 * j, row = (size_t)2 * nx;
                        const double x = first[b];
                        if (first[b + 1] != x || first[b + row] != x || first[b + row + 1] != x)
                            throw Error(std::string(path) + ": RT_Src of cell (" + std::to_string(i + 1) + "," +
                                        std::to_string(j + 1) + "), group " + std::to_string(g + 1) +
                                        " differs between corners; this build keeps RT_Src per (cell, group)");
                        v[i + (size_t)nx * j + (size_t)nx * ny * g] = x;
                    }
            }
        } else {
            v.resize(c.arrays.at(name).size());
            r.begin(k, name, v.size() * 8);
            r.get(v.data(), v.size() * 8);
        }
    }
    // the file is whole and matches the context: install it
    TRT_CHECK(hipStreamSynchronize(c.stream));
    for (int k = 1; k < kNRec; ++k) {
        const char *name = kRecords[k];
        if (!strcmp(name, "I_crn")) push_icrn(c, c.icur, vals[k].data());
        else TRT_CHECK(hipMemcpyAsync(c.arr(name), vals[k].data(), vals[k].size() * 8, hipMemcpyHostToDevice, c.stream));
    }
    rebuild_sc(c);   // the sweep's per-(cell, group) (KapE + 1/(c dt), RT_Src) pairs
    // the reference restarts on a freshly initialised state: the inflow of reflective
    // sides is RT_INIT's Planckian (RESTART_IN reads no boundary intensities, AL:234-236)
    reflect_reset(c);
    TRT_CHECK(hipStreamSynchronize(c.stream));
    c.time = time;
}

}  // namespace trt

// ---------------------------------------------------------------------------
// Host-only record I/O (no device, no context)
// ---------------------------------------------------------------------------
namespace {
thread_local std::string g_io_error;

template <typename F>
int io_guarded(F &&f)
{
    try {
        f();
        g_io_error.clear();
        return 0;
    } catch (const trt::Error &e) {
        g_io_error = e.msg;
    } catch (const std::exception &e) {
        g_io_error = e.what();
    }
    return -2;
}
}  // namespace

extern "C" {

int trt_frecords_write(const char *path, int nrec, const double *const *data, const size_t *count, int append)
{
    return io_guarded([&] {
        if (nrec < 0 || (nrec > 0 && (!data || !count))) throw trt::Error("null or negative record list");
        trt::RecordWriter w(path, append != 0);
        for (int k = 0; k < nrec; ++k) {
            if (count[k] && !data[k]) throw trt::Error("record " + std::to_string(k + 1) + " has no data pointer");
            w.begin(count[k] * 8);
            w.put(data[k], count[k] * 8);
        }
        w.finish();
    });
}

int trt_frecords_read(const char *path, int nrec, double *const *data, const size_t *count)
{
    return io_guarded([&] {
        if (nrec < 0 || (nrec > 0 && (!data || !count))) throw trt::Error("null or negative record list");
        trt::RecordReader r(path);
        for (int k = 0; k < nrec; ++k) {
            if (count[k] && !data[k]) throw trt::Error("record " + std::to_string(k + 1) + " has no data pointer");
            r.begin(k, "array", count[k] * 8);
            r.get(data[k], count[k] * 8);
        }
    });
}

int trt_restart_time_in(const char *path, double *time)
{
    return io_guarded([&] {
        if (!time) throw trt::Error("null argument");
        trt::RecordReader r(path);
        r.begin(0, "Time", 8);
        r.get(time, 8);
    });
}

const char *trt_frecords_last_error(void) { return g_io_error.c_str(); }

}  // extern "C"
