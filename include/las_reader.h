/*
 * las_reader.h -- LAS 1.0-1.4 point reader without libLAS.
 *
 * Replaces the reference's nni_io::las_reader (las_reader.h:26-58,
 * las_reader.inl:24-78, point_reader.h:26-45) and the parts of libLAS 1.6 it
 * uses: the public header block (bounding box, scale, offset) and, per
 * point, X/Y/Z and the intensity that TerraNNI reads as the point's time
 * (las_reader.inl:59).  Every point record format 0-10 starts with int32 X,
 * Y, Z and uint16 intensity, so records are read with the header's record
 * length as the stride whatever the format.  Compressed files (LAZ) are
 * rejected.
 *
 * Two ways in:
 *   - las_file::read_raw: blocks of raw int32 X / Y / Z and the intensity,
 *     for the engine, which de-scales X * scale + offset in FP64 on the GPU;
 *   - las_reader<P, Z>::next_point: one point at a time as
 *     P(x, y, z, intensity) with x = X * scale + offset in FP64, the
 *     reference's interface (point_reader.inl:38-48).
 */
#ifndef TERRANNI_LAS_READER_H
#define TERRANNI_LAS_READER_H

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace nni_io {

// The public header block fields TerraNNI needs, with libLAS-style getters
// (terranni.cpp:77-83 reads GetMin/GetMax through liblas::Header).
struct las_header {
    uint8_t version_major = 0, version_minor = 0;
    uint16_t header_size = 0;
    uint32_t offset_to_points = 0;
    uint8_t point_format = 0;
    uint16_t record_length = 0;
    uint64_t n_points = 0;
    double scale[3] = {1, 1, 1};
    double offset[3] = {0, 0, 0};
    double max[3] = {0, 0, 0};
    double min[3] = {0, 0, 0};

    double GetMinX() const { return min[0]; }
    double GetMinY() const { return min[1]; }
    double GetMinZ() const { return min[2]; }
    double GetMaxX() const { return max[0]; }
    double GetMaxY() const { return max[1]; }
    double GetMaxZ() const { return max[2]; }
    double GetScaleX() const { return scale[0]; }
    double GetScaleY() const { return scale[1]; }
    double GetScaleZ() const { return scale[2]; }
    double GetOffsetX() const { return offset[0]; }
    double GetOffsetY() const { return offset[1]; }
    double GetOffsetZ() const { return offset[2]; }
    uint64_t GetPointRecordsCount() const { return n_points; }
    uint8_t GetDataFormatId() const { return point_format; }
    uint16_t GetDataRecordLength() const { return record_length; }
};

namespace detail {
template <class T>
inline T get_le(const unsigned char *p)
{
    T v;
    std::memcpy(&v, p, sizeof v);  // LAS is little-endian, as is the host
    return v;
}
}  // namespace detail

// One LAS file: header parse and block reads of raw point fields.
class las_file {
public:
    las_file() = default;
    las_file(const las_file &) = delete;
    las_file &operator=(const las_file &) = delete;
    ~las_file() { close(); }

    // Open and parse the header; throws std::runtime_error on a file that is
    // not LAS 1.x or is compressed.
    void open(const std::string &path)
    {
        close();
        f_ = std::fopen(path.c_str(), "rb");
        if (!f_) throw std::runtime_error("could not open " + path);
        unsigned char h[375];
        std::memset(h, 0, sizeof h);
        const size_t got = std::fread(h, 1, sizeof h, f_);
        if (got < 227 || std::memcmp(h, "LASF", 4) != 0) {
            close();
            throw std::runtime_error(path + ": not a LAS file");
        }
        using detail::get_le;
        hd_ = las_header();
        hd_.version_major = h[24];
        hd_.version_minor = h[25];
        hd_.header_size = get_le<uint16_t>(h + 94);
        hd_.offset_to_points = get_le<uint32_t>(h + 96);
        const uint8_t fmt = h[104];
        if (fmt & 0xC0) {
            close();
            throw std::runtime_error(path + ": compressed (LAZ) point data is not supported");
        }
        hd_.point_format = fmt;
        hd_.record_length = get_le<uint16_t>(h + 105);
        hd_.n_points = get_le<uint32_t>(h + 107);
        for (int k = 0; k < 3; ++k) {
            hd_.scale[k] = get_le<double>(h + 131 + 8 * k);
            hd_.offset[k] = get_le<double>(h + 155 + 8 * k);
            hd_.max[k] = get_le<double>(h + 179 + 16 * k);
            hd_.min[k] = get_le<double>(h + 187 + 16 * k);
        }
        // LAS 1.4: the 64-bit point count (the legacy one may be 0)
        if (hd_.version_major == 1 && hd_.version_minor >= 4 && hd_.header_size >= 375 && got >= 255) {
            const uint64_t n64 = get_le<uint64_t>(h + 247);
            if (n64) hd_.n_points = n64;
        }
        if (hd_.version_major != 1 || fmt > 10 || hd_.record_length < 14 ||
            hd_.offset_to_points < hd_.header_size) {
            close();
            throw std::runtime_error(path + ": unsupported LAS version, point format or record length");
        }
        path_ = path;
        rewind();
    }

    void close()
    {
        if (f_) std::fclose(f_);
        f_ = nullptr;
        next_ = 0;
    }

    bool is_open() const { return f_ != nullptr; }
    const las_header &header() const { return hd_; }
    const std::string &path() const { return path_; }

    // Back to the first point record.
    void rewind()
    {
        if (!f_) return;
        if (std::fseek(f_, (long)hd_.offset_to_points, SEEK_SET) != 0)
            throw std::runtime_error(path_ + ": seek failed");
        next_ = 0;
    }

    // Up to max_n next points' raw fields (any pointer may be null).  Returns
    // how many were read; 0 at the end of the points.
    size_t read_raw(size_t max_n, int32_t *X, int32_t *Y, int32_t *Z, uint16_t *intensity)
    {
        if (!f_) return 0;
        const uint64_t left = hd_.n_points - next_;
        size_t want = (size_t)(left < (uint64_t)max_n ? left : (uint64_t)max_n);
        const size_t rl = hd_.record_length;
        size_t done = 0;
        while (done < want) {
            const size_t k = std::min(want - done, kBlockRecords);
            buf_.resize(k * rl);
            const size_t got = std::fread(buf_.data(), rl, k, f_);
            for (size_t i = 0; i < got; ++i) {
                const unsigned char *r = buf_.data() + i * rl;
                if (X) X[done + i] = detail::get_le<int32_t>(r);
                if (Y) Y[done + i] = detail::get_le<int32_t>(r + 4);
                if (Z) Z[done + i] = detail::get_le<int32_t>(r + 8);
                if (intensity) intensity[done + i] = detail::get_le<uint16_t>(r + 12);
            }
            done += got;
            if (got < k) break;  // a truncated file ends early
        }
        next_ += done;
        return done;
    }

private:
    static constexpr size_t kBlockRecords = 1 << 16;
    std::FILE *f_ = nullptr;
    std::string path_;
    las_header hd_;
    uint64_t next_ = 0;
    std::vector<unsigned char> buf_;
};

// The reference's point reader interface (point_reader.h:28-41): open,
// next_point, reset, close.  P is built as P(x, y, z, intensity), x, y, z
// de-scaled in FP64 (las_reader.inl:59).  Z is kept for source
// compatibility with the reference's template.
template <typename P, typename Z = double>
class las_reader {
public:
    virtual ~las_reader() { close(); }

    void open(const std::string &filename)
    {
        close();
        file_.open(filename);
        pos_ = cnt_ = 0;
        is_open_ = true;
    }
    void close()
    {
        if (is_open_) file_.close();
        is_open_ = false;
    }
    void reset()
    {
        file_.rewind();
        pos_ = cnt_ = 0;
    }
    const las_header &get_header() const { return file_.header(); }

    bool next_point(P &point)
    {
        if (!is_open_) return false;
        if (pos_ == cnt_) {
            X_.resize(kBatch);
            Y_.resize(kBatch);
            Z_.resize(kBatch);
            I_.resize(kBatch);
            cnt_ = file_.read_raw(kBatch, X_.data(), Y_.data(), Z_.data(), I_.data());
            pos_ = 0;
            if (cnt_ == 0) return false;
        }
        const las_header &h = file_.header();
        const double x = (double)X_[pos_] * h.scale[0] + h.offset[0];
        const double y = (double)Y_[pos_] * h.scale[1] + h.offset[1];
        const double z = (double)Z_[pos_] * h.scale[2] + h.offset[2];
        point = P(x, y, z, (int)I_[pos_]);
        ++pos_;
        return true;
    }

private:
    static constexpr size_t kBatch = 4096;
    las_file file_;
    bool is_open_ = false;
    size_t pos_ = 0, cnt_ = 0;
    std::vector<int32_t> X_, Y_, Z_;
    std::vector<uint16_t> I_;
};

}  // namespace nni_io

#endif
