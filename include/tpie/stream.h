// tpie::ami::stream<T> subset: a file-backed stream of fixed-size items with
// TPIE's read_item / write_item / seek / tell / truncate / persist calls, as
// used by nni.cpp:741-788,858 and terranni.cpp:143-151,231-241,404,413-417.
//
// Named streams are persistent (they outlive the object, like TerraNNI's
// work-package streams r<row>c<col>.tpie); unnamed streams live in a
// temporary file that is removed on destruction unless persist() is called.
#ifndef TERRANNI_TPIE_STREAM_H
#define TERRANNI_TPIE_STREAM_H

#include <sys/types.h>
#include <unistd.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "tpie.h"

namespace tpie {

enum persistence { PERSIST_DELETE = 0, PERSIST_READ_ONCE = 1, PERSIST_PERSISTENT = 2 };

namespace ami {

enum err { NO_ERROR = 0, IO_ERROR, END_OF_STREAM, READ_ONLY, OS_ERROR, BASE_METHOD, MEMORY_ERROR, PERMISSION_DENIED, INSUFFICIENT_MAIN_MEMORY, INSUFFICIENT_AVAILABLE_STREAMS, ENV_UNDEFINED, BIT_MATRIX_BOUNDS, NOT_POWER_OF_2, NULL_POINTER, GENERIC_ERROR };

enum stream_type { READ_STREAM = 1, WRITE_STREAM, APPEND_STREAM, READ_WRITE_STREAM };

template <class T>
class stream {
    static_assert(std::is_trivially_copyable<T>::value, "stream items must be trivially copyable");

public:
    // A temporary stream.
    stream() : persist_(PERSIST_DELETE)
    {
        char tmpl[4096];
        std::snprintf(tmpl, sizeof tmpl, "%s/tnni_tpie_XXXXXX", tempname::default_path().c_str());
        const int fd = mkstemp(tmpl);
        if (fd >= 0) {
            path_ = tmpl;
            f_ = fdopen(fd, "w+b");
        }
        status_ok_ = f_ != nullptr;
    }

    explicit stream(const std::string &path, stream_type st = READ_WRITE_STREAM)
        : path_(path), persist_(PERSIST_PERSISTENT)
    {
        if (st == WRITE_STREAM) {
            f_ = std::fopen(path.c_str(), "w+b");
        } else {
            f_ = std::fopen(path.c_str(), "r+b");
            if (!f_ && st != READ_STREAM) f_ = std::fopen(path.c_str(), "w+b");
        }
        status_ok_ = f_ != nullptr;
        if (f_) {
            std::fseek(f_, 0, SEEK_END);
            len_ = (TPIE_OS_OFFSET)(ftello(f_) / (off_t)sizeof(T));
            if (st == APPEND_STREAM) pos_ = len_;
        }
    }

    stream(const stream &) = delete;
    stream &operator=(const stream &) = delete;

    ~stream()
    {
        if (f_) {
            flush();
            std::fclose(f_);
        }
        if (persist_ == PERSIST_DELETE && !path_.empty()) std::remove(path_.c_str());
    }

    bool is_valid() const { return status_ok_; }
    const std::string &name() const { return path_; }

    // The item at the current position; *out points into the stream's own
    // buffer and stays valid until the next call on the stream.
    err read_item(T **out)
    {
        if (!f_) return IO_ERROR;
        if (pos_ >= length()) return END_OF_STREAM;
        if (pos_ < rbeg_ || pos_ >= rbeg_ + (TPIE_OS_OFFSET)rbuf_.size()) {
            flush();
            const size_t want = (size_t)std::min<TPIE_OS_OFFSET>(kBlock, length() - pos_);
            rbuf_.resize(want);
            if (fseeko(f_, (off_t)pos_ * (off_t)sizeof(T), SEEK_SET) != 0) return IO_ERROR;
            const size_t got = std::fread(rbuf_.data(), sizeof(T), want, f_);
            rbuf_.resize(got);
            rbeg_ = pos_;
            if (got == 0) return IO_ERROR;
        }
        *out = &rbuf_[(size_t)(pos_ - rbeg_)];
        ++pos_;
        return NO_ERROR;
    }

    // Up to *len items from the current position into mm_space; *len is set
    // to the number read (TPIE's read_array).
    err read_array(T *mm_space, TPIE_OS_SIZE_T *len)
    {
        if (!f_ || !mm_space || !len) return IO_ERROR;
        flush();
        const TPIE_OS_OFFSET want = std::min<TPIE_OS_OFFSET>((TPIE_OS_OFFSET)*len, length() - pos_);
        *len = 0;
        if (want <= 0) return END_OF_STREAM;
        if (fseeko(f_, (off_t)pos_ * (off_t)sizeof(T), SEEK_SET) != 0) return IO_ERROR;
        const size_t got = std::fread(mm_space, sizeof(T), (size_t)want, f_);
        pos_ += (TPIE_OS_OFFSET)got;
        *len = (TPIE_OS_SIZE_T)got;
        rbuf_.clear();
        return got == (size_t)want ? NO_ERROR : IO_ERROR;
    }

    err write_item(const T &item)
    {
        if (!f_) return IO_ERROR;
        if (!rbuf_.empty() && pos_ >= rbeg_ && pos_ < rbeg_ + (TPIE_OS_OFFSET)rbuf_.size())
            rbuf_.clear();  // overwriting cached data
        if (wbuf_.empty() || wbeg_ + (TPIE_OS_OFFSET)wbuf_.size() != pos_) {
            if (flush() != NO_ERROR) return IO_ERROR;
            wbeg_ = pos_;
        }
        wbuf_.push_back(item);
        ++pos_;
        if (wbuf_.size() >= (size_t)kBlock) return flush();
        return NO_ERROR;
    }

    // len items at the current position (TPIE's write_array).
    err write_array(const T *mm_space, TPIE_OS_SIZE_T len)
    {
        if (!f_ || (!mm_space && len)) return IO_ERROR;
        if (len == 0) return NO_ERROR;
        if (flush() != NO_ERROR) return IO_ERROR;
        rbuf_.clear();
        if (fseeko(f_, (off_t)pos_ * (off_t)sizeof(T), SEEK_SET) != 0) return IO_ERROR;
        const size_t put = std::fwrite(mm_space, sizeof(T), (size_t)len, f_);
        pos_ += (TPIE_OS_OFFSET)put;
        if (pos_ > len_) len_ = pos_;
        return put == (size_t)len ? NO_ERROR : IO_ERROR;
    }

    err seek(TPIE_OS_OFFSET off)
    {
        if (off < 0 || off > length()) return OS_ERROR;
        pos_ = off;
        return NO_ERROR;
    }
    TPIE_OS_OFFSET tell() const { return pos_; }
    TPIE_OS_OFFSET stream_len() const { return length(); }

    err truncate(TPIE_OS_OFFSET n)
    {
        if (!f_ || n < 0) return OS_ERROR;
        flush();
        rbuf_.clear();
        if (ftruncate(fileno(f_), (off_t)n * (off_t)sizeof(T)) != 0) return OS_ERROR;
        len_ = n;
        if (pos_ > n) pos_ = n;
        return NO_ERROR;
    }

    void persist(persistence p) { persist_ = p; }
    void persist() { persist_ = PERSIST_PERSISTENT; }
    persistence persist_state() const { return persist_; }

    err main_memory_usage(TPIE_OS_SIZE_T *usage, mem::stream_usage) const
    {
        *usage = (TPIE_OS_SIZE_T)(2 * kBlock * sizeof(T) + sizeof(*this));
        return NO_ERROR;
    }
    static TPIE_OS_SIZE_T available_streams() { return 512; }

    err flush()
    {
        if (wbuf_.empty() || !f_) return NO_ERROR;
        if (fseeko(f_, (off_t)wbeg_ * (off_t)sizeof(T), SEEK_SET) != 0) return IO_ERROR;
        const size_t put = std::fwrite(wbuf_.data(), sizeof(T), wbuf_.size(), f_);
        const TPIE_OS_OFFSET end = wbeg_ + (TPIE_OS_OFFSET)wbuf_.size();
        if (end > len_) len_ = end;
        wbuf_.clear();
        return put == 0 ? IO_ERROR : NO_ERROR;
    }

private:
    static constexpr TPIE_OS_OFFSET kBlock = 16384;

    TPIE_OS_OFFSET length() const
    {
        const TPIE_OS_OFFSET w = wbuf_.empty() ? 0 : wbeg_ + (TPIE_OS_OFFSET)wbuf_.size();
        return std::max(len_, w);
    }

    std::FILE *f_ = nullptr;
    std::string path_;
    persistence persist_;
    bool status_ok_ = false;
    TPIE_OS_OFFSET pos_ = 0, len_ = 0;
    std::vector<T> rbuf_, wbuf_;
    TPIE_OS_OFFSET rbeg_ = 0, wbeg_ = 0;
};

}  // namespace ami
}  // namespace tpie

#endif
