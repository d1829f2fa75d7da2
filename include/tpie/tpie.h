// Minimal, source-compatible subset of TPIE (the external-memory library
// that TerraNNI's nni.h and terranni.cpp build on): just what the tile API
// and the streaming driver use.  tpie_init / MM_manager / tempname,
// terranni.cpp:499-502,599.
#ifndef TERRANNI_TPIE_TPIE_H
#define TERRANNI_TPIE_TPIE_H

#include <cstddef>
#include <cstdint>
#include <string>

typedef int64_t TPIE_OS_OFFSET;
typedef size_t TPIE_OS_SIZE_T;

namespace tpie {

inline void tpie_init() {}
inline void tpie_finish() {}

namespace mem {
enum stream_usage { STREAM_USAGE_OVERHEAD, STREAM_USAGE_BUFFER, STREAM_USAGE_CURRENT, STREAM_USAGE_MAXIMUM, STREAM_USAGE_SUBSTREAM };
}

// Memory accounting of the external-memory algorithms (terranni.cpp:146,501-502).
class memory_manager {
public:
    void set_memory_limit(size_t bytes) { limit_ = bytes; }
    size_t memory_limit() const { return limit_; }
    size_t consecutive_memory_available() const { return limit_; }
    void warn_memory_limit() {}
    void enforce_memory_limit() {}

private:
    size_t limit_ = (size_t)2000 * 1024 * 1024;
};

inline memory_manager &mm_instance()
{
    static memory_manager m;
    return m;
}
#define MM_manager mm_instance()

namespace tempname {
inline std::string &default_path()
{
    static std::string p = ".";
    return p;
}
inline void set_default_path(const std::string &p) { default_path() = p; }
}  // namespace tempname

}  // namespace tpie

#endif
