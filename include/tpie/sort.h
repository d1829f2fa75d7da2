// tpie::ami::sort subset: sort a stream in place with a comparison object
// whose compare(a, b) returns <0 / 0 / >0 (terranni.cpp:119-138,305-320,450).
// Runs that fit the memory limit are sorted in memory; larger streams are
// sorted as runs written to temporary streams and k-way merged.
#ifndef TERRANNI_TPIE_SORT_H
#define TERRANNI_TPIE_SORT_H

#include <algorithm>
#include <memory>
#include <queue>
#include <vector>

#include "stream.h"

namespace tpie {
namespace ami {

template <class T, class CMP>
err sort(stream<T> *s, CMP *cmp)
{
    if (!s || !cmp) return NULL_POINTER;
    const TPIE_OS_OFFSET n = s->stream_len();
    const TPIE_OS_OFFSET run = std::max<TPIE_OS_OFFSET>(
        1024, (TPIE_OS_OFFSET)(MM_manager.consecutive_memory_available() / (2 * sizeof(T))));
    auto less = [cmp](const T &a, const T &b) { return cmp->compare(a, b) < 0; };
    std::vector<T> buf;
    if (s->seek(0) != NO_ERROR) return IO_ERROR;
    if (n <= run) {
        buf.reserve((size_t)n);
        T *p;
        while (s->read_item(&p) == NO_ERROR) buf.push_back(*p);
        std::stable_sort(buf.begin(), buf.end(), less);
        s->seek(0);
        for (const T &x : buf) s->write_item(x);
        s->flush();
        s->seek(0);
        return NO_ERROR;
    }
    // sorted runs in temporary streams
    std::vector<std::unique_ptr<stream<T>>> runs;
    T *p;
    while (true) {
        buf.clear();
        while ((TPIE_OS_OFFSET)buf.size() < run && s->read_item(&p) == NO_ERROR) buf.push_back(*p);
        if (buf.empty()) break;
        std::stable_sort(buf.begin(), buf.end(), less);
        runs.emplace_back(new stream<T>());
        for (const T &x : buf) runs.back()->write_item(x);
        runs.back()->flush();
        runs.back()->seek(0);
        if ((TPIE_OS_OFFSET)buf.size() < run) break;
    }
    buf.clear();
    buf.shrink_to_fit();
    // k-way merge back into s; ties go to the earlier run (stable)
    struct Head {
        T item;
        size_t run;
    };
    auto greater = [cmp](const Head &a, const Head &b) {
        const int c = cmp->compare(a.item, b.item);
        return c > 0 || (c == 0 && a.run > b.run);
    };
    std::priority_queue<Head, std::vector<Head>, decltype(greater)> heap(greater);
    for (size_t r = 0; r < runs.size(); ++r)
        if (runs[r]->read_item(&p) == NO_ERROR) heap.push(Head{*p, r});
    s->seek(0);
    while (!heap.empty()) {
        Head h = heap.top();
        heap.pop();
        s->write_item(h.item);
        if (runs[h.run]->read_item(&p) == NO_ERROR) heap.push(Head{*p, h.run});
    }
    s->flush();
    s->seek(0);
    return NO_ERROR;
}

}  // namespace ami
}  // namespace tpie

#endif
