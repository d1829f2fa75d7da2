// Checks of the TPIE subset (stream + sort) used by the nni.h API and the
// driver; run by tests/test_host.py.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "tpie/sort.h"
#include "tpie/stream.h"

#define CHECK(c)                                                        \
    do {                                                                \
        if (!(c)) {                                                     \
            std::fprintf(stderr, "FAIL %s:%d %s\n", __FILE__, __LINE__, #c); \
            return 1;                                                   \
        }                                                               \
    } while (0)

struct Item {
    int key, val;
};
struct ByKey {
    int compare(const Item &a, const Item &b) { return a.key - b.key; }
};

int main(int argc, char **argv)
{
    const std::string dir = argc > 1 ? argv[1] : ".";
    tpie::tempname::set_default_path(dir);
    // write, seek, read, tell
    {
        tpie::ami::stream<Item> s(dir + "/a.tpie");
        s.truncate(0);
        for (int i = 0; i < 50000; ++i) CHECK(s.write_item(Item{i, 2 * i}) == tpie::ami::NO_ERROR);
        CHECK(s.stream_len() == 50000);
        CHECK(s.seek(12345) == tpie::ami::NO_ERROR);
        Item *p;
        CHECK(s.read_item(&p) == tpie::ami::NO_ERROR && p->key == 12345 && p->val == 24690);
        CHECK(s.tell() == 12346);
        s.seek(49999);
        CHECK(s.read_item(&p) == tpie::ami::NO_ERROR && p->key == 49999);
        CHECK(s.read_item(&p) == tpie::ami::END_OF_STREAM);
        // overwrite in the middle, then read it back
        s.seek(100);
        s.write_item(Item{-1, -1});
        s.seek(99);
        CHECK(s.read_item(&p) == tpie::ami::NO_ERROR && p->key == 99);
        CHECK(s.read_item(&p) == tpie::ami::NO_ERROR && p->key == -1);
        CHECK(s.read_item(&p) == tpie::ami::NO_ERROR && p->key == 101);
        // bulk read from a position, clipped at the end
        std::vector<Item> buf(1000);
        s.seek(49500);
        TPIE_OS_SIZE_T len = buf.size();
        CHECK(s.read_array(buf.data(), &len) == tpie::ami::NO_ERROR && len == 500);
        CHECK(buf[0].key == 49500 && buf[499].key == 49999 && s.tell() == 50000);
        s.seek(98);
        len = 4;
        CHECK(s.read_array(buf.data(), &len) == tpie::ami::NO_ERROR && len == 4);
        CHECK(buf[0].key == 98 && buf[2].key == -1 && buf[3].key == 101);
        CHECK(s.read_item(&p) == tpie::ami::NO_ERROR && p->key == 102);
        // bulk write after buffered single writes, at the end and over the middle
        s.seek(50000);
        s.write_item(Item{50000, 0});
        std::vector<Item> more(3);
        for (int k = 0; k < 3; ++k) more[k] = Item{50001 + k, 0};
        CHECK(s.write_array(more.data(), 3) == tpie::ami::NO_ERROR);
        CHECK(s.stream_len() == 50004 && s.tell() == 50004);
        s.seek(200);
        more[0].key = more[1].key = -2;
        CHECK(s.write_array(more.data(), 2) == tpie::ami::NO_ERROR && s.tell() == 202);
        s.seek(199);
        len = 4;
        CHECK(s.read_array(buf.data(), &len) == tpie::ami::NO_ERROR && len == 4);
        CHECK(buf[0].key == 199 && buf[1].key == -2 && buf[2].key == -2 && buf[3].key == 202);
        s.seek(50000);
        CHECK(s.read_item(&p) == tpie::ami::NO_ERROR && p->key == 50000);
        CHECK(s.read_item(&p) == tpie::ami::NO_ERROR && p->key == 50001);
        s.seek(50003);
        CHECK(s.read_item(&p) == tpie::ami::NO_ERROR && p->key == 50003);
        CHECK(s.read_item(&p) == tpie::ami::END_OF_STREAM);
        s.truncate(50000);
    }
    // named streams persist; APPEND_STREAM starts at the end
    {
        tpie::ami::stream<Item> s(dir + "/a.tpie", tpie::ami::APPEND_STREAM);
        CHECK(s.stream_len() == 50000 && s.tell() == 50000);
        s.write_item(Item{7, 7});
        CHECK(s.stream_len() == 50001);
        s.truncate(10);
        CHECK(s.stream_len() == 10 && s.tell() == 10);
    }
    // in-memory sort, stable on equal keys
    {
        tpie::ami::stream<Item> s;
        for (int i = 0; i < 1000; ++i) s.write_item(Item{(i * 7919) % 13, i});
        ByKey cmp;
        CHECK(tpie::ami::sort(&s, &cmp) == tpie::ami::NO_ERROR);
        Item *p, prev{-1, -1};
        int n = 0;
        while (s.read_item(&p) == tpie::ami::NO_ERROR) {
            CHECK(p->key > prev.key || (p->key == prev.key && p->val > prev.val));
            prev = *p;
            ++n;
        }
        CHECK(n == 1000);
    }
    // external sort: runs of the memory limit, merged (stable)
    {
        tpie::MM_manager.set_memory_limit(64 * 1024);
        tpie::ami::stream<Item> s;
        std::vector<int> hist(97, 0);
        for (int i = 0; i < 200000; ++i) {
            const int k = (int)(((long long)i * 2654435761LL) % 97);
            hist[k]++;
            s.write_item(Item{k, i});
        }
        ByKey cmp;
        CHECK(tpie::ami::sort(&s, &cmp) == tpie::ami::NO_ERROR);
        Item *p, prev{-1, -1};
        std::vector<int> got(97, 0);
        int n = 0;
        while (s.read_item(&p) == tpie::ami::NO_ERROR) {
            CHECK(p->key > prev.key || (p->key == prev.key && p->val > prev.val));
            got[p->key]++;
            prev = *p;
            ++n;
        }
        CHECK(n == 200000 && got == hist);
    }
    std::printf("tpie subset ok\n");
    return 0;
}
