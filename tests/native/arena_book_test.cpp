// Host-side check of the scratch-arena rules (fri-cc_amd/csrc/arena.h) with a
// stand-in allocator that records every allocation, free and device wait.
#include <stdio.h>
#include <stdlib.h>

#include <set>
#include <string>

#include "arena.h"

struct FakeAlloc {
  std::set<void*> live;
  int n_alloc = 0, n_free = 0, n_sync = 0;
  int alloc(void** p, size_t b) {
    *p = malloc(b);
    live.insert(*p);
    n_alloc++;
    return 0;
  }
  int free(void* p) {
    if (!live.count(p)) return 7;  // double free / unknown pointer
    live.erase(p);
    ::free(p);
    n_free++;
    return 0;
  }
  int sync() {
    n_sync++;
    return 0;
  }
  const char* error_string(int) { return "fake"; }
};

static int fails = 0;
#define CHECK(c)                                            \
  do {                                                      \
    if (!(c)) {                                             \
      fprintf(stderr, "FAIL %s:%d: %s\n", __FILE__, __LINE__, #c); \
      fails++;                                              \
    }                                                       \
  } while (0)

int main() {
  using namespace fricc;
  FakeAlloc al;
  static ArenaBook<FakeAlloc> book;
  std::string err;
  void *p1 = nullptr, *p2 = nullptr, *p3 = nullptr;

  // a slot grows on demand and is reused while large enough
  CHECK(book.get(al, 0, 5, 1000, &p1, &err) == ARENA_OK && p1);
  CHECK(book.get(al, 0, 5, 800, &p2, &err) == ARENA_OK && p2 == p1);
  CHECK(al.n_alloc == 1 && al.n_sync == 0);
  // growth frees the old buffer after a device wait
  CHECK(book.get(al, 0, 5, 100000, &p3, &err) == ARENA_OK);
  CHECK(al.n_alloc == 2 && al.n_free == 1 && al.n_sync == 1);
  CHECK(!al.live.count(p1));

  // arenas are independent
  void* q = nullptr;
  CHECK(book.get(al, 3, 5, 100, &q, &err) == ARENA_OK && q != p3);

  // a pinned arena never frees: growth is an error, the pointer stays valid
  CHECK(book.pin(0, +1, &err) == ARENA_OK && book.pins(0) == 1);
  void* r = nullptr;
  CHECK(book.get(al, 0, 5, 100000, &r, &err) == ARENA_OK && r == p3);  // fits: same buffer
  const int frees = al.n_free;
  CHECK(book.get(al, 0, 5, 1 << 24, &r, &err) == ARENA_EINVAL);
  CHECK(err.find("recorded graph") != std::string::npos);
  CHECK(al.n_free == frees && al.live.count(p3));
  // a new slot of a pinned arena may still be allocated (nothing is freed)
  void* s = nullptr;
  CHECK(book.get(al, 0, 9, 64, &s, &err) == ARENA_OK && s);
  // release is refused while pinned
  CHECK(book.release(al, 0, &err) == ARENA_EINVAL && al.live.count(p3));
  // other arenas are unaffected by the pin
  CHECK(book.get(al, 3, 5, 1 << 20, &q, &err) == ARENA_OK);

  // unpinning restores growth and release
  CHECK(book.pin(0, -1, &err) == ARENA_OK && book.pins(0) == 0);
  CHECK(book.pin(0, -1, &err) == ARENA_EINVAL);  // not pinned
  CHECK(book.get(al, 0, 5, 1 << 24, &r, &err) == ARENA_OK && r != p3 && !al.live.count(p3));
  CHECK(book.release(al, 0, &err) == ARENA_OK);
  CHECK(book.release(al, 3, &err) == ARENA_OK);
  CHECK(al.live.empty());

  // nested pins: two graphs on one arena
  CHECK(book.pin(7, +1, &err) == ARENA_OK && book.pin(7, +1, &err) == ARENA_OK);
  CHECK(book.pin(7, -1, &err) == ARENA_OK && book.pins(7) == 1);
  CHECK(book.release(al, 7, &err) == ARENA_EINVAL);
  CHECK(book.pin(7, -1, &err) == ARENA_OK && book.release(al, 7, &err) == ARENA_OK);

  // bounds
  CHECK(book.get(al, -1, 0, 8, &r, &err) == ARENA_EINVAL);
  CHECK(book.get(al, NARENA, 0, 8, &r, &err) == ARENA_EINVAL);
  CHECK(book.get(al, 0, NSLOT, 8, &r, &err) == ARENA_EINVAL);
  CHECK(book.pin(NARENA, +1, &err) == ARENA_EINVAL);
  CHECK(book.pin(0, 2, &err) == ARENA_EINVAL);

  if (fails) return 1;
  printf("arena book ok\n");
  return 0;
}
