// Bookkeeping of the library's scratch arenas, independent of the device
// runtime so that its rules can be tested on the host with a stand-in
// allocator.
//
// Per device there are NARENA arenas of NSLOT grow-only slots.  A slot that
// must grow is freed and allocated again (its contents are not kept).  A
// recorded graph holds the addresses of the slots it was recorded with, so an
// arena can be pinned: while its pin count is above zero, a slot of that arena
// that would have to grow is an error instead (the graph would otherwise be
// left pointing at freed memory; a slot never allocated may still be), and
// releasing the arena is refused.
//
// Alloc must provide:
//   int  alloc(void** p, size_t bytes);   // 0 on success
//   int  free(void* p);                   // 0 on success
//   int  sync();                          // wait for all work on the device
//   const char* error_string(int code);
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <string>

namespace fricc {

constexpr int NARENA = 64;
constexpr int NSLOT = 64;

enum ArenaStatus { ARENA_OK = 0, ARENA_EINVAL = 1, ARENA_EDEVICE = 2, ARENA_ENOMEM = 3 };

struct ArenaSlots {
  void* ptr[NSLOT] = {nullptr};
  size_t cap[NSLOT] = {0};
  int pins = 0;
};

template <class Alloc>
struct ArenaBook {
  ArenaSlots a[NARENA];

  static bool valid_arena(int arena) { return arena >= 0 && arena < NARENA; }

  // Slot `slot` of `arena` with room for `bytes`; grows it when needed (and
  // allowed).  On failure *err receives the reason.
  int get(Alloc& al, int arena, int slot, size_t bytes, void** out, std::string* err) {
    if (!valid_arena(arena) || slot < 0 || slot >= NSLOT) {
      *err = "workspace arena or slot out of range";
      return ARENA_EINVAL;
    }
    ArenaSlots& w = a[arena];
    if (bytes == 0) bytes = 256;
    if (w.cap[slot] < bytes) {
      if (w.pins > 0 && w.ptr[slot]) {
        *err = "workspace arena " + std::to_string(arena) + " is held by a recorded graph: slot " +
               std::to_string(slot) + " would have to grow from " + std::to_string(w.cap[slot]) + " to " +
               std::to_string(bytes) + " bytes (release the graph, or run larger shapes on another arena)";
        return ARENA_EINVAL;
      }
      if (w.ptr[slot]) {
        int e = al.sync();
        if (e) {
          *err = std::string("workspace: device synchronise failed: ") + al.error_string(e);
          return ARENA_EDEVICE;
        }
        e = al.free(w.ptr[slot]);
        w.ptr[slot] = nullptr;
        w.cap[slot] = 0;
        if (e) {
          *err = std::string("workspace: free failed: ") + al.error_string(e);
          return ARENA_EDEVICE;
        }
      }
      const size_t b = bytes + bytes / 8 + 256;
      void* p = nullptr;
      int e = al.alloc(&p, b);
      if (e || !p) {
        *err = std::string("workspace allocation failed: ") + al.error_string(e);
        return ARENA_ENOMEM;
      }
      w.ptr[slot] = p;
      w.cap[slot] = b;
    }
    *out = w.ptr[slot];
    return ARENA_OK;
  }

  // Frees every slot of the arena (waits for the device first).
  int release(Alloc& al, int arena, std::string* err) {
    if (!valid_arena(arena)) {
      *err = "workspace arena out of range (0..63)";
      return ARENA_EINVAL;
    }
    ArenaSlots& w = a[arena];
    if (w.pins > 0) {
      *err = "workspace arena " + std::to_string(arena) + " is held by " + std::to_string(w.pins) +
             " recorded graph(s) and cannot be released";
      return ARENA_EINVAL;
    }
    bool any = false;
    for (int s = 0; s < NSLOT; s++) any = any || w.ptr[s];
    if (!any) return ARENA_OK;
    int e = al.sync();
    if (e) {
      *err = std::string("workspace: device synchronise failed: ") + al.error_string(e);
      return ARENA_EDEVICE;
    }
    for (int s = 0; s < NSLOT; s++) {
      if (w.ptr[s]) {
        e = al.free(w.ptr[s]);
        if (e) {
          *err = std::string("workspace: free failed: ") + al.error_string(e);
          return ARENA_EDEVICE;
        }
      }
      w.ptr[s] = nullptr;
      w.cap[s] = 0;
    }
    return ARENA_OK;
  }

  // delta = +1: a recorded graph now holds the arena; -1: it was released.
  int pin(int arena, int delta, std::string* err) {
    if (!valid_arena(arena) || (delta != 1 && delta != -1)) {
      *err = "workspace_pin: arena out of range or delta not +1/-1";
      return ARENA_EINVAL;
    }
    if (delta < 0 && a[arena].pins == 0) {
      *err = "workspace_pin: arena " + std::to_string(arena) + " is not pinned";
      return ARENA_EINVAL;
    }
    a[arena].pins += delta;
    return ARENA_OK;
  }

  int pins(int arena) const { return valid_arena(arena) ? a[arena].pins : -1; }
};

}  // namespace fricc
