// Device scratch owned by a context: named slots that grow on demand and are
// reused across frames (no per-frame hipMalloc once the sizes are reached).
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <string>

namespace splhip {

enum Slot : int {
  S_REC, S_ACT, S_HIST, S_SCAN0, S_SCAN1, S_SCAN2, S_SCAN3,
  S_TOTALS, S_SOA, S_RECSTAGE, S_PIC, S_MAP, S_AUX0, S_JOBS, S_GRAD0, S_GRAD1, S_GRAD2, S_GRAD3,
  S_GRAD4, S_GRAD5, S_GRAD6, S_GRAD7, S_GRAD8, S_GRAD9, S_SCRATCH, S_JOBS2, S_NORM,
  S_PASS  // first per-pass slot
};

// Slots of one render pass (both passes' data live at the same time).
enum PassSlot : int {
  P_POS, P_MO, P_K0, P_V0, P_K1, P_V1, P_RECS, P_CNT, P_OFF, P_EK0, P_EV0, P_EK1, P_EV1,
  P_BINOFF, P_ICNT, P_IOFF, P_IK0, P_IV0, P_IK1, P_IV1, P_TRACE,
  P_CNT16, P_OFF16, P_ROFF, P_UCNT, P_ENT, P_CCNT, P_CPRE, P_SX, P_SBASE, P_SBAT, P_SOPS, P_SQ,
  P_NSLOTS
};

constexpr int S_NSLOTS = S_PASS + 2 * P_NSLOTS;

// Slot x of pass p (0: ordered, 1: additive).
constexpr int pslot(int p, int x) { return S_PASS + p * P_NSLOTS + x; }

struct HipError {
  hipError_t code;
  const char *what;
};

inline void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) throw HipError{e, what};
}

class Workspace {
 public:
  ~Workspace() { release(); }
  // A frozen workspace never grows: a request past a slot's capacity throws
  // (no allocation or device synchronisation, e.g. under stream capture).
  void freeze(bool f) { frozen_ = f; }
  void *get(int slot, size_t bytes) {
    if (bytes == 0) bytes = 256;
    if (cap_[slot] < bytes && frozen_)
      throw HipError{hipErrorOutOfMemory, "request past the reserved workspace"};
    if (cap_[slot] < bytes) {
      if (ptr_[slot]) {
        hip_check(hipDeviceSynchronize(), "sync before workspace growth");
        hip_check(hipFree(ptr_[slot]), "hipFree");
        ptr_[slot] = nullptr;
      }
      size_t c = bytes + bytes / 8;  // headroom against small growth steps
      c = (c + 4095) & ~size_t(4095);
      hip_check(hipMalloc(&ptr_[slot], c), "hipMalloc workspace");
      cap_[slot] = c;
    }
    return ptr_[slot];
  }
  template <class T>
  T *get_as(int slot, size_t count) {
    return static_cast<T *>(get(slot, count * sizeof(T)));
  }
  size_t total_bytes() const {
    size_t s = 0;
    for (int i = 0; i < S_NSLOTS; ++i) s += cap_[i];
    return s;
  }
  void release() {
    for (int i = 0; i < S_NSLOTS; ++i) {
      if (ptr_[i]) (void)hipFree(ptr_[i]);
      ptr_[i] = nullptr;
      cap_[i] = 0;
    }
  }

 private:
  void *ptr_[S_NSLOTS] = {};
  size_t cap_[S_NSLOTS] = {};
  bool frozen_ = false;
};

}  // namespace splhip
