// runtime.hpp -- device selection, memory and cached workspaces.
#pragma once
#include <cstddef>

namespace v3d {

// Workspace slots.  Each slot is a device buffer owned by the library that
// grows on demand (hipMalloc) and is reused across calls.  Calls that use the
// same slot must not overlap on different streams.
enum WorkspaceSlot {
  kWsFarPartial = 0,
  kWsSht0,
  kWsSht1,
  kWsSht2,
  kWsSing0,
  kWsSing1,
  kWsSing2,
  kWsSing3,
  kWsSing4,
  kWsSolver,
  kWsFarRecords,
  kWsFarCoarse,
  kWsFarCoarse2,
  kWsFarBounds,
  kWsCount
};

// Returns a device buffer of at least `bytes` for `slot`, or nullptr on
// allocation failure.  Growing a slot frees the old buffer (hipFree
// synchronises the device).
void *workspace(int slot, size_t bytes);

}  // namespace v3d
