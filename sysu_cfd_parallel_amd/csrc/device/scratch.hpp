// Fixed device scratch for the entry points that allocate nothing.
#pragma once
#include <cstddef>

namespace s355 {

// doubles of the scratch array: 24 MiB per GPU, held until the process ends
constexpr size_t kScratchDoubles = size_t(3) << 20;

// Device address of the current device's scratch array.  The first call on a
// device looks the address up (the device must be current); later calls
// return the kept address.  Throws when the lookup fails.
double* scratch_for_current_device();

}  // namespace s355
