// get_max_resident_mib (memory_usage.h of JeffersonLab/qphix): the peak
// resident host memory of the process in MiB (device memory is not counted).
#pragma once

#include <sys/resource.h>

namespace QPhiX {

inline double get_max_resident_mib()
{
  struct rusage ru;
  if (getrusage(RUSAGE_SELF, &ru) != 0) return 0.0;
  return ru.ru_maxrss / 1024.0;  // ru_maxrss is in KiB on Linux
}

}  // namespace QPhiX
