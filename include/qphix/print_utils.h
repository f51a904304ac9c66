// masterPrintf / localPrintf (print_utils.h of JeffersonLab/qphix). One
// process drives one GPU here, so both print on stdout.
#pragma once

#include <cstdarg>
#include <cstdio>

namespace QPhiX {

inline void masterPrintf(const char *format, ...)
{
  va_list ap;
  va_start(ap, format);
  std::vprintf(format, ap);
  va_end(ap);
  std::fflush(stdout);
}

inline void localPrintf(const char *format, ...)
{
  va_list ap;
  va_start(ap, format);
  std::vprintf(format, ap);
  va_end(ap);
  std::fflush(stdout);
}

}  // namespace QPhiX
