// internal.h -- shared host-side helpers of the aloha_amd library.
#pragma once

#include <cstdarg>
#include <cstdio>
#include <string>

#include "aloha_amd.h"

namespace aloha {

// Error text of the last failing call that had no context (thread-local).
std::string &global_error();

// Format into s; returns code so callers can `return fail(...)`.
inline int format_error(std::string &s, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  s = buf;
  return code;
}

// Pixel bounds of the rays that can reach the boxes (aloha_project_boxes).
void project_boxes(const float cam[12], const aloha_box3f *boxes, int n, int W, int H,
                   int32_t out[4]);

}  // namespace aloha

#define ALOHA_FAIL_GLOBAL(code, ...) ::aloha::format_error(::aloha::global_error(), code, __VA_ARGS__)
