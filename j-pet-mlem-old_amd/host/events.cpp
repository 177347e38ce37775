// Event ingestion C ABI (include/strip_events.h) over host/io.hpp.
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "io.hpp"
#include "strip_events.h"

namespace {
thread_local std::string g_error;

bool ends_with_txt(const std::string& s) {
  const size_t dot = s.find_last_of('.');
  const size_t sep = s.find_last_of("\\/");
  if (dot == std::string::npos || (sep != std::string::npos && sep > dot))
    return false;
  return s.substr(dot) == ".txt";
}
}  // namespace

extern "C" {

int strip_events_read(const char* path, int is_3d, float** events,
                      int64_t* n_events) {
  g_error.clear();
  if (!path || !events || !n_events) {
    g_error = "NULL argument";
    return -1;
  }
  *events = nullptr;
  *n_events = 0;
  try {
    std::vector<float> v;
    if (ends_with_txt(path)) {
      strip_io::append_text_events(path, is_3d != 0, v);
    } else {
      if (is_3d) throw std::runtime_error("3D input must have .txt extension");
      strip_io::append_binary_events(path, v);
    }
    float* out = static_cast<float*>(std::malloc(sizeof(float) * (v.size() + 1)));
    if (!out) throw std::runtime_error("out of memory");
    std::memcpy(out, v.data(), sizeof(float) * v.size());
    *events = out;
    *n_events = (int64_t)(v.size() / 3);
    return 0;
  } catch (const std::exception& ex) {
    g_error = ex.what();
  }
  return -1;
}

void strip_events_free(float* events) { std::free(events); }

void strip_events_scale(float* events, int64_t n, int apply_length,
                        double scale_length, int dl_is_time,
                        double speed_of_light, double scale_time) {
  if (!events) return;
  if (apply_length)
    for (int64_t i = 0; i < n; ++i) {
      events[3 * i] *= scale_length;
      events[3 * i + 1] *= scale_length;
      if (!dl_is_time) events[3 * i + 2] *= scale_length;
    }
  if (dl_is_time)
    for (int64_t i = 0; i < n; ++i)
      events[3 * i + 2] = events[3 * i + 2] * speed_of_light * scale_time;
}

const char* strip_events_last_error(void) { return g_error.c_str(); }

}  // extern "C"
