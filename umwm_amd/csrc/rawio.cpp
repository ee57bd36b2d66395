#include "rawio.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>

namespace umwm {
namespace {

constexpr char kMagic[8] = {'U', 'M', 'W', 'M', 'R', 'A', 'W', '1'};

struct File {
  std::FILE* f = nullptr;
  ~File() {
    if (f) std::fclose(f);
  }
};

bool put(std::FILE* f, const void* p, size_t n) { return std::fwrite(p, 1, n, f) == n; }
bool get(std::FILE* f, void* p, size_t n) { return std::fread(p, 1, n, f) == n; }

bool put_str(std::FILE* f, const std::string& s) {
  const int32_t n = (int32_t)s.size();
  return put(f, &n, 4) && put(f, s.data(), s.size());
}

bool get_str(std::FILE* f, std::string& s) {
  int32_t n = 0;
  if (!get(f, &n, 4) || n < 0 || n > (1 << 20)) return false;
  s.resize((size_t)n);
  return n == 0 || get(f, &s[0], (size_t)n);
}

}  // namespace

int raw_write(const std::string& path, const std::vector<RawVar>& vars, std::string& err) {
  File w;
  w.f = std::fopen(path.c_str(), "wb");
  if (!w.f) {
    err = "cannot open " + path + " for writing";
    return -1;
  }
  const int32_t nv = (int32_t)vars.size();
  bool ok = put(w.f, kMagic, 8) && put(w.f, &nv, 4);
  for (const RawVar& v : vars) {
    if (!ok) break;
    int64_t n = 1;
    for (int64_t s : v.shape) n *= s;
    if ((int64_t)v.data.size() != n || v.dims.size() != v.shape.size()) {
      err = "variable " + v.name + ": data does not match its shape";
      return -1;
    }
    const int32_t nd = (int32_t)v.dims.size();
    ok = put_str(w.f, v.name) && put_str(w.f, v.units) && put_str(w.f, v.description) &&
         put(w.f, &v.type, 4) && put(w.f, &nd, 4);
    for (int32_t k = 0; ok && k < nd; ++k) ok = put_str(w.f, v.dims[k]) && put(w.f, &v.shape[k], 8);
    if (!ok) break;
    if (v.type == RAW_F64) {
      ok = put(w.f, v.data.data(), v.data.size() * 8);
    } else if (v.type == RAW_F32) {
      std::vector<float> t(v.data.begin(), v.data.end());
      ok = put(w.f, t.data(), t.size() * 4);
    } else {
      std::vector<int32_t> t(v.data.size());
      for (size_t q = 0; q < t.size(); ++q) t[q] = (int32_t)std::lround(v.data[q]);
      ok = put(w.f, t.data(), t.size() * 4);
    }
  }
  if (!ok || std::fflush(w.f) != 0) {
    err = "write to " + path + " failed";
    return -1;
  }
  return 0;
}

int raw_read(const std::string& path, std::vector<RawVar>& vars, std::string& err) {
  File r;
  r.f = std::fopen(path.c_str(), "rb");
  if (!r.f) {
    err = "cannot open " + path;
    return -1;
  }
  char magic[8];
  int32_t nv = 0;
  if (!get(r.f, magic, 8) || std::memcmp(magic, kMagic, 8) != 0 || !get(r.f, &nv, 4) || nv < 0) {
    err = path + " is not a UMWMRAW1 file";
    return -1;
  }
  vars.clear();
  for (int32_t q = 0; q < nv; ++q) {
    RawVar v;
    int32_t nd = 0;
    bool ok = get_str(r.f, v.name) && get_str(r.f, v.units) && get_str(r.f, v.description) &&
              get(r.f, &v.type, 4) && get(r.f, &nd, 4) && nd >= 0 && nd <= 8;
    int64_t n = 1;
    for (int32_t k = 0; ok && k < nd; ++k) {
      std::string dn;
      int64_t s = 0;
      ok = get_str(r.f, dn) && get(r.f, &s, 8) && s >= 0;
      v.dims.push_back(dn);
      v.shape.push_back(s);
      n *= s;
    }
    if (ok && (n < 0 || n > (int64_t)1 << 40)) ok = false;
    if (ok) {
      v.data.resize((size_t)n);
      if (v.type == RAW_F64) {
        ok = get(r.f, v.data.data(), (size_t)n * 8);
      } else if (v.type == RAW_F32 || v.type == RAW_I32) {
        std::vector<float> t((size_t)n);
        ok = get(r.f, t.data(), (size_t)n * 4);
        if (ok && v.type == RAW_F32)
          for (int64_t k = 0; k < n; ++k) v.data[k] = t[k];
        if (ok && v.type == RAW_I32)
          for (int64_t k = 0; k < n; ++k) {
            int32_t x;
            std::memcpy(&x, &t[k], 4);
            v.data[k] = x;
          }
      } else {
        ok = false;
      }
    }
    if (!ok) {
      err = path + ": truncated or malformed variable " + std::to_string(q);
      return -1;
    }
    vars.push_back(std::move(v));
  }
  return 0;
}

}  // namespace umwm
