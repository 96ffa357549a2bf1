// khrr_tune_cache.hip -- the autotuner's results on disk (host code only).
//
// khrr_tune_gravity caches its winner per power-of-two size bucket in the
// context and khrr_tune_batched its winner for the batched form; these
// entry points read them back and write them to / read them from a JSON
// file, so a process can start from shapes tuned earlier on the same kind
// of device instead of timing candidates again.  The file names the
// device's architecture and CU count; a file made on another device is
// refused.  Layout:
//   {"format": "khrr-tune", "version": 1, "arch": "...", "cu_count": 256,
//    "gravity": [{"bucket": 20, "n_max": 1048576, "i_per_lane": 8, ...,
//                 "ms": 110.4}, ...],
//    "batched": {"i_per_lane": 4, ...}}          (batched: null if untuned)
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "khrr_internal.h"

namespace {

// A JSON value: just what the cache file holds (objects, arrays, strings,
// numbers, true/false/null).
struct JVal {
  enum Kind { kNull, kBool, kNum, kStr, kArr, kObj } kind = kNull;
  double num = 0.0;
  bool b = false;
  std::string str;
  std::vector<JVal> arr;
  std::vector<std::pair<std::string, JVal>> obj;
  const JVal* get(const char* key) const {
    if (kind != kObj) return nullptr;
    for (const auto& kv : obj)
      if (kv.first == key) return &kv.second;
    return nullptr;
  }
};

class JParser {
 public:
  JParser(const char* p, const char* end) : p_(p), end_(end) {}
  bool parse(JVal& out) {
    if (!value(out, 0)) return false;
    ws();
    return p_ == end_;
  }
  size_t offset(const char* begin) const { return (size_t)(p_ - begin); }

 private:
  const char* p_;
  const char* end_;
  void ws() {
    while (p_ < end_ && (*p_ == ' ' || *p_ == '\n' || *p_ == '\r' || *p_ == '\t')) ++p_;
  }
  bool lit(const char* w) {
    const size_t k = std::strlen(w);
    if ((size_t)(end_ - p_) < k || std::strncmp(p_, w, k) != 0) return false;
    p_ += k;
    return true;
  }
  bool string(std::string& s) {
    if (p_ >= end_ || *p_ != '"') return false;
    ++p_;
    while (p_ < end_ && *p_ != '"') {
      if (*p_ == '\\') {
        if (++p_ >= end_) return false;
        const char c = *p_;
        if (c != '"' && c != '\\' && c != '/') return false;  // no other escapes are written
        s.push_back(c);
      } else {
        s.push_back(*p_);
      }
      ++p_;
    }
    if (p_ >= end_) return false;
    ++p_;
    return true;
  }
  bool value(JVal& v, int depth) {
    if (depth > 16) return false;
    ws();
    if (p_ >= end_) return false;
    const char c = *p_;
    if (c == '{') {
      ++p_;
      v.kind = JVal::kObj;
      ws();
      if (p_ < end_ && *p_ == '}') return ++p_, true;
      for (;;) {
        ws();
        std::string key;
        if (!string(key)) return false;
        ws();
        if (p_ >= end_ || *p_ != ':') return false;
        ++p_;
        JVal item;
        if (!value(item, depth + 1)) return false;
        v.obj.emplace_back(std::move(key), std::move(item));
        ws();
        if (p_ < end_ && *p_ == ',') {
          ++p_;
          continue;
        }
        if (p_ < end_ && *p_ == '}') return ++p_, true;
        return false;
      }
    }
    if (c == '[') {
      ++p_;
      v.kind = JVal::kArr;
      ws();
      if (p_ < end_ && *p_ == ']') return ++p_, true;
      for (;;) {
        JVal item;
        if (!value(item, depth + 1)) return false;
        v.arr.push_back(std::move(item));
        ws();
        if (p_ < end_ && *p_ == ',') {
          ++p_;
          continue;
        }
        if (p_ < end_ && *p_ == ']') return ++p_, true;
        return false;
      }
    }
    if (c == '"') {
      v.kind = JVal::kStr;
      return string(v.str);
    }
    if (lit("true")) return v.kind = JVal::kBool, v.b = true, true;
    if (lit("false")) return v.kind = JVal::kBool, v.b = false, true;
    if (lit("null")) return v.kind = JVal::kNull, true;
    // number
    std::string tok;
    while (p_ < end_ && (std::strchr("+-0123456789.eE", *p_) != nullptr)) tok.push_back(*p_++);
    if (tok.empty()) return false;
    char* e = nullptr;
    errno = 0;
    v.num = std::strtod(tok.c_str(), &e);
    v.kind = JVal::kNum;
    return e && *e == '\0' && errno == 0 && std::isfinite(v.num);
  }
};

const char* const kFields[] = {"i_per_lane", "waves_per_block", "j_split", "j_tile", "symmetric",
                               "superblock"};

void put_tune(std::string& out, const khrr_tune& t) {
  const int v[] = {t.i_per_lane, t.waves_per_block, t.j_split, t.j_tile, t.symmetric,
                   t.superblock};
  for (int k = 0; k < 6; ++k) {
    out += "\"";
    out += kFields[k];
    out += "\": " + std::to_string(v[k]);
    if (k < 5) out += ", ";
  }
}

// A tune object's six integer fields (all required, whole numbers).
bool get_tune(const JVal& o, khrr_tune& t, std::string& why) {
  int v[6];
  for (int k = 0; k < 6; ++k) {
    const JVal* f = o.get(kFields[k]);
    if (!f || f->kind != JVal::kNum || f->num != std::floor(f->num) || std::fabs(f->num) > 1e9) {
      why = std::string("field \"") + kFields[k] + "\" missing or not an integer";
      return false;
    }
    v[k] = (int)f->num;
  }
  t = khrr_tune{v[0], v[1], v[2], v[3], v[4], v[5]};
  return true;
}

std::string quote(const std::string& s) {
  std::string q = "\"";
  for (char c : s) {
    if (c == '"' || c == '\\') q.push_back('\\');
    q.push_back(c);
  }
  return q + "\"";
}

}  // namespace

extern "C" {

int khrr_tune_cached(khrr_ctx* ctx, int n, khrr_tune* out) {
  if (!ctx || !out) return khrr::fail(ctx, KHRR_EINVAL, "null argument");
  if (n < 1) return khrr::fail(ctx, KHRR_EINVAL, "n < 1");
  const khrr_tune* t = khrr::cached_tune(ctx, n);
  if (!t)
    return khrr::fail(ctx, KHRR_ESTATE, "no tuned shape cached for %d bodies (bucket %d)", n,
                      khrr::tune_bucket(n));
  *out = *t;
  return 0;
}

int khrr_tune_batched_cached(khrr_ctx* ctx, khrr_tune* out) {
  if (!ctx || !out) return khrr::fail(ctx, KHRR_EINVAL, "null argument");
  if (!ctx->has_batch_tuned) return khrr::fail(ctx, KHRR_ESTATE, "no tuned batched shape cached");
  *out = ctx->batch_tuned;
  return 0;
}

int khrr_tune_save(khrr_ctx* ctx, const char* path) {
  if (!ctx || !path) return khrr::fail(ctx, KHRR_EINVAL, "null argument");
  std::string s = "{\n  \"format\": \"khrr-tune\",\n  \"version\": 1,\n  \"arch\": " +
                  quote(ctx->arch) + ",\n  \"cu_count\": " + std::to_string(ctx->cu_count) +
                  ",\n  \"gravity\": [";
  bool first = true;
  for (int b = 0; b < khrr::kTuneBuckets; ++b) {
    if (!ctx->has_tuned[b]) continue;
    char ms[64];
    std::snprintf(ms, sizeof(ms), "%.6g", ctx->tuned_ms[b]);
    s += first ? "\n    {" : ",\n    {";
    s += "\"bucket\": " + std::to_string(b) + ", \"n_max\": " + std::to_string(1ll << b) + ", ";
    put_tune(s, ctx->tuned[b]);
    s += std::string(", \"ms\": ") + ms + "}";
    first = false;
  }
  s += first ? "],\n" : "\n  ],\n";
  s += "  \"batched\": ";
  if (ctx->has_batch_tuned) {
    s += "{";
    put_tune(s, ctx->batch_tuned);
    s += "}\n}\n";
  } else {
    s += "null\n}\n";
  }
  const std::string tmp = std::string(path) + ".tmp";
  FILE* f = std::fopen(tmp.c_str(), "wb");
  if (!f) return khrr::fail(ctx, KHRR_EINVAL, "can't write %s: %s", tmp.c_str(), strerror(errno));
  const bool ok = std::fwrite(s.data(), 1, s.size(), f) == s.size();
  if (std::fclose(f) != 0 || !ok || std::rename(tmp.c_str(), path) != 0) {
    std::remove(tmp.c_str());
    return khrr::fail(ctx, KHRR_EINVAL, "can't write %s", path);
  }
  return 0;
}

int khrr_tune_load(khrr_ctx* ctx, const char* path) {
  if (!ctx || !path) return khrr::fail(ctx, KHRR_EINVAL, "null argument");
  FILE* f = std::fopen(path, "rb");
  if (!f) return khrr::fail(ctx, KHRR_EINVAL, "can't open %s: %s", path, strerror(errno));
  std::string text;
  char buf[4096];
  size_t got;
  while ((got = std::fread(buf, 1, sizeof(buf), f)) > 0 && text.size() < (1u << 20))
    text.append(buf, got);
  std::fclose(f);
  if (text.size() >= (1u << 20)) return khrr::fail(ctx, KHRR_EINVAL, "%s: too large", path);
  JVal root;
  JParser parser(text.data(), text.data() + text.size());
  if (!parser.parse(root) || root.kind != JVal::kObj)
    return khrr::fail(ctx, KHRR_EINVAL, "%s: not a JSON object (at byte %zu)", path,
                      parser.offset(text.data()));
  const JVal* fmt = root.get("format");
  const JVal* ver = root.get("version");
  if (!fmt || fmt->kind != JVal::kStr || fmt->str != "khrr-tune" || !ver ||
      ver->kind != JVal::kNum || ver->num != 1.0)
    return khrr::fail(ctx, KHRR_EINVAL, "%s: not a khrr-tune version 1 file", path);
  const JVal* arch = root.get("arch");
  const JVal* cus = root.get("cu_count");
  if (!arch || arch->kind != JVal::kStr || !cus || cus->kind != JVal::kNum)
    return khrr::fail(ctx, KHRR_EINVAL, "%s: arch / cu_count missing", path);
  if (arch->str != ctx->arch || (int)cus->num != ctx->cu_count)
    return khrr::fail(ctx, KHRR_ESTATE,
                      "%s was tuned on %s with %d CUs; this device is %s with %d CUs", path,
                      arch->str.c_str(), (int)cus->num, ctx->arch.c_str(), ctx->cu_count);
  // parse and check everything first; the context changes only if all of
  // the file is valid
  std::vector<std::pair<int, khrr_tune>> grav;
  std::vector<double> grav_ms;
  const JVal* g = root.get("gravity");
  if (!g || g->kind != JVal::kArr) return khrr::fail(ctx, KHRR_EINVAL, "%s: no gravity list", path);
  for (size_t k = 0; k < g->arr.size(); ++k) {
    const JVal& e = g->arr[k];
    const JVal* b = e.get("bucket");
    std::string why;
    khrr_tune t;
    if (!b || b->kind != JVal::kNum || b->num != std::floor(b->num) || b->num < 0 ||
        b->num >= khrr::kTuneBuckets)
      return khrr::fail(ctx, KHRR_EINVAL, "%s: gravity entry %zu: bad bucket", path, k);
    if (!get_tune(e, t, why))
      return khrr::fail(ctx, KHRR_EINVAL, "%s: gravity entry %zu: %s", path, k, why.c_str());
    if (khrr::validate_tune(ctx, t) != 0) {
      const std::string msg = ctx->err;
      return khrr::fail(ctx, KHRR_EINVAL, "%s: gravity entry %zu: %s", path, k, msg.c_str());
    }
    const JVal* ms = e.get("ms");
    grav.emplace_back((int)b->num, t);
    grav_ms.push_back(ms && ms->kind == JVal::kNum ? ms->num : 0.0);
  }
  bool has_batch = false;
  khrr_tune bt{};
  const JVal* bj = root.get("batched");
  if (bj && bj->kind == JVal::kObj) {
    std::string why;
    if (!get_tune(*bj, bt, why))
      return khrr::fail(ctx, KHRR_EINVAL, "%s: batched: %s", path, why.c_str());
    if (khrr::validate_tune(ctx, bt) != 0) {
      const std::string msg = ctx->err;
      return khrr::fail(ctx, KHRR_EINVAL, "%s: batched: %s", path, msg.c_str());
    }
    has_batch = true;
  } else if (bj && bj->kind != JVal::kNull) {
    return khrr::fail(ctx, KHRR_EINVAL, "%s: batched must be an object or null", path);
  }
  for (size_t k = 0; k < grav.size(); ++k) {
    ctx->tuned[grav[k].first] = grav[k].second;
    ctx->has_tuned[grav[k].first] = true;
    ctx->tuned_ms[grav[k].first] = grav_ms[k];
  }
  if (has_batch) {
    ctx->batch_tuned = bt;
    ctx->has_batch_tuned = true;
  }
  return 0;
}

}  // extern "C"
