// instance_io.cpp -- reader for the reference's instance JSON schema.
//
// Schema (written by scripts/sample.py:43-65, read by
// simba/src/instance.cpp:7-27 and vfb/src/utils.h:15-44):
//   {"original_formula": str, "num_variables": int, "num_examples": int,
//    "examples": [{"inputs": [{"value": u, "size": 32}, ...],
//                  "output": [{"value": u, "size": 32}]}, ...]}
// Inputs are laid out [sample][var].  The reference reads values with
// get<uint32_t>(); here they are read as unsigned 64-bit and the word width
// is the "size" field of the values (32 when absent).
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/mba_api.h"

namespace {

struct JVal {
  enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
  bool neg = false;
  uint64_t u = 0;     // magnitude of an integer
  bool integral = true;
  std::string s;
  std::vector<JVal> arr;
  std::vector<std::pair<std::string, JVal>> obj;
  const JVal *get(const char *k) const {
    for (auto &kv : obj)
      if (kv.first == k) return &kv.second;
    return nullptr;
  }
};

struct Parser {
  const char *p, *e;
  bool ok = true;
  void ws() {
    while (p < e && std::isspace((unsigned char)*p)) ++p;
  }
  bool lit(const char *w) {
    size_t n = std::strlen(w);
    if ((size_t)(e - p) >= n && std::memcmp(p, w, n) == 0) {
      p += n;
      return true;
    }
    return false;
  }
  std::string str() {
    std::string out;
    ++p;  // opening quote
    while (p < e && *p != '"') {
      if (*p == '\\' && p + 1 < e) {
        ++p;
        switch (*p) {
          case 'n': out += '\n'; break;
          case 't': out += '\t'; break;
          case 'u': p += 4; out += '?'; break;  // escapes are not used by the schema
          default: out += *p;
        }
        ++p;
      } else {
        out += *p++;
      }
    }
    if (p >= e) ok = false;
    else ++p;
    return out;
  }
  JVal val(int depth) {
    JVal v;
    ws();
    if (p >= e || depth > 64) {
      ok = false;
      return v;
    }
    if (*p == '{') {
      v.kind = JVal::OBJ;
      ++p;
      ws();
      if (p < e && *p == '}') {
        ++p;
        return v;
      }
      while (ok) {
        ws();
        if (p >= e || *p != '"') {
          ok = false;
          break;
        }
        std::string k = str();
        ws();
        if (p >= e || *p != ':') {
          ok = false;
          break;
        }
        ++p;
        JVal x = val(depth + 1);
        v.obj.emplace_back(std::move(k), std::move(x));
        ws();
        if (p < e && *p == ',') {
          ++p;
          continue;
        }
        if (p < e && *p == '}') {
          ++p;
          break;
        }
        ok = false;
      }
    } else if (*p == '[') {
      v.kind = JVal::ARR;
      ++p;
      ws();
      if (p < e && *p == ']') {
        ++p;
        return v;
      }
      while (ok) {
        v.arr.push_back(val(depth + 1));
        ws();
        if (p < e && *p == ',') {
          ++p;
          continue;
        }
        if (p < e && *p == ']') {
          ++p;
          break;
        }
        ok = false;
      }
    } else if (*p == '"') {
      v.kind = JVal::STR;
      v.s = str();
    } else if (lit("true")) {
      v.kind = JVal::BOOL;
      v.u = 1;
    } else if (lit("false")) {
      v.kind = JVal::BOOL;
    } else if (lit("null")) {
      v.kind = JVal::NUL;
    } else {
      v.kind = JVal::NUM;
      if (*p == '-') {
        v.neg = true;
        ++p;
      }
      if (p >= e || !std::isdigit((unsigned char)*p)) {
        ok = false;
        return v;
      }
      unsigned __int128 acc = 0;
      while (p < e && std::isdigit((unsigned char)*p)) {
        acc = acc * 10 + (unsigned)(*p++ - '0');
        if (acc >> 64) ok = false;
      }
      if (p < e && (*p == '.' || *p == 'e' || *p == 'E')) {
        v.integral = false;
        while (p < e && (std::isdigit((unsigned char)*p) || *p == '.' || *p == 'e' ||
                         *p == 'E' || *p == '+' || *p == '-'))
          ++p;
      }
      v.u = (uint64_t)acc;
    }
    return v;
  }
};

bool as_uint(const JVal *v, uint64_t &out) {
  if (!v || v->kind != JVal::NUM || !v->integral) return false;
  out = v->neg ? (uint64_t)(0 - v->u) : v->u;
  return true;
}

}  // namespace

extern "C" int mba_instance_read_json(const char *path, mba_instance *out) {
  if (!path || !out) return MBA_ERR_ARG;
  std::memset(out, 0, sizeof *out);
  FILE *f = std::fopen(path, "rb");
  if (!f) return MBA_ERR_IO;
  std::string buf;
  char tmp[65536];
  size_t k;
  while ((k = std::fread(tmp, 1, sizeof tmp, f)) > 0) buf.append(tmp, k);
  std::fclose(f);
  Parser ps{buf.data(), buf.data() + buf.size()};
  JVal root = ps.val(0);
  if (!ps.ok || root.kind != JVal::OBJ) return MBA_ERR_IO;
  uint64_t nv = 0, ne = 0;
  if (!as_uint(root.get("num_variables"), nv) || !as_uint(root.get("num_examples"), ne))
    return MBA_ERR_IO;
  const JVal *ex = root.get("examples");
  if (!ex || ex->kind != JVal::ARR || ex->arr.size() != ne || nv > MBA_MAX_VARS ||
      ne > MBA_MAX_SAMPLES || ne == 0)
    return MBA_ERR_IO;
  std::vector<uint64_t> in, outv;
  uint32_t width = 0;
  for (const JVal &e : ex->arr) {
    const JVal *ins = e.get("inputs");
    const JVal *o = e.get("output");
    if (!ins || ins->kind != JVal::ARR || ins->arr.size() != nv || !o ||
        o->kind != JVal::ARR || o->arr.empty())
      return MBA_ERR_IO;
    for (const JVal &x : ins->arr) {
      uint64_t v;
      if (!as_uint(x.get("value"), v)) return MBA_ERR_IO;
      uint64_t sz;
      if (width == 0 && as_uint(x.get("size"), sz)) width = (uint32_t)sz;
      in.push_back(v);
    }
    uint64_t v;
    if (!as_uint(o->arr[0].get("value"), v)) return MBA_ERR_IO;
    uint64_t sz;
    if (width == 0 && as_uint(o->arr[0].get("size"), sz)) width = (uint32_t)sz;
    outv.push_back(v);
  }
  if (width == 0) width = 32;
  if (width != 32 && width != 64) return MBA_ERR_IO;
  const uint64_t m = width == 64 ? ~0ull : 0xffffffffull;
  uint64_t *pi = (uint64_t *)std::malloc(sizeof(uint64_t) * (in.size() ? in.size() : 1));
  uint64_t *po = (uint64_t *)std::malloc(sizeof(uint64_t) * outv.size());
  if (!pi || !po) {
    std::free(pi);
    std::free(po);
    return MBA_ERR_NOMEM;
  }
  for (size_t i = 0; i < in.size(); ++i) pi[i] = in[i] & m;
  for (size_t i = 0; i < outv.size(); ++i) po[i] = outv[i] & m;
  out->width = width;
  out->n_vars = (uint32_t)nv;
  out->n_samples = (uint32_t)ne;
  out->n_consts = 0;
  out->inputs = pi;
  out->outputs = po;
  out->consts = nullptr;
  return 0;
}

extern "C" void mba_instance_free(mba_instance *inst) {
  if (!inst) return;
  std::free((void *)inst->inputs);
  std::free((void *)inst->outputs);
  std::free((void *)inst->consts);
  std::memset(inst, 0, sizeof *inst);
}
