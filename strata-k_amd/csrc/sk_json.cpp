// sk_json.cpp — a serialised CoreProgram (JSON, as strata-ir's serde derives
// write it: crates/strata-ir/src/core/mod.rs:26-94, docs/ir-encoding.md)
// decoded into the flattened sk_program of the C ABI.
//
// serde's encoding: structs are objects with snake_case fields, unknown
// fields ignored, missing fields an error unless defaulted
// (`neg_weight_cycle_check`); CoreLiteral and CoreTerm are adjacently
// tagged, {"kind": ..., "data": ...}; Semiring and AggOp are bare strings;
// SymbolId is its u32.  `neg_weight_cycle_check` is read and dropped (the
// engine checks the universe + 2 round cap of every Trop stratum whatever
// the flag says).  A compound term (@terms) decodes to an
// SK_T_COMPOUND term, which preparing the program refuses like the struct
// path does.
//
// The parser is a small recursive-descent reader of RFC 8259 JSON with a
// nesting limit; numbers must be integers (a weight or id never has a
// fraction or exponent).
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "strata_k.h"

namespace {

struct JVal {
  enum Kind { NUL, BOOL, INT, STR, ARR, OBJ } kind = NUL;
  bool b = false;
  int64_t i = 0;
  bool big = false;  // an integer outside i64
  std::string s;
  std::vector<JVal> a;
  std::vector<std::pair<std::string, JVal>> o;
  const JVal* get(const char* k) const {
    for (auto& kv : o)
      if (kv.first == k) return &kv.second;
    return nullptr;
  }
};

struct JErr {
  std::string msg;
};

class Parser {
 public:
  Parser(const char* p, size_t n) : p_(p), e_(p + n), b_(p) {}
  JVal parse() {
    JVal v = value(0);
    ws();
    if (p_ != e_) fail("trailing characters after the document");
    return v;
  }

 private:
  const char *p_, *e_, *b_;
  static constexpr int kMaxDepth = 256;

  [[noreturn]] void fail(const std::string& m) {
    throw JErr{"invalid JSON at byte " + std::to_string(p_ - b_) + ": " + m};
  }
  void ws() {
    while (p_ < e_ && (*p_ == ' ' || *p_ == '\t' || *p_ == '\n' || *p_ == '\r')) p_++;
  }
  bool lit(const char* w) {
    const size_t n = std::strlen(w);
    if ((size_t)(e_ - p_) >= n && std::memcmp(p_, w, n) == 0) {
      p_ += n;
      return true;
    }
    return false;
  }
  static void utf8(std::string& out, uint32_t c) {
    if (c < 0x80) {
      out += (char)c;
    } else if (c < 0x800) {
      out += (char)(0xC0 | (c >> 6));
      out += (char)(0x80 | (c & 0x3F));
    } else if (c < 0x10000) {
      out += (char)(0xE0 | (c >> 12));
      out += (char)(0x80 | ((c >> 6) & 0x3F));
      out += (char)(0x80 | (c & 0x3F));
    } else {
      out += (char)(0xF0 | (c >> 18));
      out += (char)(0x80 | ((c >> 12) & 0x3F));
      out += (char)(0x80 | ((c >> 6) & 0x3F));
      out += (char)(0x80 | (c & 0x3F));
    }
  }
  uint32_t hex4() {
    if (e_ - p_ < 4) fail("short \\u escape");
    uint32_t c = 0;
    for (int k = 0; k < 4; k++) {
      const char h = *p_++;
      c <<= 4;
      if (h >= '0' && h <= '9') c |= (uint32_t)(h - '0');
      else if (h >= 'a' && h <= 'f') c |= (uint32_t)(h - 'a' + 10);
      else if (h >= 'A' && h <= 'F') c |= (uint32_t)(h - 'A' + 10);
      else fail("bad \\u escape");
    }
    return c;
  }
  std::string str() {
    p_++;  // opening quote
    std::string out;
    while (true) {
      if (p_ >= e_) fail("unterminated string");
      const char c = *p_++;
      if (c == '"') return out;
      if ((unsigned char)c < 0x20) fail("control character in a string");
      if (c != '\\') {
        out += c;
        continue;
      }
      if (p_ >= e_) fail("unterminated escape");
      const char x = *p_++;
      switch (x) {
        case '"': out += '"'; break;
        case '\\': out += '\\'; break;
        case '/': out += '/'; break;
        case 'b': out += '\b'; break;
        case 'f': out += '\f'; break;
        case 'n': out += '\n'; break;
        case 'r': out += '\r'; break;
        case 't': out += '\t'; break;
        case 'u': {
          uint32_t c1 = hex4();
          if (c1 >= 0xD800 && c1 < 0xDC00) {
            if (!(e_ - p_ >= 2 && p_[0] == '\\' && p_[1] == 'u')) fail("unpaired surrogate");
            p_ += 2;
            const uint32_t c2 = hex4();
            if (c2 < 0xDC00 || c2 >= 0xE000) fail("unpaired surrogate");
            c1 = 0x10000 + ((c1 - 0xD800) << 10) + (c2 - 0xDC00);
          } else if (c1 >= 0xDC00 && c1 < 0xE000) {
            fail("unpaired surrogate");
          }
          utf8(out, c1);
          break;
        }
        default: fail("bad escape");
      }
    }
  }
  JVal number() {
    JVal v;
    v.kind = JVal::INT;
    const char* s = p_;
    if (*p_ == '-') p_++;
    if (p_ >= e_ || *p_ < '0' || *p_ > '9') fail("bad number");
    if (*p_ == '0' && p_ + 1 < e_ && p_[1] >= '0' && p_[1] <= '9') fail("leading zero");
    while (p_ < e_ && *p_ >= '0' && *p_ <= '9') p_++;
    if (p_ < e_ && (*p_ == '.' || *p_ == 'e' || *p_ == 'E')) fail("a number with a fraction or exponent");
    const std::string t(s, p_);
    errno = 0;
    char* end = nullptr;
    const long long x = std::strtoll(t.c_str(), &end, 10);
    if (errno == ERANGE) v.big = true;
    v.i = (int64_t)x;
    return v;
  }
  JVal value(int depth) {
    if (depth > kMaxDepth) fail("nesting too deep");
    ws();
    if (p_ >= e_) fail("unexpected end");
    JVal v;
    const char c = *p_;
    if (c == '{') {
      v.kind = JVal::OBJ;
      p_++;
      ws();
      if (p_ < e_ && *p_ == '}') {
        p_++;
        return v;
      }
      while (true) {
        ws();
        if (p_ >= e_ || *p_ != '"') fail("expected a field name");
        std::string k = str();
        ws();
        if (p_ >= e_ || *p_ != ':') fail("expected ':'");
        p_++;
        v.o.emplace_back(std::move(k), value(depth + 1));
        ws();
        if (p_ < e_ && *p_ == ',') {
          p_++;
          continue;
        }
        if (p_ < e_ && *p_ == '}') {
          p_++;
          return v;
        }
        fail("expected ',' or '}'");
      }
    }
    if (c == '[') {
      v.kind = JVal::ARR;
      p_++;
      ws();
      if (p_ < e_ && *p_ == ']') {
        p_++;
        return v;
      }
      while (true) {
        v.a.push_back(value(depth + 1));
        ws();
        if (p_ < e_ && *p_ == ',') {
          p_++;
          continue;
        }
        if (p_ < e_ && *p_ == ']') {
          p_++;
          return v;
        }
        fail("expected ',' or ']'");
      }
    }
    if (c == '"') {
      v.kind = JVal::STR;
      v.s = str();
      return v;
    }
    if (c == '-' || (c >= '0' && c <= '9')) return number();
    if (lit("true")) {
      v.kind = JVal::BOOL;
      v.b = true;
      return v;
    }
    if (lit("false")) {
      v.kind = JVal::BOOL;
      return v;
    }
    if (lit("null")) return v;
    fail("unexpected character");
  }
};

// The decoded program and the storage its pointers refer to.
struct Decoded {
  std::vector<std::string> names;  // predicate names, then atom predicates
  std::vector<sk_pred> preds;
  std::vector<sk_rule> rules;
  std::vector<sk_atom> atoms;
  std::vector<sk_term> terms;
  std::vector<size_t> atom_name;  // index into names per atom
  sk_program prog{};
};

[[noreturn]] void bad(const std::string& where, const std::string& m) { throw JErr{where + ": " + m}; }

const JVal& field(const JVal& o, const char* k, const std::string& where) {
  if (o.kind != JVal::OBJ) bad(where, "expected an object");
  const JVal* v = o.get(k);
  if (!v) bad(where, std::string("missing field `") + k + "`");
  return *v;
}

uint32_t u32_of(const JVal& v, const std::string& where) {
  if (v.kind != JVal::INT || v.big || v.i < 0 || v.i > 0xFFFFFFFFll) bad(where, "expected a u32");
  return (uint32_t)v.i;
}

const std::string& str_of(const JVal& v, const std::string& where) {
  if (v.kind != JVal::STR) bad(where, "expected a string");
  return v.s;
}

// {"kind": ..., "data": ...}
std::pair<const std::string*, const JVal*> tagged(const JVal& v, const std::string& where) {
  const JVal& k = field(v, "kind", where);
  const JVal& d = field(v, "data", where);
  return {&str_of(k, where + ".kind"), &d};
}

void decode_term(Decoded& D, const JVal& t, const std::string& where) {
  auto [kind, d] = tagged(t, where);
  sk_term out{};
  if (*kind == "var") {
    out.kind = SK_T_VAR;
    out.a = u32_of(field(*d, "slot", where), where + ".slot");
  } else if (*kind == "const") {
    out.kind = SK_T_CONST;
    out.a = u32_of(field(*d, "sym", where), where + ".sym");
  } else if (*kind == "int") {
    const JVal& v = field(*d, "value", where);
    if (v.kind != JVal::INT || v.big) bad(where + ".value", "expected an i64");
    out.kind = SK_T_INT;
    out.b = v.i;
  } else if (*kind == "agg") {
    static const char* ops[] = {"count", "sum", "min", "max", "prob_or"};
    const std::string& op = str_of(field(*d, "op", where), where + ".op");
    int k = -1;
    for (int q = 0; q < 5; q++)
      if (op == ops[q]) k = q;
    if (k < 0) bad(where + ".op", "unknown aggregate `" + op + "`");
    out.kind = SK_T_AGG;
    out.a = (uint32_t)k;
    out.b = u32_of(field(*d, "slot", where), where + ".slot");
  } else if (*kind == "compound") {
    out.kind = SK_T_COMPOUND;  // refused when the program is prepared
  } else {
    bad(where, "unknown term kind `" + *kind + "`");
  }
  D.terms.push_back(out);
}

void decode_atom(Decoded& D, const JVal& a, bool negated, const std::string& where) {
  const std::string& pred = str_of(field(a, "pred", where), where + ".pred");
  const JVal& args = field(a, "args", where);
  if (args.kind != JVal::ARR) bad(where + ".args", "expected an array");
  sk_atom at{};
  at.negated = negated ? 1u : 0u;
  at.first_term = (uint32_t)D.terms.size();
  at.n_terms = (uint32_t)args.a.size();
  for (size_t i = 0; i < args.a.size(); i++) decode_term(D, args.a[i], where + ".args[" + std::to_string(i) + "]");
  D.names.push_back(pred);
  D.atom_name.push_back(D.names.size() - 1);
  D.atoms.push_back(at);
}

void decode(Decoded& D, const JVal& doc) {
  const JVal& preds = field(doc, "predicates", "program");
  const JVal& rules = field(doc, "rules", "program");
  if (preds.kind != JVal::ARR) bad("predicates", "expected an array");
  if (rules.kind != JVal::ARR) bad("rules", "expected an array");
  std::vector<size_t> pred_name;
  for (size_t i = 0; i < preds.a.size(); i++) {
    const std::string w = "predicates[" + std::to_string(i) + "]";
    const JVal& p = preds.a[i];
    sk_pred d{};
    D.names.push_back(str_of(field(p, "name", w), w + ".name"));
    pred_name.push_back(D.names.size() - 1);
    d.arity = u32_of(field(p, "arity", w), w + ".arity");
    const std::string& sr = str_of(field(p, "semiring", w), w + ".semiring");
    if (sr == "bool") d.semiring = SK_BOOL;
    else if (sr == "trop") d.semiring = SK_TROP;
    else bad(w + ".semiring", "expected `bool` or `trop`");
    d.stratum = u32_of(field(p, "stratum", w), w + ".stratum");
    D.preds.push_back(d);
  }
  for (size_t i = 0; i < rules.a.size(); i++) {
    const std::string w = "rules[" + std::to_string(i) + "]";
    const JVal& r = rules.a[i];
    sk_rule out{};
    out.head = (uint32_t)D.atoms.size();
    decode_atom(D, field(r, "head", w), false, w + ".head");
    const JVal& body = field(r, "body", w);
    if (body.kind != JVal::ARR) bad(w + ".body", "expected an array");
    out.first_lit = (uint32_t)D.atoms.size();
    out.n_lits = (uint32_t)body.a.size();
    for (size_t j = 0; j < body.a.size(); j++) {
      const std::string wl = w + ".body[" + std::to_string(j) + "]";
      auto [kind, a] = tagged(body.a[j], wl);
      if (*kind != "pos" && *kind != "neg") bad(wl, "unknown literal kind `" + *kind + "`");
      decode_atom(D, *a, *kind == "neg", wl);
    }
    out.stratum = u32_of(field(r, "stratum", w), w + ".stratum");
    out.var_count = u32_of(field(r, "var_count", w), w + ".var_count");
    if (const JVal* f = r.get("neg_weight_cycle_check"))
      if (f->kind != JVal::BOOL) bad(w + ".neg_weight_cycle_check", "expected a boolean");
    D.rules.push_back(out);
  }
  D.prog.num_strata = u32_of(field(doc, "num_strata", "program"), "num_strata");
  // pointers into the now final storage
  for (size_t i = 0; i < D.preds.size(); i++) D.preds[i].name = D.names[pred_name[i]].c_str();
  for (size_t i = 0; i < D.atoms.size(); i++) D.atoms[i].pred = D.names[D.atom_name[i]].c_str();
  D.prog.preds = D.preds.data();
  D.prog.n_preds = (uint32_t)D.preds.size();
  D.prog.rules = D.rules.data();
  D.prog.n_rules = (uint32_t)D.rules.size();
  D.prog.atoms = D.atoms.data();
  D.prog.n_atoms = (uint32_t)D.atoms.size();
  D.prog.terms = D.terms.data();
  D.prog.n_terms = (uint32_t)D.terms.size();
}

}  // namespace

struct sk_program_json {
  Decoded d;
};

extern "C" {

sk_status sk_program_from_json(const char* json, size_t len, sk_program_json** out, char* err, size_t err_len) {
  if (!json || !out) return SK_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  auto h = std::make_unique<sk_program_json>();
  try {
    Parser ps(json, len);
    const JVal doc = ps.parse();
    decode(h->d, doc);
  } catch (const JErr& e) {
    if (err && err_len) std::snprintf(err, err_len, "%s", e.msg.c_str());
    return SK_ERR_INVALID_ARGUMENT;
  } catch (const std::bad_alloc&) {
    if (err && err_len) std::snprintf(err, err_len, "out of host memory");
    return SK_ERR_INVALID_ARGUMENT;
  }
  if (err && err_len) err[0] = 0;
  *out = h.release();
  return SK_OK;
}

const sk_program* sk_program_json_get(const sk_program_json* p) { return p ? &p->d.prog : nullptr; }

void sk_program_json_free(sk_program_json* p) { delete p; }

}  // extern "C"
